"""HIP path vs the CPU oracle, stage by stage and over whole steps.

Bars:
  - bit-exact: seeding, particle -> cell keys, labels, RHS, pressure
    projection, G2P + advection (same inputs on both sides; both sides keep the
    reference's FP32 operation order without contraction);
  - P2G: relative 2e-6 of the face magnitude (FP32 sums per halo cell, added
    per face in a fixed order, vs the oracle's FP64 sums in particle order);
  - pressure: ||p - p_ref|| / ||p_ref|| <= 1e-5 against FP64 direct solves
    (scipy spsolve golden vectors and the oracle's band LDL^T);
  - whole steps: particle positions within 2e-5 * domain after 3 steps.
"""
import os

import numpy as np
import pytest

from fluidprjct_amd._abi import (FS_GRID_CELLKEY, FS_GRID_LABELS, FS_GRID_P, FS_GRID_P64, FS_GRID_RHS, FS_GRID_U,
                                 FS_GRID_USAVED, FS_GRID_V, FS_GRID_VSAVED, FS_GRID_W, FS_GRID_WSAVED)

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def tank(N, block=(0.25, 0.5, 1.0), shape=None):
    D, H, W = shape or (N, N, N)
    lab = np.full((D, H, W), 2, np.uint8)
    lab[0], lab[-1] = 0, 0
    lab[:, 0], lab[:, -1] = 0, 0
    lab[:, :, 0], lab[:, :, -1] = 0, 0
    bx, by, bz = block
    lab[1:max(2, int(bz * D)) - (1 if bz >= 1 else 0), 1:max(2, int(by * H)), 1:max(2, int(bx * W))] = 1
    return lab


def pair(hip, oracle, **cfg):
    base = dict(dt=1.0 / 120, dx=1.0 / 32, pic_weight=0.05, solver_tol=1e-6)
    base.update(cfg)
    return hip.solver(**base), oracle.solver(**base)


def random_particles(rng, n, W, H, D, dx, margin=0.0):
    lo = np.array([-margin, -margin, -margin]) * dx
    hi = np.array([W + margin, H + margin, D + margin]) * dx
    pos = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    vel = rng.standard_normal((n, 3)).astype(np.float32)
    return pos, vel


def test_seed_bitexact(hip, oracle):
    lab = tank(20)
    h, o = pair(hip, oracle, dx=0.05)
    for s in (h, o):
        s.set_labels(lab)
        s.seed_particles(8, 1234)
    ph, vh = h.particles()
    po, vo = o.particles()
    assert ph.shape == po.shape and ph.shape[0] == 8 * int((lab == 1).sum())
    assert np.array_equal(ph.view(np.uint32), po.view(np.uint32))
    assert not vh.any() and not vo.any()


def test_keys_and_labels_bitexact(hip, oracle):
    rng = np.random.default_rng(1)
    W, H, D = 19, 23, 17
    lab = np.full((D, H, W), 2, np.uint8)
    lab[0], lab[:, :, 0], lab[:, -1] = 0, 0, 0
    lab[5:9, 4:8, 6:10] = 0
    dx = 0.037
    pos, vel = random_particles(rng, 20000, W, H, D, dx, margin=0.3)
    # particles exactly on cell faces and on the domain faces
    pos[:500] = (rng.integers(0, 18, (500, 3)) * np.float32(dx)).astype(np.float32)
    h, o = pair(hip, oracle, dx=dx)
    for s in (h, o):
        s.set_labels(lab)
        s.set_particles(pos, vel)
        s.stage("label")
    assert np.array_equal(h.get_grid(FS_GRID_LABELS), o.get_grid(FS_GRID_LABELS))
    assert np.array_equal(h.get_grid(FS_GRID_CELLKEY), o.get_grid(FS_GRID_CELLKEY))
    # sorting must not lose or alter particles
    ph, vh = h.particles()
    assert np.array_equal(ph, pos) and np.array_equal(vh, vel)


def test_body_labels_bitexact(hip, oracle):
    """BODY cells of a rotated, off-centre box-shaped body (the SDF is not
    symmetric, so the rotation and the local centre of mass both matter)
    among FLUID and AIR cells: same labels as the oracle."""
    rng = np.random.default_rng(8)
    N = 40
    dx = 1.0 / N
    lab = np.full((N, N, N), 2, np.uint8)
    lab[0], lab[:, 0], lab[:, :, 0] = 0, 0, 0
    pos, vel = random_particles(rng, 60000, N, N, N, dx)
    cs = 0.5 * dx
    n = (36, 20, 28)                       # SDF voxels along z, y, x
    z, y, x = np.meshgrid(*[(np.arange(m) + 0.5) * cs - 0.5 * m * cs for m in n], indexing="ij")
    sdf = np.maximum(np.maximum(np.abs(x) - 0.28, np.abs(y) - 0.15), np.abs(z) - 0.37).astype(np.float32)
    sdf -= np.float32(0.02) * (x > 0.1)    # off-centre mass
    origin = [-0.5 * n[2] * cs, -0.5 * n[1] * cs, -0.5 * n[0] * cs]
    h, o = pair(hip, oracle, dx=dx)
    for s in (h, o):
        s.set_labels(lab)
        s.set_body_sdf(sdf, origin, cs, (0.47, 0.52, 0.55), 1.0)
        st = s.body_state()
        st.theta, st.psi, st.phi = 0.3, -0.5, 0.7
        s.set_body_state(st)
        s.set_particles(pos, vel)
        s.stage("label")
    lh, lo = h.get_grid(FS_GRID_LABELS), o.get_grid(FS_GRID_LABELS)
    assert (lo == 3).sum() > 200 and (lo == 1).sum() > 1000
    assert np.array_equal(lh, lo)


def test_p2g_matches_oracle(hip, oracle):
    rng = np.random.default_rng(2)
    W, H, D = 37, 21, 18   # not multiples of the LDS tile
    dx = 0.05
    lab = tank(0, shape=(D, H, W))
    pos, vel = random_particles(rng, 60000, W, H, D, dx, margin=0.2)
    h, o = pair(hip, oracle, dx=dx, dt=0.01)
    for s in (h, o):
        s.set_labels(lab)
        s.set_particles(pos, vel)
        s.stage("label")
        s.stage("p2g")
    for g in (FS_GRID_U, FS_GRID_V, FS_GRID_W, FS_GRID_USAVED, FS_GRID_VSAVED, FS_GRID_WSAVED):
        a, b = h.get_grid(g), o.get_grid(g)
        scale = max(1.0, float(np.abs(b).max()))
        assert np.abs(a - b).max() <= 2e-6 * scale, (g, np.abs(a - b).max())


# one column block, and a grid that is ragged against the 64 x 4 column blocks
# with more than one z chunk per column and BODY cells next to the faces
@pytest.mark.parametrize("W,H,D,labels", [(16, 14, 12, (0, 1, 1, 1, 2)), (70, 9, 37, (0, 1, 1, 1, 2, 3))])
def test_rhs_and_project_bitexact(hip, oracle, W, H, D, labels):
    rng = np.random.default_rng(3)
    lab = rng.choice(np.array(labels, np.uint8), size=(D, H, W))
    h, o = pair(hip, oracle, dx=0.1, dt=0.02)
    shapes = {FS_GRID_U: (D, H, W + 1), FS_GRID_V: (D, H + 1, W), FS_GRID_W: (D + 1, H, W)}
    fields = {g: rng.standard_normal(s).astype(np.float32) for g, s in shapes.items()}
    p = (rng.standard_normal((D, H, W)) * 50).astype(np.float32) * (lab == 1)
    for s in (h, o):
        s.set_labels(lab)
        if 3 in labels:   # a moving rigid body: BODY faces take V + Omega x r
            st = s.body_state()
            st.pos[:] = (0.5 * W * 0.1, 0.4 * H * 0.1, 0.6 * D * 0.1)
            st.vel[:] = (0.3, -0.2, 0.1)
            st.omega[:] = (0.5, 0.7, -0.4)
            s.set_body_state(st)
        for g, a in fields.items():
            s.set_grid(g, a)
        s.stage("rhs")
    assert np.array_equal(h.get_grid(FS_GRID_RHS).view(np.uint32), o.get_grid(FS_GRID_RHS).view(np.uint32))
    for s in (h, o):
        s.set_grid(FS_GRID_P, p)
        s.stage("project")
    for g in shapes:
        assert np.array_equal(h.get_grid(g).view(np.uint32), o.get_grid(g).view(np.uint32)), g


# 12 m/s leaves the staged LDS margin (global path); dx = 1/16 takes the
# power-of-two path (x * (1/dx) for x / dx), dx = 1/20 the division path
@pytest.mark.parametrize("speed", [0.7, 12.0])
@pytest.mark.parametrize("dx", [1.0 / 20, 1.0 / 16])
def test_g2p_advect_bitexact(hip, oracle, speed, dx):
    rng = np.random.default_rng(4)
    N = 20
    lab = tank(N, block=(0.5, 0.5, 1.0))
    lab[8:11, 3:6, 8:12] = 0   # an interior obstacle: exercises the projection
    pos, vel = random_particles(rng, 30000, N - 2, N - 2, N - 2, dx)
    pos += dx
    h, o = pair(hip, oracle, dx=dx, dt=1.0 / 60, pic_weight=0.05)
    shapes = {FS_GRID_U: (N, N, N + 1), FS_GRID_V: (N, N + 1, N), FS_GRID_W: (N + 1, N, N)}
    f = {g: (rng.standard_normal(s) * speed).astype(np.float32) for g, s in shapes.items()}
    fs = {g: (a + rng.standard_normal(a.shape).astype(np.float32) * 0.1) for g, a in f.items()}
    saved = {FS_GRID_U: FS_GRID_USAVED, FS_GRID_V: FS_GRID_VSAVED, FS_GRID_W: FS_GRID_WSAVED}
    for s in (h, o):
        s.set_labels(lab)
        s.set_particles(pos, vel)
        for g in shapes:
            s.set_grid(g, f[g])
            s.set_grid(saved[g], fs[g])
        s.stage("g2p_advect")
    ph, vh = h.particles()
    po, vo = o.particles()
    assert np.array_equal(vh.view(np.uint32), vo.view(np.uint32))
    bad = np.flatnonzero((ph.view(np.uint32) != po.view(np.uint32)).any(1))
    assert bad.size == 0, (bad[:10], ph[bad[:3]], po[bad[:3]])


# a grid ragged against the G2P tiles on every axis, with BODY cells in the
# obstacle (the projection search treats them as solid)
@pytest.mark.parametrize("dx", [1.0 / 40, 1.0 / 32])
def test_g2p_advect_bitexact_ragged_body(hip, oracle, dx):
    rng = np.random.default_rng(12)
    D, H, W = 29, 23, 37
    lab = tank(0, block=(0.6, 0.5, 1.0), shape=(D, H, W))
    lab[10:16, 4:9, 12:20] = 3
    lab[12:14, 5:7, 14:16] = 0
    pos, vel = random_particles(rng, 50000, W - 2, H - 2, D - 2, dx)
    pos += dx
    h, o = pair(hip, oracle, dx=dx, dt=1.0 / 60, pic_weight=0.05)
    shapes = {FS_GRID_U: (D, H, W + 1), FS_GRID_V: (D, H + 1, W), FS_GRID_W: (D + 1, H, W)}
    f = {g: (rng.standard_normal(s) * 3.0).astype(np.float32) for g, s in shapes.items()}
    fs = {g: (a + rng.standard_normal(a.shape).astype(np.float32) * 0.1) for g, a in f.items()}
    saved = {FS_GRID_U: FS_GRID_USAVED, FS_GRID_V: FS_GRID_VSAVED, FS_GRID_W: FS_GRID_WSAVED}
    for s in (h, o):
        s.set_labels(lab)
        s.set_particles(pos, vel)
        for g in shapes:
            s.set_grid(g, f[g])
            s.set_grid(saved[g], fs[g])
        s.stage("g2p_advect")
    ph, vh = h.particles()
    po, vo = o.particles()
    assert np.array_equal(vh.view(np.uint32), vo.view(np.uint32))
    bad = np.flatnonzero((ph.view(np.uint32) != po.view(np.uint32)).any(1))
    assert bad.size == 0, (bad[:10], ph[bad[:3]], po[bad[:3]])


@pytest.mark.parametrize("name", ["dam", "mix"])
def test_pressure_vs_golden(hip, name):
    g = np.load(os.path.join(GOLDEN, "pressure_golden.npz"))
    lab, b, pref = g[name + "_labels"], g[name + "_b"], g[name + "_p"]
    h = hip.solver()
    p, st = h.solve_pressure_system(lab, b, 1e-6)
    rel = np.linalg.norm(p - pref) / np.linalg.norm(pref)
    assert st.converged == 1 and st.rel_residual <= 1e-6
    assert rel <= 1e-5, rel
    p, st = h.solve_pressure_system(lab, b, 1e-12)
    assert np.linalg.norm(p - pref) / np.linalg.norm(pref) <= 1e-9


def test_pressure_vs_oracle_direct(hip, oracle):
    rng = np.random.default_rng(5)
    N = 40
    lab = tank(N, block=(0.6, 0.4, 1.0))
    b = (rng.standard_normal(lab.shape) * 1e3).astype(np.float32) * (lab == 1)
    o = oracle.solver()
    po, so = o.solve_pressure_system(lab, b, 1e-13)
    h = hip.solver()
    ph, sh = h.solve_pressure_system(lab, b, 1e-6)
    assert sh.converged == 1
    assert np.linalg.norm(ph - po) / np.linalg.norm(po) <= 1e-5


def test_full_steps_close(hip, oracle):
    N = 32
    lab = tank(N)
    h, o = pair(hip, oracle, dx=1.0 / N, dt=1.0 / 120, solver_tol=1e-8)
    for s in (h, o):
        s.set_labels(lab)
        s.seed_particles(8, 3)
    for step in range(3):
        for s in (h, o):
            s.step(1)
        ph, vh = h.particles()
        po, vo = o.particles()
        assert np.abs(ph - po).max() <= 2e-5, (step, np.abs(ph - po).max())
        assert np.abs(vh - vo).max() <= 2e-3 * max(1.0, np.abs(vo).max()), (step, np.abs(vh - vo).max())
    lh, lo = h.get_grid(FS_GRID_LABELS), o.get_grid(FS_GRID_LABELS)
    assert (lh != lo).sum() <= 2


def test_graph_replays_match_eager(hip):
    """The inner PCG batch replayed as a cached HIP graph gives bit-identical
    solutions to eager launches, over repeated solves with other work between."""
    from fluidprjct_amd._abi import FS_SOLVER_NO_GRAPH
    rng = np.random.default_rng(11)
    N = 48
    labs = [tank(N, block=(0.6, 0.5, 1.0)), tank(N, block=(0.3, 0.7, 0.8))]
    g = hip.solver()
    e = hip.solver(solver_flags=FS_SOLVER_NO_GRAPH)
    for rep in range(4):
        lab = labs[rep % 2]
        b = (rng.standard_normal(lab.shape) * 1e3).astype(np.float32) * (lab == 1)
        pg, sg = g.solve_pressure_system(lab, b, 1e-7)
        pe, se = e.solve_pressure_system(lab, b, 1e-7)
        assert sg.iterations == se.iterations and sg.converged == 1
        assert np.array_equal(pg, pe), rep


@pytest.mark.parametrize("N,block", [(64, (0.6, 0.5, 1.0)), (96, (1.0, 0.4, 1.0)), (128, (1.0, 0.4, 1.0))])
def test_tail_matches_level_launches(hip, N, block):
    """The V-cycle's small levels run in one workgroup with the iterates in LDS;
    the result is bit-identical to running each of those levels as separate
    kernel launches (same operations in the same floating-point order)."""
    from fluidprjct_amd._abi import FS_SOLVER_NO_TAIL
    rng = np.random.default_rng(N)
    lab = tank(N, block=block)
    b = (rng.standard_normal(lab.shape) * 1e3).astype(np.float32) * (lab == 1)
    pt, st = hip.solver().solve_pressure_system(lab, b, 1e-7)
    pn, sn = hip.solver(solver_flags=FS_SOLVER_NO_TAIL).solve_pressure_system(lab, b, 1e-7)
    assert st.converged == 1 and sn.converged == 1
    assert st.iterations == sn.iterations
    assert np.array_equal(pt.view(np.uint32), pn.view(np.uint32))


@pytest.mark.parametrize("shape", [(64, 64, 1), (40, 40, 2)])
def test_tail_matches_level_launches_thin_slab(hip, shape):
    """A solve box of at most 4096 cells runs its whole V-cycle, the fine level
    included, in the LDS tail (r.z from a separate pass): the same iterations
    as the per-level launches and the solution within the rounding of the
    differently ordered r.z partial sums."""
    from fluidprjct_amd import scenes
    from fluidprjct_amd._abi import FS_SOLVER_NO_TAIL
    W, H, D = shape
    lab = np.full((D, H, W), 2, np.uint8)
    lab[:, 0, :] = 0
    lab[:, :, 0] = 0
    lab[:, :, -1] = 0
    lab[:, 1:H // 2, 1:W // 2] = 1
    assert (lab == 1).sum() * 1.0 <= 4096
    rng = np.random.default_rng(W)
    b = (rng.standard_normal(lab.shape) * 1e3).astype(np.float32) * (lab == 1)
    pt, st = hip.solver().solve_pressure_system(lab, b, 1e-7)
    pn, sn = hip.solver(solver_flags=FS_SOLVER_NO_TAIL).solve_pressure_system(lab, b, 1e-7)
    assert st.converged == 1 and sn.converged == 1
    assert st.iterations == sn.iterations
    assert np.abs(pt - pn).max() <= 1e-9 * np.abs(pn).max()


def test_c1_slab_steps_close(hip, oracle):
    """The 2D dam break (128^2, 4 ppc) as a one-cell-deep slab: the HIP step
    and the oracle stay within the whole-step bars for 5 steps, and the slab
    stays planar (z = 0, w = 0)."""
    from fluidprjct_amd import scenes
    N = 128
    lab = scenes.dam_break_2d(N, (0.25, 0.5))
    h, o = pair(hip, oracle, dx=1.0 / N, dt=1.0 / 120, solver_tol=1e-8)
    for s in (h, o):
        s.set_labels(lab)
        s.seed_particles(4, 1)
    for step in range(5):
        for s in (h, o):
            s.step(1)
        ph, vh = h.particles()
        po, vo = o.particles()
        assert ph.shape == po.shape
        assert np.abs(ph - po).max() <= 2e-5, (step, np.abs(ph - po).max())
        assert np.abs(vh - vo).max() <= 2e-3 * max(1.0, np.abs(vo).max()), (step, np.abs(vh - vo).max())
    assert not ph[:, 2].any() and not vh[:, 2].any()
    assert not h.get_grid(FS_GRID_W).any()
    assert (h.get_grid(FS_GRID_LABELS) != o.get_grid(FS_GRID_LABELS)).sum() <= 2


# 192: a level of more than 256K cells below the fine one (per-sweep launches there)
@pytest.mark.parametrize("N,block", [(96, (1.0, 0.4, 1.0)), (128, (0.6, 0.5, 1.0)), (192, (1.0, 0.4, 1.0))])
def test_blocked_smoothing_matches_per_sweep_launches(hip, N, block):
    """On the small coarse levels all pre-smoothing sweeps + the residual, and
    the prolongation + all post-smoothing sweeps, run as one temporally blocked
    kernel each; the solution is bit-identical to one launch per sweep."""
    from fluidprjct_amd._abi import FS_SOLVER_NO_FUSE
    rng = np.random.default_rng(N + 1)
    lab = tank(N, block=block)
    b = (rng.standard_normal(lab.shape) * 1e3).astype(np.float32) * (lab == 1)
    pf, sf = hip.solver().solve_pressure_system(lab, b, 1e-7)
    pn, sn = hip.solver(solver_flags=FS_SOLVER_NO_FUSE).solve_pressure_system(lab, b, 1e-7)
    assert sf.converged == 1 and sn.converged == 1
    assert sf.iterations == sn.iterations
    assert np.array_equal(pf.view(np.uint32), pn.view(np.uint32))
