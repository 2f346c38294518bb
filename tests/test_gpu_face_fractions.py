"""The face-fraction option (FS_STEP_FACE_FRACTIONS), HIP path vs the CPU
oracle.  It has no reference counterpart; the oracle's restatement of it is
pinned in test_oracle_face_fractions.py.

Bars:
  - open fractions of every face bit-exact (the same FP32 operations on both
    sides: SDF trilinear at the corners, two triangles per face);
  - right-hand side and projection bit-exact;
  - the pressure of the fraction-weighted matrix within 1e-6 (relative L2) of
    the oracle's direct solve, with and without AIR (Dirichlet) contact;
  - a falling sphere through the water impact: body position within 0.01 dx,
    99.9% of the particles within 1e-4 and at most 1e-4 of them further than
    1e-3 (as for the voxel coupling).
"""
import numpy as np
import pytest

from body_cases import box_sdf
from fluidprjct_amd import scenes
from fluidprjct_amd._abi import (FS_GRID_LABELS, FS_GRID_OPEN_U, FS_GRID_OPEN_V, FS_GRID_OPEN_W, FS_GRID_P,
                                 FS_GRID_P64, FS_GRID_RHS, FS_GRID_U, FS_GRID_V, FS_GRID_W, FS_STEP_FACE_FRACTIONS)

pytestmark = pytest.mark.gpu

OPEN = (FS_GRID_OPEN_U, FS_GRID_OPEN_V, FS_GRID_OPEN_W)
VEL = (FS_GRID_U, FS_GRID_V, FS_GRID_W)


def body_scene(lib, N, body, level=0.6, angles=(0.2, -0.3, 0.4), vel=(0.3, -0.5, 0.2), omega=(1.0, -2.0, 0.5),
               tol=1e-8, solver_flags=0):
    dx = 1.0 / N
    s = lib.solver(dx=dx, dt=1.0 / 120, pic_weight=0.05, solver_tol=tol, step_flags=FS_STEP_FACE_FRACTIONS,
                   solver_flags=solver_flags)
    s.set_labels(scenes.water_column(N, level))
    if body == "sphere":
        r = 0.17
        sdf, origin, cs = scenes.sphere_sdf(r, dx / 2)
        s.set_body_sdf(sdf, origin, cs, (0.48, 0.52, 0.51), scenes.sphere_mass(r, 500.0))
    else:
        sdf, origin = box_sdf((0.3, 0.08, 0.22), dx / 3, angle_z=0.5)
        s.set_body_sdf(sdf, origin, dx / 3, (0.5, 0.45, 0.47), 3.0)
    st = s.body_state()
    st.vel[:] = vel
    st.omega[:] = omega
    st.theta, st.psi, st.phi = angles
    s.set_body_state(st)
    s.seed_particles(2, 9)
    s.stage("label")
    return s


@pytest.mark.parametrize("body", ["sphere", "box"])
def test_open_fractions_bitexact(hip, oracle, body):
    N = 40
    h, o = (body_scene(lib, N, body) for lib in (hip, oracle))
    assert np.array_equal(h.get_grid(FS_GRID_LABELS), o.get_grid(FS_GRID_LABELS))
    partial = 0
    for g in OPEN:
        th, to = h.get_grid(g), o.get_grid(g)
        assert np.array_equal(th.view(np.uint32), to.view(np.uint32)), g
        partial += int(((to > 0) & (to < 1)).sum())
    assert partial > 200, partial


@pytest.mark.parametrize("body", ["sphere", "box"])
def test_rhs_and_project_bitexact_with_face_fractions(hip, oracle, body):
    N = 40
    rng = np.random.default_rng(17)
    sols = [body_scene(lib, N, body) for lib in (hip, oracle)]
    lab = sols[1].get_grid(FS_GRID_LABELS)
    fields = {}
    for g in VEL:
        shape, _ = sols[1].grid_shape(g)
        fields[g] = rng.standard_normal(shape).astype(np.float32)
    p = (rng.standard_normal(lab.shape) * 50).astype(np.float32) * (lab == 1)
    for s in sols:
        for g, a in fields.items():
            s.set_grid(g, a)
        s.stage("rhs")
    h, o = sols
    assert np.array_equal(h.get_grid(FS_GRID_RHS).view(np.uint32), o.get_grid(FS_GRID_RHS).view(np.uint32))
    for s in sols:
        s.set_grid(FS_GRID_P, p)
        s.stage("project")
    for g in VEL:
        assert np.array_equal(h.get_grid(g).view(np.uint32), o.get_grid(g).view(np.uint32)), g


@pytest.mark.parametrize("level", [0.6, 1.0])   # free surface (Dirichlet) / tank full of fluid (singular)
def test_pressure_with_face_fractions_vs_oracle(hip, oracle, level):
    N = 24
    rng = np.random.default_rng(23)
    sols = [body_scene(lib, N, "sphere", level=level, tol=1e-10) for lib in (hip, oracle)]
    lab = sols[1].get_grid(FS_GRID_LABELS)
    assert np.array_equal(sols[0].get_grid(FS_GRID_LABELS), lab)
    if level == 1.0:   # cells left without particles become FLUID too: no AIR anywhere
        lab[lab == 2] = 1
        for s in sols:
            s.set_grid(FS_GRID_LABELS, lab)
    b = (rng.standard_normal(lab.shape) * 1e3).astype(np.float32) * (lab == 1)
    res = []
    for s in sols:
        s.set_grid(FS_GRID_RHS, b)
        s.stage("pressure")
        res.append((s.get_grid(FS_GRID_P64), s.solve_stats()))
    (ph, sh), (po, so) = res
    assert sh.converged == 1 and sh.singular == so.singular == (1 if level == 1.0 else 0)
    assert np.linalg.norm(ph - po) / np.linalg.norm(po) <= 1e-6, np.linalg.norm(ph - po) / np.linalg.norm(po)


def test_blocked_smoothing_bitexact_with_face_fractions(hip):
    """The temporally blocked smoothing of the small coarse levels against one
    launch per sweep (FS_SOLVER_NO_FUSE) with the fraction-weighted fine
    level of a 72^3 tank (about 230K cells; the sphere gives it partly open
    faces): the same solution bits and iteration count."""
    from fluidprjct_amd._abi import FS_SOLVER_NO_FUSE
    N = 72
    res = []
    for flags in (0, FS_SOLVER_NO_FUSE):
        s = body_scene(hip, N, "sphere", tol=1e-8, solver_flags=flags)
        lab = s.get_grid(FS_GRID_LABELS)
        b = (np.random.default_rng(29).standard_normal(lab.shape) * 1e3).astype(np.float32) * (lab == 1)
        s.set_grid(FS_GRID_RHS, b)
        s.stage("pressure")
        res.append((s.get_grid(FS_GRID_P64), s.solve_stats()))
        partial = sum(int(((g > 0) & (g < 1)).sum()) for g in (s.get_grid(o) for o in OPEN))
        assert partial > 500, partial
        s.close()
    (pf, sf), (pn, sn) = res
    assert sf.converged == 1 and sn.converged == 1
    assert sf.iterations == sn.iterations
    assert np.array_equal(pf.view(np.uint64), pn.view(np.uint64))


def test_falling_sphere_with_face_fractions(hip, oracle):
    """The falling-sphere workload at 48^3 with the option on, through the
    water impact (near step 31), with the bars of the voxel coupling's test."""
    N, steps = 48, 50
    dx = 1.0 / N
    r = 0.08
    sols = []
    for lib in (hip, oracle):
        s = lib.solver(dx=dx, dt=1.0 / 120, pic_weight=0.05, cfl=1.0, solver_tol=1e-8,
                       step_flags=FS_STEP_FACE_FRACTIONS)
        s.set_labels(scenes.water_column(N, 0.4))
        sdf, origin, cs = scenes.sphere_sdf(r, dx)
        s.set_body_sdf(sdf, origin, cs, (0.5, 0.8, 0.5), scenes.sphere_mass(r, 500.0))
        s.seed_particles(8, 3)
        sols.append(s)
    h, o = sols
    impact = None
    for step in range(steps):
        for s in sols:
            s.step(1)
        bh, bo = h.body_state(), o.body_state()
        if impact is None and bo.submerged_volume > 0:
            impact = step
        dpos = np.abs(np.array(bh.pos) - np.array(bo.pos)).max()
        assert dpos <= 0.01 * dx, (step, dpos)
        if step % 10 == 9 or step == steps - 1:
            ph, _ = h.particles()
            po, _ = o.particles()
            e = np.abs(ph - po).max(axis=1)
            q999, far = np.quantile(e, 0.999), float((e > 1e-3).mean())
            print(f"step {step}: body y {bo.pos[1]:.5f} |dbody| {dpos:.2e} |dparticles| max {e.max():.2e} "
                  f"99.9% {q999:.2e} >1e-3: {far:.1e}", flush=True)
            assert q999 <= 1e-4 and far <= 1e-4, (step, q999, far)
    assert impact is not None and impact < steps - 5, impact
