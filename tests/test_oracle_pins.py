"""Pins of the CPU oracle: fixtures of the reference, closed forms, hand-computed
tiny cases and an independent FP64 solve (scipy).  Runs without a GPU."""
import ctypes as C
import os

import numpy as np
import pytest

from fluidprjct_amd._abi import (FS_GRID_LABELS, FS_GRID_P, FS_GRID_P64, FS_GRID_RHS, FS_GRID_U, FS_GRID_USAVED,
                                 FS_GRID_V, FS_GRID_VSAVED, FS_GRID_W, FS_GRID_WSAVED)
from fluidprjct_amd import scenes
from body_cases import DIRS as _DIRS, one_face_body

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
INPUTS = os.path.join(GOLDEN, "inputdata")


# ---------------------------------------------------------------- scene input
def test_label_file_counts(oracle):
    s = oracle.solver()
    s.load_labels_file(os.path.join(INPUTS, "labels_3d_sphere.txt"))
    L = s.get_grid(FS_GRID_LABELS)
    assert s.dims == (35, 35, 35)
    assert [(L == v).sum() for v in (1, 0, 2)] == [7570, 20276, 15029]
    s.load_labels_file(os.path.join(INPUTS, "labels_bathymetry.txt"))
    L = s.get_grid(FS_GRID_LABELS)
    assert s.dims == (100, 40, 30)
    assert [(L == v).sum() for v in (1, 0, 2)] == [21944, 43612, 54444]


def test_label_file_reverse_order(oracle, tmp_path):
    # the first character is cell (W-1, H-1, D-1) (FluidSolver3D.cu:139-152); '?' leaves SOLID
    p = tmp_path / "l.txt"
    p.write_text("3 2 2\nF A S\nA A A\nS S ?\nA F A\n")
    s = oracle.solver()
    s.load_labels_file(str(p))
    L = s.get_grid(FS_GRID_LABELS)   # (D, H, W)
    seq = "FASAAASS?AFA"
    m = {"F": 1, "A": 2, "S": 0, "?": 0}
    for n, ch in enumerate(seq):
        idx = 3 * 2 * 2 - 1 - n
        k, rem = divmod(idx, 6)
        j, i = divmod(rem, 3)
        assert L[k, j, i] == m[ch], (n, ch)
    # the sphere tank's pool lies at the bottom (j = 4..13) only with reversed reading
    s.load_labels_file(os.path.join(INPUTS, "labels_3d_sphere.txt"))
    L = s.get_grid(FS_GRID_LABELS)
    js = np.nonzero((L == 1).any(axis=(0, 2)))[0]
    assert js.min() == 4 and js.max() == 22


def test_label_file_2d_reads_as_depth_one(oracle):
    s = oracle.solver()
    s.load_labels_file(os.path.join(INPUTS, "labels_simple.txt"))
    assert s.dims == (5, 4, 1)
    assert (s.get_grid(FS_GRID_LABELS) == 1).sum() == 6


def test_config_json_keys(oracle):
    cfg = oracle.default_config()
    assert (cfg.dt, cfg.dx, cfg.frame_amount, cfg.body_mass) == (pytest.approx(0.02), pytest.approx(0.1), 100, 50.0)
    assert cfg.pic_weight == pytest.approx(0.01) and cfg.iter_per_frame == 1 and cfg.active_body == 0
    rc = oracle.lib.fs_config_load_json(C.byref(cfg), os.path.join(INPUTS, "config.json").encode())
    assert rc == 0
    assert cfg.dt == pytest.approx(0.005) and cfg.dx == pytest.approx(0.02)
    assert cfg.pic_weight == pytest.approx(0.001) and cfg.frame_amount == 280 and cfg.active_body == 1
    assert cfg.input_label_data == b"InputData/labels_gaussian.txt"
    assert (cfg.body_x, cfg.body_y, cfg.body_z) == (pytest.approx(0.5), pytest.approx(0.2), pytest.approx(0.5))
    assert cfg.body_mass == pytest.approx(0.03)


# ---------------------------------------------------------------- seeding
def test_seeding_matches_reference_init_dump(oracle):
    """InputData/particles_-1.ply (FluidSolver3D.cu:201): 4 particles per FLUID
    cell, k->j->i cell order, each within +-0.49 dx of its cell centre (octant
    offset 0.25 dx + jitter U(-0.24, 0.24) dx)."""
    ref = np.load(os.path.join(GOLDEN, "particles_init.npz"))["pos"]
    dx = 0.02
    assert ref.shape == (44144, 3)
    cells = np.floor(ref / dx).astype(int)
    uniq, first, counts = np.unique(cells[:, ::-1], axis=0, return_index=True, return_counts=True)
    assert uniq.shape[0] == 11036 and (counts == 4).all()
    D, H, W = 50, 12, 50
    lab = np.zeros((D, H, W), np.uint8)
    lab[uniq[:, 0], uniq[:, 1], uniq[:, 2]] = 1
    s = oracle.solver(dx=dx)
    s.set_labels(lab)
    s.seed_particles(4, 99)
    pos, vel = s.particles()
    assert pos.shape == ref.shape and not vel.any()
    mine = np.floor(pos / dx).astype(int)
    # same cells, same order (k -> j -> i, 4 consecutive per cell)
    assert np.array_equal(mine, cells)
    for p in (pos, ref):
        off = p / dx - (np.floor(p / dx) + 0.5)
        assert np.abs(off).max() <= 0.49 + 1e-4
        assert np.abs(off).min() >= 0.01 - 1e-4
        # octants are drawn uniformly
        octant = (off > 0).astype(int) @ np.array([1, 2, 4])
        frac = np.bincount(octant, minlength=8) / octant.size
        assert np.abs(frac - 0.125).max() < 0.01


def test_seeding_is_deterministic_and_seed_dependent(oracle):
    lab = scenes.dam_break(12)
    a, b, c = (oracle.solver(dx=1 / 12) for _ in range(3))
    for s, seed in ((a, 5), (b, 5), (c, 6)):
        s.set_labels(lab)
        s.seed_particles(8, seed)
    pa, pb, pc = a.particles()[0], b.particles()[0], c.particles()[0]
    assert np.array_equal(pa, pb) and not np.array_equal(pa, pc)


# ---------------------------------------------------------------- pressure system
def test_cudss_5x5_known_answer(oracle):
    """cuDSStest (test.cu:262-290): SPD upper CSR, exact x = 1..5, tol 2e-15."""
    f = oracle.lib.oracle_spd_solve_upper_csr
    f.argtypes = [C.c_int] + [C.c_void_p] * 5
    off = np.array([0, 2, 4, 6, 7, 8], np.int32)
    col = np.array([0, 2, 1, 2, 2, 4, 3, 4], np.int32)
    val = np.array([4, 1, 3, 2, 5, 1, 1, 2], np.float64)
    b = np.array([7, 12, 25, 4, 13], np.float64)
    x = np.zeros(5)
    assert f(5, off.ctypes.data, col.ctypes.data, val.ctypes.data, b.ctypes.data, x.ctypes.data) == 0
    assert np.abs(x - np.arange(1, 6)).max() <= 2e-15


def test_reference_csr_of_enclosed_toy(oracle):
    """labels_simple3d.txt: 3x2 enclosed fluid (no AIR): unit-coefficient SPD
    matrix of MatrixFiller3D (FluidSolver3D.cu:1656-1716), upper triangle,
    diagonal first then neighbours with col > row in +x,-x,+y,-y,+z,-z order."""
    s = oracle.solver()
    s.load_labels_file(os.path.join(INPUTS, "labels_simple3d.txt"))
    lab = s.get_grid(FS_GRID_LABELS)
    f = oracle.lib.oracle_assemble_upper_csr
    f.restype = C.c_int64
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    nnz = f(lab.ctypes.data, 5, 4, 1, None, None, None)
    off = np.zeros(7, np.int32)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float32)
    f(lab.ctypes.data, 5, 4, 1, off.ctypes.data, col.ctypes.data, val.ctypes.data)
    # fluid numbering: (1,1)=0 (2,1)=1 (3,1)=2 (1,2)=3 (2,2)=4 (3,2)=5
    assert off.tolist() == [0, 3, 6, 8, 10, 12, 13]
    assert col.tolist() == [0, 1, 3, 1, 2, 4, 2, 5, 3, 4, 4, 5, 5]
    assert val.tolist() == [2, -1, -1, 3, -1, -1, 2, -1, 2, -1, 3, -1, 2]


@pytest.mark.parametrize("name", ["dam", "mix"])
def test_direct_solve_vs_scipy(oracle, name):
    g = np.load(os.path.join(GOLDEN, "pressure_golden.npz"))
    lab, b, pref = g[name + "_labels"], g[name + "_b"], g[name + "_p"]
    p, st = oracle.solver().solve_pressure_system(lab, b, 1e-13)
    assert np.abs(p - pref).max() <= 1e-10 * np.abs(pref).max()


def test_pcg_vs_direct(oracle):
    rng = np.random.default_rng(8)
    lab = scenes.dam_break(40, (0.6, 0.4, 1.0))
    b = (rng.standard_normal(lab.shape) * 1e3).astype(np.float32) * (lab == 1)
    s = oracle.solver()
    p_direct, st = s.solve_pressure_system(lab, b, 1e-13)
    assert st.iterations == 0           # band LDL^T path
    big = scenes.dam_break(44, (0.6, 0.8, 1.0))   # too large for the direct path
    b2 = (rng.standard_normal(big.shape) * 1e3).astype(np.float32) * (big == 1)
    p_cg, st2 = s.solve_pressure_system(big, b2, 1e-12)
    assert st2.iterations > 0 and st2.rel_residual <= 1e-12
    from scipy.sparse.linalg import spsolve
    import importlib.util
    spec = importlib.util.spec_from_file_location("mf", os.path.join(GOLDEN, "make_fixtures.py"))
    mf = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mf)
    A, fl = mf.assemble(big)
    x = spsolve(A.tocsc(), b2.reshape(-1)[fl].astype(np.float64))
    assert np.abs(p_cg.reshape(-1)[fl] - x).max() <= 1e-9 * np.abs(x).max()


def test_singular_enclosed_system_is_projected(oracle):
    s = oracle.solver()
    s.load_labels_file(os.path.join(INPUTS, "labels_simple3d.txt"))
    lab = s.get_grid(FS_GRID_LABELS)
    b = np.zeros(lab.shape, np.float32)
    b[lab == 1] = [1, 2, 3, -1, 0, 4]
    p, st = s.solve_pressure_system(lab, b, 1e-12)
    assert st.singular == 1 and st.converged == 1
    # A p = b - mean(b) on the fluid cells
    f = lab == 1
    bb = b[f] - b[f].mean()
    pp = p[f]
    A = np.array([[2, -1, 0, -1, 0, 0], [-1, 3, -1, 0, -1, 0], [0, -1, 2, 0, 0, -1],
                  [-1, 0, 0, 2, -1, 0], [0, -1, 0, -1, 3, -1], [0, 0, -1, 0, -1, 2]], float)
    assert np.abs(A @ pp - bb).max() <= 1e-9


def test_hydrostatic_pool(oracle):
    """Fluid layer of m cells under air at rest: b = rho g dx in the bottom
    layer, p_j = rho g dx (m - j + 1), and the projected v vanishes."""
    N, m = 12, 6
    lab = scenes.bordered(N, N, N)
    lab[1:-1, 1:1 + m, 1:-1] = 1
    dx, dt = 1.0 / N, 0.01
    s = oracle.solver(dx=dx, dt=dt, solver_tol=1e-12)
    s.set_labels(lab)
    s.seed_particles(8, 1)
    for st in ("label", "p2g", "rhs"):
        s.stage(st)
    b = s.get_grid(FS_GRID_RHS)
    rho_g_dx = 1000.0 * 9.81 * dx
    assert np.allclose(b[1:-1, 1, 1:-1], rho_g_dx, rtol=1e-5)
    assert np.abs(b[1:-1, 2:1 + m, 1:-1]).max() <= 1e-3
    s.stage("pressure")
    p = s.get_grid(FS_GRID_P64)
    for j in range(1, m + 1):
        assert np.allclose(p[1:-1, j, 1:-1], rho_g_dx * (m - j + 1), rtol=1e-6)
    s.stage("project")
    v = s.get_grid(FS_GRID_V)
    assert np.abs(v[1:-1, 1:m + 2, 1:-1]).max() <= 1e-5


# ---------------------------------------------------------------- transfers
def test_p2g_single_particle_weights(oracle):
    """One particle at (1.3, 2.6, 1.9) dx in a 4^3 grid: u-face base (1,2,1) with
    fractions (0.3, 0.1, 0.4); v-face base (0,2,1) (0.8, 0.6, 0.4); w-face base
    (0,2,1) (0.8, 0.1, 0.9) (MAC offsets of FluidSolver3D.cu:482-571)."""
    dx = 0.25
    s = oracle.solver(dx=dx, dt=0.01)
    s.set_labels(np.full((4, 4, 4), 2, np.uint8))
    s.set_particles(np.array([[1.3 * dx, 2.6 * dx, 1.9 * dx]], np.float32), np.array([[1.0, 2.0, 3.0]], np.float32))
    s.stage("label")
    s.stage("p2g")

    def den_pattern(base, frac, shape):
        out = np.zeros(shape)
        for dz in (0, 1):
            for dy in (0, 1):
                for dxx in (0, 1):
                    w = ((frac[0] if dxx else 1 - frac[0]) * (frac[1] if dy else 1 - frac[1]) *
                         (frac[2] if dz else 1 - frac[2]))
                    out[base[2] + dz, base[1] + dy, base[0] + dxx] = w
        return out

    us = s.get_grid(FS_GRID_USAVED)
    vs = s.get_grid(FS_GRID_VSAVED)
    ws = s.get_grid(FS_GRID_WSAVED)
    # a single particle reproduces its own velocity on every face it touches
    for g, val, base, frac in ((us, 1.0, (1, 2, 1), (0.3, 0.1, 0.4)), (vs, 2.0, (0, 2, 1), (0.8, 0.6, 0.4)),
                               (ws, 3.0, (0, 2, 1), (0.8, 0.1, 0.9))):
        pat = den_pattern(base, frac, g.shape)
        touched = pat > 1e-8
        assert touched.sum() == 8
        assert np.allclose(g[touched], val, rtol=1e-6) and not g[~touched].any()
        assert pat.sum() == pytest.approx(1.0)
    # gravity on every v face after the save (ApplyScalarForce, FluidSolver3D.cu:687-698)
    v = s.get_grid(FS_GRID_V)
    assert np.allclose(v - vs, np.float32(0.01) * np.float32(-9.81), rtol=1e-6)
    assert np.array_equal(s.get_grid(FS_GRID_U), us)


def test_p2g_two_particles_hat_weights(oracle):
    """Two particles of different velocities: every face value is the
    weighted mean sum(w v) / sum(w) with the trilinear hat weights
    w = prod over axes of max(0, 1 - |x_p - x_f| / dx), x_f the face centre
    (u at (i, j + 1/2, k + 1/2) dx, v and w alike), computed here from the
    positions alone (particlesToGrid FluidSolver3D.cu:607-685)."""
    dx = 0.25
    pos = np.array([[1.3, 2.6, 1.9], [1.75, 2.2, 1.45]], np.float32) * np.float32(dx)
    vel = np.array([[1.0, 2.0, 3.0], [-2.0, 0.5, -1.0]], np.float32)
    s = oracle.solver(dx=dx, dt=0.01)
    s.set_labels(np.full((4, 4, 4), 2, np.uint8))
    s.set_particles(pos, vel)
    s.stage("label")
    s.stage("p2g")
    p = pos.astype(float) / dx
    for comp, g in enumerate((FS_GRID_USAVED, FS_GRID_VSAVED, FS_GRID_WSAVED)):
        got = s.get_grid(g)
        D_, H_, W_ = got.shape
        k, j, i = np.meshgrid(np.arange(D_), np.arange(H_), np.arange(W_), indexing="ij")
        fc = [i + (0.0 if comp == 0 else 0.5), j + (0.0 if comp == 1 else 0.5), k + (0.0 if comp == 2 else 0.5)]
        num = np.zeros(got.shape)
        den = np.zeros(got.shape)
        for q in range(2):
            w = np.ones(got.shape)
            for ax in range(3):
                w *= np.maximum(0.0, 1.0 - np.abs(p[q, ax] - fc[ax]))
            num += w * vel[q, comp]
            den += w
        want = np.where(den > 1e-8, num / np.maximum(den, 1e-300), 0.0)
        assert np.abs(got - want).max() <= 1e-6 * np.abs(vel).max(), comp
        assert np.count_nonzero(den) > 8   # the windows overlap


def test_rhs_of_a_linear_field(oracle):
    """constructRHS (FluidSolver3D.cu:1360-1474): b = -(rho dx / dt) div u per
    FLUID cell, the divergence taken with the faces a SOLID cell shares at
    velocity 0 and those a BODY cell shares at the body velocity.  Face
    velocities linear in the face index (u = a i, v = b j, w = c k), one SOLID
    cell and one moving BODY cell inside the fluid."""
    dx, dt, rho = 0.125, 0.01, 1000.0
    a, b, c = 0.3, -0.7, 0.45
    V = np.array([0.8, -0.25, 0.6])
    s = one_face_body(oracle, list(_DIRS), (0, 0, 0), V, (2.5 * dx, 2.5 * dx, 2.5 * dx), dx=dx, dt=dt, N=8)
    lab = np.full((8, 8, 8), 2, np.uint8)
    lab[1:-1, 1:-1, 1:-1] = 1
    lab[2, 2, 2] = 3          # the body's cell (i, j, k) = (2, 2, 2)
    lab[4, 5, 4] = 0          # a SOLID cell at (4, 5, 4)
    s.set_grid(FS_GRID_LABELS, lab)
    for g, f in ((FS_GRID_U, lambda k, j, i: a * i), (FS_GRID_V, lambda k, j, i: b * j),
                 (FS_GRID_W, lambda k, j, i: c * k)):
        shape, _ = s.grid_shape(g)
        s.set_grid(g, np.fromfunction(f, shape).astype(np.float32))
    s.stage("rhs")
    rhs = s.get_grid(FS_GRID_RHS)
    scale = rho * dx / dt

    def face(comp, i, j, k, side):
        # velocity of the face of cell (i, j, k) on side -1 / +1 of axis comp
        n = [i, j, k]
        n[comp] += side
        L = lab[n[2], n[1], n[0]]
        if L == 0:
            return 0.0
        if L == 3:
            return V[comp]
        idx = [i, j, k]
        idx[comp] += (1 if side > 0 else 0)
        return (a, b, c)[comp] * idx[comp]

    checked = 0
    for k in range(8):
        for j in range(8):
            for i in range(8):
                if lab[k, j, i] != 1:
                    assert rhs[k, j, i] == 0.0
                    continue
                div = sum(face(q, i, j, k, 1) - face(q, i, j, k, -1) for q in range(3))
                assert rhs[k, j, i] == pytest.approx(-scale * div, rel=2e-5, abs=1e-4 * scale), (i, j, k)
                checked += 1
    assert checked > 200


def test_p2g_uniform_velocity_is_reproduced(oracle):
    rng = np.random.default_rng(3)
    N, dx = 10, 0.1
    s = oracle.solver(dx=dx)
    s.set_labels(scenes.bordered(N, N, N))
    pos = (rng.random((5000, 3)) * N * dx).astype(np.float32)
    s.set_particles(pos, np.tile(np.array([[0.7, -1.25, 2.5]], np.float32), (5000, 1)))
    s.stage("label")
    s.stage("p2g")
    for g, val in ((FS_GRID_USAVED, 0.7), (FS_GRID_VSAVED, -1.25), (FS_GRID_WSAVED, 2.5)):
        a = s.get_grid(g)
        assert np.allclose(a[a != 0], val, rtol=2e-6)


def _uniform_fields(s, N, U, dU):
    for g, gs, val, dv in ((FS_GRID_U, FS_GRID_USAVED, U[0], dU[0]), (FS_GRID_V, FS_GRID_VSAVED, U[1], dU[1]),
                           (FS_GRID_W, FS_GRID_WSAVED, U[2], dU[2])):
        shape, _ = s.grid_shape(g)
        s.set_grid(g, np.full(shape, val, np.float32))
        s.set_grid(gs, np.full(shape, val - dv, np.float32))


def test_g2p_uniform_fields_and_advection(oracle):
    """Uniform u = U with saved u = U - dU: v_p <- a U + (1 - a)(v_p + dU)
    (FluidSolver3D.cu:972-974); constant-field RK3 moves x by U dt exactly up
    to rounding (Ralston weights 2/9 + 3/9 + 4/9 = 1)."""
    N, dx, dt, alpha = 16, 1.0 / 16, 0.01, 0.05
    s = oracle.solver(dx=dx, dt=dt, pic_weight=alpha)
    s.set_labels(np.full((N, N, N), 2, np.uint8))
    rng = np.random.default_rng(4)
    pos = (0.3 + 0.4 * rng.random((200, 3))).astype(np.float32)
    vel = rng.standard_normal((200, 3)).astype(np.float32)
    s.set_particles(pos, vel)
    U, dU = (0.8, -0.5, 0.25), (0.1, 0.2, -0.3)
    _uniform_fields(s, N, U, dU)
    s.stage("g2p_advect")
    p2, v2 = s.particles()
    expect = alpha * np.array(U) + (1 - alpha) * (vel + np.array(dU))
    assert np.allclose(v2, expect, atol=1e-6)
    assert np.allclose(p2, pos + np.array(U) * dt, atol=2e-6)


def test_advection_rk3_in_a_rotation_field(oracle):
    """advectParticles (FluidSolver3D.cu:1241-1330) in a rigid rotation about z,
    which trilinear interpolation of the MAC faces reproduces exactly: the
    positions after one step equal Ralston's third-order Runge-Kutta (stages
    at 1/2 and 3/4, weights 2/9, 3/9, 4/9) with the reference's substeps
    (dT = C dx / (|v| + sqrt(5 dx 9.8)), the last one shortened to end at dt,
    the one before it halved when two would overshoot), computed here in
    FP64 from the analytic field; and they lie within the scheme's
    third-order error of the exact rotation."""
    N, w, dt, C = 32, 40.0, 0.01, 0.1
    dx = 1.0 / N
    c = np.array([0.5, 0.5])
    s = oracle.solver(dx=dx, dt=dt, cfl=0.0, pic_weight=1.0)
    s.set_labels(np.full((N, N, N), 1, np.uint8))
    ang = np.linspace(0.0, 2 * np.pi, 7, endpoint=False) + 0.3
    r0 = np.array([0.2, 0.15, 0.23, 0.18, 0.21, 0.16, 0.19])
    pos = np.stack([c[0] + r0 * np.cos(ang), c[1] + r0 * np.sin(ang), np.full(7, 0.5)], 1).astype(np.float32)
    s.set_particles(pos)
    for g, f in ((FS_GRID_U, lambda k, j, i: -w * ((j + 0.5) * dx - c[1])),
                 (FS_GRID_V, lambda k, j, i: w * ((i + 0.5) * dx - c[0]))):
        shape, _ = s.grid_shape(g)
        s.set_grid(g, np.fromfunction(f, shape).astype(np.float32))
    s.set_grid(FS_GRID_USAVED, s.get_grid(FS_GRID_U))
    s.set_grid(FS_GRID_VSAVED, s.get_grid(FS_GRID_V))
    s.stage("g2p_advect")
    got = s.particles()[0].astype(float)

    def vel(p):
        return np.array([-w * (p[1] - c[1]), w * (p[0] - c[0]), 0.0])

    floor = np.sqrt(5.0 * dx * 9.8)
    for q in range(7):
        p = pos[q].astype(float)
        sub, done = 0.0, False
        while not done:
            k1 = vel(p)
            dT = C * dx / (np.linalg.norm(k1) + floor)
            if sub + dT >= dt:
                dT, done = dt - sub, True
            elif sub + 2 * dT >= dt:
                dT *= 0.5
            k2 = vel(p + 0.5 * dT * k1)
            k3 = vel(p + 0.75 * dT * k2)
            p = p + dT * (2.0 / 9.0 * k1 + 3.0 / 9.0 * k2 + 4.0 / 9.0 * k3)
            sub += dT
        assert np.abs(got[q] - p).max() <= 3e-6, (q, got[q], p)
        a = ang[q] + w * dt
        exact = np.array([c[0] + r0[q] * np.cos(a), c[1] + r0[q] * np.sin(a), 0.5])
        assert np.abs(got[q] - exact).max() <= 1e-4


@pytest.mark.parametrize("axis", [None, 0, 1, 2])
def test_g2p_samples_faces_exactly(oracle, axis):
    """A particle on a u-face centre (i dx, (j + 1/2) dx, (k + 1/2) dx) has all
    three u fractions 0, so with PIC weight 1 its new v_x is that face's value
    bit for bit; moved by dx/2 along one axis, that fraction is 1/2 and v_x is
    the mean of the face and its neighbour along that axis, rounded once.
    Pins the base index and the orientation of each axis of the trilinear
    interpolation (a swapped axis or stride reads a different face)."""
    N, dx = 12, 1.0 / 16
    s = oracle.solver(dx=dx, dt=1e-6, pic_weight=1.0)
    s.set_labels(np.full((N, N, N), 2, np.uint8))
    rng = np.random.default_rng(9)
    f = {}
    for g, gs in ((FS_GRID_U, FS_GRID_USAVED), (FS_GRID_V, FS_GRID_VSAVED), (FS_GRID_W, FS_GRID_WSAVED)):
        shape, _ = s.grid_shape(g)
        f[g] = rng.standard_normal(shape).astype(np.float32)
        s.set_grid(g, f[g])
        s.set_grid(gs, f[g])
    ijk = rng.integers(1, N - 2, (300, 3))
    pos = np.stack([ijk[:, 0] * dx, (ijk[:, 1] + 0.5) * dx, (ijk[:, 2] + 0.5) * dx], 1)
    u = f[FS_GRID_U]
    want = u[ijk[:, 2], ijk[:, 1], ijk[:, 0]]
    if axis is not None:
        pos[:, axis] += 0.5 * dx
        nb = ijk.copy()
        nb[:, axis] += 1
        half = np.float32(0.5)
        want = want * half + u[nb[:, 2], nb[:, 1], nb[:, 0]] * half
    s.set_particles(pos.astype(np.float32), np.zeros(pos.shape, np.float32))
    s.stage("g2p_advect")
    _, v = s.particles()
    assert np.array_equal(v[:, 0], want.astype(np.float32))


def test_advection_clamps_to_domain(oracle):
    N, dx, dt = 8, 0.125, 0.1
    s = oracle.solver(dx=dx, dt=dt, pic_weight=1.0)
    s.set_labels(np.full((N, N, N), 2, np.uint8))
    s.set_particles(np.array([[0.8, 0.05, 0.5]], np.float32))
    _uniform_fields(s, N, (3.0, -3.0, 0.0), (0, 0, 0))
    s.stage("g2p_advect")
    p, _ = s.particles()
    assert p[0, 0] == np.float32((N - 1) * dx) and p[0, 1] == 0.0 and p[0, 2] == pytest.approx(0.5)


def test_advection_projects_out_of_solid(oracle):
    """A particle carried into a SOLID cell is moved by 0.5 dx toward the
    nearest FLUID cell among the 34 offsets (FluidSolver3D.cu:1161-1239)."""
    N, dx = 8, 0.125
    lab = np.full((N, N, N), 2, np.uint8)
    lab[:, :, 5:] = 0      # solid for i >= 5
    lab[4, 4, 4] = 1       # fluid cell next to the wall
    s = oracle.solver(dx=dx, dt=0.01, pic_weight=1.0)
    s.set_labels(lab)
    x0 = np.array([[4.98 * dx, 4.5 * dx, 4.5 * dx]], np.float32)
    s.set_particles(x0)
    _uniform_fields(s, N, (1.0, 0.0, 0.0), (0, 0, 0))
    s.stage("g2p_advect")
    p, _ = s.particles()
    # it crossed into cell i = 5 (SOLID) and was pushed back by -0.5 dx in x
    assert np.floor(p[0, 0] / dx) == 4
    assert p[0, 0] == pytest.approx(x0[0, 0] + 0.01 - 0.5 * dx, abs=1e-5)


# ---------------------------------------------------------------- rigid body
def test_body_sphere_labels_and_buoyancy(oracle):
    """Sphere in a flooded tank with one particle at every cell centre.  A
    voxel with phi <= 0 counts as submerged when a particle of its cell lies
    within 2 dx (FluidSolver3D.cu:2993-3004); every cell holds one here, so the
    submerged volume is the sphere's volume and the vertical force is
    (rho V - m) g (drag vanishes for a body at rest in still fluid)."""
    N, r = 40, 0.2
    dx = 1.0 / N
    lab = scenes.bordered(N, N, N)
    lab[1:-1, 1:-1, 1:-1] = 1
    sdf, origin, cs = scenes.sphere_sdf(r, dx / 2)
    mass = scenes.sphere_mass(r, 500.0)
    s = oracle.solver(dx=dx, dt=1e-3)
    s.set_labels(lab)
    s.set_body_sdf(sdf, origin, cs, (0.5, 0.5, 0.5), mass)
    L = s.get_grid(FS_GRID_LABELS)
    nb = int((L == 3).sum())
    vol = 4 / 3 * np.pi * r ** 3
    assert abs(nb * dx ** 3 / vol - 1) < 0.05
    k, j, i = np.nonzero(lab == 1)
    pos = ((np.stack([i, j, k], 1) + 0.5) * dx).astype(np.float32)
    s.set_particles(pos)
    s.step(1)
    st = s.body_state()
    # every voxel with phi <= 0 counts, with the partial weight 0.5 - 0.5 phi/cs
    # inside the surface band (FluidSolver3D.cu:2979-2983)
    inside = sdf[sdf <= 0].astype(np.float64)
    w = np.where(inside > -cs, 0.5 - 0.5 * inside / cs, 1.0)
    assert st.submerged_volume == pytest.approx(cs ** 3 * w.sum(), rel=1e-4)
    assert abs(st.submerged_volume / vol - 1) < 0.06
    expect = (1000.0 * st.submerged_volume - mass) * 9.81
    assert st.force[1] == pytest.approx(expect, rel=1e-3)
    assert abs(st.force[0]) < 1e-3 * expect and abs(st.force[2]) < 1e-3 * expect


def test_body_free_fall_in_air(oracle):
    N = 24
    dx, dt = 1.0 / N, 0.01
    lab = scenes.bordered(N, N, N)
    lab[1:-1, 1:3, 1:-1] = 1
    sdf, origin, cs = scenes.sphere_sdf(0.1, dx)
    s = oracle.solver(dx=dx, dt=dt)
    s.set_labels(lab)
    s.set_body_sdf(sdf, origin, cs, (0.5, 0.7, 0.5), 2.0)
    s.seed_particles(2, 1)
    s.step(3)
    st = s.body_state()
    # semi-implicit Euler: v_n = -n g dt, y_n = y0 - g dt^2 n(n+1)/2
    assert st.vel[1] == pytest.approx(-3 * 9.81 * dt, rel=1e-5)
    assert st.pos[1] == pytest.approx(0.7 - 9.81 * dt * dt * 6, rel=1e-5)
    assert st.submerged_volume == 0.0


# ---- drag / lift on one BODY-FLUID face (FluidSolver3D.cu:2674-2791) and the
# ---- rigid-body integrator (RigidBody::integrate utility.cuh:408-484)
def face_force_expected(n, u_rel, rho=1000.0, dx=0.125, cd=0.1, cl=0.1):
    """Drag -k|u|u and lift k_l |u| (u (u.n) - n |u|^2) / |n x u| with
    k = rho C A_eff / 2, A_eff = (n.u/|u| / 2 + 1/2) dx^2 when n.u > 0
    (the lift direction u x (n x u) / |n x u| written with the BAC-CAB
    identity).  FP64."""
    n, u = np.asarray(n, float), np.asarray(u_rel, float)
    ua = np.linalg.norm(u)
    nu = float(n @ u)
    if nu <= 0:
        return np.zeros(3)
    area = (0.5 * nu / ua + 0.5) * dx * dx
    f = -0.5 * rho * cd * area * ua * u
    nxu = np.linalg.norm(np.cross(n, u))
    if nxu > 0:
        f += 0.5 * rho * cl * area * ua * (u * nu - n * ua * ua) / nxu
    return f


@pytest.mark.parametrize("face,urel", [("+x", (2.0, 0.0, 0.0)),      # along n: drag only, lift guard
                                        ("+x", (0.0, 1.5, 0.0)),      # tangential: n.u = 0, no force
                                        ("+x", (-1.0, 0.5, 0.0)),     # away from the fluid: no force
                                        ("+x", (1.0, 1.0, 0.0)),      # 45 degrees in the xy plane
                                        ("-z", (0.3, -0.4, -1.2)),    # general direction
                                        ("+y", (-0.7, 0.9, 0.2))])
def test_drag_lift_one_face_closed_form(oracle, face, urel):
    U = np.array([0.25, -0.5, 0.75])
    V = U + np.array(urel)
    dx, mass = 0.125, 2.0
    cc = (np.array([2, 2, 2]) + 0.5) * dx
    pos = cc + np.array([0.03, -0.05, 0.02])
    s = one_face_body(oracle, [face], U, V, pos, dx=dx, mass=mass)
    s.stage("body")
    st = s.body_state()
    n = np.array(_DIRS[face], float)
    f = face_force_expected(n, urel, dx=dx)
    r = cc + 0.5 * dx * n - np.asarray(pos, np.float32).astype(float)
    # the force is m g + F_fluid summed in FP32: allow the rounding of that sum
    mg = np.float32(mass) * np.float32(-9.81)
    fluid_force = np.array(st.force, float) - np.array([0.0, mg, 0.0])
    tol = 2e-5 * np.linalg.norm(f) + 2e-7 * abs(mg)
    assert np.abs(fluid_force - f).max() <= tol, (fluid_force, f)
    assert np.abs(np.array(st.torque) - np.cross(r, f)).max() <= 2e-5 * np.linalg.norm(f)
    ua = np.array(urel)
    if np.linalg.norm(f) > 0:
        drag = -0.5 * 1000.0 * 0.1 * (0.5 * (n @ ua) / np.linalg.norm(ua) + 0.5) * dx * dx * np.linalg.norm(ua) * ua
        lift = f - drag
        # lift is normal to the relative velocity and lies in the plane of n and u_rel
        assert abs(lift @ ua) <= 1e-9 * np.linalg.norm(f) * np.linalg.norm(ua)
        assert abs(lift @ np.cross(n, ua)) <= 1e-9 * np.linalg.norm(f)
        # drag opposes the body's motion relative to the fluid
        assert drag @ ua < 0


def test_drag_lift_faces_add(oracle):
    """Two FLUID neighbours: the force and torque are the sums of the two
    one-face cases."""
    U, dx, mass = np.array([0.1, 0.2, -0.3]), 0.125, 2.0
    urel = np.array([0.8, 0.6, -0.4])
    cc = (np.array([2, 2, 2]) + 0.5) * dx
    pos = cc + np.array([-0.01, 0.04, 0.03])
    s = one_face_body(oracle, ["+x", "-z"], U, U + urel, pos, dx=dx, mass=mass)
    s.stage("body")
    st = s.body_state()
    posf = np.asarray(pos, np.float32).astype(float)
    fe, te = np.zeros(3), np.zeros(3)
    for face in ("+x", "-z"):
        n = np.array(_DIRS[face], float)
        f = face_force_expected(n, urel, dx=dx)
        fe += f
        te += np.cross(cc + 0.5 * dx * n - posf, f)
    mg = np.float32(mass) * np.float32(-9.81)
    got = np.array(st.force, float) - np.array([0.0, mg, 0.0])
    assert np.abs(got - fe).max() <= 2e-5 * np.linalg.norm(fe) + 2e-7 * abs(mg)
    assert np.abs(np.array(st.torque) - te).max() <= 2e-5 * np.linalg.norm(fe)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_integrator_constant_torque(oracle, axis):
    """A constant torque about one axis (the body's velocity and position reset
    before every step, so the drag that makes it stays the same): semi-implicit
    Euler gives omega_n = n (tau / I) dt and the angle about that axis
    (phi about x, theta about y, psi about z) = (tau / I) dt^2 n (n + 1) / 2,
    the other two angles stay 0, and the rotation matrix of the angles is the
    textbook rotation about that axis.  I = 0.4 m (w cs / 2)^2."""
    dx, dt, mass, steps = 0.125, 0.01, 2.0, 12
    a = 1.5
    # (FLUID face, relative velocity along it, offset of the body centre from the face centre)
    face, urel, off = [("+y", (0.0, a, 0.0), (0.0, 0.0, -0.04)),
                       ("+x", (a, 0.0, 0.0), (0.0, 0.0, -0.04)),
                       ("+x", (a, 0.0, 0.0), (0.0, -0.04, 0.0))][axis]
    n = np.array(_DIRS[face], float)
    cc = (np.array([2, 2, 2]) + 0.5) * dx
    pos = np.asarray(cc + 0.5 * dx * n + np.array(off), np.float32)
    U = np.zeros(3)
    s = one_face_body(oracle, [face], U, urel, pos, dx=dx, dt=dt, mass=mass)
    f = face_force_expected(n, urel, dx=dx)
    tau = np.cross(cc + 0.5 * dx * n - pos.astype(float), f)
    assert np.count_nonzero(np.abs(tau) > 1e-12) == 1 and abs(tau[axis]) > 0
    inertia = 0.4 * mass * (4 * 0.05 / 2) ** 2
    for _ in range(steps):
        st = s.body_state()
        st.pos[:] = pos
        st.vel[:] = urel
        s.set_body_state(st)
        s.stage("body")
    st = s.body_state()
    alpha = tau[axis] / inertia
    om = np.array(st.omega)
    assert om[axis] == pytest.approx(steps * alpha * dt, rel=1e-5)
    assert np.count_nonzero(om) == 1
    ang = {0: st.phi, 1: st.theta, 2: st.psi}
    want = alpha * dt * dt * steps * (steps + 1) / 2
    assert ang[axis] == pytest.approx(want, rel=1e-5)
    assert [ang[q] for q in range(3) if q != axis] == [0.0, 0.0]
    s.set_body_state(st)                 # rotation matrix of the current angles
    R = np.array(s.body_state().rotation, float).reshape(3, 3)
    c, sn = np.cos(want), np.sin(want)
    Rt = {0: [[1, 0, 0], [0, c, -sn], [0, sn, c]],
          1: [[c, 0, sn], [0, 1, 0], [-sn, 0, c]],
          2: [[c, -sn, 0], [sn, c, 0], [0, 0, 1]]}[axis]
    assert np.abs(R - np.array(Rt)).max() <= 1e-6


def test_rotation_matrix_is_a_rotation(oracle):
    """RigidBody::update_rotation_matrix (utility.cuh:470-484) of three nonzero
    angles is orthonormal with determinant 1 and equals Ry(theta) Rz(psi)
    Rx(phi)."""
    s = one_face_body(oracle, ["+x"], (0, 0, 0), (0, 0, 0), (0.3, 0.3, 0.3))
    st = s.body_state()
    st.theta, st.psi, st.phi = 0.4, -0.9, 1.3
    s.set_body_state(st)
    R = np.array(s.body_state().rotation, float).reshape(3, 3)
    assert np.abs(R @ R.T - np.eye(3)).max() <= 1e-6 and np.linalg.det(R) == pytest.approx(1.0, abs=1e-6)
    th, ps, ph = np.float32(0.4), np.float32(-0.9), np.float32(1.3)
    Ry = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    Rz = np.array([[np.cos(ps), -np.sin(ps), 0], [np.sin(ps), np.cos(ps), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, np.cos(ph), -np.sin(ph)], [0, np.sin(ph), np.cos(ph)]])
    assert np.abs(R - Ry @ Rz @ Rx).max() <= 1e-6


def test_integrator_torque_in_body_frame(oracle):
    """The torque enters the angular velocity in the body frame, omega +=
    (dt / I) R^T tau, with R = Ry(theta) Rz(psi) Rx(phi) mapping body to world:
    the transform the label marking applies to world points
    (Utility::contains utility.cu:163-206).  One step from a rotated state."""
    dx, dt, mass = 0.125, 0.01, 2.0
    cc = (np.array([2, 2, 2]) + 0.5) * dx
    pos = np.asarray(cc + np.array([0.02, -0.03, 0.05]), np.float32)
    urel = np.array([0.9, 0.5, -0.3])
    s = one_face_body(oracle, ["+x"], (0, 0, 0), urel, pos, dx=dx, dt=dt, mass=mass)
    st = s.body_state()
    th, ps, ph = 0.3, 0.5, -0.2
    st.theta, st.psi, st.phi = th, ps, ph
    s.set_body_state(st)
    R = np.array(s.body_state().rotation, float).reshape(3, 3)
    n = np.array(_DIRS["+x"], float)
    f = face_force_expected(n, urel, dx=dx)
    tau = np.cross(cc + 0.5 * dx * n - pos.astype(float), f)
    s.stage("body")
    om = np.array(s.body_state().omega, float)
    inertia = 0.4 * mass * 0.1 ** 2
    want = R.T @ tau * dt / inertia
    assert np.abs(om - want).max() <= 2e-5 * np.linalg.norm(want)
    assert np.abs(om - R @ tau * dt / inertia).max() > 1e-2 * np.linalg.norm(want)


def test_euler_angle_rates(oracle):
    """The Euler-angle update of RigidBody::integrate (utility.cuh:455-460),
    restated from the reference's source: with no torque (a body in air),
    one step adds dt * (wy cos(phi) - wz sin(phi)) to theta,
    dt * (wy sin(phi) + wz cos(phi)) / cos(theta) to psi and
    dt * (wx + dpsi sin(theta)) to phi, from the angles before the step.
    These rates are not the body-frame rates of the reference's own
    R = Ry(theta) Rz(psi) Rx(phi) (R^T dR/dt != [omega]x); the build keeps the
    reference's equations, and the second check records the difference."""
    dt = 0.01
    s = one_face_body(oracle, [], (0, 0, 0), (0, 0, 0), (0.3, 0.3, 0.3), dt=dt)
    st = s.body_state()
    st.theta, st.psi, st.phi = 0.4, -0.3, 0.7
    st.omega[:] = (0.5, -1.2, 0.8)
    s.set_body_state(st)
    s.stage("body")
    st2 = s.body_state()
    wx, wy, wz = (float(v) for v in np.float32([0.5, -1.2, 0.8]))
    th, ps, ph = (float(v) for v in np.float32([0.4, -0.3, 0.7]))
    assert tuple(st2.omega) == pytest.approx((wx, wy, wz), abs=0)
    dth = wy * np.cos(ph) - wz * np.sin(ph)
    dps = (wy * np.sin(ph) + wz * np.cos(ph)) / np.cos(th)
    dph = wx + dps * np.sin(th)
    got = np.array([st2.theta - th, st2.psi - ps, st2.phi - ph]) / dt
    assert np.abs(got - np.array([dth, dps, dph])).max() <= 2e-4
    # the body-frame rates of R = Ry Rz Rx would differ
    body_dth = (wy * np.cos(ph) - wz * np.sin(ph)) / np.cos(ps)
    body_dps = wy * np.sin(ph) + wz * np.cos(ph)
    assert abs(body_dps - dps) > 5e-3 and abs(body_dth - dth) > 5e-3


def test_isolated_fluid_cell_is_no_unknown(oracle):
    """A FLUID cell whose six neighbours are all SOLID has an empty matrix row
    (FluidSolver3D.cu:1656-1716 gives it diagonal 0): it is left out of the
    system, so p = 0 there, its b does not enter the residual, and the rest of
    the solution is that of the system without it."""
    N = 16
    lab = scenes.dam_break(N, (0.5, 0.5, 1.0))
    lab[6:11, 10:15, 6:11] = 0
    lab[8, 12, 8] = 1                      # enclosed by SOLID on all sides
    rng = np.random.default_rng(9)
    b = (rng.standard_normal(lab.shape) * 1e3).astype(np.float32) * (lab == 1)
    s = oracle.solver()
    p, st = s.solve_pressure_system(lab, b, 1e-12)
    assert st.converged == 1 and p[8, 12, 8] == 0.0
    assert st.fluid_cells == int((lab == 1).sum()) - 1
    lab2 = lab.copy()
    lab2[8, 12, 8] = 0
    p2, _ = s.solve_pressure_system(lab2, b * (lab2 == 1), 1e-12)
    assert np.abs(p - p2).max() <= 1e-9 * np.abs(p2).max()


def test_step_reports_unconverged_solve(oracle):
    """A step whose pressure solve misses the tolerance still completes (the
    reference ignores the solver status), and fs_step then returns
    FS_ERR_NOT_CONVERGED.  (More than 20000 fluid cells: the oracle solves
    iteratively.)"""
    from fluidprjct_amd._abi import FsError
    N = 32
    s = oracle.solver(dx=1.0 / N, dt=1.0 / 120, solver_tol=1e-14, solver_max_iter=1)
    s.set_labels(scenes.dam_break(N, (1.0, 0.8, 1.0)))
    s.seed_particles(1, 1)
    with pytest.raises(FsError) as e:
        s.step(2)
    assert e.value.code == -4 and "2 of 2 steps" in str(e.value)
    assert s.clock()[2] == 2


# ---- options beyond the reference's default path: inertia tensor with a
# ---- quaternion orientation (FS_STEP_BODY_INERTIA), wall collision (FS_STEP_BODY_WALLS)
def box_inertia(m, sides, angle_z=0.0):
    """Closed form of a solid box, turned by angle_z about z."""
    a, b, c = sides
    I = m / 12.0 * np.diag([b * b + c * c, a * a + c * c, a * a + b * b])
    ca, sa = np.cos(angle_z), np.sin(angle_z)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    return Rz @ I @ Rz.T


@pytest.mark.parametrize("angle", [0.0, 0.6])
def test_inertia_tensor_of_a_box(oracle, angle):
    """One step from rest under a torque tau: omega = dt I^-1 tau (the
    gyroscopic term vanishes at rest), with I the box's closed form (voxel
    discretisation at 1/32 of the shortest side: 3%).  The turned box has
    products of inertia."""
    from body_cases import box_sdf
    from fluidprjct_amd._abi import FS_STEP_BODY_INERTIA
    dx, dt, mass = 0.125, 0.01, 2.0
    sides = (0.2, 0.12, 0.08)
    cs = 0.0025
    sdf = box_sdf(sides, cs, angle)
    cc = (np.array([2, 2, 2]) + 0.5) * dx
    urel = np.array([0.0, 1.5, 0.0])
    n = np.array(_DIRS["+y"], float)
    pos = np.asarray(cc + 0.5 * dx * n + np.array([0.0, 0.0, -0.04]), np.float32)
    s = one_face_body(oracle, ["+y"], (0, 0, 0), urel, pos, dx=dx, dt=dt, mass=mass, sdf=(sdf[0], sdf[1], cs),
                      step_flags=FS_STEP_BODY_INERTIA)
    f = face_force_expected(n, urel, dx=dx)
    tau = np.cross(cc + 0.5 * dx * n - pos.astype(float), f)
    s.stage("body")
    om = np.array(s.body_state().omega, float)
    want = dt * np.linalg.solve(box_inertia(mass, sides, angle), tau)
    assert np.abs(om - want).max() <= 0.03 * np.abs(want).max(), (om, want)
    if angle:
        assert abs(want[1]) > 0.1 * abs(want[0])    # the turned box couples x and y


def test_quaternion_rotation_about_a_principal_axis(oracle):
    """Constant torque about a principal axis (x): omega_n = n (tau / Ixx) dt
    and the angle about x = sum of 2 atan(omega_k dt / 2) (the normalised
    quaternion step), which phi reports; R stays a rotation about x."""
    from body_cases import box_sdf
    from fluidprjct_amd._abi import FS_STEP_BODY_INERTIA
    dx, dt, mass, steps = 0.125, 0.01, 2.0, 10
    sides = (0.2, 0.12, 0.08)
    sdf = box_sdf(sides, 0.005)
    cc = (np.array([2, 2, 2]) + 0.5) * dx
    urel = np.array([0.0, 1.5, 0.0])
    n = np.array(_DIRS["+y"], float)
    pos = np.asarray(cc + 0.5 * dx * n + np.array([0.0, 0.0, -0.04]), np.float32)
    s = one_face_body(oracle, ["+y"], (0, 0, 0), urel, pos, dx=dx, dt=dt, mass=mass, sdf=(sdf[0], sdf[1], 0.005),
                      step_flags=FS_STEP_BODY_INERTIA)
    s.stage("body")
    w1 = s.body_state().omega[0]               # dt tau / Ixx (the tensor is pinned above)
    for _ in range(steps - 1):
        st = s.body_state()
        st.pos[:] = pos
        st.vel[:] = urel
        s.set_body_state(st)
        s.stage("body")
    st = s.body_state()
    assert st.omega[0] == pytest.approx(steps * w1, rel=1e-5)
    assert abs(st.omega[1]) <= 1e-6 * abs(st.omega[0]) and abs(st.omega[2]) <= 1e-6 * abs(st.omega[0])
    want = sum(2 * np.arctan(k * w1 * dt / 2) for k in range(1, steps + 1))
    assert st.phi == pytest.approx(want, rel=1e-4)
    assert abs(st.theta) <= 1e-6 and abs(st.psi) <= 1e-6
    R = np.array(st.rotation, float).reshape(3, 3)
    c, sn = np.cos(want), np.sin(want)
    assert np.abs(R - np.array([[1, 0, 0], [0, c, -sn], [0, sn, c]])).max() <= 1e-5


def test_torque_free_spin_conserves_angular_momentum(oracle):
    """A turned box spinning about a general axis in air (no fluid next to
    it, so no drag and no torque): the world-frame angular momentum
    R I_b R^T omega, with I_b the box's closed form, stays within 3.5% over
    40 small steps (3% for the voxel tensor against the closed form, 0.5% for
    the integration), while omega itself precesses."""
    from body_cases import box_sdf
    from fluidprjct_amd._abi import FS_STEP_BODY_INERTIA
    N, dt, mass = 16, 2e-3, 1.0
    sides = (0.3, 0.18, 0.1)
    sdf, origin = box_sdf(sides, 0.01, 0.3)
    s = oracle.solver(dx=1.0 / N, dt=dt, step_flags=FS_STEP_BODY_INERTIA)
    s.set_labels(scenes.bordered(N, N, N))
    s.set_body_sdf(sdf, origin, 0.01, (0.5, 0.5, 0.5), mass)
    st = s.body_state()
    st.omega[:] = (3.0, -4.0, 2.0)
    s.set_body_state(st)
    Ib = box_inertia(mass, sides, 0.3)

    def L(st):
        R = np.array(st.rotation, float).reshape(3, 3)
        return R @ Ib @ R.T @ np.array(st.omega, float)

    L0 = L(s.body_state())
    om0 = np.array(s.body_state().omega, float)
    for _ in range(40):
        s.stage("body")
    st = s.body_state()
    assert np.linalg.norm(L(st) - L0) <= 0.035 * np.linalg.norm(L0)
    assert np.linalg.norm(np.array(st.omega, float) - om0) > 1e-3 * np.linalg.norm(om0)


def test_wall_collision(oracle):
    """handleBodyWallCollisions (FluidSolver3D.cu:2848-2895): with r = half the
    largest SDF box side and the wall at the centre of the border cell, a
    body within r of its clamp position pMin = dx/2 + r is moved out by the
    difference, and its velocity into the wall is removed; along the other
    axes it is left alone."""
    from fluidprjct_amd._abi import FS_STEP_BODY_WALLS
    N, dt = 32, 1e-3
    dx = 1.0 / N
    sdf, origin, cs = scenes.sphere_sdf(0.1, dx)          # box side w * cs
    r = 0.5 * sdf.shape[2] * cs
    pmin = 0.5 * dx + r
    y0 = pmin + 0.3 * r                                     # 0.3 r past the clamp: pushed by 0.7 r
    s = oracle.solver(dx=dx, dt=dt, step_flags=FS_STEP_BODY_WALLS)
    s.set_labels(scenes.bordered(N, N, N))
    s.set_body_sdf(sdf, origin, cs, (0.5, y0, 0.5), 1.0)
    st = s.body_state()
    st.vel[:] = (0.2, -1.0, 0.0)
    s.set_body_state(st)
    s.stage("body")
    st = s.body_state()
    y1 = y0 + (-1.0 - 9.81 * dt) * dt                      # after the free-fall step
    assert st.pos[1] == pytest.approx(y1 + (r - (y1 - pmin)), rel=1e-5)
    assert st.vel[1] == 0.0 and st.vel[0] == pytest.approx(0.2)
    assert st.pos[0] == pytest.approx(0.5 + 0.2 * dt, rel=1e-5)
