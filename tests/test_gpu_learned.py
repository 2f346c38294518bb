"""Learned pressure preconditioner on the GPU (include/fluidsim_learned.h).

Bars:
  - forward: relative L2 <= 5e-3 against the FP64 oracle network with BF16
    rounding emulated at the same storage points, <= 3e-2 against the exact
    FP64 network; z is exactly 0 off the fluid cells;
  - loss within 1e-2 relative and gradient within 3e-2 relative L2 per
    parameter block of the BF16-emulating oracle, cosine >= 0.99 against the
    exact FP64 gradient (itself pinned by finite differences on the CPU);
  - flexible PCG with the network reaches the 1e-6 relative residual (checked
    here in FP64 with an independent numpy stencil); a network that breaks
    down hands over to multigrid (fallback flag) and still converges;
  - training lowers the loss, and the trained network needs fewer iterations
    than PCG without it (Jacobi).
"""
import numpy as np
import pytest

import lnet_ref as L
from fluidprjct_amd._abi import FS_PRECOND_JACOBI, FS_PRECOND_LEARNED, FS_PRECOND_MG

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def apply_A(lab, p):
    """A p of the reference pressure matrix (FluidSolver3D.cu:1598-1716) in FP64:
    on FLUID cells diag = #in-domain FLUID/AIR neighbours, -1 per FLUID neighbour."""
    f = lab == 1
    fa = (lab == 1) | (lab == 2)
    out = np.zeros(lab.shape, np.float64)
    for ax in range(3):
        for sgn in (1, -1):
            nb_fa = np.zeros(lab.shape, bool)
            nb_p = np.zeros(lab.shape, np.float64)
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            if sgn == 1:
                dst[ax], src[ax] = slice(0, -1), slice(1, None)
            else:
                dst[ax], src[ax] = slice(1, None), slice(0, -1)
            nb_fa[tuple(dst)] = fa[tuple(src)]
            nb_p[tuple(dst)] = np.where(f[tuple(src)], p[tuple(src)], 0.0)
            out += np.where(nb_fa, p, 0.0) - nb_p
    return np.where(f, out, 0.0)


def dam(N, ext=(0.4, 0.5, 0.6)):
    from fluidprjct_amd import scenes
    return scenes.dam_break(N, ext)


def rhs_of(lab, seed):
    rng = np.random.default_rng(seed)
    b = rng.normal(size=lab.shape) * (lab == 1)
    return b.astype(np.float32)


def test_forward_matches_oracle(hip):
    lab, r = L.random_scene((23, 17, 20), 21)
    prm = L.random_params(22)
    s = hip.solver(dx=0.05)
    s.set_learned_params(prm)
    z = s.learned_apply(lab, r)
    zb, _ = L.forward(prm, lab, r, bf16=True)
    ze, _ = L.forward(prm, lab, r, bf16=False)
    assert (z[lab != 1] == 0).all()
    assert rel(z, zb) <= 5e-3, rel(z, zb)
    assert rel(z, ze) <= 3e-2, rel(z, ze)
    assert np.array_equal(s.learned_params(), prm)


def test_gradient_matches_oracle(hip):
    lab, b = L.random_scene((18, 21, 19), 31)
    p = np.random.default_rng(32).normal(size=lab.shape) * (lab == 1)
    prm = L.random_params(33)
    s = hip.solver(dx=0.05)
    s.set_learned_params(prm)
    loss, g = s.learned_loss_grad(lab, b, p)
    lb, gb = L.loss_grad(prm, lab, b, p, bf16=True)
    le, ge = L.loss_grad(prm, lab, b, p, bf16=False)
    assert loss == pytest.approx(lb, rel=1e-2)
    assert loss == pytest.approx(le, rel=3e-2)
    for name, (off, n) in L.BLOCKS.items():
        e = rel(g[off:off + n], gb[off:off + n])
        assert e <= 3e-2, (name, e)
    cos = float(np.dot(g, ge) / (np.linalg.norm(g) * np.linalg.norm(ge)))
    assert cos >= 0.99, cos


def test_learned_fpcg_converges(hip):
    N = 40
    lab = dam(N)
    b = rhs_of(lab, 41)
    s = hip.solver(dx=1.0 / N, seed=3)
    s.set_learned_params(L.random_params(42, scale=0.5))
    s.set_preconditioner(FS_PRECOND_LEARNED, 5)
    p, st = s.solve_pressure_system(lab, b, 1e-6)
    assert st.converged == 1
    res = np.linalg.norm(b.astype(np.float64) - apply_A(lab, p)) / np.linalg.norm(b.astype(np.float64))
    assert res <= 1.2e-6, res
    assert st.rel_residual <= 1e-6
    s.set_preconditioner(FS_PRECOND_MG)
    pm, stm = s.solve_pressure_system(lab, b, 1e-6)
    assert rel(p, pm) <= 1e-4


def test_breakdown_falls_back_to_multigrid(hip):
    N = 32
    lab = dam(N)
    b = rhs_of(lab, 51)
    s = hip.solver(dx=1.0 / N)
    s.set_learned_params(np.zeros(L.NPARAM, np.float32))   # z = 0: p.Ap = 0, a breakdown
    s.set_preconditioner(FS_PRECOND_LEARNED, 5)
    p, st = s.solve_pressure_system(lab, b, 1e-6)
    assert st.fallback == 1 and st.converged == 1
    res = np.linalg.norm(b.astype(np.float64) - apply_A(lab, p)) / np.linalg.norm(b.astype(np.float64))
    assert res <= 1.2e-6


def test_graph_and_eager_learned_solves_agree(hip):
    N = 36
    lab = dam(N, (0.5, 0.4, 0.7))
    b = rhs_of(lab, 61)
    prm = L.random_params(62, scale=0.5)
    out = []
    for flags in (0, 1):
        s = hip.solver(dx=1.0 / N, solver_flags=flags)
        s.set_learned_params(prm)
        s.set_preconditioner(FS_PRECOND_LEARNED, 5)
        out.append(s.solve_pressure_system(lab, b, 1e-6))
        s.close()
    (p0, s0), (p1, s1) = out
    assert s0.iterations == s1.iterations and s0.fallback == s1.fallback
    assert np.array_equal(p0, p1)


def training_pairs(hip, N, n, seed):
    from fluidprjct_amd.learned import generate_pairs
    return generate_pairs(hip, n, N=N, seed=seed, steps_per_scene=6, log=None)


def test_training_lowers_loss_and_beats_jacobi(hip):
    from fluidprjct_amd.learned import train
    N = 32
    pairs = training_pairs(hip, N, 36, 7)
    s = hip.solver(dx=1.0 / N, seed=5)
    s.learned_init(5)
    held = pairs[-4:]
    before = np.mean([s.learned_loss_grad(l, b, p)[0] for l, b, p in held])
    hist = train(s, pairs[:-4], iters=400, lr=2e-3, log=None)
    after = np.mean([s.learned_loss_grad(l, b, p)[0] for l, b, p in held])
    assert np.isfinite(hist).all()
    assert after < 0.5 * before, (before, after)
    # the trained network as the preconditioner of a held-out solve
    lab, b, _ = held[0]
    s.set_preconditioner(FS_PRECOND_LEARNED, 8)
    p, st = s.solve_pressure_system(lab, b, 1e-6)
    s.set_preconditioner(FS_PRECOND_JACOBI)
    pj, stj = s.solve_pressure_system(lab, b, 1e-6)
    assert st.converged == 1 and stj.converged == 1
    assert st.fallback == 0
    assert st.iterations < stj.iterations, (st.iterations, stj.iterations)
