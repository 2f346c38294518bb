"""Pins of the oracle's FP64 learned-preconditioner network (oracle/lnet_oracle.cpp).

The network has no reference counterpart (the reference solves with cuDSS,
src/FluidSim/FluidSolver3D.cu:1930-2008).  Single-tap weights reduce it to
closed forms computed here independently with numpy (shifts, 2x pooling,
nearest upsampling, masks, biases), which pin the tap orientation, the channel
order of every layer, the pooling/upsampling and skip structure and the output
scaling; central finite differences of the FP64 loss pin the gradient of every
parameter block.  Runs on the CPU."""
import numpy as np
import pytest

import lnet_ref as L

SHAPE = (7, 6, 9)   # (D, H, W), odd sizes exercise the zero padding of the pooling


def scene(seed=0):
    """labels and a float32 residual; expectations are formed from r in FP64"""
    return L.random_scene(SHAPE, seed)


def f64(a):
    return np.asarray(a, np.float64)


def test_identity_path_is_relu_of_normalised_residual():
    lab, r = scene(1)
    p = np.zeros(L.NPARAM, np.float32)
    L.conv_w(p, "A", L.tap(0, 0, 0), 0, 0, 1.0)
    L.conv_w(p, "E", L.tap(0, 0, 0), 0, 0, 1.0)
    z, zn = L.forward(p, lab, r)
    s = L.fluid_rms(lab, r)
    np.testing.assert_allclose(zn, np.maximum(f64(r) / s, 0), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(z, np.maximum(f64(r) / s, 0) * s * (lab == 1), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("da,de", [((1, 0, 0), (0, 0, -1)), ((0, -1, 1), (1, 1, 0)), ((-1, 1, -1), (0, 0, 0))])
def test_tap_orientation(da, de):
    """out(v) = sum_t w[t] in(v + d_t): A at offset da, E at offset de."""
    lab, r = scene(2)
    p = np.zeros(L.NPARAM, np.float32)
    L.conv_w(p, "A", L.tap(*da), 0, 3, 1.0)     # into channel 3
    L.conv_w(p, "E", L.tap(*de), 3, 0, 1.0)
    _, zn = L.forward(p, lab, r)
    s = L.fluid_rms(lab, r)
    a = np.maximum(L.shift(f64(r) / s, *da), 0)
    np.testing.assert_allclose(zn, L.shift(a, *de), rtol=1e-12, atol=1e-12)


def test_mask_channels_and_biases():
    lab, r = scene(3)
    p = np.zeros(L.NPARAM, np.float32)
    L.conv_w(p, "A", L.tap(0, 0, 0), 1, 0, 1.0)    # fluid mask -> ch 0
    L.conv_w(p, "A", L.tap(0, 0, 0), 2, 5, 1.0)    # air mask -> ch 5
    L.conv_w(p, "E", L.tap(0, 0, 0), 0, 0, 1.0)
    L.conv_w(p, "E", L.tap(0, 0, 0), 5, 0, 2.0)
    off, _ = L.BLOCKS["E.b"]
    p[off] = 0.5
    off, _ = L.BLOCKS["A.b"]
    p[off + 7] = 3.0                                # ch 7 constant 3, unused by E
    _, zn = L.forward(p, lab, r)
    np.testing.assert_allclose(zn, (lab == 1) + 2.0 * (lab == 2) + 0.5, rtol=1e-12, atol=1e-12)


def test_pool_up_skip_structure():
    """A = relu(r), B, D identity; C identity or off: closed forms of the U-net skips."""
    lab, r = scene(4)
    s = L.fluid_rms(lab, r)
    a = np.maximum(f64(r) / s, 0)
    for with_c in (False, True):
        p = np.zeros(L.NPARAM, np.float32)
        L.conv_w(p, "A", L.tap(0, 0, 0), 0, 2, 1.0)
        L.conv_w(p, "B", L.tap(0, 0, 0), 2, 4, 1.0)
        if with_c:
            L.conv_w(p, "C", L.tap(0, 0, 0), 4, 4, 1.0)
        L.conv_w(p, "D", L.tap(0, 0, 0), 4, 9, 1.0)
        L.conv_w(p, "E", L.tap(0, 0, 0), 9, 0, 1.0)      # E reads only the decoder's channel 9
        _, zn = L.forward(p, lab, r)
        b1 = L.pool2(a)
        s1 = b1 + (L.up2(L.pool2(b1), b1.shape) if with_c else 0)
        expect = L.up2(s1, SHAPE)                          # channel 9 of a0 is 0
        np.testing.assert_allclose(zn, expect, rtol=1e-12, atol=1e-12)
        # E reading channel 2 sees the skip a0 only (d1 is 0 in channel 2)
        L.conv_w(p, "E", L.tap(0, 0, 0), 9, 0, 0.0)
        L.conv_w(p, "E", L.tap(0, 0, 0), 2, 0, 1.0)
        _, zn = L.forward(p, lab, r)
        np.testing.assert_allclose(zn, a, rtol=1e-12, atol=1e-12)


def test_scale_equivariance_and_zero_input():
    lab, r = scene(5)
    p = L.random_params(7)
    z1, zn1 = L.forward(p, lab, r)
    z2, zn2 = L.forward(p, lab, (4.0 * r).astype(np.float32))
    np.testing.assert_allclose(zn2, zn1, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(z2, 4.0 * z1, rtol=1e-12, atol=1e-12)
    z0, _ = L.forward(p, lab, np.zeros_like(r))      # rms 0 -> scale 1, finite output
    assert np.isfinite(z0).all()
    assert (z0[lab != 1] == 0).all()


def test_apply_A_matches_assembled_matrix(oracle):
    """The numpy stencil used by these tests equals the oracle's assembled
    reference matrix (upper CSR of constructA, FluidSolver3D.cu:1718-1777)."""
    import ctypes as C
    lab, _ = scene(16)
    x = np.random.default_rng(17).normal(size=SHAPE) * (lab == 1)
    D, H, W = SHAPE
    f = lab.ravel() == 1
    nf = int(f.sum())
    asm = oracle.lib.oracle_assemble_upper_csr
    asm.restype = C.c_int64
    asm.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    lab_c = np.ascontiguousarray(lab)
    nnz = asm(lab_c.ctypes.data, W, H, D, None, None, None)
    off = np.zeros(nf + 1, np.int32)
    col = np.zeros(nnz, np.int32)
    val = np.zeros(nnz, np.float32)
    asm(lab_c.ctypes.data, W, H, D, off.ctypes.data, col.ctypes.data, val.ctypes.data)
    A = np.zeros((nf, nf))
    for i in range(nf):
        for k in range(off[i], off[i + 1]):
            A[i, col[k]] = val[k]
            A[col[k], i] = val[k]
    xf = x.ravel()[f]
    np.testing.assert_allclose(L.apply_A(lab, x).ravel()[f], A @ xf, rtol=1e-12, atol=1e-12)


def test_loss_definition():
    lab, b = scene(6)
    p = L.random_params(8)
    target = np.random.default_rng(9).normal(size=SHAPE) * (lab == 1)
    loss, _ = L.loss_grad(p, lab, b, target)
    _, zn = L.forward(p, lab, b)
    f = lab == 1
    nf = f.sum()
    t = np.where(f, target, 0.0)
    sp = np.sqrt(np.sum(t * L.apply_A(lab, t)) / nf)
    d = np.where(f, zn - t / sp, 0.0)
    assert loss == pytest.approx(0.5 * np.sum(d * L.apply_A(lab, d)) / nf, rel=1e-12)
    # invariant to the scale of the target
    loss2, _ = L.loss_grad(p, lab, b, 7.0 * target)
    assert loss2 == pytest.approx(loss, rel=1e-12)


def test_gradient_matches_central_differences():
    """Every parameter block, FP64 loss, float32 parameters perturbed by +-h."""
    lab, b = scene(10)
    p = L.random_params(11)
    target = np.random.default_rng(12).normal(size=SHAPE) * (lab == 1)
    _, g = L.loss_grad(p, lab, b, target)
    rng = np.random.default_rng(13)
    bad = []
    checked = 0
    for name, (off, n) in L.BLOCKS.items():
        idx = rng.choice(n, size=min(n, 8), replace=False) + off
        for i in idx:
            h = np.float32(1e-4 * max(1.0, abs(float(p[i]))))
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            lp, _ = L.loss_grad(pp, lab, b, target)
            lm, _ = L.loss_grad(pm, lab, b, target)
            fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
            checked += 1
            if abs(fd - g[i]) > 1e-6 * max(abs(g[i]), 1e-3 * np.abs(g).max()):
                bad.append((name, int(i - off), fd, g[i]))
    # a ReLU kink inside [-h, h] can spoil an isolated entry
    assert len(bad) <= 2, bad
    assert checked >= 70


def test_bf16_emulation_is_close_to_exact():
    lab, r = scene(14)
    p = L.random_params(15)
    _, zn = L.forward(p, lab, r)
    _, znb = L.forward(p, lab, r, bf16=True)
    rel = np.linalg.norm(znb - zn) / np.linalg.norm(zn)
    assert 1e-5 < rel < 3e-2
