"""ctypes access to the oracle's FP64 network restatement (oracle/lnet_oracle.cpp)
and small numpy helpers for the learned-preconditioner tests.  Test code only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "libfs_oracle.so")

NPARAM = 27 * 3 * 16 + 16 + 3 * (27 * 16 * 16 + 16) + 27 * 16 + 1
# parameter blocks: name -> (offset, size)
BLOCKS = {}
_off = 0
for _name, _n in (("A.w", 27 * 3 * 16), ("A.b", 16), ("B.w", 27 * 256), ("B.b", 16), ("C.w", 27 * 256),
                  ("C.b", 16), ("D.w", 27 * 256), ("D.b", 16), ("E.w", 27 * 16), ("E.b", 1)):
    BLOCKS[_name] = (_off, _n)
    _off += _n
assert _off == NPARAM

_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(ORACLE_LIB)
        vp = C.c_void_p
        L.oracle_lnet_forward.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]
        L.oracle_lnet_forward.restype = C.c_int
        L.oracle_lnet_loss_grad.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.POINTER(C.c_double), vp]
        L.oracle_lnet_loss_grad.restype = C.c_int
        L.oracle_lnet_num_params.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def forward(params, labels, r, bf16=False):
    """(z, zn) of the oracle network; labels/r shaped (D, H, W)."""
    prm = np.ascontiguousarray(params, np.float32)
    lab = np.ascontiguousarray(labels, np.uint8)
    rr = np.ascontiguousarray(r, np.float32)
    D, H, W = lab.shape
    z = np.zeros(lab.shape, np.float64)
    zn = np.zeros(lab.shape, np.float64)
    assert lib().oracle_lnet_forward(_p(prm), _p(lab), _p(rr), W, H, D, int(bf16), _p(z), _p(zn)) == 0
    return z, zn


def loss_grad(params, labels, b, p, bf16=False):
    prm = np.ascontiguousarray(params, np.float32)
    lab = np.ascontiguousarray(labels, np.uint8)
    bb = np.ascontiguousarray(b, np.float32)
    pp = np.ascontiguousarray(p, np.float64)
    D, H, W = lab.shape
    g = np.zeros(NPARAM, np.float64)
    loss = C.c_double()
    assert lib().oracle_lnet_loss_grad(_p(prm), _p(lab), _p(bb), _p(pp), W, H, D, int(bf16), C.byref(loss),
                                       _p(g)) == 0
    return loss.value, g


def tap(dx, dy, dz):
    """Tap index of the offset (dx, dy, dz): out(v) reads in(v + d)."""
    return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)


def conv_w(params, name, t, ci, co, value):
    """Set weight [t][ci][co] of layer name (A: 3 inputs, E: 1 output, else 16x16)."""
    off, _ = BLOCKS[name + ".w"]
    cin = 3 if name == "A" else 16
    cout = 1 if name == "E" else 16
    params[off + (t * cin + ci) * cout + co] = value


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for name, (off, n) in BLOCKS.items():
        if name.endswith(".w"):
            cin = 3 if name[0] == "A" else 16
            p[off:off + n] = rng.normal(0, scale * np.sqrt(2.0 / (27 * cin)), n)
        else:
            p[off:off + n] = rng.normal(0, 0.05 * scale, n)
    return p


def random_scene(shape, seed, p_fluid=0.6, p_air=0.25):
    rng = np.random.default_rng(seed)
    u = rng.random(shape)
    lab = np.where(u < p_fluid, 1, np.where(u < p_fluid + p_air, 2, 0)).astype(np.uint8)
    r = rng.normal(size=shape).astype(np.float32) * (lab == 1)
    return lab, r.astype(np.float32)


def shift(a, dx, dy, dz):
    """out[z, y, x] = a[z + dz, y + dy, x + dx] with zeros outside."""
    D, H, W = a.shape
    out = np.zeros_like(a)
    zs = slice(max(0, -dz), min(D, D - dz))
    ys = slice(max(0, -dy), min(H, H - dy))
    xs = slice(max(0, -dx), min(W, W - dx))
    zs2 = slice(zs.start + dz, zs.stop + dz)
    ys2 = slice(ys.start + dy, ys.stop + dy)
    xs2 = slice(xs.start + dx, xs.stop + dx)
    out[zs, ys, xs] = a[zs2, ys2, xs2]
    return out


def pool2(a):
    """Mean of 2x2x2 children, zero padding to even sizes."""
    D, H, W = a.shape
    pad = np.zeros(((D + 1) // 2 * 2, (H + 1) // 2 * 2, (W + 1) // 2 * 2), a.dtype)
    pad[:D, :H, :W] = a
    return pad.reshape(pad.shape[0] // 2, 2, pad.shape[1] // 2, 2, pad.shape[2] // 2, 2).mean(axis=(1, 3, 5))


def up2(a, shape):
    """Nearest-neighbour upsampling to shape."""
    D, H, W = shape
    return a.repeat(2, 0).repeat(2, 1).repeat(2, 2)[:D, :H, :W]


def fluid_rms(labels, r):
    f = labels == 1
    if not f.any():
        return 1.0
    s = float(np.sqrt(np.mean(r[f].astype(np.float64) ** 2)))
    return s if s > 0 else 1.0


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
