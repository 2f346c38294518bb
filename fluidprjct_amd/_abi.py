"""ctypes declarations of the fs_* C ABI (include/fluidsim.h).

`FsLib` binds any shared library that exports the ABI; `Solver` wraps one
handle with numpy in/out.  The package itself only ever binds the HIP library
(see `fluidprjct_amd.load_library`).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

FS_OK = 0
FS_SOLID, FS_FLUID, FS_AIR, FS_BODY = 0, 1, 2, 3
FS_FMT_PLY, FS_FMT_OFF, FS_FMT_PLY_BINARY = 0, 1, 2
FS_PRECOND_MG, FS_PRECOND_LEARNED, FS_PRECOND_JACOBI, FS_PRECOND_HYBRID = 0, 1, 2, 3
FS_SOLVER_NO_GRAPH = 1
FS_SOLVER_NO_TAIL = 2
FS_SOLVER_NO_FUSE = 4
FS_SOLVER_COLD_START = 8
FS_STEP_DETERMINISTIC, FS_STEP_BODY_WALLS, FS_STEP_BODY_INERTIA, FS_STEP_FACE_FRACTIONS = 1, 2, 4, 8
(FS_GRID_U, FS_GRID_V, FS_GRID_W, FS_GRID_USAVED, FS_GRID_VSAVED, FS_GRID_WSAVED, FS_GRID_P,
 FS_GRID_LABELS, FS_GRID_RHS, FS_GRID_P64, FS_GRID_CELLKEY, FS_GRID_SOLID,
 FS_GRID_OPEN_U, FS_GRID_OPEN_V, FS_GRID_OPEN_W) = range(15)
STAGES = ("label", "p2g", "pressure", "project", "body", "g2p_advect", "sort")


class FsConfig(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("dx", C.c_float), ("pic_weight", C.c_float),
        ("iter_per_frame", C.c_int), ("output_format", C.c_int),
        ("output_template", C.c_char * 256), ("input_label_data", C.c_char * 256),
        ("frame_amount", C.c_int), ("active_body", C.c_int),
        ("body_x", C.c_float), ("body_y", C.c_float), ("body_z", C.c_float),
        ("body_mass", C.c_float), ("body_sdf_path", C.c_char * 256),
        ("particles_per_cell", C.c_int), ("seed", C.c_uint64), ("gravity_y", C.c_float),
        ("density", C.c_float), ("advect_c", C.c_float), ("solver_tol", C.c_double),
        ("solver_max_iter", C.c_int), ("precond", C.c_int), ("stall_window", C.c_int),
        ("cfl", C.c_float), ("solver_flags", C.c_int), ("step_flags", C.c_int),
    ]


class FsBodyState(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3), ("vel", C.c_float * 3), ("omega", C.c_float * 3),
        ("theta", C.c_float), ("psi", C.c_float), ("phi", C.c_float),
        ("force", C.c_float * 3), ("torque", C.c_float * 3),
        ("submerged_volume", C.c_float), ("rotation", C.c_float * 9),
    ]


class FsSolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("outer", C.c_int), ("fallback", C.c_int), ("converged", C.c_int),
        ("rel_residual", C.c_double), ("solve_ms", C.c_double), ("fluid_cells", C.c_int),
        ("singular", C.c_int),
    ]


class FsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fs error {code}: {msg}")
        self.code = code


_vp = C.c_void_p
_SIGS = {
    "fs_config_default": (None, [C.POINTER(FsConfig)]),
    "fs_config_load_json": (C.c_int, [C.POINTER(FsConfig), C.c_char_p]),
    "fs_last_error": (C.c_char_p, []),
    "fs_backend": (C.c_char_p, []),
    "fs_create": (C.c_int, [C.POINTER(FsConfig), C.POINTER(_vp)]),
    "fs_destroy": (None, [_vp]),
    "fs_load_labels_file": (C.c_int, [_vp, C.c_char_p]),
    "fs_set_labels": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int]),
    "fs_set_level_sets": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int]),
    "fs_get_dims": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "fs_load_body_sdf": (C.c_int, [_vp, C.c_char_p, C.c_float, C.c_float, C.c_float, C.c_float]),
    "fs_set_body_sdf": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                  C.c_float, C.c_float, C.c_float, C.c_float, C.c_float]),
    "fs_get_body_state": (C.c_int, [_vp, C.POINTER(FsBodyState)]),
    "fs_set_body_state": (C.c_int, [_vp, C.POINTER(FsBodyState)]),
    "fs_seed_particles": (C.c_int, [_vp, C.c_int, C.c_uint64]),
    "fs_set_particles": (C.c_int, [_vp, _vp, _vp, C.c_int64]),
    "fs_get_particles": (C.c_int, [_vp, _vp, _vp, C.c_int64, C.POINTER(C.c_int64)]),
    "fs_num_particles": (C.c_int64, [_vp]),
    "fs_step": (C.c_int, [_vp, C.c_int]),
    "fs_run": (C.c_int, [_vp, C.c_int]),
    "fs_stage_label": (C.c_int, [_vp]),
    "fs_stage_p2g": (C.c_int, [_vp]),
    "fs_stage_rhs": (C.c_int, [_vp]),
    "fs_stage_pressure": (C.c_int, [_vp]),
    "fs_stage_project": (C.c_int, [_vp]),
    "fs_stage_body": (C.c_int, [_vp]),
    "fs_stage_g2p_advect": (C.c_int, [_vp]),
    "fs_get_grid": (C.c_int, [_vp, C.c_int, _vp, C.c_int64]),
    "fs_set_grid": (C.c_int, [_vp, C.c_int, _vp, C.c_int64]),
    "fs_get_stage_times": (C.c_int, [_vp, C.POINTER(C.c_double), C.c_int]),
    "fs_get_solve_stats": (C.c_int, [_vp, C.POINTER(FsSolveStats)]),
    "fs_get_clock": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int64)]),
    "fs_write_particles": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "fs_solve_pressure_system": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double,
                                           C.POINTER(FsSolveStats)]),
}
ABI_SYMBOLS = tuple(_SIGS)

# include/fluidsim_learned.h (HIP library only)
_LEARNED_SIGS = {
    "fs_learned_init": (C.c_int, [_vp, C.c_uint64]),
    "fs_learned_get_params": (C.c_int, [_vp, _vp, C.c_int]),
    "fs_learned_set_params": (C.c_int, [_vp, _vp, C.c_int]),
    "fs_learned_apply": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int]),
    "fs_learned_loss_grad": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), _vp]),
    "fs_learned_train_step": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float,
                                        C.POINTER(C.c_double)]),
    "fs_set_preconditioner": (C.c_int, [_vp, C.c_int, C.c_int]),
    "fs_mg_vcycle": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int]),
}
LEARNED_SYMBOLS = tuple(_LEARNED_SIGS)

# include/fluidsim_checkpoint.h (HIP library only)
_CKPT_SIGS = {
    "fs_save_checkpoint": (C.c_int, [_vp, C.c_char_p]),
    "fs_load_checkpoint": (C.c_int, [_vp, C.c_char_p]),
}
CHECKPOINT_SYMBOLS = tuple(_CKPT_SIGS)

# include/fluidsim_noalloc.h (HIP library only)
_NOALLOC_SIGS = {
    "fs_solve_pressure_system_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double,
                                                   C.POINTER(FsSolveStats)]),
}
NOALLOC_SYMBOLS = tuple(_NOALLOC_SIGS)
FS_NOALLOC_MAX_CELLS = 1 << 20
FSL_NUM_PARAMS = 27 * 3 * 16 + 16 + 3 * (27 * 16 * 16 + 16) + 27 * 16 + 1


class FsLib:
    """A loaded library exporting the fs_* ABI."""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = os.path.abspath(path)
        self.lib = C.CDLL(self.path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(self.lib, name)
            fn.restype = res
            fn.argtypes = args
        self.has_learned = all(hasattr(self.lib, n) for n in _LEARNED_SIGS)
        self.has_checkpoint = all(hasattr(self.lib, n) for n in _CKPT_SIGS)
        self.has_noalloc = all(hasattr(self.lib, n) for n in _NOALLOC_SIGS)
        for present, sigs in ((self.has_learned, _LEARNED_SIGS), (self.has_checkpoint, _CKPT_SIGS),
                              (self.has_noalloc, _NOALLOC_SIGS)):
            if present:
                for name, (res, args) in sigs.items():
                    fn = getattr(self.lib, name)
                    fn.restype = res
                    fn.argtypes = args
        self.backend = self.lib.fs_backend().decode()

    def check(self, rc):
        if rc != FS_OK:
            raise FsError(rc, self.lib.fs_last_error().decode(errors="replace"))
        return rc

    def default_config(self, **kw) -> FsConfig:
        cfg = FsConfig()
        self.lib.fs_config_default(C.byref(cfg))
        for k, v in kw.items():
            if isinstance(v, str):
                v = v.encode()
            setattr(cfg, k, v)
        return cfg

    def solver(self, cfg: FsConfig | None = None, **kw) -> "Solver":
        return Solver(self, cfg if cfg is not None else self.default_config(**kw))

    def solve_pressure_system_noalloc(self, labels, b, tol):
        """fs_solve_pressure_system_noalloc (include/fluidsim_noalloc.h): no
        handle, no device allocation; (p, stats)."""
        if not self.has_noalloc:
            raise RuntimeError(f"{self.path} has no allocation-free solve")
        lab = np.ascontiguousarray(labels, dtype=np.uint8)
        bb = np.ascontiguousarray(b, dtype=np.float32)
        D, H, W = lab.shape
        p = np.zeros((D, H, W), np.float64)
        st = FsSolveStats()
        self.check(self.lib.fs_solve_pressure_system_noalloc(_ptr(lab), _ptr(bb), _ptr(p), W, H, D, float(tol),
                                                             C.byref(st)))
        return p, st


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Solver:
    """One fs_* handle with numpy-facing helpers."""

    def __init__(self, fl: FsLib, cfg: FsConfig):
        self.h = None
        self.fl, self.lib, self.cfg = fl, fl.lib, cfg
        h = C.c_void_p()
        fl.check(self.lib.fs_create(C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.fs_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # scene --------------------------------------------------------------
    def load_labels_file(self, path):
        self.fl.check(self.lib.fs_load_labels_file(self.h, str(path).encode()))

    def set_labels(self, labels: np.ndarray):
        """labels: uint8 array of shape (D, H, W) (i fastest)."""
        lab = np.ascontiguousarray(labels, dtype=np.uint8)
        D, H, W = lab.shape
        self.fl.check(self.lib.fs_set_labels(self.h, _ptr(lab), W, H, D))

    def set_level_sets(self, phi_solid, phi_fluid):
        ps = np.ascontiguousarray(phi_solid, dtype=np.float32)
        pf = np.ascontiguousarray(phi_fluid, dtype=np.float32)
        D, H, W = ps.shape
        self.fl.check(self.lib.fs_set_level_sets(self.h, _ptr(ps), _ptr(pf), W, H, D))

    @property
    def dims(self):
        W, H, D = C.c_int(), C.c_int(), C.c_int()
        self.fl.check(self.lib.fs_get_dims(self.h, C.byref(W), C.byref(H), C.byref(D)))
        return W.value, H.value, D.value

    def set_body_sdf(self, sdf: np.ndarray, origin, cell_size, pos, mass):
        s = np.ascontiguousarray(sdf, dtype=np.float32)
        d, h, w = s.shape
        self.fl.check(self.lib.fs_set_body_sdf(self.h, _ptr(s), w, h, d, *[float(o) for o in origin],
                                               float(cell_size), *[float(p) for p in pos], float(mass)))

    def load_body_sdf(self, path, pos, mass):
        self.fl.check(self.lib.fs_load_body_sdf(self.h, str(path).encode(), *[float(p) for p in pos],
                                                float(mass)))

    def body_state(self) -> FsBodyState:
        st = FsBodyState()
        self.fl.check(self.lib.fs_get_body_state(self.h, C.byref(st)))
        return st

    def set_body_state(self, st: FsBodyState):
        self.fl.check(self.lib.fs_set_body_state(self.h, C.byref(st)))

    # particles ------------------------------------------------------------
    def seed_particles(self, ppc, seed):
        self.fl.check(self.lib.fs_seed_particles(self.h, int(ppc), int(seed)))

    def set_particles(self, pos, vel=None):
        p = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        v = np.zeros_like(p) if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
        self.fl.check(self.lib.fs_set_particles(self.h, _ptr(p), _ptr(v), p.shape[0]))

    @property
    def num_particles(self):
        return int(self.lib.fs_num_particles(self.h))

    def particles(self):
        n = self.num_particles
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        got = C.c_int64()
        self.fl.check(self.lib.fs_get_particles(self.h, _ptr(pos), _ptr(vel), n, C.byref(got)))
        return pos, vel

    # stepping -------------------------------------------------------------
    def step(self, n=1):
        self.fl.check(self.lib.fs_step(self.h, int(n)))

    def run(self, frames):
        self.fl.check(self.lib.fs_run(self.h, int(frames)))

    def stage(self, name):
        self.fl.check(getattr(self.lib, "fs_stage_" + name)(self.h))

    # grids ----------------------------------------------------------------
    def grid_shape(self, which):
        W, H, D = self.dims
        if which in (FS_GRID_U, FS_GRID_USAVED, FS_GRID_OPEN_U):
            return (D, H, W + 1), np.float32
        if which in (FS_GRID_V, FS_GRID_VSAVED, FS_GRID_OPEN_V):
            return (D, H + 1, W), np.float32
        if which in (FS_GRID_W, FS_GRID_WSAVED, FS_GRID_OPEN_W):
            return (D + 1, H, W), np.float32
        if which in (FS_GRID_P, FS_GRID_RHS):
            return (D, H, W), np.float32
        if which in (FS_GRID_LABELS, FS_GRID_SOLID):
            return (D, H, W), np.uint8
        if which == FS_GRID_P64:
            return (D, H, W), np.float64
        if which == FS_GRID_CELLKEY:
            return (self.num_particles,), np.int32
        raise ValueError(which)

    def get_grid(self, which):
        shape, dt = self.grid_shape(which)
        out = np.empty(shape, dt)
        self.fl.check(self.lib.fs_get_grid(self.h, which, _ptr(out), out.size))
        return out

    def set_grid(self, which, arr):
        shape, dt = self.grid_shape(which)
        a = np.ascontiguousarray(arr, dtype=dt).reshape(shape)
        self.fl.check(self.lib.fs_set_grid(self.h, which, _ptr(a), a.size))

    def stage_times(self):
        buf = (C.c_double * len(STAGES))()
        self.fl.check(self.lib.fs_get_stage_times(self.h, buf, len(STAGES)))
        return dict(zip(STAGES, list(buf)))

    def clock(self):
        t, dt, n = C.c_double(), C.c_float(), C.c_int64()
        self.fl.check(self.lib.fs_get_clock(self.h, C.byref(t), C.byref(dt), C.byref(n)))
        return t.value, dt.value, n.value

    def solve_stats(self) -> FsSolveStats:
        st = FsSolveStats()
        self.fl.check(self.lib.fs_get_solve_stats(self.h, C.byref(st)))
        return st

    def write_particles(self, path, fmt=FS_FMT_PLY):
        self.fl.check(self.lib.fs_write_particles(self.h, str(path).encode(), fmt))

    def solve_pressure_system(self, labels, b, tol):
        lab = np.ascontiguousarray(labels, dtype=np.uint8)
        bb = np.ascontiguousarray(b, dtype=np.float32)
        D, H, W = lab.shape
        p = np.zeros((D, H, W), np.float64)
        st = FsSolveStats()
        self.fl.check(self.lib.fs_solve_pressure_system(self.h, _ptr(lab), _ptr(bb), _ptr(p), W, H, D,
                                                        float(tol), C.byref(st)))
        return p, st

    # ---- learned preconditioner (include/fluidsim_learned.h)
    def _need_learned(self):
        if not self.fl.has_learned:
            raise RuntimeError(f"{self.fl.path} has no learned preconditioner")

    def set_preconditioner(self, precond, stall_window=0):
        self._need_learned()
        self.fl.check(self.lib.fs_set_preconditioner(self.h, int(precond), int(stall_window)))

    def mg_vcycle(self, labels, r):
        """One multigrid V-cycle z = V(r) from zero (fs_mg_vcycle)."""
        self._need_learned()
        lab = np.ascontiguousarray(labels, dtype=np.uint8)
        rr = np.ascontiguousarray(r, dtype=np.float32)
        D, H, W = lab.shape
        z = np.zeros((D, H, W), np.float32)
        self.fl.check(self.lib.fs_mg_vcycle(self.h, _ptr(lab), _ptr(rr), _ptr(z), W, H, D))
        return z

    def learned_init(self, seed):
        self._need_learned()
        self.fl.check(self.lib.fs_learned_init(self.h, int(seed)))

    def learned_params(self):
        self._need_learned()
        out = np.zeros(FSL_NUM_PARAMS, np.float32)
        self.fl.check(self.lib.fs_learned_get_params(self.h, _ptr(out), out.size))
        return out

    def set_learned_params(self, params):
        self._need_learned()
        a = np.ascontiguousarray(params, dtype=np.float32)
        self.fl.check(self.lib.fs_learned_set_params(self.h, _ptr(a), a.size))

    def learned_apply(self, labels, r):
        self._need_learned()
        lab = np.ascontiguousarray(labels, dtype=np.uint8)
        rr = np.ascontiguousarray(r, dtype=np.float32)
        D, H, W = lab.shape
        z = np.zeros((D, H, W), np.float32)
        self.fl.check(self.lib.fs_learned_apply(self.h, _ptr(lab), _ptr(rr), _ptr(z), W, H, D))
        return z

    def learned_loss_grad(self, labels, b, p):
        self._need_learned()
        lab = np.ascontiguousarray(labels, dtype=np.uint8)
        bb = np.ascontiguousarray(b, dtype=np.float32)
        pp = np.ascontiguousarray(p, dtype=np.float64)
        D, H, W = lab.shape
        g = np.zeros(FSL_NUM_PARAMS, np.float32)
        loss = C.c_double()
        self.fl.check(self.lib.fs_learned_loss_grad(self.h, _ptr(lab), _ptr(bb), _ptr(pp), W, H, D, C.byref(loss),
                                                    _ptr(g)))
        return loss.value, g

    def learned_train_step(self, labels, b, p, lr):
        self._need_learned()
        lab = np.ascontiguousarray(labels, dtype=np.uint8)
        bb = np.ascontiguousarray(b, dtype=np.float32)
        pp = np.ascontiguousarray(p, dtype=np.float64)
        D, H, W = lab.shape
        loss = C.c_double()
        self.fl.check(self.lib.fs_learned_train_step(self.h, _ptr(lab), _ptr(bb), _ptr(pp), W, H, D, float(lr),
                                                     C.byref(loss)))
        return loss.value

    # ---- checkpoint / resume (include/fluidsim_checkpoint.h)
    def save_checkpoint(self, path):
        if not self.fl.has_checkpoint:
            raise RuntimeError(f"{self.fl.path} has no checkpoint support")
        self.fl.check(self.lib.fs_save_checkpoint(self.h, str(path).encode()))

    def load_checkpoint(self, path):
        if not self.fl.has_checkpoint:
            raise RuntimeError(f"{self.fl.path} has no checkpoint support")
        self.fl.check(self.lib.fs_load_checkpoint(self.h, str(path).encode()))
