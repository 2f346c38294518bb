"""fluidprjct_amd -- MI355X-native PIC/FLIP step behind the FluidSim API.

The compute path is the HIP library ``libfluidsim_hip.so`` built from
``csrc/`` for gfx950 (``make -C fluidprjct_amd``); its C ABI is
``include/fluidsim.h``.  This module binds it with ctypes and offers
``FluidSolver3D``, a mirror of the reference class
(src/FluidSim/FluidSolver3D.cuh:13-186): public fields PIC_WEIGHT,
iterPerFrame, outputFormat, outputTemplate, activeBody, initialBodyPos,
bodyMass, and the methods init(labels_path) and run(frames).

There is no CPU fallback: if the HIP library is missing, loading raises.
"""
from __future__ import annotations

import os

from ._abi import (FS_AIR, FS_BODY, FS_FLUID, FS_FMT_OFF, FS_FMT_PLY, FS_SOLID, FsError, FsLib,  # noqa: F401
                   Solver)
from . import _abi as abi  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfluidsim_hip.so")

_LIB = None


def load_library() -> FsLib:
    """Load the HIP library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` "
                               "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        _LIB = FsLib(LIB_PATH)
        if _LIB.backend != "hip":
            raise RuntimeError(f"{LIB_PATH} is not the HIP backend ({_LIB.backend})")
    return _LIB


class FluidSolver3D:
    """Python mirror of the reference's FluidSolver3D (FluidSolver3D.cuh:159-186).

    FluidSolver3D(W, H, D, dx, dt) -- the grid dimensions are replaced by the
    label file in init(), as in the reference (FluidSolver3D.cu:88).
    """

    def __init__(self, width: int, height: int, depth: int, dx: float, dt: float):
        self._lib = load_library()
        self.dx, self.dt = float(dx), float(dt)
        self.PIC_WEIGHT = 0.5          # FluidSolver3D.cuh:160
        self.iterPerFrame = 10         # :161
        self.outputFormat = FS_FMT_PLY # :162
        self.outputTemplate = "InputData/particles_"
        self.activeBody = False
        self.initialBodyPos = (0.5, 0.5, 0.5)
        self.bodyMass = 50.0
        self.bodySdfPath = "InputData/ball.sdf"
        self.particlesPerCell = 4      # FluidSolver3D.cuh:72
        self.seed = 1
        self.solver = None

    def _config(self):
        cfg = self._lib.default_config(
            dt=self.dt, dx=self.dx, pic_weight=float(self.PIC_WEIGHT), iter_per_frame=int(self.iterPerFrame),
            output_format=int(self.outputFormat), output_template=self.outputTemplate,
            active_body=int(bool(self.activeBody)), body_mass=float(self.bodyMass),
            body_sdf_path=self.bodySdfPath, particles_per_cell=int(self.particlesPerCell), seed=int(self.seed))
        cfg.body_x, cfg.body_y, cfg.body_z = (float(v) for v in self.initialBodyPos)
        return cfg

    def init(self, label_path: str):
        """FluidSolver3D::init (FluidSolver3D.cu:81-204)."""
        self.solver = Solver(self._lib, self._config())
        self.solver.load_labels_file(label_path)
        if self.activeBody:
            self.solver.load_body_sdf(self.bodySdfPath, self.initialBodyPos, self.bodyMass)
        self.solver.seed_particles(self.particlesPerCell, self.seed)
        os.makedirs("InputData", exist_ok=True)
        self.solver.write_particles("InputData/particles_-1.ply")

    def run(self, max_steps: int):
        """FluidSolver3D::run (FluidSolver3D.cu:3100-3163)."""
        os.makedirs("OutputData", exist_ok=True)
        self.solver.run(max_steps)
        return self.solver.stage_times()
