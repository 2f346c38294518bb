"""Synthetic scenes of the benchmark configurations, plus writers for the
reference's scene formats.

Label grids are uint8 arrays of shape (D, H, W), i fastest, with the values of
utility.cuh:65-68 (SOLID 0, FLUID 1, AIR 2).  Every scene has a one-cell SOLID
border (the reference's edge reads assume one, FluidSolver3D.cu:1406,2109).

Formats written here are the ones the reference reads:
  - label files: "W H D" then W*H*D characters S/F/A, first character = cell
    (W-1,H-1,D-1) (FluidSolver3D.cu:88-153);
  - SDF files: "w h d", origin line, cell-size line, then k->j->i values
    (RigidBody::loadSDF utility.cuh:231-338).  The reference needs
    InputData/ball.sdf and InputData/labels_gaussian.txt but ships neither;
    `write_sdf(sphere_sdf(...))` and `shallow_basin(...)` generate them.
"""
from __future__ import annotations

import numpy as np

SOLID, FLUID, AIR = 0, 1, 2


def _centres(n, dx):
    return (np.arange(n, dtype=np.float64) + 0.5) * dx


def bordered(W, H, D):
    lab = np.full((D, H, W), AIR, np.uint8)
    lab[0], lab[-1] = SOLID, SOLID
    lab[:, 0], lab[:, -1] = SOLID, SOLID
    lab[:, :, 0], lab[:, :, -1] = SOLID, SOLID
    return lab


def dam_break(N, extent=(0.25, 0.5, 1.0), dx=None):
    """Block x<=ex, y<=ey, z<=ez of fluid in a unit box with solid walls."""
    dx = dx or 1.0 / N
    lab = bordered(N, N, N)
    x, y, z = _centres(N, dx), _centres(N, dx), _centres(N, dx)
    m = (z[:, None, None] <= extent[2]) & (y[None, :, None] <= extent[1]) & (x[None, None, :] <= extent[0])
    lab[m & (lab == AIR)] = FLUID
    return lab


def dam_break_2d(N, extent=(0.25, 0.5), dx=None):
    """The 2D dam break (block x<=ex, y<=ey) as a slab one cell deep: solid walls
    in x and y, D = 1 (no z neighbours, so w stays 0 and advection clamps z to
    0 - the 3D step restricted to a plane)."""
    dx = dx or 1.0 / N
    lab = np.full((1, N, N), AIR, np.uint8)
    lab[:, 0], lab[:, -1] = SOLID, SOLID
    lab[:, :, 0], lab[:, :, -1] = SOLID, SOLID
    x, y = _centres(N, dx), _centres(N, dx)
    m = (y[None, :, None] <= extent[1]) & (x[None, None, :] <= extent[0])
    lab[m & (lab == AIR)] = FLUID
    return lab


def water_column(N, level=0.4, dx=None):
    """Fluid layer y < level in a unit box with solid walls."""
    dx = dx or 1.0 / N
    lab = bordered(N, N, N)
    y = _centres(N, dx)
    m = np.broadcast_to((y < level)[None, :, None], lab.shape)
    lab[m & (lab == AIR)] = FLUID
    return lab


def sphere_sdf(radius, cell_size, margin_cells=2):
    """Signed distance |x - c| - r sampled on voxel corners of a cube around the
    sphere; returns (sdf (d,h,w) float32, origin (3,), cell_size).  The voxel
    grid is centred on the origin of the body frame."""
    n = int(np.ceil(2 * radius / cell_size)) + 2 * margin_cells
    half = 0.5 * n * cell_size
    c = np.arange(n) * cell_size - half
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    sdf = (np.sqrt(xx * xx + yy * yy + zz * zz) - radius).astype(np.float32)
    return sdf, np.array([-half, -half, -half], np.float32), float(cell_size)


def shallow_basin(N, seed=5, n_gauss=32, level=0.3, tilt=0.05, dx=None):
    """Floor = sum of n_gauss Gaussians (amplitude U[0.02,0.1], width
    U[0.03,0.15]) as a solid SDF; fluid below y = level + tilt*(x - 0.5)."""
    dx = dx or 1.0 / N
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.02, 0.1, n_gauss)
    wid = rng.uniform(0.03, 0.15, n_gauss)
    cx, cz = rng.uniform(0, 1, n_gauss), rng.uniform(0, 1, n_gauss)
    x = _centres(N, dx)
    X, Z = np.meshgrid(x, x, indexing="xy")   # (z, x)
    h = np.zeros_like(X)
    for a, w, x0, z0 in zip(amp, wid, cx, cz):
        h += a * np.exp(-((X - x0) ** 2 + (Z - z0) ** 2) / (2 * w * w))
    y = _centres(N, dx)
    lab = bordered(N, N, N)
    phi_solid = y[None, :, None] - h[:, None, :]                 # (z, y, x)
    surf = level + tilt * (x[None, None, :] - 0.5)
    phi_fluid = y[None, :, None] - surf
    inner = lab == AIR
    lab[inner & (phi_solid < 0)] = SOLID
    lab[inner & (phi_solid >= 0) & (phi_fluid < 0)] = FLUID
    return lab


def write_labels(path, lab):
    D, H, W = lab.shape
    ch = np.array([b"S", b"F", b"A", b"S"])[lab.reshape(-1)[::-1]]
    with open(path, "wb") as f:
        f.write(f"{W} {H} {D}\n".encode())
        for row in range(0, ch.size, W):
            f.write(b" ".join(ch[row:row + W]) + b"\n")


def write_sdf(path, sdf, origin, cell_size):
    d, h, w = sdf.shape
    with open(path, "w") as f:
        f.write(f"{w} {h} {d}\n{origin[0]:.9g} {origin[1]:.9g} {origin[2]:.9g}\n{cell_size:.9g}\n")
        flat = sdf.reshape(-1)
        for i in range(0, flat.size, w):
            f.write(" ".join(f"{v:.9g}" for v in flat[i:i + w]) + "\n")


def sphere_mass(radius, density):
    return float(density * 4.0 / 3.0 * np.pi * radius ** 3)
