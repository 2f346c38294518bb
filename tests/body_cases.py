"""Hand-built rigid-body scenes shared by the oracle pins and the GPU parity
tests: one BODY cell with chosen FLUID neighbours in a uniform velocity field."""
import numpy as np

from fluidprjct_amd._abi import FS_GRID_LABELS, FS_GRID_U, FS_GRID_V, FS_GRID_W

DIRS = {"+x": (1, 0, 0), "-x": (-1, 0, 0), "+y": (0, 1, 0), "-y": (0, -1, 0), "+z": (0, 0, 1), "-z": (0, 0, -1)}


def box_sdf(sides, cs, angle_z=0.0, n=None):
    """SDF of a box with the given full side lengths (x, y, z), turned by
    angle_z about z, sampled at the voxel corners of a cube of n voxels of cs
    centred on the body frame's origin; (sdf (d, h, w), origin)."""
    n = n or int(np.ceil(1.5 * max(sides) / cs))
    c = np.arange(n) * cs - 0.5 * n * cs
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    ca, sa = np.cos(angle_z), np.sin(angle_z)
    xl, yl = ca * x + sa * y, -sa * x + ca * y      # box-frame coordinates
    sdf = np.maximum(np.maximum(np.abs(xl) - sides[0] / 2, np.abs(yl) - sides[1] / 2), np.abs(z) - sides[2] / 2)
    return sdf.astype(np.float32), np.array([-0.5 * n * cs] * 3, np.float32)


def one_face_body(lib, faces, U, V, body_pos, dx=0.125, dt=0.01, mass=2.0, N=6, cell=(2, 2, 2), sdf=None,
                  step_flags=0):
    """One BODY cell at `cell` (i, j, k) with FLUID neighbours through `faces`
    (keys of DIRS), AIR elsewhere, uniform grid velocity U, body velocity V.
    By default the SDF is positive everywhere (no voxel is inside the body, so
    no buoyancy and no cell is re-marked); its box is 4 voxels of 0.05, so
    the scalar inertia is 0.4 m (0.1)^2 (FluidSolver3D.cu:167-169).  sdf:
    (values, origin, cell size) of another body; the labels are set after
    it, so it marks no cell."""
    lab = np.full((N, N, N), 2, np.uint8)
    i, j, k = cell
    for f in faces:
        d = DIRS[f]
        lab[k + d[2], j + d[1], i + d[0]] = 1
    s = lib.solver(dx=dx, dt=dt, step_flags=step_flags)
    s.set_labels(lab)
    if sdf is None:
        s.set_body_sdf(np.ones((4, 4, 4), np.float32), (-0.1, -0.1, -0.1), 0.05, body_pos, mass)
    else:
        s.set_body_sdf(sdf[0], sdf[1], sdf[2], body_pos, mass)
    lab[k, j, i] = 3
    s.set_grid(FS_GRID_LABELS, lab)
    for g, val in ((FS_GRID_U, U[0]), (FS_GRID_V, U[1]), (FS_GRID_W, U[2])):
        shape, _ = s.grid_shape(g)
        s.set_grid(g, np.full(shape, val, np.float32))
    st = s.body_state()
    st.vel[:] = V
    s.set_body_state(st)
    return s
