"""Regenerate the committed fixtures under tests/golden/ (run in the build
container, where the reference tree is mounted at /root/reference).

- inputdata/: the reference's scene data files (InputData/*.txt, config.json)
  copied verbatim; they are data, read by the tests in place of the reference.
- particles_init.npz: positions of InputData/particles_-1.ply (the reference's
  init dump, FluidSolver3D.cu:201) as float32.
- pressure_golden.npz: small label grids, right-hand sides and their FP64
  solutions by scipy.sparse.linalg.spsolve of the pressure matrix assembled
  from the rule of constructA (FluidSolver3D.cu:1598-1777).
"""
import os
import shutil

import numpy as np

REF = "/root/reference/InputData"
HERE = os.path.dirname(os.path.abspath(__file__))


def copy_inputs():
    out = os.path.join(HERE, "inputdata")
    os.makedirs(out, exist_ok=True)
    for f in ("config.json", "labels_3d_sphere.txt", "labels_bathymetry.txt", "labels_colimn.txt",
              "labels_simple.txt", "labels_simple3d.txt"):
        shutil.copyfile(os.path.join(REF, f), os.path.join(out, f))


def ply_positions():
    with open(os.path.join(REF, "particles_-1.ply")) as f:
        lines = f.read().splitlines()
    n = int(lines[2].split()[2])
    start = lines.index("end_header") + 1
    pos = np.array([[float(t) for t in ln.split()] for ln in lines[start:start + n]], np.float32)
    np.savez_compressed(os.path.join(HERE, "particles_init.npz"), pos=pos)


def assemble(lab):
    """SPD matrix on FLUID cells (linear i-fastest numbering): diag = number of
    in-domain FLUID/AIR neighbours, -1 per FLUID neighbour."""
    import scipy.sparse as sp
    D, H, W = lab.shape
    fl = np.argwhere(lab.reshape(-1) == 1).ravel()
    num = -np.ones(lab.size, np.int64)
    num[fl] = np.arange(fl.size)
    rows, cols, vals = [], [], []
    for r, c in enumerate(fl):
        k, rem = divmod(c, W * H)
        j, i = divmod(rem, W)
        diag = 0
        for di, dj, dk in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            ni, nj, nk = i + di, j + dj, k + dk
            if not (0 <= ni < W and 0 <= nj < H and 0 <= nk < D):
                continue
            t = lab[nk, nj, ni]
            if t == 1:
                rows.append(r); cols.append(num[ni + nj * W + nk * W * H]); vals.append(-1.0)
                diag += 1
            elif t == 2:
                diag += 1
        rows.append(r); cols.append(r); vals.append(float(diag))
    return sp.csr_matrix((vals, (rows, cols)), shape=(fl.size, fl.size)), fl


def pressure_golden():
    from scipy.sparse.linalg import spsolve
    rng = np.random.default_rng(2024)
    cases = {}
    # tank with a dam-break block, open air above
    lab = np.full((10, 12, 9), 2, np.uint8)
    lab[0], lab[-1] = 0, 0
    lab[:, 0], lab[:, :, 0], lab[:, :, -1] = 0, 0, 0
    lab[1:-1, 1:6, 1:5] = 1
    cases["dam"] = lab
    # random fluid/air/solid mixture with a solid border (many Dirichlet contacts)
    lab = rng.choice(np.array([1, 1, 1, 2, 0], np.uint8), size=(8, 9, 10))
    lab[0], lab[-1], lab[:, 0], lab[:, -1], lab[:, :, 0], lab[:, :, -1] = 0, 0, 0, 0, 0, 0
    lab[1, 1, 1] = 2
    cases["mix"] = lab
    out = {}
    for name, lab in cases.items():
        A, fl = assemble(lab)
        b = np.zeros(lab.size, np.float32)
        b[fl] = rng.standard_normal(fl.size).astype(np.float32) * 100
        x = spsolve(A.tocsc(), b[fl].astype(np.float64))
        p = np.zeros(lab.size, np.float64)
        p[fl] = x
        out[name + "_labels"] = lab
        out[name + "_b"] = b.reshape(lab.shape)
        out[name + "_p"] = p.reshape(lab.shape)
    np.savez_compressed(os.path.join(HERE, "pressure_golden.npz"), **out)


if __name__ == "__main__":
    copy_inputs()
    ply_positions()
    pressure_golden()
    print("fixtures written to", HERE)
