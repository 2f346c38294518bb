"""Scene generators and the reference's file formats (label files, SDF files,
PLY/OFF output), checked through the oracle's readers and writers.  No GPU."""
import os

import numpy as np
import pytest

from fluidprjct_amd import scenes
from fluidprjct_amd._abi import FS_GRID_LABELS


def test_label_writer_roundtrip(oracle, tmp_path):
    lab = scenes.shallow_basin(24, seed=3)
    p = tmp_path / "labels_gaussian.txt"
    scenes.write_labels(str(p), lab)
    s = oracle.solver()
    s.load_labels_file(str(p))
    assert s.dims == (24, 24, 24)
    assert np.array_equal(s.get_grid(FS_GRID_LABELS), lab)


def test_sdf_writer_roundtrip(oracle, tmp_path):
    sdf, origin, cs = scenes.sphere_sdf(0.1, 0.02)
    p = tmp_path / "ball.sdf"
    scenes.write_sdf(str(p), sdf, origin, cs)
    lab = scenes.bordered(20, 20, 20)
    a, b = oracle.solver(dx=0.05), oracle.solver(dx=0.05)
    for s in (a, b):
        s.set_labels(lab)
    a.load_body_sdf(str(p), (0.5, 0.5, 0.5), 1.0)
    b.set_body_sdf(sdf, origin, cs, (0.5, 0.5, 0.5), 1.0)
    la, lb = a.get_grid(FS_GRID_LABELS), b.get_grid(FS_GRID_LABELS)
    assert np.array_equal(la, lb) and (la == 3).sum() > 0


def test_scene_shapes():
    lab = scenes.dam_break(32)
    # cell centres x<=0.25, y<=0.5, z<=1 inside the one-cell border
    assert (lab == 1).sum() == 7 * 15 * 30
    lab = scenes.water_column(40, 0.4)
    assert (lab == 1).sum() == 15 * 38 * 38
    assert scenes.sphere_mass(0.08, 500.0) == pytest.approx(1.0723, abs=1e-4)


def test_ply_and_off_output(oracle, tmp_path):
    s = oracle.solver(dx=0.1)
    s.set_labels(scenes.bordered(6, 6, 6))
    pos = np.array([[0.15, 0.25, 0.35], [0.123456789, 0.3, 0.45]], np.float32)
    s.set_particles(pos)
    ply, off = tmp_path / "p.ply", tmp_path / "p.off"
    s.write_particles(str(ply), 0)
    s.write_particles(str(off), 1)
    lines = ply.read_text().splitlines()
    # save3dParticlesToPLY (utility.cu:91-109): 7-line header, default 6 significant digits
    assert lines[:7] == ["ply", "format ascii 1.0", "element vertex 2", "property float x", "property float y",
                         "property float z", "end_header"]
    assert lines[7] == "0.15 0.25 0.35" and lines[8] == "0.123457 0.3 0.45"
    assert off.read_text().splitlines()[:2] == ["OFF", "2 0 0"]


def read_ply(path):
    """(format, vertices (n, 3) float32) of an ASCII or binary little-endian PLY"""
    data = open(path, "rb").read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    header = data[:end].decode().split("\n")
    fmt = header[1].split()[1]
    n = int(header[2].split()[2])
    body = data[end:]
    if fmt == "binary_little_endian":
        v = np.frombuffer(body, dtype="<f4", count=3 * n).reshape(n, 3)
    else:
        v = np.array([list(map(float, ln.split())) for ln in body.decode().splitlines()], np.float32).reshape(n, 3)
    return fmt, v


def test_binary_and_ascii_particle_files(oracle, tmp_path):
    """utility.cu:91-109 ASCII PLY (6 significant digits) and the binary PLY
    option hold the same particles in the same order."""
    from fluidprjct_amd._abi import FS_FMT_PLY, FS_FMT_PLY_BINARY
    s = oracle.solver(dx=0.05)
    lab = np.full((6, 7, 8), 2, np.uint8)
    lab[0], lab[:, 0], lab[:, :, 0] = 0, 0, 0
    lab[1:4, 1:4, 1:5] = 1
    s.set_labels(lab)
    s.seed_particles(4, 9)
    pos, _ = s.particles()
    s.write_particles(tmp_path / "a.ply", FS_FMT_PLY)
    s.write_particles(tmp_path / "b.ply", FS_FMT_PLY_BINARY)
    fa, va = read_ply(tmp_path / "a.ply")
    fb, vb = read_ply(tmp_path / "b.ply")
    assert (fa, fb) == ("ascii", "binary_little_endian")
    assert np.array_equal(vb, pos)
    np.testing.assert_allclose(va, pos, rtol=1e-5, atol=1e-6)


def test_dam_break_2d_slab(oracle):
    """The 2D case (c1) is a slab one cell deep: 31 x 63 FLUID cells inside the
    walls, 4 particles each; stepping keeps it planar (z = 0, no w)."""
    from fluidprjct_amd._abi import FS_GRID_W
    lab = scenes.dam_break_2d(128, (0.25, 0.5))
    assert lab.shape == (1, 128, 128)
    assert int((lab == 1).sum()) == 31 * 63
    assert (lab[0, 0] == 0).all() and (lab[0, :, 0] == 0).all()
    s = oracle.solver(dt=1.0 / 120, dx=1.0 / 128, pic_weight=0.05, cfl=1.0, particles_per_cell=4, seed=1)
    s.set_labels(lab)
    s.seed_particles(4, 1)
    s.step(3)
    pos, vel = s.particles()
    assert pos.shape == (4 * 31 * 63, 3)
    assert not pos[:, 2].any() and not vel[:, 2].any() and not s.get_grid(FS_GRID_W).any()
    assert vel[:, 1].min() < 0.0   # falling under gravity
    assert (pos[:, :2] >= 0).all() and (pos[:, :2] <= 127 / 128).all()
