"""fs_run (FluidSolver3D::run, FluidSolver3D.cu:3100-3163) on the GPU: one
particle file per frame named outputTemplate + frame, written on a host thread
while the next steps run; the last frame equals the final state."""
import numpy as np
import pytest

from fluidprjct_amd import scenes
from fluidprjct_amd._abi import FS_FMT_PLY, FS_FMT_PLY_BINARY
from test_scenes_io import read_ply

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", [FS_FMT_PLY_BINARY, FS_FMT_PLY])
def test_run_writes_every_frame(hip, tmp_path, fmt):
    N = 24
    tmpl = str(tmp_path / "particles_")
    s = hip.solver(dt=1.0 / 120, dx=1.0 / N, pic_weight=0.05, cfl=1.0, iter_per_frame=2, output_format=fmt,
                   output_template=tmpl, particles_per_cell=4, seed=5)
    s.set_labels(scenes.dam_break(N, (0.3, 0.5, 0.8)))
    s.seed_particles(4, 5)
    s.run(3)
    files = sorted(tmp_path.glob("particles_*.ply"))
    assert [f.name for f in files] == [f"particles_{i}.ply" for i in range(4)]
    assert s.clock()[2] == 3 * 2 + 1
    s.write_particles(tmp_path / "final.ply", fmt)
    assert (tmp_path / "particles_3.ply").read_bytes() == (tmp_path / "final.ply").read_bytes()
    pos, _ = s.particles()
    _, v = read_ply(tmp_path / "particles_3.ply")
    if fmt == FS_FMT_PLY_BINARY:
        assert np.array_equal(v, pos)
    else:
        np.testing.assert_allclose(v, pos, rtol=1e-5, atol=1e-6)
    _, v0 = read_ply(files[0])
    assert v0.shape == pos.shape and not np.array_equal(v0, pos)
