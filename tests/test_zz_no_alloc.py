"""The allocation-free pressure solve (include/fluidsim_noalloc.h) against the
handle's fs_solve_pressure_system on the same systems: the same kernels in the
same order, so the same bits and iteration counts (bar: bit-exact).  Two calls
in a row, the second on a smaller system (the shared workspace keeps the first
call's contents), then a system past the size limit (rejected with a clear
error) and a call after it."""
import numpy as np
import pytest

from fluidprjct_amd import scenes
from fluidprjct_amd._abi import FS_NOALLOC_MAX_CELLS, FsError

pytestmark = pytest.mark.gpu


def system(rng, shape, block, tank=True):
    D, H, W = shape
    lab = scenes.bordered(W, H, D)
    lab[1:max(2, int(block[2] * D)), 1:max(2, int(block[1] * H)), 1:max(2, int(block[0] * W))] = 1
    b = (rng.standard_normal(lab.shape) * 1e3).astype(np.float32) * (lab == 1)
    return lab, b


def test_noalloc_solve_matches_handle_solve(hip):
    assert hip.has_noalloc
    rng = np.random.default_rng(23)
    cases = [((56, 48, 64), (0.6, 0.5, 1.0), 1e-7),     # first call
             ((20, 30, 36), (1.0, 0.4, 0.7), 1e-9)]     # second call, smaller
    for shape, block, tol in cases:
        lab, b = system(rng, shape, block)
        pr, sr = hip.solver().solve_pressure_system(lab, b, tol)
        pn, sn = hip.solve_pressure_system_noalloc(lab, b, tol)
        assert sr.converged == 1 and sn.converged == 1, shape
        assert (sn.iterations, sn.outer, sn.fluid_cells) == (sr.iterations, sr.outer, sr.fluid_cells), shape
        assert np.array_equal(pn.view(np.uint64), pr.view(np.uint64)), shape
        assert not pn[lab != 1].any()


def test_noalloc_solve_rejects_oversize_systems(hip):
    rng = np.random.default_rng(24)
    lab, b = system(rng, (64, 128, 129), (0.5, 0.5, 1.0))
    assert lab.size > FS_NOALLOC_MAX_CELLS
    with pytest.raises(FsError) as e:
        hip.solve_pressure_system_noalloc(lab, b, 1e-6)
    assert e.value.code == -1 and "at most" in str(e.value)
    # a thin line of cells under the cell limit whose workspace does not fit
    lab = np.full((1, 1, FS_NOALLOC_MAX_CELLS), 2, np.uint8)
    lab[0, 0, :3] = 0
    lab[0, 0, 3:1000] = 1
    b = np.zeros(lab.shape, np.float32)
    b[lab == 1] = 1.0
    try:
        p, st = hip.solve_pressure_system_noalloc(lab, b, 1e-6)
        pr, sr = hip.solver().solve_pressure_system(lab, b, 1e-6)
        assert np.array_equal(p.view(np.uint64), pr.view(np.uint64))
    except FsError as err:
        assert err.code == -1 and "bytes of workspace" in str(err)
    # the entry point still works after a rejected call
    lab, b = system(rng, (24, 24, 24), (0.5, 0.5, 1.0))
    pn, sn = hip.solve_pressure_system_noalloc(lab, b, 1e-7)
    pr, sr = hip.solver().solve_pressure_system(lab, b, 1e-7)
    assert sn.converged == 1 and np.array_equal(pn.view(np.uint64), pr.view(np.uint64))
