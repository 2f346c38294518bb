"""The C ABI: both libraries load and export every function include/fluidsim.h
declares; the ctypes layer matches the header; the HIP library refuses to run
without a device instead of falling back to the CPU.  Runs without a GPU."""
import ctypes as C
import os
import re

import pytest

from fluidprjct_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fluidsim.h")
LEARNED_HEADER = os.path.join(ROOT, "include", "fluidsim_learned.h")
CKPT_HEADER = os.path.join(ROOT, "include", "fluidsim_checkpoint.h")
NOALLOC_HEADER = os.path.join(ROOT, "include", "fluidsim_noalloc.h")


def header_functions(path=HEADER):
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fs_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_ctypes_agree():
    assert header_functions() == sorted(_abi.ABI_SYMBOLS)
    assert header_functions(LEARNED_HEADER) == sorted(_abi.LEARNED_SYMBOLS)
    assert header_functions(CKPT_HEADER) == sorted(_abi.CHECKPOINT_SYMBOLS)
    assert header_functions(NOALLOC_HEADER) == sorted(_abi.NOALLOC_SYMBOLS)
    txt = open(NOALLOC_HEADER).read()
    assert int(eval(__import__("re").search(r"#define FS_NOALLOC_MAX_CELLS (.*)", txt).group(1))) == \
        _abi.FS_NOALLOC_MAX_CELLS


def test_learned_parameter_count_matches_header():
    txt = open(LEARNED_HEADER).read()
    expr = re.search(r"#define FSL_NUM_PARAMS (.*)", txt).group(1)
    assert eval(expr) == _abi.FSL_NUM_PARAMS == 22529


def test_hip_library_exports_learned_symbols():
    path = os.path.join(ROOT, "fluidprjct_amd", "libfluidsim_hip.so")
    dll = C.CDLL(path)
    missing = [f for f in header_functions(LEARNED_HEADER) + header_functions(CKPT_HEADER) +
               header_functions(NOALLOC_HEADER) if not hasattr(dll, f)]
    assert not missing, missing
    fl = _abi.FsLib(path)
    assert fl.has_learned and fl.has_checkpoint and fl.has_noalloc


@pytest.mark.parametrize("lib", ["hip", "oracle"])
def test_library_exports_every_symbol(lib):
    path = os.path.join(ROOT, "fluidprjct_amd", "libfluidsim_hip.so") if lib == "hip" else \
        os.path.join(ROOT, "oracle", "_build", "libfs_oracle.so")
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (python -c 'import __graft_entry__ as g; g.build()')")
    dll = C.CDLL(path)
    missing = [f for f in header_functions() if not hasattr(dll, f)]
    assert not missing, missing
    fl = _abi.FsLib(path)
    assert fl.backend == lib


def test_config_struct_layout_and_defaults(oracle):
    """main.cpp:18-52 defaults through the C struct (checks the ctypes layout)."""
    cfg = oracle.default_config()
    assert cfg.output_template == b"InputData/particles_"
    assert cfg.input_label_data == b"InputData/labels.txt"
    assert cfg.particles_per_cell == 4 and cfg.density == 1000.0
    assert cfg.gravity_y == pytest.approx(-9.81) and cfg.advect_c == pytest.approx(0.1)
    assert cfg.solver_tol == pytest.approx(1e-6) and cfg.stall_window == 5 and cfg.cfl == 0.0


def test_hip_library_fails_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from fluidprjct_amd import load_library
    lib = load_library()
    with pytest.raises(_abi.FsError) as e:
        lib.solver()
    assert e.value.code == -5


def test_package_refuses_missing_library(monkeypatch, tmp_path):
    import fluidprjct_amd
    monkeypatch.setattr(fluidprjct_amd, "_LIB", None)
    monkeypatch.setattr(fluidprjct_amd, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError):
        fluidprjct_amd.load_library()


def test_oracle_error_paths(oracle, tmp_path):
    s = oracle.solver()
    with pytest.raises(_abi.FsError) as e:
        s.load_labels_file(str(tmp_path / "nope.txt"))
    assert e.value.code == -2 and "Cannot open file" in str(e.value)
    with pytest.raises(_abi.FsError):
        s.step(1)
