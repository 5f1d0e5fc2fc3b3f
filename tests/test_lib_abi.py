"""libgkhip.so loads on the CPU host and exports every entry point declared in
include/gk.h (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gk.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gk_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_functions()
    for required in ("gk_create", "gk_assemble", "gk_solve", "gk_eval", "gk_cholesky_f32", "gk_lu_f64",
                     "gk_matvec", "gk_combine_lith"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from gempy_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.fail("libgkhip.so is not built (run python -m gempy_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    # the Python binding covers the same set
    assert set(declared_functions()) == set(_lib.SIGNATURES)
    bound = _lib.load()
    assert bound.gk_version() >= 100


def test_header_constants_match_binding():
    """Every enum / #define constant of include/gk.h has the same value in the
    Python binding (the ctypes layer cannot see the header)."""
    from gempy_amd import _lib
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    consts = {}
    for body in re.findall(r"enum\s*\{([^}]*)\}", text):
        for name, val in re.findall(r"(GK_[A-Z0-9_]+)\s*=\s*(-?\d+)", body):
            consts[name] = int(val)
    for name, val in re.findall(r"#define\s+(GK_[A-Z0-9_]+)\s+(-?\d+)\b", text):
        consts[name] = int(val)
    assert len(consts) >= 15
    for name, val in consts.items():
        if name.startswith("GK_ERR_"):
            assert _lib.GK_ERR[val] == name
        elif name == "GK_OK":
            assert _lib.GK_OK == val
        else:
            assert getattr(_lib, name) == val, name
    # the gk_solve_info layout carries every field the library writes
    assert [f[0] for f in _lib.GkSolveInfo._fields_][-1] == "shift"


def test_error_reporting_without_gpu():
    from gempy_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.gk_create(0, ctypes.byref(h))
    # no device in the build container: a clean error, not a crash
    if rc != 0:
        assert lib.gk_last_error()
    else:
        lib.gk_destroy(h)


def test_backend_dispatch_errors():
    from gempy_amd import AvailableBackends, BackendTensor
    with pytest.raises(ValueError):
        BackendTensor.change_backend_gempy(AvailableBackends.numpy)
    with pytest.raises(RuntimeError):
        BackendTensor.change_backend_gempy(AvailableBackends.hip, use_gpu=False)
