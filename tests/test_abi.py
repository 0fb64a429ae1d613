"""C-ABI library: loads, exports every declared symbol, rejects bad arguments
without touching a GPU (all checks run before any launch)."""

import ctypes as C
import re
from pathlib import Path

import pytest

import flowmc_amd._native as n

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(fmc_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_loads_and_version():
    lib = n.load()
    assert lib.fmc_abi_version() == n.ABI_VERSION


def test_every_declared_symbol_is_exported_and_bound():
    lib = n.load()
    decl = declared_functions()
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"
        assert name in n.exported_names(), f"{name} has no ctypes signature"


def test_invalid_arguments_return_status_and_message():
    lib = n.load()
    # NULL target
    rc = lib.fmc_mala_run(None, None, None, None, 10, 1, 0, 0.1, 0, 0, None, None, None, None, None, 1, 0, 1, None)
    assert rc == 1
    assert "target is NULL" in n.last_error()
    # bad target kind
    t = n.fmc_target(99, 4, 0, 0, 0x1000, 8)
    rc = lib.fmc_mala_run(None, C.byref(t), 0x1000, 0x1000, 10, 1, 0, 0.1, 0, 0, None, None, None, None, None, 1, 0,
                          1, None)
    assert rc == 1 and "unknown target kind" in n.last_error()
    # not enough params
    t = n.fmc_target(0, 4, 0, 0, 0x1000, 3)
    rc = lib.fmc_target_eval(None, C.byref(t), 0x1000, 5, 0x1000, None, None)
    assert rc == 1 and "needs 8 params" in n.last_error()
    # layout that cannot hold d
    t = n.fmc_target(2, 100, 0, 0, 0x1000, 1)
    cfg = n.fmc_launch_cfg(4, 1, 256, 0)
    rc = lib.fmc_target_eval(None, C.byref(t), 0x1000, 5, 0x1000, None, C.byref(cfg))
    assert rc == 1 and "holds 16 dims" in n.last_error()
    # buffer row overflow
    t = n.fmc_target(2, 8, 0, 0, 0x1000, 1)
    rc = lib.fmc_mala_run(None, C.byref(t), 0x1000, 0x1000, 10, 64, 0, 0.1, 0, 0, None, None, 0x1000, None, None, 10, 0,
                          1, None)
    assert rc == 1 and "out of range" in n.last_error()
    # thinning
    rc = lib.fmc_nf_mh(None, 1, 1, 1, 1, 1, 1, 4, 5, 2, 0, 0, None, 1, None, None, None, None, None, 10, 0, 0)
    assert rc == 1 and "thinning" in n.last_error()


def test_python_wrapper_raises_on_failure():
    with pytest.raises(n.FlowMCNativeError) as ei:
        n.call("fmc_hmc_drift", None, None, None, None, 1, 1, 0.1)
    assert ei.value.status == 1
