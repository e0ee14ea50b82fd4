"""The C-ABI library: it loads without a GPU and exports every function that
include/amcafe.h declares; the error path works without touching a device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "amcafe.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(amcafe_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_seams():
    names = declared_functions()
    for must in ("amcafe_domain_create", "amcafe_analytic_temp_curr", "amcafe_update_voxels", "amcafe_sim_create",
                 "amcafe_sim_step", "amcafe_last_error"):
        assert must in names


def test_library_exports_every_declared_function():
    import amcafe_amd
    L = amcafe_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"declared in amcafe.h but not exported: {missing}"


def test_argument_errors_need_no_device():
    import amcafe_amd
    L = amcafe_amd.lib()
    L.amcafe_last_error.restype = ctypes.c_char_p
    L.amcafe_sim_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    out = ctypes.c_void_p()
    rc = L.amcafe_sim_create(None, 0, ctypes.byref(out))
    assert rc == -1  # AMCAFE_ERR_ARG
    assert L.amcafe_last_error()
    L.amcafe_domain_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    assert L.amcafe_domain_create(None, ctypes.byref(out)) == -1
    # the driver-level queries refuse a null simulation or output before any device call
    buf = ctypes.create_string_buffer(256)
    for name in ("amcafe_sim_tuning", "amcafe_sim_stats", "amcafe_sim_info"):
        f = getattr(L, name)
        if name == "amcafe_sim_tuning":
            f.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            rc = f(None, buf)
        else:
            f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
            rc = f(None, 0, buf)
        assert rc == -1, name
        assert L.amcafe_last_error(), name


def test_product_path_has_no_cpu_fallback(monkeypatch, tmp_path):
    """Without the built library the package refuses to run rather than falling back."""
    import importlib
    import amcafe_amd._lib as lib_mod
    monkeypatch.setattr(lib_mod, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(lib_mod, "_lib", None)
    with pytest.raises(lib_mod.LibraryMissing):
        lib_mod.lib()
