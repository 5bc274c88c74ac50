"""The C ABI library exists, loads without a GPU and exports every function
include/ugks_hip.h declares.  No compute call is made here."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "ugks_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ugks_[a-z0-9_]+)\s*\(", text)))


def test_header_and_library_exports():
    import ugks2d_high_order_amd as U
    from ugks2d_high_order_amd import build
    if not os.path.exists(U.LIB_PATH):
        build.build(force=True)
    lib = ctypes.CDLL(U.LIB_PATH)
    names = _declared()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(U.EXPORTS) == set(names)


def test_package_fails_loudly_without_library(tmp_path):
    import ugks2d_high_order_amd as U
    old = U._lib
    U._lib = None
    try:
        try:
            U.load_library(str(tmp_path / "missing.so"))
            raise AssertionError("expected failure")
        except RuntimeError as e:
            assert "missing" in str(e)
    finally:
        U._lib = old


def test_fortran_binding_declares_every_core_function():
    """The Fortran module binds the core entry points under their C names."""
    import re
    from ugks2d_high_order_amd import build as B
    src = open(os.path.join(B.FORTRAN, "ugks_hip_binding.f90")).read()
    bound = set(re.findall(r'bind\(C, name="(ugks_\w+)"\)', src))
    core = {"ugks_create", "ugks_destroy", "ugks_last_error", "ugks_set_state", "ugks_get_state",
            "ugks_cold_start", "ugks_evolution", "ugks_norms", "ugks_implicit_step", "ugks_run",
            "ugks_observables", "ugks_sync"}
    assert core <= bound, core - bound
    assert bound <= set(_declared()), bound - set(_declared())
