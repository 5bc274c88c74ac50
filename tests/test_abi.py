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
