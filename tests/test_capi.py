"""The C-ABI library loads on the CPU host and exports every entry point that
include/mbem_hip.h declares (no compute calls: there is no GPU here)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mbem_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t)\s+(mbem_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert len(names) >= 20, names


def test_library_exports_every_declared_symbol():
    from mbem_amd import _lib
    lib = _lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    from mbem_amd import _lib
    assert set(_declared()) == set(_lib.exported_names())


def test_src_fields_constant():
    from mbem_amd import _lib
    assert _lib.load().mbem_src_fields() == _lib.SRC_FIELDS
