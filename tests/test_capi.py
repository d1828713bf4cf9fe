"""The C-ABI library loads and exports every function include/yuzen_gpu.h
declares, and the ctypes mirrors of its structures have the compiled sizes
(no device call is made here)."""
import ctypes as C
import os
import re

from yuzen_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "yuzen_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(yz_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    lib = capi.load()
    names = declared_functions()
    assert "yz_step_advance" in names and "yz_linsolve" in names
    for n in names:
        assert hasattr(lib, n), n


def test_structure_sizes_match():
    lib = capi.load()
    assert lib.yz_sizeof_case() == C.sizeof(capi.YzCase)
    assert lib.yz_sizeof_lin_settings() == C.sizeof(capi.YzLinSettings)
    assert lib.yz_sizeof_cz_case() == C.sizeof(capi.YzCzCase)


def test_defaults_agree():
    lib = capi.load()
    c = capi.YzCase()
    lib.yz_case_defaults(C.byref(c))
    d = capi.case_defaults()
    for name, _ in capi.YzCase._fields_:
        a, b = getattr(c, name), getattr(d, name)
        if name == "bc":
            assert list(a) == list(b)
        else:
            assert a == b, name


def test_cz_defaults_agree():
    lib = capi.load()
    c = capi.YzCzCase()
    lib.yz_cz_case_defaults(C.byref(c))
    d = capi.cz_case_defaults()
    for name, _ in capi.YzCzCase._fields_:
        a, b = getattr(c, name), getattr(d, name)
        if name == "bc":
            assert list(a) == list(b)
        else:
            assert a == b, name
