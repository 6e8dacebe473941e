"""C ABI boundary checks that need no GPU: the library loads and exports
exactly what include/oavbd.h declares; bindings match the header."""
import ctypes
import os
import re

import pytest

from ornis_amd import avbd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "oavbd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(oavbd_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("oavbd_create", "oavbd_destroy", "oavbd_set_bodies", "oavbd_add_ball_joint", "oavbd_step",
                 "oavbd_step_n", "oavbd_get_state", "oavbd_get_pairs", "oavbd_get_colours", "oavbd_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(avbd.LIB_PATH):
        pytest.skip("liboavbd.so not built")
    lib = ctypes.CDLL(avbd.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_python_bindings_cover_the_header():
    assert set(declared_functions()) == set(avbd.SIGNATURES)


def test_load_library_binds_without_a_device():
    if not os.path.exists(avbd.LIB_PATH):
        pytest.skip("liboavbd.so not built")
    L = avbd.load_library()
    p = avbd.OavbdParams()
    L.oavbd_default_params(ctypes.byref(p))
    assert p.iters == 10 and abs(p.gravity[1] + 9.81) < 1e-6
    assert L.oavbd_last_error(None) == b"null context"
    # argument validation happens before any device call
    assert L.oavbd_create(None, 0, 0, 0, None) == avbd.EINVAL
    assert L.oavbd_step(None, ctypes.c_float(1.0 / 60.0)) == avbd.EINVAL
