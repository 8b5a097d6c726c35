"""The C-ABI library loads and exports every symbol include/*.h declares
(CPU only: nothing here launches device work)."""
import ctypes
import os
import re

from generic_mobility_interface_class_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(gmi_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    bound = {n for n, _, _ in capi.SIGNATURES}
    assert declared_symbols() <= bound


def test_create_rejects_bad_parameters_without_device():
    lib = capi.load()
    ns = (ctypes.c_int * 1)(20)
    nb = (ctypes.c_int * 1)(42)
    # nseg_per_swimmer not a multiple of nseg_per_filament -> GMI_ERR_ARG,
    # rejected before any device call
    p = capi.GmiParams(1, ctypes.cast(ns, ctypes.POINTER(ctypes.c_int)),
                       ctypes.cast(nb, ctypes.POINTER(ctypes.c_int)), 3, 1, 1.0,
                       1.0, 1.0, 0, 0)
    h = ctypes.c_void_p()
    assert lib.gmi_create(ctypes.byref(h), ctypes.byref(p)) == capi.GMI_ERR_ARG
    assert not h.value
    p.nseg_per_filament = 20
    p.mu = -1.0
    assert lib.gmi_create(ctypes.byref(h), ctypes.byref(p)) == capi.GMI_ERR_ARG
    p.mu = 1.0
    p.precision = 7
    assert lib.gmi_create(ctypes.byref(h), ctypes.byref(p)) == capi.GMI_ERR_ARG


def test_device_count_without_gpu_is_a_number():
    assert capi.device_count() >= 0
