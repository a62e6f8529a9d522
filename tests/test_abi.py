"""The C-ABI library loads and exports every entry point include/piny_amd.h
declares (no device needed), and the header and the binding agree."""
import ctypes
import os
import re

from piny_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_and_binding_agree():
    hdr = open(os.path.join(ROOT, "include", "piny_amd.h")).read()
    declared = set(re.findall(r"\b(pmx_[a-z_0-9]+)\s*\(", hdr))
    assert declared == set(engine.EXPORTED)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(engine.LIB_PATH)
    for name in engine.EXPORTED:
        assert hasattr(lib, name), name
    hdr = open(os.path.join(ROOT, "include", "piny_amd.h")).read()
    ver = int(re.search(r"#define PMX_ABI_VERSION (\d+)", hdr).group(1))
    assert lib.pmx_abi_version() == ver


def test_coefficient_helpers_match_oracle():
    import numpy as np
    from oracle import oracle as orc
    L = engine.lib()
    for args in ((0.21, 1.1), (0.03, 2.3)):
        a, b = np.zeros(7), np.zeros(7)
        L.pmx_bond_coef_morse(*args, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        orc.lib().or_bond_coef_morse(*args, b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert np.array_equal(a, b)
    c1, s1, c2, s2 = (np.zeros(7) for _ in range(4))
    P = ctypes.POINTER(ctypes.c_double)
    L.pmx_bend_coef_harm(0.12, 1.87, c1.ctypes.data_as(P), s1.ctypes.data_as(P))
    orc.lib().or_bend_coef_harm(0.12, 1.87, c2.ctypes.data_as(P), s2.ctypes.data_as(P))
    assert np.array_equal(c1, c2) and np.array_equal(s1, s2)


def test_create_without_device_fails_cleanly():
    L = engine.lib()
    if L.pmx_device_count() > 0:
        return
    from piny_amd import abi, water as w
    c = w.build_config(w.SPC_FW, 8, w.density_box(8), 1, 1, w.Settings(cutoff=3.0, kmax_ewd=4, kmax_pme=6))
    st = abi.SystemStruct(c.system)
    err = ctypes.create_string_buffer(256)
    h = L.pmx_create(ctypes.byref(st.sys), ctypes.byref(st.par), None, None, None, 0, err, 256)
    assert not h and b"no HIP device" in err.value
