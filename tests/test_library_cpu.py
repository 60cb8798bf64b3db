"""C-ABI library checks that need no GPU: it loads, exports every symbol the
public header declares, and fails cleanly (status code, no abort) when no
device is present."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT
from minimal_amd import fmm

HEADER = os.path.join(ROOT, "include", "exafmm_mi355x.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(exafmm_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert set(declared()) == set(fmm.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = fmm.load()
    for name in declared():
        assert hasattr(lib.lib, name), name


def test_config_defaults():
    cfg = fmm.default_config()
    assert (cfg.P, cfg.ncrit, cfg.theta) == (10, 64, 0.4)
    assert cfg.m2l_truncation == fmm.M2L_FULL
    assert cfg.p2p_precision == fmm.PREC_FP32
    assert C.sizeof(fmm.Config) == 72


def test_body_layout():
    from minimal_amd.inputs import BODY_DTYPE
    assert BODY_DTYPE.itemsize == 64
    assert fmm.CELL_DTYPE.itemsize == 56


def test_create_reports_errors_without_device():
    lib = fmm.load()
    if lib.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(fmm.ExaFMMError) as e:
        fmm.FMM()
    assert e.value.code == -7


def test_bad_config_is_rejected():
    lib = fmm.load()
    for bad in ({"P": 1}, {"P": 40}, {"ncrit": 0}, {"theta": 0.0}, {"m2l_truncation": 5}, {"images": -1}):
        cfg = fmm.default_config(**bad)
        h = C.c_void_p()
        code = lib.lib.exafmm_create(C.byref(cfg), C.byref(h))
        assert code in (-1, -7), bad
        assert not h.value


def test_body_arrays_are_validated():
    """The library copies n * 64 bytes through a body pointer: arrays that
    are strided, of another dtype, read-only or shorter than n raise before
    any call."""
    import numpy as np
    from minimal_amd.inputs import BODY_DTYPE
    b = np.zeros(10, BODY_DTYPE)
    p, n = fmm._bodies(b)
    assert n == 10 and p.value == b.ctypes.data
    with pytest.raises(ValueError):
        fmm._bodies(b[::2])
    with pytest.raises(TypeError):
        fmm._bodies(np.zeros((10, 8)))
    with pytest.raises(ValueError):
        fmm._bodies(b, n=11)
    ro = b.copy()
    ro.flags.writeable = False
    with pytest.raises(ValueError):
        fmm._bodies(ro)
    assert fmm._bodies(ro, write=False)[1] == 10
    with pytest.raises(ValueError):
        fmm._bodies(12345)  # a device pointer needs a count
    assert fmm._bodies(12345, n=4)[1] == 4
