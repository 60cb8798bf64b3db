"""The allocation-free Ewald entry point (exafmm_ewald_noalloc) against the
allocating one (exafmm_ewald with device pointers), through the C ABI.

Both run the same wave-part and real-part kernels over the same wave
vectors in the same order (the allocation-free one builds its wave table on
the device in a fixed per-device array), so p and F agree bit for bit.  Two
calls in a row, the second smaller, check that the shared array's leftover
contents from the first call do not leak into the second.
"""
import ctypes as C

import numpy as np
import pytest

from minimal_amd import fmm
from minimal_amd.inputs import BODY_DTYPE, uniform_cube

pytestmark = pytest.mark.gpu

# torch first: its HIP runtime must be loaded before the library's
torch = pytest.importorskip("torch")

ALPHA, SIGMA, CUTOFF = 10.0, 0.25 / np.pi, 0.5  # 3dp/fmm.cxx's Ewald parameters


def _lib(gpu):
    L = gpu.lib
    L.exafmm_ewald_noalloc.restype = C.c_int
    L.exafmm_ewald_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int,
                                       C.c_double, C.c_double, C.c_double]
    return L


def _dev(b, torch=torch):
    t = torch.from_numpy(b.view(np.uint8).copy()).cuda()
    torch.cuda.synchronize()
    return t


def test_ewald_noalloc_matches_ewald(gpu):
    L = _lib(gpu)
    f = fmm.FMM(P=10, ncrit=64, theta=0.4, images=3, cycle=1.0)
    for n, ksize in ((4096, 11), (1500, 6)):  # the second call is smaller
        src = uniform_cube(n, 7, "sym")
        tg = src[: n // 3].copy()
        tg["p"] = 0
        tg["F"] = 0
        ds, da, db = _dev(src, torch), _dev(tg, torch), _dev(tg, torch)
        f.ewald(da.data_ptr(), ds.data_ptr(), ksize, ALPHA, SIGMA, CUTOFF, nt=len(tg), ns=n, on_device=True)
        code = L.exafmm_ewald_noalloc(f.h, db.data_ptr(), len(tg), ds.data_ptr(), n, ksize, ALPHA, SIGMA, CUTOFF)
        assert code == 0, L.exafmm_last_error(f.h)
        torch.cuda.synchronize()
        a = da.cpu().numpy().view(BODY_DTYPE)
        b = db.cpu().numpy().view(BODY_DTYPE)
        print(f"n={n} ksize={ksize}: max |p| {np.max(np.abs(a['p'])):.3e}")
        assert np.all(np.isfinite(a["p"])) and np.max(np.abs(a["p"])) > 0
        assert np.array_equal(a["p"], b["p"]) and np.array_equal(a["F"], b["F"])


def test_ewald_noalloc_rejects_oversized_wave_table(gpu):
    L = _lib(gpu)
    f = fmm.FMM(P=10, ncrit=64, theta=0.4, images=3, cycle=1.0)
    src = uniform_cube(64, 7, "sym")
    ds = _dev(src, torch)
    dt = _dev(src.copy(), torch)
    # ksize 26 needs more than the 32,768 waves the fixed array holds (25 needs 32,633)
    code = L.exafmm_ewald_noalloc(f.h, dt.data_ptr(), 64, ds.data_ptr(), 64, 26, ALPHA, SIGMA, CUTOFF)
    assert code == -6  # EXAFMM_ERR_CAPACITY
    # non-periodic context
    g = fmm.FMM(P=10, ncrit=64, theta=0.4)
    assert L.exafmm_ewald_noalloc(g.h, dt.data_ptr(), 64, ds.data_ptr(), 64, 6, ALPHA, SIGMA, CUTOFF) == -1
