"""ctypes bindings for the FP64 CPU oracle and the reference shims.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BODY_DTYPE = np.dtype([("X", "<f8", (3,)), ("q", "<f8"), ("p", "<f8"), ("F", "<f8", (3,))])
CELL_DTYPE = np.dtype([("nchild", "<i4"), ("nbody", "<i4"), ("child", "<i8"), ("body", "<i8"),
                       ("X", "<f8", (3,)), ("R", "<f8")])

_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_L = C.c_int64


def build(reference: str | None = None) -> None:
    """Compile liborc.so (and the reference shims when the sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    ref = reference or os.environ.get("EXAFMM_REFERENCE", "/root/reference")
    if os.path.isdir(os.path.join(ref, "3d")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REFERENCE={ref}"], check=True)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """FP64 restatement of the reference 3d/ path (own code, liborc.so)."""

    def __init__(self) -> None:
        path = os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        L = self.lib
        L.orc_create.restype = _P
        L.orc_create.argtypes = [_I, _I, _D, _I]
        L.orc_destroy.argtypes = [_P]
        for name in ("orc_build_tree",):
            getattr(L, name).argtypes = [_P, _P, _L]
        L.orc_ncells.restype = _L
        L.orc_ncells.argtypes = [_P]
        L.orc_get_cells.argtypes = [_P, _P]
        L.orc_get_levels.argtypes = [_P, _P]
        L.orc_upward.argtypes = [_P, _P]
        L.orc_make_lists.argtypes = [_P]
        L.orc_list_total.restype = _L
        L.orc_list_total.argtypes = [_P, _I]
        L.orc_get_lists.argtypes = [_P, _I, _P, _P]
        L.orc_evaluate.argtypes = [_P, _P]
        L.orc_downward.argtypes = [_P, _P]
        L.orc_get_M.argtypes = [_P, _P]
        L.orc_get_L.argtypes = [_P, _P]
        L.orc_set_M.argtypes = [_P, _P]
        L.orc_set_L.argtypes = [_P, _P]
        L.orc_fmm.argtypes = [_P, _L, _I, _I, _D, _I, _P]
        L.orc_fmm_periodic.argtypes = [_P, _L, _I, _I, _D, _I, _I, _D, _P]
        L.orc_direct_periodic.argtypes = [_P, _L, _P, _L, _I, _D]
        L.orc_ewald.argtypes = [_P, _L, _P, _L, _I, _D, _D, _D, _D, _I]
        L.orc_p2p.argtypes = [_P, _L, _P, _L]
        L.orc_p2p_shift.argtypes = [_P, _L, _P, _L, _P]
        L.orc_p2m.argtypes = [_I, _P, _P, _L, _P]
        L.orc_m2m.argtypes = [_I, _P, _P, _P, _P]
        L.orc_m2l.argtypes = [_I, _I, _P, _P, _P, _P]
        L.orc_l2l.argtypes = [_I, _P, _P, _P, _P]
        L.orc_l2p.argtypes = [_I, _P, _P, _P, _L]
        L.orc_kat.argtypes = [_I, _I, _P]
        L.orc_eval_multipole.argtypes = [_I, _D, _D, _D, _P, _P]
        L.orc_eval_local.argtypes = [_I, _I, _D, _D, _D, _P]
        L.orc_num_threads.restype = _I

    # ---- stateless operators; coefficient arrays are complex128[NTERM]
    def p2p(self, targets: np.ndarray, sources: np.ndarray) -> None:
        self.lib.orc_p2p(_ptr(targets), len(targets), _ptr(sources), len(sources))

    def p2p_shift(self, targets, sources, shift) -> None:
        s = np.ascontiguousarray(shift, dtype=np.float64)
        self.lib.orc_p2p_shift(_ptr(targets), len(targets), _ptr(sources), len(sources), _ptr(s))

    def p2m(self, P, center, bodies, M=None):
        M = np.zeros(P * (P + 1) // 2, np.complex128) if M is None else M.copy()
        c = np.ascontiguousarray(center, dtype=np.float64)
        self.lib.orc_p2m(P, _ptr(c), _ptr(bodies), len(bodies), _ptr(M))
        return M

    def m2m(self, P, cparent, cchild, Mchild, Mparent=None):
        Mp = np.zeros(P * (P + 1) // 2, np.complex128) if Mparent is None else Mparent.copy()
        a = np.ascontiguousarray(cparent, dtype=np.float64)
        b = np.ascontiguousarray(cchild, dtype=np.float64)
        Mc = np.ascontiguousarray(Mchild, dtype=np.complex128)
        self.lib.orc_m2m(P, _ptr(a), _ptr(b), _ptr(Mc), _ptr(Mp))
        return Mp

    def m2l(self, P, full, Xi, Xj, Mj, Li=None):
        Lo = np.zeros(P * (P + 1) // 2, np.complex128) if Li is None else Li.copy()
        a = np.ascontiguousarray(Xi, dtype=np.float64)
        b = np.ascontiguousarray(Xj, dtype=np.float64)
        Mc = np.ascontiguousarray(Mj, dtype=np.complex128)
        self.lib.orc_m2l(P, int(full), _ptr(a), _ptr(b), _ptr(Mc), _ptr(Lo))
        return Lo

    def l2l(self, P, cparent, Lparent, cchild, Lchild=None):
        Lc = np.zeros(P * (P + 1) // 2, np.complex128) if Lchild is None else Lchild.copy()
        a = np.ascontiguousarray(cparent, dtype=np.float64)
        b = np.ascontiguousarray(cchild, dtype=np.float64)
        Lp = np.ascontiguousarray(Lparent, dtype=np.complex128)
        self.lib.orc_l2l(P, _ptr(a), _ptr(Lp), _ptr(b), _ptr(Lc))
        return Lc

    def l2p(self, P, center, L, bodies) -> None:
        c = np.ascontiguousarray(center, dtype=np.float64)
        Lc = np.ascontiguousarray(L, dtype=np.complex128)
        self.lib.orc_l2p(P, _ptr(c), _ptr(Lc), _ptr(bodies), len(bodies))

    def kat(self, P, full=False):
        out = np.zeros(10)
        self.lib.orc_kat(P, int(full), _ptr(out))
        return out

    def eval_multipole(self, P, rho, alpha, beta):
        Y = np.zeros(P * P, np.complex128)
        Yt = np.zeros(P * P, np.complex128)
        self.lib.orc_eval_multipole(P, rho, alpha, beta, _ptr(Y), _ptr(Yt))
        return Y, Yt

    def eval_local(self, P, fillP, rho, alpha, beta):
        Y = np.zeros(4 * P * P, np.complex128)
        self.lib.orc_eval_local(P, fillP, rho, alpha, beta, _ptr(Y))
        return Y

    # ---- whole evaluation
    def fmm(self, bodies, P, ncrit, theta, full=False):
        times = np.zeros(5)
        st = self.lib.orc_fmm(_ptr(bodies), len(bodies), P, ncrit, theta, int(full), _ptr(times))
        if st != 0:
            raise RuntimeError(f"orc_fmm failed: {st}")
        return times

    def fmm_periodic(self, bodies, P, ncrit, theta, images, cycle, full=False):
        times = np.zeros(5)
        st = self.lib.orc_fmm_periodic(_ptr(bodies), len(bodies), P, ncrit, theta, int(full),
                                       images, cycle, _ptr(times))
        if st != 0:
            raise RuntimeError(f"orc_fmm_periodic failed: {st}")
        return times

    def direct_periodic(self, targets, sources, images, cycle):
        self.lib.orc_direct_periodic(_ptr(targets), len(targets), _ptr(sources), len(sources), images, cycle)

    def ewald(self, targets, sources, ksize, alpha, sigma, cutoff, cycle, ncrit=64):
        self.lib.orc_ewald(_ptr(targets), len(targets), _ptr(sources), len(sources), ksize, alpha, sigma,
                           cutoff, cycle, ncrit)

    def session(self, P, ncrit, theta, full=False):
        return OracleSession(self, P, ncrit, theta, full)

    def threads(self) -> int:
        return int(self.lib.orc_num_threads())


class OracleSession:
    def __init__(self, orc: Oracle, P, ncrit, theta, full):
        self.o = orc
        self.P = P
        self.nterm = P * (P + 1) // 2
        self.h = orc.lib.orc_create(P, ncrit, theta, int(full))
        if not self.h:
            raise ValueError("bad oracle configuration")

    def __del__(self):
        if getattr(self, "h", None):
            self.o.lib.orc_destroy(self.h)
            self.h = None

    def build_tree(self, bodies) -> int:
        return self.o.lib.orc_build_tree(self.h, _ptr(bodies), len(bodies))

    def cells(self) -> np.ndarray:
        n = self.o.lib.orc_ncells(self.h)
        out = np.zeros(n, CELL_DTYPE)
        self.o.lib.orc_get_cells(self.h, _ptr(out))
        return out

    def levels(self) -> np.ndarray:
        n = self.o.lib.orc_ncells(self.h)
        out = np.zeros(n, np.int32)
        self.o.lib.orc_get_levels(self.h, _ptr(out))
        return out

    def upward(self, bodies):
        return self.o.lib.orc_upward(self.h, _ptr(bodies))

    def make_lists(self):
        return self.o.lib.orc_make_lists(self.h)

    def lists(self, which: int):
        n = self.o.lib.orc_ncells(self.h)
        tot = self.o.lib.orc_list_total(self.h, which)
        off = np.zeros(n + 1, np.int64)
        src = np.zeros(max(tot, 1), np.int64)
        self.o.lib.orc_get_lists(self.h, which, _ptr(off), _ptr(src))
        return off, src[:tot]

    def evaluate(self, bodies):
        return self.o.lib.orc_evaluate(self.h, _ptr(bodies))

    def downward(self, bodies):
        return self.o.lib.orc_downward(self.h, _ptr(bodies))

    def M(self):
        n = self.o.lib.orc_ncells(self.h)
        out = np.zeros(n * self.nterm, np.complex128)
        self.o.lib.orc_get_M(self.h, _ptr(out))
        return out.reshape(n, self.nterm)

    def L(self):
        n = self.o.lib.orc_ncells(self.h)
        out = np.zeros(n * self.nterm, np.complex128)
        self.o.lib.orc_get_L(self.h, _ptr(out))
        return out.reshape(n, self.nterm)

    def set_M(self, M):
        M = np.ascontiguousarray(M, dtype=np.complex128)
        self.o.lib.orc_set_M(self.h, _ptr(M))

    def set_L(self, L):
        L = np.ascontiguousarray(L, dtype=np.complex128)
        self.o.lib.orc_set_L(self.h, _ptr(L))


class Reference:
    """The reference's own 3d/ code through oracle/_ref/libref3d.so."""

    def __init__(self) -> None:
        path = os.path.join(HERE, "_ref", "libref3d.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        L = self.lib
        L.ref3d_set.argtypes = [_I, _I, _D]
        L.ref3d_build_tree.restype = _L
        L.ref3d_build_tree.argtypes = [_P, _L]
        L.ref3d_get_cells.argtypes = [_P, _P, _P, _P, _P]
        L.ref3d_make_lists.argtypes = [_P]
        L.ref3d_get_lists.argtypes = [_I, _P, _P]
        L.ref3d_fmm.restype = _I
        L.ref3d_fmm.argtypes = [_P, _L, _P, _P]
        L.ref3d_get_L.argtypes = [_P]
        L.ref3d_fmm_two.argtypes = [_P, _L, _P, _L, _P]
        L.ref3d_get_lists_two.argtypes = [_I, _P, _P]
        L.ref3d_num_cells.restype = _L
        L.ref3d_num_cells.argtypes = [_I]
        L.ref3d_direct.argtypes = [_P, _L, _P, _L]
        L.ref3d_p2p.argtypes = [_P, _L, _P, _L]
        L.ref3d_p2m.argtypes = [_P, _P, _L, _P]
        L.ref3d_m2m.argtypes = [_P, _P, _P, _P]
        L.ref3d_m2l.argtypes = [_P, _P, _P, _P]
        L.ref3d_l2l.argtypes = [_P, _P, _P, _P]
        L.ref3d_l2p.argtypes = [_P, _P, _P, _L]
        L.ref3d_eval_local.argtypes = [_D, _D, _D, _P]
        L.ref3d_eval_multipole.argtypes = [_D, _D, _D, _P, _P]
        self.P = None
        self.ncells = 0

    def set(self, P, ncrit, theta):
        self.P = P
        self.lib.ref3d_set(P, ncrit, theta)

    def build_tree(self, bodies) -> np.ndarray:
        n = self.lib.ref3d_build_tree(_ptr(bodies), len(bodies))
        self.ncells = n
        return self._cells(n)

    def _cells(self, n):
        nchild = np.zeros(n, np.int32)
        nbody = np.zeros(n, np.int32)
        child = np.zeros(n, np.int64)
        body = np.zeros(n, np.int64)
        XR = np.zeros((n, 4))
        self.lib.ref3d_get_cells(_ptr(nchild), _ptr(nbody), _ptr(child), _ptr(body), _ptr(XR))
        out = np.zeros(n, CELL_DTYPE)
        out["nchild"], out["nbody"], out["child"], out["body"] = nchild, nbody, child, body
        out["X"], out["R"] = XR[:, :3], XR[:, 3]
        return out

    def lists(self):
        tot = np.zeros(2, np.int64)
        self.lib.ref3d_make_lists(_ptr(tot))
        res = []
        for w in (0, 1):
            off = np.zeros(self.ncells + 1, np.int64)
            src = np.zeros(max(int(tot[w]), 1), np.int64)
            self.lib.ref3d_get_lists(w, _ptr(off), _ptr(src))
            res.append((off, src[: int(tot[w])]))
        return res

    def fmm(self, bodies, want_M=False):
        times = np.zeros(5)
        n = self.lib.ref3d_fmm(_ptr(bodies), len(bodies), _ptr(times), None)
        self.ncells = n
        return times

    def fmm_two(self, targets, sources):
        """Two-tree evaluation (horizontalPass(icells, jcells)); both arrays
        are permuted in place, the targets receive p and F.  Returns the
        (offsets, sources) lists of M2L and P2P, sources numbered in the
        source tree."""
        tot = np.zeros(2, np.int64)
        self.lib.ref3d_fmm_two(_ptr(targets), len(targets), _ptr(sources), len(sources), _ptr(tot))
        nc = int(self.lib.ref3d_num_cells(0))
        res = []
        for w in (0, 1):
            off = np.zeros(nc + 1, np.int64)
            src = np.zeros(max(int(tot[w]), 1), np.int64)
            self.lib.ref3d_get_lists_two(w, _ptr(off), _ptr(src))
            res.append((off, src[: int(tot[w])]))
        return res

    def L(self):
        nt = self.P * (self.P + 1) // 2
        out = np.zeros(self.ncells * nt, np.complex128)
        self.lib.ref3d_get_L(_ptr(out))
        return out.reshape(self.ncells, nt)

    def direct(self, targets, sources):
        self.lib.ref3d_direct(_ptr(targets), len(targets), _ptr(sources), len(sources))

    def p2p(self, targets, sources):
        self.lib.ref3d_p2p(_ptr(targets), len(targets), _ptr(sources), len(sources))

    def p2m(self, center, bodies, M=None):
        nt = self.P * (self.P + 1) // 2
        M = np.zeros(nt, np.complex128) if M is None else M.copy()
        c = np.ascontiguousarray(center, dtype=np.float64)
        self.lib.ref3d_p2m(_ptr(c), _ptr(bodies), len(bodies), _ptr(M))
        return M

    def m2m(self, cparent, cchild, Mchild, Mparent=None):
        nt = self.P * (self.P + 1) // 2
        Mp = np.zeros(nt, np.complex128) if Mparent is None else Mparent.copy()
        a = np.ascontiguousarray(cparent, dtype=np.float64)
        b = np.ascontiguousarray(cchild, dtype=np.float64)
        Mc = np.ascontiguousarray(Mchild, dtype=np.complex128)
        self.lib.ref3d_m2m(_ptr(a), _ptr(b), _ptr(Mc), _ptr(Mp))
        return Mp

    def m2l(self, Xi, Xj, Mj, Li=None):
        nt = self.P * (self.P + 1) // 2
        Lo = np.zeros(nt, np.complex128) if Li is None else Li.copy()
        a = np.ascontiguousarray(Xi, dtype=np.float64)
        b = np.ascontiguousarray(Xj, dtype=np.float64)
        Mc = np.ascontiguousarray(Mj, dtype=np.complex128)
        self.lib.ref3d_m2l(_ptr(a), _ptr(b), _ptr(Mc), _ptr(Lo))
        return Lo

    def l2l(self, cparent, Lparent, cchild, Lchild=None):
        nt = self.P * (self.P + 1) // 2
        Lc = np.zeros(nt, np.complex128) if Lchild is None else Lchild.copy()
        a = np.ascontiguousarray(cparent, dtype=np.float64)
        b = np.ascontiguousarray(cchild, dtype=np.float64)
        Lp = np.ascontiguousarray(Lparent, dtype=np.complex128)
        self.lib.ref3d_l2l(_ptr(a), _ptr(Lp), _ptr(b), _ptr(Lc))
        return Lc

    def l2p(self, center, L, bodies):
        c = np.ascontiguousarray(center, dtype=np.float64)
        Lc = np.ascontiguousarray(L, dtype=np.complex128)
        self.lib.ref3d_l2p(_ptr(c), _ptr(Lc), _ptr(bodies), len(bodies))

    def eval_local(self, rho, alpha, beta):
        Y = np.zeros(4 * self.P * self.P, np.complex128)
        self.lib.ref3d_eval_local(rho, alpha, beta, _ptr(Y))
        return Y

    def eval_multipole(self, rho, alpha, beta):
        Y = np.zeros(self.P * self.P, np.complex128)
        Yt = np.zeros(self.P * self.P, np.complex128)
        self.lib.ref3d_eval_multipole(rho, alpha, beta, _ptr(Y), _ptr(Yt))
        return Y, Yt


class ReferencePeriodic:
    """The reference's own 3dp/ code through oracle/_ref/libref3dp.so."""

    def __init__(self) -> None:
        path = os.path.join(HERE, "_ref", "libref3dp.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        self.lib.ref3dp_fmm.restype = _I
        self.lib.ref3dp_fmm.argtypes = [_P, _L, _I, _I, _D, _I, _D]
        self.lib.ref3dp_ewald.argtypes = [_P, _L, _I, _I, _D, _D, _D, _D]

    def fmm(self, bodies, P, ncrit, theta, images, cycle):
        return self.lib.ref3dp_fmm(_ptr(bodies), len(bodies), P, ncrit, theta, images, cycle)

    def ewald(self, bodies, ncrit, ksize, alpha, sigma, cutoff, cycle):
        self.lib.ref3dp_ewald(_ptr(bodies), len(bodies), ncrit, ksize, alpha, sigma, cutoff, cycle)


def reference_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libref3d.so"))
