"""Host-side mirror of the reference's evaluation interface over the HIP C ABI.

The reference drives the path with free functions on global state
(3d/fmm.cxx:40-53): ``buildTree(bodies)``, ``upwardPass(cells)``,
``horizontalPass(cells, cells)``, ``downwardPass(cells)``,
``direct(targets, sources)``.  :class:`FMM` keeps that call sequence and
meaning (bodies are permuted in place by ``build_tree``, p and F are
accumulated by the passes), with the globals P / ncrit / theta moved into a
configuration, over ``libexafmm_mi355x.so``.  There is no CPU fallback: if
the library or a device is missing every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .inputs import BODY_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libexafmm_mi355x.so")

CELL_DTYPE = np.dtype([("nchild", "<i4"), ("nbody", "<i4"), ("child", "<i8"), ("body", "<i8"),
                       ("X", "<f8", (3,)), ("R", "<f8")])

M2L_REFERENCE, M2L_FULL = 0, 1
PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_FP32_NEWTON = 0, 1, 2, 3
ORDER_FIXED, ORDER_LEARNED = 0, 1

ERRORS = {-1: "EXAFMM_ERR_ARG", -2: "EXAFMM_ERR_STATE", -3: "EXAFMM_ERR_DEPTH", -4: "EXAFMM_ERR_ALLOC",
          -5: "EXAFMM_ERR_HIP", -6: "EXAFMM_ERR_CAPACITY", -7: "EXAFMM_ERR_NODEVICE",
          -8: "EXAFMM_ERR_UNSUPPORTED"}

EXPORTS = [
    "exafmm_config_default", "exafmm_device_count", "exafmm_last_error", "exafmm_create", "exafmm_destroy",
    "exafmm_synchronize", "exafmm_build_tree", "exafmm_num_cells", "exafmm_get_cells", "exafmm_upward",
    "exafmm_horizontal", "exafmm_downward", "exafmm_evaluate", "exafmm_direct", "exafmm_list_sizes",
    "exafmm_get_lists", "exafmm_get_multipoles", "exafmm_get_locals", "exafmm_stage_times", "exafmm_p2p",
    "exafmm_p2m", "exafmm_m2m", "exafmm_m2l", "exafmm_l2l", "exafmm_l2p", "exafmm_work_counts",
    "exafmm_dipole_correction", "exafmm_ewald", "exafmm_ewald_noalloc", "exafmm_order_model_train", "exafmm_set_order_model",
    "exafmm_get_order_model", "exafmm_order_training_set", "exafmm_mlp_forward", "exafmm_mlp_gradient",
    "exafmm_order_stats", "exafmm_graph_create", "exafmm_graph_create_grouped", "exafmm_graph_launch",
    "exafmm_graph_synchronize", "exafmm_graph_nodes", "exafmm_graph_gemm_shapes", "exafmm_graph_destroy",
    "exafmm_evaluate_batch", "exafmm_evaluate_forest",
]

MLP_PARAMS = 1579  # EXAFMM_MLP_PARAMS
MLP_IN, MLP_OUT, ORDER_MIN = 4, 11, 4
ORDER_FIXED, ORDER_LEARNED = 0, 1


class Problem(C.Structure):
    """exafmm_problem: one problem of exafmm_evaluate_batch."""
    _fields_ = [("bodies", C.c_void_p), ("n", C.c_int64), ("P", C.c_int32), ("on_device", C.c_int32)]


BATCH_REUSE_GEOMETRY = 1
BATCH_FOREST = 2
FOREST_MAX = 512  # problems per forest (8^3 cells above the problems' roots)


class Config(C.Structure):
    _fields_ = [("P", C.c_int32), ("ncrit", C.c_int32), ("theta", C.c_double), ("m2l_truncation", C.c_int32),
                ("m2l_precision", C.c_int32), ("p2p_precision", C.c_int32), ("images", C.c_int32),
                ("cycle", C.c_double), ("order_policy", C.c_int32), ("device", C.c_int32),
                ("autotune", C.c_int32), ("reserved", C.c_int32 * 5)]


class ExaFMMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_P, _I, _L, _D = C.c_void_p, C.c_int, C.c_int64, C.c_double


class Library:
    def __init__(self, path: str = LIBPATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run __graft_entry__.build())")
        self.path = path
        self.lib = C.CDLL(path)
        L = self.lib
        L.exafmm_last_error.restype = C.c_char_p
        L.exafmm_last_error.argtypes = [_P]
        L.exafmm_config_default.argtypes = [C.POINTER(Config)]
        L.exafmm_device_count.argtypes = [C.POINTER(C.c_int)]
        L.exafmm_create.argtypes = [C.POINTER(Config), C.POINTER(_P)]
        L.exafmm_destroy.argtypes = [_P]
        L.exafmm_synchronize.argtypes = [_P]
        L.exafmm_build_tree.argtypes = [_P, _P, _L, _I]
        L.exafmm_num_cells.argtypes = [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
        L.exafmm_get_cells.argtypes = [_P, _P, _L]
        L.exafmm_upward.argtypes = [_P]
        L.exafmm_horizontal.argtypes = [_P, _P]
        L.exafmm_downward.argtypes = [_P]
        L.exafmm_evaluate.argtypes = [_P, _P, _L, _I]
        L.exafmm_direct.argtypes = [_P, _P, _L, _P, _L, _I]
        L.exafmm_list_sizes.argtypes = [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.exafmm_get_lists.argtypes = [_P, _I, _P, _P]
        L.exafmm_get_multipoles.argtypes = [_P, _P]
        L.exafmm_get_locals.argtypes = [_P, _P]
        L.exafmm_stage_times.argtypes = [_P, _P, _I]
        L.exafmm_work_counts.argtypes = [_P, _P, _I]
        L.exafmm_dipole_correction.argtypes = [_P]
        L.exafmm_ewald.argtypes = [_P, _P, _L, _P, _L, _I, _D, _D, _D, _I]
        L.exafmm_ewald_noalloc.argtypes = [_P, _P, _L, _P, _L, _I, _D, _D, _D]
        L.exafmm_p2p.argtypes = [_P, _P, _L, _P, _L]
        L.exafmm_p2m.argtypes = [_P, _P, _P, _L, _P]
        L.exafmm_m2m.argtypes = [_P, _P, _P, _P, _P]
        L.exafmm_m2l.argtypes = [_P, _P, _P, _P, _P]
        L.exafmm_l2l.argtypes = [_P, _P, _P, _P, _P]
        L.exafmm_l2p.argtypes = [_P, _P, _P, _P, _L]
        L.exafmm_order_model_train.argtypes = [_P, _I, _I, _D, C.c_uint64, _P]
        L.exafmm_set_order_model.argtypes = [_P, _P, _I, _D]
        L.exafmm_get_order_model.argtypes = [_P, _P, _I]
        L.exafmm_order_training_set.argtypes = [_P, _P, _P, _P, _L, C.POINTER(C.c_int64)]
        L.exafmm_mlp_forward.argtypes = [_P, _P, _P, _L, _P]
        L.exafmm_mlp_gradient.argtypes = [_P, _P, _P, _P, _L, _P, _P]
        L.exafmm_order_stats.argtypes = [_P, _P, _I]
        L.exafmm_graph_create.argtypes = [_P, _I, _I, C.POINTER(_P)]
        L.exafmm_graph_create_grouped.argtypes = [_P, _I, C.POINTER(_P)]
        L.exafmm_graph_launch.argtypes = [_P]
        L.exafmm_graph_synchronize.argtypes = [_P]
        L.exafmm_graph_nodes.argtypes = [_P, C.POINTER(C.c_int64)]
        L.exafmm_graph_gemm_shapes.argtypes = [_P, _P, _I, C.POINTER(C.c_int)]
        L.exafmm_evaluate_batch.argtypes = [_P, _P, _I, _I]
        L.exafmm_evaluate_forest.argtypes = [_P, _P, _P, _I, _I]
        L.exafmm_graph_destroy.argtypes = [_P]

    def device_count(self) -> int:
        n = C.c_int(0)
        self.lib.exafmm_device_count(C.byref(n))
        return n.value

    def check(self, code: int, ctx=None):
        if code != 0:
            msg = self.lib.exafmm_last_error(ctx)
            raise ExaFMMError(code, msg.decode() if msg else "")


_LIB: Library | None = None


def load() -> Library:
    global _LIB
    if _LIB is None:
        _LIB = Library()
    return _LIB


def _ptr(a):
    if isinstance(a, int):
        return C.c_void_p(a)
    return a.ctypes.data_as(C.c_void_p)


def _bodies(b, n=None, write=True, what="bodies"):
    """(pointer, n) of a body array: a device pointer (int, with n given) or
    a numpy array of the 64-byte Body layout, C-contiguous, writeable when
    the library writes into it, holding at least n bodies.  The library
    copies n * 64 bytes in and out through the pointer, so a strided view, a
    different dtype or a short array would corrupt memory: they raise."""
    if isinstance(b, int):
        if n is None:
            raise ValueError(f"{what}: a device pointer needs an explicit count")
        return C.c_void_p(b), int(n)
    if not isinstance(b, np.ndarray):
        raise TypeError(f"{what}: expected a numpy array of BODY_DTYPE or a device pointer")
    if b.dtype != BODY_DTYPE:
        raise TypeError(f"{what}: dtype {b.dtype} is not the 64-byte Body layout")
    if not b.flags["C_CONTIGUOUS"]:
        raise ValueError(f"{what}: not C-contiguous")
    if write and not b.flags["WRITEABLE"]:
        raise ValueError(f"{what}: read-only array")
    n = len(b) if n is None else int(n)
    if n < 0 or n > len(b):
        raise ValueError(f"{what}: n = {n} exceeds the array's {len(b)} bodies")
    return b.ctypes.data_as(C.c_void_p), n


def default_config(**kw) -> Config:
    cfg = Config()
    load().lib.exafmm_config_default(C.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    return cfg


class FMM:
    """One evaluation context (the reference's globals + Cells, on one device)."""

    def __init__(self, **config):
        self.L = load()
        self.cfg = default_config(**config)
        h = C.c_void_p()
        self.L.check(self.L.lib.exafmm_create(C.byref(self.cfg), C.byref(h)))
        self.h = h
        self.P = self.cfg.P
        self.nterm = self.P * (self.P + 1) // 2
        self._keep = None

    def close(self):
        if getattr(self, "h", None):
            self.L.lib.exafmm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, code):
        self.L.check(code, self.h)

    # ---- passes, host numpy arrays or device pointers (int)
    def build_tree(self, bodies, n=None, on_device=False):
        p, n = _bodies(bodies, n)
        self._keep = bodies
        self._chk(self.L.lib.exafmm_build_tree(self.h, p, n, int(on_device)))

    def upward(self):
        self._chk(self.L.lib.exafmm_upward(self.h))

    def horizontal(self, source: "FMM | None" = None):
        self._chk(self.L.lib.exafmm_horizontal(self.h, (source or self).h))

    def downward(self):
        self._chk(self.L.lib.exafmm_downward(self.h))

    def evaluate(self, bodies, n=None, on_device=False):
        p, n = _bodies(bodies, n)
        self._keep = bodies
        self._chk(self.L.lib.exafmm_evaluate(self.h, p, n, int(on_device)))

    def direct(self, targets, sources, nt=None, ns=None, on_device=False):
        pt, nt = _bodies(targets, nt, what="targets")
        ps, ns = _bodies(sources, ns, write=False, what="sources")
        self._chk(self.L.lib.exafmm_direct(self.h, pt, nt, ps, ns, int(on_device)))

    def evaluate_forest(self, bodies, offsets, on_device=False):
        """Independent problems in one call (exafmm_evaluate_forest): problem
        p is bodies[offsets[p]:offsets[p+1]]; a host array or a device
        pointer (int)."""
        off = np.ascontiguousarray(offsets, np.int64)
        if off.ndim != 1 or len(off) < 2:
            raise ValueError("offsets: need nprob + 1 >= 2 entries")
        p, n = _bodies(bodies, int(off[-1]))
        self._keep = bodies
        self._chk(self.L.lib.exafmm_evaluate_forest(self.h, p, off.ctypes.data_as(C.c_void_p), len(off) - 1,
                                                    int(on_device)))

    def evaluate_batch(self, problems, reuse_geometry=False, forest=False):
        """Evaluate many problems [(bodies, P) or (device pointer, n, P)] with
        this context's configuration (exafmm_evaluate_batch).  Host numpy
        arrays or device pointers (then every problem must be on the device).
        With reuse_geometry, later calls with the same problems replay a HIP
        graph (positions kept, charges read afresh); with forest, the
        problems of each order are evaluated as one forest (trees and lists
        built afresh by every call)."""
        arr = (Problem * len(problems))()
        keep = []
        for k, pr in enumerate(problems):
            if len(pr) == 2:
                b, P = pr
                p, n = _bodies(b, what=f"problem {k}")
                keep.append(b)
                arr[k] = Problem(p.value, n, P, 0)
            else:
                ptr, n, P = pr
                arr[k] = Problem(ptr, n, P, 1)
        self._keep_batch = (arr, keep)
        flags = (BATCH_REUSE_GEOMETRY if reuse_geometry else 0) | (BATCH_FOREST if forest else 0)
        self._chk(self.L.lib.exafmm_evaluate_batch(self.h, C.cast(arr, C.c_void_p), len(problems), flags))

    def dipole_correction(self):
        self._chk(self.L.lib.exafmm_dipole_correction(self.h))

    def ewald(self, targets, sources, ksize, alpha, sigma, cutoff, nt=None, ns=None, on_device=False):
        pt, nt = _bodies(targets, nt, what="targets")
        ps, ns = _bodies(sources, ns, write=False, what="sources")
        self._chk(self.L.lib.exafmm_ewald(self.h, pt, nt, ps, ns, ksize, alpha, sigma, cutoff, int(on_device)))

    def synchronize(self):
        self._chk(self.L.lib.exafmm_synchronize(self.h))

    # ---- inspection
    def num_cells(self):
        n = C.c_int64()
        lv = C.c_int32()
        self._chk(self.L.lib.exafmm_num_cells(self.h, C.byref(n), C.byref(lv)))
        return n.value, lv.value

    def cells(self) -> np.ndarray:
        n, _ = self.num_cells()
        out = np.zeros(n, CELL_DTYPE)
        self._chk(self.L.lib.exafmm_get_cells(self.h, _ptr(out), n))
        return out

    def lists(self, which: int):
        nm = C.c_int64()
        npp = C.c_int64()
        self._chk(self.L.lib.exafmm_list_sizes(self.h, C.byref(nm), C.byref(npp)))
        n, _ = self.num_cells()
        tot = nm.value if which == 0 else npp.value
        off = np.zeros(n + 1, np.int64)
        src = np.zeros(max(tot, 1), np.int64)
        self._chk(self.L.lib.exafmm_get_lists(self.h, which, _ptr(off), _ptr(src)))
        return off, src[:tot]

    def multipoles(self):
        n, _ = self.num_cells()
        out = np.zeros(n * self.nterm, np.complex128)
        self._chk(self.L.lib.exafmm_get_multipoles(self.h, _ptr(out)))
        return out.reshape(n, self.nterm)

    def locals(self):
        n, _ = self.num_cells()
        out = np.zeros(n * self.nterm, np.complex128)
        self._chk(self.L.lib.exafmm_get_locals(self.h, _ptr(out)))
        return out.reshape(n, self.nterm)

    def stage_times(self) -> dict:
        v = np.zeros(8)
        self._chk(self.L.lib.exafmm_stage_times(self.h, _ptr(v), 8))
        keys = ["tree", "lists", "upward", "m2l", "p2p", "downward", "total", "m2l_kernel"]
        return {k: float(x) for k, x in zip(keys, v)}

    def work_counts(self) -> dict:
        v = np.zeros(8, np.int64)
        self._chk(self.L.lib.exafmm_work_counts(self.h, _ptr(v), 8))
        keys = ["m2l_pairs", "p2p_pairs", "p2p_interactions", "p2p_source_bodies", "cells", "leaves", "m2l_keys",
                "depth"]
        return {k: int(x) for k, x in zip(keys, v)}

    # ---- learned order
    def train_order_model(self, nsamples=16384, epochs=400, lr=3e-3, seed=1) -> float:
        loss = np.zeros(1)
        self._chk(self.L.lib.exafmm_order_model_train(self.h, nsamples, epochs, lr, seed, _ptr(loss)))
        return float(loss[0])

    def set_order_model(self, weights, log10_tol):
        w = np.ascontiguousarray(weights, np.float32)
        self._chk(self.L.lib.exafmm_set_order_model(self.h, _ptr(w), w.size, float(log10_tol)))

    def order_model(self) -> np.ndarray:
        w = np.zeros(MLP_PARAMS, np.float32)
        self._chk(self.L.lib.exafmm_get_order_model(self.h, _ptr(w), MLP_PARAMS))
        return w

    def order_training_set(self):
        """(X [n,4], T [n,11], pairs [n,3] = target cell, source cell, image)."""
        n = C.c_int64()
        self._chk(self.L.lib.exafmm_order_training_set(self.h, None, None, None, 0, C.byref(n)))
        X = np.zeros((n.value, MLP_IN), np.float32)
        T = np.zeros((n.value, MLP_OUT), np.float32)
        pairs = np.zeros((n.value, 3), np.int32)
        self._chk(self.L.lib.exafmm_order_training_set(self.h, _ptr(X), _ptr(T), _ptr(pairs), n.value,
                                                       C.byref(n)))
        return X, T, pairs

    def mlp_forward(self, weights, X) -> np.ndarray:
        w = np.ascontiguousarray(weights, np.float32)
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((len(X), MLP_OUT), np.float32)
        self._chk(self.L.lib.exafmm_mlp_forward(self.h, _ptr(w), _ptr(X), len(X), _ptr(Y)))
        return Y

    def mlp_gradient(self, weights, X, T):
        w = np.ascontiguousarray(weights, np.float32)
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(self.L.lib.exafmm_mlp_gradient(self.h, _ptr(w), _ptr(X), _ptr(T), len(X), _ptr(g), _ptr(loss)))
        return g, float(loss[0])

    def order_stats(self) -> np.ndarray:
        h = np.zeros(16, np.int64)
        self._chk(self.L.lib.exafmm_order_stats(self.h, _ptr(h), 16))
        return h

    # ---- kernel.h operator surface (host arrays)
    def p2p(self, targets, sources):
        pt, nt = _bodies(targets, what="targets")
        ps, ns = _bodies(sources, write=False, what="sources")
        self._chk(self.L.lib.exafmm_p2p(self.h, pt, nt, ps, ns))

    def p2m(self, center, bodies, M=None):
        M = np.zeros(self.nterm, np.complex128) if M is None else np.array(M, np.complex128)
        c = np.ascontiguousarray(center, np.float64)
        pb, nb = _bodies(bodies, write=False)
        self._chk(self.L.lib.exafmm_p2m(self.h, _ptr(c), pb, nb, _ptr(M)))
        return M

    def m2m(self, cparent, cchild, Mchild, Mparent=None):
        Mp = np.zeros(self.nterm, np.complex128) if Mparent is None else np.array(Mparent, np.complex128)
        a, b = np.ascontiguousarray(cparent, np.float64), np.ascontiguousarray(cchild, np.float64)
        Mc = np.ascontiguousarray(Mchild, np.complex128)
        self._chk(self.L.lib.exafmm_m2m(self.h, _ptr(a), _ptr(b), _ptr(Mc), _ptr(Mp)))
        return Mp

    def m2l(self, Xi, Xj, Mj, Li=None):
        Lo = np.zeros(self.nterm, np.complex128) if Li is None else np.array(Li, np.complex128)
        a, b = np.ascontiguousarray(Xi, np.float64), np.ascontiguousarray(Xj, np.float64)
        Mc = np.ascontiguousarray(Mj, np.complex128)
        self._chk(self.L.lib.exafmm_m2l(self.h, _ptr(a), _ptr(b), _ptr(Mc), _ptr(Lo)))
        return Lo

    def l2l(self, cparent, Lparent, cchild, Lchild=None):
        Lc = np.zeros(self.nterm, np.complex128) if Lchild is None else np.array(Lchild, np.complex128)
        a, b = np.ascontiguousarray(cparent, np.float64), np.ascontiguousarray(cchild, np.float64)
        Lp = np.ascontiguousarray(Lparent, np.complex128)
        self._chk(self.L.lib.exafmm_l2l(self.h, _ptr(a), _ptr(Lp), _ptr(b), _ptr(Lc)))
        return Lc

    def l2p(self, center, L, bodies):
        c = np.ascontiguousarray(center, np.float64)
        Lc = np.ascontiguousarray(L, np.complex128)
        pb, nb = _bodies(bodies)
        self._chk(self.L.lib.exafmm_l2p(self.h, _ptr(c), _ptr(Lc), pb, nb))


class Graph:
    """A hipGraph that re-evaluates many evaluated contexts (device-resident
    bodies) with one launch; see exafmm_graph_create.  The contexts must
    outlive the graph."""

    def __init__(self, contexts, streams: int = 8, grouped: bool = False):
        """grouped: one launch per stage over all problems
        (exafmm_graph_create_grouped) instead of per-problem launches on
        `streams` parallel branches."""
        self.L = load()
        self.contexts = list(contexts)
        arr = (_P * len(self.contexts))(*[c.h for c in self.contexts])
        h = _P()
        if grouped:
            rc = self.L.lib.exafmm_graph_create_grouped(arr, len(self.contexts), C.byref(h))
        else:
            rc = self.L.lib.exafmm_graph_create(arr, len(self.contexts), streams, C.byref(h))
        self.L.check(rc)
        self.h = h

    def launch(self):
        self.L.check(self.L.lib.exafmm_graph_launch(self.h))

    def synchronize(self):
        self.L.check(self.L.lib.exafmm_graph_synchronize(self.h))

    def nodes(self) -> int:
        n = C.c_int64(0)
        self.L.check(self.L.lib.exafmm_graph_nodes(self.h, C.byref(n)))
        return n.value

    def gemm_shapes(self) -> list:
        """[(P, rows per workgroup, keys per LDS round)] of a grouped graph."""
        buf = np.zeros(48, np.int32)
        n = C.c_int(0)
        self.L.check(self.L.lib.exafmm_graph_gemm_shapes(self.h, _ptr(buf), 16, C.byref(n)))
        return [tuple(int(x) for x in buf[3 * k:3 * k + 3]) for k in range(min(n.value, 16))]

    def close(self):
        if getattr(self, "h", None):
            self.L.lib.exafmm_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
