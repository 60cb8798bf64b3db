"""horizontalPass(icells, jcells) with two distinct trees (3d/traverse_eager.h:
147-172, list form 3d/traverse_lazy.h:59-62) through the C ABI, against the
reference's own two-tree run (tests/golden/two_trees_n3000_n5000.npz: 3,000
targets in [0,1)^3, 5,000 sources in [-0.1,1.1)^3, P = 8, ncrit 32).

Two trees with different root boxes share no lattice of cell offsets, so
their M2L pairs run through the per-pair FP64 M2L kernel whatever M2L
precision the context asks for; P2P keeps the context's precision.

Bars:
  both trees, permutations         bit-exact
  M2L / P2P lists                  equal as sets per target cell
  FP64 mode                        |diff| <= 1e-11 * max |ref| (reference truncation)
  BF16x3 M2L asked, FP32 P2P       relative L2 <= 2e-6 (the FP32 P2P bar: the M2L is FP64)
"""
import numpy as np
import pytest

from conftest import golden
from minimal_amd import fmm
from minimal_amd.inputs import rel_l2

pytestmark = pytest.mark.gpu


def _pair(g, **kw):
    mk = lambda: fmm.FMM(P=int(g["P"]), ncrit=int(g["ncrit"]), theta=float(g["theta"]),
                         m2l_truncation=fmm.M2L_REFERENCE, **kw)
    ft, fs = mk(), mk()
    t, s = g["targets"].copy(), g["sources"].copy()
    ft.build_tree(t)
    fs.build_tree(s)
    fs.upward()
    ft.upward()
    ft.horizontal(fs)
    return ft, fs, t, s


def test_two_trees_tree_and_lists(gpu):
    g = golden("two_trees_n3000_n5000.npz")
    ft, fs, t, s = _pair(g, m2l_precision=fmm.PREC_FP64, p2p_precision=fmm.PREC_FP64)
    assert np.array_equal(t["X"], g["sorted_targets_X"])
    assert np.array_equal(s["X"], g["sorted_sources_X"])
    for f, name in ((ft, "tcells"), (fs, "scells")):
        c = f.cells()
        for k in ("nchild", "nbody", "child", "body", "X", "R"):
            assert np.array_equal(c[k], g[f"{name}_{k}"]), (name, k)
    for which, (ok, sk) in ((0, ("m2l_off", "m2l_src")), (1, ("p2p_off", "p2p_src"))):
        off, src = ft.lists(which)
        go, gs = g[ok], g[sk]
        assert len(src) == len(gs)
        for c in range(len(off) - 1):
            assert np.array_equal(np.sort(src[off[c]:off[c + 1]]), np.sort(gs[go[c]:go[c + 1]])), (which, c)


def test_two_trees_fp64_matches_reference(gpu):
    g = golden("two_trees_n3000_n5000.npz")
    ft, fs, t, s = _pair(g, m2l_precision=fmm.PREC_FP64, p2p_precision=fmm.PREC_FP64)
    ft.downward()
    assert np.max(np.abs(t["p"] - g["p"])) <= 1e-11 * np.max(np.abs(g["p"]))
    assert np.max(np.abs(t["F"] - g["F"])) <= 1e-11 * np.max(np.abs(g["F"]))


def test_two_trees_fast_context(gpu):
    g = golden("two_trees_n3000_n5000.npz")
    ft, fs, t, s = _pair(g, m2l_precision=fmm.PREC_BF16X3, p2p_precision=fmm.PREC_FP32)
    ft.downward()
    assert rel_l2(t["p"], g["p"]) <= 2e-6
    assert rel_l2(t["F"], g["F"]) <= 2e-6
