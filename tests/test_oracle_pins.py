"""Pins of the FP64 oracle beyond the reference's reference-truncation
fixtures: the full-truncation M2L mode (the yardstick of every fast-path GPU
test), the exact shift identities, the P2P formula, and the oracle at the
config-2 size.

What pins each part:
  full M2L, config-1 set      errors against the FP64 direct sum equal the
                              values measured from the reference's own
                              kernels with evalLocal filled to degree 2P-1
                              (4.512e-8 / 1.177e-6), to 1%
  full M2L, KAT geometry      the same modified reference chain of
                              3d/kernel.cxx: 1.4815e-3 / 4.6464e-5 /
                              1.6570e-6 at P = 10 / 20 / 30, to 1%
  full M2L, one pair          P2M -> M2L -> L2P converges to the direct
                              sum geometrically in P down to rounding (a
                              dropped term or a wrong sign leaves a floor)
  full periodic evaluation    closer to the reference's P = 18 periodic run
                              than the reference truncation at P = 10
  M2M, L2L                    exact re-expansions: M2M(P2M about the child)
                              = P2M about the parent; L2P of L2L(L) = L2P
                              of L, both to rounding
  P2P                         an independent numpy evaluation of
                              sum q/r and -sum q dX/r^3 (r = 0 skipped) and
                              Newton's third law
  tree, lists at N = 2^20     digest and counts of the reference's tree and
                              lists (tests/golden/config2_n2p20.npz)
"""
import hashlib

import numpy as np
import pytest

from conftest import golden
import oracle as O
from minimal_amd.inputs import rel_l2, sample_targets, uniform_cube

rng = np.random.default_rng(7)


def _bodies(n, c, r):
    b = np.zeros(n, O.BODY_DTYPE)
    b["X"] = np.asarray(c) + rng.uniform(-r, r, (n, 3))
    b["q"] = rng.uniform(-1, 1, n)
    return b


def _direct_errors(orc, b, ntargets):
    idx = sample_targets(len(b), ntargets)
    tg = b[idx].copy()
    tg["p"] = 0
    tg["F"] = 0
    orc.p2p(tg, b)
    return rel_l2(b["p"][idx], tg["p"]), rel_l2(b["F"][idx], tg["F"])


@pytest.mark.parametrize("full,expect", [(True, (4.512e-8, 1.177e-6)), (False, (2.357e-7, 1.126e-5))])
def test_config1_errors(orc, full, expect):
    """Config 1 (N = 1e4 in [0,1)^3, q = U[0,1)/N, P = 10, theta 0.4,
    ncrit 64): relative L2 of p and F against the FP64 direct sum on 100
    stride targets, normalised by the direct values."""
    b = golden("config1_n10000.npz")["input"].copy()
    orc.fmm(b, 10, 64, 0.4, full=full)
    ep, eF = _direct_errors(orc, b, 100)
    assert ep == pytest.approx(expect[0], rel=0.01)
    assert eF == pytest.approx(expect[1], rel=0.01)


@pytest.mark.parametrize("P,expect", [(10, 1.4815e-3), (20, 4.6464e-5), (30, 1.6570e-6)])
def test_kat_full(orc, P, expect):
    """3d/kernel.cxx chain with every M2L term j+n <= 2P-2 kept."""
    e = orc.kat(P, full=True)
    assert e[0] == pytest.approx(expect, rel=0.01)


def test_full_m2l_converges_to_direct(orc):
    Xj = np.array([0.1, 0.2, 0.3])
    Xi = np.array([1.3, 0.9, 1.1])
    src = _bodies(40, Xj, 0.15)
    tgt = _bodies(20, Xi, 0.15)
    ref = tgt.copy()
    ref["p"] = 0
    ref["F"] = 0
    orc.p2p(ref, src)
    errs = {}
    for full in (True, False):
        for P in range(4, 17, 2):
            L = orc.m2l(P, full, Xi, Xj, orc.p2m(P, Xj, src))
            t = tgt.copy()
            t["p"] = 0
            t["F"] = 0
            orc.l2p(P, Xi, L, t)
            errs[full, P] = (rel_l2(t["p"], ref["p"]), rel_l2(t["F"], ref["F"]))
    for P in range(6, 17, 2):
        # geometric convergence: (sqrt(3)*0.3/|d|)^2 = 0.105 per two orders
        assert errs[True, P][0] < 0.2 * errs[True, P - 2][0]
        assert errs[True, P][1] < 0.2 * errs[True, P - 2][1]
        # the full sum is the more accurate one at every order
        assert errs[True, P][0] < errs[False, P][0]
    assert errs[True, 16][0] < 1e-13 and errs[True, 16][1] < 1e-11


def test_periodic_full_mode(orc):
    """3dp driver set (N = 1000, P = 10, images 4, cycle 2 pi): the
    full-truncation periodic evaluation against the reference's own run at
    P = 18."""
    g = golden("periodic_n1000.npz")
    hi = golden("periodic_n1000_p18.npz")
    res = {}
    for P, full in ((10, True), (10, False), (14, True)):
        b = g["input"].copy()
        orc.fmm_periodic(b, P, 64, 0.4, 4, 2 * np.pi, full=full)
        res[P, full] = (rel_l2(b["p"], hi["p"]), rel_l2(b["F"], hi["F"]))
    assert res[10, True][0] < 3e-5 and res[10, True][1] < 2e-5
    assert res[10, True][0] < 0.25 * res[10, False][0]
    assert res[10, True][1] < 0.25 * res[10, False][1]
    assert res[14, True][0] < res[10, True][0] and res[14, True][1] < res[10, True][1]


@pytest.mark.parametrize("P", [4, 10, 14])
def test_m2m_exact(orc, P):
    cc = np.array([0.3, 0.3, 0.3])
    cp = np.array([0.2, 0.4, 0.25])
    s = _bodies(30, cc, 0.1)
    M1 = orc.m2m(P, cp, cc, orc.p2m(P, cc, s))
    M2 = orc.p2m(P, cp, s)
    assert np.max(np.abs(M1 - M2)) <= 1e-13 * np.max(np.abs(M2))


@pytest.mark.parametrize("P", [4, 10, 14])
def test_l2l_exact(orc, P):
    nt = P * (P + 1) // 2
    L = rng.normal(size=nt) + 1j * rng.normal(size=nt)
    m0 = [n * (n + 1) // 2 for n in range(P)]
    L[m0] = L[m0].real
    cp = np.array([0.2, 0.4, 0.25])
    ch = cp + np.array([0.05, -0.05, 0.05])
    Lc = orc.l2l(P, cp, L, ch)
    t = _bodies(20, ch, 0.05)
    t["p"] = 0
    t["F"] = 0
    a, b = t.copy(), t.copy()
    orc.l2p(P, cp, L, a)
    orc.l2p(P, ch, Lc, b)
    assert np.max(np.abs(a["p"] - b["p"])) <= 1e-13 * np.max(np.abs(a["p"]))
    assert np.max(np.abs(a["F"] - b["F"])) <= 1e-13 * np.max(np.abs(a["F"]))


def test_p2p_formula_and_third_law(orc):
    src = _bodies(50, (0.5, 0.5, 0.5), 0.5)
    tgt = src[:30].copy()
    tgt["p"] = 0
    tgt["F"] = 0
    orc.p2p(tgt, src)
    d = tgt["X"][:, None, :] - src["X"][None, :, :]
    r2 = np.sum(d * d, axis=2)
    with np.errstate(divide="ignore"):
        rinv = np.where(r2 > 0, 1 / np.sqrt(r2), 0.0)
    p = np.sum(src["q"][None, :] * rinv, axis=1)
    F = -np.sum((src["q"][None, :] * rinv ** 3)[:, :, None] * d, axis=1)
    assert np.allclose(tgt["p"], p, rtol=1e-13, atol=0)
    assert np.allclose(tgt["F"], F, rtol=1e-12, atol=1e-12 * np.max(np.abs(F)))
    # A on B and B on A: sum q p equal, sum q F opposite.
    A, B = _bodies(25, (0.2, 0.2, 0.2), 0.1), _bodies(35, (0.6, 0.4, 0.5), 0.1)
    for x in (A, B):
        x["p"] = 0
        x["F"] = 0
    orc.p2p(A, B)
    orc.p2p(B, A)
    assert np.sum(A["q"] * A["p"]) == pytest.approx(np.sum(B["q"] * B["p"]), rel=1e-13)
    fa = np.sum(A["q"][:, None] * A["F"], axis=0)
    fb = np.sum(B["q"][:, None] * B["F"], axis=0)
    assert np.max(np.abs(fa + fb)) <= 1e-12 * np.max(np.abs(fa))


def test_oracle_tree_and_lists_at_config2(orc):
    """The oracle's tree and lists at N = 2^20 against the reference's
    digest and counts."""
    g = golden("config2_n2p20.npz")
    b = uniform_cube(int(g["n"]), 0, "sym")
    s = orc.session(10, 64, 0.4)
    assert s.build_tree(b) == 0
    c = s.cells()
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(b["X"], "<f8").tobytes())
    h.update(np.ascontiguousarray(b["q"], "<f8").tobytes())
    for f in ("nchild", "nbody", "child", "body", "X", "R"):
        h.update(np.ascontiguousarray(c[f]).tobytes())
    assert h.hexdigest() == str(g["digest"])
    s.make_lists()
    mo, ms = s.lists(0)
    po, ps = s.lists(1)
    assert len(ms) == int(g["m2l_pairs"]) and len(ps) == int(g["p2p_pairs"])
    nb = c["nbody"].astype(np.int64)
    tgt = np.repeat(np.arange(len(c)), np.diff(po))
    assert int(np.sum(nb[tgt] * nb[ps])) == int(g["p2p_interactions"])
