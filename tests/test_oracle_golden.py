"""The FP64 oracle against fixtures written by the reference's own code.

Bit-exact: tree, body permutation, interaction lists, final p/F.
Operators: to rounding (the oracle and the reference evaluate the same
expressions; equal bits are expected but libm across machines may differ
in the last place, so a 1e-14 relative bound is used).
"""
import numpy as np
import pytest

from conftest import golden
import oracle as O


def _cells_match(s, g):
    c = s.cells()
    assert len(c) == len(g["cells_nbody"])
    assert np.array_equal(c["nchild"], g["cells_nchild"])
    assert np.array_equal(c["nbody"], g["cells_nbody"])
    assert np.array_equal(c["child"], g["cells_child"])
    assert np.array_equal(c["body"], g["cells_body"])
    assert np.array_equal(c["X"], g["cells_X"])
    assert np.array_equal(c["R"], g["cells_R"])


@pytest.mark.parametrize("name", ["config1_n10000.npz", "uniform_n1000_p4.npz", "plummer_n4096_p6.npz"])
def test_tree_lists_fmm_bitexact(orc, name):
    g = golden(name)
    P, ncrit, theta = int(g["P"]), int(g["ncrit"]), float(g["theta"])
    b = g["input"].copy()
    s = orc.session(P, ncrit, theta)
    assert s.build_tree(b) == 0
    xq = np.stack([b["X"][:, 0], b["X"][:, 1], b["X"][:, 2], b["q"]], 1)
    assert np.array_equal(xq, g["sorted_xq"])
    _cells_match(s, g)
    s.make_lists()
    mo, ms = s.lists(0)
    po, ps = s.lists(1)
    assert np.array_equal(mo, g["m2l_off"]) and np.array_equal(ms, g["m2l_src"])
    assert np.array_equal(po, g["p2p_off"]) and np.array_equal(ps, g["p2p_src"])
    b2 = g["input"].copy()
    orc.fmm(b2, P, ncrit, theta, full=False)
    assert np.array_equal(b2["p"], g["p"])
    assert np.array_equal(b2["F"], g["F"])


def test_driver_errors(orc):
    """test/test_fmm.cxx thresholds and the values the reference prints."""
    g = golden("driver_n10000.npz")
    for P, bound in ((10, 1e-3), (20, 1e-6)):
        b = g["input"].copy()
        orc.fmm(b, P, 64, 0.4)
        tg = b[np.arange(10) * 1000].copy()
        assert np.array_equal(tg["p"], g[f"P{P}_fmm_p"])
        assert np.array_equal(tg["F"], g[f"P{P}_fmm_F"])
        fp = tg["p"].copy()
        tg["p"] = 0
        tg["F"] = 0
        orc.p2p(tg, b)
        assert np.array_equal(tg["p"], g[f"P{P}_dir_p"])
        perr = np.sqrt(np.sum((tg["p"] - fp) ** 2) / np.sum(fp ** 2))
        assert perr == pytest.approx(float(g[f"P{P}_perr"]), rel=1e-12)
        assert perr < bound


@pytest.mark.parametrize("P,bound,expect", [(10, 1e-3, 9.77517e-4), (20, 1e-6, 9.53675e-7), (30, 1e-9, 9.31325e-10)])
def test_kernel_kat(orc, P, bound, expect):
    """3d/kernel.cxx / test/test_kernel.cxx: chain error vs P2P."""
    e = orc.kat(P, full=False)
    assert e[0] < bound
    assert e[0] == pytest.approx(expect, rel=1e-5)


@pytest.mark.parametrize("P", [4, 10, 14])
def test_operators(orc, P):
    g = golden(f"ops_P{P}.npz")
    rt = 1e-14

    def close(a, b):
        a, b = np.asarray(a), np.asarray(b)
        assert np.max(np.abs(a - b)) <= rt * max(np.max(np.abs(b)), 1e-300)

    close(orc.p2m(P, g["p2m_center"], g["p2m_bodies"]), g["p2m_M"])
    close(orc.m2m(P, g["m2m_parent"], g["m2m_child"], g["m2m_Mchild"], g["m2m_M0"]), g["m2m_M"])
    close(orc.m2l(P, False, g["m2l_Xi"], g["m2l_Xj"], g["m2m_M"], g["m2l_L0"]), g["m2l_L"])
    close(orc.l2l(P, g["m2l_Xi"], g["m2l_L"], g["l2l_child"], g["l2l_Lc0"]), g["l2l_L"])
    tb = g["l2p_bodies_in"].copy()
    orc.l2p(P, g["l2l_child"], g["l2l_L"], tb)
    close(tb["p"], g["l2p_bodies_out"]["p"])
    close(tb["F"], g["l2p_bodies_out"]["F"])
    tg = g["p2p_tgt_in"].copy()
    orc.p2p(tg, g["p2p_src"])
    assert np.array_equal(tg["p"], g["p2p_tgt_out"]["p"])
    assert np.array_equal(tg["F"], g["p2p_tgt_out"]["F"])
    close(orc.eval_local(P, P, *g["harm_rab"]), g["eval_local"])
    Y, Yt = orc.eval_multipole(P, 0.6, 1.1, 0.4)
    close(Y, g["eval_multipole"])
    close(Yt, g["eval_multipole_theta"])


def test_periodic_fmm_and_ewald(orc):
    """3dp driver set: periodic FMM bit-exact; Ewald within its own accuracy."""
    g = golden("periodic_n1000.npz")
    b = g["input"].copy()
    orc.fmm_periodic(b, 10, 64, 0.4, 4, 2 * np.pi)
    ref = g["fmm_sorted"]
    assert np.array_equal(b["X"], ref["X"])
    assert np.array_equal(b["p"], ref["p"])
    assert np.array_equal(b["F"], ref["F"])
    ew = g["ewald"]
    t = b.copy()
    t["p"] = 0
    t["F"] = 0
    orc.ewald(t, b, 11, 11 / (2 * np.pi), 0.25 / np.pi, np.pi, 2 * np.pi)
    # The reference picks one periodic image per leaf pair from the leaf
    # centres (3dp/ewald.h:104-114); with leaves a quarter of the cycle wide
    # it drops some body pairs whose nearest image is inside the cutoff.
    # The oracle uses each pair's nearest image, so the two agree only to
    # the size of those dropped terms.
    assert np.max(np.abs(t["p"] - ew["p"])) <= 1e-5 * np.max(np.abs(ew["p"]))
    assert np.max(np.abs(t["F"] - ew["F"])) <= 1e-4 * np.max(np.abs(ew["F"]))


def test_ewald_alpha_independence(orc):
    """The Ewald split is exact: the sum must not depend on alpha once the
    real and wave series have converged (a dropped term, a wrong factor or
    sign in either part breaks this)."""
    g = golden("periodic_n1000.npz")
    b = g["fmm_sorted"][:200].copy()
    b["q"] -= b["q"].mean()
    cyc = 2 * np.pi
    res = []
    for alpha, ks in ((1.2, 14), (1.6, 18)):
        t = b.copy()
        t["p"] = 0
        t["F"] = 0
        orc.ewald(t, b, ks, alpha, 0.25 / np.pi, cyc / 2, cyc)
        res.append(t)
    assert np.max(np.abs(res[0]["p"] - res[1]["p"])) <= 1e-6 * np.max(np.abs(res[1]["p"]))
    assert np.max(np.abs(res[0]["F"] - res[1]["F"])) <= 1e-5 * np.max(np.abs(res[1]["F"]))
