"""Writes the golden fixtures of this directory from the reference's own code.

Runs only where the reference sources exist (they are compiled into
oracle/_ref/ by oracle/Makefile); the fixtures it writes are committed so the
tests can run anywhere.  Every value in a fixture is an output of the
reference's buildTree / getList / kernels / passes (3d/ and 3dp/ headers).

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py config2    # one section by name
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from minimal_amd import inputs  # noqa: E402


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"wrote {name}: {os.path.getsize(path)} bytes")


def tree_and_lists(ref, bodies, P, ncrit, theta, with_fmm=True, with_lists=True):
    ref.set(P, ncrit, theta)
    b = bodies.copy()
    cells = ref.build_tree(b)
    out = dict(input=bodies, sorted_xq=np.stack([b["X"][:, 0], b["X"][:, 1], b["X"][:, 2], b["q"]], 1),
               cells_nchild=cells["nchild"], cells_nbody=cells["nbody"], cells_child=cells["child"],
               cells_body=cells["body"], cells_X=cells["X"], cells_R=cells["R"])
    if with_lists:
        (mo, ms), (po, ps) = ref.lists()
        out.update(m2l_off=mo.astype(np.int32), m2l_src=ms.astype(np.int32),
                   p2p_off=po.astype(np.int32), p2p_src=ps.astype(np.int32))
    if with_fmm:
        b2 = bodies.copy()
        ref.fmm(b2)
        out.update(p=b2["p"], F=b2["F"])
    return out


def tree_digest(bodies: np.ndarray, cells: np.ndarray) -> str:
    """SHA-256 over the tree-ordered X and q (little-endian FP64) and the
    cell table (nchild, nbody, child, body, X, R per cell, in the
    reference's numbering); the same digest is computed from the device's
    outputs by the tests."""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(bodies["X"], "<f8").tobytes())
    h.update(np.ascontiguousarray(bodies["q"], "<f8").tobytes())
    for f in ("nchild", "nbody", "child", "body", "X", "R"):
        h.update(np.ascontiguousarray(cells[f]).tobytes())
    return h.hexdigest()


def config2(ref):
    """Config 2 at its stated size: N = 2^20 uniform in [0,1)^3, q = U[-1,1)
    (drand48 seed 0, the reference's draw order), P = 10, theta = 0.4,
    ncrit = 64.  Stores the digest of the reference's tree, its list sizes
    and interaction counts, and the reference's FP64 direct sum on 1,000
    stride-sampled targets of the tree-ordered bodies (3d/fmm.cxx:57-69)."""
    n = 1 << 20
    b = inputs.uniform_cube(n, 0, "sym")
    ref.set(10, 64, 0.4)
    cells = ref.build_tree(b)
    (mo, ms), (po, ps) = ref.lists()
    nb = cells["nbody"].astype(np.int64)
    tgt = np.repeat(np.arange(len(cells)), np.diff(po))
    inter = int(np.sum(nb[tgt] * nb[ps]))
    idx = inputs.sample_targets(n, 1000)
    tg = b[idx].copy()
    tg["p"] = 0
    tg["F"] = 0
    ref.direct(tg, b)
    save("config2_n2p20.npz", n=n, P=10, ncrit=64, theta=0.4, digest=np.array(tree_digest(b, cells)),
         ncells=len(cells), nleaves=int(np.sum(cells["nchild"] == 0)), m2l_pairs=len(ms), p2p_pairs=len(ps),
         p2p_interactions=inter, p2p_source_bodies=int(np.sum(nb[ps])), targets=idx,
         direct_p=tg["p"], direct_F=tg["F"])


def periodic_high_order(refp):
    """The 3dp driver's set evaluated by the reference at P = 18 (other
    parameters as shipped): the near-converged periodic sum that pins the
    full-truncation periodic evaluation."""
    drv = inputs.reference_driver_bodies(1000, 0)
    b2 = drv.copy()
    refp.fmm(b2, 18, 64, 0.4, 4, 2 * np.pi)
    save("periodic_n1000_p18.npz", p=b2["p"], F=b2["F"])


def config5_bodies(n: int) -> np.ndarray:
    """Config 5's bodies: uniform in the unit periodic cube (drand48 seed 0,
    the reference's draw order), charges +1 / -1 alternating by index."""
    b = inputs.uniform_cube(n, 0, "sym")
    b["q"] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return b


def config5(refp, n=4096):
    """Config 5's parameters at a reduced N: P = 10, theta = 0.4, ncrit =
    64, images = 3, cycle = 1; the reference's periodic FMM (as written,
    3dp/fmm.cxx:46-61) and its Ewald sum with alpha = 10, ksize = 11,
    sigma = 0.25/pi, cutoff = 0.5 (3dp/fmm.cxx:77-94) on the tree-ordered
    bodies."""
    b = config5_bodies(n)
    refp.fmm(b, 10, 64, 0.4, 3, 1.0)
    ew = b.copy()
    refp.ewald(ew, 64, 11, 10.0, 0.25 / np.pi, 0.5, 1.0)
    save(f"config5_n{n}.npz", n=n, P=10, ncrit=64, theta=0.4, images=3, cycle=1.0, ksize=11, alpha=10.0,
         sigma=0.25 / np.pi, cutoff=0.5, fmm_p=b["p"], fmm_F=b["F"], ewald_p=ew["p"], ewald_F=ew["F"])


def two_trees(ref):
    """horizontalPass(icells, jcells) with two distinct trees: 3,000 target
    bodies uniform in [0,1)^3 and 5,000 source bodies uniform in
    [-0.1,1.1)^3 (different root boxes), q = U[-1,1), P = 8, theta 0.4,
    ncrit 32.  The reference's two cell tables, both permutations, the
    two-tree lists and the targets' p and F."""
    P, ncrit, theta = 8, 32, 0.4
    t = inputs.uniform_cube(3000, 11, "sym")
    s = inputs.uniform_cube(5000, 12, "sym")
    s["X"] = s["X"] * 1.2 - 0.1
    ref.set(P, ncrit, theta)
    tcells = ref.build_tree(t.copy())
    scells = ref.build_tree(s.copy())
    t2, s2 = t.copy(), s.copy()
    (mo, ms), (po, ps) = ref.fmm_two(t2, s2)
    out = dict(targets=t, sources=s, P=P, ncrit=ncrit, theta=theta, sorted_targets_X=t2["X"],
               sorted_sources_X=s2["X"], p=t2["p"], F=t2["F"],
               m2l_off=mo.astype(np.int32), m2l_src=ms.astype(np.int32),
               p2p_off=po.astype(np.int32), p2p_src=ps.astype(np.int32))
    for name, c in (("tcells", tcells), ("scells", scells)):
        for f in ("nchild", "nbody", "child", "body", "X", "R"):
            out[f"{name}_{f}"] = c[f]
    save("two_trees_n3000_n5000.npz", **out)


def main():
    oracle.build()
    ref = oracle.Reference()
    only = sys.argv[1:]
    if only:
        if "config2" in only:
            config2(ref)
        if "periodic_p18" in only:
            periodic_high_order(oracle.ReferencePeriodic())
        if "two_trees" in only:
            two_trees(ref)
        if "config5" in only:
            config5(oracle.ReferencePeriodic())
        return
    rng = np.random.default_rng(20261202)

    # Config-1 body set (N=1e4, [0,1)^3, q=U[0,1)/N), P=10, theta=0.4, ncrit=64.
    b = inputs.uniform_cube(10000, 0, "unit_over_n")
    save("config1_n10000.npz", P=10, ncrit=64, theta=0.4, **tree_and_lists(ref, b, 10, 64, 0.4))

    # Config-2 distribution at a small N (q=U[-1,1)), P=4 and an adaptive
    # Plummer cluster with background (deep, unbalanced tree).
    b = inputs.uniform_cube(1000, 7, "sym")
    save("uniform_n1000_p4.npz", P=4, ncrit=16, theta=0.4, **tree_and_lists(ref, b, 4, 16, 0.4))
    n = 4096
    pos = np.concatenate([inputs.plummer(n - 410, 0.05, 0.5, rng), rng.random((410, 3))])
    pb = np.zeros(n, inputs.BODY_DTYPE)
    pb["X"] = pos
    pb["q"] = rng.uniform(-1, 1, n)
    save("plummer_n4096_p6.npz", P=6, ncrit=32, theta=0.4, **tree_and_lists(ref, pb, 6, 32, 0.4))

    # Driver body set of 3d/fmm.cxx and test/test_fmm.h (X in [-pi,pi)^3,
    # neutral q), N=1e4, ncrit=64: p, F of the FMM at P=10 and P=20 and the
    # FP64 direct sum at the 10 stride-sampled targets (test/test_fmm.h:73-104).
    drv = inputs.reference_driver_bodies(10000, 0)
    res = {}
    for P in (10, 20):
        ref.set(P, 64, 0.4)
        b2 = drv.copy()
        ref.fmm(b2)
        stride = len(b2) // 10
        tg = b2[np.arange(10) * stride].copy()
        fmm_p, fmm_F = tg["p"].copy(), tg["F"].copy()
        tg["p"] = 0
        tg["F"] = 0
        ref.direct(tg, b2)
        perr = np.sqrt(np.sum((tg["p"] - fmm_p) ** 2) / np.sum(fmm_p ** 2))
        res[f"P{P}_fmm_p"], res[f"P{P}_fmm_F"] = fmm_p, fmm_F
        res[f"P{P}_dir_p"], res[f"P{P}_dir_F"] = tg["p"].copy(), tg["F"].copy()
        res[f"P{P}_perr"] = perr
        print(f"test_fmm({P}) = {perr:.6e}")
    save("driver_n10000.npz", input=drv, **res)

    # Operator vectors (one reference kernel call each) at P = 4, 10, 14.
    for P in (4, 10, 14):
        ref.set(P, 64, 0.4)
        nt = P * (P + 1) // 2
        o = {}
        cb = np.zeros(24, inputs.BODY_DTYPE)
        c = np.array([0.3, -0.2, 0.1])
        cb["X"] = c + rng.uniform(-0.25, 0.25, (24, 3))
        cb["q"] = rng.uniform(-1, 1, 24)
        o["p2m_center"], o["p2m_bodies"] = c, cb
        o["p2m_M"] = ref.p2m(c, cb)
        cp = c + np.array([0.25, -0.25, 0.25])
        Mc = o["p2m_M"]
        M0 = rng.normal(size=nt) + 1j * rng.normal(size=nt)
        o["m2m_parent"], o["m2m_child"], o["m2m_Mchild"], o["m2m_M0"] = cp, c, Mc, M0
        o["m2m_M"] = ref.m2m(cp, c, Mc, M0)
        xi = np.array([2.1, 0.7, -1.3])
        o["m2l_Xi"], o["m2l_Xj"] = xi, cp
        L0 = rng.normal(size=nt) + 1j * rng.normal(size=nt)
        o["m2l_L0"] = L0
        o["m2l_L"] = ref.m2l(xi, cp, o["m2m_M"], L0)
        xc = xi + np.array([-0.125, 0.125, 0.125])
        Lc0 = rng.normal(size=nt) + 1j * rng.normal(size=nt)
        o["l2l_child"], o["l2l_Lc0"] = xc, Lc0
        o["l2l_L"] = ref.l2l(xi, o["m2l_L"], xc, Lc0)
        tb = np.zeros(20, inputs.BODY_DTYPE)
        tb["X"] = xc + rng.uniform(-0.12, 0.12, (20, 3))
        tb["q"] = 1.0
        tb["p"] = rng.normal(size=20)
        tb["F"] = rng.normal(size=(20, 3))
        o["l2p_bodies_in"] = tb.copy()
        ref.l2p(xc, o["l2l_L"], tb)
        o["l2p_bodies_out"] = tb
        src = np.zeros(40, inputs.BODY_DTYPE)
        src["X"] = rng.random((40, 3))
        src["q"] = rng.uniform(-1, 1, 40)
        tgt = src[:15].copy()
        tgt["X"][5] += 0.5
        o["p2p_src"], o["p2p_tgt_in"] = src, tgt.copy()
        ref.p2p(tgt, src)
        o["p2p_tgt_out"] = tgt
        rab = np.array([1.7, 0.9, -2.2])
        o["harm_rab"] = rab
        o["eval_local"] = ref.eval_local(*rab)
        o["eval_multipole"], o["eval_multipole_theta"] = ref.eval_multipole(0.6, 1.1, 0.4)
        save(f"ops_P{P}.npz", P=P, **o)

    # Periodic (3dp): the shipped driver's set (N=1000, X in [-pi,pi)^3,
    # neutral q, P=10, ncrit=64, theta=0.4, images=4, cycle=2pi) and the
    # Ewald sum with its parameters (3dp/fmm.cxx:13-23).
    refp = oracle.ReferencePeriodic()
    drv = inputs.reference_driver_bodies(1000, 0)
    b2 = drv.copy()
    refp.fmm(b2, 10, 64, 0.4, 4, 2 * np.pi)
    ew = b2.copy()
    refp.ewald(ew, 64, 11, 11 / (2 * np.pi), 0.25 / np.pi, np.pi, 2 * np.pi)
    save("periodic_n1000.npz", input=drv, fmm_sorted=b2, ewald=ew)
    periodic_high_order(refp)
    config5(refp)
    two_trees(ref)
    config2(ref)


if __name__ == "__main__":
    main()
