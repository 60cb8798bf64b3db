"""Forest evaluation (exafmm_evaluate_forest, and exafmm_evaluate_batch with
EXAFMM_BATCH_FOREST): independent problems in one tree build, one traversal
and one launch per pass, against each problem evaluated alone.

Bars:
  X, q of every problem after the call vs exafmm_evaluate of that
      problem alone (same root cube and tree per problem, hence the
      same permutation of its bodies)                                bit for bit
  p, F vs the problem alone (same lists; the keyed GEMM packs other
      problems' cells into its tiles, so the M2L sums run in another
      FP32 order)                                                   rel L2 <= 2e-6 (p), 1e-5 (F)
  p, F vs the FP64 direct sum of each problem (P = 8)               rel L2 <= 5e-5 (p), 2e-4 (F)
"""
import numpy as np
import pytest

from minimal_amd import fmm
from minimal_amd.inputs import BODY_DTYPE, batched_problems, rel_l2, sample_targets


def _tiny(n, seed):
    rng = np.random.default_rng(seed)
    b = np.zeros(n, BODY_DTYPE)
    b["X"] = rng.random((n, 3)) * 3.0 - 7.0
    b["q"] = rng.uniform(-1, 1, n)
    return b


def _close(a, b, tol):
    if not np.any(b):
        return np.array_equal(a, b)
    return rel_l2(a, b) <= tol


@pytest.mark.gpu
def test_forest_matches_problems_alone(gpu):
    probs = [b for b, _ in batched_problems(10, 11)]
    probs[3:3] = [_tiny(1, 1), _tiny(40, 2), _tiny(65, 3)]  # one body; one leaf; a root that splits
    off = np.concatenate([[0], np.cumsum([len(b) for b in probs])])
    allb = np.concatenate(probs)
    f = fmm.FMM(P=8, ncrit=64, theta=0.4, autotune=0)
    f.evaluate_forest(allb, off)
    for i, b in enumerate(probs):
        alone = b.copy()
        g = fmm.FMM(P=8, ncrit=64, theta=0.4, autotune=0)
        g.evaluate(alone)
        got = allb[off[i]:off[i + 1]]
        assert np.array_equal(got["X"], alone["X"]) and np.array_equal(got["q"], alone["q"]), i
        assert _close(got["p"], alone["p"], 2e-6) and _close(got["F"], alone["F"], 1e-5), (
            i, rel_l2(got["p"], alone["p"]), rel_l2(got["F"], alone["F"]))
        if len(b) > 100:
            idx = sample_targets(len(b), 200)
            ref = got[idx].copy()
            ref["p"] = 0
            ref["F"] = 0
            g.direct(ref, got.copy())
            ep, eF = rel_l2(got["p"][idx], ref["p"]), rel_l2(got["F"][idx], ref["F"])
            assert ep <= 5e-5 and eF <= 2e-4, (i, len(b), ep, eF)
        g.close()
    # a second call on the same context (buffers reused, trees rebuilt)
    again = np.concatenate(probs)
    again["p"] = 0
    again["F"] = 0
    f.evaluate_forest(again, off)
    assert np.array_equal(again["p"], allb["p"]) and np.array_equal(again["F"], allb["F"])
    f.close()


@pytest.mark.gpu
def test_zero_radius_problems(gpu):
    """One body, or bodies at one point: the root radius is 0 (as the
    reference's getBounds makes it, 3d/build_tree.h:7-21), nothing
    interacts, p = F = 0 -- alone and inside a forest."""
    one = _tiny(1, 6)
    same = np.repeat(_tiny(1, 7), 5)
    for b in (one, same):
        x = b.copy()
        f = fmm.FMM(P=6)
        f.evaluate(x)
        assert np.all(x["p"] == 0) and np.all(x["F"] == 0)
        f.close()
    allb = np.concatenate([one, _tiny(300, 8), same])
    off = [0, 1, 301, 306]
    f = fmm.FMM(P=6)
    f.evaluate_forest(allb, off)
    assert np.all(allb["p"][:1] == 0) and np.all(allb["F"][:1] == 0)
    assert np.all(allb["p"][301:] == 0) and np.all(allb["F"][301:] == 0)
    assert np.all(np.isfinite(allb["p"])) and np.any(allb["p"][1:301] != 0)
    f.close()


@pytest.mark.gpu
def test_batch_forest_matches_batch(gpu):
    """EXAFMM_BATCH_FOREST (problems of orders 6, 8, 12 as three forests)
    against the per-problem batch evaluation."""
    probs = batched_problems(16, 12)
    assert len({P for _, P in probs}) == 3
    ref = [b.copy() for b, _ in probs]
    a = fmm.FMM(ncrit=64, theta=0.4, autotune=0)
    a.evaluate_batch([(r, P) for r, (_, P) in zip(ref, probs)])
    got = [b.copy() for b, _ in probs]
    f = fmm.FMM(ncrit=64, theta=0.4, autotune=0)
    f.evaluate_batch([(g, P) for g, (_, P) in zip(got, probs)], forest=True)
    for i, (g, r) in enumerate(zip(got, ref)):
        assert np.array_equal(g["X"], r["X"]), i
        assert rel_l2(g["p"], r["p"]) <= 2e-6 and rel_l2(g["F"], r["F"]) <= 1e-5, (
            i, probs[i][1], rel_l2(g["p"], r["p"]), rel_l2(g["F"], r["F"]))
    a.close()
    f.close()


@pytest.mark.gpu
def test_forest_argument_errors(gpu):
    b = np.concatenate([_tiny(10, 4), _tiny(20, 5)])
    f = fmm.FMM(P=6)
    with pytest.raises(fmm.ExaFMMError) as e:
        f.evaluate_forest(b, [0, 20, 10])  # decreasing
    assert e.value.code == -1
    with pytest.raises(fmm.ExaFMMError) as e:
        f.evaluate_forest(b, [5, 30])  # offsets[0] != 0
    assert e.value.code == -1
    many = np.concatenate([_tiny(1, i) for i in range(fmm.FOREST_MAX + 1)])
    with pytest.raises(fmm.ExaFMMError) as e:
        f.evaluate_forest(many, np.arange(fmm.FOREST_MAX + 2))
    assert e.value.code == -1
    f.close()
    # a forest context cannot be recorded in a graph
    g = fmm.FMM(P=6)
    g.evaluate_forest(b.copy(), [0, 10, 30])
    with pytest.raises(fmm.ExaFMMError) as e:
        fmm.Graph([g])
    assert e.value.code == -8
    g.close()
    p = fmm.FMM(P=6, images=1, cycle=1.0)
    with pytest.raises(fmm.ExaFMMError) as e:
        p.evaluate_forest(b, [0, 10, 30])
    assert e.value.code == -8
    p.close()
