"""Batched evaluation of independent problems in one HIP graph
(exafmm_graph_*) against the same contexts evaluated one by one.

Bars (written in each test):
  graph replay vs eager evaluation       bit-exact p and F (same kernels, same
                                         accumulation order, fill kernels for
                                         the zeroing)
  charges doubled between replays        p and F exactly doubled (every pass is
                                         linear in q; scaling by 2 is exact)
  replay vs FP64 direct sum              relative L2 <= 1e-5 (P = 12), 1e-4
                                         (P = 8), 1e-3 (P = 6)
  grouped graph vs eager evaluation      relative L2 of the difference <= 1e-6
                                         (one GEMM shape per order instead of
                                         the autotuned per-problem shapes: the
                                         FP32 partial sums of M2L differ in
                                         order); bit-exact between replays
"""
import numpy as np
import pytest

from minimal_amd import fmm
from minimal_amd.inputs import BODY_DTYPE, batched_problems, rel_l2, sample_targets

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _problems():
    """Six problems of the batched workload's kinds, small sizes: uniform and
    Plummer, N from 1.5k to 20k, P in {6, 8, 12}."""
    ps = batched_problems(256, 7)
    pick = [0, 1, 2, 3, 4, 5]  # even uniform, odd Plummer
    out = []
    for k, i in enumerate(pick):
        b, _ = ps[i]
        n = min(len(b), 20000)
        out.append((b[:n].copy(), (6, 8, 12)[k % 3]))
    return out


def _setup(problems):
    devs, ctxs = [], []
    for b, P in problems:
        d = torch.from_numpy(b.view(np.uint8).copy()).cuda()
        f = fmm.FMM(P=P, ncrit=64, theta=0.4)
        f.evaluate(d.data_ptr(), n=len(b), on_device=True)
        f.synchronize()
        devs.append(d)
        ctxs.append(f)
    return devs, ctxs


def _host(d):
    return d.cpu().numpy().view(BODY_DTYPE).copy()


def test_graph_replay_matches_eager(gpu):
    probs = _problems()
    devs, ctxs = _setup(probs)
    eager = [_host(d) for d in devs]
    g = fmm.Graph(ctxs, streams=3)
    assert g.nodes() > 6 * 8
    for rep in range(3):
        for d in devs:  # scribble over the outputs: the graph must overwrite them
            h = d.cpu().numpy().view(BODY_DTYPE)
            h["p"] = 7.0
            h["F"] = -3.0
            d.copy_(torch.from_numpy(h.view(np.uint8).copy()).cuda())
        torch.cuda.synchronize()
        g.launch()
        g.synchronize()
        for i, d in enumerate(devs):
            h = _host(d)
            assert np.array_equal(h["X"], eager[i]["X"]), (rep, i)
            assert np.array_equal(h["p"], eager[i]["p"]), (rep, i)
            assert np.array_equal(h["F"], eager[i]["F"]), (rep, i)
        x = torch.ones(1 << 20, device="cuda").sum()  # eager work between replays
        assert float(x) == float(1 << 20)
    g.close()


def test_graph_uses_current_charges(gpu):
    probs = _problems()
    devs, ctxs = _setup(probs)
    eager = [_host(d) for d in devs]
    g = fmm.Graph(ctxs, streams=2)
    for d in devs:
        h = d.cpu().numpy().view(BODY_DTYPE)
        h["q"] *= 2.0
        d.copy_(torch.from_numpy(h.view(np.uint8).copy()).cuda())
    torch.cuda.synchronize()
    g.launch()
    g.synchronize()
    for i, d in enumerate(devs):
        h = _host(d)
        assert np.array_equal(h["p"], 2.0 * eager[i]["p"]), i
        assert np.array_equal(h["F"], 2.0 * eager[i]["F"]), i
    g.close()


def test_graph_accuracy_vs_direct(gpu):
    probs = _problems()
    devs, ctxs = _setup(probs)
    g = fmm.Graph(ctxs, streams=4)
    g.launch()
    g.synchronize()
    bound = {6: 1e-3, 8: 1e-4, 12: 1e-5}
    for (b, P), d, f in zip(probs, devs, ctxs):
        h = _host(d)
        idx = sample_targets(len(h), 200)
        ref = h[idx].copy()
        ref["p"] = 0
        ref["F"] = 0
        f.direct(ref, h)
        assert rel_l2(h["p"][idx], ref["p"]) <= bound[P], P
        assert rel_l2(h["F"][idx], ref["F"]) <= bound[P], P
    g.close()


def test_graph_holds_contexts(gpu):
    probs = _problems()[:2]
    devs, ctxs = _setup(probs)
    g = fmm.Graph(ctxs, streams=1)
    with pytest.raises(fmm.ExaFMMError):
        ctxs[0].evaluate(devs[0].data_ptr(), n=len(probs[0][0]), on_device=True)
    with pytest.raises(fmm.ExaFMMError):
        fmm.Graph(ctxs[:1])  # already held
    g.close()
    ctxs[0].evaluate(devs[0].data_ptr(), n=len(probs[0][0]), on_device=True)  # released
    ctxs[0].synchronize()


def test_graph_rejects_unevaluated_and_host_contexts(gpu):
    f = fmm.FMM(P=6)
    with pytest.raises(fmm.ExaFMMError):
        fmm.Graph([f])
    b, _ = _problems()[0]
    h = fmm.FMM(P=6)
    hb = b.copy()
    h.evaluate(hb)  # host-resident bodies
    with pytest.raises(fmm.ExaFMMError):
        fmm.Graph([h])


def test_grouped_graph_matches_eager(gpu):
    probs = _problems()
    devs, ctxs = _setup(probs)
    eager = [_host(d) for d in devs]
    g = fmm.Graph(ctxs, grouped=True)
    assert g.nodes() < 60  # one launch per stage, per level and per order
    first = None
    for rep in range(3):
        for d in devs:
            h = d.cpu().numpy().view(BODY_DTYPE)
            h["p"] = 7.0
            h["F"] = -3.0
            d.copy_(torch.from_numpy(h.view(np.uint8).copy()).cuda())
        torch.cuda.synchronize()
        g.launch()
        g.synchronize()
        res = [_host(d) for d in devs]
        for i, h in enumerate(res):
            assert rel_l2(h["p"], eager[i]["p"]) <= 1e-6, (rep, i)
            assert rel_l2(h["F"], eager[i]["F"]) <= 1e-6, (rep, i)
        if first is None:
            first = res
        else:
            for a, b in zip(first, res):
                assert np.array_equal(a["p"], b["p"]) and np.array_equal(a["F"], b["F"])
        x = torch.ones(1 << 20, device="cuda").sum()
        assert float(x) == float(1 << 20)
    g.close()


def test_grouped_graph_charges_and_accuracy(gpu):
    probs = _problems()
    devs, ctxs = _setup(probs)
    g = fmm.Graph(ctxs, grouped=True)
    g.launch()
    g.synchronize()
    base = [_host(d) for d in devs]
    bound = {6: 1e-3, 8: 1e-4, 12: 1e-5}
    for (b, P), h, f in zip(probs, base, ctxs):
        idx = sample_targets(len(h), 200)
        ref = h[idx].copy()
        ref["p"] = 0
        ref["F"] = 0
        f.direct(ref, h)
        assert rel_l2(h["p"][idx], ref["p"]) <= bound[P], P
        assert rel_l2(h["F"][idx], ref["F"]) <= bound[P], P
    for d in devs:
        h = d.cpu().numpy().view(BODY_DTYPE)
        h["q"] *= 2.0
        d.copy_(torch.from_numpy(h.view(np.uint8).copy()).cuda())
    torch.cuda.synchronize()
    g.launch()
    g.synchronize()
    for i, d in enumerate(devs):
        h = _host(d)
        assert np.array_equal(h["p"], 2.0 * base[i]["p"]), i
        assert np.array_equal(h["F"], 2.0 * base[i]["F"]), i
    g.close()


def test_grouped_graph_rejects_other_precisions(gpu):
    b, _ = _problems()[0]
    d = torch.from_numpy(b.view(np.uint8).copy()).cuda()
    f = fmm.FMM(P=6, m2l_precision=fmm.PREC_FP32)
    f.evaluate(d.data_ptr(), n=len(b), on_device=True)
    with pytest.raises(fmm.ExaFMMError):
        fmm.Graph([f], grouped=True)
    fmm.Graph([f]).close()  # the per-problem graph takes it


def test_evaluate_batch_matches_single_contexts(gpu):
    """exafmm_evaluate_batch on host arrays equals per-problem contexts bit
    for bit (both without the GEMM autotuner, whose timed pick between
    near-equal shapes may differ between contexts and with it the FP32 order
    of the M2L sums); on device arrays with reused geometry the results are
    the recorded graph's (1e-6 of the eager ones), and its replay after
    doubling the charges doubles p and F (1e-6)."""
    probs = _problems()
    single = []
    for b, P in probs:
        f = fmm.FMM(P=P, ncrit=64, theta=0.4, autotune=0)
        bb = b.copy()
        f.evaluate(bb)
        single.append(bb)
    host = [b.copy() for b, _ in probs]
    batch = fmm.FMM(ncrit=64, theta=0.4, autotune=0)
    batch.evaluate_batch([(h, P) for h, (_, P) in zip(host, probs)])
    for h, s in zip(host, single):
        assert np.array_equal(h["X"], s["X"])
        assert np.array_equal(h["p"], s["p"]) and np.array_equal(h["F"], s["F"])
    devs = [torch.from_numpy(b.view(np.uint8).copy()).cuda() for b, _ in probs]
    spec = [(d.data_ptr(), len(b), P) for d, (b, P) in zip(devs, probs)]
    batch.evaluate_batch(spec, reuse_geometry=True)
    # recording the graph ends with one pass of it, so these are the graph's
    # results (grouped GEMM shapes: FP32 partial sums in another order)
    first = [_host(d) for d in devs]
    for h, s in zip(first, single):
        assert np.array_equal(h["X"], s["X"])
        assert rel_l2(h["p"], s["p"]) <= 1e-6 and rel_l2(h["F"], s["F"]) <= 1e-6
    for d in devs:
        h = d.cpu().numpy().view(BODY_DTYPE)
        h["q"] *= 2.0
        d.copy_(torch.from_numpy(h.view(np.uint8).copy()).cuda())
    torch.cuda.synchronize()
    batch.evaluate_batch(spec, reuse_geometry=True)  # graph replay
    for d, f in zip(devs, first):
        h = _host(d)
        assert rel_l2(h["p"], 2.0 * f["p"]) <= 1e-6 and rel_l2(h["F"], 2.0 * f["F"]) <= 1e-6
    batch.close()
