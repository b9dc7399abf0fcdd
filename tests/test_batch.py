"""Batched form (include/rmq_batch.h): many arrays, one persistent launch, HIP-graph replay.

Batched answers must equal per-array plain launches (rmq_build + rmq_query) and the oracle's
CPU BASE bit-exactly, for every execution variant and scan threshold."""

import ctypes
import sys

import numpy as np
import pytest

import oracle
import rmq_amd
from rmq_amd import _lib, workloads


def test_batch_abi_rejects_bad_arguments():
    L = rmq_amd.lib()
    out = ctypes.c_void_p()
    arr = (_lib.BatchArray * 1)()
    arr[0].offset, arr[0].n, arr[0].t, arr[0].variant = 0, 100, 64, 8
    assert L.rmq_batch_build(None, arr, 1, None, ctypes.byref(out)) == rmq_amd.RMQ_EINVAL
    assert L.rmq_batch_build(0x1000, arr, 0, None, ctypes.byref(out)) == rmq_amd.RMQ_EINVAL
    arr[0].t = 63
    assert L.rmq_batch_build(0x1000, arr, 1, None, ctypes.byref(out)) == rmq_amd.RMQ_ETHRESHOLD
    arr[0].t, arr[0].variant = 64, 5
    assert L.rmq_batch_build(0x1000, arr, 1, None, ctypes.byref(out)) == rmq_amd.RMQ_EUNSUPPORTED
    arr[0].variant, arr[0].offset = 8, 2
    assert L.rmq_batch_build(0x1000, arr, 1, None, ctypes.byref(out)) == rmq_amd.RMQ_EINVAL
    assert L.rmq_batch_query(None, None, 0x1000, 0x1000, None, None) == rmq_amd.RMQ_EINVAL
    assert L.rmq_fill_uniform(None, 10, 1, 0, None) == rmq_amd.RMQ_EINVAL
    L.rmq_batch_free(None)
    L.rmq_batch_plan_free(None)
    L.rmq_graph_free(None)


def test_workload_generator_shapes():
    w = workloads.batch_workload(arrays=50, log_n=(6, 14), log_q=(3, 9), seed=3)
    assert len(w.sizes) == 50 and w.qoffsets[-1] == w.queries.shape[0]
    for a in range(50):
        q = w.queries[w.qoffsets[a]:w.qoffsets[a + 1]]
        assert (q[:, 0] <= q[:, 1]).all() and (q[:, 1] < w.sizes[a]).all()
    assert (w.offsets % 64 == 0).all()


@pytest.mark.gpu
def test_fill_uniform_matches_host_restatement(device):
    import torch

    out = torch.empty(100003, dtype=torch.float32, device="cuda")
    rmq_amd.fill_uniform(out, seed=12345, start=777)
    torch.cuda.synchronize()
    want = workloads.splitmix_uniform(12345, 777, 100003)
    assert np.array_equal(out.cpu().numpy().view(np.uint32), want.view(np.uint32))
    assert 0.0 <= float(out.min()) and float(out.max()) < 1.0


def _device_workload(arrays, log_n, log_q, seed, kind="per-array"):
    import torch

    w = workloads.batch_workload(arrays=arrays, log_n=log_n, log_q=log_q, kind=kind, seed=seed)
    total = int(w.offsets[-1] + w.sizes[-1] + 64)
    dA = torch.empty(total, dtype=torch.float32, device="cuda")
    rmq_amd.fill_uniform(dA, seed=seed)
    dQ = torch.from_numpy(w.queries.view(np.int32)).cuda()
    torch.cuda.synchronize()
    return w, dA, dQ


@pytest.mark.gpu
@pytest.mark.parametrize("want", ["index", "value"])
def test_batch_equals_plain_and_oracle(device, want):
    import torch

    w, dA, dQ = _device_workload(96, (4, 17), (2, 11), seed=11)
    A_host = dA.cpu().numpy()
    # plant ties and signed zeros in a few arrays
    rng = np.random.default_rng(5)
    for a in range(0, 96, 7):
        o, n = int(w.offsets[a]), int(w.sizes[a])
        A_host[o:o + n] = rng.integers(0, 3, size=n).astype(np.float32) * \
            np.where(rng.random(n) < 0.5, np.float32(-1), np.float32(1))
    dA.copy_(torch.from_numpy(A_host))
    variants = ["group8", "group4", "group16", "lane", "lane8", "group32", "wave", "lane4"]
    t = [64 << (a % 5) for a in range(96)]
    var = [variants[a % 8] for a in range(96)]
    wv, wi = want in ("value",), want in ("index",)
    with rmq_amd.Batch(dA, w.offsets, w.sizes, t, var) as b:
        plan = b.plan(w.qoffsets)
        bv, bi = plan.query(dQ, want_val=True, want_idx=wi)
        torch.cuda.synchronize()
    bv = bv.cpu().numpy()
    bi = bi.cpu().numpy().view(np.uint32) if wi else None
    for a in range(96):
        o, n = int(w.offsets[a]), int(w.sizes[a])
        q0, q1 = int(w.qoffsets[a]), int(w.qoffsets[a + 1])
        Aa = A_host[o:o + n]
        lr = w.queries[q0:q1]
        ref_i = oracle.cpu_base_index(Aa, lr)
        assert np.array_equal(bv[q0:q1].view(np.uint32), Aa[ref_i].view(np.uint32)), (a, var[a])
        if wi:
            assert np.array_equal(bi[q0:q1], ref_i), (a, var[a])
        # per-array plain launch through the C ABI gives the same answers
        plain = {"lane8": dict(alg=20, cg_amount_log=1), "lane4": dict(alg=20, cg_amount_log=0)}.get(
            var[a], dict(alg=16))
        with rmq_amd.Hierarchy(dA[o:o + n], t[a], **plain) as h:
            pv, pi = h.query(dQ[q0:q1], True, True)
            torch.cuda.synchronize()
        assert np.array_equal(pv.cpu().numpy().view(np.uint32), bv[q0:q1].view(np.uint32))
        if wi:
            assert np.array_equal(pi.cpu().numpy().view(np.uint32), bi[q0:q1])


@pytest.mark.gpu
def test_batch_lane_variants_only(device):
    """A batch whose arrays all run lane variants takes the kernel built without the group
    solvers: answers equal CPU BASE in value and index mode, and its graph replays match."""
    import torch

    w, dA, dQ = _device_workload(120, (4, 18), (2, 12), seed=31, kind="mixed")
    A_host = dA.cpu().numpy()
    rng = np.random.default_rng(31)
    for a in range(0, 120, 9):  # ties and signed zeros in some arrays
        o, n = int(w.offsets[a]), int(w.sizes[a])
        A_host[o:o + n] = rng.integers(0, 3, size=n).astype(np.float32) * \
            np.where(rng.random(n) < 0.5, np.float32(-1), np.float32(1))
    dA.copy_(torch.from_numpy(A_host))
    var = [("lane", "lane8", "lane4")[a % 3] for a in range(120)]
    t = [64 << (a % 6) for a in range(120)]
    with rmq_amd.Batch(dA, w.offsets, w.sizes, t, var) as b:
        plan = b.plan(w.qoffsets)
        bv, bi = plan.query(dQ, want_val=True, want_idx=True)
        vv, _ = plan.query(dQ, want_val=True, want_idx=False)
        torch.cuda.synchronize()
        gv = torch.full_like(bv, -1.0)
        gi = torch.full_like(bi, -1)
        graph = plan.capture(dQ, gv, gi)
        for rep in range(3):
            gv.fill_(-1.0)
            gi.fill_(-1)
            torch.cuda.synchronize()
            graph.launch()
            torch.cuda.synchronize()
            assert torch.equal(gv.view(torch.int32), bv.view(torch.int32)), rep
            assert torch.equal(gi, bi), rep
    bv = bv.cpu().numpy()
    vv = vv.cpu().numpy()
    bi = bi.cpu().numpy().view(np.uint32)
    for a in range(120):
        o, n = int(w.offsets[a]), int(w.sizes[a])
        q0, q1 = int(w.qoffsets[a]), int(w.qoffsets[a + 1])
        Aa = A_host[o:o + n]
        ref_i = oracle.cpu_base_index(Aa, w.queries[q0:q1])
        assert np.array_equal(bi[q0:q1], ref_i), (a, var[a])
        assert np.array_equal(bv[q0:q1].view(np.uint32), Aa[ref_i].view(np.uint32)), (a, var[a])
        assert np.array_equal(vv[q0:q1].view(np.uint32), Aa[ref_i].view(np.uint32)), (a, var[a])


@pytest.mark.gpu
def test_batch_graph_replays(device):
    """The captured launch replays correctly several times, with eager work in between."""
    import torch

    w, dA, dQ = _device_workload(200, (8, 18), (4, 12), seed=21, kind="mixed")
    with rmq_amd.Batch(dA, w.offsets, w.sizes, [1 << 10] * 200, ["group8"] * 200) as b:
        plan = b.plan(w.qoffsets)
        ev, ei = plan.query(dQ, True, True)
        torch.cuda.synchronize()
        eager_v, eager_i = ev.clone(), ei.clone()
        gv = torch.full_like(ev, -1.0)
        gi = torch.full_like(ei, -1)
        graph = plan.capture(dQ, gv, gi)
        for rep in range(4):
            gv.fill_(-1.0)
            gi.fill_(-1)
            torch.cuda.synchronize()
            graph.launch()
            torch.cuda.synchronize()
            assert torch.equal(gv.view(torch.int32), eager_v.view(torch.int32)), rep
            assert torch.equal(gi, eager_i), rep
            junk = torch.rand(1 << 20, device="cuda").sum()  # eager work between replays
            torch.cuda.synchronize()
            assert float(junk) > 0


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_config4_full_workload(device):
    """The batched workload at full size, exactly as bench.py runs it: 4096 arrays with sizes
    log-U[2^10, 2^24] (about 7.3 G elements), log-U[2^6, 2^16] queries per array (about 39 M),
    mixed lengths, variant and t autotuned per array. The eager launch (value + index, value
    only) and two graph replays equal the per-array plain launches (rmq_build_each /
    rmq_query_each) on every query; every array's answers equal CPU BASE (the oracle) on a sample
    of its queries, and every 16th array's equal the oracle hierarchy on all of its queries."""
    import torch

    from rmq_amd import autotune

    seed = 27722
    w = workloads.batch_workload(arrays=4096, log_n=(10, 24), log_q=(6, 16), kind="mixed", seed=seed)
    total = int(w.offsets[-1] + w.sizes[-1] + 64)
    dA = torch.empty(total, dtype=torch.float32, device="cuda")
    rmq_amd.fill_uniform(dA, seed=seed)
    dQ = torch.from_numpy(w.queries.view(np.int32)).cuda()
    choices = []
    for a in range(4096):
        q = w.queries[w.qoffsets[a]:w.qoffsets[a + 1]]
        choices.append(autotune.choose(int(w.sizes[a]), q[:, 1].astype(np.int64) - q[:, 0] + 1))
    var = [c[0] for c in choices]
    ts = [c[1] for c in choices]
    Q = int(w.qoffsets[-1])
    torch.cuda.synchronize()
    sys.stderr.write(f"[gpu-case] config 4: {Q} queries, variants {sorted(set(var))}\n")
    with rmq_amd.Batch(dA, w.offsets, w.sizes, ts, var) as b:
        plan = b.plan(w.qoffsets)
        bv, bi = plan.query(dQ, True, True)
        vv, _ = plan.query(dQ, True, False)
        torch.cuda.synchronize()
        gi = torch.full_like(bi, -1)
        graph = plan.capture(dQ, None, gi)
        for rep in range(2):
            gi.fill_(-1)
            torch.cuda.synchronize()
            graph.launch()
            torch.cuda.synchronize()
            assert torch.equal(gi, bi), rep
            junk = torch.rand(1 << 20, device="cuda").sum()  # eager work between replays
            torch.cuda.synchronize()
            assert float(junk) > 0
        del gi
        # the batched build's levels of every 64th array against the oracle's level build
        rlog = {"lane": 5, "lane8": 3, "lane4": 2}
        for a in range(0, 4096, 64):
            o, n = int(w.offsets[a]), int(w.sizes[a])
            ref = oracle.CpuHierarchy(dA[o:o + n].cpu().numpy(), ts[a], rlog.get(var[a], 5))
            lv, li = b.levels(a, ref.floats)
            torch.cuda.synchronize()
            assert np.array_equal(lv.cpu().numpy().view(np.uint32), ref.lv[:ref.floats].view(np.uint32)), a
            assert np.array_equal(li.cpu().numpy().view(np.uint32), ref.li[:ref.floats]), a
    sys.stderr.write("[gpu-case] config 4: per-array plain launches\n")
    with rmq_amd.PlainMany(dA, w.offsets, w.sizes, ts) as pm:
        out_p = torch.empty(Q, dtype=torch.int32, device="cuda")
        pm.query(dQ, w.qoffsets, out_p)
        torch.cuda.synchronize()
    assert torch.equal(out_p, bi)
    del out_p
    base = torch.from_numpy(np.repeat(w.offsets, w.qcounts)).cuda()
    at = dA[base + (bi.to(torch.int64) & 0xFFFFFFFF)]
    assert torch.equal(at.view(torch.int32), bv.view(torch.int32))
    assert torch.equal(at.view(torch.int32), vv.view(torch.int32))
    del base, at, vv, bv
    bi_h = bi.cpu().numpy().view(np.uint32)
    sys.stderr.write("[gpu-case] config 4: oracle\n")
    rng = np.random.default_rng(seed + 1)
    for a in range(4096):
        o, n = int(w.offsets[a]), int(w.sizes[a])
        q0, q1 = int(w.qoffsets[a]), int(w.qoffsets[a + 1])
        Aa = dA[o:o + n].cpu().numpy()
        lr = w.queries[q0:q1]
        got = bi_h[q0:q1]
        s = rng.choice(q1 - q0, size=min(16, q1 - q0), replace=False)
        assert np.array_equal(got[s], oracle.cpu_base_index(Aa, lr[s])), (a, var[a], ts[a])
        if a % 16 == 0:
            assert np.array_equal(got, oracle.CpuHierarchy(Aa, 64, 5, 16).query(lr)[1]), (a, var[a], ts[a])
    del dA, dQ, bi
    torch.cuda.empty_cache()


@pytest.mark.gpu
def test_batch_nan_arrays(device):
    """Arrays holding NaN in the batched form: a NaN at A[l] answers l, elsewhere a NaN never wins
    (CPU BASE, the oracle), for every variant, value and index."""
    import torch

    w, dA, dQ = _device_workload(64, (4, 16), (2, 11), seed=71, kind="mixed")
    A_host = dA.cpu().numpy()
    rng = np.random.default_rng(71)
    for a in range(64):
        o, n = int(w.offsets[a]), int(w.sizes[a])
        k = a % 4
        if k == 0:
            A_host[o + rng.choice(n, max(1, n // 20), replace=False)] = np.nan
        elif k == 1:
            s = int(rng.integers(0, n))
            A_host[o + s:o + min(n, s + 40 * (a + 1))] = np.nan
        elif k == 2:
            A_host[o:o + n] = np.nan
        q0, q1 = int(w.qoffsets[a]), int(w.qoffsets[a + 1])
        nanpos = np.nonzero(np.isnan(A_host[o:o + n]))[0]
        if nanpos.size:  # some queries start on a NaN
            sel = w.queries[q0:q1:3]
            sel[:, 0] = np.minimum(nanpos[rng.integers(0, nanpos.size, sel.shape[0])], sel[:, 1])
            w.queries[q0:q1:3] = sel
    dA.copy_(torch.from_numpy(A_host))
    dQ = torch.from_numpy(w.queries.view(np.int32)).cuda()
    variants = ["lane", "lane8", "lane4", "group8", "group4", "group16", "wave"]
    var = [variants[a % 7] for a in range(64)]
    t = [64 << (a % 3) for a in range(64)]
    for lanes_only in (False, True):
        vv = [("lane", "lane8", "lane4")[a % 3] for a in range(64)] if lanes_only else var
        with rmq_amd.Batch(dA, w.offsets, w.sizes, t, vv) as b:
            plan = b.plan(w.qoffsets)
            bv, bi = plan.query(dQ, True, True)
            ov, _ = plan.query(dQ, True, False)
            torch.cuda.synchronize()
        bv, ov = bv.cpu().numpy().view(np.uint32), ov.cpu().numpy().view(np.uint32)
        bi = bi.cpu().numpy().view(np.uint32)
        for a in range(64):
            o, n = int(w.offsets[a]), int(w.sizes[a])
            q0, q1 = int(w.qoffsets[a]), int(w.qoffsets[a + 1])
            Aa = A_host[o:o + n]
            ref = oracle.cpu_base_index(Aa, w.queries[q0:q1])
            assert np.array_equal(bi[q0:q1], ref), (a, vv[a])
            assert np.array_equal(bv[q0:q1], Aa[ref].view(np.uint32)), (a, vv[a])
            assert np.array_equal(ov[q0:q1], Aa[ref].view(np.uint32)), (a, vv[a], "value")
