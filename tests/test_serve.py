"""Serving loop (include/rmq_serve.h): resident hierarchies, free-memory admission, continuous
batching. Every answer must equal the oracle's, and a request must get the same answer alone
and when batched with others."""

import ctypes
import sys

import numpy as np
import pytest

import oracle
import rmq_amd
from rmq_amd import _lib, workloads


def test_server_abi_rejects_bad_arguments():
    L = rmq_amd.lib()
    assert L.rmq_server_create(None, None) == rmq_amd.RMQ_EINVAL
    aid = ctypes.c_int32()
    assert L.rmq_server_admit(None, 0, None, 10, 0, 64, 1, ctypes.byref(aid)) == rmq_amd.RMQ_EINVAL
    t = ctypes.c_uint64()
    assert L.rmq_server_submit(None, 0, None, 1, None, None, ctypes.byref(t)) == rmq_amd.RMQ_EINVAL
    assert L.rmq_server_poll(None, None, 0, None) == rmq_amd.RMQ_EINVAL
    assert L.rmq_server_evict(None, 0) == rmq_amd.RMQ_EINVAL
    L.rmq_server_destroy(None)


def _host_arrays(rng, count, lo, hi):
    arrays = []
    for a in range(count):
        n = int(rng.integers(lo, hi))
        A = rng.random(n, dtype=np.float32)
        if a % 3 == 0:  # ties and signed zeros
            A = rng.integers(0, 3, size=n).astype(np.float32)
            A[rng.random(n) < 0.4] *= np.float32(-1)
        arrays.append(A)
    return arrays


@pytest.mark.gpu
def test_serve_matches_oracle_and_is_order_independent(device):
    rng = np.random.default_rng(17)
    arrays = _host_arrays(rng, 12, 1000, 300000)
    variants = ["lane", "group8", "group4", "group16", "lane8", "wave", "lane4"]
    with rmq_amd.Server(slots=3, max_batch_queries=1 << 16) as srv:
        ids = [srv.admit_host(A, t=64 << (i % 4), variant=variants[i % 7])
               for i, A in enumerate(arrays)]
        reqs = []
        for k in range(150):
            a = int(rng.integers(0, len(arrays)))
            n = arrays[a].size
            q = int(np.exp2(rng.uniform(0, 13)))
            kind = ["small", "medium", "large"][k % 3]
            reqs.append((a, workloads.make_queries(rng, n, q, kind), k % 2 == 0))
        # every probe alone first
        alone = []
        for a, lr, wi in reqs[:20]:
            t, v, i = srv.submit(ids[a], lr, True, wi)
            assert srv.drain() == [t]
            alone.append((v.copy(), None if i is None else i.copy()))
        # then everything in a mix: submit in bursts while earlier batches are in flight
        out = []
        done = set()
        for j, (a, lr, wi) in enumerate(reqs):
            out.append(srv.submit(ids[a], lr, True, wi))
            if j % 7 == 0:
                done.update(srv.poll())
        done.update(srv.drain())
        assert done == {t for t, _, _ in out}
        st = srv.stats()
        assert st["completed"] == st["submitted"] == 170 and st["in_flight"] == 0
        assert st["batches"] < 170  # requests were batched together
        for j, ((a, lr, wi), (t, v, i)) in enumerate(zip(reqs, out)):
            ref = oracle.cpu_base_index(arrays[a], lr)
            assert np.array_equal(v.view(np.uint32), arrays[a][ref].view(np.uint32)), j
            if wi:
                assert np.array_equal(i, ref), j
            if j < 20:
                assert np.array_equal(v.view(np.uint32), alone[j][0].view(np.uint32))
                if wi:
                    assert np.array_equal(i, alone[j][1])


@pytest.mark.gpu
def test_admission_by_free_memory_and_eviction(device):
    import torch

    free, total = torch.cuda.mem_get_info()
    # a reserve that leaves ~1 GiB admissible
    with rmq_amd.Server(slots=2, reserve_bytes=max(free - (1 << 30), 1)) as srv:
        small = srv.admit_uniform(1 << 20, seed=1, t=64)
        with pytest.raises(rmq_amd.RMQError) as e:
            srv.admit_uniform(1 << 29, seed=2, t=64)  # 2 GiB array + levels: refused
        assert e.value.status == rmq_amd.RMQ_ENOMEM
        assert srv.stats()["resident"] == 1
        lr = workloads.make_queries(np.random.default_rng(3), 1 << 20, 5000, "medium")
        t, v, _ = srv.submit(small, lr, True, False)
        with pytest.raises(rmq_amd.RMQError):
            srv.evict(small)  # a request still uses it
        srv.drain()
        host = torch.empty(1 << 20, dtype=torch.float32, device="cuda")
        rmq_amd.fill_uniform(host, seed=1)
        A = host.cpu().numpy()
        ref = oracle.cpu_base_index(A, lr)
        assert np.array_equal(v.view(np.uint32), A[ref].view(np.uint32))
        srv.evict(small)
        assert srv.stats()["resident"] == 0
        with pytest.raises(rmq_amd.RMQError):
            srv.submit(small, lr, True, False)  # evicted


@pytest.mark.gpu
def test_submit_refuses_pairs_outside_the_array(device):
    """A request holding one pair with l > r or r >= n is refused with RMQ_EINVAL and nothing is
    queued (the shared batch kernel never sees it); valid requests around it are answered."""
    rng = np.random.default_rng(8)
    A = rng.random(5000, dtype=np.float32)
    with rmq_amd.Server(slots=2, max_batch_queries=1 << 12) as srv:
        aid = srv.admit_host(A, t=64, variant="lane")
        good = workloads.make_queries(rng, A.size, 300, "medium")
        for bad in ([7, 6], [0, 5000], [4999, 5000], [0xFFFFFFFF, 0xFFFFFFFF]):
            lr = good.copy()
            lr[123] = bad
            with pytest.raises(rmq_amd.RMQError) as e:
                srv.submit(aid, lr, True, True)
            assert e.value.status == rmq_amd.RMQ_EINVAL, bad
        st = srv.stats()
        assert st["submitted"] == 0 and st["batches"] == 0
        t, v, i = srv.submit(aid, good, True, True)
        assert srv.drain() == [t]
        ref = oracle.cpu_base_index(A, good)
        assert np.array_equal(i, ref) and np.array_equal(v.view(np.uint32), A[ref].view(np.uint32))


def _explain_mismatch(dA, A, lr, v, x, ref, bad, a, i, srv, aid):
    """Which side differs on the mismatching queries of one request: GPU BASE and a plain
    rmq_query (lane solver, R = 8, t = 64) on the same device array and the request submitted
    again alone, printed beside the server's and the oracle's answers."""
    import torch

    sel = lr[bad[:8]]
    dQ = torch.from_numpy(np.ascontiguousarray(sel).view(np.int32)).cuda()
    bv, bi = rmq_amd.gpu_base(dA, dQ, True, True)
    with rmq_amd.Hierarchy(dA, 64, 20, cg_amount_log=1) as h:
        pv, pi = h.query(dQ, True, True)
        torch.cuda.synchronize()
    _, av, _ = srv.submit(aid, lr, True, False)
    srv.drain()
    sys.stderr.write(f"[gpu-case] config 5 mismatch: array {a} (n={A.size}) request {i}: {bad.size} of "
                     f"{lr.shape[0]} values differ (index requested: {x is not None}); first {bad[:8].tolist()}\n")
    for k in range(sel.shape[0]):
        j = int(bad[k])
        sys.stderr.write(f"  q={j} lr={sel[k].tolist()} server={v[j]!r} ({v.view(np.uint32)[j]:#x}) "
                         f"oracle={A[ref[j]]!r} @ {int(ref[j])} gpu_base={float(bv[k])!r} @ {int(bi[k])} "
                         f"plain={float(pv[k])!r} @ {int(pi[k])} resubmitted alone={av[j]!r}\n")
    sys.stderr.write(f"  resubmitted alone: {int(np.sum(av.view(np.uint32) != A[ref].view(np.uint32)))} differ\n")
    sys.stderr.flush()


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_config5_full_workload(device):
    """The serving workload at full size, as bench.py runs it: 64 resident arrays with sizes
    log-U[2^20, 2^28] (about 3 G elements), each tuned for the mix of lengths it serves;
    1500 requests with Zipf(1.1) targets and 1..2^16 queries (log-uniform) of small, medium or
    large lengths, submitted in bursts while earlier batches are in flight, so that they coalesce.
    Every answer (index and value) equals the oracle hierarchy's, a sample of the short queries
    equals CPU BASE's, and 20 probe requests get the same answers alone as in the mix."""
    import torch

    from rmq_amd import autotune

    seed = 27722
    rng = np.random.default_rng(seed)
    sizes = np.exp2(rng.uniform(20, 28, size=64)).astype(np.int64)
    mrng = np.random.default_rng(seed + 7)
    dev = []
    with rmq_amd.Server(slots=3, max_batch_queries=1 << 21) as srv:
        ids = []
        for a, n in enumerate(sizes):
            d = torch.empty(int(n), dtype=torch.float32, device="cuda")
            rmq_amd.fill_uniform(d, seed=seed + a)
            mix = np.concatenate([workloads.query_lengths(mrng, int(n), 256, k)
                                  for k in ("small", "medium", "large")])
            var, t = autotune.choose(int(n), mix)
            ids.append(srv.admit_device(d, t=t, variant=var))
            dev.append(d)
        ranks = np.arange(1, 65, dtype=np.float64)
        p = ranks ** -1.1
        p /= p.sum()
        nreq = 1500
        targets = rng.choice(64, size=nreq, p=p)
        counts = np.exp2(rng.uniform(0, 16, size=nreq)).astype(np.int64)
        kinds = rng.choice(["small", "medium", "large"], size=nreq)
        reqs = [workloads.make_queries(rng, int(sizes[targets[i]]), int(counts[i]), str(kinds[i]))
                for i in range(nreq)]
        sys.stderr.write(f"[gpu-case] config 5: {int(counts.sum())} queries in {nreq} requests\n")
        alone = []
        for i in range(20):
            t, v, x = srv.submit(ids[targets[i]], reqs[i], True, True)
            assert srv.drain() == [t]
            alone.append((v, x))
        out = []
        done = set()
        for i in range(nreq):
            out.append(srv.submit(ids[targets[i]], reqs[i], True, i % 2 == 0))
            if i % 5 == 4:
                done.update(srv.poll())
        done.update(srv.drain())
        assert done == {t for t, _, _ in out}
        st = srv.stats()
        assert st["completed"] == st["submitted"] == nreq + 20 and st["in_flight"] == 0
        assert st["batches"] < 20 + nreq // 2, st  # requests were batched together
        for i in range(20):
            assert np.array_equal(out[i][1].view(np.uint32), alone[i][0].view(np.uint32)), i
            if out[i][2] is not None:
                assert np.array_equal(out[i][2], alone[i][1]), i
        sys.stderr.write("[gpu-case] config 5: oracle\n")
        for a in range(64):
            mine = np.nonzero(targets == a)[0]
            if mine.size == 0:
                continue
            A = dev[a].cpu().numpy()
            lr = np.concatenate([reqs[i] for i in mine])
            ref = oracle.CpuHierarchy(A, 64, 5, 16).query(lr)[1]
            k = 0
            for i in mine:
                q = reqs[i].shape[0]
                _, v, x = out[i]
                bad = np.nonzero(v.view(np.uint32) != A[ref[k:k + q]].view(np.uint32))[0]
                if bad.size:
                    _explain_mismatch(dev[a], A, reqs[i], v, x, ref[k:k + q], bad, a, i, srv, ids[a])
                assert bad.size == 0, (a, i)
                if x is not None:
                    assert np.array_equal(x, ref[k:k + q]), (a, i)
                k += q
            short = np.nonzero((lr[:, 1].astype(np.int64) - lr[:, 0]) < (1 << 16))[0][:500]
            assert np.array_equal(oracle.cpu_base_index(A, lr[short]), ref[short]), a
    del dev
    torch.cuda.empty_cache()


@pytest.mark.gpu
def test_staging_ring_full_and_wrapping(device):
    """The staging ring: a request that does not fit is refused with RMQ_EAGAIN and nothing is
    queued; once earlier requests are done it is accepted. Then 400 requests of 1..4096 queries
    go around a ring of 5000 queries many times (starting over at its beginning whenever a
    request would not fit before its end), with polls in between and retries after RMQ_EAGAIN;
    every answer equals the oracle's."""
    rng = np.random.default_rng(12)
    arrays = _host_arrays(rng, 4, 2000, 200000)
    with rmq_amd.Server(slots=2, max_batch_queries=4096, ring_queries=5000) as srv:
        ids = [srv.admit_host(A, t=64, variant=("lane", "lane8", "lane4", "group8")[i]) for i, A in enumerate(arrays)]
        lr = workloads.make_queries(rng, arrays[0].size, 3000, "medium")
        t1, v1, _ = srv.submit(ids[0], lr, True, False)
        with pytest.raises(rmq_amd.RMQError) as e:
            srv.submit(ids[0], lr, True, False)
        assert e.value.status == rmq_amd.RMQ_EAGAIN
        assert srv.stats()["submitted"] == 1
        assert srv.drain() == [t1]
        t2, v2, _ = srv.submit(ids[0], lr, True, False)
        assert srv.drain() == [t2]
        ref = oracle.cpu_base_index(arrays[0], lr)
        assert np.array_equal(v1.view(np.uint32), arrays[0][ref].view(np.uint32))
        assert np.array_equal(v2.view(np.uint32), arrays[0][ref].view(np.uint32))
        reqs, out = [], []
        refused = 0
        for k in range(400):
            a = int(rng.integers(0, 4))
            q = int(np.exp2(rng.uniform(0, 12)))
            lr = workloads.make_queries(rng, arrays[a].size, q, ("small", "medium", "large")[k % 3])
            for attempt in range(1000):  # bounded: every request fits once earlier ones are done
                try:
                    out.append(srv.submit(ids[a], lr, True, k % 2 == 0))
                    break
                except rmq_amd.RMQError as err:
                    assert err.status == rmq_amd.RMQ_EAGAIN
                    refused += 1
                    if attempt % 2:
                        srv.drain()
                    else:
                        srv.poll()
            else:
                raise AssertionError(f"request {k} of {q} queries never fitted the ring")
            reqs.append((a, lr))
            if k % 3 == 0:
                srv.poll()
        srv.drain()
        assert refused > 0 and srv.stats()["completed"] == 402
        for (a, lr), (_, v, x) in zip(reqs, out):
            ref = oracle.cpu_base_index(arrays[a], lr)
            assert np.array_equal(v.view(np.uint32), arrays[a][ref].view(np.uint32))
            if x is not None:
                assert np.array_equal(x, ref)
