"""Serving loop (include/rmq_serve.h): resident hierarchies, free-memory admission, continuous
batching. Every answer must equal the oracle's, and a request must get the same answer alone
and when batched with others."""

import ctypes

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
    variants = ["lane", "group8", "group4", "group16", "lane8", "wave"]
    with rmq_amd.Server(slots=3, max_batch_queries=1 << 16) as srv:
        ids = [srv.admit_host(A, t=64 << (i % 4), variant=variants[i % 6])
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
