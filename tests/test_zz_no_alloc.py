"""rmq_batch_query_noalloc: the batched query whose work queue (tiles, queue starts, counters)
lives in the library's static device array, so that the call allocates nothing. Its answers
equal rmq_batch_plan_queries + rmq_batch_query bit for bit on two calls in a row (the second
with a smaller layout), in value, index and value + index mode, with and without the group
solvers; a graph captured around it replays correctly; sizes past its limits are refused
before anything runs."""

import sys

import numpy as np
import pytest

import rmq_amd
from rmq_amd import workloads

pytestmark = pytest.mark.gpu

MAX_ARRAYS = 65536
MAX_TILES = 262144


def _say(msg):
    sys.stderr.write(f"[gpu-case] {msg}\n")
    sys.stderr.flush()


def _noalloc(b, dqoff, total, dQ, out_v, out_i, stream=None):
    return rmq_amd.lib().rmq_batch_query_noalloc(
        b._b, dqoff.data_ptr(), total, dQ.data_ptr(),
        out_v.data_ptr() if out_v is not None else None,
        out_i.data_ptr() if out_i is not None else None, stream)


def _workload(seed, arrays):
    import torch

    w = workloads.batch_workload(arrays=arrays, log_n=(4, 18), log_q=(0, 12), kind="mixed", seed=seed)
    total = int(w.offsets[-1] + w.sizes[-1] + 64)
    dA = torch.empty(total, dtype=torch.float32, device="cuda")
    rmq_amd.fill_uniform(dA, seed=seed)
    A = dA.cpu().numpy()
    rng = np.random.default_rng(seed)
    for a in range(0, arrays, 5):  # ties and both signs of zero in some arrays
        o, n = int(w.offsets[a]), int(w.sizes[a])
        A[o:o + n] = rng.integers(0, 3, size=n).astype(np.float32) * \
            np.where(rng.random(n) < 0.5, np.float32(-1), np.float32(1))
    dA.copy_(torch.from_numpy(A))
    torch.cuda.synchronize()
    return w, dA


def _half_layout(w):
    """Every array keeps the first half of its queries: a smaller layout over the same arrays."""
    keep = w.qcounts // 2
    qoff = np.concatenate([[0], np.cumsum(keep)]).astype(np.int64)
    lr = np.concatenate([w.queries[w.qoffsets[a]:w.qoffsets[a] + keep[a]] for a in range(len(keep))])
    return qoff, np.ascontiguousarray(lr)


@pytest.mark.parametrize("variants", [("lane", "lane8", "lane4", "group8", "group4", "wave"),
                                      ("lane", "lane8", "lane4")], ids=["all", "lanes"])
def test_noalloc_equals_planned_query(device, variants):
    import torch

    arrays = 150
    w, dA = _workload(41, arrays)
    var = [variants[a % len(variants)] for a in range(arrays)]
    ts = [64 << (a % 5) for a in range(arrays)]
    qoff2, lr2 = _half_layout(w)
    layouts = [(w.qoffsets.astype(np.int64), w.queries), (qoff2, lr2)]
    with rmq_amd.Batch(dA, w.offsets, w.sizes, ts, var) as b:
        for call, (qoff, lr) in enumerate(layouts):  # the second call holds fewer queries
            Q = int(qoff[-1])
            dQ = torch.from_numpy(np.ascontiguousarray(lr).view(np.int32)).cuda()
            plan = b.plan(qoff)
            pv, pi = plan.query(dQ, True, True)
            dqo = torch.from_numpy(qoff).cuda()
            for want_v, want_i in ((True, True), (True, False), (False, True)):
                _say(f"noalloc {variants} call {call} Q={Q} value={want_v} index={want_i}")
                nv = torch.full((Q,), -1.0, device="cuda") if want_v else None
                ni = torch.full((Q,), -1, dtype=torch.int32, device="cuda") if want_i else None
                assert _noalloc(b, dqo, Q, dQ, nv, ni) == rmq_amd.RMQ_OK
                torch.cuda.synchronize()
                if want_v:
                    assert torch.equal(nv.view(torch.int32), pv.view(torch.int32)), (call, want_v, want_i)
                if want_i:
                    assert torch.equal(ni, pi), (call, want_v, want_i)


def test_noalloc_graph_replays(device):
    """The call captured in a HIP graph (every step is a kernel launch) replays correctly several
    times, with eager work and synchronisations between the replays."""
    import torch

    arrays = 120
    w, dA = _workload(43, arrays)
    var = [("lane8", "lane4", "lane")[a % 3] for a in range(arrays)]
    ts = [64 << (a % 4) for a in range(arrays)]
    Q = int(w.qoffsets[-1])
    dQ = torch.from_numpy(w.queries.view(np.int32)).cuda()
    dqo = torch.from_numpy(w.qoffsets.astype(np.int64)).cuda()
    with rmq_amd.Batch(dA, w.offsets, w.sizes, ts, var) as b:
        ev, ei = b.plan(w.qoffsets).query(dQ, True, True)
        gv = torch.full((Q,), -1.0, device="cuda")
        gi = torch.full((Q,), -1, dtype=torch.int32, device="cuda")
        assert _noalloc(b, dqo, Q, dQ, gv, gi) == rmq_amd.RMQ_OK  # first call on the device: eager
        torch.cuda.synchronize()
        s = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            st = _noalloc(b, dqo, Q, dQ, gv, gi, stream=s.cuda_stream)
        assert st == rmq_amd.RMQ_OK
        for rep in range(4):
            gv.fill_(-1.0)
            gi.fill_(-1)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            _say(f"noalloc graph replay {rep}")
            assert torch.equal(gv.view(torch.int32), ev.view(torch.int32)), rep
            assert torch.equal(gi, ei), rep
            junk = torch.rand(1 << 20, device="cuda").sum()  # eager work between replays
            torch.cuda.synchronize()
            assert float(junk) > 0
        del g


def test_noalloc_limits(device):
    """Calls past the limits are refused with RMQ_EINVAL before anything runs; zero queries is a
    no-op; a misaligned offset array is refused."""
    import torch

    arrays = 40
    w, dA = _workload(47, arrays)
    dQ = torch.from_numpy(w.queries.view(np.int32)).cuda()
    dqo = torch.from_numpy(w.qoffsets.astype(np.int64)).cuda()
    out = torch.empty(int(w.qoffsets[-1]), dtype=torch.int32, device="cuda")
    with rmq_amd.Batch(dA, w.offsets, w.sizes, [64] * arrays, ["lane8"] * arrays) as b:
        too_many = (MAX_TILES - arrays + 1) * 256  # total / 256 + count = MAX_TILES + 1
        assert _noalloc(b, dqo, too_many, dQ, None, out) == rmq_amd.RMQ_EINVAL
        assert _noalloc(b, dqo, 0, dQ, None, out) == rmq_amd.RMQ_OK
        L = rmq_amd.lib()
        assert L.rmq_batch_query_noalloc(b._b, dqo.data_ptr() + 4, int(w.qoffsets[-1]), dQ.data_ptr(), None,
                                         out.data_ptr(), None) == rmq_amd.RMQ_EINVAL
        assert L.rmq_batch_query_noalloc(b._b, dqo.data_ptr(), int(w.qoffsets[-1]), dQ.data_ptr(), None,
                                         None, None) == rmq_amd.RMQ_EINVAL
        torch.cuda.synchronize()
    # more arrays than the static queue holds
    n = MAX_ARRAYS + 1
    small = torch.empty(n * 16, dtype=torch.float32, device="cuda")
    rmq_amd.fill_uniform(small, seed=3)
    with rmq_amd.Batch(small, np.arange(n) * 16, [16] * n, [64] * n, ["lane"] * n) as b:
        qo = torch.zeros(n + 1, dtype=torch.int64, device="cuda")
        qo[1:] = 1
        one = torch.zeros((1, 2), dtype=torch.int32, device="cuda")
        o1 = torch.empty(1, dtype=torch.int32, device="cuda")
        assert _noalloc(b, qo, 1, one, None, o1) == rmq_amd.RMQ_EINVAL
    torch.cuda.synchronize()
