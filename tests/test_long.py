"""The 64-bit form (the reference's IS_LONG build, src/main.cu:21-27): arrays held as chunks, each
its own hierarchy, queried with 64-bit {l, r} pairs through rmq_query64. Small arrays built with a
small chunk size exercise every chunk-crossing case against CPU BASE; one array of 2^31 + 4099
elements runs the real thing."""

import sys

import numpy as np
import pytest

import oracle
import rmq_amd

pytestmark = pytest.mark.gpu


def _say(msg):
    sys.stderr.write(f"[gpu-case] {msg}\n")
    sys.stderr.flush()


def _cpu_ref(A, lr64):
    """CPU BASE through the oracle (32-bit pairs: these arrays are small)."""
    return oracle.cpu_base_index(A, lr64.astype(np.uint32)).astype(np.uint64)


def _chunk_queries(rng, n, chunk):
    b = np.arange(chunk, n, chunk, dtype=np.int64)
    pairs = [(0, n - 1), (0, 0), (n - 1, n - 1)]
    for x in b[: 8].tolist() + b[-4:].tolist():
        pairs += [(x - 1, x), (x - 1, x - 1), (x, x), (x - chunk, x - 1), (x - 3, min(n - 1, x + chunk + 2)),
                  (max(0, x - chunk - 1), min(n - 1, x + chunk))]
    rnd = np.sort(rng.integers(0, n, size=(3000, 2)), axis=1)
    short_l = rng.integers(0, n - 100, size=1000)
    short = np.stack([short_l, short_l + rng.integers(0, 100, size=1000)], axis=1)
    return np.concatenate([np.array(pairs, dtype=np.int64), rnd, short]).astype(np.uint64)


@pytest.mark.parametrize("alg,opts", [(20, {}), (20, dict(cg_amount_log=1)), (16, {}),
                                      (24, dict(group_lanes=4)), (20, dict(cg_amount_log=0))],
                         ids=str)
def test_chunked_small_against_cpu_base(device, alg, opts):
    import torch

    rng = np.random.default_rng(alg * 10 + len(opts))
    n, clog = 100003, 12
    for kind in ("uniform", "ties", "nan"):
        A = rng.random(n, dtype=np.float32)
        if kind == "ties":
            A = rng.integers(0, 3, size=n).astype(np.float32)
            A[rng.random(n) < 0.5] *= np.float32(-1.0)  # +0 and -0 in many chunks
        elif kind == "nan":  # NaN at every chunk start, one chunk all NaN, 1% scattered
            A[:: 1 << clog] = np.nan
            A[3 << clog:4 << clog] = np.nan
            A[rng.random(n) < 0.01] = np.nan
        lr = _chunk_queries(rng, n, 1 << clog)
        ref = _cpu_ref(A, lr)
        dA = torch.from_numpy(A).cuda()
        dQ = torch.from_numpy(lr.view(np.int64)).cuda()
        for t in (64 if alg != 20 else 16, 1 << 15):
            _say(f"chunked alg={alg} opts={opts} kind={kind} t={t}")
            with rmq_amd.Hierarchy(dA, t, alg, chunk_log=clog, **opts) as h:
                v, i = h.query64(dQ, True, True)
                torch.cuda.synchronize()
                with pytest.raises(rmq_amd.RMQError):  # 32-bit queries do not reach a chunked hierarchy
                    h.query(torch.zeros(2, dtype=torch.int32, device="cuda"), True, False)
            i = i.cpu().numpy().view(np.uint64)
            v = v.cpu().numpy()
            bad = np.nonzero(i != ref)[0]
            assert bad.size == 0, (kind, t, lr[bad[:3]].tolist(), i[bad[:3]], ref[bad[:3]])
            assert np.array_equal(v.view(np.uint32), A[ref.astype(np.int64)].view(np.uint32))


def test_many_chunks(device):
    """391 chunks of 2^8: more than the query kernel copies into LDS, so the chunks' level tables
    are read from their views (the other form of the kernel)."""
    import torch

    rng = np.random.default_rng(391)
    n, clog = 100003, 8
    A = rng.integers(0, 5, size=n).astype(np.float32)
    lr = _chunk_queries(rng, n, 1 << clog)
    ref = _cpu_ref(A, lr)
    dA = torch.from_numpy(A).cuda()
    dQ = torch.from_numpy(lr.view(np.int64)).cuda()
    for alg, opts, t in ((20, dict(cg_amount_log=0), 16), (16, {}, 64)):
        _say(f"many chunks alg={alg} opts={opts} t={t}")
        with rmq_amd.Hierarchy(dA, t, alg, chunk_log=clog, **opts) as h:
            v, i = h.query64(dQ, True, True)
            torch.cuda.synchronize()
        i = i.cpu().numpy().view(np.uint64)
        bad = np.nonzero(i != ref)[0]
        assert bad.size == 0, (alg, lr[bad[:3]].tolist(), i[bad[:3]], ref[bad[:3]])
        assert np.array_equal(v.cpu().numpy().view(np.uint32), A[ref.astype(np.int64)].view(np.uint32))


def test_query64_on_one_piece_equals_query(device):
    import torch

    n = 77777
    A = rmq_amd.gen_array(n, 5)
    lr = rmq_amd.gen_queries(5000, -6, n, 5)
    dA = torch.from_numpy(A).cuda()
    with rmq_amd.Hierarchy(dA, 1024, 20, cg_amount_log=1) as h:
        v32, i32 = h.query(torch.from_numpy(lr.view(np.int32)).cuda(), True, True)
        v64, i64 = h.query64(torch.from_numpy(lr.astype(np.uint64).view(np.int64)).cuda(), True, True)
        torch.cuda.synchronize()
    assert torch.equal(v32.view(torch.int32), v64.view(torch.int32))
    assert torch.equal(i32.to(torch.int64), i64)


def test_gpu_base64(device):
    import torch

    rng = np.random.default_rng(8)
    n = 5000
    A = rng.integers(0, 4, size=n).astype(np.float32)
    lr = np.sort(rng.integers(0, n, size=(2000, 2)), axis=1).astype(np.uint64)
    v, i = rmq_amd.gpu_base64(torch.from_numpy(A).cuda(), torch.from_numpy(lr.view(np.int64)).cuda(),
                              True, True)
    torch.cuda.synchronize()
    ref = _cpu_ref(A, lr)
    assert np.array_equal(i.cpu().numpy().view(np.uint64), ref)


def test_array_above_2e31(device):
    """n = 2^31 + 4099 (8.6 GB): queries from the IS_LONG generator (mixed lengths, up to the
    whole array); short ones against GPU BASE, long ones spanning the 2^30 / 2^31 chunk edges
    against torch.argmin (first occurrence), every answer inside its range with value ==
    A[index]."""
    import torch

    n = (1 << 31) + 4099
    _say("long: fill")
    dA = torch.empty(n, dtype=torch.float32, device="cuda")
    rmq_amd.fill_uniform(dA, seed=31)  # 24-bit values: every long range holds ties at its minimum
    lr = rmq_amd.gen_queries64(1 << 16, -6, n, 27722)
    extra = np.array([[(1 << 30) - 5, (1 << 30) + 5], [(1 << 31) - 1, 1 << 31], [0, n - 1],
                      [(1 << 30) - 1000, (1 << 31) + 1000], [n - 10, n - 1], [1 << 31, n - 1]],
                     dtype=np.uint64)
    lr = np.concatenate([extra, lr])
    dQ = torch.from_numpy(lr.view(np.int64)).cuda()
    _say("long: build + query")
    with rmq_amd.Hierarchy(dA, 1 << 13, 20, cg_amount_log=1) as h:
        assert h.metadata().sizes[0] == n
        v, i = h.query64(dQ, True, True)
        torch.cuda.synchronize()
    ih = i.cpu().numpy().view(np.uint64)
    assert (ih >= lr[:, 0]).all() and (ih <= lr[:, 1]).all()
    assert torch.equal(v.view(torch.int32), dA[i].view(torch.int32))
    short = np.nonzero((lr[:, 1] - lr[:, 0]) < (1 << 20))[0]
    _say(f"long: GPU BASE on {short.size} short queries")
    sq = torch.from_numpy(np.ascontiguousarray(lr[short]).view(np.int64)).cuda()
    _, bi = rmq_amd.gpu_base64(dA, sq, False, True)
    torch.cuda.synchronize()
    assert np.array_equal(bi.cpu().numpy().view(np.uint64), ih[short])
    _say("long: torch.argmin (first occurrence) on the long edge queries")
    for k in (0, 2, 3, 5):
        l, r = int(lr[k, 0]), int(lr[k, 1])
        assert int(ih[k]) == l + int(torch.argmin(dA[l:r + 1])), (k, l, r)
