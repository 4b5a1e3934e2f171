"""The allocation-free brute-force matcher (orbfe_match_bf_noalloc_async,
static device scratch) against the allocating entry point
(orbfe_match_bf_async), through the C ABI: identical (idx, d1, d2) on two
calls in a row that share the static scratch, the second smaller; searches
beyond the scratch are refused with ORBFE_ERANGE before any launch."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from uni_slam_amd import orbfe, synth

pytestmark = pytest.mark.gpu


def _run(torch, ctx, fn, q, t):
    nq, nt = len(q), len(t)
    dq = torch.from_numpy(q).cuda()
    dt = torch.from_numpy(t).cuda()
    out = [torch.full((nq,), 7, dtype=torch.int32, device="cuda") for _ in range(3)]
    fn(dq, nq, dt, nt, *out, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    return [o.cpu().numpy().view(np.uint32) for o in out]


def test_noalloc_matches_allocating_entry_point(torch_gpu):
    torch = torch_gpu
    with orbfe.Context(64, 48) as ctx:
        # 65536 x 262144 fills the static scratch (32 splits), 65536 x 65536
        # half of it (16 splits); each smaller call leaves the previous
        # call's partials beyond its own in the array, which it must not
        # read.
        for nq, nt, seed in ((65536, 262144, 20), (65536, 65536, 21), (3000, 20000, 22)):
            q, t, _, _ = synth.descriptors(seed, nq, nt)
            print(f"noalloc vs allocating matcher {nq}x{nt}", flush=True)
            want = _run(torch, ctx, ctx.match_bf_async, q, t)
            got = _run(torch, ctx, ctx.match_bf_noalloc_async, q, t)
            for a, b in zip(got, want):
                assert np.array_equal(a, b), (nq, nt)
            if nq * nt <= 2 ** 26:
                ref = O.match_bf(q, t)
                for a, b in zip(got, ref):
                    assert np.array_equal(a, b), (nq, nt)


def test_noalloc_refuses_searches_beyond_its_scratch(torch_gpu):
    with orbfe.Context(64, 48) as ctx:
        null = C.c_void_p(8)  # never dereferenced: the size check comes first
        # 65536 queries x ceil(300000 / 8192) = 37 splits > 2^21 entries
        rc = orbfe.lib().orbfe_match_bf_noalloc_async(ctx._h, null, 65536, null, 300000, null,
                                                      null, null, None)
        assert rc == orbfe.ERANGE
        assert b"static scratch" in orbfe.lib().orbfe_last_error(ctx._h)
        assert 65536 * -(-262144 // 8192) * 8 == orbfe.NOALLOC_SCRATCH_BYTES  # the limit
