"""whir_dev_sumcheck_round_noalloc, the sumcheck round that allocates
nothing (its partials live in the library's file-scope device scratch), gives
what whir_dev_sumcheck_round gives, through the C ABI, on two calls in a row
with the second one smaller (so the scratch still holds the first call's
partials beyond the second's), for base-field and Ext3 first vectors."""
import ctypes

import numpy as np
import pytest

import oracle
import whirplonk_amd as W

pytestmark = pytest.mark.gpu


def _inputs(n, a_base, seed):
    a, st = W.lcg_uniform(seed, n if a_base else 3 * n)
    b, st = W.lcg_uniform(st, 3 * n)
    r, _ = W.lcg_uniform(st, 3)
    return a, b, r


@pytest.mark.parametrize("a_base", [True, False])
def test_noalloc_round_equals_round(a_base):
    lib = W.load()
    with W.Context() as ctx:
        for call, n in enumerate((1 << 21, 6000)):
            a, b, r = _inputs(n, a_base, 0x4E0A + 7 * call + int(a_base))
            d_a, d_b = ctx.upload(a), ctx.upload(b)
            half = 3 * (n // 2)
            outs = [ctx.alloc(8 * half) for _ in range(4)]
            d_c = ctx.alloc(8 * 6)
            print("noalloc round", "base" if a_base else "ext3", n, flush=True)
            # the existing entry point (allocates its partials and copies c0, c2 to the host)
            c0 = np.zeros(3, dtype=np.uint64)
            c2 = np.zeros(3, dtype=np.uint64)
            W._check(lib.whir_dev_sumcheck_round(ctx.h, d_a, int(a_base), d_b, n, W._p64(r), outs[0], outs[1],
                                                 W._p64(c0), W._p64(c2)))
            # the allocation-free one: c0, c2 stay on the device
            W._check(lib.whir_dev_sumcheck_round_noalloc(ctx.h, d_a, int(a_base), d_b, n, W._p64(r), outs[2],
                                                         outs[3], d_c))
            ctx.sync()
            got = ctx.download(d_c, 6)
            assert np.array_equal(got[:3], c0) and np.array_equal(got[3:], c2), call
            assert np.array_equal(ctx.download(outs[2], half), ctx.download(outs[0], half)), call
            assert np.array_equal(ctx.download(outs[3], half), ctx.download(outs[1], half)), call
            if n <= 8192:  # the small call against the oracle too
                lifted = a
                if a_base:
                    lifted = np.zeros(3 * n, dtype=np.uint64)
                    lifted[0::3] = a
                w0, w2 = oracle.sumcheck_compute_ext3(lifted, b)
                assert np.array_equal(got[:3], w0) and np.array_equal(got[3:], w2)
                assert np.array_equal(ctx.download(outs[3], half), oracle.sumcheck_fold_ext3(b, r))
            for p in (d_a, d_b, d_c, *outs):
                ctx.free(p)


def test_noalloc_round_rejects():
    lib = W.load()
    with W.Context() as ctx:
        d = ctx.alloc(64)
        r = np.zeros(3, dtype=np.uint64)
        # odd length, base-field fold in place, too long, null output: refused before any launch
        assert lib.whir_dev_sumcheck_round_noalloc(ctx.h, d, 1, d, 7, W._p64(r), d + 0, d, d) == 3
        assert lib.whir_dev_sumcheck_round_noalloc(ctx.h, d, 1, d, 2, W._p64(r), d, d, d) == 3
        assert lib.whir_dev_sumcheck_round_noalloc(ctx.h, d, 0, d, (1 << 32) + 2, W._p64(r), d, d, d) == 3
        assert lib.whir_dev_sumcheck_round_noalloc(ctx.h, d, 0, d, 2, W._p64(r), d, d, None) == 3
        ctx.free(d)
