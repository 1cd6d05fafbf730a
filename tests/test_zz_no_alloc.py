"""The allocation-free star product (vrc_star_product_noalloc) against vrc_star_product.

Both run the same kernels; the no-alloc entry point takes its temporaries from the
library's fixed per-GPU scratch array instead of the context's pool. Checked through the
C ABI on two calls in a row, the second smaller than the first (the array keeps the first
call's contents), with the context's pool staying empty, and a request above the array's
size is refused before any work is queued.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_TOL = 1e-12  # same kernels on the same inputs: agreement to FP64 rounding


def random_pair(N, B, seed):
    """S and s as (4, B, N, N) complex128 blocks [Tdd, Tuu, Rud, Rdu]; reflection blocks
    scaled so that I - Rud rdu is well conditioned."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)

    def blk(scale):
        re = torch.randn(B, N, N, generator=g, dtype=torch.float64)
        im = torch.randn(B, N, N, generator=g, dtype=torch.float64)
        return (torch.complex(re, im) * (scale / np.sqrt(2 * N))).cuda()

    S = torch.stack([blk(1.0), blk(1.0), blk(0.3), blk(0.3)])
    s = torch.stack([blk(1.0), blk(1.0), blk(0.3), blk(0.3)])
    return S, s


def call(ctx, fn, S, s):
    from varrcwa_amd._lib import ptr
    _, B, N, _ = S.shape
    return fn(ctx.h, N, B, ptr(S[0]), ptr(S[1]), ptr(S[2]), ptr(S[3]), ptr(s[0]), ptr(s[1]), ptr(s[2]), ptr(s[3]),
              N * N)


def rel_err(a, b):
    return float((a - b).abs().max() / b.abs().max())


def test_noalloc_star_product_matches_pooled_entry_point():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from varrcwa_amd._lib import Context
    ctx = Context(0)  # a fresh context: its pool starts empty
    try:
        lib = ctx.lib
        assert lib.vrc_pool_bytes(ctx.h) == 0
        for N, B, seed in [(882, 1, 11), (242, 3, 12)]:
            print(f"no-alloc star product N={N} batch={B} seed={seed}", flush=True)
            S, s = random_pair(N, B, seed)
            S_na = S.clone()
            before = lib.vrc_pool_bytes(ctx.h)
            rc = call(ctx, lib.vrc_star_product_noalloc, S_na, s)
            torch.cuda.synchronize()
            assert rc == 0, lib.vrc_last_error(ctx.h).decode()
            # nothing was taken from the context's pool
            assert lib.vrc_pool_bytes(ctx.h) == before
            S_ref = S.clone()
            assert call(ctx, lib.vrc_star_product, S_ref, s) == 0, lib.vrc_last_error(ctx.h).decode()
            torch.cuda.synchronize()
            assert bool(S_na.isfinite().all())
            for k in range(4):
                assert rel_err(S_na[k], S_ref[k]) <= REL_TOL, (N, B, k, rel_err(S_na[k], S_ref[k]))
            # the product changed S (it is not a no-op on these inputs)
            assert rel_err(S_na[0], S[0]) > 1e-3
            pooled = lib.vrc_pool_bytes(ctx.h)
            # a second no-alloc call on the same sizes leaves the pool as it is
            S_na2 = S.clone()
            assert call(ctx, lib.vrc_star_product_noalloc, S_na2, s) == 0
            torch.cuda.synchronize()
            assert lib.vrc_pool_bytes(ctx.h) == pooled
            for k in range(4):
                assert rel_err(S_na2[k], S_ref[k]) <= REL_TOL
        # above the array's size: refused with the sizes in the message, inputs untouched
        S, s = random_pair(882, 2, 13)
        keep = S.clone()
        rc = call(ctx, lib.vrc_star_product_noalloc, S, s)
        msg = lib.vrc_last_error(ctx.h).decode()
        assert rc == -1 and "scratch" in msg and "882" in msg, (rc, msg)
        torch.cuda.synchronize()
        assert torch.equal(S, keep)
        # bad arguments
        assert lib.vrc_star_product_noalloc(ctx.h, 0, 1, *([None] * 8), 0) == -1
        assert lib.vrc_star_product_noalloc(None, 4, 1, *([None] * 8), 0) == -1
    finally:
        ctx.close()
