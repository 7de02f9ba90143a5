"""bmg_face_flux_noalloc (staging in the library's fixed device array, no
allocation) against bmg_face_flux (staging in allocated device buffers):
bit for bit, on two calls in a row with the second smaller, with and without
lambda arrays, and the size limit of the fixed array."""
import ctypes

import numpy as np
import pytest

from briscola_amd import workloads as W

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _noalloc(ctx, x, lam):
    out = [np.full(ctx.padded(0), np.nan) for _ in range(3)]
    ls = [None] * 3 if lam is None else [_dp(a) for a in lam]
    code = ctx.lib.bmg_face_flux_noalloc(ctx.ptr, _dp(x), *ls,
                                         *[_dp(a) for a in out])
    return code, out


def _case(gpu, n, seed, with_lam):
    p = W.spheres(n, seed=seed)
    ctx = gpu.Context(p.vertices, p.bc, p.bc_value, rel_tol=1e-8,
                      tolerance=0.0)
    rng = np.random.default_rng(seed)
    x = np.ascontiguousarray(rng.standard_normal(ctx.padded(0)))
    lam = ([np.ascontiguousarray(a) for a in p.lam] if with_lam else None)
    return ctx, x, lam


@pytest.mark.parametrize("with_lam", [True, False])
def test_noalloc_flux_matches_allocating_entry(gpu, with_lam):
    # first call (48 x 32 x 64), then a smaller one (16 x 24 x 8): the
    # second sees the array as the first left it
    for n, seed in (((48, 32, 64), 1), ((16, 24, 8), 2)):
        print(f"[noalloc] {n} lambda={with_lam}", flush=True)
        ctx, x, lam = _case(gpu, n, seed, with_lam)
        ref = ctx.face_flux(x, lam)
        code, out = _noalloc(ctx, x, lam)
        assert code == 0, gpu.load().bmg_last_error()
        for d in range(3):
            assert np.array_equal(out[d], ref[d]), (n, d)
        ctx.close()


def test_noalloc_flux_size_limit(gpu):
    # 96^3: 98^3 padded cells; 4 staged fields fit the 4 194 304 doubles,
    # 7 (with lambda arrays) do not
    print("[noalloc] 96^3 limit", flush=True)
    ctx, x, lam = _case(gpu, 96, 3, True)
    code, _ = _noalloc(ctx, x, lam)
    assert code == -1   # BMG_ERR_ARG
    assert b"staging" in gpu.load().bmg_last_error()
    ref = ctx.face_flux(x, None)
    code, out = _noalloc(ctx, x, None)
    assert code == 0, gpu.load().bmg_last_error()
    for d in range(3):
        assert np.array_equal(out[d], ref[d])
    ctx.close()
