"""mlz_zp_gemm_noalloc (the gemm entry point that allocates nothing) against
mlz_zp_gemm on the same device inputs, bit for bit, for every kernel family
and storage: two calls in a row, the second smaller, so the second reuses the
static scratch the first filled.  A request beyond the scratch is rejected
before any launch and leaves C unchanged.  Calls go through the C ABI with
device pointers."""
import ctypes

import numpy as np
import pytest

import lela_amd as la
import oracle_lib as orc
from lela_amd import _lib, synth

pytestmark = pytest.mark.gpu

FAMILIES = [("int8", la.ALGO_INT8), ("fp64", la.ALGO_FP64), ("plain", la.ALGO_PLAIN), ("auto", la.ALGO_AUTO)]


def _call(ctx, fn, F, a, dA, dB, b, dC, m, n, k, algo):
    return fn(ctx.handle, F.elem, F.p, m, n, k, a, dA.ptr, dA.ld, dB.ptr, dB.ld, b, dC.ptr, dC.ld, algo)


def _pair(ctx, F, m, n, k, seed, algo):
    """Same inputs through both entry points; returns (noalloc C, mlz_zp_gemm C, oracle C or None)."""
    p = F.p
    A = synth.zp_matrix(seed, m, k, p)
    B = synth.zp_matrix(seed + 1, k, n, p)
    C = synth.zp_matrix(seed + 2, m, n, p)
    a, b = 3 % p, (p - 2) % p
    dA, dB = la.DeviceMatrix.from_host(ctx, F.from_canonical(A)), la.DeviceMatrix.from_host(ctx, F.from_canonical(B))
    dC1, dC2 = (la.DeviceMatrix.from_host(ctx, F.from_canonical(C)) for _ in range(2))
    try:
        ctx.set_ptr_mode(la.PTR_DEVICE)
        rc = _call(ctx, ctx.lib.mlz_zp_gemm_noalloc, F, a, dA, dB, b, dC1, m, n, k, algo)
        ctx.check(rc, "gemm_noalloc")
        rc = _call(ctx, ctx.lib.mlz_zp_gemm, F, a, dA, dB, b, dC2, m, n, k, algo)
        ctx.check(rc, "gemm")
        ctx.synchronize()
        got, ref = F.to_canonical(dC1.download()), F.to_canonical(dC2.download())
    finally:
        for d in (dA, dB, dC1, dC2):
            d.free()
        ctx.set_ptr_mode(la.PTR_HOST)
    want = orc.zp_gemm(p, a, A, B, b, C) if m * n * k <= 1 << 26 else None
    return got, ref, want


@pytest.mark.parametrize("family,algo", FAMILIES)
@pytest.mark.parametrize("storage,p", [("uint16", 65521), ("uint32", 65521), ("double", 67108859),
                                       ("uint32", 4294967291)])
def test_noalloc_matches_gemm_twice(gpu_ctx, family, algo, storage, p):
    if family == "int8" and p > 1 << 16:
        pytest.skip("INT8 limbs need p <= 2^16")
    if family == "fp64" and p >= 1 << 26:
        pytest.skip("FP64 carrier needs p < 2^26")
    F = la.Modular(p, storage)
    gpu_ctx.F = F
    for (m, n, k), seed in [((1000, 900, 1100), 71), ((300, 200, 257), 81)]:
        print(f"noalloc {family} {storage} p={p} {m}x{n}x{k}", flush=True)
        got, ref, want = _pair(gpu_ctx, F, m, n, k, seed, algo)
        assert np.array_equal(got, ref), (family, storage, p, m, n, k)
        if want is not None:
            assert np.array_equal(got, want), (family, storage, p, m, n, k)


def test_noalloc_largest_int8_fits(gpu_ctx):
    """(4096 + 4096) x 4096 x 2 bytes of packed limbs fill the 64 MiB exactly."""
    p = 65521
    F = la.Modular(p, "uint16")
    gpu_ctx.F = F
    print("noalloc int8 4096^3", flush=True)
    got, ref, _ = _pair(gpu_ctx, F, 4096, 4096, 4096, 91, la.ALGO_INT8)
    assert np.array_equal(got, ref)


def test_noalloc_rejects_beyond_scratch(gpu_ctx):
    """A product whose packed operands exceed the scratch returns MLZ_E_ARG and leaves C as it was."""
    p = 65521
    F = la.Modular(p, "uint16")
    gpu_ctx.F = F
    m, n, k = 4096, 4224, 4096          # n pads to 4352: (4096 + 4352) * 4096 * 2 > 64 MiB
    dA = la.DeviceMatrix(gpu_ctx, m, k, np.uint16)
    dB = la.DeviceMatrix(gpu_ctx, k, n, np.uint16)
    C = synth.zp_matrix(5, m, n, p, np.uint16)
    dC = la.DeviceMatrix.from_host(gpu_ctx, C)
    try:
        rc = _call(gpu_ctx, gpu_ctx.lib.mlz_zp_gemm_noalloc, F, 1, dA, dB, 1, dC, m, n, k, la.ALGO_INT8)
        assert rc == _lib.MLZ_E_ARG
        assert "scratch" in gpu_ctx.last_error()
        gpu_ctx.synchronize()
        assert np.array_equal(dC.download(), C)
    finally:
        for d in (dA, dB, dC):
            d.free()
        gpu_ctx.set_ptr_mode(la.PTR_HOST)
