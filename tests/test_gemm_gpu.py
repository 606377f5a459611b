"""GPU parity of mlz_zp_gemm against the oracle (bit-exact), all three kernel
families, all storages, ragged shapes, strided views, k-chunking, and the
full 8192^3 workload through size-independent checks (Freivalds, cross-kernel
equality)."""
import numpy as np
import pytest

import lela_amd as la
import oracle_lib as orc
from lela_amd import synth

pytestmark = pytest.mark.gpu

ALGOS = {"plain": la.ALGO_PLAIN, "int8": la.ALGO_INT8, "fp64": la.ALGO_FP64}


def _valid(algo, storage, p):
    if storage == "uint16" and p >= 1 << 16:
        return False
    if storage == "double" and p >= 1 << 26:
        return False
    if algo == "int8" and p > 1 << 16:
        return False
    if algo == "fp64" and p >= 1 << 26:
        return False
    return True


def run_gemm(ctx, p, storage, algo, a, A, B, b, C):
    F = la.Modular(p, storage)
    ctx.F = F
    Cs = F.from_canonical(C)
    la.BLAS3.gemm(ctx, a, F.from_canonical(A), F.from_canonical(B), b, Cs, algo=ALGOS[algo])
    return F.to_canonical(Cs)


@pytest.mark.parametrize("algo", ["plain", "int8", "fp64"])
@pytest.mark.parametrize("storage", ["uint16", "uint32", "double"])
@pytest.mark.parametrize("p", [2, 3, 251, 32749, 65521, 65535, 67108859, 2147483647, 4294967291])
def test_gemm_matches_oracle(gpu_ctx, algo, storage, p):
    if not _valid(algo, storage, p):
        pytest.skip("modulus outside this kernel/storage")
    print(f"gemm {algo} {storage} p={p}", flush=True)
    for (m, n, k), seed in [((37, 45, 53), 1), ((128, 128, 64), 2), ((200, 130, 300), 3),
                            ((1, 1, 1), 4)]:
        A, B, C = (synth.zp_matrix(seed * 10 + i, *s, p) for i, s in
                   enumerate([(m, k), (k, n), (m, n)]))
        for a, b in [(1, 1), (p - 1, 3 % p), (5 % p, 0)]:
            got = run_gemm(gpu_ctx, p, storage, algo, a, A, B, b, C)
            want = orc.zp_gemm(p, a, A, B, b, C)
            assert np.array_equal(got, want), (algo, storage, p, m, n, k, a, b)


@pytest.mark.parametrize("algo", ["plain", "int8", "fp64"])
def test_gemm_edge_shapes(gpu_ctx, algo):
    p = 65521
    for (m, n, k) in [(0, 5, 3), (5, 0, 3), (4, 6, 0), (1, 300, 1), (300, 1, 257), (65, 129, 127)]:
        A, B, C = synth.zp_matrix(1, m, k, p), synth.zp_matrix(2, k, n, p), synth.zp_matrix(3, m, n, p)
        got = run_gemm(gpu_ctx, p, "uint32", algo, 3, A, B, 7, C)
        want = orc.zp_gemm(p, 3, A, B, 7, C)
        assert np.array_equal(got, want), (algo, m, n, k)


@pytest.mark.parametrize("algo", ["plain", "int8", "fp64"])
def test_gemm_strided_views(gpu_ctx, algo):
    """Sub-matrix views with a parent stride, as LELA passes them (dense.h:425-429)."""
    p = 65521
    F = la.Modular(p, "uint32")
    gpu_ctx.F = F
    big = synth.zp_matrix(9, 300, 300, p)
    A = big[10:110, 5:77]          # 100 x 72, ld 300
    B = big[120:192, 40:131]       # 72 x 91
    Cp = synth.zp_matrix(8, 140, 200, p)
    C = Cp[20:120, 100:191]        # 100 x 91 view, writes in place
    want = orc.zp_gemm(p, 1, A, B, 1, C)
    outside = Cp.copy()
    la.BLAS3.gemm(gpu_ctx, 1, A, B, 1, C, algo=ALGOS[algo])
    assert np.array_equal(C, want)
    mask = np.ones_like(Cp, dtype=bool)
    mask[20:120, 100:191] = False
    assert np.array_equal(Cp[mask], outside[mask])


@pytest.mark.parametrize("variant", list(range(1, 12)))
@pytest.mark.parametrize("storage", ["uint16", "uint32", "double"])
def test_gemm_int8_tile_variants(gpu_ctx, variant, storage):
    """Every INT8 workgroup-tile variant gives the same exact product.  The k
    values give 2 .. 32 k-steps, so the multi-buffered variants (7-10) run
    with fewer steps than buffers, whole groups only, and groups plus every
    possible remainder."""
    p = 65521
    gpu_ctx.set_option("gemm.i8.variant", variant)
    try:
        for (m, n, k) in [(300, 517, 130), (129, 257, 64), (1, 1, 33), (200, 300, 1000), (64, 64, 448),
                          (130, 260, 96)]:
            A, B, C = synth.zp_matrix(41, m, k, p), synth.zp_matrix(42, k, n, p), synth.zp_matrix(43, m, n, p)
            got = run_gemm(gpu_ctx, p, storage, "int8", 7, A, B, 3, C)
            assert np.array_equal(got, orc.zp_gemm(p, 7, A, B, 3, C)), (variant, m, n, k)
    finally:
        gpu_ctx.set_option("gemm.i8.variant", 0)


def test_gemm_int8_k_chunks(gpu_ctx):
    # k > 65504 splits the int32-exact accumulation into chunks through C.
    p = 65521
    m, n, k = 8, 9, 70001
    A, B, C = synth.zp_matrix(11, m, k, p), synth.zp_matrix(12, k, n, p), synth.zp_matrix(13, m, n, p)
    got = run_gemm(gpu_ctx, p, "uint16", "int8", 1, A, B, 1, C)
    assert np.array_equal(got, orc.zp_gemm(p, 1, A, B, 1, C))
    # extreme representatives: 32639 (largest kept as is) times 32640 (the
    # smallest moved to 32640 - p), over a k that needs two chunks
    A2 = np.full((4, 65600), 32639, dtype=np.uint32)
    B2 = np.full((65600, 4), 32640, dtype=np.uint32)
    C2 = np.zeros((4, 4), dtype=np.uint32)
    got = run_gemm(gpu_ctx, p, "uint32", "int8", 1, A2, B2, 0, C2)
    assert np.array_equal(got, orc.zp_gemm(p, 1, A2, B2, 0, C2))


def test_gemm_fp64_k_chunks(gpu_ctx):
    p = 67108859
    m, n, k = 5, 7, 70003
    A, B, C = synth.zp_matrix(14, m, k, p), synth.zp_matrix(15, k, n, p), synth.zp_matrix(16, m, n, p)
    for storage in ("uint32", "double"):
        got = run_gemm(gpu_ctx, p, storage, "fp64", 1, A, B, 1, C)
        assert np.array_equal(got, orc.zp_gemm(p, 1, A, B, 1, C)), storage


def test_gemm_device_pointers(gpu_ctx):
    p = 65521
    F = la.Modular(p, "uint16")
    gpu_ctx.F = F
    A, B, C = (synth.zp_matrix(20 + i, 96, 96, p, np.uint16) for i in range(3))
    dA, dB, dC = (la.DeviceMatrix.from_host(gpu_ctx, X) for X in (A, B, C))
    try:
        la.BLAS3.gemm(gpu_ctx, 2, dA, dB, 3, dC, algo=la.ALGO_INT8)
        gpu_ctx.synchronize()
        got = dC.download()
    finally:
        for d in (dA, dB, dC):
            d.free()
        gpu_ctx.set_ptr_mode(la.PTR_HOST)
    want = orc.zp_gemm(p, 2, A, B, 3, C, storage="uint16")
    assert np.array_equal(got, want)


def matvec_mod(M, x, p):
    """(M @ x) mod p exactly in uint64 for M entries < p and x < p: x is split
    into limbs narrow enough that no partial sum reaches 2^64."""
    P = np.uint64(p)
    bits = max(1, int(p - 1).bit_length())
    kbits = max(1, int(M.shape[1]).bit_length())
    lb = max(1, 63 - bits - kbits)
    Mu = M.astype(np.uint64)
    out = np.zeros(M.shape[0], dtype=np.uint64)
    x = x.astype(np.uint64)
    shift = 0
    while shift < bits:
        limb = (x >> np.uint64(shift)) & np.uint64((1 << lb) - 1)
        part = (Mu @ limb) % P
        out = (out + part * np.uint64(pow(2, shift, p))) % P
        shift += lb
    return out


def _freivalds(A, B, C0, Cnew, p, a, b, seed):
    """Checks Cnew == a*A*B + b*C0 (mod p) on random vectors: exact uint64 math."""
    rng = np.random.default_rng(seed)
    P = np.uint64(p)
    for _ in range(2):
        x = rng.integers(0, p, size=B.shape[1], dtype=np.uint64)
        ABx = matvec_mod(A, matvec_mod(B, x, p), p)
        lhs = matvec_mod(Cnew, x, p)
        rhs = (np.uint64(a) * ABx + np.uint64(b) * matvec_mod(C0, x, p)) % P
        assert np.array_equal(lhs, rhs)


def test_matvec_mod_is_exact():
    """The limb split of matvec_mod against Python integers, at the largest entries."""
    p = 4294967291
    rng = np.random.default_rng(3)
    M = rng.integers(p - 1000, p, size=(3, 9000), dtype=np.uint64)
    x = rng.integers(p - 1000, p, size=9000, dtype=np.uint64)
    want = [sum(int(a) * int(b) for a, b in zip(row, x)) % p for row in M]
    assert [int(v) for v in matvec_mod(M, x, p)] == want


@pytest.mark.parametrize("algo", ["plain", "int8", "fp64"])
def test_gemm_512_cubed_uint32(gpu_ctx, algo):
    """The reference's CPU-runnable case: C <- A*B + C at 512^3 over
    Modular<uint32>, p = 65521, on its seeded inputs, bit-exact against the
    oracle's ZpModule<uint32> product (level3-generic.tcc:285-322)."""
    p = 65521
    A = synth.zp_matrix(synth.seed_for(1, 0), 512, 512, p)
    B = synth.zp_matrix(synth.seed_for(1, 1), 512, 512, p)
    C = synth.zp_matrix(synth.seed_for(1, 2), 512, 512, p)
    got = run_gemm(gpu_ctx, p, "uint32", algo, 1, A, B, 1, C)
    assert np.array_equal(got, orc.zp_gemm(p, 1, A, B, 1, C))


def test_gemm_int8_full_size_8192(gpu_ctx):
    """The 8192^3, p=65521, uint16 workload: Freivalds on the result, and the
    int8 and plain kernels agree on a 1024-row slice."""
    p, n = 65521, 8192
    F = la.Modular(p, "uint16")
    gpu_ctx.F = F
    A = synth.zp_matrix(synth.seed_for(2, 0), n, n, p, np.uint16)
    B = synth.zp_matrix(synth.seed_for(2, 1), n, n, p, np.uint16)
    C0 = synth.zp_matrix(synth.seed_for(2, 2), n, n, p, np.uint16)
    C = C0.copy()
    la.BLAS3.gemm(gpu_ctx, 1, A, B, 1, C, algo=la.ALGO_INT8)
    _freivalds(A, B, C0, C, p, 1, 1, seed=5)
    Cs = C0[:1024].copy()
    la.BLAS3.gemm(gpu_ctx, 1, A[:1024], B, 1, Cs, algo=la.ALGO_PLAIN)
    assert np.array_equal(Cs, C[:1024])


def test_gemm_fp64_large_prime_2048(gpu_ctx):
    p, n = 67108859, 2048
    F = la.Modular(p, "uint32")
    gpu_ctx.F = F
    A, B, C0 = (synth.zp_matrix(30 + i, n, n, p) for i in range(3))
    C = C0.copy()
    la.BLAS3.gemm(gpu_ctx, 1, A, B, 1, C, algo=la.ALGO_FP64)
    _freivalds(A, B, C0, C, p, 1, 1, seed=6)
    Cs = C0[:256].copy()
    la.BLAS3.gemm(gpu_ctx, 1, A[:256], B, 1, Cs, algo=la.ALGO_PLAIN)
    assert np.array_equal(Cs, C[:256])


@pytest.mark.timeout(600)
def test_gemm_fp64_full_size_8192(gpu_ctx):
    """The FP64-carrier leg at full size: 8192^3, p = 67108859, uint32 storage.
    Freivalds on the whole product, and a 256-row slice from the plain kernel
    compared for equality."""
    p, n = 67108859, 8192
    F = la.Modular(p, "uint32")
    gpu_ctx.F = F
    A = synth.zp_matrix(synth.seed_for(2, 5), n, n, p)
    B = synth.zp_matrix(synth.seed_for(2, 6), n, n, p)
    C0 = synth.zp_matrix(synth.seed_for(2, 7), n, n, p)
    C = C0.copy()
    print("fp64 8192^3 product", flush=True)
    la.BLAS3.gemm(gpu_ctx, 1, A, B, 1, C, algo=la.ALGO_FP64)
    print("freivalds", flush=True)
    _freivalds(A, B, C0, C, p, 1, 1, seed=7)
    Cs = C0[:256].copy()
    la.BLAS3.gemm(gpu_ctx, 1, A[:256], B, 1, Cs, algo=la.ALGO_PLAIN)
    assert np.array_equal(Cs, C[:256])
