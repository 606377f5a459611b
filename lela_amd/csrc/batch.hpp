// lela_amd/csrc/batch.hpp -- device-side descriptors of the batched executor
// (many independent Z/p gemm and echelon-form instances in persistent,
// work-list driven launches).  Shared by batch.hip, zp_gemm.hip, zp_rref.hip.
#pragma once

#include "common.hpp"

namespace mlz {

// Kernel variants a gemm instance can run with (all exact).
enum {
	BG_PLAIN = 0,    // VALU uint64 multiply-accumulate, 64 x 64 tiles
	BG_SPARSE = 1,   // VALU, skips zero entries of A (ballot over 64 k), 16 x 256 tiles
	BG_I8S = 2,      // INT8 MFMA limbs, 64 x 64 tiles (packed operands)
	BG_I8L = 3,      // INT8 MFMA limbs, 128 x 128 tiles (packed operands)
	BG_NVAR = 4
};
// Kernel variants an echelon-form instance can run with (one workgroup per
// instance; the trailing update of each panel on VALU or on INT8 MFMA).
enum {
	BR_W64 = 0,      // 64-column panels, FP64 (or uint64) VALU update
	BR_W32 = 1,      // 32-column panels, VALU update
	BR_I8_W64 = 2,   // 64-column panels, INT8 MFMA update (p <= 2^16)
	BR_I8_W32 = 3,   // 32-column panels, INT8 MFMA update (p <= 2^16)
	BR_NVAR = 4
};

constexpr int BR_MAXM = 4096;    // rows of a batched echelon-form instance (pivot flags in LDS)

// One gemm instance C <- a*A*B + b*C (device pointers, element storage E).
struct BatchGemm {
	const void *A, *B;
	void *C;
	uint64_t lda, ldb, ldc;
	int m, n, k;
	uint32_t a, b;
	int8_t *Ap, *Bp;       // packed INT8 operands (INT8 variants), else null
	int KB;                // 32-k blocks of the packed operands
	int pad;
};

// One echelon-form instance, reduced in place (standard layout).
struct BatchRref {
	void *A;               // caller matrix, m x n, leading dimension lda
	uint64_t lda;
	uint32_t *pivots;      // caller's pivot-column array (min(m, n)) or null
	void *Wm;              // m x n canonical working copy (uint16 for the MFMA variants, else uint32)
	uint32_t *Wb;          // panel-width x n rows of P_J^-1 * P (VALU variants), else null
	int8_t *Wbp;           // Wb as INT8 limb fragments (MFMA variants), else null
	int *piv_rows, *piv_cols;   // min(m, n) each
	int m, n;
};

// Work-list entry: task index and up to three coordinates.
struct BatchItem {
	int task, x, y, z;
};

// Common launch arguments of the persistent kernels.
struct BatchLaunch {
	const BatchItem *items;
	int nitems;
	int *queue;                    // work counter, zero before the launch
	unsigned long long *tcost;     // per-task sum of workgroup durations (10 ns ticks), or null
};

// Next work items for the whole workgroup: the first of n consecutive items
// claimed with one atomic (-1 when the list is exhausted; the claim may run
// past the end, the caller stops at L.nitems).  Claiming several small items
// at once keeps the single counter from serialising the launch.  Ends with a
// barrier, so slot may be rewritten on the next call only after every thread
// has read it.
__device__ __forceinline__ int batch_next(const BatchLaunch &L, int *slot, int n = 1)
{
	__syncthreads();
	if (threadIdx.x == 0)
		*slot = atomicAdd(L.queue, n);
	__syncthreads();
	// uniform across the workgroup: descriptor loads can go through the scalar unit
	const int it = __builtin_amdgcn_readfirstlane(*slot);
	return it < L.nitems ? it : -1;
}

__device__ __forceinline__ void batch_charge(const BatchLaunch &L, int task, uint64_t t0)
{
	if (L.tcost && threadIdx.x == 0)
		atomicAdd(&L.tcost[task], (unsigned long long) (wall_clock64() - t0));
}

// Launchers (queue on `stream`; grid = number of persistent workgroups).
int batch_pack_i8_launch(hipStream_t stream, int elem, uint32_t p, const BatchGemm *tasks, BatchLaunch L, int grid);
int batch_gemm_launch(hipStream_t stream, int elem, int variant, uint32_t p, uint64_t kmax, const BatchGemm *tasks,
		      BatchLaunch L, int grid);
int batch_rref_launch(hipStream_t stream, int elem, int variant, uint32_t p, const BatchRref *tasks, int *ranks,
		      int *bad, BatchLaunch L, int grid);
// Resident workgroups per CU of a variant's kernel (variant BG_NVAR: the INT8
// pack kernel), for grid sizing and for turning durations into CU time.
int batch_gemm_wgs_per_cu(int elem, int variant, uint32_t p, uint64_t kmax);
int batch_rref_wgs_per_cu(int elem, int variant, uint32_t p);

// Tile geometry of the gemm variants.
inline int bg_tile_m(int v) { return v == BG_SPARSE ? 16 : v == BG_I8L ? 128 : 64; }
inline int bg_tile_n(int v) { return v == BG_SPARSE ? 256 : v == BG_I8L ? 128 : 64; }
inline int br_width(int v) { return v == BR_W32 || v == BR_I8_W32 ? 32 : 64; }
inline bool br_mfma(int v) { return v == BR_I8_W64 || v == BR_I8_W32; }
// Row stride of an instance's working copy: padded to a multiple of 4 entries
// for the MFMA variants.
__host__ __device__ inline int br_ldw(int n, bool mfma) { return mfma ? (n + 3) & ~3 : n; }
// Bytes of the packed Wb of an MFMA variant: [32-column block][k block][limb][64 lanes][16 bytes].
inline size_t br_wbp_bytes(int v, size_t n) { return (size_t) br_width(v) * ((n + 31) / 32 * 32) * 2; }

}  // namespace mlz
