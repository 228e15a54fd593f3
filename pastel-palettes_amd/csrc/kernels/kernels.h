// Host-side declarations of the device kernels' launch wrappers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vm.h"

namespace pp {

// One leaf of an H-matrix for the matrix-vector product.
struct MatvecBlock {
  int32_t type;  // 1 dense, 2 low rank
  int32_t row0, col0, ny, nx, rank;
  int64_t off, off_u, off_v;  // element offsets in the arena the plan is applied to
  int64_t t_off;              // low rank: offset of its rank x nrhs V^T x block
  // low rank: V^T x is formed in nparts partial products over row parts of
  // V (at most kPartRows rows each), part s at p_off + s * p_stride of the
  // partial area
  int32_t part0, nparts;
  int64_t p_off, p_stride;
};

// A row range of the product and the leaves covering it.
struct MatvecTile {
  int32_t row0, rows, first, count;  // contributions [first, first + count)
};

struct MatvecPlanDev {
  const MatvecBlock* blocks;
  const MatvecTile* tiles;
  const int* contrib;  // pairs (block index, unused)
  const int* lr_list;  // indices of low-rank blocks
  const int* parts;    // pairs (block index, part number) of the V^T x partial products
  int n_tiles, n_lowrank, n_parts;
  int max_rank;         // largest rank of a low-rank block
  int64_t t_elems = 0;  // V^T x area (max t_off + rank * nrhs)
  int64_t p_elems = 0;  // partial area that follows it
};

// Rows of V per V^T x partial product: a part is one 16 x 16 MFMA output tile
// per wave pass whose loads for the whole part are in flight at once, and an
// 8192-row factor is spread over 64 workgroups instead of streaming through
// one CU.
constexpr int kPartRows = 128;
// Rows of a row tile of the product (a diagonal leaf's rows split four ways
// at leaf 256); each tile is handled by up to four workgroups of 16 rows.
constexpr int kTileRows = 64;
// Rows per workgroup of the first stage of the per-member column norms.
constexpr int kNormRows = 256;

// ---- H-matrix construction (build.hip)
constexpr int kBuildMaxRank = 256;
constexpr int kBuildMaxRows = 65536;
constexpr int kBuildSlice = 256;  // rows / columns per workgroup

// Kernel function: kind 0: exp(-|x_i - x_j| / param) on 1-D points, 1:
// 1 / (|p_i - p_j| + param) on 2-D points, + shift on the global diagonal.
struct BuildArgs {
  const double* pts;
  double param, shift;
};
// A dense leaf: entries at `off` of the arena (ld nx).
struct BuildLeaf {
  int32_t row0, col0, ny, nx, rank, ld;
  int64_t off, off_v, aux;
};
// A cross approximation of rank k into U (ny x ld) and V (nx x ld).
struct AcaLeaf {
  int32_t row0, col0, ny, nx, k, ld;
  double* U;
  double* V;
  int32_t* piv;    // 2k pivot indices or null
  uint32_t* used;  // ceil(ny / 32) words, zero at the start
  int32_t col_part, ncol, row_part, nrow;  // its slice partials
};
struct AcaState {
  int32_t i, jp, zero, next;  // pivot row (-1: done), pivot column, empty step, row after an empty step
  double d;                   // pivot value
};
struct BuildSlice {
  int32_t leaf, first, count, part;
};
struct BuildPartial {
  double m, v;
  int32_t idx, pad;
};
// Truncation of a rank-k approximation (pitch kp) to rank r.
struct TruncFactor {  // U_k or V_k of one leaf
  double *G, *Rt, *T, *d, *Ri;
  double* parts;  // column-norm partials, slice q at parts + q * part_stride
  double* gparts;  // Gram partials, slice q at gparts + q * gpart_stride
  int64_t part_stride, gpart_stride;
  int32_t kp, nslices, m, status;
};
struct TruncSlice {
  double* X;  // the factor (m x kp)
  double* part;
  double* gpart;
  const double* d;
  int32_t first, count, kp, factor;  // factor index 2 leaf + (0: U, 1: V)
};
struct TruncLeaf {
  double *N, *J, *out_u, *out_v;
  int32_t* sel;
  int32_t kp, r, status, rotated;
};

hipError_t launch_dense_fill(int kind, const BuildLeaf* dense, int n_dense, int max_elems, const BuildArgs& a,
                             double* arena, hipStream_t s);
// max_k steps of five launches each over all leaves
hipError_t launch_aca(int kind, const AcaLeaf* leaves, AcaState* st, int n_leaves, int max_k,
                      const BuildSlice* col_slices, int n_col, const BuildSlice* row_slices, int n_row,
                      BuildPartial* part, const BuildArgs& a, hipStream_t s);
// synchronises s once per Jacobi sweep (flag: one device int)
hipError_t launch_truncate(TruncLeaf* leaves, int n_leaves, int max_kp, TruncFactor* factors,
                           const TruncSlice* slices, int n_slices, int* flag, hipStream_t s);

// partial-pivoted LU of `count` row-major ny x nx blocks (stride elements
// apart); pivots: count x ny permutations, info: count values
hipError_t launch_getrf_pivoted(int fp64, void* a, int ld, int ny, int nx, int count, int64_t stride,
                                int32_t* pivots, int32_t* info, hipStream_t s);
// rows of M (ny x nx) permuted by pivots (recover: by the inverse); M's
// column chunks pass through LDS, ny * element size <= kPivotLdsBytes
constexpr int kPivotLdsBytes = 48 * 1024;
int apply_pivot_chunk(int fp64, int ny);
hipError_t launch_apply_pivot(int fp64, void* m, int ld, int nx, int ny, const int32_t* pivots, int recover,
                              hipStream_t s);

// factor: launch under the factorization symbol (vm_factor_kernel)
hipError_t launch_vm(const VmArgs& a, int precision, int workers, int threads, bool factor, uint32_t* epoch_dev,
                     int32_t* status, hipStream_t s);
int vm_blocks_per_cu(int precision, int threads);

hipError_t launch_fill_zero(void* p, int64_t n, int elem_bytes, hipStream_t s);
hipError_t launch_convert(void* dst, int dst_bytes, const void* src, int src_bytes, int64_t n, hipStream_t s);
hipError_t launch_copy2d(void* dst, int dst_bytes, int ldd, const void* src, int src_bytes, int lds_, int n, int nrhs,
                         hipStream_t s);
hipError_t launch_add_correction(double* x, int ldx, const void* d, int d_bytes, int ldd, int n, int nrhs,
                                 hipStream_t s);
// Optional outputs of an H-matrix-vector product besides out (which may be
// null): a second copy in 4- or 8-byte elements (out2, pitch ld2); an FP64
// target xo set to (xmode 1) or incremented by (xmode 2) the product; the
// per-workgroup column sums of squares of the product (nrm_part: for column
// c, the n_tiles kTileRows / 16 workgroup sums at nrm_part + c n_tiles
// kTileRows / 16) for launch_rownorm_sum.
struct HmvExtra {
  void* out2 = nullptr;
  int out2_bytes = 0, ld2 = 0;
  double* xo = nullptr;
  int ldxo = 0, xmode = 0;
  double* nrm_part = nullptr;
};
// out = y - A_H x (FP64 residual)
hipError_t launch_hmatvec(const MatvecPlanDev& p, const double* A, const double* x, int ldx, const double* y, int ldy,
                          double* out, int ldo, int nrhs, double* t, const HmvExtra& ex, hipStream_t s);
// out = sum(D x) - sum(U (V^T x)) over the plan (a hierarchical inverse whose
// off-diagonal blocks are stored as -U V^T), in the arena's element type
// (8: FP64, 4: FP32).
hipError_t launch_happly(const MatvecPlanDev& p, int elem_bytes, const void* A, const void* x, int ldx, void* out,
                         int ldo, int nrhs, void* t, const HmvExtra& ex, hipStream_t s);
// out[g nrhs + c] = sum of squares of column c over member g's row tiles
// (seg_tiles: pairs first tile, tile count) from a product's nrm_part.
hipError_t launch_rownorm_sum(const double* part, const int* seg_tiles, int n_tiles, int nseg, int nrhs, double* out,
                              hipStream_t s);
// Per (member, column) sums of squares of r: a first launch sums rows in
// chunks of kNormRows into part (nseg x nchunks x nrhs), a second adds the
// chunks of each member in order into out (nseg x nrhs).
hipError_t launch_segnorm2(const double* r, int ldr, const int* seg, int nseg, int nrhs, int nchunks, double* part,
                           double* out, hipStream_t s);
hipError_t launch_mfma_probe(int mode, const double* a, const double* b, double* out, hipStream_t s);
hipError_t launch_single_op(int precision, int threads, int reps, const SingleOp& op, void* scratch, const void* rnd,
                            int32_t* status, hipStream_t s);
// Sketch values [-1, 1) of the library's generator (splitmix64 from `seed`,
// value i from state seed + (i + 1) * 0x9E3779B97F4A7C15), as FP64
// (elem_bytes 8) or rounded to FP32 (4), computed on the device.
hipError_t launch_sketch_fill(void* out, int64_t count, int elem_bytes, uint64_t seed, hipStream_t s);
hipError_t launch_clear_i32(int32_t* p, int n, hipStream_t s);

}  // namespace pp
