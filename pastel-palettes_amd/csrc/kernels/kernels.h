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
  int n_tiles, n_lowrank;
};

// One leaf of an H-matrix to construct.  Dense: entries at `off` (ld nx);
// low rank: U (ny x rank) at `off`, V (nx x rank) at `off_v`, ld rank, and
// its 2 * rank pivot indices at `aux` of the pivot array.
struct BuildLeaf {
  int32_t row0, col0, ny, nx, rank, ld;  // ld: row pitch of U and V (>= rank)
  int64_t off, off_v, aux;
};
// A cross approximation of rank k (pitch kp, a multiple of 32) in scratch,
// truncated to rank r into the arena: U (ny x r) at out_u, V (nx x r) at
// out_v.  work: 6 kp^2 + 2 kp scratch elements; status: 0, or 1 when a
// Cholesky pivot failed.
struct RecompressLeaf {
  int32_t ny, nx, k, kp, r, status;
  int64_t u, v, work, out_u, out_v;
};
constexpr int kBuildMaxRank = 256;
constexpr int kBuildMaxRows = 65536;

// Fills an arena from kernel entries (kind 0: exp(-|x_i - x_j| / param) on
// 1-D points, 1: 1 / (|p_i - p_j| + param) on 2-D points, + shift on the
// global diagonal): dense leaves entry by entry, low-rank leaves by
// fixed-rank partially pivoted cross approximation.
hipError_t launch_build(int kind, const BuildLeaf* dense, int n_dense, int max_dense_elems, const BuildLeaf* lowrank,
                        int n_lowrank, const double* pts, double param, double shift, double* arena,
                        int32_t* pivots, hipStream_t s);

hipError_t launch_recompress(RecompressLeaf* leaves, int n, double* scratch, double* arena, hipStream_t s);

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
hipError_t launch_hmatvec(const MatvecPlanDev& p, const double* A, const double* x, int ldx, const double* y, int ldy,
                          double* out, int ldo, int nrhs, double* t, hipStream_t s);
// out = sum(D x) - sum(U (V^T x)) over the plan (a hierarchical inverse whose
// off-diagonal blocks are stored as -U V^T), in the arena's element type
// (8: FP64, 4: FP32).
hipError_t launch_happly(const MatvecPlanDev& p, int elem_bytes, const void* A, const void* x, int ldx, void* out,
                         int ldo, int nrhs, void* t, hipStream_t s);
hipError_t launch_segnorm2(const double* r, int ldr, const int* seg, int nseg, int nrhs, double* out, hipStream_t s);
hipError_t launch_colnorm2(const double* r, int ldr, int n, int nrhs, double* out, hipStream_t s);
hipError_t launch_mfma_probe(int mode, const double* a, const double* b, double* out, hipStream_t s);
hipError_t launch_single_op(int precision, int threads, int reps, const int32_t* words, void* const* ptrs,
                            void* scratch, const void* rnd, int32_t* status, hipStream_t s);

}  // namespace pp
