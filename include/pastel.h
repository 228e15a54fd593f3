/*
 * pastel.h -- C ABI of the MI355X H-LU factor + solve library (libpastel.so).
 *
 * Plain pointers and sizes only.  Every function returns PP_OK (0) or a
 * negative PP_ERR_* code and leaves a message for pp_last_error() (thread
 * local).  No C++ exception crosses this boundary.
 *
 * The reference (QxGbM/Pastel-Palettes) is a C++/CUDA library with no FFI of
 * its own; each entry point below names the reference interface whose role
 * it takes over.  Device pointers passed in ("dptr") are HIP device
 * addresses on the context's device.
 */
#ifndef PASTEL_H
#define PASTEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PP_OK 0
#define PP_ERR_INVALID (-1)  /* bad argument or shape */
#define PP_ERR_HIP (-2)      /* HIP runtime error */
#define PP_ERR_TIMEOUT (-3)  /* VM watchdog fired or host wait timed out */
#define PP_ERR_NOMEM (-4)    /* HBM budget or allocation failure */
#define PP_ERR_INTERNAL (-5) /* anything else */
#define PP_ERR_DEVICE (-6)   /* kernel reported a bad opcode / operation */

/* precision modes */
#define PP_FP64 0  /* FP64 storage, FP64 MFMA everywhere (the drop-in) */
#define PP_FP32 1  /* FP32 storage, FP32 MFMA */
#define PP_MIXED 2 /* FP32 storage, FP32 leaf getrf/trsm, BF16 MFMA (FP32 accumulate) for gemm and accum */

const char* pp_last_error(void);
/* Library version string. */
const char* pp_version(void);

/* ------------------------------------------------------------------ rts */

/* DataMap2D::checkOverlap (rts/ops.h:18, rts/ops.cpp:7-40).  Byte units.
 * Returns 1 / 0, or a negative error code. */
int pp_rts_check_overlap(uint64_t ptr_a, uint64_t pitch_a, uint64_t width_a, uint64_t height_a, uint64_t ptr_b,
                         uint64_t pitch_b, uint64_t width_b, uint64_t height_b);

/* Operation::checkDependencyFrom (rts/ops.cpp:42-61): maps are n x 4 arrays
 * (ptr, pitch, width, height), read_only n flags.  Returns the 3-bit mask
 * (flow 1, anti 2, out 4) of op_to depending on op_from. */
int pp_rts_check_dependency(int n_from, const uint64_t* maps_from, const int32_t* ro_from, int n_to,
                            const uint64_t* maps_to, const int32_t* ro_to);

/* dependency_table (rts/dependency.h:24-53, rts/dependency.cpp:11-76): the
 * edge set with its from / to lists and mask per edge.  update with mask 0
 * (dependency_t::no) removes the edge; lists come in insertion order
 * (to != 0: predecessors of inst, else successors); list returns the count
 * and fills at most cap entries of out (may be NULL). */
typedef struct pp_dep_table pp_dep_table;
int pp_dep_table_create(pp_dep_table** out);
void pp_dep_table_destroy(pp_dep_table* t);
int pp_dep_table_update(pp_dep_table* t, int64_t inst_from, int64_t inst_to, int dep);
int pp_dep_table_lookup(const pp_dep_table* t, int64_t inst_from, int64_t inst_to);
int pp_dep_table_list(const pp_dep_table* t, int64_t inst, int to, int64_t* out, int cap);
int64_t pp_dep_table_count(const pp_dep_table* t, int64_t inst, int to);
int pp_dep_table_clear(pp_dep_table* t);

/* instructions_scheduler (rts/scheduler.h:73-126) over an explicit DAG:
 * n instructions with FLOP estimates, n_edges edges (edges[2e] -> edges[2e+1])
 * with masks.  From- and to-lists are taken in ascending order. */
typedef struct pp_schedule pp_schedule;
int pp_schedule_create(int n, const int64_t* flops, int n_edges, const int32_t* edges, const int32_t* masks,
                       int workers, pp_schedule** out);
int pp_schedule_worker_length(const pp_schedule* s, int worker);
/* Entries of one worker: instruction ids, 1 = execute / 0 = wait, and the
 * worker's FLOP clock at each entry.  Any output pointer may be NULL. */
int pp_schedule_worker_entries(const pp_schedule* s, int worker, int32_t* insts, int32_t* ex_w, int64_t* start_flops);
int pp_schedule_comm_waits(const pp_schedule* s);
void pp_schedule_destroy(pp_schedule* s);
/* _MIN_INST_FLOPS used by the scheduler. */
int64_t pp_min_inst_flops(void);

/* ------------------------------------------------------------------ H-matrix structure */

/* Hierarchical (matrix/Hierarchical.h; API implied by src/launcher.cu:126-139). */
typedef struct pp_hmatrix pp_hmatrix;
/* HODLR: `levels` binary splits, off-diagonal blocks of rank `rank`. */
int pp_hmatrix_create_hodlr(int n, int levels, int rank, pp_hmatrix** out);
/* All-dense binary partition down to `leaf`. */
int pp_hmatrix_create_dense(int n, int leaf, pp_hmatrix** out);
void pp_hmatrix_destroy(pp_hmatrix* h);
int pp_hmatrix_size(const pp_hmatrix* h);
/* Arena size in elements. */
int64_t pp_hmatrix_storage(const pp_hmatrix* h);
int pp_hmatrix_num_leaves(const pp_hmatrix* h);
/* info = {type (1 dense, 2 low rank), row0, col0, ny, nx, rank, off (dense) or off_u, off_v} */
int pp_hmatrix_leaf(const pp_hmatrix* h, int i, int64_t info[8]);

/* ------------------------------------------------------------------ programs */

/* An instruction stream for one launch: h-ops (h_ops tree, src/launcher.cu:139)
 * -> DAG (h_ops_dag, src/launcher.cu:53) -> schedule (rts/scheduler.cpp)
 * -> per-worker words (instructions_manager, src/launcher.cu:72).
 * Pointer-table slots: 0 = matrix arena, 1 = temp arena, 2 = right-hand sides.
 * Batched programs use slots 3*i, 3*i+1, 3*i+2 for matrix i. */
typedef struct pp_program pp_program;
/* H-LU factorization (dev_hierarchical_GETRF, src/launcher.cu:117). */
int pp_program_factor(const pp_hmatrix* h, int workers, pp_program** out);
/* Forward + backward substitution on an n x nrhs row-major block (pitch ld). */
int pp_program_solve(const pp_hmatrix* h, int nrhs, int ld, int rhs_split, int workers, pp_program** out);
/* One program factoring (solve != 0: solving) `count` matrices at once. */
int pp_program_factor_batch(const pp_hmatrix* const* hs, int count, int workers, pp_program** out);
/* Windowed factorization (replaces the kernel_size loop of
 * dev_hierarchical_GETRF, src/launcher.cu:146-167, and the start_index /
 * length_max arguments of generateLaunchArgsFromTree, src/launcher.cu:46-53):
 * the factorization's h-ops cut into ceil(ops / window_ops) consecutive
 * windows (window_ops <= 0: one), each with its own DAG, schedule and
 * streams; launching the windows in order computes what the whole program
 * computes.  Every window addresses the whole temp area (slot 1,
 * pp_program_temp_elems(window, 0) elements): pass the same buffer to every
 * window's stream; the first window's launch clears it, later windows keep
 * what earlier ones wrote (their upload fails without that buffer). */
int pp_program_factor_num_windows(const pp_hmatrix* h, int window_ops);
int pp_program_factor_window(const pp_hmatrix* h, int workers, int window_ops, int index, pp_program** out);
int pp_program_solve_batch(const pp_hmatrix* const* hs, int count, const int32_t* nrhs, const int32_t* ld,
                           int rhs_split, int workers, pp_program** out);

/* Solve through explicit inverses (no counterpart in the reference, which
 * stops at the factorization).  The factor program below also forms, for
 * every dense diagonal leaf, L^{-1} of its unit lower factor and U^{-1} of its
 * upper factor, and for HODLR-shaped matrices the low-rank off-diagonal
 * blocks of the full L^{-1} and U^{-1}, in an inverse area that starts at element
 * pp_hmatrix_inverse_base(h) of the matrix slot and holds
 * pp_hmatrix_inverse_elems(h) elements; the area must hold the identity
 * blocks of pp_hmatrix_inverse_identity before each run.  The solve program
 * then applies each leaf with one dense product (operation trmm). */
int64_t pp_hmatrix_inverse_base(const pp_hmatrix* h);
int64_t pp_hmatrix_inverse_elems(const pp_hmatrix* h);
/* Writes the identity blocks (zeros elsewhere) into a host area of
 * pp_hmatrix_inverse_elems(h) doubles. */
int pp_hmatrix_inverse_identity(const pp_hmatrix* h, double* area);
/* Layout of the inverse area.  Leaf i (solve order): info = {row0, nb,
 * offset of L^{-1} (U^{-1} follows at + nb*nb rounded up to 32), arena offset
 * of the leaf}.  For HODLR-shaped matrices the area also holds, per internal
 * node i (children before parents), the low-rank factors of the off-diagonal
 * blocks of L^{-1} (-l_u l_v^T, rows of the second child x columns of the
 * first) and U^{-1} (-u_u u_v^T, rows of the first child x columns of the
 * second): info = {row0, na, nb, rl, ru, l_u, l_v, u_u, u_v} (offsets within
 * the area, ld = rank).  Non-HODLR-shaped matrices have no nodes. */
int pp_hmatrix_inverse_count(const pp_hmatrix* h, int* leaves, int* nodes);
int pp_hmatrix_inverse_leaf(const pp_hmatrix* h, int i, int64_t info[4]);
int pp_hmatrix_inverse_node(const pp_hmatrix* h, int i, int64_t info[9]);
int pp_program_factor_inv_batch(const pp_hmatrix* const* hs, int count, int workers, pp_program** out);
int pp_program_solve_inv_batch(const pp_hmatrix* const* hs, int count, const int32_t* nrhs, const int32_t* ld,
                               int rhs_split, int workers, pp_program** out);
void pp_program_destroy(pp_program* p);

typedef struct pp_program_info {
  int32_t n_insts, workers, n_slots, pad;
  int64_t n_edges, flops, temp_elems, scratch_elems, total_words, comm_waits;
  double dag_density, schedule_ms, dag_ms, encode_ms;
} pp_program_info;
int pp_program_get_info(const pp_program* p, pp_program_info* info);
/* Temp-arena size (elements) of batch member i (slot 3*i+1). */
int64_t pp_program_temp_elems(const pp_program* p, int member);
int pp_program_worker_length(const pp_program* p, int worker); /* words */
int pp_program_worker_words(const pp_program* p, int worker, int32_t* out, int cap);
/* execute words of instruction `inst` (id slot filled); returns length */
int pp_program_inst_words(const pp_program* p, int inst, int32_t* out, int cap);
int64_t pp_program_inst_flops(const pp_program* p, int inst);
/* Algorithmic bytes of instruction `inst` at 8 bytes per element: operands
 * read once, operands written counted twice (read and written); 0 for a
 * decoded program (operand maps are not part of the byte-string format). */
int64_t pp_program_inst_bytes(const pp_program* p, int inst);
/* predecessors (to-list) of inst; returns count (fills up to cap) */
int pp_program_inst_preds(const pp_program* p, int inst, int32_t* preds, int32_t* masks, int cap);
/* scheduler entries of one worker (like pp_schedule_worker_entries) */
int pp_program_schedule_length(const pp_program* p, int worker);
int pp_program_schedule_entries(const pp_program* p, int worker, int32_t* insts, int32_t* ex_w, int64_t* start_flops);

/* ------------------------------------------------------------------ device runtime */

typedef struct pp_ctx pp_ctx;
/* Context on `device`; allocations through it are capped at budget_bytes
 * (0 = no cap).  The context owns a non-blocking HIP stream ("the context's
 * stream", used wherever a call takes hip_stream = NULL): it does not wait
 * for work on the legacy null stream, so data produced there (for example by
 * a framework's default stream) must be complete, or the caller's stream
 * passed instead. */
int pp_ctx_create(int device, uint64_t budget_bytes, pp_ctx** out);
int pp_ctx_destroy(pp_ctx* ctx);
int pp_ctx_cu_count(const pp_ctx* ctx);
uint64_t pp_ctx_bytes_in_use(const pp_ctx* ctx);
int pp_alloc(pp_ctx* ctx, uint64_t bytes, void** dptr);
int pp_free(pp_ctx* ctx, void* dptr);
int pp_copy_h2d(pp_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int pp_copy_d2h(pp_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* Uploaded instruction stream (launchKernelWithArgs' arguments,
 * src/launcher.cu:92-115: insts, ptrs, comm_space, block_tmps, rnd, clocks). */
typedef struct pp_stream pp_stream;
/* Generic upload: per-worker words, pointer table (device addresses),
 * scratch bytes per worker.  `precision` selects the element type of the
 * sketch matrix (generated from `seed`) and of the scratch. */
int pp_stream_upload(pp_ctx* ctx, const int32_t* const* words, const int32_t* lengths, int workers,
                     const void* const* ptr_table, int n_ptrs, uint64_t scratch_bytes_per_worker, int n_insts,
                     int precision, pp_stream** out);
/* Upload a program; ptr_table must hold n_slots device addresses.  Temp slots
 * left NULL are allocated (and zero-filled before every launch) by the
 * stream. */
int pp_stream_upload_program(pp_ctx* ctx, const pp_program* prog, const void* const* ptr_table, int n_ptrs,
                             int precision, pp_stream** out);
/* Rebind pointer-table slot i. */
int pp_stream_set_slot(pp_ctx* ctx, pp_stream* s, int slot, void* dptr);
/* Launch the persistent VM (kernel_dynamic, src/kernel.cu:1375) on hip_stream
 * (NULL = the context stream); threads = 64 * wavefronts per worker (4 wavefronts:
 * one 256-thread workgroup per worker). */
int pp_launch(pp_ctx* ctx, pp_stream* s, int wavefronts_per_inst, void* hip_stream);
/* Wait for the launch; on a VM error reports the first unsignalled instruction
 * (or the faulting one) in *first_unsignalled (may be NULL). */
int pp_wait(pp_ctx* ctx, pp_stream* s, int timeout_ms, int32_t* first_unsignalled);
/* Enable per-entry timestamps for later launches (1) or disable (0). */
int pp_stream_enable_trace(pp_ctx* ctx, pp_stream* s, int enable);
/* Per-worker entry timestamps (100 MHz wall clock) of the last traced
 * launch, concatenated by worker, each worker holding 1 + its entry count;
 * returns the number of values (instructions_scheduler::analyzeClocks). */
int64_t pp_read_trace(pp_ctx* ctx, pp_stream* s, uint64_t* out, int64_t cap);
int pp_stream_destroy(pp_ctx* ctx, pp_stream* s);
/* Watchdog limit for waits in milliseconds (default 20000). */
int pp_stream_set_watchdog(pp_ctx* ctx, pp_stream* s, int ms);
/* Kernel time of the last launch in ms (HIP events around the VM kernel). */
double pp_stream_last_ms(const pp_stream* s);
/* accum operations of the last waited launch whose CholeskyQR3 failed and
 * fell back to Householder QR (a numerical-health count; no reference
 * counterpart, the reference always uses Givens QR). */
int pp_stream_qr_fallbacks(const pp_stream* s);
/* Maximum workers that fit co-resident for a precision / wave count. */
int pp_max_workers(pp_ctx* ctx, int precision, int wavefronts_per_inst);

/* Run one execute entry (words, id slot ignored) on a single workgroup;
 * device pointer table of n_ptrs slots.  Test harness for the device
 * routines. */
int pp_exec_single(pp_ctx* ctx, int precision, const int32_t* words, int n_words, const void* const* ptr_table,
                   int n_ptrs, uint64_t scratch_bytes, uint64_t seed);
/* Same, `reps` times back to back in one launch of 64 * wavefronts threads;
 * *ms receives the mean kernel time per execution (per-op latency). */
int pp_exec_single_timed(pp_ctx* ctx, int precision, const int32_t* words, int n_words, const void* const* ptr_table,
                         int n_ptrs, uint64_t scratch_bytes, uint64_t seed, int wavefronts, int reps, double* ms);
/* pp_exec_single without any allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * worker scratch (scratch_bytes), the accum sketch and the status word live
 * in one static device array per GPU, resolved once per device by
 * pp_ctx_create; a call whose scratch, sketch and status need more than
 * pp_exec_single_noalloc_max_bytes() (4 MiB) fails with PP_ERR_INVALID.
 * ptr_table holds at most 16 device addresses.  Synchronous on hip_stream
 * (NULL = the context's stream).  Every call on one GPU shares that array
 * and leaves its contents there: calls that overlap on one GPU race. */
int pp_exec_single_noalloc(pp_ctx* ctx, int precision, const int32_t* words, int n_words,
                           const void* const* ptr_table, int n_ptrs, uint64_t scratch_bytes, uint64_t seed,
                           void* hip_stream);
uint64_t pp_exec_single_noalloc_max_bytes(void);
/* Fill a device buffer with the sketch matrix used by accum (count values,
 * uniform in [-1, 1) from the in-library generator). */
int pp_fill_sketch(pp_ctx* ctx, void* dptr, int64_t count, int precision, uint64_t seed);
/* MFMA lane-layout probe (mode 0 f64 16x16x4, 1 f32 16x16x4, 2 bf16 16x16x32):
 * out[4*lane + r] = accumulator r of lane. a, b, out are host arrays. */
int pp_mfma_probe(pp_ctx* ctx, int mode, const double* a, int na, const double* b, int nb, double* out);

/* ------------------------------------------------------------------ solver (factor + solve + refinement) */

typedef struct pp_solver pp_solver;
typedef struct pp_solve_report {
  int32_t iterations, converged, pad0, pad1;
  double residual;      /* max over columns of ||b - A x|| / ||b|| (FP64) */
  double factor_ms, solve_ms, total_ms;
} pp_solve_report;

/* Build a solver for H-matrix h whose FP64 arena is at host `storage`
 * (pp_hmatrix_storage(h) doubles).  Keeps the FP64 matrix resident for the
 * residual and a working copy in the precision's storage type for the
 * factors. */
int pp_solver_create(pp_ctx* ctx, const pp_hmatrix* h, const double* storage, int precision, int workers,
                     int wavefronts_per_inst, int nrhs_max, int rhs_split, pp_solver** out);
/* Same, with the FP64 arena already on the device (copied, not adopted). */
int pp_solver_create_device(pp_ctx* ctx, const pp_hmatrix* h, const double* storage_dev, int precision, int workers,
                            int wavefronts_per_inst, int nrhs_max, int rhs_split, pp_solver** out);
/* One solver over `count` H-matrices handled as a block-diagonal system:
 * one persistent launch factors all of them, one solves all of them
 * (batched execution).  storages[i] is member i's FP64 arena (host memory,
 * or device memory when on_device != 0).  Right-hand sides and solutions
 * are the members' rows stacked in order (pp_solver_member gives offsets). */
int pp_solver_create_batch(pp_ctx* ctx, const pp_hmatrix* const* hs, int count, const double* const* storages,
                           int on_device, int precision, int workers, int wavefronts_per_inst, int nrhs_max,
                           int rhs_split, pp_solver** out);
/* Solve modes (no counterpart in the reference).  PP_SOLVE_HINV (the
 * default of the functions above): the factorization launch also forms the
 * full hierarchical inverses of HODLR-shaped members and each solve is two
 * H-matrix-vector products; PP_SOLVE_LEAFINV: leaf inverses only, solve
 * programs apply them (trmm); PP_SOLVE_SUBST: plain factorization, solve
 * programs by substitution (trsml / trsmu).  The first moves work from every
 * solve into each factorization, which pays off when solves are latency
 * bound, the last is cheapest when one launch carries many matrices. */
#define PP_SOLVE_HINV 0
#define PP_SOLVE_LEAFINV 1
#define PP_SOLVE_SUBST 2
int pp_solver_create_batch_mode(pp_ctx* ctx, const pp_hmatrix* const* hs, int count, const double* const* storages,
                                int on_device, int precision, int workers, int wavefronts_per_inst, int nrhs_max,
                                int rhs_split, int solve_mode, pp_solver** out);
int pp_solver_num_members(const pp_solver* s);
int64_t pp_solver_rows(const pp_solver* s);
/* Row offset, size and last refined residual of member i. */
int pp_solver_member(const pp_solver* s, int i, int64_t* row_off, int* n, double* last_residual);
/* Element offset of member i's arena (its pp_hmatrix_storage elements, then
 * its inverse area) within the factor arena of pp_solver_factor_arena and
 * within the solver's FP64 copy of the matrices. */
int pp_solver_member_arena(const pp_solver* s, int i, int64_t* arena_off);
/* Elements of the factor arena (every member's arena and inverse area). */
int64_t pp_solver_storage_elems(const pp_solver* s);
/* Factor through a captured hipGraph (captured on first use per stream). */
int pp_solver_use_graph(pp_solver* s, int enable);
/* Factor in place (restores the working copy from the FP64 matrix first). */
int pp_solver_factor(pp_solver* s, void* hip_stream);
/* x = A^{-1} b by the factors, refined in FP64 until the relative residual is
 * <= tol or max_iter refinement sweeps ran (max_iter 0 = plain solve).  b and
 * x are device n x nrhs row-major FP64 with pitch ld. */
int pp_solver_solve(pp_solver* s, const double* b, double* x, int nrhs, int ld, double tol, int max_iter,
                    void* hip_stream, pp_solve_report* rep);
/* r = b - A x with the FP64 H-matrix (device buffers, pitch ld). */
int pp_solver_residual(pp_solver* s, const double* b, const double* x, double* r, int nrhs, int ld, void* hip_stream);
/* Device address of the working (factor) arena and its element size. */
int pp_solver_factor_arena(pp_solver* s, void** dptr, int* elem_bytes);
int pp_solver_destroy(pp_solver* s);
/* Stream of the factorization (for tracing / timing). */
pp_stream* pp_solver_factor_stream(pp_solver* s);
double pp_solver_factor_flops(const pp_solver* s);
double pp_solver_last_factor_ms(const pp_solver* s);
double pp_solver_solve_flops(const pp_solver* s, int nrhs);

/* ------------------------------------------------------------------ construction on the device
 * Fills an H-matrix arena (device, FP64, pp_hmatrix_storage elements) from a
 * kernel function of points (device, n x dim row-major FP64):
 *   PP_KERNEL_EXP1D  A_ij = exp(-|x_i - x_j| / param)          (dim 1)
 *   PP_KERNEL_INV2D  A_ij = 1 / (|p_i - p_j|_2 + param)        (dim 2)
 * plus `shift` on the diagonal.  Dense leaves are evaluated entry by entry.
 * A low-rank leaf of rank r gets a partially pivoted adaptive cross
 * approximation of rank k = min(max(4r, 64, min(rows, columns) / 16), 256,
 * rows, columns), truncated to rank r (shifted CholeskyQR3 of both factors,
 * one-sided Jacobi SVD of the k x k core) when k > r; every leaf is spread over
 * 256-row slices, one workgroup each.  `pivots` (device,
 * optional, pp_hmatrix_build_pivots entries) receives the (row, column) pivot
 * pairs of every cross approximation in leaf order, -1 for steps left empty.
 * Synchronous on hip_stream (NULL = the context's stream); allocates its
 * scratch through the context.
 * Limits: rank <= 256 and <= 65536 rows per low-rank leaf.  No counterpart in
 * the reference, whose H-matrices come from an external compressor; its SVD /
 * sampling routines (src/kernel.cu:1084-1372) are unused there. */
#define PP_KERNEL_EXP1D 0
#define PP_KERNEL_INV2D 1
int pp_hmatrix_build_kernel(pp_ctx* ctx, const pp_hmatrix* h, int kind, const double* points, double param,
                            double shift, double* arena, int32_t* pivots, void* hip_stream);
/* Length of the pivot array of pp_hmatrix_build_kernel (2 x sum of ranks). */
int64_t pp_hmatrix_build_pivots(const pp_hmatrix* h);

/* ------------------------------------------------------------------ pivoting
 * Partial-pivoted LU for dense blocks that are not diagonally dominant (the
 * H-LU path factors its diagonal leaves without pivoting).  Replaces the
 * reference's pivot helpers, src/kernel.cu:958-1051 (warpReduceMax_Index,
 * blockReduceMax_Index, blockSwapRows, blockApplyPivot, resetPivot), which
 * no opcode of the reference dispatches.
 *
 * pp_dense_getrf_pivoted: `count` row-major ny x nx blocks (device memory,
 * pitch ld, `stride` elements apart) are factored in place, P A = L U (L unit
 * lower, U upper).  pivots (device, count x ny) receives each block's
 * permutation in the reference's convention: row i of P A is row p[i] of A.
 * Column j's pivot is its largest |a_ij|, i >= j, the smallest row on ties.
 * info (device, count values): 0, or 1 + the first column whose pivot is
 * zero (that column is left as it is).  precision PP_FP64 or PP_FP32 (the
 * block's element type).  Asynchronous on hip_stream (NULL = the context's
 * stream); one 256-thread workgroup per block. */
int pp_dense_getrf_pivoted(pp_ctx* ctx, int precision, void* a, int ld, int ny, int nx, int count, int64_t stride,
                           int32_t* pivots, int32_t* info, void* hip_stream);
/* Rows of the row-major ny x nx block m (device, pitch ld) permuted in place
 * by pivots (device, ny entries): row r <- row p[r] (recover = 0), or row p[r]
 * <- row r (recover = 1, the inverse) -- blockApplyPivot, src/kernel.cu:
 * 1023-1043.  pivots is copied to the host and checked to be a permutation
 * first (the call waits for hip_stream); ny * element size <= 48 KiB. */
int pp_dense_apply_pivot(pp_ctx* ctx, int precision, void* m, int ld, int nx, int ny, const int32_t* pivots,
                         int recover, void* hip_stream);

/* ------------------------------------------------------------------ byte-string format
 * H-matrices, instruction programs and factorizations as byte strings, for
 * files, caches and the wire.  No counterpart in the reference, which has no
 * persistent format (its docs/CONTRIBUTING.md:27-31 only mentions an external
 * compressor); layout: magic "PPHL", u32 version 1, u32 section count, u32 0,
 * then sections (u32 tag, u32 0, u64 length, payload padded to 8 bytes, u64
 * FNV-1a 64 of the payload), little-endian.  Encoders return the string's
 * length in bytes (negative = error) and write it only when buf is non-NULL
 * and cap is at least that length: call once to size, once to fill.
 * Decoders check the header, every length and checksum and the structure
 * before they build anything, and fail with PP_ERR_INVALID otherwise. */

/* Structure (TREE) and, when arena is non-NULL, the FP64 arena (ARNA). */
int64_t pp_hmatrix_encode(const pp_hmatrix* h, const double* arena, uint8_t* buf, uint64_t cap);
/* Rebuilds the structure; copies the arena into `arena` (capacity
 * arena_cap elements) when non-NULL, which fails if the string holds none.
 * *has_arena (optional) tells whether it does. */
int pp_hmatrix_decode(const uint8_t* buf, uint64_t len, double* arena, uint64_t arena_cap, int* has_arena,
                      pp_hmatrix** out);
/* Instructions, dependencies, schedule and per-worker streams (PROG).  A
 * decoded program answers every pp_program_* query as the encoded one did and
 * uploads with pp_stream_upload_program without rebuilding its DAG and
 * schedule.  Its streams are checked against its instructions; operand
 * offsets are trusted as words given to pp_stream_upload are. */
int64_t pp_program_encode(const pp_program* p, uint8_t* buf, uint64_t cap);
int pp_program_decode(const uint8_t* buf, uint64_t len, pp_program** out);
/* Factor arenas of a solver (FACT, in its storage precision, including the
 * inverse areas): a factorization computed once and loaded later into a
 * solver made over the same matrices, precision and solve mode, which can then
 * solve without factoring.  The decoder rejects a string whose layout,
 * structures or FP64 matrix entries (by checksum) differ from the solver's. */
int64_t pp_solver_encode_factors(pp_solver* s, uint8_t* buf, uint64_t cap);
int pp_solver_decode_factors(pp_solver* s, const uint8_t* buf, uint64_t len);

#ifdef __cplusplus
}
#endif

#endif /* PASTEL_H */
