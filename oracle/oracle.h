/*
 * CPU oracle for the H-LU factor + solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatement of the reference's device-routine semantics
 * (src/kernel.cu) and of its instruction decoder (src/kernel.cu:1384-1522),
 * plus a dense no-pivot LU used as the accuracy oracle.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never does.
 *
 * Conventions: row-major, element (i, j) of X at X[i * ld + j]; nx = columns,
 * ny = rows; op(X) = X^T when its flag is set.
 */
#ifndef PP_ORACLE_H
#define PP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* In-place LU without pivoting, unit-lower L (src/kernel.cu:104-147). */
void or_getrf(double* M, int nx, int ny, int ld);
/* B <- L^{-1} B, L unit lower ny_b x ny_b; b_T: B stored transposed
 * (src/kernel.cu:150-182). */
void or_trsml(double* B, const double* L, int nx_b, int ny_b, int ld_b, int ld_l, int b_T);
/* B <- B U^{-1}, U upper non-unit nx_b x nx_b; b_T: B stored transposed
 * (src/kernel.cu:185-285). */
void or_trsmr(double* B, const double* U, int nx_b, int ny_b, int ld_b, int ld_u, int b_T);
/* B <- U^{-1} B, U upper non-unit ny_b x ny_b (backward substitution). */
void or_trsmu(double* B, const double* U, int nx_b, int ny_b, int ld_b, int ld_u);
void or_trmm(double* B, const double* Tm, int nx_b, int ny_b, int ld_b, int ld_t);
/* M <- beta M + alpha op(A) op(B) (src/kernel.cu:288-664); beta == 0 does
 * not read M. */
void or_gemm(double alpha, double beta, double* M, const double* A, const double* B, int m, int n, int k, int ld_m,
             int ld_a, int ld_b, int a_T, int b_T);
/* Givens QR of Y (ny x k, pitch ld) and the explicit ny x k Q (pitch ldq)
 * (blockGivensRotation / blockGivensRecoverQ, src/kernel.cu:1160-1253). */
void or_givens_qr(double* Y, int ld, int ny, int k, double* Q, int ldq);
/* U1 V1^T <- Q Q^T (U1 V1^T + U2 V2^T), Q = orth((U1 V1^T + U2 V2^T) Omega),
 * Omega = rnd viewed as nx x k1 with pitch k1 (src/kernel.cu:1255-1280). */
void or_accum(double* U1, double* V1, const double* U2, const double* V2, int nx, int ny, int k1, int k2, int ld_u1,
              int ld_v1, int ld_u2, int ld_v2, const double* rnd);

/* Execute one instruction entry (words as in the stream) against a pointer
 * table of host arrays; scratch is per-call workspace (doubles).  Returns the
 * entry length or -1. */
int or_exec(const int32_t* words, double* const* ptrs, double* scratch, const double* rnd);

/* Run per-worker streams honouring their waits.  threads <= 1: one thread
 * interleaves the workers deterministically; otherwise `threads` OS threads
 * share the workers (worker w on thread w % threads, which needs
 * workers <= threads for progress).  flags: n_insts ints, zeroed here.
 * Returns 0, or -1 on deadlock / bad entry. */
int or_run_streams(const int32_t* const* streams, int workers, double* const* ptrs, double* scratch,
                   int64_t scratch_stride, const double* rnd, int n_insts, int threads);

/* Dense FP64 LU (no pivoting) of an n x n row-major matrix in place, blocked,
 * using up to `threads` threads. */
void or_dense_lu(double* A, int n, int threads);
/* X (n x nrhs, pitch ldx) <- (LU)^{-1} X with the packed factors of or_dense_lu. */
void or_dense_lu_solve(const double* LU, int n, double* X, int nrhs, int ldx);
/* R = B - A X for dense A (n x n); returns max over columns of
 * ||R_c|| / ||B_c||. */
double or_dense_residual(const double* A, int n, const double* X, const double* B, int nrhs, int ld);

/* splitmix64 sketch values in [-1, 1) (same stream as the library's). */
void or_sketch(double* out, int64_t count, uint64_t seed);

/* Kernel entry (kind 0: exp(-|x_i - x_j| / param), 1-D points; kind 1:
 * 1 / (|p_i - p_j| + param), 2-D points) + shift where i == j. */
double or_kernel_entry(int kind, const double* pts, int i, int j, double param, double shift);
/* ny x nx block of entries at global (r0, c0), row-major. */
void or_kernel_block(int kind, const double* pts, int r0, int ny, int c0, int nx, double param, double shift,
                     double* out);
/* Fixed-rank partially pivoted cross approximation A ~ U V^T of the kernel
 * block (U ny x r, V nx x r, row-major); piv (optional) gets 2r pivot
 * indices (row, column), -1 for columns left zero. */
void or_aca(int kind, const double* pts, int r0, int ny, int c0, int nx, double param, double shift, int r, double* U,
            double* V, int* piv);
/* The same on an explicit ny x nx row-major matrix. */
void or_aca_dense(const double* A, int ny, int nx, int r, double* U, double* V, int* piv);
/* Truncation of a rank-k cross approximation (U ny x kp, V nx x kp, pitch
 * kp, columns >= k zero; both overwritten) to rank r into Uo (ny x r) and Vo
 * (nx x r) by the device builder's algorithm; 1 when a Cholesky pivot fails. */
int or_truncate(double* U, double* V, int ny, int nx, int kp, int r, double* Uo, double* Vo);

/* Partial-pivoted LU (row-major ny x nx, pitch ld) in place; p: permutation,
   row i of P A is row p[i] of A.  Returns 0 or 1 + first zero-pivot column. */
int or_getrf_pivoted(double* A, int ny, int nx, int ld, int32_t* p);
int or_getrf_pivoted_f32(float* A, int ny, int nx, int ld, int32_t* p);

#ifdef __cplusplus
}
#endif

#endif
