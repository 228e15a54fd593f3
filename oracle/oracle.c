/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Each routine restates the semantics of the reference routine it cites in
 * straightforward loops; no blocking tricks, so a reader can check it against
 * the reference line by line.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

static inline double get_op(const double* X, int ld, int t, int i, int k) {
  return t ? X[(int64_t)k * ld + i] : X[(int64_t)i * ld + k];
}

/* DenseGetrf, src/kernel.cu:104-147: for each pivot i, rows below divide by
 * the pivot and subtract the scaled pivot row from their trailing part. */
void or_getrf(double* M, int nx, int ny, int ld) {
  const int mn = nx < ny ? nx : ny;
  for (int i = 0; i < mn; ++i) {
    const double piv = M[(int64_t)i * ld + i];
    for (int r = i + 1; r < ny; ++r) {
      double* row = &M[(int64_t)r * ld];
      const double l = row[i] / piv;
      row[i] = l;
      for (int c = i + 1; c < nx; ++c) row[c] = fma(M[(int64_t)i * ld + c], -l, row[c]);
    }
  }
}

/* DenseTrsmL, src/kernel.cu:150-182 (forward substitution, unit diagonal). */
void or_trsml(double* B, const double* L, int nx_b, int ny_b, int ld_b, int ld_l, int b_T) {
  for (int i = 0; i < ny_b; ++i)
    for (int r = i + 1; r < ny_b; ++r) {
      const double l = L[(int64_t)r * ld_l + i];
      for (int c = 0; c < nx_b; ++c) {
        double* dst = b_T ? &B[(int64_t)c * ld_b + r] : &B[(int64_t)r * ld_b + c];
        const double src = b_T ? B[(int64_t)c * ld_b + i] : B[(int64_t)i * ld_b + c];
        *dst = fma(src, -l, *dst);
      }
    }
}

/* DenseTrsmR / DenseTrsmR_transposeB, src/kernel.cu:185-285: column i of X is
 * B's column i divided by U[i][i] after subtracting earlier columns. */
void or_trsmr(double* B, const double* U, int nx_b, int ny_b, int ld_b, int ld_u, int b_T) {
  for (int i = 0; i < nx_b; ++i) {
    const double d = U[(int64_t)i * ld_u + i];
    for (int r = 0; r < ny_b; ++r) {
      double* x = b_T ? &B[(int64_t)i * ld_b + r] : &B[(int64_t)r * ld_b + i];
      *x = *x / d;
    }
    for (int j = i + 1; j < nx_b; ++j) {
      const double u = U[(int64_t)i * ld_u + j];
      for (int r = 0; r < ny_b; ++r) {
        const double xi = b_T ? B[(int64_t)i * ld_b + r] : B[(int64_t)r * ld_b + i];
        double* dst = b_T ? &B[(int64_t)j * ld_b + r] : &B[(int64_t)r * ld_b + j];
        *dst = fma(xi, -u, *dst);
      }
    }
  }
}

/* Backward substitution with the non-unit upper factor (no reference
 * counterpart: the reference has no solve). */
void or_trsmu(double* B, const double* U, int nx_b, int ny_b, int ld_b, int ld_u) {
  for (int i = ny_b - 1; i >= 0; --i) {
    const double d = U[(int64_t)i * ld_u + i];
    for (int c = 0; c < nx_b; ++c) B[(int64_t)i * ld_b + c] /= d;
    for (int r = 0; r < i; ++r) {
      const double u = U[(int64_t)r * ld_u + i];
      for (int c = 0; c < nx_b; ++c) B[(int64_t)r * ld_b + c] = fma(B[(int64_t)i * ld_b + c], -u, B[(int64_t)r * ld_b + c]);
    }
  }
}

/* B <- T B with a dense ny_b x ny_b T (no reference counterpart: applies the
 * explicit leaf inverses of the inverse-based solve). */
void or_trmm(double* B, const double* Tm, int nx_b, int ny_b, int ld_b, int ld_t) {
  double* X = (double*)malloc(sizeof(double) * (size_t)ny_b * (size_t)(nx_b > 0 ? nx_b : 1));
  for (int i = 0; i < ny_b; ++i)
    for (int c = 0; c < nx_b; ++c) X[(int64_t)i * nx_b + c] = B[(int64_t)i * ld_b + c];
  for (int i = 0; i < ny_b; ++i)
    for (int c = 0; c < nx_b; ++c) {
      double s = 0.;
      for (int k = 0; k < ny_b; ++k) s += Tm[(int64_t)i * ld_t + k] * X[(int64_t)k * nx_b + c];
      B[(int64_t)i * ld_b + c] = s;
    }
  free(X);
}

/* blockDenseGemm, src/kernel.cu:415-664: M = beta M + alpha op(A) op(B). */
void or_gemm(double alpha, double beta, double* M, const double* A, const double* B, int m, int n, int k, int ld_m,
             int ld_a, int ld_b, int a_T, int b_T) {
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0.;
      for (int kk = 0; kk < k; ++kk) s += get_op(A, ld_a, a_T, i, kk) * get_op(B, ld_b, !b_T, j, kk);
      double* c = &M[(int64_t)i * ld_m + j];
      *c = (beta == 0.) ? alpha * s : beta * *c + alpha * s;
    }
}

/* Givens QR, src/kernel.cu:1160-1253: every column is cleared bottom-up by
 * rotations of adjacent rows (the reference runs the same rotations in a
 * wavefront order across columns; rotations on disjoint rows commute).  The
 * explicit Q applies the transposed rotations to [I; 0] in reverse. */
void or_givens_qr(double* Y, int ld, int ny, int k, double* Q, int ldq) {
  const int nrot_max = (ny > 0 ? ny : 1) * (k > 0 ? k : 1);
  double* cs = (double*)malloc(sizeof(double) * 2 * nrot_max);
  int* rows = (int*)malloc(sizeof(int) * nrot_max);
  int nrot = 0;
  for (int c = 0; c < k && c < ny; ++c)
    for (int r = ny - 2; r >= c; --r) {
      const double a = Y[(int64_t)r * ld + c], b = Y[(int64_t)(r + 1) * ld + c];
      double co, si;
      if (b == 0.) {
        co = 1.;
        si = 0.;
      } else {
        const double h = hypot(a, b);
        co = a / h;
        si = b / h;
      }
      /* [a'; b'] = [co si; -si co] [a; b] zeroes b */
      for (int j = c; j < k; ++j) {
        const double x = Y[(int64_t)r * ld + j], y = Y[(int64_t)(r + 1) * ld + j];
        Y[(int64_t)r * ld + j] = co * x + si * y;
        Y[(int64_t)(r + 1) * ld + j] = -si * x + co * y;
      }
      cs[2 * nrot] = co;
      cs[2 * nrot + 1] = si;
      rows[nrot++] = r;
    }
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < k; ++j) Q[(int64_t)i * ldq + j] = (i == j) ? 1. : 0.;
  for (int t = nrot - 1; t >= 0; --t) {
    const int r = rows[t];
    const double co = cs[2 * t], si = cs[2 * t + 1];
    for (int j = 0; j < k; ++j) {
      const double x = Q[(int64_t)r * ldq + j], y = Q[(int64_t)(r + 1) * ldq + j];
      Q[(int64_t)r * ldq + j] = co * x - si * y;
      Q[(int64_t)(r + 1) * ldq + j] = si * x + co * y;
    }
  }
  free(cs);
  free(rows);
}

/* blockLowRankAccum, src/kernel.cu:1255-1280. */
void or_accum(double* U1, double* V1, const double* U2, const double* V2, int nx, int ny, int k1, int k2, int ld_u1,
              int ld_v1, int ld_u2, int ld_v2, const double* rnd) {
  const int kmax = k1 > k2 ? k1 : k2;
  double* S = (double*)malloc(sizeof(double) * k1 * kmax);
  double* Y = (double*)malloc(sizeof(double) * (size_t)ny * k1);
  double* Q = (double*)malloc(sizeof(double) * (size_t)ny * k1);
  double* Vn = (double*)malloc(sizeof(double) * (size_t)nx * k1);
  /* Y = (U1 V1^T + U2 V2^T) Omega */
  or_gemm(1., 0., S, rnd, V1, k1, k1, nx, k1, k1, ld_v1, 1, 0);
  or_gemm(1., 0., Y, U1, S, ny, k1, k1, k1, ld_u1, k1, 0, 1);
  or_gemm(1., 0., S, rnd, V2, k1, k2, nx, k2, k1, ld_v2, 1, 0);
  or_gemm(1., 1., Y, U2, S, ny, k1, k2, k1, ld_u2, k2, 0, 1);
  or_givens_qr(Y, k1, ny, k1, Q, k1);
  /* V1 <- V1 (Q^T U1)^T + V2 (Q^T U2)^T, U1 <- Q */
  or_gemm(1., 0., S, Q, U1, k1, k1, ny, k1, k1, ld_u1, 1, 0);
  or_gemm(1., 0., Vn, V1, S, nx, k1, k1, k1, ld_v1, k1, 0, 1);
  or_gemm(1., 0., S, Q, U2, k1, k2, ny, k2, k1, ld_u2, 1, 0);
  or_gemm(1., 1., Vn, V2, S, nx, k1, k2, k1, ld_v2, k2, 0, 1);
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < k1; ++j) U1[(int64_t)i * ld_u1 + j] = Q[(int64_t)i * k1 + j];
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < k1; ++j) V1[(int64_t)i * ld_v1 + j] = Vn[(int64_t)i * k1 + j];
  free(S);
  free(Y);
  free(Q);
  free(Vn);
}

/* Decoder of src/kernel.cu:1393-1501 with the word offsets of the
 * instruction layout (opcode, id, operation, operands...). */
int or_exec(const int32_t* w, double* const* P, double* tmp, const double* rnd) {
  switch (w[2]) {
    case 0: /* nop */
      return 3;
    case 1: /* getrf */
      or_getrf(P[w[3]] + w[4], w[5], w[6], w[7]);
      return 8;
    case 2: /* trsml */
      or_trsml(P[w[3]] + w[5], P[w[4]] + w[6], w[7], w[8], w[10], w[11], w[12]);
      return 13;
    case 3: /* trsmr */
      or_trsmr(P[w[3]] + w[5], P[w[4]] + w[6], w[7], w[8], w[10], w[11], w[12]);
      return 13;
    case 9: /* trsmu */
      or_trsmu(P[w[3]] + w[5], P[w[4]] + w[6], w[7], w[8], w[10], w[11]);
      return 13;
    case 10: /* trmm */
      or_trmm(P[w[3]] + w[5], P[w[4]] + w[6], w[7], w[8], w[9], w[10]);
      return 11;
    case 4: /* gemm: M -= A B */
    case 5: /* gemm_plus: M += A B */
      or_gemm(w[2] == 4 ? -1. : 1., 1., P[w[3]] + w[6], P[w[4]] + w[7], P[w[5]] + w[8], w[9], w[10], w[11], w[12],
              w[13], w[14], w[15], w[16]);
      return 17;
    case 6: { /* gemm_3x: M -= A B C */
      double *M = P[w[3]] + w[7], *A = P[w[4]] + w[8], *B = P[w[5]] + w[9], *C = P[w[6]] + w[10];
      const int m = w[11], n = w[12], k = w[13], l = w[14];
      if (w[22]) {
        or_gemm(1., 0., tmp, A, B, m, l, k, l, w[16], w[17], w[19], w[20]);
        or_gemm(-1., 1., M, tmp, C, m, n, l, w[15], l, w[18], 0, w[21]);
      } else {
        or_gemm(1., 0., tmp, B, C, k, n, l, n, w[17], w[18], w[20], w[21]);
        or_gemm(-1., 1., M, A, tmp, m, n, k, w[15], w[16], n, w[19], 0);
      }
      return 23;
    }
    case 7: { /* gemm_4x: M -= A B C D, association by control (src/kernel.cu:914-953) */
      double *M = P[w[3]] + w[8], *A = P[w[4]] + w[9], *B = P[w[5]] + w[10], *C = P[w[6]] + w[11], *D = P[w[7]] + w[12];
      const int m = w[13], n = w[14], k = w[15], l = w[16], o = w[17];
      const int ldm = w[18], lda = w[19], ldb = w[20], ldc = w[21], ldd = w[22];
      const int aT = w[23], bT = w[24], cT = w[25], dT = w[26];
      double *t1 = tmp, *t2 = tmp + w[28];
      switch (w[27]) {
        case 0:
          or_gemm(1., 0., t1, A, B, m, l, k, l, lda, ldb, aT, bT);
          or_gemm(1., 0., t2, t1, C, m, o, l, o, l, ldc, 0, cT);
          or_gemm(-1., 1., M, t2, D, m, n, o, ldm, o, ldd, 0, dT);
          break;
        case 1:
          or_gemm(1., 0., t1, B, C, k, o, l, o, ldb, ldc, bT, cT);
          or_gemm(1., 0., t2, A, t1, m, o, k, o, lda, o, aT, 0);
          or_gemm(-1., 1., M, t2, D, m, n, o, ldm, o, ldd, 0, dT);
          break;
        case 2:
          or_gemm(1., 0., t1, B, C, k, o, l, o, ldb, ldc, bT, cT);
          or_gemm(1., 0., t2, t1, D, k, n, o, n, o, ldd, 0, dT);
          or_gemm(-1., 1., M, A, t2, m, n, k, ldm, lda, n, aT, 0);
          break;
        case 3:
          or_gemm(1., 0., t1, C, D, l, n, o, n, ldc, ldd, cT, dT);
          or_gemm(1., 0., t2, B, t1, k, n, l, n, ldb, n, bT, 0);
          or_gemm(-1., 1., M, A, t2, m, n, k, ldm, lda, n, aT, 0);
          break;
        case 4:
          or_gemm(1., 0., t1, A, B, m, l, k, l, lda, ldb, aT, bT);
          or_gemm(1., 0., t2, C, D, l, n, o, n, ldc, ldd, cT, dT);
          or_gemm(-1., 1., M, t1, t2, m, n, l, ldm, l, n, 0, 0);
          break;
        default:
          return -1;
      }
      return 29;
    }
    case 8: /* accum */
      or_accum(P[w[3]] + w[7], P[w[4]] + w[8], P[w[5]] + w[9], P[w[6]] + w[10], w[11], w[12], w[13], w[14], w[15],
               w[16], w[17], w[18], rnd);
      return 21;
    default:
      return -1;
  }
}

/* ------------------------------------------------------------------ stream runner */

typedef struct {
  const int32_t* const* streams;
  int workers, nthreads, tid;
  double* const* ptrs;
  double* scratch;
  int64_t stride;
  const double* rnd;
  atomic_int* flags;
  atomic_int* failed;
} runner_t;

/* Advance worker w from *pc until it blocks on an unmet wait (returns 0),
 * finishes (returns 1), or fails (-1).  Progress flag set when anything ran. */
static int step_worker(runner_t* r, int w, int64_t* pc, int* progressed) {
  const int32_t* s = r->streams[w];
  for (;;) {
    const int32_t op = s[*pc];
    if (op == 0) return 1;
    if (op == 2) {
      if (atomic_load_explicit(&r->flags[s[*pc + 1]], memory_order_acquire) == 0) return 0;
      *pc += 3;
      *progressed = 1;
      continue;
    }
    if (op != 1) return -1;
    const int len = or_exec(&s[*pc], r->ptrs, r->scratch + (int64_t)w * r->stride, r->rnd);
    if (len < 0) return -1;
    atomic_store_explicit(&r->flags[s[*pc + 1]], 1, memory_order_release);
    *pc += len;
    *progressed = 1;
  }
}

static void* thread_main(void* arg) {
  runner_t* r = (runner_t*)arg;
  int nmine = 0;
  for (int w = r->tid; w < r->workers; w += r->nthreads) ++nmine;
  int64_t* pcs = (int64_t*)calloc(nmine > 0 ? nmine : 1, sizeof(int64_t));
  int* done = (int*)calloc(nmine > 0 ? nmine : 1, sizeof(int));
  int left = nmine;
  long idle = 0;
  while (left > 0 && !atomic_load(r->failed)) {
    int progressed = 0;
    for (int q = 0; q < nmine; ++q) {
      if (done[q]) continue;
      const int st = step_worker(r, r->tid + q * r->nthreads, &pcs[q], &progressed);
      if (st < 0) {
        atomic_store(r->failed, 1);
        break;
      }
      if (st == 1) {
        done[q] = 1;
        --left;
      }
    }
    if (!progressed) {
      if (r->nthreads == 1 || ++idle > 200000000L) {
        atomic_store(r->failed, 1); /* single thread without progress: deadlock */
        break;
      }
      sched_yield();
    } else {
      idle = 0;
    }
  }
  free(pcs);
  free(done);
  return NULL;
}

int or_run_streams(const int32_t* const* streams, int workers, double* const* ptrs, double* scratch,
                   int64_t scratch_stride, const double* rnd, int n_insts, int threads) {
  if (threads < 1) threads = 1;
  if (threads > workers) threads = workers;
  atomic_int* flags = (atomic_int*)calloc(n_insts > 0 ? n_insts : 1, sizeof(atomic_int));
  atomic_int failed;
  atomic_init(&failed, 0);
  runner_t* rs = (runner_t*)calloc(threads, sizeof(runner_t));
  pthread_t* th = (pthread_t*)calloc(threads, sizeof(pthread_t));
  for (int t = 0; t < threads; ++t) {
    rs[t] = (runner_t){streams, workers, threads, t, ptrs, scratch, scratch_stride, rnd, flags, &failed};
    if (threads == 1) thread_main(&rs[t]);
    else pthread_create(&th[t], NULL, thread_main, &rs[t]);
  }
  if (threads > 1)
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  const int bad = atomic_load(&failed);
  free(flags);
  free(rs);
  free(th);
  return bad ? -1 : 0;
}

/* ------------------------------------------------------------------ dense oracle */

typedef struct {
  double* A;
  int n, p, pb, t0, t1;
} lu_task_t;

static void* lu_trailing(void* arg) {
  lu_task_t* t = (lu_task_t*)arg;
  double* A = t->A;
  const int n = t->n, p = t->p, pe = t->p + t->pb;
  for (int r = t->t0; r < t->t1; ++r) {
    double* row = &A[(int64_t)r * n];
    for (int k = p; k < pe; ++k) {
      const double l = row[k];
      const double* prow = &A[(int64_t)k * n];
      for (int c = pe; c < n; ++c) row[c] = fma(prow[c], -l, row[c]);
    }
  }
  return NULL;
}

void or_dense_lu(double* A, int n, int threads) {
  const int pb = 64;
  if (threads < 1) threads = 1;
  pthread_t th[256];
  lu_task_t tk[256];
  if (threads > 256) threads = 256;
  for (int p = 0; p < n; p += pb) {
    const int w = (n - p) < pb ? (n - p) : pb;
    /* panel: unblocked LU of A[p:n, p:p+w] plus the U rows A[p:p+w, p+w:n] */
    for (int i = p; i < p + w; ++i) {
      const double piv = A[(int64_t)i * n + i];
      for (int r = i + 1; r < n; ++r) {
        double* row = &A[(int64_t)r * n];
        const double l = row[i] / piv;
        row[i] = l;
        const int cend = r < p + w ? n : p + w; /* rows inside the panel update everything */
        for (int c = i + 1; c < cend; ++c) row[c] = fma(A[(int64_t)i * n + c], -l, row[c]);
      }
    }
    /* trailing update of rows p+w.. with the panel */
    const int rows = n - p - w;
    if (rows <= 0) continue;
    const int nt = rows < threads ? rows : threads;
    for (int t = 0; t < nt; ++t) {
      tk[t] = (lu_task_t){A, n, p, w, p + w + (int)((int64_t)rows * t / nt), p + w + (int)((int64_t)rows * (t + 1) / nt)};
      if (nt == 1) lu_trailing(&tk[t]);
      else pthread_create(&th[t], NULL, lu_trailing, &tk[t]);
    }
    if (nt > 1)
      for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
  }
}

void or_dense_lu_solve(const double* LU, int n, double* X, int nrhs, int ldx) {
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < i; ++k) {
      const double l = LU[(int64_t)i * n + k];
      for (int c = 0; c < nrhs; ++c) X[(int64_t)i * ldx + c] -= l * X[(int64_t)k * ldx + c];
    }
  for (int i = n - 1; i >= 0; --i) {
    for (int k = i + 1; k < n; ++k) {
      const double u = LU[(int64_t)i * n + k];
      for (int c = 0; c < nrhs; ++c) X[(int64_t)i * ldx + c] -= u * X[(int64_t)k * ldx + c];
    }
    const double d = LU[(int64_t)i * n + i];
    for (int c = 0; c < nrhs; ++c) X[(int64_t)i * ldx + c] /= d;
  }
}

double or_dense_residual(const double* A, int n, const double* X, const double* B, int nrhs, int ld) {
  double worst = 0.;
  double* rn = (double*)calloc(nrhs, sizeof(double));
  double* bn = (double*)calloc(nrhs, sizeof(double));
  for (int i = 0; i < n; ++i)
    for (int c = 0; c < nrhs; ++c) {
      double s = B[(int64_t)i * ld + c];
      for (int k = 0; k < n; ++k) s -= A[(int64_t)i * n + k] * X[(int64_t)k * ld + c];
      rn[c] += s * s;
      bn[c] += B[(int64_t)i * ld + c] * B[(int64_t)i * ld + c];
    }
  for (int c = 0; c < nrhs; ++c) {
    const double r = bn[c] > 0 ? sqrt(rn[c] / bn[c]) : sqrt(rn[c]);
    if (r > worst) worst = r;
  }
  free(rn);
  free(bn);
  return worst;
}

void or_sketch(double* out, int64_t count, uint64_t seed) {
  uint64_t st = seed;
  for (int64_t i = 0; i < count; ++i) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    out[i] = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
  }
}

/* ------------------------------------------------------------------ H-matrix construction
 * CPU restatement of the device builder (pastel-palettes_amd/csrc/kernels/
 * build.hip): kernel entries and fixed-rank partially pivoted adaptive cross
 * approximation.  The reference has no construction code on its path (its
 * compressors, src/kernel.cu:1084-1372, are never dispatched), so this
 * follows the textbook algorithm (Bebendorf 2000, "Approximation of boundary
 * element matrices") with this build's tie and zero-row rules.  Compiled
 * with contraction off (-std=c11), with fma() where the device uses it. */

double or_kernel_entry(int kind, const double* pts, int i, int j, double param, double shift) {
  double a;
  if (kind == 0) {
    const double d = fabs(pts[i] - pts[j]);
    a = exp(-d / param);
  } else {
    const double dx = pts[2 * i] - pts[2 * j], dy = pts[2 * i + 1] - pts[2 * j + 1];
    const double s = dx * dx + dy * dy;
    a = 1.0 / (sqrt(s) + param);
  }
  if (i == j) a = a + shift;
  return a;
}

void or_kernel_block(int kind, const double* pts, int r0, int ny, int c0, int nx, double param, double shift,
                     double* out) {
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < nx; ++j) out[(int64_t)i * nx + j] = or_kernel_entry(kind, pts, r0 + i, c0 + j, param, shift);
}

typedef struct {
  int kind, r0, c0;
  const double* pts;
  double param, shift;
  const double* dense; /* explicit matrix (ld nx) instead of a kernel */
  int ld;
} aca_src;

static double aca_entry(const aca_src* s, int t, int j) {
  if (s->dense) return s->dense[(int64_t)t * s->ld + j];
  return or_kernel_entry(s->kind, s->pts, s->r0 + t, s->c0 + j, s->param, s->shift);
}

static void aca_run(const aca_src* s, int ny, int nx, int r, double* U, double* V, int* piv) {
  char* used = (char*)calloc((size_t)ny, 1);
  int i = 0, k = 0;
  while (k < r) {
    /* residual row i into V[:, k]; column of largest magnitude (first on ties) */
    int jp = -1;
    double best = -1.0;
    for (int j = 0; j < nx; ++j) {
      double a = aca_entry(s, i, j);
      for (int l = 0; l < k; ++l) a = fma(-U[(int64_t)i * r + l], V[(int64_t)j * r + l], a);
      V[(int64_t)j * r + k] = a;
      if (fabs(a) > best) {
        best = fabs(a);
        jp = j;
      }
    }
    used[i] = 1;
    if (best == 0.0) { /* row represented exactly: the step stays empty (zero
                          column, pivots -1) and the next starts at the
                          smallest unused row */
      for (int t = 0; t < ny; ++t) U[(int64_t)t * r + k] = 0.0;
      if (piv) piv[2 * k] = piv[2 * k + 1] = -1;
      ++k;
      i = -1;
      for (int t = 0; t < ny; ++t)
        if (!used[t]) {
          i = t;
          break;
        }
      if (i < 0) break;
      continue;
    }
    const double d = V[(int64_t)jp * r + k];
    for (int j = 0; j < nx; ++j) V[(int64_t)j * r + k] = V[(int64_t)j * r + k] / d;
    /* residual column jp into U[:, k]; next row: largest unused (first on ties) */
    int nxt = -1;
    best = -1.0;
    for (int t = 0; t < ny; ++t) {
      double a = aca_entry(s, t, jp);
      for (int l = 0; l < k; ++l) a = fma(-U[(int64_t)t * r + l], V[(int64_t)jp * r + l], a);
      U[(int64_t)t * r + k] = a;
      if (!used[t] && fabs(a) > best) {
        best = fabs(a);
        nxt = t;
      }
    }
    if (piv) {
      piv[2 * k] = i;
      piv[2 * k + 1] = jp;
    }
    ++k;
    i = nxt;
    if (i < 0) break;
  }
  for (int c = k; c < r; ++c) {
    for (int t = 0; t < ny; ++t) U[(int64_t)t * r + c] = 0.0;
    for (int j = 0; j < nx; ++j) V[(int64_t)j * r + c] = 0.0;
    if (piv) piv[2 * c] = piv[2 * c + 1] = -1;
  }
  free(used);
}

void or_aca(int kind, const double* pts, int r0, int ny, int c0, int nx, double param, double shift, int r, double* U,
            double* V, int* piv) {
  aca_src s = {kind, r0, c0, pts, param, shift, NULL, 0};
  aca_run(&s, ny, nx, r, U, V, piv);
}

void or_aca_dense(const double* A, int ny, int nx, int r, double* U, double* V, int* piv) {
  aca_src s = {0, 0, 0, NULL, 0.0, 0.0, A, nx};
  aca_run(&s, ny, nx, r, U, V, piv);
}

/* Truncation of a rank-k cross approximation (U: ny x kp, V: nx x kp,
 * pitch kp, columns >= k zero; both overwritten) to rank r (Uo: ny x r,
 * Vo: nx x r), as the device builder does it: unit columns, shifted
 * CholeskyQR3 of each factor, one-sided Jacobi SVD of the kp x kp core with
 * cyclic round-robin pairs, the r columns of largest norm.  Column norms and
 * Gram matrices are sums over 256-row slices (each slice summed first, the
 * slices then added in order); X <- X R^{-1} through R^{-1} formed column by
 * column; the Jacobi dot products are 64 lane sums combined by an xor
 * butterfly.  Returns 1 when a Cholesky pivot is not positive, else 0. */
#define OR_SLICE 256 /* rows per partial sum of the device builder */

/* sum of 64 lane partials (lane l: elements l, l + 64, ... in order) by an
 * xor butterfly, as a 64-wide wavefront reduces them */
static double lane_dot(const double* x, const double* y, int n) {
  double p[64], q[64];
  for (int l = 0; l < 64; ++l) {
    double acc = 0.0;
    for (int i = l; i < n; i += 64) acc = fma(x[i], y[i], acc);
    p[l] = acc;
  }
  for (int o = 32; o >= 1; o >>= 1) {
    for (int l = 0; l < 64; ++l) q[l] = p[l] + p[l ^ o];
    memcpy(p, q, sizeof(p));
  }
  return p[0];
}

static int trunc_orth(double* X, int m, int kp, double* d, double* G, double* Rt, double* T, double* Ri) {
  for (int c = 0; c < kp; ++c) {
    double tot = 0.0;
    for (int s0 = 0; s0 < m; s0 += OR_SLICE) {
      double acc = 0.0;
      for (int t = s0; t < m && t < s0 + OR_SLICE; ++t) acc = fma(X[(int64_t)t * kp + c], X[(int64_t)t * kp + c], acc);
      tot = tot + acc;
    }
    d[c] = sqrt(tot);
  }
  for (int t = 0; t < m; ++t)
    for (int c = 0; c < kp; ++c)
      if (d[c] > 0.0) X[(int64_t)t * kp + c] = X[(int64_t)t * kp + c] / d[c];
  for (int e = 0; e < kp * kp; ++e) Rt[e] = (e / kp == e % kp) ? 1.0 : 0.0;
  for (int pass = 0; pass < 3; ++pass) {
    for (int a = 0; a < kp; ++a)
      for (int b = 0; b < kp; ++b) {
        double tot = 0.0;
        for (int s0 = 0; s0 < m; s0 += OR_SLICE) {
          double acc = 0.0;
          for (int t = s0; t < m && t < s0 + OR_SLICE; ++t)
            acc = fma(X[(int64_t)t * kp + a], X[(int64_t)t * kp + b], acc);
          tot = tot + acc;
        }
        G[a * kp + b] = tot;
      }
    double tr = 0.0;
    for (int c = 0; c < kp; ++c) tr = tr + G[c * kp + c];
    const double shift = pass == 0 ? 11.0 * ((double)m * kp + (double)kp * (kp + 1)) * 0x1.0p-53 * tr : 0.0;
    for (int c = 0; c < kp; ++c) G[c * kp + c] = d[c] > 0.0 ? G[c * kp + c] + shift : 1.0;
    for (int j = 0; j < kp; ++j) {
      const double piv = G[j * kp + j];
      if (!(piv > 0.0)) return 1;
      const double rjj = sqrt(piv);
      for (int c = j; c < kp; ++c) G[j * kp + c] = c == j ? rjj : G[j * kp + c] / rjj;
      for (int a = j + 1; a < kp; ++a)
        for (int c = a; c < kp; ++c) G[a * kp + c] = fma(-G[j * kp + a], G[j * kp + c], G[a * kp + c]);
    }
    /* R^{-1} column by column, then X <- X R^{-1} (sums over a <= c) */
    for (int c = 0; c < kp; ++c) {
      for (int a = c + 1; a < kp; ++a) Ri[a * kp + c] = 0.0;
      Ri[c * kp + c] = 1.0 / G[c * kp + c];
      for (int a = c - 1; a >= 0; --a) {
        double v = 0.0;
        for (int b = a + 1; b <= c; ++b) v = fma(G[a * kp + b], Ri[b * kp + c], v);
        Ri[a * kp + c] = -v / G[a * kp + a];
      }
    }
    double* row = (double*)malloc(sizeof(double) * kp);
    for (int t = 0; t < m; ++t) {
      double* x = X + (int64_t)t * kp;
      memcpy(row, x, sizeof(double) * kp);
      for (int c = 0; c < kp; ++c) {
        double v = 0.0;
        for (int a = 0; a <= c; ++a) v = fma(row[a], Ri[a * kp + c], v);
        x[c] = v;
      }
    }
    free(row);
    for (int a = 0; a < kp; ++a)
      for (int c = 0; c < kp; ++c) {
        double v = 0.0;
        for (int b = a; b <= c; ++b) v = fma(G[a * kp + b], Rt[b * kp + c], v);
        T[a * kp + c] = c >= a ? v : 0.0;
      }
    memcpy(Rt, T, sizeof(double) * kp * kp);
  }
  return 0;
}

int or_truncate(double* U, double* V, int ny, int nx, int kp, int r, double* Uo, double* Vo) {
  const size_t q = (size_t)kp * kp;
  double* w = (double*)calloc(7 * q + 2 * (size_t)kp, sizeof(double));
  double *Ru = w, *Rv = w + q, *G = w + 2 * q, *T = w + 3 * q, *N = w + 4 * q, *J = w + 5 * q, *Ri = w + 6 * q,
         *du = w + 7 * q, *dv = du + kp;
  if (trunc_orth(U, ny, kp, du, G, Ru, T, Ri) || trunc_orth(V, nx, kp, dv, G, Rv, T, Ri)) {
    free(w);
    return 1;
  }
  for (int b = 0; b < kp; ++b)
    for (int a = 0; a < kp; ++a) {
      double v = 0.0;
      for (int c = a > b ? a : b; c < kp; ++c) v = fma(Ru[a * kp + c] * du[c], Rv[b * kp + c] * dv[c], v);
      N[b * kp + a] = v;
      J[b * kp + a] = a == b ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 30; ++sweep) {
    int any = 0;
    for (int rd = 0; rd < kp - 1; ++rd)
      for (int i0 = 0; i0 < kp / 2; ++i0) {
        const int i1 = kp - 1 - i0;
        const int p0 = i0 == 0 ? 0 : 1 + (i0 - 1 + rd) % (kp - 1);
        const int p1 = 1 + (i1 - 1 + rd) % (kp - 1);
        const int p = p0 < p1 ? p0 : p1, qq = p0 < p1 ? p1 : p0;
        double* mp = N + (size_t)p * kp;
        double* mq = N + (size_t)qq * kp;
        const double al = lane_dot(mp, mp, kp), be = lane_dot(mq, mq, kp), ga = lane_dot(mp, mq, kp);
        if (ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
          const double zeta = (be - al) / (2.0 * ga);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
          double* jp = J + (size_t)p * kp;
          double* jq = J + (size_t)qq * kp;
          for (int i = 0; i < kp; ++i) {
            const double x = mp[i], y = mq[i];
            mp[i] = c * x - sn * y;
            mq[i] = sn * x + c * y;
            const double u = jp[i], v = jq[i];
            jp[i] = c * u - sn * v;
            jq[i] = sn * u + c * v;
          }
          any = 1;
        }
      }
    if (!any) break;
  }
  double* sig = (double*)malloc(sizeof(double) * kp);
  int* sel = (int*)malloc(sizeof(int) * (r > 0 ? r : 1));
  for (int c = 0; c < kp; ++c) {
    double acc = 0.0;
    for (int i = 0; i < kp; ++i) acc = fma(N[(size_t)c * kp + i], N[(size_t)c * kp + i], acc);
    sig[c] = sqrt(acc);
  }
  for (int o = 0; o < r; ++o) {
    int b = -1;
    for (int c = 0; c < kp; ++c)
      if (sig[c] >= 0.0 && (b < 0 || sig[c] > sig[b])) b = c;
    sel[o] = b;
    if (b >= 0) sig[b] = -1.0;
  }
  for (int t = 0; t < ny; ++t)
    for (int o = 0; o < r; ++o) {
      double v = 0.0;
      if (sel[o] >= 0)
        for (int a = 0; a < kp; ++a) v = fma(U[(int64_t)t * kp + a], N[(size_t)sel[o] * kp + a], v);
      Uo[(int64_t)t * r + o] = v;
    }
  for (int t = 0; t < nx; ++t)
    for (int o = 0; o < r; ++o) {
      double v = 0.0;
      if (sel[o] >= 0)
        for (int a = 0; a < kp; ++a) v = fma(V[(int64_t)t * kp + a], J[(size_t)sel[o] * kp + a], v);
      Vo[(int64_t)t * r + o] = v;
    }
  free(sig);
  free(sel);
  free(w);
  return 0;
}

/* ---------------------------------------------------------------- pivoting
 * Partial-pivoted LU of a row-major ny x nx block (pitch ld) in place:
 * P A = L U, L unit lower (multipliers below the diagonal), U upper.
 * The pivot vector is a permutation in the reference's convention
 * (blockApplyPivot, src/kernel.cu:1023-1043: row i of P A is row p[i] of A;
 * resetPivot, src/kernel.cu:1046-1050, starts it as the identity).  Column j's
 * pivot is the row of largest |a_ij|, i >= j (blockReduceMax_Index,
 * src/kernel.cu:974-1005); ties go to the smallest row index (the
 * reference's reduction breaks ties by lane, which depends on its block
 * size).  Rows are exchanged whole (blockSwapRows, src/kernel.cu:1008-1012).
 * Multipliers are l = a_ij / u_jj (correctly rounded division) and each
 * trailing element takes one fused multiply-add per step, a_ik <- fma(-l,
 * u_jk, a_ik), so any implementation that keeps those operations is
 * bit-identical.  A zero pivot column is left as it is (every entry at or
 * below the diagonal is zero, so the update would change nothing).
 * Returns 0, or 1 + the first column whose pivot is zero. */
#define OR_GETRF_PIVOTED(NAME, T, FMA, ABS)                                            \
  int NAME(T* A, int ny, int nx, int ld, int32_t* p) {                                 \
    int info = 0;                                                                      \
    const int K = ny < nx ? ny : nx;                                                   \
    for (int i = 0; i < ny; ++i) p[i] = i;                                             \
    for (int j = 0; j < K; ++j) {                                                      \
      int best = j;                                                                    \
      T bv = ABS(A[(int64_t)j * ld + j]);                                              \
      for (int i = j + 1; i < ny; ++i) {                                               \
        const T v = ABS(A[(int64_t)i * ld + j]);                                       \
        if (v > bv) {                                                                  \
          bv = v;                                                                      \
          best = i;                                                                    \
        }                                                                              \
      }                                                                                \
      if (best != j) {                                                                 \
        for (int k = 0; k < nx; ++k) {                                                 \
          const T t = A[(int64_t)j * ld + k];                                          \
          A[(int64_t)j * ld + k] = A[(int64_t)best * ld + k];                          \
          A[(int64_t)best * ld + k] = t;                                               \
        }                                                                              \
        const int32_t t = p[j];                                                        \
        p[j] = p[best];                                                                \
        p[best] = t;                                                                   \
      }                                                                                \
      const T piv = A[(int64_t)j * ld + j];                                            \
      if (piv == (T)0) {                                                               \
        if (info == 0) info = j + 1;                                                   \
        continue;                                                                      \
      }                                                                                \
      for (int i = j + 1; i < ny; ++i) {                                               \
        T* row = &A[(int64_t)i * ld];                                                  \
        const T l = row[j] / piv;                                                      \
        row[j] = l;                                                                    \
        for (int k = j + 1; k < nx; ++k) row[k] = FMA(-l, A[(int64_t)j * ld + k], row[k]); \
      }                                                                                \
    }                                                                                  \
    return info;                                                                       \
  }
OR_GETRF_PIVOTED(or_getrf_pivoted, double, fma, fabs)
OR_GETRF_PIVOTED(or_getrf_pivoted_f32, float, fmaf, fabsf)
