// Expansion of one execute entry into primitive workgroup steps.
//
// Every operation of the instruction stream (reference opcode semantics:
// src/kernel.cu:1375-1526 dispatch, routines cited in dev_lu.h) is a short
// sequence of primitive steps: a GEMM on MFMA, a panel LU, the inverses of
// diagonal blocks, a Householder QR or a 2-D copy.  make_step(s) returns the
// s-th step of the entry, computed from the instruction words alone, and the
// VM executes the steps through one switch (exec_step).  Each primitive is
// therefore inlined once per kernel: no device-function calls (and no
// callee-saved register spills) on the hot path, and every LDS access is
// addressed as LDS.
#pragma once

#include "../h_ops/h_ops.h"
#include "dev_accum.h"
#include "dev_lds_lu.h"
#include "dev_lu.h"

namespace pp {

__device__ __forceinline__ int32_t word(const int32_t* pc, int i) { return __builtin_amdgcn_readfirstlane(pc[i]); }

template <typename T>
__device__ __forceinline__ T* slot(void* const* ptrs, const int32_t* pc, int i) {
  return reinterpret_cast<T*>(ptrs[word(pc, i)]);
}

enum : int {
  ST_NONE = 0,
  ST_GEMM = 1,
  ST_PANEL = 2,
  ST_DIAGINV = 3,
  ST_QR = 4,
  ST_COPY = 5,
  ST_CHOL = 6,
  ST_TSOLVE = 7,
  ST_LUUPD = 8,
  ST_LDSLU = 9,
  ST_LDSLU2 = 10
};

// One primitive step.  Field use per kind:
//   GEMM     C = beta C + alpha op(A) op(B), m x n x k; leaf selects the
//            precision mode of the LU routines (else the update mode)
//   PANEL    panel_lu(M = C, ld = ldc, p = m, pw = n, ny = k, linv = A)
//   DIAGINV  diag_inverses(F = A, ldf = lda, n = m, out = C, upper = aT)
//   QR       householder_qr_q(Y = C, ld = ldc, ny = m, k = n, Q = B, ldq = ldb,
//            tau = A, w = A + k)
//   COPY     C[i][j] = A[i][j] for i < m, j < n (pitches ldc, lda)
//   CHOL     chol_upper(G = C, k = m, rows = k, first = aT)
//   LUUPD    lu_update(M = C, ld = ldc, p = m, pw = n, nx = lda, ny = k, linv = A)
//   TSOLVE   narrow_solve(W = C, ldw = ldc, w_trans = bT, m, w = n, F = A,
//            ldf = lda, f_trans = aT, dinv = B, forward = k != 0)
//   LDSLU    getrf_lds(M = C, ld = ldc, n = m): a square leaf factored in LDS
//   LDSLU2   getrf_lds2(M = C, ld = ldc, n = m): the two-block form (FP32,
//            128 < n <= 256)
template <typename T>
struct Step {
  int kind;
  int m, n, k, ldc, lda, ldb;
  bool aT, bT, leaf;
  T alpha, beta;
  T* C;
  const T* A;
  const T* B;
};

template <typename T>
__device__ __forceinline__ void set_gemm(Step<T>& st, bool leaf, int m, int n, int k, T alpha, T beta, T* C, int ldc,
                                         const T* A, int lda, bool aT, const T* B, int ldb, bool bT) {
  st.kind = (m > 0 && n > 0) ? ST_GEMM : ST_NONE;
  st.leaf = leaf;
  st.m = m;
  st.n = n;
  st.k = k;
  st.alpha = alpha;
  st.beta = beta;
  st.C = C;
  st.ldc = ldc;
  st.A = A;
  st.lda = lda;
  st.aT = aT;
  st.B = B;
  st.ldb = ldb;
  st.bT = bT;
}

// A trsm with at most 64 right-hand-side columns whose right-hand side fits
// in LDS runs as DIAGINV + one TSOLVE step.
template <typename T>
__device__ __forceinline__ bool narrow_ok(int m, int w) {
  return w <= 64 && narrow_solve_bytes<T>(m, w) <= kLdsBytes - 1024;
}

template <typename T>
__device__ __forceinline__ void set_tsolve(Step<T>& st, T* W, int ldw, bool w_trans, int m, int w, const T* F, int ldf,
                                           bool f_trans, const T* dinv, bool forward) {
  st.kind = ST_TSOLVE;
  st.leaf = true;
  st.C = W;
  st.ldc = ldw;
  st.bT = w_trans;
  st.m = m;
  st.n = w;
  st.A = F;
  st.lda = ldf;
  st.aT = f_trans;
  st.B = dinv;
  st.k = forward ? 1 : 0;
  st.ldb = 0;  // narrow_solve's dl: inverses from dinv (exec_entry may set 1 or 2)
}

// Entry length in words, or -1 for an unknown operation.
__device__ __forceinline__ int entry_length(int op) {
  switch (op) {
    case nop: return nop_l;
    case getrf: return getrf_l;
    case trsml: return trsml_l;
    case trsmr: return trsmr_l;
    case trsmu: return trsmu_l;
    case trmm: return trmm_l;
    case gemm: return gemm_l;
    case gemm_plus: return gemm_plus_l;
    case gemm_3x: return gemm_3x_l;
    case gemm_4x: return gemm_4x_l;
    case accum: return accum_l;
    default: return -1;
  }
}

// Step s of the entry at pc.  Returns 1 (step in st), 0 (no more steps) or
// -1 (malformed entry).
template <typename T>
__device__ __forceinline__ int make_step(const int32_t* pc, void* const* P, int s, T* tmp, const T* rnd, Step<T>& st) {
  st.kind = ST_NONE;
  switch (word(pc, 2)) {
    case nop:
      return 0;
    case getrf: {
      // square leaves that fit in LDS: one step (getrf_lds); otherwise
      // right-looking blocked LU: per 32-column panel, PANEL then LUUPD
      // (U12 = L11^{-1} A12 and A22 -= L21 U12)
      T* M = slot<T>(P, pc, 3) + word(pc, 4);
      const int nx = word(pc, 5), ny = word(pc, 6), ld = word(pc, 7);
      if (nx == ny && nx <= (lds_lu_max<T>() > lds2_max<T>() ? lds_lu_max<T>() : lds2_max<T>())) {
        if (s > 0 || nx == 0) return 0;
        st.kind = nx <= lds_lu_max<T>() ? ST_LDSLU : ST_LDSLU2;
        st.leaf = true;
        st.C = M;
        st.ldc = ld;
        st.m = nx;
        return 1;
      }
      const int mn = nx < ny ? nx : ny, np = (mn + NB - 1) / NB;
      if (s >= 2 * np) return 0;
      const int q = s >> 1, r = s & 1, p = q * NB;
      const int pw = (mn - p) < NB ? (mn - p) : NB;
      st.kind = r == 0 ? ST_PANEL : ST_LUUPD;
      st.leaf = true;
      st.C = M;
      st.ldc = ld;
      st.m = p;
      st.n = pw;
      st.k = ny;
      st.lda = nx;
      st.A = tmp;
      return 1;
    }
    case trsml: {
      // B <- L^{-1} B (b_T: Bs <- Bs L^{-T}); DIAGINV, then per row block:
      // X_i = inv_i B_i and B_{>i} -= L_{>i,i} X_i
      T* B = slot<T>(P, pc, 3) + word(pc, 5);
      const T* L = slot<T>(P, pc, 4) + word(pc, 6);
      const int nx_b = word(pc, 7), n = word(pc, 8), ld_b = word(pc, 10), ld_l = word(pc, 11);
      const bool b_T = word(pc, 12) != 0;
      const int nblk = (n + NB - 1) / NB;
      if (s == 0) {
        st.kind = ST_DIAGINV;
        st.A = L;
        st.lda = ld_l;
        st.m = n;
        st.C = tmp;
        st.aT = false;
        return 1;
      }
      if (narrow_ok<T>(n, nx_b)) {
        if (s > 1) return 0;
        set_tsolve<T>(st, B, ld_b, b_T, n, nx_b, L, ld_l, false, tmp, true);
        return 1;
      }
      if (s > 2 * nblk) return 0;
      const int i = (s - 1) >> 1, r = (s - 1) & 1;
      const int r0 = i * NB, bs = (n - r0) < NB ? (n - r0) : NB, rest = n - r0 - bs;
      const T* Li = tmp + (int64_t)i * NB * NB;
      const T* Lr = L + (int64_t)(r0 + bs) * ld_l + r0;
      if (!b_T) {
        T* Bi = B + (int64_t)r0 * ld_b;
        if (r == 0) set_gemm<T>(st, true, bs, nx_b, bs, T(1), T(0), Bi, ld_b, Li, NB, false, Bi, ld_b, false);
        else set_gemm<T>(st, true, rest, nx_b, bs, T(-1), T(1), Bi + (int64_t)bs * ld_b, ld_b, Lr, ld_l, false, Bi, ld_b, false);
      } else {
        T* Bi = B + r0;
        if (r == 0) set_gemm<T>(st, true, nx_b, bs, bs, T(1), T(0), Bi, ld_b, Bi, ld_b, false, Li, NB, true);
        else set_gemm<T>(st, true, nx_b, rest, bs, T(-1), T(1), Bi + bs, ld_b, Bi, ld_b, false, Lr, ld_l, true);
      }
      return 1;
    }
    case trsmr: {
      // B <- B U^{-1} (b_T: Bs <- U^{-T} Bs); DIAGINV, then per column block:
      // X_j = B_j inv_j and B_{>j} -= X_j U_{j,>j}
      T* B = slot<T>(P, pc, 3) + word(pc, 5);
      const T* U = slot<T>(P, pc, 4) + word(pc, 6);
      const int n = word(pc, 7), ny_b = word(pc, 8), ld_b = word(pc, 10), ld_u = word(pc, 11);
      const bool b_T = word(pc, 12) != 0;
      const int nblk = (n + NB - 1) / NB;
      if (s == 0) {
        st.kind = ST_DIAGINV;
        st.A = U;
        st.lda = ld_u;
        st.m = n;
        st.C = tmp;
        st.aT = true;
        return 1;
      }
      if (narrow_ok<T>(n, ny_b)) {
        // W = B^T: a forward solve with U^T
        if (s > 1) return 0;
        set_tsolve<T>(st, B, ld_b, !b_T, n, ny_b, U, ld_u, true, tmp, true);
        return 1;
      }
      if (s > 2 * nblk) return 0;
      const int j = (s - 1) >> 1, r = (s - 1) & 1;
      const int c0 = j * NB, bs = (n - c0) < NB ? (n - c0) : NB, rest = n - c0 - bs;
      const T* Ui = tmp + (int64_t)j * NB * NB;
      const T* Ur = U + (int64_t)c0 * ld_u + c0 + bs;
      if (!b_T) {
        T* Bj = B + c0;
        if (r == 0) set_gemm<T>(st, true, ny_b, bs, bs, T(1), T(0), Bj, ld_b, Bj, ld_b, false, Ui, NB, false);
        else set_gemm<T>(st, true, ny_b, rest, bs, T(-1), T(1), Bj + bs, ld_b, Bj, ld_b, false, Ur, ld_u, false);
      } else {
        T* Bj = B + (int64_t)c0 * ld_b;
        if (r == 0) set_gemm<T>(st, true, bs, ny_b, bs, T(1), T(0), Bj, ld_b, Ui, NB, true, Bj, ld_b, false);
        else set_gemm<T>(st, true, rest, ny_b, bs, T(-1), T(1), Bj + (int64_t)bs * ld_b, ld_b, Ur, ld_u, true, Bj, ld_b, false);
      }
      return 1;
    }
    case trsmu: {
      // B <- U^{-1} B (backward); DIAGINV, then from the last row block:
      // X_i = inv_i B_i and B_{<i} -= U_{<i,i} X_i
      T* B = slot<T>(P, pc, 3) + word(pc, 5);
      const T* U = slot<T>(P, pc, 4) + word(pc, 6);
      const int nx_b = word(pc, 7), n = word(pc, 8), ld_b = word(pc, 10), ld_u = word(pc, 11);
      const int nblk = (n + NB - 1) / NB;
      if (s == 0) {
        st.kind = ST_DIAGINV;
        st.A = U;
        st.lda = ld_u;
        st.m = n;
        st.C = tmp;
        st.aT = true;
        return 1;
      }
      if (narrow_ok<T>(n, nx_b)) {
        if (s > 1) return 0;
        set_tsolve<T>(st, B, ld_b, false, n, nx_b, U, ld_u, false, tmp, false);
        return 1;
      }
      if (s > 2 * nblk) return 0;
      const int i = nblk - 1 - ((s - 1) >> 1), r = (s - 1) & 1;
      const int r0 = i * NB, bs = (n - r0) < NB ? (n - r0) : NB;
      T* Bi = B + (int64_t)r0 * ld_b;
      if (r == 0) set_gemm<T>(st, true, bs, nx_b, bs, T(1), T(0), Bi, ld_b, tmp + (int64_t)i * NB * NB, NB, false, Bi, ld_b, false);
      else set_gemm<T>(st, true, r0, nx_b, bs, T(-1), T(1), B, ld_b, U + r0, ld_u, false, Bi, ld_b, false);
      return 1;
    }
    case trmm: {
      // B <- T B: copy B to the scratch, then B = T * copy
      T* B = slot<T>(P, pc, 3) + word(pc, 5);
      const T* Tm = slot<T>(P, pc, 4) + word(pc, 6);
      const int nx_b = word(pc, 7), ny_b = word(pc, 8), ld_b = word(pc, 9), ld_t = word(pc, 10);
      if (s == 0) {
        st.kind = ST_COPY;
        st.C = tmp;
        st.ldc = nx_b;
        st.A = B;
        st.lda = ld_b;
        st.m = ny_b;
        st.n = nx_b;
        return 1;
      }
      if (s > 1) return 0;
      set_gemm<T>(st, true, ny_b, nx_b, ny_b, T(1), T(0), B, ld_b, Tm, ld_t, false, tmp, nx_b, false);
      return 1;
    }
    case gemm:
    case gemm_plus: {
      if (s > 0) return 0;
      T* M = slot<T>(P, pc, 3) + word(pc, 6);
      const T* A = slot<T>(P, pc, 4) + word(pc, 7);
      const T* B = slot<T>(P, pc, 5) + word(pc, 8);
      const T alpha = word(pc, 2) == gemm ? T(-1) : T(1);
      set_gemm<T>(st, false, word(pc, 9), word(pc, 10), word(pc, 11), alpha, T(1), M, word(pc, 12), A, word(pc, 13),
                  word(pc, 15) != 0, B, word(pc, 14), word(pc, 16) != 0);
      return 1;
    }
    case gemm_3x: {
      // association chosen by the host (src/kernel.cu:887-911)
      if (s > 1) return 0;
      T* M = slot<T>(P, pc, 3) + word(pc, 7);
      const T* A = slot<T>(P, pc, 4) + word(pc, 8);
      const T* B = slot<T>(P, pc, 5) + word(pc, 9);
      const T* C = slot<T>(P, pc, 6) + word(pc, 10);
      const int m = word(pc, 11), n = word(pc, 12), k = word(pc, 13), l = word(pc, 14);
      const int ldm = word(pc, 15), lda = word(pc, 16), ldb = word(pc, 17), ldc = word(pc, 18);
      const bool aT = word(pc, 19) != 0, bT = word(pc, 20) != 0, cT = word(pc, 21) != 0;
      if (word(pc, 22)) {  // (A B) C
        if (s == 0) set_gemm<T>(st, false, m, l, k, T(1), T(0), tmp, l, A, lda, aT, B, ldb, bT);
        else set_gemm<T>(st, false, m, n, l, T(-1), T(1), M, ldm, tmp, l, false, C, ldc, cT);
      } else {  // A (B C)
        if (s == 0) set_gemm<T>(st, false, k, n, l, T(1), T(0), tmp, n, B, ldb, bT, C, ldc, cT);
        else set_gemm<T>(st, false, m, n, k, T(-1), T(1), M, ldm, A, lda, aT, tmp, n, false);
      }
      return 1;
    }
    case gemm_4x: {
      // association chosen by the host (src/kernel.cu:914-953)
      const int control = word(pc, 27);
      if (control < 0 || control > 4) return -1;
      if (s > 2) return 0;
      T* M = slot<T>(P, pc, 3) + word(pc, 8);
      const T* A = slot<T>(P, pc, 4) + word(pc, 9);
      const T* B = slot<T>(P, pc, 5) + word(pc, 10);
      const T* C = slot<T>(P, pc, 6) + word(pc, 11);
      const T* D = slot<T>(P, pc, 7) + word(pc, 12);
      const int m = word(pc, 13), n = word(pc, 14), k = word(pc, 15), l = word(pc, 16), o = word(pc, 17);
      const int ldm = word(pc, 18), lda = word(pc, 19), ldb = word(pc, 20), ldc = word(pc, 21), ldd = word(pc, 22);
      const bool aT = word(pc, 23) != 0, bT = word(pc, 24) != 0, cT = word(pc, 25) != 0, dT = word(pc, 26) != 0;
      T* t1 = tmp;
      T* t2 = tmp + word(pc, 28);
      switch (control * 3 + s) {
        // ((A B) C) D
        case 0: set_gemm<T>(st, false, m, l, k, T(1), T(0), t1, l, A, lda, aT, B, ldb, bT); break;
        case 1: set_gemm<T>(st, false, m, o, l, T(1), T(0), t2, o, t1, l, false, C, ldc, cT); break;
        case 2: set_gemm<T>(st, false, m, n, o, T(-1), T(1), M, ldm, t2, o, false, D, ldd, dT); break;
        // (A (B C)) D
        case 3: set_gemm<T>(st, false, k, o, l, T(1), T(0), t1, o, B, ldb, bT, C, ldc, cT); break;
        case 4: set_gemm<T>(st, false, m, o, k, T(1), T(0), t2, o, A, lda, aT, t1, o, false); break;
        case 5: set_gemm<T>(st, false, m, n, o, T(-1), T(1), M, ldm, t2, o, false, D, ldd, dT); break;
        // A ((B C) D)
        case 6: set_gemm<T>(st, false, k, o, l, T(1), T(0), t1, o, B, ldb, bT, C, ldc, cT); break;
        case 7: set_gemm<T>(st, false, k, n, o, T(1), T(0), t2, n, t1, o, false, D, ldd, dT); break;
        case 8: set_gemm<T>(st, false, m, n, k, T(-1), T(1), M, ldm, A, lda, aT, t2, n, false); break;
        // A (B (C D))
        case 9: set_gemm<T>(st, false, l, n, o, T(1), T(0), t1, n, C, ldc, cT, D, ldd, dT); break;
        case 10: set_gemm<T>(st, false, k, n, l, T(1), T(0), t2, n, B, ldb, bT, t1, n, false); break;
        case 11: set_gemm<T>(st, false, m, n, k, T(-1), T(1), M, ldm, A, lda, aT, t2, n, false); break;
        // (A B) (C D)
        case 12: set_gemm<T>(st, false, m, l, k, T(1), T(0), t1, l, A, lda, aT, B, ldb, bT); break;
        case 13: set_gemm<T>(st, false, l, n, o, T(1), T(0), t2, n, C, ldc, cT, D, ldd, dT); break;
        default: set_gemm<T>(st, false, m, n, l, T(-1), T(1), M, ldm, t1, l, false, t2, n, false); break;
      }
      return 1;
    }
    case accum: {
      // U1 (ny x k1), V1 (nx x k1), U2 (ny x k2), V2 (nx x k2).
      //   Y  = U1 (Omega^T V1)^T + U2 (Omega^T V2)^T        (sketch, 4 GEMMs)
      //   Q  = orth(Y) by shifted CholeskyQR3: three passes of
      //        G = Y^T Y, R = chol(G [+ shift]), Y <- Y R^{-1}
      //        (R^{-1} applied as diagonal-block inverses + block updates);
      //        range(Q) = range(Y), so Q Q^T (U1 V1^T + U2 V2^T) is the
      //        reference's projection.  If a factorisation fails even with
      //        the shift, Y is recomputed and orthonormalised by Householder.
      //   Vn = V1 (Q^T U1)^T + V2 (Q^T U2)^T; U1 <- Q, V1 <- Vn.
      // Scratch: Y [0, off1), Vn [off1, off2), Q [off2, ...), then S, G, D
      // (inverses of G's diagonal blocks), tau / w.
      T* U1 = slot<T>(P, pc, 3) + word(pc, 7);
      T* V1 = slot<T>(P, pc, 4) + word(pc, 8);
      const T* U2 = slot<T>(P, pc, 5) + word(pc, 9);
      const T* V2 = slot<T>(P, pc, 6) + word(pc, 10);
      const int nx = word(pc, 11), ny = word(pc, 12), k1 = word(pc, 13), k2 = word(pc, 14);
      const int ld_u1 = word(pc, 15), ld_v1 = word(pc, 16), ld_u2 = word(pc, 17), ld_v2 = word(pc, 18);
      if (k1 > 64) return -1;
      const int kmax = k1 > k2 ? k1 : k2;
      auto a32 = [](int64_t x) { return (x + 31) / 32 * 32; };
      T* Y = tmp;
      T* Vn = tmp + word(pc, 19);
      T* Q = tmp + word(pc, 20);
      T* S = Q + a32((int64_t)ny * k1);
      T* G = S + a32((int64_t)k1 * kmax);
      const int nblk = (k1 + NB - 1) / NB;
      T* D = G + a32((int64_t)k1 * k1);
      T* tau = D + (int64_t)nblk * NB * NB;
      constexpr int kPasses = 3;
      const int per_pass = 3 + 2 * nblk;
      const int f0 = 4, b0 = f0 + kPasses * per_pass, p0 = b0 + 6;
      if (s >= p0 + 6) return 0;
      auto sketch = [&](int q) {
        switch (q) {
          case 0: set_gemm<T>(st, false, k1, k1, nx, T(1), T(0), S, k1, rnd, k1, true, V1, ld_v1, false); break;
          case 1: set_gemm<T>(st, false, ny, k1, k1, T(1), T(0), Y, k1, U1, ld_u1, false, S, k1, true); break;
          case 2: set_gemm<T>(st, false, k1, k2, nx, T(1), T(0), S, k2, rnd, k1, true, V2, ld_v2, false); break;
          default: set_gemm<T>(st, false, ny, k1, k2, T(1), T(1), Y, k1, U2, ld_u2, false, S, k2, true); break;
        }
      };
      auto copy = [&](T* dst, int ldd, const T* src, int ld_src, int rows) {
        st.kind = ST_COPY;
        st.C = dst;
        st.ldc = ldd;
        st.A = src;
        st.lda = ld_src;
        st.m = rows;
        st.n = k1;
      };
      if (s < f0) {
        sketch(s);
        return 1;
      }
      if (s < b0) {
        const int p = (s - f0) / per_pass, r = (s - f0) % per_pass;
        // the flag is valid once the first factorisation of this entry ran
        if ((p > 0 || r > 1) && *chol_fail_flag() != 0) return 1;
        if (r == 0) {  // G = Y^T Y
          set_gemm<T>(st, true, k1, k1, ny, T(1), T(0), G, k1, Y, k1, true, Y, k1, false);
        } else if (r == 1) {
          st.kind = ST_CHOL;
          st.C = G;
          st.m = k1;
          st.k = ny;
          st.aT = p == 0;
        } else if (r == 2) {
          st.kind = ST_DIAGINV;
          st.A = G;
          st.lda = k1;
          st.m = k1;
          st.C = D;
          st.aT = true;
        } else {  // Y <- Y R^{-1}, column block j (as trsmr)
          const int j = (r - 3) >> 1, c0 = j * NB;
          const int bs = (k1 - c0) < NB ? (k1 - c0) : NB, rest = k1 - c0 - bs;
          if (((r - 3) & 1) == 0)
            set_gemm<T>(st, true, ny, bs, bs, T(1), T(0), Y + c0, k1, Y + c0, k1, false, D + (int64_t)j * NB * NB, NB,
                        false);
          else
            set_gemm<T>(st, true, ny, rest, bs, T(-1), T(1), Y + c0 + bs, k1, Y + c0, k1, false,
                        G + (int64_t)c0 * k1 + c0 + bs, k1, false);
        }
        return 1;
      }
      if (s < p0) {  // Householder fallback
        if (*chol_fail_flag() == 0) return 1;
        const int q = s - b0;
        if (q < 4) {
          sketch(q);
        } else if (q == 4) {
          st.kind = ST_QR;
          st.C = Y;
          st.ldc = k1;
          st.m = ny;
          st.n = k1;
          st.B = Q;
          st.ldb = k1;
          st.A = tau;
          st.k = kmax;
        } else {
          copy(Y, k1, Q, k1, ny);
        }
        return 1;
      }
      switch (s - p0) {  // projection with Q held in Y
        case 0: set_gemm<T>(st, false, k1, k1, ny, T(1), T(0), S, k1, Y, k1, true, U1, ld_u1, false); break;
        case 1: set_gemm<T>(st, false, nx, k1, k1, T(1), T(0), Vn, k1, V1, ld_v1, false, S, k1, true); break;
        case 2: set_gemm<T>(st, false, k1, k2, ny, T(1), T(0), S, k2, Y, k1, true, U2, ld_u2, false); break;
        case 3: set_gemm<T>(st, false, nx, k1, k2, T(1), T(1), Vn, k1, V2, ld_v2, false, S, k2, true); break;
        case 4: copy(U1, ld_u1, Y, k1, ny); break;
        default: copy(V1, ld_v1, Vn, k1, nx); break;
      }
      return 1;
    }
    default:
      return -1;
  }
}

// dst and src never overlap (a copy step moves a block between an operand
// and a worker's scratch), so each thread requests eight elements before it
// stores them: element by element, every load would wait for the store
// before it (they may alias as far as the compiler knows).
template <typename T>
__device__ __forceinline__ void copy_block(T* dst, int ldd, const T* src, int lds_, int m, int n) {
  const int total = m * n, nt = blockDim.x;
  for (int e0 = threadIdx.x; e0 < total; e0 += 8 * nt) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * nt < total ? e0 + u * nt : e0, i = e / n, c = e - i * n;
      v[u] = gload<T>(src + (int64_t)i * lds_ + c);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * nt, i = e / n, c = e - i * n;
      if (e < total) gstore<T>(dst + (int64_t)i * ldd + c, v[u]);
    }
  }
  __syncthreads();
}

template <typename T, int FMODE, int GMODE>
__device__ __forceinline__ void exec_step(const Step<T>& st) {
  switch (st.kind) {
    case ST_GEMM:
      if constexpr (FMODE == GMODE) {
        gemm_block<T, FMODE>(st.m, st.n, st.k, st.alpha, st.beta, st.C, st.ldc, st.A, st.lda, st.aT, st.B, st.ldb, st.bT);
      } else {
        if (st.leaf)
          gemm_block<T, FMODE>(st.m, st.n, st.k, st.alpha, st.beta, st.C, st.ldc, st.A, st.lda, st.aT, st.B, st.ldb,
                               st.bT);
        else
          gemm_block<T, GMODE>(st.m, st.n, st.k, st.alpha, st.beta, st.C, st.ldc, st.A, st.lda, st.aT, st.B, st.ldb,
                               st.bT);
      }
      break;
    case ST_PANEL:
      panel_lu<T>(st.C, st.ldc, st.m, st.n, st.k, const_cast<T*>(st.A));
      break;
    case ST_DIAGINV:
      if (st.aT) diag_inverses<T, true>(st.A, st.lda, st.m, st.C);
      else diag_inverses<T, false>(st.A, st.lda, st.m, st.C);
      break;
    case ST_QR:
      householder_qr_q<T>(st.C, st.ldc, st.m, st.n, const_cast<T*>(st.B), st.ldb, const_cast<T*>(st.A),
                          const_cast<T*>(st.A) + st.k, (PP_LDS T*)(lds_base() + 40 * 1024));
      break;
    case ST_COPY:
      copy_block<T>(st.C, st.ldc, st.A, st.lda, st.m, st.n);
      break;
    case ST_CHOL:
      chol_upper<T>(st.C, st.m, st.k, st.aT, GMODE == MODE_BF16);
      break;
    case ST_LUUPD:
      lu_update<T, FMODE>(st.C, st.ldc, st.m, st.n, st.lda, st.k, st.A);
      break;
    case ST_TSOLVE:
      narrow_solve<T, FMODE>(st.C, st.ldc, st.bT, st.m, st.n, st.A, st.lda, st.aT, st.B, st.k != 0, st.ldb);
      break;
    case ST_LDSLU:
      getrf_lds<T, FMODE>(st.C, st.ldc, st.m);
      break;
    case ST_LDSLU2:
      if constexpr (lds2_max<T>() > 0) getrf_lds2<T, FMODE>(st.C, st.ldc, st.m);
      break;
    default:
      break;
  }
}

// Executes one execute entry; returns its length in words or -1.
// fallbacks (optional): counts the Householder fallbacks of accum.
// ---------------------------------------------------------------- fused gemm_4x
// M -= (A (op(B) op(C))) op(D), the (A (B C)) D association of gemm_4x, with
// k = o = NB (l a multiple of the MFMA chunk), A row-major and op(D) = D^T
// (D n x o row-major): t1 = op(B) op(C) (NB x NB, inner dimension l) is
// formed in LDS with its chunks split over the waves and the partial sums
// added in wave order (as the step form's split-K GEMM does), then t2 = A t1
// (m x NB) tile by tile straight into the staged operand rows of
// gemm_staged_run while the copies of D are in flight, and M -= t2 op(D)
// runs as the staged update.
template <typename T, int MODE>
__device__ __forceinline__ bool gemm4x_fused_ok(int m, int n, int k, int l, int o, bool aT, const T* D, int ldd, bool dT,
                                                const T* M, int ldm) {
  const int nw = blockDim.x >> 6;
  return k == NB && o == NB && l >= Mma<T, MODE>::KW && l % Mma<T, MODE>::KW == 0 && !aT && dT &&
         gemm_staged_ok<T, MODE>(m, n, o, D, o, true, D, ldd, true, M, ldm, nw) &&
         ((int64_t)(m + n) * o + NB * (NB + 1)) * (int64_t)sizeof(T) <= kStagedBytes &&
         m * NB >= 4 * 64 * 16;  // t1's partial sums fit in the t2 rows
}

// Not inlined: inlined, its register demand made the VM kernels spill more
// on every path.
template <typename T, int MODE>
__device__ __attribute__((noinline)) void gemm4x_fused(int m, int n, int l, T* M, int ldm, const T* A, int lda,
                                                       const T* B, int ldb, bool bT, const T* Cm, int ldc, bool cT,
                                                       const T* D, int ldd) {
  using P = Mma<T, MODE>;
  using acc_t = typename P::acc_t;
  constexpr int KW = P::KW, EPL = P::EPL, CH = NB / KW, V = 16 / (int)sizeof(T), TP = NB + 1, CPR = NB / V;
  const int nw = blockDim.x >> 6, w = wave_id(), l0 = lane_id(), g = l0 >> 4, li = l0 & 15;
  auto fill = [&](PP_LDS T* sa) {
    PP_LDS T* t1 = sa + (m + n) * NB;  // after the staged rows, pitch TP
    PP_LDS T* red = sa;                // partial sums, in the t2 rows before t2 is formed
    // t1: chunks w, w + gk, ... of the inner dimension on wave w < gk
    const int nch = l / KW;
    int gk = 1;
    while (2 * gk <= nw && 2 * gk <= nch) gk *= 2;
    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i) acc[i][0] = acc[i][1] = P::zero();
    if (w < gk) {
      OperandLane<T, EPL> la[2], lb[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        la[h].init(B, ldb, !bT, 16 * h + li, NB);
        lb[h].init(Cm, ldc, cT, 16 * h + li, NB);
      }
      for (int c = w; c < nch; c += gk) {
        T fa[2][EPL], fb[2][EPL];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          la[h].template load<false>(fa[h], c * KW + EPL * g, l);
          lb[h].template load<false>(fb[h], c * KW + EPL * g, l);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = P::mma_chunk(fa[i], fb[j], acc[i][j]);
      }
      if (w > 0) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) red[(w * 64 + l0) * 16 + (i * 2 + j) * 4 + r] = acc[i][j][r];
      }
    }
    __syncthreads();
    if (w == 0) {
      for (int q = 1; q < gk; ++q) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] += red[(q * 64 + l0) * 16 + (i * 2 + j) * 4 + r];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) t1[(16 * i + P::orow(l0, r)) * TP + 16 * j + li] = acc[i][j][r];
    }
    __syncthreads();
    // t2 = A t1, tiles of 32 rows dealt to the waves, into the staged rows
    for (int t = w; t < (m >> 5); t += nw) {
      const int r0 = 32 * t;
      OperandLane<T, EPL> la[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) la[h].init(A, lda, true, r0 + 16 * h + li, m);
      acc_t a2[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a2[i][0] = a2[i][1] = P::zero();
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        T fa[2][EPL], fb[2][EPL];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          la[h].template load<false>(fa[h], c * KW + EPL * g, NB);
#pragma unroll
          for (int e = 0; e < EPL; ++e) fb[h][e] = t1[(c * KW + EPL * g + e) * TP + 16 * h + li];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) a2[i][j] = P::mma_chunk(fa[i], fb[j], a2[i][j]);
      }
      // t2's row r, column c into the staged layout (pitch NB, chunk c / V
      // at c / V ^ (r mod CPR))
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = r0 + 16 * i + P::orow(l0, r), col = 16 * j + li;
            sa[row * NB + V * ((col / V) ^ (row & (CPR - 1))) + col % V] = a2[i][j][r];
          }
    }
  };
  gemm_staged_run<T, MODE, true, false>(m, n, NB, T(-1), T(1), M, ldm, static_cast<const T*>(nullptr), NB, D, ldd,
                                        fill);
}

template <typename T, int FMODE, int GMODE>
__device__ __forceinline__ int exec_entry(const int32_t* pc, void* const* ptrs, const void* rnd, T* tmp,
                                          int32_t* fallbacks = nullptr) {
  const int len = entry_length(word(pc, 2));
  if (len < 0) return -1;
  if (word(pc, 2) == accum) {
    // the whole accumulation in one routine: the sketch in LDS (small
    // blocks) or streamed through LDS from the worker scratch; only when its
    // factorisations fail does the step sequence below run, from the start
    // (its Cholesky passes fail the same way, then Householder)
    const int nx = word(pc, 11), ny = word(pc, 12), k1 = word(pc, 13), k2 = word(pc, 14);
    T* U1 = slot<T>(ptrs, pc, 3) + word(pc, 7);
    T* V1 = slot<T>(ptrs, pc, 4) + word(pc, 8);
    const T* U2 = slot<T>(ptrs, pc, 5) + word(pc, 9);
    const T* V2 = slot<T>(ptrs, pc, 6) + word(pc, 10);
    const int l1 = word(pc, 15), l2 = word(pc, 16), l3 = word(pc, 17), l4 = word(pc, 18);
    const T* om = reinterpret_cast<const T*>(rnd);
    if (accum_lds_ok<T>(nx, ny, k1, k2)) {
      accum_lds<T, FMODE, GMODE>(U1, l1, V1, l2, U2, l3, V2, l4, nx, ny, k1, k2, om, GMODE == MODE_BF16);
      if (*chol_fail_flag() == 0) return len;
    } else if (accum_big_ok<T>(nx, ny, k1, k2)) {
      accum_big<T, FMODE, GMODE>(U1, l1, V1, l2, U2, l3, V2, l4, nx, ny, k1, k2, om, tmp, GMODE == MODE_BF16);
      if (*chol_fail_flag() == 0) return len;
    }
  }
  if (word(pc, 2) == gemm_4x && word(pc, 27) == 1) {
    // the (A (B C)) D association with rank-32 outer factors: the two small
    // products in LDS, feeding the staged update
    const int m = word(pc, 13), n = word(pc, 14), k = word(pc, 15), l = word(pc, 16), o = word(pc, 17);
    T* M = slot<T>(ptrs, pc, 3) + word(pc, 8);
    const T* A = slot<T>(ptrs, pc, 4) + word(pc, 9);
    const T* B = slot<T>(ptrs, pc, 5) + word(pc, 10);
    const T* C = slot<T>(ptrs, pc, 6) + word(pc, 11);
    const T* D = slot<T>(ptrs, pc, 7) + word(pc, 12);
    const int ldm = word(pc, 18), lda = word(pc, 19), ldb = word(pc, 20), ldc = word(pc, 21), ldd = word(pc, 22);
    const bool aT = word(pc, 23) != 0, bT = word(pc, 24) != 0, cT = word(pc, 25) != 0, dT = word(pc, 26) != 0;
    if (gemm4x_fused_ok<T, GMODE>(m, n, k, l, o, aT, D, ldd, dT, M, ldm)) {
      gemm4x_fused<T, GMODE>(m, n, l, M, ldm, A, lda, B, ldb, bT, C, ldc, cT, D, ldd);
      return len;
    }
  }
  // a narrow triangular solve whose diagonal-block inverses fit in LDS next
  // to its right-hand side forms them itself (dl): the separate step and the
  // round trip through the worker scratch are skipped
  int dl = 0;
  {
    const int opw = word(pc, 2);
    if (opw == trsml || opw == trsmr || opw == trsmu) {
      const int n = opw == trsmr ? word(pc, 7) : word(pc, 8), wd = opw == trsmr ? word(pc, 8) : word(pc, 7);
      if (narrow_ok<T>(n, wd) && narrow_dl_ok<T>(n, wd)) dl = opw == trsml ? 1 : 2;
    }
  }
  for (int s = 0;; ++s) {
    Step<T> st;
    const int r = make_step<T>(pc, ptrs, s, tmp, reinterpret_cast<const T*>(rnd), st);
    if (r < 0) return -1;
    if (r == 0) break;
    if (st.kind == ST_QR && fallbacks != nullptr && threadIdx.x == 0) atomicAdd(fallbacks, 1);
    if (dl != 0) {
      if (st.kind == ST_DIAGINV) continue;
      if (st.kind == ST_TSOLVE) st.ldb = dl;
    }
    exec_step<T, FMODE, GMODE>(st);
  }
  return len;
}

}  // namespace pp
