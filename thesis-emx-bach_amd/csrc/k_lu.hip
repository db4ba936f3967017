// k_lu.hip — the global 292x292 FP64 solve, bit-identical to the reference's
// Numerical Recipes ludcmp + lubksb (src/bachUtil.cpp:319-411).
//
// Right-looking blocked elimination with the reference's per-element order:
// element (i,l) receives -L_ik U_kl for k ascending, exactly as Crout forms it,
// and rows are swapped by the same pivot sequence (implicit scaling vv, `>=`
// rule: the last row attaining the maximum wins; NaN never wins; no candidate
// keeps the previous row).  Applying a panel's row swaps to a trailing column
// before that panel's updates is equivalent to interleaving them, because the
// swaps only exchange rows that the panel's updates treat identically.
//
//   k_lu_begin   vv, zero-row test (ludcmp returns 1 before factorising), and
//                the first 16-column panel factorised in LDS;
//   k_lu_step(p) one workgroup per 16 trailing columns: apply the panel's
//                swaps, the in-panel triangular update, the rank-16 update of
//                the rows below (each element summed in k order); workgroup 0
//                owns the next panel's columns and factorises that panel
//                right after updating it (one launch per panel).  A last
//                workgroup carries the right-hand side as one more column:
//                lubksb's forward substitution, including its `ii` rule
//                (updates start at the first nonzero entry), fused here;
//   k_lu_backsub back substitution; its ascending-j sums are one sequential
//                chain per row in the reference order.
#include <algorithm>

#include "xb_dev.h"

namespace xb {

constexpr int LB = 16;  // panel width = trailing columns per workgroup
constexpr int NPANEL = (MATSIZE + LB - 1) / LB;
constexpr int LU_THREADS = 256;
constexpr int STEP_THREADS = 64;  // one wave: the panel's barriers stay wave-local

struct LuState {
  int zero;      // a zero row: ludcmp returned 1 without factorising
  int lastMaxI;  // pivot row of the previous column (kept when none qualifies)
  int ii;        // lubksb's first nonzero index
  int fallback;  // a column without a finite pivot kept an earlier row: replay
  int skip;      // empty stamp list: the reference never solves
  int pad[3];
};

// Factorise the panel held in T (rows r0..n of the matrix at T rows
// off..off+R-1, columns 0..cB-1).  c0 = matrix column of T column 0.
__device__ void panel_factor(double (*T)[LB + 1], int off, int c0, int cB, double* vvL,
                             int* piv, LuState* st) {
  const int n = MATSIZE;
  const int tid = threadIdx.x;
  const int lane = lane_id(), wid = tid / WAVE, nw = blockDim.x / WAVE;
  __shared__ double wv[16];
  __shared__ int wi[16];
  __shared__ int s_maxI;
  if (tid == 0) s_maxI = st->lastMaxI;
  __syncthreads();
  for (int q = 0; q < cB; q++) {
    const int k = c0 + q;
    double bv = -1.0;
    int bi = -1;
    for (int i = k + tid; i <= n; i += blockDim.x) {
      double dum = vvL[i - c0] * fabs(T[off + i - c0][q]);
      if (dum >= 0.0 && dum >= bv) { bv = dum; bi = i; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      double ov = __shfl_xor(bv, o);
      int oi = __shfl_xor(bi, o);
      if (ov > bv || (ov == bv && oi > bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) { wv[wid] = bv; wi[wid] = bi; }
    __syncthreads();
    if (tid == 0) {
      double b2 = wv[0];
      int i2 = wi[0];
      for (int w = 1; w < nw; w++)
        if (wv[w] > b2 || (wv[w] == b2 && wi[w] > i2)) { b2 = wv[w]; i2 = wi[w]; }
      if (i2 >= 0) s_maxI = i2;
    }
    __syncthreads();
    const int maxI = s_maxI;
    if (maxI < k) {
      // No row qualified and the kept pivot row lies above this column: the
      // reference swaps with an already-factorised row, which only a literal
      // replay reproduces (k_lu_fallback).
      if (tid == 0) st->fallback = 1;
      __syncthreads();
      return;
    }
    if (maxI != k) {
      if (tid < cB) {
        double t = T[off + maxI - c0][tid];
        T[off + maxI - c0][tid] = T[off + k - c0][tid];
        T[off + k - c0][tid] = t;
      }
      if (tid == 0) vvL[maxI - c0] = vvL[k - c0];
    }
    __syncthreads();
    if (tid == 0) {
      piv[k] = maxI;
      if (T[off + k - c0][q] == 0.0) T[off + k - c0][q] = 1.0e-20;
    }
    __syncthreads();
    if (k != n) {
      const double dum = 1.0 / T[off + k - c0][q];
      for (int i = k + 1 + tid; i <= n; i += blockDim.x) T[off + i - c0][q] *= dum;
      __syncthreads();
      const int m = cB - q - 1;  // remaining panel columns
      if (m > 0) {
        const int rows = n - k;
        for (int e = tid; e < rows * m; e += blockDim.x) {
          int i = k + 1 + e / m, l = q + 1 + e % m;
          T[off + i - c0][l] = T[off + i - c0][l] - T[off + i - c0][q] * T[off + k - c0][l];
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) st->lastMaxI = s_maxI;
  __syncthreads();
}

__global__ __launch_bounds__(LU_THREADS) void k_lu_begin(double* M, double* vv, int* piv,
                                                         LuState* st, const int* skipIfZero) {
  const int n = MATSIZE, N = MATN;
  const int tid = threadIdx.x;
  __shared__ int s_zero;
  __shared__ double T[MATSIZE][LB + 1];
  __shared__ double vvL[MATSIZE];
  if (tid == 0) s_zero = 0;
  for (int i = tid; i < N; i += blockDim.x) piv[i] = 0;
  const int skip = (skipIfZero && *skipIfZero == 0) ? 1 : 0;
  if (tid == 0) {
    st->skip = skip;
    st->fallback = 0;
    st->zero = 0;
  }
  if (skip) return;  // testFit returns 666 before solving (cdkscUtil.cpp:101-103)
  __syncthreads();
  for (int i = 1 + tid; i <= n; i += blockDim.x) {
    double big = 0.0;
    for (int j = 1; j <= n; j++) {
      double t = fabs(M[(size_t)i * N + j]);
      if (t > big) big = t;
    }
    if (big == 0.0) atomicOr(&s_zero, 1);
    double v = 1.0 / big;
    vv[i] = v;
    vvL[i - 1] = v;
  }
  __syncthreads();
  if (tid == 0) {
    st->zero = s_zero;
    st->lastMaxI = 0;
    st->ii = 0;
  }
  if (s_zero) return;
  const int cB = min(LB, n);
  for (int e = tid; e < n * cB; e += blockDim.x) {
    int r = e / cB, q = e % cB;
    T[r][q] = M[(size_t)(1 + r) * N + 1 + q];
  }
  __syncthreads();
  panel_factor(T, 0, 1, cB, vvL, piv, st);
  for (int e = tid; e < n * cB; e += blockDim.x) {
    int r = e / cB, q = e % cB;
    M[(size_t)(1 + r) * N + 1 + q] = T[r][q];
  }
  for (int i = tid; i < n; i += blockDim.x) vv[1 + i] = vvL[i];
}

__global__ __launch_bounds__(STEP_THREADS) void k_lu_step(double* M, double* vv, int* piv,
                                                        double* b, LuState* st, int p,
                                                        int nColBlocks) {
  if (st->zero || st->skip || st->fallback) return;
  const int n = MATSIZE, N = MATN;
  const int tid = threadIdx.x;
  const int c0 = 1 + p * LB;
  const int cB = min(LB, n - c0 + 1);
  const int R = n - c0 + 1;  // rows c0..n
  __shared__ double T[MATSIZE][LB + 1];
  __shared__ double vvL[MATSIZE];
  __shared__ int s_ii;
  if ((int)blockIdx.x == nColBlocks) {
    // right-hand side column: permutation + forward substitution (lubksb)
    double* y = &T[0][0];  // R doubles
    for (int r = tid; r < R; r += blockDim.x) y[r] = b[c0 + r];
    if (tid == 0) s_ii = st->ii;
    __syncthreads();
    if (tid == 0)
      for (int q = 0; q < cB; q++) {
        int k = c0 + q, m = piv[k];
        if (m != k) {
          double t = y[m - c0];
          y[m - c0] = y[k - c0];
          y[k - c0] = t;
        }
      }
    __syncthreads();
    for (int q = 0; q < cB; q++) {
      const int k = c0 + q;
      if (tid == 0 && s_ii == 0 && !(y[k - c0] == 0.0)) s_ii = k;
      __syncthreads();
      if (s_ii != 0) {
        const double yk = y[k - c0];
        for (int r = k + 1 + tid; r <= n; r += blockDim.x)
          y[r - c0] = y[r - c0] - M[(size_t)r * N + k] * yk;
      }
      __syncthreads();
    }
    for (int r = tid; r < R; r += blockDim.x) b[c0 + r] = y[r];
    if (tid == 0) st->ii = s_ii;
    return;
  }
  // trailing column block
  const int l0 = c0 + cB + blockIdx.x * LB;
  const int nc = min(LB, n - l0 + 1);
  if (nc <= 0) return;
  for (int e = tid; e < R * nc; e += blockDim.x) {
    int r = e / nc, t = e % nc;
    T[r][t] = M[(size_t)(c0 + r) * N + l0 + t];
  }
  __syncthreads();
  if (tid < nc)
    for (int q = 0; q < cB; q++) {
      int k = c0 + q, m = piv[k];
      if (m != k) {
        double t = T[m - c0][tid];
        T[m - c0][tid] = T[k - c0][tid];
        T[k - c0][tid] = t;
      }
    }
  __syncthreads();
  // in-panel rows: U12 = L11^-1 A12 with k ascending per element
  for (int q = 0; q + 1 < cB; q++) {
    const int k = c0 + q;
    const int rows = cB - q - 1;
    for (int e = tid; e < rows * nc; e += blockDim.x) {
      int r = k + 1 + e / nc, t = e % nc;
      T[r - c0][t] = T[r - c0][t] - M[(size_t)r * N + k] * T[q][t];
    }
    __syncthreads();
  }
  // rows below the panel: rank-cB update, k ascending per element
  for (int e = tid; e < (R - cB) * nc; e += blockDim.x) {
    int r = c0 + cB + e / nc, t = e % nc;
    const double* Lr = M + (size_t)r * N + c0;
    double s = T[r - c0][t];
    for (int q = 0; q < cB; q++) s = s - Lr[q] * T[q][t];
    T[r - c0][t] = s;
  }
  __syncthreads();
  for (int e = tid; e < R * nc; e += blockDim.x) {
    int r = e / nc, t = e % nc;
    M[(size_t)(c0 + r) * N + l0 + t] = T[r][t];
  }
  if (blockIdx.x == 0 && nc > 0) {
    // lookahead: this block holds the next panel (columns l0..l0+nc-1) for
    // rows c0..n; factorise its rows l0..n now.
    const int c1 = l0;
    for (int i = tid; i <= n - c1; i += blockDim.x) vvL[i] = vv[c1 + i];
    __syncthreads();
    panel_factor(T, c1 - c0, c1, nc, vvL, piv, st);
    for (int e = tid; e < (n - c1 + 1) * nc; e += blockDim.x) {
      int r = e / nc, t = e % nc;
      M[(size_t)(c1 + r) * N + l0 + t] = T[c1 - c0 + r][t];
    }
    for (int i = tid; i <= n - c1; i += blockDim.x) vv[c1 + i] = vvL[i];
  }
}

// Back substitution (and, after a zero row, the reference's lubksb on the
// unfactorised matrix with an all-zero index).
// Literal replay of ludcmp (src/bachUtil.cpp:319-384) on the saved matrix,
// one thread, only after a column had no finite pivot candidate.
__global__ __launch_bounds__(64) void k_lu_fallback(double* M, const double* M0, int* piv,
                                                    LuState* st) {
  if (!st->fallback || st->skip || st->zero) return;
  const int n = MATSIZE, N = MATN;
  for (int i = threadIdx.x; i < N * N; i += blockDim.x) M[i] = M0[i];
  __syncthreads();
  if (threadIdx.x != 0) return;
  double vv[MATN];
  for (int i = 0; i < N; i++) piv[i] = 0;
  vv[0] = 0.0;
  for (int i = 1; i <= n; i++) {
    double big = 0.0;
    for (int j = 1; j <= n; j++) {
      double t = fabs(M[(size_t)i * N + j]);
      if (t > big) big = t;
    }
    vv[i] = 1.0 / big;
  }
  int maxI = 0;
  for (int j = 1; j <= n; j++) {
    for (int i = 1; i < j; i++) {
      double sum = M[(size_t)i * N + j];
      for (int k = 1; k < i; k++) sum -= M[(size_t)i * N + k] * M[(size_t)k * N + j];
      M[(size_t)i * N + j] = sum;
    }
    double big = 0.0;
    for (int i = j; i <= n; i++) {
      double sum = M[(size_t)i * N + j];
      for (int k = 1; k < j; k++) sum -= M[(size_t)i * N + k] * M[(size_t)k * N + j];
      M[(size_t)i * N + j] = sum;
      double dum = vv[i] * fabs(sum);
      if (dum >= big) {
        big = dum;
        maxI = i;
      }
    }
    if (j != maxI) {
      for (int k = 1; k <= n; k++) {
        double d = M[(size_t)maxI * N + k];
        M[(size_t)maxI * N + k] = M[(size_t)j * N + k];
        M[(size_t)j * N + k] = d;
      }
      vv[maxI] = vv[j];
    }
    piv[j] = maxI;
    if (M[(size_t)j * N + j] == 0.0) M[(size_t)j * N + j] = 1.0e-20;
    if (j != n) {
      double dum = 1.0 / M[(size_t)j * N + j];
      for (int i = j + 1; i <= n; i++) M[(size_t)i * N + j] *= dum;
    }
  }
}

__global__ __launch_bounds__(64) void k_lu_backsub(const double* M, const double* M0,
                                                   const int* piv, double* b, const double* b0,
                                                   const LuState* st, double* out) {
  const int n = MATSIZE, N = MATN;
  __shared__ double x[NSOL];
  if (st->skip) {
    for (int i = threadIdx.x; i < NSOL; i += blockDim.x)
      if (out) out[i] = 0.0;
    return;
  }
  const bool full = st->zero || st->fallback;  // forward pass not fused
  if (st->zero) M = M0;                         // ludcmp returned before factorising
  for (int i = threadIdx.x; i < NSOL; i += blockDim.x) x[i] = full ? b0[i] : b[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    if (full) {
      int ii = 0;
      for (int i = 1; i <= n; i++) {
        int ip = piv[i];
        double sum = x[ip];
        x[ip] = x[i];
        if (ii) {
          for (int j = ii; j <= i - 1; j++) sum -= M[(size_t)i * N + j] * x[j];
        } else if (sum != 0.0) {
          ii = i;
        }
        x[i] = sum;
      }
    }
  }
  __syncthreads();
  // Back substitution x_i = (y_i - sum_{j>i} U_ij x_j) / U_ii with the sum
  // in ascending j (lubksb).  The products need only final x values, so all
  // lanes form them (row i-1 is prefetched into registers meanwhile); one lane
  // runs the ordered subtraction chain, the only sequential part.
  __shared__ double P[MATN];
  constexpr int PER = (MATSIZE + 63) / 64;  // row entries per lane
  const int lane = threadIdx.x;
  double cur[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    int j = n + 1 + lane + 64 * q;  // row n has no entries right of the diagonal
    cur[q] = 0.0;
    (void)j;
  }
  double diag = M[(size_t)n * N + n];
  for (int i = n; i >= 1; i--) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      int j = i + 1 + lane + 64 * q;
      if (j <= n) P[j] = cur[q] * x[j];
    }
    // prefetch row i-1 (columns i..n) and its diagonal
    double nxt[PER];
    double nd = 0.0;
    if (i > 1) {
      const double* r1 = M + (size_t)(i - 1) * N;
#pragma unroll
      for (int q = 0; q < PER; q++) {
        int j = i + lane + 64 * q;
        nxt[q] = j <= n ? r1[j] : 0.0;
      }
      nd = r1[i - 1];
    }
    __syncthreads();
    if (lane == 0) {
      double sum = x[i];
      for (int j = i + 1; j <= n; j++) sum = sum - P[j];
      x[i] = sum / diag;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) cur[q] = nxt[q];
    diag = nd;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NSOL; i += blockDim.x) {
    b[i] = x[i];
    if (out) out[i] = x[i];
  }
}

hipError_t launch_global_solve(Ctx& c, double* d_out, const int* d_skipIfZero) {
  LuState* st = reinterpret_cast<LuState*>(c.luState);
  prof_begin(c, XB_PROF_SOLVE);
  hipLaunchKernelGGL(k_lu_begin, dim3(1), dim3(LU_THREADS), 0, c.stream, c.M, c.luVV, c.luIndex,
                     st, d_skipIfZero);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  for (int p = 0; p < NPANEL; p++) {
    const int c0 = 1 + p * LB;
    const int cB = std::min(LB, MATSIZE - c0 + 1);
    const int trailing = MATSIZE - (c0 + cB) + 1;
    const int nColBlocks = (trailing + LB - 1) / LB;
    hipLaunchKernelGGL(k_lu_step, dim3(nColBlocks + 1), dim3(STEP_THREADS), 0, c.stream, c.M,
                       c.luVV, c.luIndex, c.rhs, st, p, nColBlocks);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_lu_fallback, dim3(1), dim3(64), 0, c.stream, c.M, c.M0, c.luIndex, st);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_lu_backsub, dim3(1), dim3(64), 0, c.stream, c.M, c.M0, c.luIndex, c.rhs,
                     c.rhs0, st, d_out);
  prof_end(c, XB_PROF_SOLVE);
  return hipGetLastError();
}

}  // namespace xb
