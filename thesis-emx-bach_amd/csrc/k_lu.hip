// k_lu.hip — the global 292x292 FP64 solve, bit-identical to the reference's
// Numerical Recipes ludcmp + lubksb (src/bachUtil.cpp:319-411).
//
// Right-looking blocked elimination with the reference's per-element order:
// element (i,l) receives -L_ik U_kl for k ascending, exactly as Crout forms it,
// and rows are interchanged by the same pivot sequence (implicit scaling vv,
// `>=` rule: the last row attaining the maximum wins; NaN never wins; no
// candidate keeps the previous row).  Applying a panel's row swaps to a
// trailing column before that panel's updates is equivalent to interleaving
// them, because the swaps only exchange rows that the panel's updates treat
// identically.
//
//   k_lu_scale   vv (ludcmp's implicit scaling) and zero-row flags, a wave
//                per row;
//   k_lu_begin   the zero-row test (ludcmp returns 1 before factorising) and
//                the first 16-column panel;
//   k_lu_step(p) one workgroup per 16 trailing columns: load them with the
//                panel's interchanges applied (each row straight to its final
//                position), the in-panel triangular update, the rank-16 update of
//                the rows below (each element summed in k order), with the
//                factorised panel staged once in LDS; workgroup 0 owns the
//                next panel's columns and factorises that panel right after
//                updating it (one launch per panel).  A last workgroup carries
//                the right-hand side as one more column: lubksb's forward
//                substitution, including its `ii` rule (updates start at the
//                first nonzero entry), fused here;
//   k_lu_backsub back substitution; its ascending-j sums are one sequential
//                chain per row in the reference order.
//
// Panels are factorised by one wave holding the panel's rows in registers
// (wave_panel_factor): the pivot vote is a DPP reduction and the pivot row a
// v_readlane broadcast, so a column costs no barrier and no LDS round trip.
// A row interchange relabels the two rows (logical position numbers) instead
// of moving them; vv stays attached to its row, as ludcmp's vv[imax] = vv[j]
// keeps it for every position still searched.  The factorised rows are
// stored at their final positions, and the map from each row's position
// before the panel to its final one goes to the next step, which applies the
// interchanges to the trailing columns and b as it loads them.
#include <algorithm>

#include "xb_dev.h"

namespace xb {

constexpr int LB = 16;          // panel width = trailing columns per workgroup
constexpr int LDP = LB + 1;     // LDS pitch of a 16-column slab
constexpr int NPANEL = (MATSIZE + LB - 1) / LB;
constexpr int LU_THREADS = 256;
constexpr int NSLOT = (MATSIZE + WAVE - 1) / WAVE;  // row entries per lane

struct LuState {
  int zero;      // a zero row: ludcmp returned 1 without factorising
  int lastMaxI;  // pivot row of the previous column (kept when none qualifies)
  int ii;        // lubksb's first nonzero index
  int fallback;  // a column without a finite pivot kept an earlier row: replay
  int skip;      // empty stamp list: the reference never solves
  int pad[3];
};

// One step of a (value, index) maximum across lanes (ties to the larger
// index), the partner chosen by a DPP control within each row of 16 lanes.
template <int CTRL>
__device__ __forceinline__ void argmax_dpp(double& bv, int& bi) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(bv), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(bv), CTRL, 0xF, 0xF, false);
  const int oi = __builtin_amdgcn_update_dpp(0, bi, CTRL, 0xF, 0xF, false);
  const double ov = __hiloint2double(hi, lo);
  if (ov > bv || (ov == bv && oi > bi)) {
    bv = ov;
    bi = oi;
  }
}

// Wave-wide (value, index) maximum: rotations by 1, 2, 4, 8 within each row
// of 16 lanes (DPP row_ror), then the two cross-row exchanges.  Every lane
// ends with the result (a maximum under a total order: any combining order).
__device__ __forceinline__ void wave_argmax(double& bv, int& bi) {
  argmax_dpp<0x121>(bv, bi);
  argmax_dpp<0x122>(bv, bi);
  argmax_dpp<0x124>(bv, bi);
  argmax_dpp<0x128>(bv, bi);
#pragma unroll
  for (int o = 16; o <= 32; o <<= 1) {
    const double ov = __shfl_xor(bv, o);
    const int oi = __shfl_xor(bi, o);
    if (ov > bv || (ov == bv && oi > bi)) {
      bv = ov;
      bi = oi;
    }
  }
}

constexpr int WSLOT = (MATSIZE + WAVE - 1) / WAVE;  // panel rows per lane (one wave)
static_assert(WSLOT == 5, "the step dispatches 1..5 row slots");

// DPP move of a double (both halves), lanes outside RMASK's rows keep v.
template <int CTRL, int RMASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, RMASK,
                                             0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, RMASK,
                                             0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Wave-wide maxima, returned wave-uniform: rotations by 1, 2, 4, 8 within
// each row of 16 lanes give every lane its row's maximum; the row broadcasts
// of lanes 15 and 31 carry them into the last row, whose lane 63 holds the
// maximum of all 64 lanes.  A maximum is exact in any combining order.
__device__ __forceinline__ double wave_max_f64(double v) {
  v = fmax(v, dpp_f64<0x121, 0xF>(v));
  v = fmax(v, dpp_f64<0x122, 0xF>(v));
  v = fmax(v, dpp_f64<0x124, 0xF>(v));
  v = fmax(v, dpp_f64<0x128, 0xF>(v));
  v = fmax(v, dpp_f64<0x142, 0xA>(v));
  v = fmax(v, dpp_f64<0x143, 0xC>(v));
  return readlane_f64(v, 63);
}

__device__ __forceinline__ int wave_max_i32(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x121, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x122, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x124, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x128, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));
  v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));
  return __builtin_amdgcn_readlane(v, 63);
}

// Factorise columns c1..c1+cB-1 over rows c1..n with ONE wave (lanes of wave
// 0 only; no workgroup barrier inside).  S[r - c1][q] holds matrix element
// (r, c1 + q) (updated by all earlier columns); vin[r - c1] the row scale of
// position r; lastMaxI the pivot row of column c1 - 1.  Lane l holds rows
// c1 + l + 64 s (s < NS) in registers.  Per column: each lane's best
// candidate (selects, no branches), the wave maximum of vv*|a| (DPP), then
// the largest position attaining it (ludcmp's `>=` keeps the last row with
// the maximum; NaN never qualifies): when a single lane holds the maximum,
// that lane's row, else a second DPP maximum over the positions.  The pivot
// row is read from its owner lane (v_readlane) into wave-uniform registers,
// so the scaling and the rank-1 update of the later panel columns need no
// memory traffic.  A row interchange relabels the two rows (logical position
// numbers) instead of moving them; vv stays attached to its row.  When a row
// becomes the pivot of column k it is copied to LDS (its L and U parts are
// final then); its registers are never read again.  Outputs: the pivot rows
// at positions c1.., the other rows' panel columns at their final positions,
// vv by final position, piv, and dst[physical row] = final position, with
// which the next step loads the trailing columns.
// NS = row slots per lane: 64 * NS >= n - c1 + 1 (slots past the last row
// would only hold empty rows).
template <int NS>
__device__ __forceinline__ void wave_panel_factor(const double (*S)[LDP], int c1, int cB,
                                                  const double* vin, int lastMaxI, double* M,
                                                  double* vout, int* piv, int* dst,
                                                  LuState* st) {
  const int n = MATSIZE, N = MATN;
  const int lane = lane_id();
  __shared__ double prow[LB][LB];
  __shared__ int ppiv[LB];
  double a[NS][LB];
  double vr[NS];
  int lidx[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int r = c1 + lane + WAVE * s;
    const bool ok = r <= n;
    lidx[s] = ok ? r : -1;
    vr[s] = ok ? vin[r - c1] : 0.0;
#pragma unroll
    for (int q = 0; q < LB; q++) a[s][q] = (ok && q < cB) ? S[r - c1][q] : 0.0;
  }
  int maxI = lastMaxI;
  bool fb = false;
  // The branch-free part of column q's pivot vote (max vv*|a| over positions
  // >= c1 + q, ties to the larger position): each lane's best candidate and
  // the wave maximum.  Column q + 1's vote runs as soon as that column has
  // its last update, before the rest of column q's updates, which then fill
  // the DPP chain's latency; every element still gets its updates in column
  // order.
  double bv, big;
  int bi;
  auto vote = [&](const int q) __attribute__((always_inline)) {
    const int k = c1 + q;
    bv = -1.0;
    bi = -1;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const double dum = vr[s] * fabs(a[s][q]);
      const bool take = lidx[s] >= k && dum >= 0.0 && (dum > bv || (dum == bv && lidx[s] > bi));
      bv = take ? dum : bv;
      bi = take ? lidx[s] : bi;
    }
    big = wave_max_f64(bv);
  };
  vote(0);
#pragma unroll
  for (int q = 0; q < LB; q++) {
    const int k = c1 + q;
    if (q < cB && !fb) {
      const bool tie = bi >= 0 && bv == big;
      const unsigned long long cand = __ballot(tie);
      int best = -1;
      if (cand != 0ull)
        best = (cand & (cand - 1ull)) == 0ull
                   ? __builtin_amdgcn_readlane(bi, __builtin_ctzll(cand))
                   : wave_max_i32(tie ? bi : -1);
      if (best >= 0) maxI = best;
      if (maxI < k) {
        // No row qualified and the kept pivot row lies above this column:
        // the reference swaps with an already-factorised row, which only a
        // literal replay reproduces (at the start of k_lu_backsub).
        fb = true;
      } else {
        double u[LB];
        bool found = false;
#pragma unroll
        for (int s = 0; s < NS; s++) {
          const unsigned long long own = __ballot(lidx[s] == maxI);
          if (!found && own != 0ull) {
            const int src = __builtin_ctzll(own);
            // columns q.. only: the row's L part (columns before q) is
            // final in its owner lane's registers and stored from there
#pragma unroll
            for (int l = q; l < LB; l++) u[l] = readlane_f64(a[s][l], src);
            found = true;
          }
        }
        if (u[q] == 0.0) u[q] = 1.0e-20;
        if (lane == 0) {
#pragma unroll
          for (int l = q; l < LB; l++) prow[q][l] = u[l];
          ppiv[q] = maxI;
        }
        if (maxI != k) {
#pragma unroll
          for (int s = 0; s < NS; s++)
            lidx[s] = lidx[s] == k ? maxI : (lidx[s] == maxI ? k : lidx[s]);
        }
        if (k != n) {
          const double dinv = 1.0 / u[q];
#pragma unroll
          for (int s = 0; s < NS; s++) a[s][q] *= dinv;
          if (q + 1 < LB) {
#pragma unroll
            for (int s = 0; s < NS; s++) a[s][q + 1] = a[s][q + 1] - a[s][q] * u[q + 1];
            vote(q + 1);
#pragma unroll
            for (int l = q + 2; l < LB; l++) {
#pragma unroll
              for (int s = 0; s < NS; s++) a[s][l] = a[s][l] - a[s][q] * u[l];
            }
          }
        }
      }
    }
  }
  // lane 0's LDS writes precede these reads in the wave's LDS order; only
  // the compiler needs fencing
  __builtin_amdgcn_wave_barrier();
  __asm__ __volatile__("" ::: "memory");
  if (fb) {
    if (lane == 0) st->fallback = 1;
    return;
  }
  // the pivot rows' U parts (columns from their own on) from LDS ...
  for (int e = lane; e < LB * LB; e += WAVE) {
    const int q = e / LB, l = e % LB;
    if (q < cB && l < cB && l >= q) M[(size_t)(c1 + q) * N + c1 + l] = prow[q][l];
  }
  if (lane < cB) piv[c1 + lane] = ppiv[lane];
  // ... and every other element from its row's registers: all panel
  // columns of a row below the panel, and a pivot row's L part (the columns
  // before its own, untouched since their scaling)
#pragma unroll
  for (int s = 0; s < NS; s++) {
    if (lidx[s] >= c1) {
      const int own = lidx[s] - c1;  // the row's pivot column, or >= cB
      double* row = M + (size_t)lidx[s] * N + c1;
#pragma unroll
      for (int q = 0; q < LB; q++)
        if (q < cB && q < own) row[q] = a[s][q];
    }
    if (lidx[s] > 0) {
      vout[lidx[s]] = vr[s];
      dst[c1 + lane + WAVE * s] = lidx[s];
    }
  }
  if (lane == 0) st->lastMaxI = maxI;
}

__device__ __forceinline__ int* lu_zero_flags(LuState* st) {
  return reinterpret_cast<int*>(reinterpret_cast<char*>(st) + 64);
}

// final position of each physical row after panel p's interchanges (a step
// reads its panel's while its workgroup 0 writes the next panel's); all panels'
// maps are kept for the iterative refinement's forward substitution
static_assert(NPANEL == LU_PANELS, "one row map per panel");
__device__ __forceinline__ int* lu_dst(LuState* st, int p) {
  return lu_zero_flags(st) + MATN * (1 + p);
}
__device__ __forceinline__ const int* lu_dst(const LuState* st, int p) {
  return reinterpret_cast<const int*>(reinterpret_cast<const char*>(st) + 64) + MATN * (1 + p);
}

// vv (ludcmp's implicit row scaling) and the zero-row flags: one wave per row,
// lanes across the columns.  The maximum of |a| with NaN never taken is the
// same in any order.
__device__ __forceinline__ void lu_scale_dev(const double* M, double* vv, LuState* st,
                                             const int* skipIfZero) {
  if (skipIfZero && *skipIfZero == 0) return;
  const int n = MATSIZE, N = MATN;
  const int lane = lane_id();
  const int i = 1 + (int)blockIdx.x * (LU_THREADS / WAVE) + (int)threadIdx.x / WAVE;
  if (i > n) return;
  const double* row = M + (size_t)i * N;
  double v[NSLOT];
#pragma unroll
  for (int s = 0; s < NSLOT; s++) {
    const int j = 1 + lane + WAVE * s;
    v[s] = j <= n ? fabs(row[j]) : 0.0;
  }
  double big = 0.0;
#pragma unroll
  for (int s = 0; s < NSLOT; s++)
    if (v[s] > big) big = v[s];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(big, o);
    if (t > big) big = t;
  }
  if (lane == 0) {
    vv[i] = 1.0 / big;
    lu_zero_flags(st)[i] = big == 0.0;
  }
}

// Staging of rows r0..r0+nr-1, columns c..c+LB-1 (zero past cB) of M into an
// LDS slab, with every thread's loads in flight together.
__device__ __forceinline__ void stage_slab(double (*D)[LDP], const double* M, int r0, int nr,
                                           int c, int cB) {
  constexpr int IT = (MATSIZE * LB + LU_THREADS - 1) / LU_THREADS;
  double v[IT];
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int e = (int)threadIdx.x + it * LU_THREADS;
    const int r = e / LB, q = e % LB;
    v[it] = (r < nr && q < cB) ? M[(size_t)(r0 + r) * MATN + c + q] : 0.0;
  }
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int e = (int)threadIdx.x + it * LU_THREADS;
    const int r = e / LB, q = e % LB;
    if (r < nr) D[r][q] = v[it];
  }
}

__device__ __forceinline__ void lu_begin_dev(double* M, double* vv, int* piv, LuState* st,
                                             const int* skipIfZero) {
  const int n = MATSIZE, N = MATN;
  const int tid = threadIdx.x;
  __shared__ int s_zero;
  __shared__ double S[MATSIZE][LDP];
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
  int z = 0;
  for (int i = 1 + tid; i <= n; i += blockDim.x) z |= lu_zero_flags(st)[i];
  if (z) atomicOr(&s_zero, 1);
  stage_slab(S, M, 1, n, 1, LB);
  __syncthreads();
  if (tid == 0) {
    st->zero = s_zero;
    st->lastMaxI = 0;
    st->ii = 0;
  }
  if (s_zero) return;  // ludcmp returns before factorising
  if (tid < WAVE) wave_panel_factor<WSLOT>(S, 1, LB, vv + 1, 0, M, vv, piv, lu_dst(st, 0), st);
}

__device__ __forceinline__ void lu_step_dev(double* M, double* vv, int* piv, double* b,
                                            LuState* st, int p, int nColBlocks) {
  if (st->zero || st->skip || st->fallback) return;
  const int n = MATSIZE, N = MATN;
  const int tid = threadIdx.x;
  const int c0 = 1 + p * LB;
  const int cB = min(LB, n - c0 + 1);
  const int R = n - c0 + 1;  // rows c0..n
  const bool rhsBlock = (int)blockIdx.x == nColBlocks;
  // Workgroups go to the 8 XCDs round robin.  Column block 1 (the panel after
  // next) is updated by workgroup 8, on workgroup 0's XCD, so the next step's
  // workgroup 0 finds those columns in its own L2; workgroup 1 takes block 8.
  int cbk = (int)blockIdx.x;
  if (nColBlocks > 8 && (cbk == 1 || cbk == 8)) cbk = 9 - cbk;
  const int l0 = c0 + cB + cbk * LB;  // first column of this block
  const int nc = rhsBlock ? 0 : min(LB, n - l0 + 1);
  __shared__ double Lp[MATSIZE][LDP];  // the factorised panel, rows c0..n
  __shared__ double T[MATSIZE][LDP];   // this block's columns, rows c0..n
  __shared__ double vS[MATSIZE];       // block 0: row scales; rhs block: b (rows c0..n)
  __shared__ int s_int;                // block 0: last pivot row; rhs block: ii
  {
    // every load of the step in flight together: one memory round trip.  The
    // panel's row interchanges are applied here: the trailing columns and b
    // are still in the order before them, and each of their rows goes to its
    // final position dst[row] (written when the panel was factorised); the
    // panel's own columns were stored at their final positions then.
    constexpr int IT = (MATSIZE * LB + LU_THREADS - 1) / LU_THREADS;
    constexpr int IV = (MATSIZE + LU_THREADS - 1) / LU_THREADS;
    const int* dmap = lu_dst(st, p);
    double va[IT], vt[IT], vr[IV];
    int dt[IT], dv[IV];
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int e = tid + it * LU_THREADS;
      const int r = e / LB, q = e % LB;
      const bool in = r < R;
      va[it] = (in && q < cB) ? M[(size_t)(c0 + r) * N + c0 + q] : 0.0;
      vt[it] = (in && q < nc) ? M[(size_t)(c0 + r) * N + l0 + q] : 0.0;
      dt[it] = (in && nc > 0) ? dmap[c0 + r] : c0 + r;
    }
    const bool wantV = rhsBlock || blockIdx.x == 0;
#pragma unroll
    for (int it = 0; it < IV; it++) {
      const int r = tid + it * LU_THREADS;
      vr[it] = (wantV && r < R) ? (rhsBlock ? b[c0 + r] : vv[c0 + r]) : 0.0;
      dv[it] = (rhsBlock && r < R) ? dmap[c0 + r] : c0 + r;
    }
    const int si = tid == 0 ? (rhsBlock ? st->ii : st->lastMaxI) : 0;
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int e = tid + it * LU_THREADS;
      const int r = e / LB, q = e % LB;
      if (r < R) {
        Lp[r][q] = va[it];
        T[min(max(dt[it] - c0, 0), R - 1)][q] = vt[it];
      }
    }
#pragma unroll
    for (int it = 0; it < IV; it++) {
      const int r = tid + it * LU_THREADS;
      if (r < R) vS[min(max(dv[it] - c0, 0), R - 1)] = vr[it];  // vv is stored by final position
    }
    if (tid == 0) s_int = si;
  }
  __syncthreads();
  if (rhsBlock) {
    // right-hand side column (already interchanged): forward substitution (lubksb)
    double* y = vS;
    for (int q = 0; q < cB; q++) {
      if (tid == 0 && s_int == 0 && !(y[q] == 0.0)) s_int = c0 + q;
      __syncthreads();
      if (s_int != 0) {
        const double yk = y[q];
        for (int r = q + 1 + tid; r < R; r += LU_THREADS) y[r] = y[r] - Lp[r][q] * yk;
      }
      __syncthreads();
    }
    for (int r = tid; r < R; r += LU_THREADS) b[c0 + r] = y[r];
    if (tid == 0) st->ii = s_int;
    return;
  }
  // trailing column block (panels with trailing columns are full: cB == LB)
  {
    // U12 = L11^-1 A12, one thread per element: row r = tid % 16 and column
    // t = tid / 16, so the 16 lanes of a DPP row hold one column.  At step q
    // lane q's value is final (it took steps 0..q-1), every lane of the row
    // receives it by DPP row broadcast, and the rows below q subtract
    // L_rq u_q: each element takes k ascending, as the reference's Crout sums.
    // The product is fma(u_q, L_rq, -0.0), which is the rounded product
    // L_rq u_q exactly (signed zeros included); the subtraction is separate.
    // The s_nop covers the DPP read of s after the previous step's VALU write.
    static_assert(LB * LB == LU_THREADS, "one thread per U12 element");
    const int r = tid & (LB - 1), t = tid / LB;
    double s = T[r][t];
    double Lr[LB];
#pragma unroll
    for (int q = 0; q < LB - 1; q++) Lr[q] = Lp[r][q];
#define XB_U12_STEP(K)                                                              \
  {                                                                                 \
    double p = -0.0;                                                                \
    __asm__ __volatile__("s_nop 1\n\tv_fmac_f64_dpp %0, %1, %2 row_newbcast:" #K      \
                         " row_mask:0xf bank_mask:0xf"                              \
                         : "+v"(p)                                                  \
                         : "v"(s), "v"(Lr[K]));                                     \
    if (r > K) s = s - p;                                                           \
  }
    XB_U12_STEP(0) XB_U12_STEP(1) XB_U12_STEP(2) XB_U12_STEP(3) XB_U12_STEP(4)
    XB_U12_STEP(5) XB_U12_STEP(6) XB_U12_STEP(7) XB_U12_STEP(8) XB_U12_STEP(9)
    XB_U12_STEP(10) XB_U12_STEP(11) XB_U12_STEP(12) XB_U12_STEP(13) XB_U12_STEP(14)
#undef XB_U12_STEP
    T[r][t] = s;
  }
  __syncthreads();
  {
    // rows below the panel: rank-16 update, k ascending per element
    const int t = tid % LB, rg = tid / LB;
    double U[LB];
#pragma unroll
    for (int q = 0; q < LB; q++) U[q] = T[q][t];
    for (int r = LB + rg; r < R; r += LU_THREADS / LB) {
      double s = T[r][t];
#pragma unroll
      for (int q = 0; q < LB; q++) s = s - Lp[r][q] * U[q];
      T[r][t] = s;
    }
  }
  __syncthreads();
  if (blockIdx.x == 0) {
    // lookahead: these columns are the next panel; store its U rows, then
    // factorise rows l0..n
    for (int e = tid; e < LB * LB; e += LU_THREADS) {
      const int r = e / LB, t = e % LB;
      if (t < nc) M[(size_t)(c0 + r) * N + l0 + t] = T[r][t];
    }
    if (tid < WAVE) {
      // rows l0..n of the next panel: as many register slots as they need
      int* dnext = lu_dst(st, p + 1);
      const int rows = n - l0 + 1;
      if (rows > 4 * WAVE)
        wave_panel_factor<5>(&T[LB], l0, nc, vS + LB, s_int, M, vv, piv, dnext, st);
      else if (rows > 3 * WAVE)
        wave_panel_factor<4>(&T[LB], l0, nc, vS + LB, s_int, M, vv, piv, dnext, st);
      else if (rows > 2 * WAVE)
        wave_panel_factor<3>(&T[LB], l0, nc, vS + LB, s_int, M, vv, piv, dnext, st);
      else if (rows > WAVE)
        wave_panel_factor<2>(&T[LB], l0, nc, vS + LB, s_int, M, vv, piv, dnext, st);
      else
        wave_panel_factor<1>(&T[LB], l0, nc, vS + LB, s_int, M, vv, piv, dnext, st);
    }
    return;
  }
  for (int e = tid; e < R * LB; e += LU_THREADS) {
    const int r = e / LB, t = e % LB;
    if (t < nc) M[(size_t)(c0 + r) * N + l0 + t] = T[r][t];
  }
}

// Literal replay of ludcmp (src/bachUtil.cpp:319-384) on the saved matrix,
// only after a column had no finite pivot candidate: the back substitution's
// workgroup copies the matrix back, then its thread 0 replays.
__device__ void lu_fallback_replay(double* M, const double* M0, int* piv) {
  const int n = MATSIZE, N = MATN;
  for (int i = threadIdx.x; i < N * N; i += blockDim.x) M[i] = M0[i];
  __syncthreads();
  __shared__ double vv[MATN];  // LDS, not per-thread scratch
  if (threadIdx.x != 0) return;
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

// Back substitution (and, after a zero row or a replay, the reference's whole
// lubksb: its forward pass was not fused into the elimination then).
//
// x_t = (y_t - sum_{j>t} U_tj x_j) / U_tt with the sum as one chain in
// ascending j, the reference's order.  Wave 0 runs the chains; waves 1-3
// prepare the next row one phase ahead, so a phase (one row) costs the chain
// plus one division and one barrier:
//   * the products P_j = U_tj x_j (j >= t+3; +0.0 elsewhere) go to LDS, and
//     wave 0 loads them into registers laid out by 16-column block (lane l of
//     every row of 16 lanes holds column 16 m + l of block m), so each chain
//     step is one v_fmac_f64 whose operand arrives by DPP row broadcast:
//     fma(P_j, -1, s) is s - P_j rounded once, the reference's subtraction;
//     the +0.0 entries before t+3 let every chain start at a block boundary
//     (s - (+0.0) is s bit for bit);
//   * the first two terms, U_t,t+1 x_t+1 and U_t,t+2 x_t+2, wave 0 forms
//     itself: so a phase opens with the barrier, then wave 0 divides for
//     x_t+1 while its loads of row t's products are in flight, and the
//     producers of row t-1 need no x newer than x_t+2;
//   * U rows stream from global memory into an 8-row LDS ring, loaded four
//     phases ahead into four register sets by row mod 4.
// The phase barrier waits only for LDS traffic (the ring's global loads stay
// in flight across it).
constexpr int BS_THREADS = 256;
constexpr int BS_RING = 8;                             // staged U rows
constexpr int BS_COLS = 320;                           // padded row length
constexpr int BS_BLK = (MATSIZE + 2 + 15) / 16;        // 19 blocks of 16 columns
constexpr int BS_PL = BS_THREADS - WAVE;               // 192 producer lanes
constexpr int BS_PER = (BS_COLS + BS_PL - 1) / BS_PL;  // 2 columns per producer lane
static_assert(16 * BS_BLK <= BS_COLS && MATSIZE % 4 == 0, "back substitution layout");

__device__ __forceinline__ void lds_barrier() {
  __asm__ __volatile__("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

#define XB_CHAIN(K)                                                                   \
  __asm__ __volatile__("v_fmac_f64_dpp %0, %1, %2 row_newbcast:" #K                   \
                       " row_mask:0xf bank_mask:0xf"                                  \
                       : "+v"(s)                                                      \
                       : "v"(p), "v"(m1));

// refine: the iterative-refinement solve (k_lu_refine_*): lubksb of b0 with
// the factors already in Mw (forward and back substitution, no replay), the
// correction left in b; nothing to correct after a skipped or zero-row solve.
__device__ __forceinline__ void lu_backsub_dev(double* Mw, const double* M0, int* piv, double* b,
                                               const double* b0, const LuState* st, double* out,
                                               int refine) {
  const double* M = Mw;
  const int n = MATSIZE, N = MATN;
  const int tid = threadIdx.x, lane = lane_id();
  __shared__ double xs[NSOL];                // y, replaced by x as rows finish
  __shared__ double Pb[2][BS_COLS];          // products of a row, by row parity
  __shared__ double ring[BS_RING][BS_COLS];  // U rows, row r in slot r % 8
  __shared__ double scb[2][4];               // y_t, U_tt, U_t,t+1, U_t,t+2 by row parity
  if (st->skip || (refine && st->zero)) {
    for (int i = tid; i < NSOL; i += BS_THREADS) {
      if (out) out[i] = 0.0;
      if (refine) b[i] = 0.0;
    }
    return;
  }
  if (!refine && st->fallback && !st->zero) {
    lu_fallback_replay(Mw, M0, piv);
    __syncthreads();
  }
  // forward pass not fused: after a zero row or a replay (whose factors are in
  // the reference's layout, so the refinement uses its lubksb too)
  const bool full = refine ? st->fallback != 0 : (st->zero || st->fallback);
  if (!refine && st->zero) M = M0;  // ludcmp returned before factorising
  for (int i = tid; i < NSOL; i += BS_THREADS) xs[i] = full ? b0[i] : b[i];
  __syncthreads();
  if (full && tid == 0) {
    int ii = 0;
    for (int i = 1; i <= n; i++) {
      int ip = piv[i];
      double sum = xs[ip];
      xs[ip] = xs[i];
      if (ii) {
        for (int j = ii; j <= i - 1; j++) sum -= M[(size_t)i * N + j] * xs[j];
      } else if (sum != 0.0) {
        ii = i;
      }
      xs[i] = sum;
    }
  }
  __syncthreads();
  if (tid < WAVE) {
    lds_barrier();  // ring filled
    lds_barrier();  // row n prepared
    const double m1 = -1.0;
    double s = scb[n & 1][0];  // row n: no terms
    double ud = scb[n & 1][1];
    double xn1 = 0.0;  // x_t+1 (none for t = n)
#pragma unroll
    for (int mb = BS_BLK - 1; mb >= 0; mb--) {
      // phases whose chain row r has its products from block mb on:
      // (r + 3) / 16 == mb; phase r finishes row t = r + 1 and runs row r
      const int rHi = min(n - 1, 16 * mb + 12), rLo = max(0, 16 * mb - 3);
      for (int r = rHi; r >= rLo; r--) {
        lds_barrier();
        const int par = r & 1;
        double P[BS_BLK];
#pragma unroll
        for (int m = mb; m < BS_BLK; m++) P[m] = Pb[par][16 * m + (lane & 15)];
        const double y = scb[par][0], udr = scb[par][1], u1 = scb[par][2], u2 = scb[par][3];
        const double xt = s / ud;  // x_{r+1}
        if (lane == 0) xs[r + 1] = xt;
        if (r >= 1) {
          // row r: j = r+1, r+2 here, then the staged products in j order
          s = y - u1 * xt;
          s = s - u2 * xn1;
#pragma unroll
          for (int m = mb; m < BS_BLK; m++) {
            const double p = P[m];
            XB_CHAIN(0) XB_CHAIN(1) XB_CHAIN(2) XB_CHAIN(3) XB_CHAIN(4) XB_CHAIN(5)
            XB_CHAIN(6) XB_CHAIN(7) XB_CHAIN(8) XB_CHAIN(9) XB_CHAIN(10) XB_CHAIN(11)
            XB_CHAIN(12) XB_CHAIN(13) XB_CHAIN(14) XB_CHAIN(15)
          }
          ud = udr;
        }
        xn1 = xt;
      }
    }
  } else {
    const int pl = tid - WAVE;
    // every phase issues exactly BS_PER loads per lane (clamped addresses
    // inside the matrix, masked when stored), so the wait for a set loaded
    // four phases earlier is vmcnt(3 * BS_PER)
    auto loadRow = [&](int r, double (&v)[BS_PER]) {
      const int rr = min(max(r, 1), n);
#pragma unroll
      for (int h = 0; h < BS_PER; h++) {
        const int j = min(pl + BS_PL * h, n);
        v[h] = M[(size_t)rr * N + j];
      }
    };
    auto storeRow = [&](int r, const double (&v)[BS_PER]) {
#pragma unroll
      for (int h = 0; h < BS_PER; h++) {
        const int j = pl + BS_PL * h;
        if (r >= 1 && j < BS_COLS) ring[r % BS_RING][j] = (j >= r && j <= n) ? v[h] : 0.0;
      }
    };
    // products (j >= r+3) and scalars of row r (r >= 1), from its staged U
    // row; x_j for j >= r+3 is published by then
    auto prep = [&](int r) {
      const double* U = ring[r % BS_RING];
#pragma unroll
      for (int h = 0; h < BS_PER; h++) {
        const int j = pl + BS_PL * h;
        if (j < BS_COLS) Pb[r & 1][j] = (j >= r + 3 && j <= n) ? U[j] * xs[j] : 0.0;
      }
      if (pl == 0) {
        scb[r & 1][0] = xs[r];
        scb[r & 1][1] = U[r];
        scb[r & 1][2] = r + 1 <= n ? U[r + 1] : 0.0;
        scb[r & 1][3] = r + 2 <= n ? U[r + 2] : 0.0;
      }
    };
    double ld[4][BS_PER];  // row r in flight in set r % 4
    {
      double v[BS_RING][BS_PER];
#pragma unroll
      for (int q = 0; q < BS_RING; q++) loadRow(n - q, v[q]);
#pragma unroll
      for (int r = n - BS_RING; r > n - BS_RING - 4; r--) loadRow(r, ld[r & 3]);
#pragma unroll
      for (int q = 0; q < BS_RING; q++) storeRow(n - q, v[q]);
    }
    lds_barrier();
    prep(n);
    lds_barrier();
    // Phase k (between barriers 1 + k and 2 + k) prepares row t - 1 = n-1-k,
    // which wave 0 runs in the next phase; its products need x_j for
    // j >= n-k+2, all written by wave 0 before barrier 1 + k.
    for (int c = 0; c < n / 4; c++) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int t = n - 4 * c - u;
        const int S = (n - u) & 3;  // == t % 4, a constant after unrolling
        if (t - 1 >= 1) prep(t - 1);
        // vmcnt(6): the set loaded four phases ago (expcnt, lgkmcnt not waited)
        static_assert(3 * BS_PER == 6, "vmcnt below");
        __builtin_amdgcn_s_waitcnt(0x0F76);
        storeRow(t - BS_RING, ld[S]);
        loadRow(t - BS_RING - 4, ld[S]);
        lds_barrier();
      }
    }
  }
  lds_barrier();  // x_1, written after the last phase barrier
  for (int i = tid; i < NSOL; i += BS_THREADS) {
    b[i] = xs[i];
    if (out) out[i] = xs[i];
  }
}
#undef XB_CHAIN

// ---------------------------------------------------------------------------
// Entry points: one system (the arguments), or a batch of systems of the same
// size, one per blockIdx.y, their buffers passed by value in the kernel
// arguments (read with scalar loads; each workgroup works on one system, with
// the same per-element arithmetic as alone, so every system's result is
// bit-identical to its single solve).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(LU_THREADS) void k_lu_scale(const double* M, double* vv,
                                                         LuState* st, const int* skipIfZero) {
  lu_scale_dev(M, vv, st, skipIfZero);
}
__global__ __launch_bounds__(LU_THREADS) void k_lu_begin(double* M, double* vv, int* piv,
                                                         LuState* st, const int* skipIfZero) {
  lu_begin_dev(M, vv, piv, st, skipIfZero);
}
__global__ __launch_bounds__(LU_THREADS) void k_lu_step(double* M, double* vv, int* piv,
                                                        double* b, LuState* st, int p,
                                                        int nColBlocks) {
  lu_step_dev(M, vv, piv, b, st, p, nColBlocks);
}
__global__ __launch_bounds__(BS_THREADS) void k_lu_backsub(double* Mw, const double* M0,
                                                           int* piv, double* b,
                                                           const double* b0, const LuState* st,
                                                           double* out, int refine) {
  lu_backsub_dev(Mw, M0, piv, b, b0, st, out, refine);
}

__global__ __launch_bounds__(LU_THREADS) void k_lu_scale_b(LuBatch B) {
  const LuSys& S = B.s[blockIdx.y];
  lu_scale_dev(S.M, S.vv, S.st, S.skip);
}
__global__ __launch_bounds__(LU_THREADS) void k_lu_begin_b(LuBatch B) {
  const LuSys& S = B.s[blockIdx.y];
  lu_begin_dev(S.M, S.vv, S.piv, S.st, S.skip);
}
__global__ __launch_bounds__(LU_THREADS) void k_lu_step_b(LuBatch B, int p, int nColBlocks) {
  const LuSys& S = B.s[blockIdx.y];
  lu_step_dev(S.M, S.vv, S.piv, S.rhs, S.st, p, nColBlocks);
}
__global__ __launch_bounds__(BS_THREADS) void k_lu_backsub_b(LuBatch B) {
  const LuSys& S = B.s[blockIdx.y];
  lu_backsub_dev(S.M, S.M0, S.piv, S.rhs, S.rhs0, S.st, S.out, 0);
}

// ---------------------------------------------------------------------------
// FP64 iterative refinement (XB_FORM_REFINE): r = b - A x with the system as
// built (M0, rhs0), each product's rounding error kept (fma) and the sums in
// double-double (TwoSum), so r is the residual rounded once; then A d = r with
// the same LU factors (k_lu_backsub in refine mode) and x += d.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void dd_add(double& hi, double& lo, double h2, double l2) {
  double s, e;
  two_sum(hi, h2, s, e);
  e = e + (lo + l2);
  hi = s + e;
  lo = e - (hi - s);
}

// one wave per row i = 1..292
__global__ __launch_bounds__(LU_THREADS) void k_lu_refine_residual(const double* M0,
                                                                   const double* rhs0,
                                                                   const double* x,
                                                                   double* r,
                                                                   const LuState* st) {
  const int n = MATSIZE, N = MATN;
  const int lane = lane_id();
  const int i = 1 + (int)blockIdx.x * (LU_THREADS / WAVE) + (int)threadIdx.x / WAVE;
  if (i > n) return;
  if (st->skip || st->zero) {
    if (lane == 0) r[i] = 0.0;
    return;
  }
  double hi = lane == 0 ? rhs0[i] : 0.0, lo = 0.0;
#pragma unroll
  for (int sl = 0; sl < NSLOT; sl++) {
    const int j = 1 + lane + WAVE * sl;
    if (j <= n) {
      const double a = M0[(size_t)i * N + j], xj = x[j];
      const double p = a * xj;
      const double e = fma(a, xj, -p);  // a*xj = p + e exactly
      dd_add(hi, lo, -p, -e);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double h2 = __shfl_xor(hi, o), l2 = __shfl_xor(lo, o);
    dd_add(hi, lo, h2, l2);
  }
  if (lane == 0) r[i] = hi + lo;
}

// The refinement's forward substitution with the blocked factors: panel by
// panel, the residual's rows go to their positions after that panel's
// interchanges (its row map), then the panel's L columns (stored in that row
// order) update the rows below, as the factorisation's right-hand-side block
// does.  One workgroup; nothing to do after a replay (lubksb then) or when the
// solve was skipped or hit a zero row.
__global__ __launch_bounds__(LU_THREADS) void k_lu_refine_forward(const double* M, double* y,
                                                                  const LuState* st) {
  if (st->skip || st->zero || st->fallback) return;
  const int n = MATSIZE, N = MATN;
  const int tid = threadIdx.x;
  __shared__ double v[MATN], t[MATN];
  for (int i = 1 + tid; i <= n; i += LU_THREADS) v[i] = y[i];
  __syncthreads();
  for (int p = 0; p < NPANEL; p++) {
    const int c0 = 1 + p * LB, cB = min(LB, n - c0 + 1), R = n - c0 + 1;
    const int* dmap = lu_dst(st, p);
    for (int r = tid; r < R; r += LU_THREADS) t[min(max(dmap[c0 + r] - c0, 0), R - 1)] = v[c0 + r];
    __syncthreads();
    for (int r = tid; r < R; r += LU_THREADS) v[c0 + r] = t[r];
    __syncthreads();
    for (int q = 0; q < cB; q++) {
      const double yk = v[c0 + q];
      for (int r = q + 1 + tid; r < R; r += LU_THREADS)
        v[c0 + r] = v[c0 + r] - M[(size_t)(c0 + r) * N + c0 + q] * yk;
      __syncthreads();
    }
  }
  for (int i = 1 + tid; i <= n; i += LU_THREADS) y[i] = v[i];
}

__global__ void k_lu_refine_add(double* x, const double* d, const LuState* st) {
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i > MATSIZE || st->skip || st->zero) return;
  x[i] = x[i] + d[i];
}

// The solve's kernels in order on stream s.
static hipError_t enqueue_solve(hipStream_t s, const LuSet& L, double* d_out,
                                const int* d_skipIfZero) {
  LuState* st = reinterpret_cast<LuState*>(L.luState);
  hipLaunchKernelGGL(k_lu_scale, dim3((MATSIZE + LU_THREADS / WAVE - 1) / (LU_THREADS / WAVE)),
                     dim3(LU_THREADS), 0, s, L.M, L.luVV, st, d_skipIfZero);
  hipLaunchKernelGGL(k_lu_begin, dim3(1), dim3(LU_THREADS), 0, s, L.M, L.luVV, L.luIndex, st,
                     d_skipIfZero);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  for (int p = 0; p < NPANEL; p++) {
    const int c0 = 1 + p * LB;
    const int cB = std::min(LB, MATSIZE - c0 + 1);
    const int trailing = MATSIZE - (c0 + cB) + 1;
    const int nColBlocks = (trailing + LB - 1) / LB;
    hipLaunchKernelGGL(k_lu_step, dim3(nColBlocks + 1), dim3(LU_THREADS), 0, s, L.M, L.luVV,
                       L.luIndex, L.rhs, st, p, nColBlocks);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_lu_backsub, dim3(1), dim3(BS_THREADS), 0, s, L.M, L.M0, L.luIndex, L.rhs,
                     L.rhs0, st, d_out, 0);
  return hipGetLastError();
}

// One step of iterative refinement of the solution x (d_out, else the rhs
// buffer the substitution left it in) on stream s.
static hipError_t enqueue_refine(hipStream_t s, const LuSet& L, double* d_out) {
  LuState* st = reinterpret_cast<LuState*>(L.luState);
  double* x = d_out ? d_out : L.rhs;
  hipLaunchKernelGGL(k_lu_refine_residual,
                     dim3((MATSIZE + LU_THREADS / WAVE - 1) / (LU_THREADS / WAVE)),
                     dim3(LU_THREADS), 0, s, L.M0, L.rhs0, x, L.rr, st);
  hipLaunchKernelGGL(k_lu_refine_forward, dim3(1), dim3(LU_THREADS), 0, s, L.M, L.rr, st);
  hipLaunchKernelGGL(k_lu_backsub, dim3(1), dim3(BS_THREADS), 0, s, L.M, L.M0, L.luIndex, L.rr,
                     L.rr, st, nullptr, 1);
  hipLaunchKernelGGL(k_lu_refine_add, dim3((MATSIZE + 255) / 256), dim3(256), 0, s, x, L.rr, st);
  return hipGetLastError();
}

// The global solves of several systems (of different pairs) as one batch on
// stream s: the same 23 launches as one solve, each covering every system.
hipError_t launch_global_solve_batch(hipStream_t s, const LuBatch& B, bool ldlt, bool refine,
                                     int ldltGroup) {
  if (ldlt) return launch_ldlt_solve_batch(s, B, refine, ldltGroup);
  if (B.n <= 0) return hipSuccess;
  if (B.n > LU_BATCH_MAX) return hipErrorInvalidValue;
  const int nb = B.n;
  hipLaunchKernelGGL(k_lu_scale_b,
                     dim3((MATSIZE + LU_THREADS / WAVE - 1) / (LU_THREADS / WAVE), nb),
                     dim3(LU_THREADS), 0, s, B);
  hipLaunchKernelGGL(k_lu_begin_b, dim3(1, nb), dim3(LU_THREADS), 0, s, B);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  for (int p = 0; p < NPANEL; p++) {
    const int c0 = 1 + p * LB;
    const int cB = std::min(LB, MATSIZE - c0 + 1);
    const int trailing = MATSIZE - (c0 + cB) + 1;
    const int nColBlocks = (trailing + LB - 1) / LB;
    hipLaunchKernelGGL(k_lu_step_b, dim3(nColBlocks + 1, nb), dim3(LU_THREADS), 0, s, B, p,
                       nColBlocks);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_lu_backsub_b, dim3(1, nb), dim3(BS_THREADS), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_refine(hipStream_t s, const LuSet& L, double* d_out) {
  return enqueue_refine(s, L, d_out);
}

LuSys lu_sys(const LuSet& L, double* d_out, const int* d_skipIfZero) {
  LuSys S;
  S.M = L.M;
  S.vv = L.luVV;
  S.piv = L.luIndex;
  S.rhs = L.rhs;
  S.st = reinterpret_cast<LuState*>(L.luState);
  S.M0 = L.M0;
  S.rhs0 = L.rhs0;
  S.out = d_out;
  S.skip = d_skipIfZero;
  S.F = L.F;
  return S;
}

// Two systems' solves as one batched launch sequence on the current stream
// (a one-stream context's two cd trial fits side by side): the batched L D L^T
// launch, or the batched ludcmp + lubksb kernels (bit-identical to single
// solves) and each system's refinement; launched directly, not from a graph.
hipError_t launch_global_solve_pair(Ctx& c, const LuSet& L0, double* out0, const int* skip0,
                                    const LuSet& L1, double* out1, const int* skip1) {
  prof_begin(c, XB_PROF_SOLVE);
  LuBatch B;
  B.n = 2;
  B.s[0] = lu_sys(L0, out0, skip0);
  B.s[1] = lu_sys(L1, out1, skip1);
  const bool refine = (c.form & XB_FORM_REFINE) != 0;
  hipError_t e;
  if (c.form & XB_FORM_LDLT) {
    e = launch_ldlt_solve_batch(c.cur, B, refine, c.shapes.ldlt_group);
  } else {
    e = launch_global_solve_batch(c.cur, B, false, false);
    if (e == hipSuccess && refine) e = enqueue_refine(c.cur, L0, out0);
    if (e == hipSuccess && refine) e = enqueue_refine(c.cur, L1, out1);
  }
  prof_end(c, XB_PROF_SOLVE);
  return e;
}

// The global solve on the context's current stream: replayed from a hipGraph
// captured at the first call with these buffers (kernels only, no memset
// nodes), or launched kernel by kernel under XB_CTX_EAGER.
hipError_t launch_global_solve(Ctx& c, const LuSet& L, double* d_out, const int* d_skipIfZero) {
  prof_begin(c, XB_PROF_SOLVE);
  hipError_t e = hipSuccess;
  if (c.form & XB_FORM_LDLT) {
    // one launch that factors, solves (and refines): nothing to capture
    e = launch_ldlt_solve(c.cur, lu_sys(L, d_out, d_skipIfZero), (c.form & XB_FORM_REFINE) != 0,
                          c.shapes.ldlt_group);
    prof_end(c, XB_PROF_SOLVE);
    return e;
  }
  if (c.eagerSolve) {
    e = enqueue_solve(c.cur, L, d_out, d_skipIfZero);
  } else {
    hipGraphExec_t g = nullptr;
    for (const auto& sg : c.solveGraphs)
      if (sg.M == L.M && sg.out == d_out && sg.skip == d_skipIfZero) {
        g = sg.exec;
        break;
      }
    if (!g) {
      hipGraph_t graph = nullptr;
      {
        // no thread allocates or frees while the capture is open
        std::shared_lock<std::shared_mutex> lk(alloc_lock());
        e = hipStreamBeginCapture(c.cur, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
          const hipError_t e1 = enqueue_solve(c.cur, L, d_out, d_skipIfZero);
          const hipError_t e2 = hipStreamEndCapture(c.cur, &graph);
          e = e1 != hipSuccess ? e1 : e2;
        }
      }
      if (e == hipSuccess) {
        std::unique_lock<std::shared_mutex> lk(alloc_lock());
        e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
      }
      if (graph) (void)hipGraphDestroy(graph);
      if (e == hipSuccess) c.solveGraphs.push_back({L.M, d_out, d_skipIfZero, g});
      if (e != hipSuccess) {
        // The capture failed (for instance invalidated by another thread's
        // work on the device): nothing of it ran.  A capture still open on
        // the stream is ended, then this solve's kernels are launched
        // directly.  The failure is counted (xb_trace.captureFailures); after
        // two the context launches every solve directly.
        g = nullptr;
        if (++c.captureFailures >= 2) c.eagerSolve = true;
        (void)hipGetLastError();
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(c.cur, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
          hipGraph_t left = nullptr;
          (void)hipStreamEndCapture(c.cur, &left);
          if (left) (void)hipGraphDestroy(left);
          (void)hipGetLastError();
        }
        e = enqueue_solve(c.cur, L, d_out, d_skipIfZero);
      }
    }
    if (e == hipSuccess && g) e = hipGraphLaunch(g, c.cur);
  }
  if (e == hipSuccess && (c.form & XB_FORM_REFINE)) e = enqueue_refine(c.cur, L, d_out);
  prof_end(c, XB_PROF_SOLVE);
  return e;
}

}  // namespace xb
