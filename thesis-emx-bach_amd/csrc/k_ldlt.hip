// k_ldlt.hip — the global FP64 solve of the fast forms (XB_FORM_LDLT).
//
// The fit's normal matrix (built by k_global_build, cl_kern/cd.cl:152-352) is
// symmetric and positive definite: a sum over stamps of Gram matrices of the
// model's derivatives with positive weights.  The plain form solves it with
// Numerical Recipes' ludcmp + lubksb in their exact order (k_lu.hip,
// src/bachUtil.cpp:319-411): partial pivoting makes 292 dependent pivot votes
// (on the fit's systems it swaps rows at ~240 of the 292 columns) and lubksb's
// back substitution is one ordered chain of ~42,500 subtractions.  A form that
// does not have to round as the reference does can factor the same matrix as
// L D L^T without pivoting, which is stable for a positive definite matrix and
// on the fit's systems as accurate as ludcmp (~1e-8 normwise), and refine the
// solution once in FP64 with a compensated residual (~1e-11).
//
// One workgroup (8 waves) solves one system in one launch:
//   copy-in  the lower triangle of M0 (1-based, 293 x 293) into F (0-based,
//            padded to 304 = 19 blocks of 16 with an identity tail); a row
//            without a nonzero entry sends the system down the reference's
//            path (below);
//   factor   19 block steps: the panel (block column k) in LDS, its diagonal
//            block's L D L^T by one wave, the panel rows below by one thread
//            each, then the trailing lower tiles on the FP64 matrix cores
//            (v_mfma_f64_16x16x4f64): A_IJ -= L_Ik W_Jk^T with W = L D;
//   solve    L y = b, z = D^-1 y, L^T x = z with one thread per row: per block
//            the diagonal block by 16 lanes of one wave (values passed by
//            shuffles), then one barrier, then every later row (earlier column)
//            takes the block's contribution; each thread loads the next block's
//            factors while the current block runs;
//   refine   (with XB_FORM_REFINE) r = b - M0 x with each product's rounding
//            error kept (fma) and the sums compensated, then L D L^T d = r and
//            x += d.
//
// A row of M0 with no nonzero finite entry is where ludcmp stops ("singular
// matrix": it returns before factorising, src/bachUtil.cpp:326-334) and
// lubksb then runs on the unfactorised matrix with an index array of zeros.
// That case runs the same literal lubksb here, so such systems give what the
// plain form gives, bit for bit.  A skipped solve (empty stamp list) zeroes the
// output as the plain form does.
#include <type_traits>

#include "xb_dev.h"

namespace xb {

constexpr int LT_B = 16;                            // block order (one MFMA tile)
constexpr int LT_NB = (MATSIZE + LT_B - 1) / LT_B;  // 19 blocks
constexpr int LT_NP = LT_NB * LT_B;                 // 304: padded order
constexpr int LT_THREADS = 512;                    // eight tile waves
constexpr int LT_WAVES = LT_THREADS / WAVE;
constexpr int LT_BLOCK = LT_THREADS + WAVE;         // and the panel wave
constexpr int LT_SPIN_MAX = 1 << 22;               // bound of the panel wave's wait
constexpr int LT_LDW = LT_B + 1;  // LDS pitch of a panel row (doubles)
// trailing tiles per load group: LT_GRP (2 or 4), a launch shape the batch
// autotuner picks (the same products in the same order either way)
constexpr int LT_TILES = LT_NB * (LT_NB + 1) / 2;  // 190 lower tiles
constexpr int LT_RK = (MATSIZE + 7) / 8;  // residual terms per lane of a row
constexpr double LT_TINY = 1.0e-20;  // an exact zero pivot, as ludcmp's TINY
static_assert(LT_NP <= LT_THREADS, "one thread per row in the substitutions");
static_assert(LT_TILES <= LT_THREADS, "one thread per tile entry");
static_assert(LT_NP * LT_NP + LT_TILES * LT_B * LT_B == LDLT_WORK_DOUBLES, "workspace size");

typedef double v4d __attribute__((ext_vector_type(4)));

struct LtShared {
  double P[2][LT_NP][LT_LDW];  // panels: w = l d (unscaled), this step's and the next
  double G[LT_NP][LT_LDW];     // this panel's -l (the trailing products' A operand)
  double d[LT_NP];             // D
  double rinv[LT_NP];          // 1 / D
  double y[LT_NP];             // block values published by the substitutions
  double zb[LT_NP];            // the bordered row: b, then z = D^-1 L^-1 b
  double x[LT_NP];             // the solution (0-based)
  double r[LT_NP];             // the refinement's residual
  int nz[LT_NP];               // row has a nonzero finite entry
  unsigned char tI[LT_TILES], tJ[LT_TILES];  // lower tiles, column by column
  int zero;
  int dflag;    // the next panel's diagonal block is in place (its step number)
  int spinFail;  // the panel wave's wait ran out (never expected)
};

__device__ __forceinline__ double lt_readlane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

// A 32-bit DPP move of each half of a double (CTRL: a DPP control code).
template <int CTRL>
__device__ __forceinline__ double lt_dpp(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void lt_two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void lt_dd_add(double& hi, double& lo, double h2, double l2) {
  double s, e;
  lt_two_sum(hi, h2, s, e);
  e = e + (lo + l2);
  hi = s + e;
  lo = e - (hi - s);
}

// One step of a 16-lane diagonal-block substitution: v += v[lane K of the
// lane's row of 16] * m (64-bit DPP broadcast).  The s_nop covers the DPP read
// of v after the previous step's write.
#define LT_BCAST_FMA(K, M)                                                           \
  __asm__ __volatile__("s_nop 1\n\tv_fmac_f64_dpp %0, %0, %1 row_newbcast:" #K         \
                       " row_mask:0xf bank_mask:0xf"                                 \
                       : "+v"(v)                                                     \
                       : "v"(M));

// A barrier for LDS traffic only: outstanding global loads stay in flight.
__device__ __forceinline__ void lt_lds_barrier() {
  __asm__ __volatile__("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// x = (L D L^T)^-1 b for the factors in F and sh.rinv.  Thread r (< LT_NP)
// passes b_r and receives x_r; the other threads pass and receive 0.  Every
// thread of the workgroup calls it (it holds barriers).  A block's 16 rows are
// 16 lanes of one wave, one DPP row, so the block's own triangle is solved by a
// chain of broadcast multiply-adds without memory traffic.  Each thread keeps
// the factors of the next two blocks in flight (a ring of three), and the
// barriers wait for LDS traffic only.
template <bool FORWARD>
__device__ __forceinline__ double lt_substitute(const double* __restrict__ F, LtShared& sh,
                                                double v) {
  const int r = threadIdx.x;
  const bool act = r < LT_NP;
  const int kr = r / LT_B, ir = r % LT_B;
  double R[3][LT_B];
  if constexpr (FORWARD) {
  // forward: L y = b.  Row r needs L[r][16k .. 16k + 15] for blocks k <= kr,
  // read from the transposed copy above the diagonal (F[16k + j][r]), so that
  // neighbouring rows read neighbouring addresses.
  // (Every load is unconditional from an address inside F; rows that do not
  // need a block never read its values.  A guarded load would become a branch
  // whose join waits for the load.)
  // A wave loads a block only when one of its rows needs it (a wave-uniform
  // test: its last row is at or below the block).
  const int wLast = min((r | (WAVE - 1)), LT_NP - 1);  // the wave's last row
  auto loadRow = [&](int k, double (&o)[LT_B]) __attribute__((always_inline)) {
    if (k < LT_NB && wLast >= LT_B * k) {
      const double* p = F + (size_t)(LT_B * k) * LT_NP + min(r, LT_NP - 1);
#pragma unroll
      for (int j = 0; j < LT_B; j++) o[j] = p[(size_t)j * LT_NP];
    }
  };
  auto fwd = [&](auto slot, int k) __attribute__((always_inline)) {
    constexpr int SL = decltype(slot)::value;
    loadRow(k + 2, R[(SL + 2) % 3]);
    const double(&Lk)[LT_B] = R[SL];
    if (act && kr == k) {
      double m[LT_B];
#pragma unroll
      for (int j = 0; j < LT_B; j++) m[j] = ir > j ? -Lk[j] : 0.0;
      LT_BCAST_FMA(0, m[0]) LT_BCAST_FMA(1, m[1]) LT_BCAST_FMA(2, m[2])
      LT_BCAST_FMA(3, m[3]) LT_BCAST_FMA(4, m[4]) LT_BCAST_FMA(5, m[5])
      LT_BCAST_FMA(6, m[6]) LT_BCAST_FMA(7, m[7]) LT_BCAST_FMA(8, m[8])
      LT_BCAST_FMA(9, m[9]) LT_BCAST_FMA(10, m[10]) LT_BCAST_FMA(11, m[11])
      LT_BCAST_FMA(12, m[12]) LT_BCAST_FMA(13, m[13]) LT_BCAST_FMA(14, m[14])
      sh.y[r] = v;
    }
    lt_lds_barrier();
    if (act && kr > k) {
      const double* yb = sh.y + LT_B * k;
      double v2 = 0.0;
#pragma unroll
      for (int j = 0; j < LT_B; j += 2) {
        v = fma(-Lk[j], yb[j], v);
        v2 = fma(-Lk[j + 1], yb[j + 1], v2);
      }
      v = v + v2;
    }
  };
  loadRow(0, R[0]);
  loadRow(1, R[1]);
  for (int k = 0; k < LT_NB; k += 3) {
    fwd(std::integral_constant<int, 0>(), k);
    if (k + 1 < LT_NB) fwd(std::integral_constant<int, 1>(), k + 1);
    if (k + 2 < LT_NB) fwd(std::integral_constant<int, 2>(), k + 2);
  }
  if (act) v = v * sh.rinv[r];
  lt_lds_barrier();  // the forward pass's reads of sh.y are done
  }
  // backward: L^T x = z.  Row r needs L[16k .. 16k + 15][r] for blocks k >= kr
  // (below the diagonal, row-major).
  const int wFirst = r & ~(WAVE - 1);  // the wave's first row
  auto loadCol = [&](int k, double (&o)[LT_B]) __attribute__((always_inline)) {
    if (k >= 0 && wFirst <= LT_B * k + LT_B - 1 && wFirst < LT_NP) {
      const double* p = F + (size_t)(LT_B * k) * LT_NP + min(r, LT_NP - 1);
#pragma unroll
      for (int j = 0; j < LT_B; j++) o[j] = p[(size_t)j * LT_NP];
    }
  };
  auto bwd = [&](auto slot, int k) __attribute__((always_inline)) {
    constexpr int SL = decltype(slot)::value;
    loadCol(k - 2, R[(SL + 2) % 3]);
    const double(&Lk)[LT_B] = R[SL];
    if (act && kr == k) {
      double m[LT_B];
#pragma unroll
      for (int j = 0; j < LT_B; j++) m[j] = ir < j ? -Lk[j] : 0.0;
      LT_BCAST_FMA(15, m[15]) LT_BCAST_FMA(14, m[14]) LT_BCAST_FMA(13, m[13])
      LT_BCAST_FMA(12, m[12]) LT_BCAST_FMA(11, m[11]) LT_BCAST_FMA(10, m[10])
      LT_BCAST_FMA(9, m[9]) LT_BCAST_FMA(8, m[8]) LT_BCAST_FMA(7, m[7])
      LT_BCAST_FMA(6, m[6]) LT_BCAST_FMA(5, m[5]) LT_BCAST_FMA(4, m[4])
      LT_BCAST_FMA(3, m[3]) LT_BCAST_FMA(2, m[2]) LT_BCAST_FMA(1, m[1])
      sh.y[r] = v;
    }
    lt_lds_barrier();
    if (act && kr < k) {
      const double* xb = sh.y + LT_B * k;
      double v2 = 0.0;
#pragma unroll
      for (int j = 0; j < LT_B; j += 2) {
        v = fma(-Lk[j], xb[j], v);
        v2 = fma(-Lk[j + 1], xb[j + 1], v2);
      }
      v = v + v2;
    }
  };
  loadCol(LT_NB - 1, R[0]);
  loadCol(LT_NB - 2, R[1]);
  for (int k = LT_NB - 1; k >= 0; k -= 3) {
    bwd(std::integral_constant<int, 0>(), k);
    if (k - 1 >= 0) bwd(std::integral_constant<int, 1>(), k - 1);
    if (k - 2 >= 0) bwd(std::integral_constant<int, 2>(), k - 2);
  }
  lt_lds_barrier();  // sh.y free again
  return act ? v : 0.0;
}
#undef LT_BCAST_FMA

// The reference's path for a system with a zero row: lubksb on the
// unfactorised matrix with a zero index array (ludcmp returned first).  One
// thread, in lubksb's order (src/bachUtil.cpp:384-411).
__device__ void lt_zero_row_path(const double* __restrict__ M0, double* xs) {
  const int n = MATSIZE, N = MATN;
  int ii = 0;
  for (int i = 1; i <= n; i++) {
    const int ip = 0;
    double sum = xs[ip];
    xs[ip] = xs[i];
    if (ii) {
      for (int j = ii; j <= i - 1; j++) sum -= M0[(size_t)i * N + j] * xs[j];
    } else if (sum != 0.0) {
      ii = i;
    }
    xs[i] = sum;
  }
  for (int i = n; i >= 1; i--) {
    double sum = xs[i];
    for (int j = i + 1; j <= n; j++) sum -= M0[(size_t)i * N + j] * xs[j];
    xs[i] = sum / M0[(size_t)i * N + i];
  }
}

// Lane K's x within each row of 16 lanes (64-bit DPP broadcast).  The s_nop
// covers a DPP read of a register the previous instruction wrote.
template <int K>
__device__ __forceinline__ double lt_bcast(double x) {
  double y;
  __asm__ __volatile__(
      "s_nop 1\n\tv_mov_b64_dpp %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf"
      : "=v"(y)
      : "v"(x), "i"(K));
  return y;
}

// 1 / d: the hardware reciprocal and two Newton steps
__device__ __forceinline__ double lt_recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return r;
}

// Column J of the diagonal block's L D L^T, in registers of wave 0 (lane i
// holds row i & 15, its entries right of the diagonal taken from the lower
// triangle by symmetry and kept up to date as the matrix stays symmetric).
// d_J and 1/d_J are in dj, ri.  Row J's entries right of J (by symmetry column
// J) reach every lane from lane J; the next column's entry is updated first
// and its pivot and reciprocal started, then the rest of the row.
template <int J>
__device__ __forceinline__ void lt_diag_col(double (&a)[LT_B], double& dj, double& ri,
                                            double* sd, double* sr, int lane) {
  if (lane == 0) {
    sd[J] = dj;
    sr[J] = ri;
  }
  const double l = a[J] * ri;
  if constexpr (J + 1 < LT_B) {
    a[J + 1] = fma(-l, lt_bcast<J>(a[J + 1]), a[J + 1]);
    double dn = lt_bcast<J + 1>(a[J + 1]);
    if (dn == 0.0) dn = LT_TINY;
    const double rn = lt_recip(dn);
#pragma unroll
    for (int q = J + 2; q < LT_B; q++) a[q] = fma(-l, lt_bcast<J>(a[q]), a[q]);
    dj = dn;
    ri = rn;
  }
}

// The diagonal block of the panel (its rows 0..15) as L D L^T by wave 0.
// Leaves W (unscaled) below the diagonal in the panel and D, 1/D in sd, sr.
__device__ __forceinline__ void lt_diag_block(double (*P)[LT_LDW], double* sd, double* sr,
                                              int lane) {
  const int i = lane & (LT_B - 1);
  double a[LT_B];
#pragma unroll
  for (int q = 0; q < LT_B; q++) a[q] = q <= i ? P[i][q] : P[q][i];
  double dj = lt_bcast<0>(a[0]);
  if (dj == 0.0) dj = LT_TINY;
  double ri = lt_recip(dj);
  lt_diag_col<0>(a, dj, ri, sd, sr, lane);
  lt_diag_col<1>(a, dj, ri, sd, sr, lane);
  lt_diag_col<2>(a, dj, ri, sd, sr, lane);
  lt_diag_col<3>(a, dj, ri, sd, sr, lane);
  lt_diag_col<4>(a, dj, ri, sd, sr, lane);
  lt_diag_col<5>(a, dj, ri, sd, sr, lane);
  lt_diag_col<6>(a, dj, ri, sd, sr, lane);
  lt_diag_col<7>(a, dj, ri, sd, sr, lane);
  lt_diag_col<8>(a, dj, ri, sd, sr, lane);
  lt_diag_col<9>(a, dj, ri, sd, sr, lane);
  lt_diag_col<10>(a, dj, ri, sd, sr, lane);
  lt_diag_col<11>(a, dj, ri, sd, sr, lane);
  lt_diag_col<12>(a, dj, ri, sd, sr, lane);
  lt_diag_col<13>(a, dj, ri, sd, sr, lane);
  lt_diag_col<14>(a, dj, ri, sd, sr, lane);
  lt_diag_col<15>(a, dj, ri, sd, sr, lane);
  if (lane < LT_B) {
#pragma unroll
    for (int q = 0; q < LT_B; q++)
      if (q < i) P[i][q] = a[q];
  }
}

template <int LT_GRP>
__device__ __forceinline__ void ldlt_solve_dev(const LuSys& S, int refine) {
  extern __shared__ double lt_dyn[];
  LtShared& sh = *reinterpret_cast<LtShared*>(lt_dyn);
  const int tid = threadIdx.x, lane = lane_id(), wid = tid / WAVE;
  const double* __restrict__ M0 = S.M0;
  const double* __restrict__ b0 = S.rhs0;
  double* __restrict__ F = S.F;
  double* __restrict__ Ft = S.F + (size_t)LT_NP * LT_NP;  // the tile store
  if (S.skip && *S.skip == 0) {
    // testFit returns 666 before solving (src/cdkscUtil.cpp:101-103)
    for (int i = tid; i < NSOL; i += LT_THREADS)
      if (S.out) S.out[i] = 0.0;
    return;
  }
  // ---- copy-in, the zero-row check, the tile list -------------------------
  for (int u = tid; u < LT_NP; u += LT_THREADS) {
    sh.nz[u] = 0;
    sh.zb[u] = u < MATSIZE ? b0[u + 1] : 0.0;
  }
  if (tid == 0) {
    sh.zero = 0;
    sh.dflag = -1;
    sh.spinFail = 0;
  }
  if (tid < LT_TILES) {
    int J = 0, rem = tid;
    while (rem >= LT_NB - J) {
      rem -= LT_NB - J;
      J++;
    }
    sh.tI[tid] = (unsigned char)(J + rem);
    sh.tJ[tid] = (unsigned char)J;
  }
  __syncthreads();
  {
    // Every lower tile into the tile store Ft (each lane's accumulator values
    // v = 2h, 2h + 1 at Ft[((2 t + h) 64 + lane) 2]: a tile is two contiguous
    // 1 KB halves), block column 0 also into the first panel.  A wave takes one
    // tile per round: lane l gets rows l / 16 + 4 v, column l % 16 of it.
    const int lr = lane >> 4, lc = lane & 15;
    for (int t = wid; t < LT_TILES && wid < LT_WAVES; t += LT_WAVES) {
      const int I = sh.tI[t], J = sh.tJ[t];
      double val[4];
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = LT_B * I + lr + 4 * v, col = LT_B * J + lc;
        const double m =
            M0[(size_t)(min(row, MATSIZE - 1) + 1) * MATN + min(col, MATSIZE - 1) + 1];
        val[v] = (row < MATSIZE && col < MATSIZE) ? m : (row == col ? 1.0 : 0.0);
      }
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int row = LT_B * I + lr + 4 * v, col = LT_B * J + lc;
        if (row < MATSIZE && col < MATSIZE && fabs(val[v]) > 0.0) {
          sh.nz[row] = 1;  // NaN is not taken, as ludcmp's row maximum
          sh.nz[col] = 1;
        }
        if (J == 0) sh.P[0][row][lc] = val[v];
      }
      double2* d = reinterpret_cast<double2*>(Ft) + (size_t)2 * t * WAVE + lane;
      d[0] = make_double2(val[0], val[1]);
      d[WAVE] = make_double2(val[2], val[3]);
    }
  }
  __syncthreads();
  if (tid < MATSIZE && !sh.nz[tid]) sh.zero = 1;
  __syncthreads();
  if (sh.zero) {
    double* xs = sh.x;  // NSOL entries
    for (int i = tid; i < NSOL; i += LT_THREADS) xs[i] = b0[i];
    __syncthreads();
    if (tid == 0) lt_zero_row_path(M0, xs);
    __syncthreads();
    for (int i = tid; i < NSOL; i += LT_THREADS) {
      S.rhs[i] = xs[i];
      if (S.out) S.out[i] = xs[i];
    }
    return;
  }
  // ---- factorisation --------------------------------------------------------
  // The panel wave factors the diagonal block of panel 0 here and that of
  // panel k + 1 during step k, as soon as the owner of tile (k + 1, k + 1) has
  // left it in the next panel (an LDS flag), beside the trailing products.
  if (wid == LT_WAVES) lt_diag_block(sh.P[0], sh.d, sh.rinv, lane);
  lt_lds_barrier();
  // Tile t of the column-by-column list of lower tiles belongs to wave t % 8 in
  // every step, so a wave reads back only tiles it wrote itself and the steps
  // need no global-memory fence; panels pass through LDS.
  for (int k = 0; k < LT_NB; k++) {
    const int c0 = LT_B * k;
    const int R = LT_NP - c0;  // panel rows c0 .. LT_NP - 1
    double (*Pc)[LT_LDW] = sh.P[k & 1];
    double (*Pn)[LT_LDW] = sh.P[(k + 1) & 1];
    if (tid >= LT_B && tid <= R) {
      // panel row i below the diagonal block, or (i = R) the bordered row, the
      // right-hand side as an extra row of the matrix: its L part is
      // z = D^-1 L^-1 b, so the forward substitution is done here.  In column
      // order: w, and -l into G (z for the bordered row).
      const int i = tid;
      const bool brow = i == R;
      double a[LT_B];
#pragma unroll
      for (int q = 0; q < LT_B; q++) a[q] = brow ? sh.zb[c0 + q] : Pc[i][q];
      double lv[LT_B];
#pragma unroll
      for (int j = 0; j < LT_B; j++) {
        const double l = a[j] * sh.rinv[c0 + j];
        lv[j] = l;
#pragma unroll
        for (int m = j + 1; m < LT_B; m++) a[m] = fma(-l, Pc[m][j], a[m]);
      }
      if (brow) {
#pragma unroll
        for (int q = 0; q < LT_B; q++) sh.zb[c0 + q] = lv[q];
      } else {
#pragma unroll
        for (int q = 0; q < LT_B; q++) {
          Pc[i][q] = a[q];
          sh.G[i][q] = -lv[q];
        }
      }
    }
    lt_lds_barrier();
    // the panel's L (and D on the diagonal) into F by all threads, neighbouring
    // threads on neighbouring addresses: row-major below the diagonal (the
    // backward pass), and transposed above it (the forward pass); the rows
    // below the diagonal block in 16-byte pairs
    for (int e = tid; e < LT_B * LT_B && wid < LT_WAVES; e += LT_THREADS) {
      const int i = e / LT_B, q = e % LT_B;
      if (i > q) {
        const double l = Pc[i][q] * sh.rinv[c0 + q];
        F[(size_t)(c0 + i) * LT_NP + c0 + q] = l;
        F[(size_t)(c0 + q) * LT_NP + c0 + i] = l;
      } else if (i == q) {
        F[(size_t)(c0 + i) * LT_NP + c0 + q] = sh.d[c0 + q];
      }
    }
    const int RB = R - LT_B;  // rows below the diagonal block (even)
    for (int e = tid; e < RB * (LT_B / 2) && wid < LT_WAVES; e += LT_THREADS) {
      const int i = LT_B + e / (LT_B / 2), q = 2 * (e % (LT_B / 2));
      *reinterpret_cast<double2*>(F + (size_t)(c0 + i) * LT_NP + c0 + q) =
          make_double2(-sh.G[i][q], -sh.G[i][q + 1]);
    }
    for (int e = tid; e < LT_B * (RB / 2) && wid < LT_WAVES; e += LT_THREADS) {
      const int q = e / (RB / 2), i = LT_B + 2 * (e % (RB / 2));
      *reinterpret_cast<double2*>(F + (size_t)(c0 + q) * LT_NP + c0 + i) =
          make_double2(-sh.G[i][q], -sh.G[i + 1][q]);
    }
    // the bordered row right of the panel: zb_c -= sum_j z_j W_cj
    for (int c = c0 + LT_B + tid; c < LT_NP; c += LT_THREADS) {
      double acc = sh.zb[c];
#pragma unroll
      for (int j = 0; j < LT_B; j++) acc = fma(-sh.zb[c0 + j], Pc[c - c0][j], acc);
      sh.zb[c] = acc;
    }
    // trailing lower tiles (I, J), k < J <= I: A_IJ -= L_I W_J^T on the matrix
    // cores.  MFMA operands: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]
    // of a 16 x 4 by 4 x 16 product; the accumulator holds element
    // (l / 16 + 4 v, l % 16) in value v.  LT_GRP tiles at a time, the next
    // group's loads in flight while the current group's products run; the
    // tiles of block column k + 1 go to the next panel in LDS, the others back
    // to F.
    const int tStart = LT_NB * (k + 1) - k * (k + 1) / 2;  // first tile with J > k
    const int t0 = tStart + (((wid - tStart) % LT_WAVES) + LT_WAVES) % LT_WAVES;
    const int ns = (t0 < LT_TILES && wid < LT_WAVES) ? (LT_TILES - t0 + LT_WAVES - 1) / LT_WAVES : 0;
    const int lr = lane >> 4, lc = lane & 15;
    // the lane's two halves of tile s of this wave's list in the tile store
    auto tileLoad = [&](int s) __attribute__((always_inline)) -> v4d {
      const int t = min(t0 + LT_WAVES * s, LT_TILES - 1);
      const double2* d = reinterpret_cast<const double2*>(Ft) + (size_t)2 * t * WAVE + lane;
      const double2 a = d[0], b = d[WAVE];
      return v4d{a.x, a.y, b.x, b.y};
    };
    v4d cur[LT_GRP], nxt[LT_GRP];
#pragma unroll
    for (int g = 0; g < LT_GRP; g++) cur[g] = tileLoad(g);
    for (int s0 = 0; s0 < ns; s0 += LT_GRP) {
#pragma unroll
      for (int g = 0; g < LT_GRP; g++) nxt[g] = tileLoad(s0 + LT_GRP + g);
      int pi[LT_GRP], pj[LT_GRP];
#pragma unroll
      for (int g = 0; g < LT_GRP; g++) {
        const int t = min(t0 + LT_WAVES * (s0 + g), LT_TILES - 1);
        pi[g] = LT_B * sh.tI[t] - c0 + lc;
        pj[g] = LT_B * sh.tJ[t] - c0 + lc;
      }
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int kk = 4 * c + lr;
#pragma unroll
        for (int g = 0; g < LT_GRP; g++)
          cur[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(sh.G[pi[g]][kk], Pc[pj[g]][kk], cur[g],
                                                       0, 0, 0);
      }
#pragma unroll
      for (int g = 0; g < LT_GRP; g++) {
        if (s0 + g < ns) {
          const int t = t0 + LT_WAVES * (s0 + g);
          if (sh.tJ[t] == k + 1) {
            const int pr = LT_B * (sh.tI[t] - k - 1) + lr;
#pragma unroll
            for (int v = 0; v < 4; v++) Pn[pr + 4 * v][lc] = cur[g][v];
            if (sh.tI[t] == k + 1)
              __hip_atomic_store(&sh.dflag, k + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
          } else {
            double2* d = reinterpret_cast<double2*>(Ft) + (size_t)2 * t * WAVE + lane;
            d[0] = make_double2(cur[g][0], cur[g][1]);
            d[WAVE] = make_double2(cur[g][2], cur[g][3]);
          }
        }
      }
#pragma unroll
      for (int g = 0; g < LT_GRP; g++) cur[g] = nxt[g];
    }
    if (wid == LT_WAVES && k + 1 < LT_NB) {
      int it = 0;
      while (__hip_atomic_load(&sh.dflag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != k + 1 &&
             ++it < LT_SPIN_MAX)
        __builtin_amdgcn_s_sleep(1);
      if (it >= LT_SPIN_MAX && lane == 0) sh.spinFail = 1;
      lt_diag_block(Pn, sh.d + c0 + LT_B, sh.rinv + c0 + LT_B, lane);
    }
    lt_lds_barrier();
  }
  __syncthreads();  // L and D in F are read by other waves from here on
  // ---- solve and one refinement step --------------------------------------
  const int r = tid;
  // the forward substitution and D^-1 were done with the bordered row
  double x = lt_substitute<false>(F, sh, r < LT_NP ? sh.zb[r] : 0.0);
  if (r < LT_NP) sh.x[r] = x;
  __syncthreads();
  if (refine) {
  // r = b - M0 x: 8 lanes per row (8 rows per wave at a time), compensated
  // (Dot2); the 8 partial sums of a row meet by DPP exchanges within the lane
  // octet in double-double
  {
    const int sub = lane & 7;
    for (int u0 = wid * 8; u0 < MATSIZE && wid < LT_WAVES; u0 += LT_WAVES * 8) {
      const int u = u0 + (lane >> 3);
      double hi = 0.0, lo = 0.0;
      if (u < MATSIZE) {
        const double* row = M0 + (size_t)(u + 1) * MATN + 1;
        double av[LT_RK];
#pragma unroll
        for (int i = 0; i < LT_RK; i++) av[i] = row[min(sub + 8 * i, MATSIZE - 1)];
        // Dot2 (Ogita, Rump and Oishi, 2005): each product split exactly by
        // an fma, each addition's error taken exactly by TwoSum, the errors
        // summed apart
        double sum = sub == 0 ? b0[u + 1] : 0.0, err = 0.0;
#pragma unroll
        for (int i = 0; i < LT_RK; i++) {
          const int v = sub + 8 * i;
          if (v < MATSIZE) {
            const double a = av[i], xv = sh.x[v];
            const double pr = a * xv;
            const double pe = fma(a, xv, -pr);  // a xv = pr + pe exactly
            double t, te;
            lt_two_sum(sum, -pr, t, te);
            sum = t;
            err = err + (te - pe);
          }
        }
        hi = sum;
        lo = err;
      }
      lt_dd_add(hi, lo, lt_dpp<0xB1>(hi), lt_dpp<0xB1>(lo));    // quad_perm 1,0,3,2
      lt_dd_add(hi, lo, lt_dpp<0x4E>(hi), lt_dpp<0x4E>(lo));    // quad_perm 2,3,0,1
      lt_dd_add(hi, lo, lt_dpp<0x141>(hi), lt_dpp<0x141>(lo));  // row_half_mirror
      if (sub == 0 && u < MATSIZE) sh.r[u] = hi + lo;
    }
  }
  __syncthreads();
  const double dx = lt_substitute<true>(F, sh, r < MATSIZE ? sh.r[r] : 0.0);
  x = x + dx;
  if (r < MATSIZE) sh.x[r] = x;
  __syncthreads();
  }
  // 1-based output, the other slots as the system's rhs holds them (as lubksb
  // leaves them)
  for (int i = tid; i < NSOL; i += LT_THREADS) {
    double val = (i >= 1 && i <= MATSIZE) ? sh.x[i - 1] : b0[i];
    if (sh.spinFail && i >= 1 && i <= MATSIZE) val = __builtin_nan("");
    S.rhs[i] = val;
    if (S.out) S.out[i] = val;
  }
}

template <int GRP>
__global__ __launch_bounds__(LT_BLOCK) void k_ldlt_solve(LuSys S, int refine) {
  ldlt_solve_dev<GRP>(S, refine);
}

template <int GRP>
__global__ __launch_bounds__(LT_BLOCK) void k_ldlt_solve_b(LuBatch B, int refine) {
  ldlt_solve_dev<GRP>(B.s[blockIdx.x], refine);
}

template <int GRP>
static hipError_t lt_attr() {
  static const hipError_t rc1 =
      hipFuncSetAttribute((const void*)k_ldlt_solve<GRP>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(LtShared));
  static const hipError_t rc2 =
      hipFuncSetAttribute((const void*)k_ldlt_solve_b<GRP>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(LtShared));
  return rc1 != hipSuccess ? rc1 : rc2;
}

template <int GRP>
static hipError_t launch_ldlt_grp(hipStream_t s, const LuSys& S, bool refine) {
  const hipError_t e = lt_attr<GRP>();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_ldlt_solve<GRP>, dim3(1), dim3(LT_BLOCK), sizeof(LtShared), s, S,
                     refine ? 1 : 0);
  return hipGetLastError();
}

template <int GRP>
static hipError_t launch_ldlt_batch_grp(hipStream_t s, const LuBatch& B, bool refine) {
  const hipError_t e = lt_attr<GRP>();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_ldlt_solve_b<GRP>, dim3(B.n), dim3(LT_BLOCK), sizeof(LtShared), s, B,
                     refine ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_ldlt_solve(hipStream_t s, const LuSys& S, bool refine, int grp) {
  return grp == 2 ? launch_ldlt_grp<2>(s, S, refine) : launch_ldlt_grp<4>(s, S, refine);
}

hipError_t launch_ldlt_solve_batch(hipStream_t s, const LuBatch& B, bool refine, int grp) {
  if (B.n <= 0) return hipSuccess;
  if (B.n > LU_BATCH_MAX) return hipErrorInvalidValue;
  return grp == 2 ? launch_ldlt_batch_grp<2>(s, B, refine)
                  : launch_ldlt_batch_grp<4>(s, B, refine);
}

}  // namespace xb
