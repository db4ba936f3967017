// k_fit.hip — kernel fit: per-stamp test solves, global normal equations,
// FP64 LU, chi2 merits and the clip-refit checks (the reference's cd + ksc).
//
// The global solve lives in k_lu.hip.
//
// LU note.  The reference factorises with Numerical Recipes' Crout ludcmp
// (implicit scaling, partial pivoting; src/bachUtil.cpp:319-384 on the host,
// cl_kern/bach.cl:68-154 per stamp).  Crout forms element (i,l) as
// a_il - sum_k L_ik U_kl with k ascending; right-looking elimination applies
// exactly the same updates to each element in the same ascending-k order and
// the same pivot row swaps, so the factors are bit-identical while every step
// is a parallel rank-1 update.  Pivot choice keeps the reference's `>=` rule
// (the last row attaining the maximum wins; NaN never wins).
#include <algorithm>

#include "xb_dev.h"

namespace xb {

// ---------------------------------------------------------------------------
// per-stamp 50x50 test system (cdkscUtil.cpp:42-72, bach.cl:68-194)
// ---------------------------------------------------------------------------
// Wave-wide (value, row) maximum, ties to the larger row, wave-uniform
// result: DPP rotations within rows of 16 lanes, then the row broadcasts of
// lanes 15 and 31 into the last row (lane 63).  A maximum under a total
// order: any combining order gives the same row.
template <int CTRL, int RM>
__device__ __forceinline__ void argmax_last_step(double& v, int& idx) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, RM, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, RM, 0xF, false);
  const int oi = __builtin_amdgcn_update_dpp(idx, idx, CTRL, RM, 0xF, false);
  const double ov = __hiloint2double(hi, lo);
  if (ov > v || (ov == v && oi > idx)) {
    v = ov;
    idx = oi;
  }
}
__device__ __forceinline__ int stamp_argmax_last(double v, int idx) {
  argmax_last_step<0x121, 0xF>(v, idx);
  argmax_last_step<0x122, 0xF>(v, idx);
  argmax_last_step<0x124, 0xF>(v, idx);
  argmax_last_step<0x128, 0xF>(v, idx);
  argmax_last_step<0x142, 0xA>(v, idx);
  argmax_last_step<0x143, 0xC>(v, idx);
  return __builtin_amdgcn_readlane(idx, 63);
}

__device__ __forceinline__ double readlane_f64_at(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

struct StampLuArgs {
  const double* Gall;
  const double* Bcall;
  const int* nKept;
  double* ksum;
};

template <int NI>
__global__ __launch_bounds__(64) void k_stamp_lu(ArgsN<StampLuArgs, NI> P) {
  const StampLuArgs& SA = args_of(P);
  const double* __restrict__ Gall = SA.Gall;
  const double* __restrict__ Bcall = SA.Bcall;
  const int* nKept = SA.nKept;
  double* ksum = SA.ksum;
  const int s = blockIdx.x;
  if (s >= *nKept) return;
  const int tid = threadIdx.x;
  __shared__ double A[QN][QN + 1];
  __shared__ double b[QN], vv[QN];
  __shared__ int index[QN];
  const double* G = Gall + (size_t)s * WROWS * WROWS;
  // createTestMat / createTestVec: A[i][j] = Q[max][min], Q[a][b] = G[a-1][b-1]
  // (every load of a lane issued before its stores: one memory round trip)
  {
    constexpr int PER = (QN * QN + WAVE - 1) / WAVE;
    double v[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = min(tid + WAVE * u, QN * QN - 1);
      const int i = e / QN, j = e % QN;
      const int a = max(max(i, j), 1), bb = max(min(i, j), 1);
      v[u] = G[(a - 1) * WROWS + (bb - 1)];
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = tid + WAVE * u;
      if (e < QN * QN) {
        const int i = e / QN, j = e % QN;
        A[i][j] = (i > 0 && j > 0) ? v[u] : 0.0;
      }
    }
  }
  for (int i = tid; i < QN; i += blockDim.x) b[i] = i > 0 ? Bcall[(size_t)s * WROWS + i - 1] : 0.0;
  __syncthreads();
  // ludcmpBig: vv[i] = 1/max_j |a_ij| (no zero-row exit on the device)
  for (int i = tid; i < QN; i += blockDim.x) {
    double v = 0.0;
    if (i > 0) {
      double big = 0.0;
      for (int j = 1; j < QN; j++) {
        double m = fabs(A[i][j]);
        if (m > big) big = m;
      }
      v = 1.0 / big;
    }
    vv[i] = v;
  }
  __syncthreads();
  const int n = QN - 1;
  // One wave: its LDS accesses complete in issue order, so between the steps
  // only the compiler needs fencing.
  auto fence = [] {
    __builtin_amdgcn_wave_barrier();
    __asm__ __volatile__("" ::: "memory");
  };
  // ludcmp in registers: lane l holds row l (columns 1..n) and its vv.  A
  // row interchange relabels the two rows (position numbers) instead of
  // moving them, vv staying with its row (NR's vv[imax] = vv[j] only touches
  // positions that are never searched again); the column update is masked to
  // the rows still below the pivot, so a pivoted row keeps its values and
  // even a swap with an already factorised row (no candidate, the previous
  // pivot kept) is the reference's.  Per column: the DPP vote over the
  // labels >= j (ties to the larger, NaN never), the pivot row's columns
  // from its lane by v_readlane, then scaling and the rank-1 update with the
  // reference's rounding (multiply, then subtract).
  {
    static_assert(QN <= WAVE, "one row per lane");
    // lane 0 holds row 0 (zeros, never a candidate): the reference swaps
    // with it when the first column has no qualifying row (imax still 0)
    double x[QN];
#pragma unroll
    for (int c = 0; c < QN; c++) x[c] = tid <= n ? A[tid][c] : 0.0;
    const double vr = (tid >= 1 && tid <= n) ? vv[tid] : 0.0;
    int lab = tid <= n ? tid : -1;
    int maxI = 0;
#pragma unroll
    for (int j = 1; j <= n; j++) {
      double bv = -1.0;
      int bi = -1;
      if (lab >= j) {
        const double dum = vr * fabs(x[j]);
        if (dum >= 0.0) {
          bv = dum;
          bi = lab;
        }
      }
      const int best = stamp_argmax_last(bv, bi);
      if (best >= 0) maxI = best;
      // the lane holding position maxI becomes position j
      const int src = __builtin_ctzll(__ballot(lab == maxI));
      if (j != maxI) lab = lab == j ? maxI : (lab == maxI ? j : lab);
      if (tid == 0) index[j] = maxI;
      double piv = readlane_f64_at(x[j], src);
      if (piv == 0.0) {
        piv = 1.0e-20;
        if (tid == src) x[j] = 1.0e-20;
      }
      if (j != n) {
        const double dinv = 1.0 / piv;
        const bool below = lab > j;
        if (below) x[j] *= dinv;
#pragma unroll
        for (int l = j + 1; l <= n; l++) {
          const double u = readlane_f64_at(x[l], src);
          if (below) x[l] = x[l] - x[j] * u;
        }
      }
    }
    // rows to their final positions for the substitution
    fence();
    if (lab >= 1) {
#pragma unroll
      for (int c = 1; c <= n; c++) A[lab][c] = x[c];
    }
    fence();
  }
  // lubksb (bach.cl:156-194) by the wave: lane j holds b[j].  Each row's
  // products A[i][j] * b[j] are formed by their lanes and staged in LDS, then
  // subtracted in ascending j as one chain: fma(P, -1, s) is s - P rounded
  // once, with P broadcast by DPP from lane K of each 16-lane row (every row
  // of the wave holds the same block, so every lane runs the same chain).
  // Slots outside a row's range hold +0.0, which leaves the chain unchanged
  // (it never holds -0.0 there: s - (+0.0) is s).
  __shared__ double Pl[64];
  const double m1 = -1.0;
  double bj = tid <= n ? b[tid] : 0.0;
#define XB_SCHAIN(K)                                                                \
  __asm__ __volatile__("v_fmac_f64_dpp %0, %1, %2 row_newbcast:" #K                 \
                       " row_mask:0xf bank_mask:0xf"                                \
                       : "+v"(sv)                                                   \
                       : "v"(pk), "v"(m1));
  // sv - (sum of the staged products of blocks m0..m1b), in ascending order
  auto chain = [&](double sv, int m0, int m1b) -> double {
    for (int m = m0; m <= m1b; m++) {
      const double pk = Pl[16 * m + (tid & 15)];
      // two wait states before the first DPP read of pk (a copy into another
      // register would be a VALU write)
      __asm__ __volatile__("s_nop 1");
      XB_SCHAIN(0) XB_SCHAIN(1) XB_SCHAIN(2) XB_SCHAIN(3) XB_SCHAIN(4) XB_SCHAIN(5)
      XB_SCHAIN(6) XB_SCHAIN(7) XB_SCHAIN(8) XB_SCHAIN(9) XB_SCHAIN(10) XB_SCHAIN(11)
      XB_SCHAIN(12) XB_SCHAIN(13) XB_SCHAIN(14) XB_SCHAIN(15)
    }
    return sv;
  };
  const int jc = min(tid, n);  // a real column for every lane (masked below)
  int ii = 0;
  for (int i = 1; i <= n; i++) {
    const int ip = __builtin_amdgcn_readfirstlane(index[i]);
    double sum = readlane_f64_at(bj, ip);
    const double bi = readlane_f64_at(bj, i);
    if (tid == ip) bj = bi;  // b[ip] = b[i] (ip >= i: the b[j], j < i, stay)
    if (ii != 0) {
      Pl[tid] = (tid >= ii && tid <= i - 1) ? A[i][jc] * bj : 0.0;
      fence();
      sum = chain(sum, ii >> 4, (i - 1) >> 4);
      fence();
    } else if (sum != 0.0) {
      ii = i;
    }
    if (tid == i) bj = sum;
  }
  for (int i = n; i >= 1; i--) {
    double sum = readlane_f64_at(bj, i);
    Pl[tid] = (tid > i && tid <= n) ? A[i][jc] * bj : 0.0;
    fence();
    sum = chain(sum, (i + 1) >> 4, n >> 4);
    fence();
    const double x = sum / A[i][i];
    if (tid == i) bj = x;
  }
#undef XB_SCHAIN
  const double k1 = readlane_f64_at(bj, 1);
  if (tid == 0) ksum[s] = k1;
}

// ---------------------------------------------------------------------------
// cd: kernel-sum clip and test-stamp selection (cdkscUtil.cpp:74-103)
// ---------------------------------------------------------------------------
struct CdSelectArgs {
  const double* ksum;
  const int* nKeptP;
  int* list;
  DevScalars* sc;
  int im;
  double alpha, sigKernFit;
};

template <int NI>
__global__ __launch_bounds__(1024) void k_cd_select(ArgsN<CdSelectArgs, NI> P) {
  const CdSelectArgs& CA = args_of(P);
  const double* ksum = CA.ksum;
  const int* nKeptP = CA.nKeptP;
  int* list = CA.list;
  DevScalars* sc = CA.sc;
  const int im = CA.im;
  const double alpha = CA.alpha, sigKernFit = CA.sigKernFit;
  __shared__ double data[XB_MAX_STAMPS];
  __shared__ uint8_t clip[XB_MAX_STAMPS];
  __shared__ double part[64];
  __shared__ double res[2];
  __shared__ int isc[20];
  __shared__ int s_n;
  const int n = *nKeptP;
  for (int i = threadIdx.x; i < n; i += blockDim.x) data[i] = ksum[i];
  __syncthreads();
  sigma_clip_small(data, n, 10, alpha, clip, part, res, isc);
  const double mean = res[0], sd = res[1];
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  for (int base = 0; base < n; base += blockDim.x) {
    int s = base + threadIdx.x;
    bool keep = false;
    if (s < n) {
      double diff = fabs((data[s] - mean) / sd);
      keep = diff < sigKernFit;
    }
    int tot;
    int pos = block_prefix_count(keep, isc, &tot);
    int c0 = s_n;
    if (keep) list[c0 + pos] = s;
    __syncthreads();
    if (threadIdx.x == 0) s_n = c0 + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sc->nTest[im] = s_n;
    sc->nList[im] = s_n;
  }
}

// ---------------------------------------------------------------------------
// global normal matrix + right-hand side
//   cd form:  cl_kern/cd.cl:152-352 over the test stamps
//   ksc form: src/cdkscUtil.cpp:307-475 over all stamps with a substamp left
// The per-element sums (over stamps, in order) are the same expressions in
// both forms; only the background right-hand side order differs (row-major in
// the cd form, column-major in the ksc form).
// ---------------------------------------------------------------------------
struct BuildArgs {
  const double* G;
  const double* Bc;
  const double* Br;
  const int* kept;
  const int* cur;
  const int* count;
  const int* ssx;
  const int* ssy;
  const int* list;   // cd: test list; ksc: nullptr (all kept)
  const int* nList;  // count of the list / kept
  int w, h;
  int cdForm;
  double* M;
  double* rhs;
  double* M0;    // untouched copies for the LU's replay paths
  double* rhs0;
};

__device__ __forceinline__ void stamp_weights(int w, int h, int x, int y, double* wt) {
  // xf, yf in FP32 (cd.cl:170-171, cdkscUtil.cpp:317-332); monomials in the
  // order of the kernelXy list (bach.cpp:278-282)
  float hw = (float)(0.5 * w), hh = (float)(0.5 * h);
  double fx = (double)(((float)x - hw) / hw);
  double fy = (double)(((float)y - hh) / hh);
  double fx2 = fx * fx, fy2 = fy * fy;
  wt[0] = 1.0 * 1.0;
  wt[1] = 1.0 * fy;
  wt[2] = 1.0 * fy2;
  wt[3] = fx * 1.0;
  wt[4] = fx * fy;
  wt[5] = fx2 * 1.0;
}

constexpr int BUILD_STAMPS_LDS = 256;
constexpr int WT_ONE = NC2;     // slot of the staged weight row that holds 1.0
constexpr int WT_STRIDE = NC2 + 2;

// Where global element (r, c) (r >= c), or rhs entry rr (r < 0), takes its
// per-stamp term from. Every term of the reference's expressions has the form
// (wt[ia] * wt[ib]) * src[s * stride + off] with wt[WT_ONE] = 1.0:
//   wt[i2] * wt[j2] * G[..]  is  (wt[i2] * wt[j2]) * G[..]
//   wt[i2] * G[..], G[..] * wt[i2], Bc[i] * wt[j]  are  (wt[.] * 1.0) * g
//   G[..], Bc[..], Br[..]  are  (1.0 * 1.0) * g
// (x * 1.0 == x and products commute, so each term is bit-identical to the
// expression written out). Elements whose terms are all zero set `zero`.
struct TermPlan {
  const double* src;
  size_t stride;
  int off, ia, ib;
  bool zero;
};

__device__ __forceinline__ TermPlan plan_term(const BuildArgs& A, int r, int c, int rr) {
  TermPlan P{A.G, (size_t)WROWS * WROWS, 0, WT_ONE, WT_ONE, false};
  if (r >= 0) {
    if (r == 0 || c == 0) {
      P.zero = true;
    } else if (r == 1) {  // c == 1
      P.off = 0;
    } else if (r <= NCOMP + 1) {
      // basis rows 2..289
      const int i = r - 2, i1 = i / NC2, i2 = i - i1 * NC2;
      P.ia = i2;
      if (c == 1) {
        P.off = (i1 + 1) * WROWS;
      } else {
        const int j = c - 2, j1 = j / NC2, j2 = j - j1 * NC2;
        P.off = (i1 + 1) * WROWS + (j1 + 1);
        P.ib = j2;
      }
    } else {
      // background rows 290..292: W row iv = 49 + iBG
      const int iv = NC1 + (r - (NCOMP + 2)) + 1;
      if (c == 1) {
        P.off = iv * WROWS;
      } else if (c <= NCOMP + 1) {
        const int jj = c - 1;
        P.off = iv * WROWS + (jj - 1) / NC2 + 1;
        P.ia = (jj - 1) % NC2;
      } else {
        P.off = iv * WROWS + NC1 + (c - (NCOMP + 2)) + 1;
      }
    }
  } else if (rr >= 1 && rr < NCOMP + 2 + NBG) {
    P.src = A.Bc;
    P.stride = WROWS;
    if (rr == 1) {
      P.off = 0;
    } else if (rr <= NCOMP + 1) {
      P.off = (rr - 2) / NC2 + 1;
      P.ia = (rr - 2) % NC2;
    } else if (A.cdForm) {
      P.src = A.Br;
      P.stride = NBG;
      P.off = rr - (NCOMP + 2);
    } else {
      P.off = NPSF + rr - (NCOMP + 2);
    }
  } else {
    P.zero = true;  // rhs slot 0, slots past the background, threads past the end
  }
  return P;
}

template <int NI>
__global__ __launch_bounds__(256) void k_global_build(ArgsN<BuildArgs, NI> PA) {
  const BuildArgs& A = args_of(PA);
  const int n = *A.nList;
  const int nLower = MATN * (MATN + 1) / 2;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double wts[BUILD_STAMPS_LDS][WT_STRIDE];
  __shared__ int sidx[BUILD_STAMPS_LDS];
  __shared__ int svalid[BUILD_STAMPS_LDS];
  // element coordinates (r >= c) of the lower triangle, or rhs entries after
  int r = -1, c = -1, rr = -1;
  if (e < nLower) {
    // r(r+1)/2 <= e
    int t = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
    while ((t + 1) * (t + 2) / 2 <= e) t++;
    while (t * (t + 1) / 2 > e) t--;
    r = t;
    c = e - t * (t + 1) / 2;
  } else if (e < nLower + NSOL) {
    rr = e - nLower;
  }
  const TermPlan P = plan_term(A, r, c, rr);
  double acc = 0.0;
  for (int s0 = 0; s0 < n; s0 += BUILD_STAMPS_LDS) {
    const int cn = min(BUILD_STAMPS_LDS, n - s0);
    __syncthreads();
    for (int u = threadIdx.x; u < cn; u += blockDim.x) {
      int s = A.list ? A.list[s0 + u] : (s0 + u);
      int gk = A.kept[s];
      int cu = A.cur[s];
      bool valid = cu < A.count[gk];
      sidx[u] = s;
      svalid[u] = valid ? 1 : 0;
      double wt[NC2];
      if (valid) stamp_weights(A.w, A.h, A.ssx[gk * MAXSS + cu], A.ssy[gk * MAXSS + cu], wt);
      else
        for (int q = 0; q < NC2; q++) wt[q] = 0.0;
      for (int q = 0; q < NC2; q++) wts[u][q] = wt[q];
      wts[u][WT_ONE] = 1.0;
      wts[u][WT_ONE + 1] = 0.0;
    }
    __syncthreads();
    // Terms of 16 stamps loaded together, then added in stamp order; the
    // loads are unconditional (every stamp index in the list addresses real
    // rows), so nothing between them waits on memory. A stamp whose substamps
    // are exhausted adds nothing.
    for (int u0 = 0; u0 < cn; u0 += 16) {
      double t[16];
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int u = min(u0 + v, cn - 1);
        const double g = P.src[(size_t)sidx[u] * P.stride + P.off];
        t[v] = (wts[u][P.ia] * wts[u][P.ib]) * g;
      }
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const bool take = u0 + v < cn && svalid[min(u0 + v, cn - 1)];
        acc = take ? acc + t[v] : acc;
      }
    }
  }
  if (P.zero) acc = 0.0;
  if (r >= 0) {
    A.M[(size_t)r * MATN + c] = acc;
    A.M[(size_t)c * MATN + r] = acc;
    A.M0[(size_t)r * MATN + c] = acc;
    A.M0[(size_t)c * MATN + r] = acc;
  } else if (rr >= 0) {
    A.rhs[rr] = acc;
    A.rhs0[rr] = acc;
  }
}

// ---------------------------------------------------------------------------
// kernel coefficients at (xf, yf) (cd.cl:354-383, bachUtil.cpp:508-526)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double kernel_coeff(const double* sol, int i, double xf, double yf) {
  if (i == 0) return sol[1];
  int k = 2 + (i - 1) * NC2;
  double c0 = 0.0, aX = 1.0;
  for (int x = 0; x <= 2; x++) {
    double aY = 1.0;
    for (int y = 0; y <= 2 - x; y++) {
      c0 += sol[k++] * aX * aY;
      aY *= yf;
    }
    aX *= xf;
  }
  return c0;
}

// getBackground (cd.cl:444-467)
__device__ __forceinline__ double background_at(const double* sol, int x, int y, int w, int h) {
  double xf = (x - 0.5 * w) / (0.5 * w);
  double yf = (y - 0.5 * h) / (0.5 * h);
  double bg = 0.0, ax = 1.0;
  int k = NCOMP + 2;
  for (int i = 0; i <= 1; i++) {
    double ay = 1.0;
    for (int j = 0; j <= 1 - i; j++) {
      bg += sol[k++] * ax * ay;
      ay *= yf;
    }
    ax *= xf;
  }
  return bg;
}

// ---------------------------------------------------------------------------
// calcSigs (cdkscUtil.cpp:477-583, cd.cl:469-655)
// ---------------------------------------------------------------------------
struct SigArgs {
  const float* tImg;
  const float* sImg;
  uint16_t* mask;
  int w, h;
  const double* W;
  const int* kept;
  const int* cur;
  const int* count;
  const int* ssx;
  const int* ssy;
  const int* list;   // nullptr: all kept stamps
  const int* nList;
  const Basis* basis;
  const double* sol;
  double* sigOut;    // indexed by list position
  uint16_t nanMark;  // mask bit of a NaN model pixel
};

// One workgroup per stamp, NT threads striding over the 992 >= 31 x 31 pixel
// slots of the reference's 32 x 31 reduction layout (one slot per thread at
// 1024), the 48 basis terms of a pixel loaded together before its
// accumulation chain.  The slots' partial sums are added in a fixed order, so
// every NT gives the same bits.  NT = 1024 is faster for a lone pair; 256 is
// faster when many pairs share the GPU (a smaller workgroup finds room on a
// busy CU sooner), so one-stream contexts (batch workers) launch 256.
constexpr int SIG_THREADS = 1024, SIG_THREADS_SHARED = 256;
template <int NI, int NT>
__global__ __launch_bounds__(NT) void k_calc_sigs(ArgsN<SigArgs, NI> P) {
  const SigArgs& A = args_of(P);
  const int u = blockIdx.x;
  if (u >= *A.nList) return;
  const int tid = threadIdx.x;
  const int s = A.list ? A.list[u] : u;
  const int gk = A.kept[s];
  const int cu = A.cur[s];
  const bool valid = cu < A.count[gk];
  __shared__ double co[NPSF];
  __shared__ double sv[32 * 31 + 32];
  __shared__ int cv[32 * 31 + 32];
  __shared__ double part[32];
  __shared__ int pcnt[32];
  __shared__ double s_bg;
  int ssx = 0, ssy = 0;
  if (valid) {
    ssx = A.ssx[gk * MAXSS + cu];
    ssy = A.ssy[gk * MAXSS + cu];
  }
  const double* sol = A.sol;
  if (valid) {
    double xf = (ssx - A.w * 0.5) / (A.w * 0.5);
    double yf = (ssy - A.h * 0.5) / (A.h * 0.5);
    if (tid < NPSF) co[tid] = kernel_coeff(sol, tid, xf, yf);
    if (tid == 64) s_bg = background_at(sol, ssx, ssy, A.w, A.h);
  } else if (tid == 64) {
    s_bg = 0.0;
  }
  __syncthreads();
  const double* W = A.W + (size_t)s * WROWS * NPIX;
  const int* evenSub = A.basis->evenSub;
  const int total = 32 * 31;  // reduceCount * localSize, cdkscUtil.cpp:484-541
  for (int gi = tid; gi < total; gi += blockDim.x) {
    const int gx = gi % SSW, gy = gi / SSW;
    double s0 = 0.0;
    int cnt = 0;
    if (valid && gy < SSW) {
      // makeModel (cd.cl:489-531): FP32 accumulator, FP64 terms
      float m0 = (float)(sol[1] * W[gi]);
#pragma unroll
      for (int i = 1; i < NPSF; i++) {
        double wv = W[(size_t)i * NPIX + gi];
        if (evenSub[i]) wv = wv - W[gi];
        m0 = (float)((double)m0 + co[i] * wv);
      }
      const int absX = min(max(gx - SSW / 2 + ssx, 0), A.w - 1);
      const int absY = min(max(gy - SSW / 2 + ssy, 0), A.h - 1);
      const size_t ai = (size_t)absX + (size_t)absY * A.w;
      const double tDat = (double)m0;
      const double sDat = (double)A.sImg[ai];
      const double diff = tDat - sDat + s_bg;
      if ((A.mask[ai] & XB_MASK_BAD_INPUT) == 0 && fabs(sDat) > 1e-10) {
        if (isnan(tDat) || isnan(sDat)) {
          mask_or_atomic(A.mask, ai, A.nanMark);
        } else {
          cnt = 1;
          s0 = diff * diff / (fabs((double)A.tImg[ai]) + fabs(sDat));
        }
      }
    }
    sv[gi] = s0;
    cv[gi] = cnt;
  }
  __syncthreads();
  if (tid < 31) {
    double ls = 0.0;
    int lc = 0;
    for (int li = 0; li < 32; li++) {
      ls += sv[tid * 32 + li];
      lc += cv[tid * 32 + li];
    }
    part[tid] = ls;
    pcnt[tid] = lc;
  }
  __syncthreads();
  if (tid == 0) {
    double sum = 0.0;
    int lc = 0;
    for (int g = 0; g < 31; g++) sum += part[g];
    for (int g = 0; g < 31; g++) lc += pcnt[g];
    if (lc == 0) {
      sum = -1.0;
    } else {
      sum /= lc;
      if (sum >= 1e10) sum = -1.0;
    }
    A.sigOut[u] = sum;
  }
}

// ---------------------------------------------------------------------------
// cd: kernel sum at (0, 0) (cd.cl:385-442) and the merit (cdkscUtil.cpp:221-266)
// ---------------------------------------------------------------------------
struct CdFinishArgs {
  const Basis* basis;
  const double* sol;
  const double* merits;
  DevScalars* sc;
  int im, w, h;
  double alpha;
  DevScalars* hsc;  // the host's pinned mirror (its merit) or null
};

template <int NI>
__global__ __launch_bounds__(1024) void k_cd_finish(ArgsN<CdFinishArgs, NI> P) {
  const CdFinishArgs& FA = args_of(P);
  const Basis* basis = FA.basis;
  const double* sol = FA.sol;
  const double* merits = FA.merits;
  DevScalars* sc = FA.sc;
  const int im = FA.im, w = FA.w, h = FA.h;
  const double alpha = FA.alpha;
  DevScalars* hsc = FA.hsc;
  __shared__ double co[NPSF];
  __shared__ double kern[KK];
  __shared__ double part[64];
  __shared__ double data[XB_MAX_STAMPS];
  __shared__ uint8_t clip[XB_MAX_STAMPS];
  __shared__ double res[2];
  __shared__ int isc[20];
  __shared__ int s_n;
  const int tid = threadIdx.x;
  // the merit, also into the host's pinned mirror (one thread)
  auto setMerit = [&](double m) {
    sc->merit[im] = m;
    if (hsc) {
      hsc->merit[im] = m;
      __threadfence_system();
    }
  };
  const int nTest = sc->nTest[im];
  if (nTest == 0) {
    if (tid == 0) {
      sc->kernelSumCd[im] = 0.0;
      setMerit(666);
    }
    return;
  }
  const double hw = 0.5 * w, hh = 0.5 * h;
  const double xf = (0 - hw) / hw, yf = (0 - hh) / hh;
  if (tid < NPSF) co[tid] = kernel_coeff(sol, tid, xf, yf);
  __syncthreads();
  if (tid < KK) {
    double k0 = 0.0;
    for (int p = 0; p < NPSF; p++) k0 += co[p] * basis->vec[p * KK + tid];
    kern[tid] = k0;
  }
  __syncthreads();
  const int nb = (KK + 31) / 32;
  if (tid < nb) {
    double sm = 0.0;
    int lc = min(32, KK - tid * 32);
    for (int i = 0; i < lc; i++) sm += kern[tid * 32 + i];
    part[tid] = sm;
  }
  __syncthreads();
  __shared__ double s_ksum;
  if (tid == 0) {
    double sm = 0.0;
    for (int g = 0; g < nb; g++) sm += part[g];
    s_ksum = sm;
    sc->kernelSumCd[im] = sm;
  }
  // clean merits (removeBadSigs, cd.cl:657-695) in canonical order
  if (tid == 0) s_n = 0;
  __syncthreads();
  for (int base = 0; base < nTest; base += blockDim.x) {
    int i = base + tid;
    bool keep = i < nTest && merits[i] >= 0.0;
    int tot;
    int pos = block_prefix_count(keep, isc, &tot);
    int c0 = s_n;
    if (keep) data[c0 + pos] = merits[i];
    __syncthreads();
    if (tid == 0) s_n = c0 + tot;
    __syncthreads();
  }
  const int n = s_n;
  if (n == 0) {
    if (tid == 0) setMerit(666);
    return;
  }
  sigma_clip_small(data, n, 10, alpha, clip, part, res, isc);
  if (tid == 0) setMerit(res[0] / s_ksum);
}

// ---------------------------------------------------------------------------
// ksc: checkBadSubStamps + sigma clip + checkBadSubStampsFromSigmaClip
// (ksc.cl:1-72, cdkscUtil.cpp:686-802).  Both rejection passes use the same
// chi2 values; stamps rejected by either get their next substamp (refilled
// once with the final index, which equals the reference's two refills).
// rec (may be null): receives every stamp's substamp index as the pass found
// it (the screen's record of the pass); zeroBits (may be null): cleared, one
// byte per stamp (the screen's flags, at a loop's first pass).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_ksc_check(const double* sig, const int* kept,
                                                    int* cur, const int* count, int* list,
                                                    const int* nKept, KscScalars* ks,
                                                    int first, double alpha,
                                                    double sigKernFit, int* rec,
                                                    uint8_t* zeroBits, KscScalars* hks,
                                                    int* label) {
  __shared__ double data[XB_MAX_STAMPS];
  __shared__ uint8_t clip[XB_MAX_STAMPS];
  __shared__ uint8_t refill[XB_MAX_STAMPS];
  __shared__ double part[64];
  __shared__ double res[2];
  __shared__ int isc[20];
  __shared__ int s_n;
  const int tid = threadIdx.x;
  const int n = *nKept;
  if (tid == 0) s_n = 0;
  // each thread records the entries it changes below (s = tid + k blockDim)
  for (int s = tid; s < n; s += blockDim.x) {
    refill[s] = 0;
    if (rec) rec[s] = cur[s];
    if (zeroBits) zeroBits[s] = 0;
  }
  __syncthreads();
  for (int base = 0; base < n; base += blockDim.x) {
    int s = base + tid;
    bool isChi = false;
    double v = 0.0;
    if (s < n && cur[s] < count[kept[s]]) {
      v = sig[s];
      if (v == -1.0) {
        cur[s] = cur[s] + 1;
        refill[s] = 1;
      } else {
        isChi = true;
      }
    }
    int tot;
    int pos = block_prefix_count(isChi, isc, &tot);
    int c0 = s_n;
    if (isChi) data[c0 + pos] = v;
    __syncthreads();
    if (tid == 0) s_n = c0 + tot;
    __syncthreads();
  }
  const int nChi = s_n;
  sigma_clip_small(data, nChi, 10, alpha, clip, part, res, isc);
  const double mean = res[0], sd = res[1];
  for (int s = tid; s < n; s += blockDim.x) {
    if (cur[s] < count[kept[s]]) {
      if ((sig[s] - mean) > sigKernFit * sd) {
        cur[s] = cur[s] + 1;
        refill[s] = 1;
      }
    }
  }
  __syncthreads();
  if (tid == 0) s_n = 0;
  __syncthreads();
  for (int base = 0; base < n; base += blockDim.x) {
    int s = base + tid;
    bool r = s < n && refill[s];
    int tot;
    int pos = block_prefix_count(r, isc, &tot);
    int c0 = s_n;
    if (r) list[c0 + pos] = s;
    __syncthreads();
    if (tid == 0) s_n = c0 + tot;
    __syncthreads();
  }
  // with rec: each stamp's label for the screen's record (2 = this pass
  // replaced its substamp, 0 = kept so far, -1 = no substamp), by the thread
  // that recorded and moved it
  if (rec && label)
    for (int s = tid; s < n; s += blockDim.x) {
      const int pc = rec[s];
      label[s] = pc < count[kept[s]] ? (cur[s] != pc ? 2 : 0) : -1;
    }
  if (tid == 0) {
    ks->nList = s_n;
    ks->check = s_n > 0 ? 1 : 0;
    ks->nChi2 = nChi;
    ks->clipMean = mean;
    ks->clipStd = sd;
    if (first) {  // the first pass's clip statistics (the diagnostics report them)
      ks->firstClip[0] = mean;
      ks->firstClip[1] = sd;
      ks->nScreened = 0;
      ks->nRestored = 0;
      ks->nConfirmed = 0;
    }
    publish_pass(*ks, hks);
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static StampLuArgs stamp_lu_args(Ctx& c, int im) {
  return StampLuArgs{c.st[im].G, c.st[im].Bc, &c.sc->nKept[im], c.st[im].ksum};
}

hipError_t launch_stamp_lu(Ctx& c, int im) {
  if (c.nKept[im] <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stamp_lu<1>, dim3(c.nKept[im]), dim3(64), 0, c.cur,
                     ArgsN<StampLuArgs, 1>{{stamp_lu_args(c, im)}});
  return hipGetLastError();
}

static CdSelectArgs cd_select_args(Ctx& c, int im) {
  return CdSelectArgs{c.st[im].ksum, &c.sc->nKept[im], c.st[im].list, c.sc, im,
                      c.m.sigClipAlpha, c.m.sigKernFit};
}

hipError_t launch_cd_select(Ctx& c, int im) {
  hipLaunchKernelGGL(k_cd_select<1>, dim3(1), dim3(1024), 0, c.cur,
                     ArgsN<CdSelectArgs, 1>{{cd_select_args(c, im)}});
  return hipGetLastError();
}

static BuildArgs build_args(Ctx& c, int im, bool cdForm) {
  BuildArgs A;
  A.G = c.st[im].G;
  A.Bc = c.st[im].Bc;
  A.Br = c.st[im].Br;
  A.kept = c.st[im].kept;
  A.cur = c.st[im].cur;
  A.count = c.st[im].count;
  A.ssx = c.st[im].ssx;
  A.ssy = c.st[im].ssy;
  A.list = cdForm ? c.st[im].list : nullptr;
  A.nList = cdForm ? &c.sc->nList[im] : &c.sc->nKept[im];
  A.w = c.w;
  A.h = c.h;
  A.cdForm = cdForm ? 1 : 0;
  const LuSet& L = cdForm ? c.lu[im] : c.ksLu[im];
  A.M = L.M;
  A.rhs = L.rhs;
  A.M0 = L.M0;
  A.rhs0 = L.rhs0;
  return A;
}

constexpr int BUILD_TOTAL = MATN * (MATN + 1) / 2 + NSOL;

hipError_t launch_global_build(Ctx& c, int im, bool cdForm) {
  hipLaunchKernelGGL(k_global_build<1>, dim3((BUILD_TOTAL + 255) / 256), dim3(256), 0, c.cur,
                     ArgsN<BuildArgs, 1>{{build_args(c, im, cdForm)}});
  return hipGetLastError();
}

static SigArgs sig_args(Ctx& c, int im, const int* d_list, const int* d_nList,
                        const double* d_sol, double* out, uint16_t nanMark) {
  SigArgs A;
  A.tImg = c.img[im];
  A.sImg = c.img[1 - im];
  A.mask = c.mask;
  A.w = c.w;
  A.h = c.h;
  A.W = c.st[im].W;
  A.kept = c.st[im].kept;
  A.cur = c.st[im].cur;
  A.count = c.st[im].count;
  A.ssx = c.st[im].ssx;
  A.ssy = c.st[im].ssy;
  A.list = d_list;
  A.nList = d_nList;
  A.basis = c.basis;
  A.sol = d_sol;
  A.sigOut = out;
  A.nanMark = nanMark;
  return A;
}

// threads per stamp of k_calc_sigs: 256 in one-stream contexts (batch
// workers), 1024 otherwise
template <int NI>
static void launch_sigs_n(Ctx& c, int n, const ArgsN<SigArgs, NI>& P) {
  if (c.side == c.stream)
    hipLaunchKernelGGL((k_calc_sigs<NI, SIG_THREADS_SHARED>), dim3(n, 1, NI),
                       dim3(SIG_THREADS_SHARED), 0, c.cur, P);
  else
    hipLaunchKernelGGL((k_calc_sigs<NI, SIG_THREADS>), dim3(n, 1, NI), dim3(SIG_THREADS), 0,
                       c.cur, P);
}

hipError_t launch_calc_sigs(Ctx& c, int im, bool useList, const double* d_sol,
                            uint16_t nanMark) {
  if (c.nKept[im] <= 0) return hipSuccess;
  const SigArgs A =
      sig_args(c, im, useList ? c.st[im].list : nullptr,
               useList ? &c.sc->nList[im] : &c.sc->nKept[im], d_sol,
               useList ? c.st[im].sig : c.st[im].sigK, nanMark);
  launch_sigs_n<1>(c, c.nKept[im], ArgsN<SigArgs, 1>{{A}});
  return hipGetLastError();
}

hipError_t launch_calc_sigs_list(Ctx& c, int im, const int* d_list, const int* d_nList,
                                 const double* d_sol, double* out) {
  if (c.nKept[im] <= 0) return hipSuccess;
  const SigArgs A = sig_args(c, im, d_list, d_nList, d_sol, out, XB_MASK_NAN_PIXEL);
  launch_sigs_n<1>(c, c.nKept[im], ArgsN<SigArgs, 1>{{A}});
  return hipGetLastError();
}

static CdFinishArgs cd_finish_args(Ctx& c, int im) {
  return CdFinishArgs{c.basis, c.cdSol + (size_t)im * NSOL, c.st[im].sig, c.sc, im, c.w, c.h,
                      c.m.sigClipAlpha, c.hscDev};
}

hipError_t launch_cd_finish(Ctx& c, int im) {
  hipLaunchKernelGGL(k_cd_finish<1>, dim3(1), dim3(1024), 0, c.cur,
                     ArgsN<CdFinishArgs, 1>{{cd_finish_args(c, im)}});
  return hipGetLastError();
}

// cd's per-stamp fits, test-stamp selection and global build of both images
// in one dispatch each (image z = blockIdx.z); both images must have stamps
hipError_t launch_cd_begin_pair(Ctx& c) {
  const int n = std::max(c.nKept[0], c.nKept[1]);
  hipLaunchKernelGGL(k_stamp_lu<2>, dim3(n, 1, 2), dim3(64), 0, c.cur,
                     ArgsN<StampLuArgs, 2>{{stamp_lu_args(c, 0), stamp_lu_args(c, 1)}});
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cd_select<2>, dim3(1, 1, 2), dim3(1024), 0, c.cur,
                     ArgsN<CdSelectArgs, 2>{{cd_select_args(c, 0), cd_select_args(c, 1)}});
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_global_build<2>, dim3((BUILD_TOTAL + 255) / 256, 1, 2), dim3(256), 0,
                     c.cur, ArgsN<BuildArgs, 2>{{build_args(c, 0, true), build_args(c, 1, true)}});
  return hipGetLastError();
}

// cd's chi2 of the test stamps and the merits of both images in one dispatch
// each (after their solves)
hipError_t launch_cd_end_pair(Ctx& c) {
  const int n = std::max(c.nKept[0], c.nKept[1]);
  SigArgs A[2];
  for (int im = 0; im < 2; im++)
    A[im] = sig_args(c, im, c.st[im].list, &c.sc->nList[im], c.cdSol + (size_t)im * NSOL,
                     c.st[im].sig, XB_MASK_NAN_PIXEL);
  launch_sigs_n<2>(c, n, ArgsN<SigArgs, 2>{{A[0], A[1]}});
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_cd_finish<2>, dim3(1, 1, 2), dim3(1024), 0, c.cur,
                     ArgsN<CdFinishArgs, 2>{{cd_finish_args(c, 0), cd_finish_args(c, 1)}});
  return hipGetLastError();
}

hipError_t launch_ksc_check(Ctx& c, int im, bool first, int* rec, uint8_t* zeroBits,
                            bool publish, int* label) {
  hipLaunchKernelGGL(k_ksc_check, dim3(1), dim3(1024), 0, c.cur, c.st[im].sigK,
                     c.st[im].kept, c.st[im].cur, c.st[im].count, c.st[im].listK,
                     &c.sc->nKept[im], &c.sc->k[im], first ? 1 : 0, c.m.sigClipAlpha,
                     c.m.sigKernFit, rec, zeroBits, publish ? &c.hscDev->k[im] : nullptr,
                     label);
  return hipGetLastError();
}

// every kept stamp moved on from its first substamp goes back to it and into
// the list; one workgroup, the list in ascending stamp order
__global__ __launch_bounds__(1024) void k_restore_list(const int* nKept, int* cur, int* list,
                                                       KscScalars* ks) {
  __shared__ int isc[20];
  __shared__ int s_n;
  const int tid = threadIdx.x;
  const int n = *nKept;
  if (tid == 0) s_n = 0;
  __syncthreads();
  for (int base = 0; base < n; base += blockDim.x) {
    const int s = base + tid;
    const bool r = s < n && cur[s] != 0;
    if (r) cur[s] = 0;
    int tot;
    const int pos = block_prefix_count(r, isc, &tot);
    const int c0 = s_n;
    if (r) list[c0 + pos] = s;
    __syncthreads();
    if (tid == 0) s_n = c0 + tot;
    __syncthreads();
  }
  if (tid == 0) ks->nList = s_n;
}

hipError_t launch_restore_fill(Ctx& c, int im) {
  if (c.nKept[im] <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_restore_list, dim3(1), dim3(1024), 0, c.cur, &c.sc->nKept[im],
                     c.st[im].cur, c.st[im].listK, &c.sc->k[im]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_fill(c, im, c.st[im].listK, &c.sc->k[im].nList, c.nKept[im]);
}

}  // namespace xb
