// k_cmv.hip — basis convolution of substamps and their normal equations (cmv).
//
// k_fill_conv: one workgroup per (stamp, W row).  The 51x51 image window of
// the stamp's current substamp is staged in LDS; the Y pass (FP32 accumulator,
// FP64 taps, cl_kern/cmv.cl:70-106) and the X pass (FP64, cmv.cl:108-135) run
// from LDS and the raw W row is written once.  Background rows are the
// monomials of cmv.cl:156-193 (FP32 xf, yf).  Even-even rows keep the raw
// value; every consumer applies W[n] - W[0] on load (cmv.cl:137-154), a single
// subtraction exactly as the reference does in place.  Each row is written
// twice: in raster order (W, for the Gram matrix and the fits) and in
// column-major pixel order (Wc, for the target products, which the reference
// sums in that order), so both consumers read contiguous rows.
//
// k_fill_gram: the 52x52 Gram matrix G = W^T W of the final rows, one wave
// per 16x16 tile of rows/columns 0..47 (64-pixel chunks staged through LDS)
// and one per strip of rows 48..51 (32-pixel chunks of all rows), each entry
// a sequential sum over the 961 pixels (cmv.cl:195-215 order).  The next chunk
// is loaded raw (its even-row subtraction happens when it is stored to LDS,
// so no wait sits before the products), and the products read LDS one 8-step
// block ahead (a sched_barrier keeps the reads early): a refill of a few
// stamps is bound by one tile's latency.
// The same launch's last work item per stamp forms the target dot products
// with one wave: column-major order for all 52 rows (createB, cmv.cl:217-248,
// and the ksc right-hand side, src/cdkscUtil.cpp:459-470) and row-major order
// for the 3 background rows (cd right-hand side, cl_kern/cd.cl:318-344), beside
// the Gram items instead of after them.  Q (51x51, 1-based) is G shifted by one.
#include <algorithm>

#include "xb_dev.h"

namespace xb {

// entries of a listed refill covered by one pass of its grid
constexpr int FILL_LIST_STAMPS = 16;

struct FillConvLds {
  float win[TW][TW + 1];
  float tmp[TH * TW];
  double fx[KW], fy[KW];
};

// W row n of stamp s (substamp centre ssx, ssy)
__device__ __forceinline__ void fill_conv_row(const FillArgs& A, FillConvLds& L, int s, int ssx,
                                              int ssy, int n) {
  const int g = A.kept[s];
  const bool valid = A.cur[s] < A.count[g];
  const int tid = threadIdx.x;
  double* Wrow = A.W + ((size_t)s * WROWS + n) * NPIX;
  double* Wcrow = A.Wc + ((size_t)s * WROWS + n) * NPIX;  // column-major copy
  auto put = [&](int p, double v) {
    Wrow[p] = v;
    Wcrow[(p % SSW) * SSW + p / SSW] = v;
  };
  if (!valid) {
    for (int p = tid; p < NPIX; p += blockDim.x) put(p, 0.0);
    return;
  }
  if (n >= NPSF) {
    // background monomials, bgXY order (0,0), (0,1), (1,0)
    const int b = n - NPSF;
    const float hw = (float)A.w * 0.5f, hh = (float)A.h * 0.5f;
    for (int p = tid; p < NPIX; p += blockDim.x) {
      int x = ssx - HSS + p % SSW;
      int y = ssy - HSS + p / SSW;
      double xf = (double)(((float)x - hw) / hw);
      double yf = (double)(((float)y - hh) / hh);
      double ax = b == 2 ? xf : 1.0;
      double ay = b == 1 ? yf : 1.0;
      put(p, ax * ay);
    }
    return;
  }
  auto& win = L.win;
  auto& tmp = L.tmp;
  auto& fx = L.fx;
  auto& fy = L.fy;
  if (tid < KW) {
    fx[tid] = A.basis->fx[n * KW + tid];
    fy[tid] = A.basis->fy[n * KW + tid];
  }
  const int x0 = ssx - (HSS + HKW), y0 = ssy - (HSS + HKW);
  for (int i = tid; i < TW * TW; i += blockDim.x) {
    int r = i / TW, c = i % TW;
    // clamped reads never trigger for real substamps (their centres lie
    // >= 40 px inside the image: border mask + window check)
    int xr = min(max(x0 + c, 0), A.w - 1), yr = min(max(y0 + r, 0), A.h - 1);
    win[r][c] = A.img[(size_t)xr + (size_t)yr * A.w];
  }
  __syncthreads();
  // Y pass: tmp[r][c], image row ssy + r - 15, column ssx + c - 25
  for (int p = tid; p < TH * TW; p += blockDim.x) {
    int r = p / TW, c = p % TW;
    float v = 0.0f;
    for (int yy = -HKW; yy <= HKW; yy++) {
      float iv = win[r + HKW + yy][c];
      v = (float)((double)v + (double)iv * fy[HKW - yy]);
    }
    tmp[p] = v;
  }
  __syncthreads();
  // X pass
  for (int p = tid; p < NPIX; p += blockDim.x) {
    int j = p / SSW - HSS, i = p % SSW - HSS;
    double w0 = 0.0;
    for (int xx = -HKW; xx <= HKW; xx++) {
      int idx = (i + xx) + TW / 2 + (j + HSS) * TW;
      w0 += (double)tmp[idx] * fx[HKW - xx];
    }
    put(p, w0);
  }
}

// One workgroup per (list entry, W row).  LOOP (listed refills): the grid
// covers FILL_LIST_STAMPS entries and loops over the rest (the list's length
// is on the device), so a short list does not dispatch a workgroup per kept
// stamp.
template <bool LOOP, int NI>
__global__ __launch_bounds__(256) void k_fill_conv(ArgsN<FillArgs, NI> P) {
  const FillArgs& A = args_of(P);
  __shared__ FillConvLds L;
  for (int k = blockIdx.x;; k += gridDim.x) {
    int s, ssx, ssy;
    if (!fill_stamp(A, k, &s, &ssx, &ssy)) return;
    fill_conv_row(A, L, s, ssx, ssy, blockIdx.y);
    if (!LOOP) return;
    __syncthreads();  // the next entry reuses the LDS
  }
}

constexpr int GT = 16;   // Gram tile
constexpr int GK = 64;   // pixel chunk

// Per stamp, the lower triangle of the 52x52 Gram matrix is split into NGT
// work items of equal length (961 pixel steps each): the 6 lower-triangular
// 16x16 tiles of rows/columns 0..47, and one strip holding rows 48..51 against
// columns 0..51.
constexpr int NGT = 7;
constexpr int NGI = NGT + 1;   // work items per stamp: the Gram items and the target products
constexpr int GP = GT + 2;     // tile LDS pitch: [pixel][row], 16-byte aligned row pairs
constexpr int SR0 = 3 * GT;    // first strip row
constexpr int SK = 32;         // strip pixel chunk
constexpr int SH = WROWS / 2;  // W rows fetched by each half of the wave
constexpr int SP = WROWS + 2;  // strip LDS pitch: [pixel][row] over all W rows
static_assert(SR0 + 4 == WROWS, "the strip holds the last four rows");
static_assert(SK * SP <= 2 * GK * GP, "a strip chunk fits the tile LDS");
static_assert(SP % 2 == 0, "row pairs stay 16-byte aligned");

// Adds 8*NB pixel steps to a lane's 2x2 block: the row pair at xb and the
// column pair at yb, consecutive pixels P doubles apart in LDS.  Block j + 1's
// LDS reads are issued before block j is computed (two register sets, the
// index resolved by the unrolling), so the reads overlap the products.  Every
// entry adds its products in ascending pixel order.
template <int P, int NB>
__device__ __forceinline__ void gram_block(const double* xb, const double* yb, double& c00,
                                           double& c01, double& c10, double& c11) {
  double2 xa[2][8], ya[2][8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    xa[0][j] = *reinterpret_cast<const double2*>(xb + j * P);
    ya[0][j] = *reinterpret_cast<const double2*>(yb + j * P);
  }
#pragma unroll
  for (int blk = 0; blk < NB; blk++) {
    const int cu = blk & 1, nx = cu ^ 1;
    if (blk + 1 < NB) {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        xa[nx][j] = *reinterpret_cast<const double2*>(xb + (8 * (blk + 1) + j) * P);
        ya[nx][j] = *reinterpret_cast<const double2*>(yb + (8 * (blk + 1) + j) * P);
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the next block's reads here
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const double2 x = xa[cu][j], y = ya[cu][j];
      const double p00 = x.x * y.x, p01 = x.x * y.y, p10 = x.y * y.x, p11 = x.y * y.y;
      c00 += p00;
      c01 += p01;
      c10 += p10;
      c11 += p11;
    }
  }
}

// The same for a partial chunk of kn steps.
template <int P>
__device__ __forceinline__ void gram_steps(const double* xb, const double* yb, int kn,
                                           double& c00, double& c01, double& c10, double& c11) {
  for (int kk = 0; kk < kn; kk++) {
    const double2 x = *reinterpret_cast<const double2*>(xb + kk * P);
    const double2 y = *reinterpret_cast<const double2*>(yb + kk * P);
    const double p00 = x.x * y.x, p01 = x.x * y.y, p10 = x.y * y.x, p11 = x.y * y.y;
    c00 += p00;
    c01 += p01;
    c10 += p10;
    c11 += p11;
  }
}

// Bit r set when W row r (r < 49) is an even basis term: one load per lane of
// the wave and a ballot (a loop of guarded loads would wait on each in turn).
__device__ __forceinline__ unsigned long long even_rows(const int* evenSub, int lane) {
  return __ballot(lane < NPSF && evenSub[min(lane, NPSF - 1)] != 0);
}

// The target products of one stamp by one wave (the last work item of
// k_fill_gram, beside the Gram items): lane r < 52 sums row r of W (minus row
// 0 for even basis terms) times the target in column-major pixel order (createB,
// cmv.cl:217-248, and the ksc right-hand side, src/cdkscUtil.cpp:459-470);
// lanes 52..54 sum the three background monomials times the target in
// row-major order (the cd right-hand side, cl_kern/cd.cl:318-344).  Every sum
// runs in its reference order.  32-position chunks are staged through LDS,
// [position][column]: the 52 column-major W rows, the target in both orders,
// and the background weights of each row-major position (1, y monomial, x
// monomial: (1.0 * 1.0), (1.0 * yf) and (xf * 1.0) of the reference are these
// values exactly), so all 55 chains run one uniform loop.  Two chunks' loads
// are in flight while one is summed.
constexpr int TQ_C = WROWS;       // target, column-major order
constexpr int TQ_R = WROWS + 1;   // target, row-major order
constexpr int TQ_B = WROWS + 2;   // background weights 1, yf, xf (3 columns)
constexpr int TP = WROWS + 6;     // LDS pitch (16-byte aligned rows)
constexpr int TQ_N = SH + 2;      // values a lane loads per chunk
static_assert(TQ_B + NBG <= TP && SK * TP <= 2 * GK * GP, "target chunks fit the LDS");

__device__ __forceinline__ void fill_target_wave(const FillArgs& A, int s, int ssx, int ssy,
                                                 bool valid, double* Wl) {
  const int lane = threadIdx.x;
  if (!valid) {
    if (lane < WROWS) A.Bc[(size_t)s * WROWS + lane] = 0.0;
    if (lane < NBG) A.Br[(size_t)s * NBG + lane] = 0.0;
    return;
  }
  double(*Ts)[TP] = reinterpret_cast<double(*)[TP]>(Wl);
  const int px = lane % SK, h = lane / SK;
  // bit r: row h*SH + r is an even basis term
  const unsigned long long sub = even_rows(A.basis->evenSub, lane) >> (h * SH);
  const double* Wc = A.Wc + (size_t)s * WROWS * NPIX;
  const float hw = (float)A.w * 0.5f, hh = (float)A.h * 0.5f;
  // raw loads of chunk k0 into v: 26 rows, row 0, and the target value at
  // position k0 + px (half 0 column-major, half 1 row-major), clamped
  auto fetch = [&](int k0, double (&v)[TQ_N]) {
    const int k = min(k0 + px, NPIX - 1);
#pragma unroll
    for (int r = 0; r < SH; r++) v[r] = Wc[(size_t)(h * SH + r) * NPIX + k];
    v[SH] = Wc[k];
    const int xx = h == 0 ? k / SSW : k % SSW, yy = h == 0 ? k % SSW : k / SSW;
    const int xr = min(max(ssx - HSS + xx, 0), A.w - 1);
    const int yr = min(max(ssy - HSS + yy, 0), A.h - 1);
    v[SH + 1] = (double)A.img2[(size_t)xr + (size_t)yr * A.w];
  };
  auto store = [&](int k0, const double (&v)[TQ_N]) {
#pragma unroll
    for (int r = 0; r < SH; r++) {
      double x = v[r];
      if ((sub >> r) & 1) x = x - v[SH];
      Ts[px][h * SH + r] = x;
    }
    Ts[px][TQ_C + h] = v[SH + 1];
    // background weights of row-major position k: monomials as the reference
    // forms them (FP32 xf, yf)
    const int k = k0 + px;
    if (h == 0) {
      Ts[px][TQ_B] = 1.0;
      Ts[px][TQ_B + 1] = (double)(((float)(ssy - HSS + k / SSW) - hh) / hh);
    } else {
      Ts[px][TQ_B + 2] = (double)(((float)(ssx - HSS + k % SSW) - hw) / hw);
    }
  };
  // lane r: W row r times the column-major target; lane 52 + b: background
  // weight b times the row-major target
  const int ca = lane < WROWS ? lane : TQ_B + min(lane - WROWS, NBG - 1);
  const int cb = lane < WROWS ? TQ_C : TQ_R;
  double acc = 0.0;
  auto sum = [&](int k0) {
    const int kn = min(SK, NPIX - k0);
    if (kn == SK) {
#pragma unroll
      for (int kk = 0; kk < SK; kk++) acc += Ts[kk][ca] * Ts[kk][cb];
    } else {
      for (int kk = 0; kk < kn; kk++) acc += Ts[kk][ca] * Ts[kk][cb];
    }
  };
  constexpr int NCH = (NPIX + SK - 1) / SK;
  double va[TQ_N], vb[TQ_N];
  fetch(0, va);
  fetch(SK, vb);
  store(0, va);
  __syncthreads();  // chunk 0 (a single-wave block)
  for (int ch = 0; ch < NCH; ch += 2) {
    // chunk ch is in LDS, ch + 1 in vb; va is free
    if (ch + 2 < NCH) fetch((ch + 2) * SK, va);
    sum(ch * SK);
    if (ch + 1 >= NCH) break;
    store((ch + 1) * SK, vb);
    __syncthreads();
    // chunk ch + 1 is in LDS, ch + 2 in va; vb is free
    if (ch + 3 < NCH) fetch((ch + 3) * SK, vb);
    sum((ch + 1) * SK);
    if (ch + 2 >= NCH) break;
    store((ch + 2) * SK, va);
    __syncthreads();
  }
  if (lane < WROWS) A.Bc[(size_t)s * WROWS + lane] = acc;
  else if (lane < WROWS + NBG) A.Br[(size_t)s * NBG + lane - WROWS] = acc;
}

// One wave per (stamp, work item).  Blocks b and b + 8 run on the same XCD
// (blocks are dealt to the 8 XCDs in turn), so each XCD receives a contiguous
// range of work items and a stamp's W is read through a single L2.
//   tile (bi, bj): lane (lr, lc) = (lane / 8, lane % 8) owns rows
//     bi*16 + 2lr + i, columns bj*16 + 2lc + j; 64-pixel chunks of 16 + 16 rows.
//   strip: lane (rp, cp) = (lane / 26, lane % 26), lanes 0..51, owns rows
//     48 + 2rp + i, columns 2cp + j; 32-pixel chunks of all 52 rows, fetched
//     by lane (pixel, half) = (lane % 32, lane / 32), 26 rows per half.
// Rows of even basis terms have W row 0 subtracted as they are stored to LDS.
// The next pixel chunk is loaded into registers while the current one is
// summed from LDS.  One LDS buffer suffices (a single wave's LDS reads and
// writes run in program order), which leaves room for two waves per SIMD.
__device__ __forceinline__ void gram_item(const FillArgs& A, int s, int ssx, int ssy, int item,
                                          double* Wl) {
  const int g = A.kept[s];
  const bool valid = A.cur[s] < A.count[g];
  const int lane = threadIdx.x;
  if (item == NGT) {
    fill_target_wave(A, s, ssx, ssy, valid, Wl);
    return;
  }
  const bool strip = item == NGT - 1;
  int bi = 0, bj = item;
  while (bj > bi) { bj -= bi + 1; bi++; }
  double* G = A.G + (size_t)s * WROWS * WROWS;
  if (!valid) {
    if (strip) {
      for (int e = lane; e < 4 * WROWS; e += WAVE) {
        const int row = SR0 + e / WROWS, col = e % WROWS;
        if (col <= row) G[row * WROWS + col] = 0.0;
      }
    } else {
      for (int e = lane; e < GT * GT; e += WAVE) {
        const int row = bi * GT + e / GT, col = bj * GT + e % GT;
        if (col <= row) G[row * WROWS + col] = 0.0;
      }
    }
    return;
  }
  const double* W = A.W + (size_t)s * WROWS * NPIX;
  const int* evenSub = A.basis->evenSub;
  double c00 = 0.0, c01 = 0.0, c10 = 0.0, c11 = 0.0;
  if (strip) {
    double(*Ws)[SP] = reinterpret_cast<double(*)[SP]>(Wl);
    const int px = lane % SK, h = lane / SK;
    // bit r: row h*SH + r is an even basis term
    const unsigned long long sub = even_rows(evenSub, lane) >> (h * SH);
    double v[SH], w0;
    // raw loads only (pixels past the window read a clamped address and are
    // zeroed at store time)
    auto fetch = [&](int k0) {
      const int k = min(k0 + px, NPIX - 1);
      w0 = W[k];
#pragma unroll
      for (int r = 0; r < SH; r++) v[r] = W[(size_t)(h * SH + r) * NPIX + k];
    };
    auto store = [&](int k0) {
      const bool in = k0 + px < NPIX;
#pragma unroll
      for (int r = 0; r < SH; r++) {
        double x = in ? v[r] : 0.0;
        if ((sub >> r) & 1) x = x - (in ? w0 : 0.0);
        Ws[px][h * SH + r] = x;
      }
    };
    const int rp = (lane / SH) & 1, cp = lane % SH;
    const double* xb = &Ws[0][SR0 + 2 * rp];
    const double* yb = &Ws[0][2 * cp];
    fetch(0);
    store(0);
    __syncthreads();
    for (int k0 = 0; k0 < NPIX; k0 += SK) {
      const int kn = min(SK, NPIX - k0);
      const bool more = k0 + SK < NPIX;
      if (more) fetch(k0 + SK);
      if (kn == SK)
        gram_block<SP, SK / 8>(xb, yb, c00, c01, c10, c11);
      else
        gram_steps<SP>(xb, yb, kn, c00, c01, c10, c11);
      if (more) {
        store(k0 + SK);
        __syncthreads();
      }
    }
    if (lane < 2 * SH) {
      const int r0 = SR0 + 2 * rp, q0 = 2 * cp;
      const double cv[2][2] = {{c00, c01}, {c10, c11}};
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          if (q0 + j <= r0 + i) G[(r0 + i) * WROWS + q0 + j] = cv[i][j];
    }
    return;
  }
  double(*Wa)[GP] = reinterpret_cast<double(*)[GP]>(Wl);
  double(*Wb)[GP] = reinterpret_cast<double(*)[GP]>(Wl + GK * GP);
  bool subA[GT], subB[GT];
  const unsigned long long em = even_rows(evenSub, lane);
#pragma unroll
  for (int rr = 0; rr < GT; rr++) {
    subA[rr] = (em >> (bi * GT + rr)) & 1ull;
    subB[rr] = (em >> (bj * GT + rr)) & 1ull;
  }
  double va[GT], vb[GT], w0;
  // raw loads only (pixels past the window read a clamped address and are
  // zeroed at store time)
  auto fetch = [&](int k0) {
    const int k = min(k0 + lane, NPIX - 1);
    w0 = W[k];
#pragma unroll
    for (int rr = 0; rr < GT; rr++) {
      va[rr] = W[(size_t)(bi * GT + rr) * NPIX + k];
      vb[rr] = W[(size_t)(bj * GT + rr) * NPIX + k];
    }
  };
  auto store = [&](int k0) {
    const bool in = k0 + lane < NPIX;
#pragma unroll
    for (int rr = 0; rr < GT; rr++) {
      double x = in ? va[rr] : 0.0;
      double y = in ? vb[rr] : 0.0;
      if (subA[rr]) x = x - (in ? w0 : 0.0);
      if (subB[rr]) y = y - (in ? w0 : 0.0);
      Wa[lane][rr] = x;
      Wb[lane][rr] = y;
    }
  };
  const int lr = lane / 8, lc = lane % 8;
  const double* xb = &Wa[0][2 * lr];
  const double* yb = &Wb[0][2 * lc];
  fetch(0);
  store(0);
  __syncthreads();
  for (int k0 = 0; k0 < NPIX; k0 += GK) {
    const int kn = min(GK, NPIX - k0);
    const bool more = k0 + GK < NPIX;
    if (more) fetch(k0 + GK);
    if (kn == GK)
      gram_block<GP, GK / 8>(xb, yb, c00, c01, c10, c11);
    else
      gram_steps<GP>(xb, yb, kn, c00, c01, c10, c11);
    if (more) {
      store(k0 + GK);
      __syncthreads();
    }
  }
  const int r0 = bi * GT + 2 * lr, q0 = bj * GT + 2 * lc;
  const double cv[2][2] = {{c00, c01}, {c10, c11}};
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
      if (q0 + j <= r0 + i) G[(r0 + i) * WROWS + q0 + j] = cv[i][j];
}

// nCover: list entries covered by one pass of the grid (the kept stamps for
// a full fill, FILL_LIST_STAMPS at most for a listed refill, LOOP); later
// entries are taken by further passes.
template <bool LOOP, int NI>
__global__ __launch_bounds__(64) void k_fill_gram(ArgsN<FillArgs, NI> P, int nCover) {
  const FillArgs& A = args_of(P);
  const int total = nCover * NGI;
  const int per = (total + 7) / 8;
  const int wk = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (wk >= total) return;
  __shared__ __align__(16) double Wl[2 * GK * GP];
  for (int base = 0;; base += nCover) {
    int s, ssx, ssy;
    if (!fill_stamp(A, base + wk / NGI, &s, &ssx, &ssy)) return;
    gram_item(A, s, ssx, ssy, wk % NGI, Wl);
    if (!LOOP) return;
    __syncthreads();  // the next entry reuses the LDS
  }
}

static FillArgs fill_args(Ctx& c, int im, const int* d_list, const int* d_nList) {
  FillArgs A;
  A.img = c.img[im];
  A.img2 = c.img[1 - im];
  A.w = c.w;
  A.h = c.h;
  A.list = d_list;
  A.nList = d_list ? d_nList : &c.sc->nKept[im];
  A.kept = c.st[im].kept;
  A.cur = c.st[im].cur;
  A.count = c.st[im].count;
  A.ssx = c.st[im].ssx;
  A.ssy = c.st[im].ssy;
  A.basis = c.basis;
  A.W = c.st[im].W;
  A.Wc = c.st[im].Wc;
  A.G = c.st[im].G;
  A.Bc = c.st[im].Bc;
  A.Br = c.st[im].Br;
  return A;
}

// The FP64 fill: W rows, then the Gram matrix and target products.  nCover:
// the entries one pass of the grids covers; LOOP kernels take longer lists in
// further passes.
template <bool LOOP, int NI>
static hipError_t launch_fill_fp64(Ctx& c, const ArgsN<FillArgs, NI>& P, int nCover) {
  hipLaunchKernelGGL((k_fill_conv<LOOP, NI>), dim3(nCover, WROWS, NI), dim3(256), 0, c.cur, P);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_fill_gram<LOOP, NI>), dim3(8 * ((nCover * NGI + 7) / 8), 1, NI),
                     dim3(WAVE), 0, c.cur, P, nCover);
  return hipGetLastError();
}

hipError_t launch_fill(Ctx& c, int im, const int* d_list, const int* d_nList, int nMax) {
  if (nMax <= 0) return hipSuccess;
  const FillArgs A = fill_args(c, im, d_list, d_nList);
  prof_begin(c, XB_PROF_FILL);
  hipError_t e;
  // The mixed form fills every kept stamp on the matrix cores (one workgroup
  // per stamp for ~80 us: throughput over 100+ stamps).  A listed refill (the
  // few stamps a clip-refit pass moved on, a restore, the screen's probes)
  // takes the FP64 fill instead, a few microseconds for a short list, and no
  // less accurate.
  if ((c.form & XB_FORM_MIXED) && !d_list)
    e = launch_fill_mixed(c, A, nMax);
  else if (d_list)
    e = launch_fill_fp64<true, 1>(c, ArgsN<FillArgs, 1>{{A}}, std::min(nMax, FILL_LIST_STAMPS));
  else
    e = launch_fill_fp64<false, 1>(c, ArgsN<FillArgs, 1>{{A}}, nMax);
  prof_end(c, XB_PROF_FILL);
  return e;
}

hipError_t launch_fill_pair(Ctx& c) {
  const int nMax = std::max(c.nKept[0], c.nKept[1]);
  if (nMax <= 0) return hipSuccess;
  const FillArgs A0 = fill_args(c, 0, nullptr, nullptr), A1 = fill_args(c, 1, nullptr, nullptr);
  prof_begin(c, XB_PROF_FILL);
  const hipError_t e = (c.form & XB_FORM_MIXED)
                           ? launch_fill_mixed_pair(c, A0, A1, nMax)
                           : launch_fill_fp64<false, 2>(c, ArgsN<FillArgs, 2>{{A0, A1}}, nMax);
  prof_end(c, XB_PROF_FILL);
  return e;
}

}  // namespace xb
