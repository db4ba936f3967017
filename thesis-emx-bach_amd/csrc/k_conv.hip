// k_conv.hip — spatially varying convolution and the difference image.
//
// Tile kernels (src/bach.cpp:336-355 realises them on the host, 9,604 calls at
// 2048^2): each workgroup forms the kernels at 16 tile centres in FP64 with the
// reference's sums (src/bachUtil.cpp:502-541), stores them (FP64, read by the
// convolution and the mask test), and sum |k| in the convolution loop's order
// (cl_kern/conv.cl:409-426).
//
// Convolution (cl_kern/conv.cl:31-97): two waves per 21 x 21 tile over the
// tile's input window staged in LDS, lane = (column, row group), FP64
// accumulation in the reference's term order (multiply and add unfused), so
// the convolved and difference images equal the CPU restatement bit for bit.
// The output mask needs sum |k| over unmasked pixels only when the window
// holds a flagged pixel; that rare case re-walks the window in the reference
// order in FP64 so the BAD_CONV decision is exact.  The flags of the 41 x 41
// window are tested once per pixel (one ballot per window row at the start, a
// bit mask per row kept in one lane), so neither the accumulation loop nor
// the masked re-walk tests pixel values; the re-walk reads the tile kernel
// from LDS.  maskAfterConv (conv.cl:99-113) is fused.
//
// sub (cl_kern/sub.cl:15-32) is formed in the convolution's epilogue, from the
// unrounded FP64 value of each pixel.
#include <algorithm>

#include "xb_dev.h"

namespace xb {

__device__ __forceinline__ double kernel_coeff_c(const double* sol, int i, double xf, double yf) {
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

__device__ __forceinline__ double background_c(const double* sol, int x, int y, int w, int h) {
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

// Tile kernels: one workgroup per TK_TILES consecutive tiles (the last tile
// index, nTiles, is the image centre whose kernel sum the sub stage needs).
// Each basis value vec[p][k] is read once per workgroup and used for all its
// tiles; every kernel element is still the sum over p = 0..48 in order
// (unfused multiply, then add), as src/bachUtil.cpp:502-541.
constexpr int TK_TILES = 16;

__global__ __launch_bounds__(512) void k_tile_kernels(const Basis* __restrict__ basis,
                                                      const double* __restrict__ sol, int w,
                                                      int h, int xSteps, int nTiles,
                                                      double* tileK64, double* tileAks,
                                                      double* kernSumOut, double* kernSumHost) {
  const int t0 = blockIdx.x * TK_TILES;
  const int nt = min(TK_TILES, nTiles + 1 - t0);
  const int tid = threadIdx.x;
  __shared__ double co[TK_TILES][NPSF];
  __shared__ double kern[TK_TILES][KK];
  const double hw = 0.5 * w, hh = 0.5 * h;
  for (int e = tid; e < nt * NPSF; e += blockDim.x) {
    const int q = e / NPSF, n = e % NPSF, t = t0 + q;
    int x, y;
    if (t < nTiles) {
      x = (t % xSteps) * KW + HKW + HKW;
      y = (t / xSteps) * KW + HKW + HKW;
    } else {
      x = w / 2;
      y = h / 2;
    }
    co[q][n] = kernel_coeff_c(sol, n, (x - hw) / hw, (y - hh) / hh);
  }
  __syncthreads();
  if (tid < KK) {
    double acc[TK_TILES];
#pragma unroll
    for (int q = 0; q < TK_TILES; q++) acc[q] = 0.0;
    for (int p = 0; p < NPSF; p++) {
      const double v = basis->vec[p * KK + tid];
#pragma unroll
      for (int q = 0; q < TK_TILES; q++) acc[q] += co[q][p] * v;
    }
#pragma unroll
    for (int q = 0; q < TK_TILES; q++) {
      if (q < nt) {
        kern[q][tid] = acc[q];
        if (t0 + q < nTiles) tileK64[(size_t)(t0 + q) * KK + tid] = acc[q];
      }
    }
  }
  __syncthreads();
  if (tid < nt) {  // one thread per tile: the sums in the reference's orders
    const int t = t0 + tid;
    if (t < nTiles) {
      double aks = 0.0;
      for (int i = KK - 1; i >= 0; i--) aks += fabs(kern[tid][i]);
      tileAks[t] = aks;
    } else {
      double sum = 0.0;
      for (int i = 0; i < KK; i++) sum += kern[tid][i];
      *kernSumOut = sum;
      if (kernSumHost) {  // the host's pinned mirror too
        *kernSumHost = sum;
        __threadfence_system();
      }
    }
  }
}

// The mixed form's tile kernels: the same product, tile kernel = sum_n
// coefficient_n(tile) * vec_n, as one FP64 matrix-core product per 16 tiles,
// K (16 tiles x 441 taps) = C (16 x 49) V (49 x 441): wave w of four takes the
// 16-tap column blocks w, w + 4, ... (28 blocks, 13 products of depth 4
// each; the depth padded 49 -> 52 with zeros).  Sum |k| per tile and the
// kernel sum at the image centre (tile index nTiles) in any order (the lane
// sums of a row of 16 lanes by shuffles, the four waves' partials added in
// wave order).
typedef double v4d_t __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_tile_kernels_mfma(const Basis* __restrict__ basis,
                                                           const double* __restrict__ sol, int w,
                                                           int h, int xSteps, int nTiles,
                                                           double* tileK64, double* tileAks,
                                                           double* kernSumOut,
                                                           double* kernSumHost) {
  constexpr int KD = 52;                // depth padded to a multiple of 4
  constexpr int NBLK = (KK + 15) / 16;  // 28 column blocks of 16 taps
  const int t0 = blockIdx.x * 16;
  const int tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
  const int lr = lane >> 4, lc = lane & 15;
  __shared__ double co[16][KD + 1];
  __shared__ double part[2][4][16];  // per wave and tile: sum |k|, sum k
  const double hw = 0.5 * w, hh = 0.5 * h;
  for (int e = tid; e < 16 * KD; e += blockDim.x) {
    const int q = e / KD, n = e % KD, t = t0 + q;
    double v = 0.0;
    if (n < NPSF && t <= nTiles) {
      int x, y;
      if (t < nTiles) {
        x = (t % xSteps) * KW + HKW + HKW;
        y = (t / xSteps) * KW + HKW + HKW;
      } else {
        x = w / 2;
        y = h / 2;
      }
      v = kernel_coeff_c(sol, n, (x - hw) / hw, (y - hh) / hh);
    }
    co[q][n] = v;
  }
  __syncthreads();
  double aks[4] = {0.0, 0.0, 0.0, 0.0}, ksum[4] = {0.0, 0.0, 0.0, 0.0};
  for (int nb = wave; nb < NBLK; nb += 4) {
    const int tap = 16 * nb + lc;
    const int tapc = min(tap, KK - 1);
    double bv[KD / 4];
#pragma unroll
    for (int c = 0; c < KD / 4; c++) {
      const int n = 4 * c + lr;
      const double b = basis->vec[min(n, NPSF - 1) * KK + tapc];
      bv[c] = (n < NPSF && tap < KK) ? b : 0.0;
    }
    v4d_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < KD / 4; c++)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(co[lc][4 * c + lr], bv[c], acc, 0, 0, 0);
    // element (tile lr + 4 v, tap 16 nb + lc) in acc[v]
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int t = t0 + lr + 4 * v;
      if (tap < KK) {
        if (t < nTiles) tileK64[(size_t)t * KK + tap] = acc[v];
        aks[v] += fabs(acc[v]);
        ksum[v] += acc[v];
      }
    }
  }
  // sums over the 16 lanes of each row of lanes (the block's taps)
#pragma unroll
  for (int v = 0; v < 4; v++) {
    double a = aks[v], k = ksum[v];
    a += __shfl_xor(a, 8);
    k += __shfl_xor(k, 8);
    a += __shfl_xor(a, 4);
    k += __shfl_xor(k, 4);
    a += __shfl_xor(a, 2);
    k += __shfl_xor(k, 2);
    a += __shfl_xor(a, 1);
    k += __shfl_xor(k, 1);
    if (lc == 0) {
      part[0][wave][lr + 4 * v] = a;
      part[1][wave][lr + 4 * v] = k;
    }
  }
  __syncthreads();
  if (tid < 16) {
    const int t = t0 + tid;
    const double a = ((part[0][0][tid] + part[0][1][tid]) + part[0][2][tid]) + part[0][3][tid];
    const double k = ((part[1][0][tid] + part[1][1][tid]) + part[1][2][tid]) + part[1][3][tid];
    if (t < nTiles) {
      tileAks[t] = a;
    } else if (t == nTiles) {
      *kernSumOut = k;
      if (kernSumHost) {  // the host's pinned mirror too
        *kernSumHost = k;
        __threadfence_system();
      }
    }
  }
}

struct ConvArgs {
  const float* T;      // image being convolved
  const float* S;      // target image
  uint16_t* mask;
  float* out;
  float* diff;         // difference image (sub, fused: formed from the FP64 value)
  const double* tileK64;
  const double* tileAks;
  const double* sol;
  const double* kernSum;
  int w, h, xSteps, yTiles;
  int scaleConv;
  double hi, lo;
};

__device__ __forceinline__ uint16_t conv_flags(float t, double hi, double lo) {
  double d = (double)t;
  uint16_t m = 0;
  if (d == 0.0) m |= XB_MASK_BAD_INPUT | XB_MASK_BAD_PIX_VAL;
  if (d >= hi) m |= XB_MASK_BAD_INPUT | XB_MASK_SAT_PIXEL;
  if (d <= lo) m |= XB_MASK_BAD_INPUT | XB_MASK_LOW_PIXEL;
  return m;
}

__device__ __forceinline__ uint16_t after_flags(float t, double hi, double lo) {
  double d = (double)t;
  uint16_t m = 0;
  if (d == 0.0) m |= XB_MASK_BAD_OUTPUT | XB_MASK_BAD_INPUT | XB_MASK_BAD_PIX_VAL;
  if (d >= hi) m |= XB_MASK_BAD_OUTPUT | XB_MASK_BAD_INPUT | XB_MASK_SAT_PIXEL;
  if (d <= lo) m |= XB_MASK_BAD_OUTPUT | XB_MASK_BAD_INPUT | XB_MASK_LOW_PIXEL;
  return m;
}

// The convolution: one wave per 21 x 21 output tile, NT tiles per workgroup
// (1, 2, 4 or 8: a launch shape the batch autotuner picks; every shape gives
// the same bits).  The tile's 41 x 41 input window is staged in LDS (FP32, the same
// clamped reads as the reference's edge handling).  Lane = (output row y,
// column group g): 21 rows x 3 groups of CG = 7 adjacent columns, so each lane
// computes 7 outputs of one row (lane 63 idles).  Output (x, y) sums
// K[20-d][20-i] * T[y-10+d][x-10+i] over d, then i, ascending (the
// reference's term order: window rows, then columns, ascending; multiply and
// add unfused).  The tile kernel is the same for every lane of the wave: its
// rows are read by scalar loads into SGPRs (the constant address space) and
// each product takes its coefficient from there.  Along a window row the 7
// outputs share their inputs: tap i of output c reads window column
// 7g + c + i, so each step i reads one new window value (widened to FP64
// once, exact) into a 7-slot register ring and applies it to all 7
// accumulators: 27 LDS reads and 147 products per lane and kernel row.
constexpr int CTW = KW + 2 * HKW;  // 41: the tile's input window
constexpr int CTP = CTW;           // its LDS pitch (odd: rows fall on distinct banks)
constexpr int CG = 7;              // output columns per lane
static_assert(3 * CG == KW, "three column groups cover a tile row");

typedef const __attribute__((address_space(4))) double* KConst;

template <bool FUSED>
__device__ __forceinline__ void conv_row_group(KConst K, const float (*win)[CTP], int y, int xc0,
                                               double (&acc)[CG]) {
#pragma unroll
  for (int c = 0; c < CG; c++) acc[c] = 0.0;
  for (int d = 0; d <= 2 * HKW; d++) {
    const float* src = win[y + d] + xc0;
    const KConst Kr = K + (2 * HKW - d) * KW;
    double t[CG];  // window column xc0 + j sits in slot j % CG
#pragma unroll
    for (int j = 0; j + 1 < CG; j++) t[j] = (double)src[j];
#pragma unroll
    for (int i = 0; i < KW; i++) {
      t[(i + CG - 1) % CG] = (double)src[i + CG - 1];
      const double k = Kr[2 * HKW - i];
      if constexpr (FUSED) {
#pragma unroll
        for (int c = 0; c < CG; c++) acc[c] = fma(k, t[(i + c) % CG], acc[c]);
      } else {
        double pr[CG];
#pragma unroll
        for (int c = 0; c < CG; c++) pr[c] = k * t[(i + c) % CG];
#pragma unroll
        for (int c = 0; c < CG; c++) acc[c] = acc[c] + pr[c];
      }
    }
  }
}

// FUSED (the mixed form): each term a fused multiply-add, in the same order;
// otherwise the reference's unfused multiply, then add.
template <bool FUSED, int NT>
__global__ __launch_bounds__(WAVE * NT) void k_conv(ConvArgs A) {
  __shared__ float win[NT][CTW][CTP];                 // each tile's input window
  __shared__ unsigned long long rowFlag[NT][CTW];     // bit c: window pixel (l, c) flagged
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / WAVE);
  const int lane = threadIdx.x % WAVE;
  const int xs = blockIdx.x * NT + wave, ys = blockIdx.y;
  const int x0 = HKW + xs * KW, y0t = HKW + ys * KW;
  const int w = A.w, h = A.h;
  const bool tileOn = xs < A.xSteps && x0 < w - HKW && y0t < h - HKW;
  const int tile = xs + ys * A.xSteps;
  float(*W)[CTP] = win[wave];
  if (tileOn) {
    // the window, clamped as the reference's reads: row min(y, h - 1),
    // column min(x, w - 1) (both start inside the image)
    // all of a lane's loads in flight together
    constexpr int NS = (CTW * CTW + WAVE - 1) / WAVE;
    float v[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const int e = min(lane + k * WAVE, CTW * CTW - 1);
      const int l = e / CTW, c = e % CTW;
      const int yy = min(y0t - HKW + l, h - 1), xx = min(x0 - HKW + c, w - 1);
      v[k] = A.T[(size_t)yy * w + xx];
    }
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const int e = lane + k * WAVE;
      if (e < CTW * CTW) W[e / CTW][e % CTW] = v[k];
    }
  }
  if (!tileOn) return;
  // each wave reads only its own tile's LDS, and a wave's LDS accesses run
  // in program order: a wave barrier (no workgroup barrier) orders them
  __builtin_amdgcn_wave_barrier();
  // flag masks of the window rows: lane l < 41 tests row l
  {
    const float* row = W[min(lane, CTW - 1)];
    unsigned long long m = 0;
#pragma unroll
    for (int c = 0; c < CTW; c++)
      if (conv_flags(row[c], A.hi, A.lo) != 0) m |= 1ull << c;
    if (lane < CTW) rowFlag[wave][lane] = m;
  }
  const int y = min(lane / 3, KW - 1), g = lane % 3;  // lane 63: row 21, idle
  const int xc0 = g * CG;
  const KConst K = (KConst)(A.tileK64 + (size_t)tile * KK);  // uniform: scalar loads
  const int yo = y0t + y;
  // the target pixels of the lane's outputs, read before the products
  float sv[CG];
#pragma unroll
  for (int c = 0; c < CG; c++) {
    const int x = min(x0 + xc0 + c, w - 1);
    sv[c] = A.S[(size_t)min(yo, h - 1) * w + x];
  }
  double acc[CG];
  conv_row_group<FUSED>(K, W, y, xc0, acc);
  __builtin_amdgcn_wave_barrier();  // rowFlag, written by lanes 0..40 above
  if (lane >= 3 * KW) return;
  if (yo >= h - HKW) return;
  // flagged pixels anywhere in the lane's 21 window rows
  unsigned long long vOr = 0;
  for (int jj = 0; jj <= 2 * HKW; jj++) vOr |= rowFlag[wave][y + jj];
  const double ks = *A.kernSum;
  const double inv = A.scaleConv ? 1.0 / ks : 1.0;
  // sub (cl_kern/sub.cl:15-32, src/bach.cpp:415-440) on the unrounded FP64
  // convolution, as the reference's sub reads its FP64 conv buffer
  const double convFactor = A.scaleConv ? ks : 1.0;
  const double finalFactor = A.scaleConv ? -(1.0 / ks) : 1.0;
  const double aks = A.tileAks[tile];
  const double thr = (double)0.99f;
#pragma unroll
  for (int c = 0; c < CG; c++) {
    const int xo = xc0 + c;  // output column in the tile
    const int x = x0 + xo;
    if (x >= w - HKW) break;
    const size_t id = (size_t)yo * w + x;
    double v = acc[c];
    v = v + background_c(A.sol, x, yo, w, h);
    v = v * inv;
    A.out[id] = (float)v;
    const uint16_t cm = conv_flags(W[y + HKW][xo + HKW], A.hi, A.lo);
    uint16_t nm = cm;
    if (cm & XB_MASK_BAD_INPUT) nm |= XB_MASK_BAD_OUTPUT;
    if ((vOr >> xo) & ((1ull << KW) - 1ull)) {
      // sum |k| over unflagged pixels in the reference's loop order (rows
      // j ascending, columns i ascending); a flagged pixel adds +0.0, which
      // leaves the sum (never -0) unchanged
      double uks = 0.0;
      for (int jj = 0; jj <= 2 * HKW; jj++) {
        const unsigned long long m = rowFlag[wave][y + jj] >> xo;
        const KConst Kr = K + (2 * HKW - jj) * KW;
        for (int ii = 0; ii <= 2 * HKW; ii++)
          uks = uks + (((m >> ii) & 1ull) ? 0.0 : fabs(Kr[2 * HKW - ii]));
      }
      if ((uks / aks) < thr) nm |= XB_MASK_BAD_OUTPUT | XB_MASK_BAD_CONV;
      else nm |= XB_MASK_OK_CONV;
    }
    const uint16_t fm = nm | after_flags(sv[c], A.hi, A.lo);
    A.mask[id] = fm;
    double dd = 1e-30;
    if ((fm & XB_MASK_BAD_OUTPUT) == 0) dd = (v * convFactor - (double)sv[c]) * finalFactor;
    A.diff[id] = (float)dd;
  }
}

// border pixels: conv and difference = 1e-30, mask keeps its value plus
// maskAfterConv flags.
// NaN marks held back by a clip-refit loop that ran before the direction was
// known become XB_MASK_NAN_PIXEL for the chosen direction (pend) and vanish for
// the other; k_conv rewrites every interior pixel's mask from scratch, so the
// border is the only place they can remain.
__device__ __forceinline__ void border_pixel(const float* __restrict__ S, uint16_t* __restrict__ mask,
                                             float* __restrict__ out, float* __restrict__ diff,
                                             size_t id, double hi, double lo, uint16_t pend) {
  out[id] = 1e-30f;
  diff[id] = 1e-30f;
  uint16_t m = mask[id];
  if (m & (mask_pend(0) | mask_pend(1)))
    m = (uint16_t)((m & XB_MASK_ALL) | ((m & pend) ? XB_MASK_NAN_PIXEL : 0));
  mask[id] = m | after_flags(S[id], hi, lo);
}

// Index e over the border pixels only (w, h > 2 HKW): the HKW top rows, the
// HKW bottom rows, then the HKW left and HKW right pixels of each row between.
__global__ __launch_bounds__(256) void k_conv_border(const float* __restrict__ S,
                                                     uint16_t* __restrict__ mask,
                                                     float* __restrict__ out,
                                                     float* __restrict__ diff, int w, int h,
                                                     double hi, double lo, uint16_t pend) {
  const size_t band = (size_t)HKW * w;
  const size_t n = 2 * band + (size_t)2 * HKW * (h - 2 * HKW);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    int x, y;
    if (e < band) {
      y = (int)(e / w);
      x = (int)(e % w);
    } else if (e < 2 * band) {
      const size_t j = e - band;
      y = h - HKW + (int)(j / w);
      x = (int)(j % w);
    } else {
      const size_t j = e - 2 * band;
      y = HKW + (int)(j / (2 * HKW));
      const int k = (int)(j % (2 * HKW));
      x = k < HKW ? k : w - 2 * HKW + k;
    }
    border_pixel(S, mask, out, diff, (size_t)y * w + x, hi, lo, pend);
  }
}

// The same over every pixel, skipping the interior (images of at most
// 2 HKW rows or columns, where the border covers all of it).
__global__ __launch_bounds__(256) void k_conv_border_all(const float* __restrict__ S,
                                                         uint16_t* __restrict__ mask,
                                                         float* __restrict__ out,
                                                         float* __restrict__ diff, int w, int h,
                                                         double hi, double lo, uint16_t pend) {
  const size_t n = (size_t)w * h;
  for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n;
       id += (size_t)gridDim.x * blockDim.x) {
    int x = (int)(id % w), y = (int)(id / w);
    bool interior = x >= HKW && x < w - HKW && y >= HKW && y < h - HKW;
    if (interior) continue;
    border_pixel(S, mask, out, diff, id, hi, lo, pend);
  }
}

// Diagnostics of the difference image (the quality figures HOTPANTS writes to
// its header, tests/test01_sub.fits): over good pixels (interior, no
// BAD_OUTPUT) and over OK pixels (good and OK_CONV), the count, sum and sum of
// squares of d, the sum of the noise sqrt(|conv| + |S|) — the per-pixel noise
// of the reference's chi2 (cl_kern/cd.cl:578) — and the sum of (d / noise)^2.
// Block partials in FP64, added in block order by k_diag_final.
__global__ __launch_bounds__(256) void k_diag(const float* __restrict__ D,
                                              const float* __restrict__ conv,
                                              const float* __restrict__ S,
                                              const uint16_t* __restrict__ mask, int w, int h,
                                              double* __restrict__ part) {
  double q[10];
#pragma unroll
  for (int k = 0; k < 10; k++) q[k] = 0.0;
  const size_t n = (size_t)w * h;
  for (size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x; id < n;
       id += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(id % w), y = (int)(id / w);
    if (x < HKW || x >= w - HKW || y < HKW || y >= h - HKW) continue;
    const uint16_t m = mask[id];
    if (m & XB_MASK_BAD_OUTPUT) continue;
    const double d = (double)D[id];
    const double nz = sqrt(fabs((double)conv[id]) + fabs((double)S[id]));
    const double r = nz > 0.0 ? d / nz : 0.0;
    q[0] += 1.0; q[1] += d; q[2] += d * d; q[3] += nz; q[4] += r * r;
    if (m & XB_MASK_OK_CONV) {
      q[5] += 1.0; q[6] += d; q[7] += d * d; q[8] += nz; q[9] += r * r;
    }
  }
  __shared__ double red[10][256];
#pragma unroll
  for (int k = 0; k < 10; k++) red[k][threadIdx.x] = q[k];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
#pragma unroll
      for (int k = 0; k < 10; k++) red[k][threadIdx.x] += red[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 10) part[(size_t)blockIdx.x * 10 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_diag_final(const double* __restrict__ part, int nb, double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= 10) return;
  double a = 0.0;
  for (int b = 0; b < nb; b++) a += part[(size_t)b * 10 + k];
  out[k] = a;
}

hipError_t launch_diag(Ctx& c, double* d_part, double* d_out) {
  hipLaunchKernelGGL(k_diag, dim3(DIAG_BLOCKS), dim3(256), 0, c.stream, c.diff, c.conv,
                     c.img[1 - c.tIdx], c.mask, c.w, c.h, d_part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, c.stream, d_part, DIAG_BLOCKS, d_out);
  return hipGetLastError();
}

template <bool FUSED, int NT>
void launch_conv_shape(Ctx& c, const ConvArgs& A, int xSteps, int ySteps) {
  hipLaunchKernelGGL((k_conv<FUSED, NT>), dim3((xSteps + NT - 1) / NT, ySteps), dim3(WAVE * NT),
                     0, c.cur, A);
}

// tiles per workgroup from the context's launch shapes (checked when set)
template <bool FUSED>
void launch_conv_nt(Ctx& c, const ConvArgs& A, int xSteps, int ySteps) {
  switch (c.shapes.conv_tiles) {
    case 1: launch_conv_shape<FUSED, 1>(c, A, xSteps, ySteps); break;
    case 2: launch_conv_shape<FUSED, 2>(c, A, xSteps, ySteps); break;
    case 8: launch_conv_shape<FUSED, 8>(c, A, xSteps, ySteps); break;
    default: launch_conv_shape<FUSED, 4>(c, A, xSteps, ySteps); break;
  }
}

hipError_t launch_conv(Ctx& c) {
  const int xSteps = (c.w + KW - 1) / KW, ySteps = (c.h + KW - 1) / KW;
  const int nTiles = xSteps * ySteps;
  const int ti = c.tIdx;
  double* kernSum = &c.sc->kernSum;
  const double* sol = c.ksSol + (size_t)ti * NSOL;  // the chosen direction's fit
  if (c.form & XB_FORM_MIXED)
    hipLaunchKernelGGL(k_tile_kernels_mfma, dim3((nTiles + 1 + 15) / 16), dim3(256), 0, c.cur,
                       c.basis, sol, c.w, c.h, xSteps, nTiles, c.tileK64, c.tileAks, kernSum,
                       &c.hscDev->kernSum);
  else
    hipLaunchKernelGGL(k_tile_kernels, dim3((nTiles + 1 + TK_TILES - 1) / TK_TILES), dim3(512),
                       0, c.cur, c.basis, sol, c.w, c.h, xSteps, nTiles, c.tileK64, c.tileAks,
                       kernSum, &c.hscDev->kernSum);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  ConvArgs A;
  A.T = c.img[ti];
  A.S = c.img[1 - ti];
  A.mask = c.mask;
  A.out = c.conv;
  A.diff = c.diff;
  A.tileK64 = c.tileK64;
  A.tileAks = c.tileAks;
  A.sol = sol;
  A.kernSum = kernSum;
  A.w = c.w;
  A.h = c.h;
  A.xSteps = xSteps;
  A.yTiles = ySteps;
  A.scaleConv = (c.m.normalizeTemplate && c.convTemplate) || (!c.m.normalizeTemplate && !c.convTemplate);
  A.hi = c.m.threshHigh;
  A.lo = c.m.threshLow;
  prof_begin(c, XB_PROF_CONV);
  if (c.form & XB_FORM_MIXED) launch_conv_nt<true>(c, A, xSteps, ySteps);
  else launch_conv_nt<false>(c, A, xSteps, ySteps);
  e = hipGetLastError();
  prof_end(c, XB_PROF_CONV);
  if (e != hipSuccess) return e;
  if (c.w > 2 * HKW && c.h > 2 * HKW) {
    const size_t nb = (size_t)2 * HKW * c.w + (size_t)2 * HKW * (c.h - 2 * HKW);
    const int blocks = (int)std::min<size_t>((nb + 255) / 256, 1024);
    hipLaunchKernelGGL(k_conv_border, dim3(blocks), dim3(256), 0, c.cur, c.img[1 - ti], c.mask,
                       c.conv, c.diff, c.w, c.h, c.m.threshHigh, c.m.threshLow, mask_pend(ti));
  } else {
    hipLaunchKernelGGL(k_conv_border_all, dim3(1024), dim3(256), 0, c.cur, c.img[1 - ti],
                       c.mask, c.conv, c.diff, c.w, c.h, c.m.threshHigh, c.m.threshLow,
                       mask_pend(ti));
  }
  return hipGetLastError();
}

}  // namespace xb
