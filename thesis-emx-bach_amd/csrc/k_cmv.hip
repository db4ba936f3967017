// k_cmv.hip — basis convolution of substamps and their normal equations (cmv).
//
// k_fill_conv: one workgroup per (stamp, W row).  The 51x51 image window of
// the stamp's current substamp is staged in LDS; the Y pass (FP32 accumulator,
// FP64 taps, cl_kern/cmv.cl:70-106) and the X pass (FP64, cmv.cl:108-135) run
// from LDS and the raw W row is written once.  Background rows are the
// monomials of cmv.cl:156-193 (FP32 xf, yf).  Even-even rows keep the raw
// value; every consumer applies W[n] - W[0] on load (cmv.cl:137-154), a single
// subtraction exactly as the reference does in place.
//
// k_fill_gram: the 52x52 Gram matrix G = W^T W of the final rows, 16x16 tiles
// staged through LDS in 64-pixel chunks, each entry a sequential sum over the
// 961 pixels (cmv.cl:195-215 order); one extra workgroup per stamp forms the
// target dot products: column-major order for all 52 rows (createB,
// cmv.cl:217-248, and the ksc right-hand side, src/cdkscUtil.cpp:459-470) and
// row-major order for the 3 background rows (cd right-hand side,
// cl_kern/cd.cl:318-344).  Q (51x51, 1-based) is G shifted by one.
#include "xb_dev.h"

namespace xb {

struct FillArgs {
  const float* img;   // image being convolved
  const float* img2;  // target image
  int w, h;
  const int* list;    // stamp list (kept indices) or nullptr for identity
  const int* nList;   // device count of the list
  const int* kept;
  const int* cur;
  const int* count;
  const int* ssx;
  const int* ssy;
  const Basis* basis;
  double* W;
  double* G;
  double* Bc;
  double* Br;
};

__device__ __forceinline__ bool fill_stamp(const FillArgs& A, int k, int* s, int* x, int* y) {
  int n = *A.nList;
  if (k >= n) return false;
  int st = A.list ? A.list[k] : k;
  *s = st;
  int g = A.kept[st];
  int c = A.cur[st];
  if (c < A.count[g]) {
    *x = A.ssx[g * MAXSS + c];
    *y = A.ssy[g * MAXSS + c];
    return true;
  }
  *x = 0;
  *y = 0;
  return true;
}

__global__ __launch_bounds__(256) void k_fill_conv(FillArgs A) {
  int s, ssx, ssy;
  if (!fill_stamp(A, blockIdx.x, &s, &ssx, &ssy)) return;
  const int g = A.kept[s];
  const bool valid = A.cur[s] < A.count[g];
  const int n = blockIdx.y;
  const int tid = threadIdx.x;
  double* Wrow = A.W + ((size_t)s * WROWS + n) * NPIX;
  if (!valid) {
    for (int p = tid; p < NPIX; p += blockDim.x) Wrow[p] = 0.0;
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
      Wrow[p] = ax * ay;
    }
    return;
  }
  __shared__ float win[TW][TW + 1];
  __shared__ float tmp[TH * TW];
  __shared__ double fx[KW], fy[KW];
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
    Wrow[p] = w0;
  }
}

constexpr int GT = 16;   // Gram tile
constexpr int GK = 64;   // pixel chunk

__global__ __launch_bounds__(256) void k_fill_gram(FillArgs A) {
  int s, ssx, ssy;
  if (!fill_stamp(A, blockIdx.x, &s, &ssx, &ssy)) return;
  const int g = A.kept[s];
  const bool valid = A.cur[s] < A.count[g];
  const int tid = threadIdx.x;
  const double* W = A.W + (size_t)s * WROWS * NPIX;
  const int* evenSub = A.basis->evenSub;
  if (blockIdx.y == 10) {
    // target dot products
    __shared__ float tgt[SSW * SSW];
    if (valid)
      for (int i = tid; i < NPIX; i += blockDim.x) {
        int xx = i % SSW - HSS, yy = i / SSW - HSS;
        int xr = min(max(ssx + xx, 0), A.w - 1), yr = min(max(ssy + yy, 0), A.h - 1);
        tgt[i] = A.img2[(size_t)xr + (size_t)yr * A.w];
      }
    __syncthreads();
    // Bc: column-major pixel order (x outer, y inner): chunks of 64 such
    // positions staged through LDS, one thread per row summing in order.
    __shared__ double wc[WROWS][GK + 1];
    double acc = 0.0;
    const bool sub = tid < NPSF && evenSub[tid];
    for (int c0 = 0; c0 < NPIX && valid; c0 += GK) {
      const int kn = min(GK, NPIX - c0);
      for (int i = tid; i < WROWS * GK; i += blockDim.x) {
        int r = i / GK, cc = i % GK;
        double v = 0.0;
        if (cc < kn) {
          int c = c0 + cc, k = (c / SSW) + SSW * (c % SSW);
          v = W[(size_t)r * NPIX + k];
          if (r < NPSF && evenSub[r]) v = v - W[k];
        }
        wc[r][cc] = v;
      }
      __syncthreads();
      if (tid < WROWS)
        for (int cc = 0; cc < kn; cc++) {
          int c = c0 + cc, k = (c / SSW) + SSW * (c % SSW);
          acc += wc[tid][cc] * (double)tgt[k];
        }
      __syncthreads();
    }
    (void)sub;
    if (tid < WROWS) A.Bc[(size_t)s * WROWS + tid] = valid ? acc : 0.0;
    // Br: row-major order over the background rows, whose values are the
    // monomials 1, yf, xf written by k_fill_conv (recomputed identically).
    if (tid >= 64 && tid < 64 + NBG) {
      const int b = tid - 64;
      const float hw = (float)A.w * 0.5f, hh = (float)A.h * 0.5f;
      double a2 = 0.0;
      if (valid)
        for (int k = 0; k < NPIX; k++) {
          int x = ssx - HSS + k % SSW, y = ssy - HSS + k / SSW;
          double xf = (double)(((float)x - hw) / hw);
          double yf = (double)(((float)y - hh) / hh);
          double bgv = (b == 2 ? xf : 1.0) * (b == 1 ? yf : 1.0);
          a2 += bgv * (double)tgt[k];
        }
      A.Br[(size_t)s * NBG + b] = a2;
    }
    return;
  }
  // lower-triangular tile pair (bi, bj), bj <= bi, over 4 row blocks
  int t = blockIdx.y, bi = 0;
  while (t > bi) { t -= bi + 1; bi++; }
  const int bj = t;
  __shared__ double ra[GT][GK + 1], rb[GT][GK + 1];
  __shared__ double r0[GK];
  const int tr = tid / GT, tc = tid % GT;
  const int row = bi * GT + tr, col = bj * GT + tc;
  const bool act = row < WROWS && col <= row;
  double acc = 0.0;
  for (int k0 = 0; k0 < NPIX; k0 += GK) {
    const int kn = min(GK, NPIX - k0);
    for (int i = tid; i < GK; i += blockDim.x) r0[i] = i < kn ? W[k0 + i] : 0.0;
    __syncthreads();
    for (int i = tid; i < GT * GK; i += blockDim.x) {
      int rr = i / GK, kk = i % GK;
      double va = 0.0, vb = 0.0;
      int ga = bi * GT + rr, gb = bj * GT + rr;
      if (kk < kn) {
        if (ga < WROWS) {
          va = W[(size_t)ga * NPIX + k0 + kk];
          if (ga < NPSF && evenSub[ga]) va = va - r0[kk];
        }
        if (gb < WROWS) {
          vb = W[(size_t)gb * NPIX + k0 + kk];
          if (gb < NPSF && evenSub[gb]) vb = vb - r0[kk];
        }
      }
      ra[rr][kk] = va;
      rb[rr][kk] = vb;
    }
    __syncthreads();
    if (act)
      for (int kk = 0; kk < kn; kk++) acc += ra[tr][kk] * rb[tc][kk];
    __syncthreads();
  }
  if (act) A.G[((size_t)s * WROWS + row) * WROWS + col] = valid ? acc : 0.0;
}

hipError_t launch_fill(Ctx& c, int im, const int* d_list, int nMax) {
  if (nMax <= 0) return hipSuccess;
  const int src = im, tgt = 1 - im;
  FillArgs A;
  A.img = c.img[src];
  A.img2 = c.img[tgt];
  A.w = c.w;
  A.h = c.h;
  A.list = d_list;
  A.nList = d_list ? &c.sc->nList : &c.sc->nKept[im];
  A.kept = c.st[im].kept;
  A.cur = c.st[im].cur;
  A.count = c.st[im].count;
  A.ssx = c.st[im].ssx;
  A.ssy = c.st[im].ssy;
  A.basis = c.basis;
  A.W = c.st[im].W;
  A.G = c.st[im].G;
  A.Bc = c.st[im].Bc;
  A.Br = c.st[im].Br;
  prof_begin(c, XB_PROF_FILL);
  hipLaunchKernelGGL(k_fill_conv, dim3(nMax, WROWS), dim3(256), 0, c.stream, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill_gram, dim3(nMax, 11), dim3(256), 0, c.stream, A);
  prof_end(c, XB_PROF_FILL);
  return hipGetLastError();
}

}  // namespace xb
