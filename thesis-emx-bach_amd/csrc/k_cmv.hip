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
// k_fill_gram: the 52x52 Gram matrix G = W^T W of the final rows, one wave
// per 16x16 tile, staged through LDS in 64-pixel chunks, each entry a
// sequential sum over the 961 pixels (cmv.cl:195-215 order).
// k_fill_target: one workgroup per stamp forms the target dot products:
// column-major order for all 52 rows (createB, cmv.cl:217-248, and the ksc
// right-hand side, src/cdkscUtil.cpp:459-470) and row-major order for the 3
// background rows (cd right-hand side, cl_kern/cd.cl:318-344).  Q (51x51,
// 1-based) is G shifted by one.
#include "xb_dev.h"

namespace xb {

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

__global__ __launch_bounds__(256) void k_fill_target(FillArgs A) {
  int s, ssx, ssy;
  if (!fill_stamp(A, blockIdx.x, &s, &ssx, &ssy)) return;
  const int g = A.kept[s];
  const bool valid = A.cur[s] < A.count[g];
  const int tid = threadIdx.x;
  const double* W = A.W + (size_t)s * WROWS * NPIX;
  const int* evenSub = A.basis->evenSub;
  // target dot products.  tc[c]: the target at column-major position c
  // (x = c / 31, y = c % 31); tr[k]: at row-major position k.
  __shared__ double tc[NPIX], tr[NPIX];
  __shared__ double mono[2][SSW];  // xf by column, yf by row
  __shared__ int rsub[WROWS];
  if (!valid) {
    if (tid < WROWS) A.Bc[(size_t)s * WROWS + tid] = 0.0;
    if (tid < NBG) A.Br[(size_t)s * NBG + tid] = 0.0;
    return;
  }
  for (int i = tid; i < NPIX; i += blockDim.x) {
    const int xx = i % SSW - HSS, yy = i / SSW - HSS;
    const int xr = min(max(ssx + xx, 0), A.w - 1), yr = min(max(ssy + yy, 0), A.h - 1);
    const double v = (double)A.img2[(size_t)xr + (size_t)yr * A.w];
    tr[i] = v;
    tc[(i % SSW) * SSW + i / SSW] = v;
  }
  if (tid < SSW) {
    // the background monomials of k_fill_conv, same FP32 expressions
    const float hw = (float)A.w * 0.5f, hh = (float)A.h * 0.5f;
    mono[0][tid] = (double)(((float)(ssx - HSS + tid) - hw) / hw);
    mono[1][tid] = (double)(((float)(ssy - HSS + tid) - hh) / hh);
  }
  if (tid < WROWS) rsub[tid] = tid < NPSF && evenSub[tid];
  __syncthreads();
  // Bc: column-major pixel order (x outer, y inner): chunks of 64 such
  // positions staged through LDS, one thread per row summing in order.
  __shared__ double wc[WROWS][GK + 1];
  double acc = 0.0;
  for (int c0 = 0; c0 < NPIX; c0 += GK) {
    const int kn = min(GK, NPIX - c0);
    for (int i = tid; i < WROWS * GK; i += blockDim.x) {
      const int r = i / GK, cc = i % GK;
      double v = 0.0;
      if (cc < kn) {
        const int c = c0 + cc, k = (c / SSW) + SSW * (c % SSW);
        v = W[(size_t)r * NPIX + k];
        if (rsub[r]) v = v - W[k];
      }
      wc[r][cc] = v;
    }
    __syncthreads();
    if (tid < WROWS) {
      const double* t = tc + c0;
      int cc = 0;
      for (; cc + 4 <= kn; cc += 4) {
        const double p0 = wc[tid][cc] * t[cc], p1 = wc[tid][cc + 1] * t[cc + 1];
        const double p2 = wc[tid][cc + 2] * t[cc + 2], p3 = wc[tid][cc + 3] * t[cc + 3];
        acc += p0;
        acc += p1;
        acc += p2;
        acc += p3;
      }
      for (; cc < kn; cc++) acc += wc[tid][cc] * t[cc];
    }
    __syncthreads();
  }
  if (tid < WROWS) A.Bc[(size_t)s * WROWS + tid] = acc;
  // Br: row-major order over the background rows 1, yf, xf
  if (tid >= 64 && tid < 64 + NBG) {
    const int b = tid - 64;
    double a2 = 0.0;
    for (int y = 0; y < SSW; y++) {
      const double yf = b == 1 ? mono[1][y] : 1.0;
      for (int x = 0; x < SSW; x++) {
        const double bgv = (b == 2 ? mono[0][x] : 1.0) * yf;
        a2 += bgv * tr[y * SSW + x];
      }
    }
    A.Br[(size_t)s * NBG + b] = a2;
  }
}

constexpr int NGT = 10;  // lower-triangular 16x16 tiles of the 52x52 Gram matrix
constexpr int GP = GT + 2;  // LDS pitch: [pixel][row], 16-byte aligned row pairs

// One wave per (stamp, tile).  Blocks b and b + 8 run on the same XCD (blocks
// are dealt to the 8 XCDs in turn), so each XCD receives a contiguous range of
// (stamp, tile) work and a stamp's W is read through a single L2.  Lane
// (lr, lc) = (lane / 8, lane % 8) owns the 2x2 entries rows bi*16 + 2lr + i,
// columns bj*16 + 2lc + j.  Pixel chunks are double-buffered: the next chunk
// is loaded into registers while the current one is summed from LDS.
__global__ __launch_bounds__(64) void k_fill_gram(FillArgs A, int nMax) {
  const int total = nMax * NGT;
  const int per = (total + 7) / 8;
  const int wk = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (wk >= total) return;
  int s, ssx, ssy;
  if (!fill_stamp(A, wk / NGT, &s, &ssx, &ssy)) return;
  const int g = A.kept[s];
  const bool valid = A.cur[s] < A.count[g];
  const int lane = threadIdx.x;
  int t = wk % NGT, bi = 0;
  while (t > bi) { t -= bi + 1; bi++; }
  const int bj = t;
  if (!valid) {
    for (int e = lane; e < GT * GT; e += WAVE) {
      const int row = bi * GT + e / GT, col = bj * GT + e % GT;
      if (row < WROWS && col <= row) A.G[((size_t)s * WROWS + row) * WROWS + col] = 0.0;
    }
    return;
  }
  const double* W = A.W + (size_t)s * WROWS * NPIX;
  const int* evenSub = A.basis->evenSub;
  __shared__ __align__(16) double Wa[2][GK][GP];
  __shared__ __align__(16) double Wb[2][GK][GP];
  // per-lane staging: pixel k0 + lane of rows bi*16 + rr and bj*16 + rr
  bool subA[GT], subB[GT];
#pragma unroll
  for (int rr = 0; rr < GT; rr++) {
    const int ga = bi * GT + rr, gb = bj * GT + rr;
    subA[rr] = ga < NPSF && evenSub[ga];
    subB[rr] = gb < NPSF && evenSub[gb];
  }
  double va[GT], vb[GT];
  auto fetch = [&](int k0) {
    const int k = k0 + lane;
    const bool in = k < NPIX;
    const double w0 = in ? W[k] : 0.0;
#pragma unroll
    for (int rr = 0; rr < GT; rr++) {
      const int ga = bi * GT + rr, gb = bj * GT + rr;
      double x = (in && ga < WROWS) ? W[(size_t)ga * NPIX + k] : 0.0;
      double y = (in && gb < WROWS) ? W[(size_t)gb * NPIX + k] : 0.0;
      if (subA[rr]) x = x - w0;
      if (subB[rr]) y = y - w0;
      va[rr] = x;
      vb[rr] = y;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int rr = 0; rr < GT; rr++) {
      Wa[buf][lane][rr] = va[rr];
      Wb[buf][lane][rr] = vb[rr];
    }
  };
  const int lr = lane / 8, lc = lane % 8;
  double c00 = 0.0, c01 = 0.0, c10 = 0.0, c11 = 0.0;
  fetch(0);
  store(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < NPIX; k0 += GK) {
    const int kn = min(GK, NPIX - k0);
    const bool more = k0 + GK < NPIX;
    if (more) fetch(k0 + GK);
    for (int kk = 0; kk < kn; kk++) {
      const double2 x = *reinterpret_cast<const double2*>(&Wa[buf][kk][2 * lr]);
      const double2 y = *reinterpret_cast<const double2*>(&Wb[buf][kk][2 * lc]);
      const double p00 = x.x * y.x, p01 = x.x * y.y, p10 = x.y * y.x, p11 = x.y * y.y;
      c00 += p00;
      c01 += p01;
      c10 += p10;
      c11 += p11;
    }
    if (more) {
      store(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
  const int r0 = bi * GT + 2 * lr, q0 = bj * GT + 2 * lc;
  const double cv[2][2] = {{c00, c01}, {c10, c11}};
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int row = r0 + i, col = q0 + j;
      if (row < WROWS && col <= row) A.G[((size_t)s * WROWS + row) * WROWS + col] = cv[i][j];
    }
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
  A.nList = d_list ? &c.sc->nList[im] : &c.sc->nKept[im];
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
  if (c.form & XB_FORM_MIXED) {
    const hipError_t em = launch_fill_mixed(c, A, nMax);
    prof_end(c, XB_PROF_FILL);
    return em;
  }
  hipLaunchKernelGGL(k_fill_conv, dim3(nMax, WROWS), dim3(256), 0, c.cur, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill_gram, dim3(8 * ((nMax * NGT + 7) / 8)), dim3(WAVE), 0, c.cur, A,
                     nMax);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill_target, dim3(nMax), dim3(256), 0, c.cur, A);
  prof_end(c, XB_PROF_FILL);
  return hipGetLastError();
}

}  // namespace xb
