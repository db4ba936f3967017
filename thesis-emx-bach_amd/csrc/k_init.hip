// k_init.hip — input masking (the reference's init stage).
//
// One pass replaces cl_kern/ini.cl:15-35 (maskInput) and 37-66 (spreadMask):
// each 64x16 output tile loads its T/S pixels plus a halo, forms the BAD_INPUT
// plane in LDS, ORs it separably over the clamped (2*w2+1)^2 window and writes
// the final 16-bit mask once.  HBM traffic: 8 B in + 2 B out per pixel.
#include "xb_dev.h"

namespace xb {

constexpr int MT_X = 64, MT_Y = 16, MAX_W2 = 8;

__global__ __launch_bounds__(256) void k_mask_input(const float* __restrict__ T,
                                                    const float* __restrict__ S,
                                                    uint16_t* __restrict__ mask, int w, int h,
                                                    int border, double hi, double lo, int w2) {
  __shared__ uint8_t bad[MT_Y + 2 * MAX_W2][MT_X + 2 * MAX_W2];
  __shared__ uint8_t hor[MT_Y + 2 * MAX_W2][MT_X];
  __shared__ uint16_t own[MT_Y][MT_X];
  const int x0 = blockIdx.x * MT_X, y0 = blockIdx.y * MT_Y;
  const int lw = MT_X + 2 * w2, lh = MT_Y + 2 * w2;
  for (int i = threadIdx.x; i < lw * lh; i += blockDim.x) {
    int ly = i / lw, lx = i % lw;
    int x = x0 + lx - w2, y = y0 + ly - w2;
    uint16_t m = 0;
    if (x >= 0 && x < w && y >= 0 && y < h) {
      size_t id = (size_t)y * w + x;
      double t = (double)T[id], s = (double)S[id];
      if (t == 0.0 || s == 0.0) m |= XB_MASK_BAD_INPUT | XB_MASK_BAD_PIX_VAL;
      if (t >= hi || s >= hi) m |= XB_MASK_BAD_INPUT | XB_MASK_SAT_PIXEL;
      if (t <= lo || s <= lo) m |= XB_MASK_BAD_INPUT | XB_MASK_LOW_PIXEL;
      if (x < border || x >= w - border || y < border || y >= h - border)
        m |= XB_MASK_BAD_PIXEL_S | XB_MASK_BAD_PIXEL_T;
    }
    bad[ly][lx] = (m & XB_MASK_BAD_INPUT) ? 1 : 0;
    int ox = lx - w2, oy = ly - w2;
    if (ox >= 0 && ox < MT_X && oy >= 0 && oy < MT_Y) own[oy][ox] = m;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < lh * MT_X; i += blockDim.x) {
    int ly = i / MT_X, ox = i % MT_X;
    uint8_t a = 0;
    for (int d = 0; d <= 2 * w2; d++) a |= bad[ly][ox + d];
    hor[ly][ox] = a;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < MT_Y * MT_X; i += blockDim.x) {
    int oy = i / MT_X, ox = i % MT_X;
    int x = x0 + ox, y = y0 + oy;
    if (x >= w || y >= h) continue;
    uint16_t m = own[oy][ox];
    if ((m & XB_MASK_BAD_INPUT) == 0) {
      uint8_t a = 0;
      for (int d = 0; d <= 2 * w2; d++) a |= hor[oy + d][ox];
      if (a) m |= XB_MASK_OK_CONV;
    }
    mask[(size_t)y * w + x] = m;
  }
}

hipError_t launch_mask_input(Ctx& c) {
  const int border = c.p.hSStampWidth + c.p.hKernelWidth;
  const int spreadWidth = (int)(c.p.hKernelWidth * c.p.inSpreadMaskFactor);
  const int w2 = spreadWidth / 2;
  if (w2 < 0 || w2 > MAX_W2) return hipErrorInvalidValue;
  dim3 grid((c.w + MT_X - 1) / MT_X, (c.h + MT_Y - 1) / MT_Y);
  prof_begin(c, XB_PROF_MASK);
  hipLaunchKernelGGL(k_mask_input, grid, dim3(256), 0, c.cur, c.img[0], c.img[1], c.mask,
                     c.w, c.h, border, c.m.threshHigh, c.m.threshLow, w2);
  prof_end(c, XB_PROF_MASK);
  return hipGetLastError();
}

}  // namespace xb
