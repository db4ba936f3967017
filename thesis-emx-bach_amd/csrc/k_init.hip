// k_init.hip — input masking (the reference's init stage).
//
// One pass replaces cl_kern/ini.cl:15-35 (maskInput) and 37-66 (spreadMask).
// A workgroup of four waves makes a 64-column output tile of MT_Y rows (16,
// 32 or 64: a launch shape the batch autotuner picks; every shape gives the
// same bits).  Each row of the
// tile plus its halo (w2 rows above and below, w2 columns either side) is
// read once, every pixel's flags formed, and the row's BAD_INPUT pixels
// collected as a bit mask by ballots; the horizontal spread over
// 2 w2 + 1 columns is a few 64-bit shifts and ORs of that mask (wave-uniform,
// on the scalar unit), the vertical spread an OR of 2 w2 + 1 row masks from
// LDS, and each output pixel tests one bit.  HBM traffic: 8 B in per pixel
// (times (64 + 2 w2) / 64 for the halo rows) + 2 B out.
#include "xb_dev.h"

namespace xb {

constexpr int MT_X = 64, MAX_W2 = 8;
constexpr int MT_WAVES = 4;

// flags of pixel (x, y) from its two values (0 outside the image)
__device__ __forceinline__ uint16_t input_flags(float tf, float sf, int x, int y, int w, int h,
                                                int border, double hi, double lo) {
  uint16_t m = 0;
  if (x >= 0 && x < w && y >= 0 && y < h) {
    const double t = (double)tf, s = (double)sf;
    if (t == 0.0 || s == 0.0) m |= XB_MASK_BAD_INPUT | XB_MASK_BAD_PIX_VAL;
    if (t >= hi || s >= hi) m |= XB_MASK_BAD_INPUT | XB_MASK_SAT_PIXEL;
    if (t <= lo || s <= lo) m |= XB_MASK_BAD_INPUT | XB_MASK_LOW_PIXEL;
    if (x < border || x >= w - border || y < border || y >= h - border)
      m |= XB_MASK_BAD_PIXEL_S | XB_MASK_BAD_PIXEL_T;
  }
  return m;
}

// scZero (block (0, 0)): the device scalars of the run, cleared here so the
// pipeline's first kernel needs no separate fill
template <int MT_Y>
__global__ __launch_bounds__(WAVE * MT_WAVES) void k_mask_input(
    const float* __restrict__ T, const float* __restrict__ S, uint16_t* __restrict__ mask, int w,
    int h, int border, double hi, double lo, int w2, uint32_t* __restrict__ scZero,
    int scWords) {
  constexpr int MT_RPW = (MT_Y + 2 * MAX_W2 + MT_WAVES - 1) / MT_WAVES;  // rows per wave
  __shared__ unsigned long long hor[MT_Y + 2 * MAX_W2];  // spread row masks
  __shared__ uint16_t own[MT_Y][MT_X];
  const int lane = threadIdx.x % WAVE;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / WAVE);
  if (blockIdx.x == 0 && blockIdx.y == 0)
    for (int i = threadIdx.x; i < scWords; i += blockDim.x) scZero[i] = 0;
  const int x0 = blockIdx.x * MT_X, y0 = blockIdx.y * MT_Y;
  const int lh = MT_Y + 2 * w2;
  // halo columns: lane k < w2 holds x0 - w2 + k, lane w2 <= k < 2 w2 holds
  // x0 + 64 + k - w2
  const int hx = lane < w2 ? x0 - w2 + lane : x0 + MT_X + lane - w2;
  // (lanes past the halo read their own column again: no extra traffic)
  const int xc = min(max(x0 + lane, 0), w - 1);
  const int hxc = lane < 2 * w2 ? min(max(hx, 0), w - 1) : xc;
  // every row's values of the wave loaded together (clamped addresses; the
  // flags test the true coordinates)
  float tv[MT_RPW], sv[MT_RPW], th[MT_RPW], sh[MT_RPW];
#pragma unroll
  for (int r = 0; r < MT_RPW; r++) {
    const int y = min(max(y0 + wave + r * MT_WAVES - w2, 0), h - 1);
    tv[r] = T[(size_t)y * w + xc];
    sv[r] = S[(size_t)y * w + xc];
    th[r] = T[(size_t)y * w + hxc];
    sh[r] = S[(size_t)y * w + hxc];
  }
#pragma unroll
  for (int r = 0; r < MT_RPW; r++) {
    const int ly = wave + r * MT_WAVES;
    if (ly >= lh) break;
    const int y = y0 + ly - w2;
    const uint16_t m = input_flags(tv[r], sv[r], x0 + lane, y, w, h, border, hi, lo);
    const int oy = ly - w2;
    if (oy >= 0 && oy < MT_Y) own[oy][lane] = m;
    const unsigned long long mid = __ballot((m & XB_MASK_BAD_INPUT) != 0);
    unsigned long long hs = mid;
    if (w2 > 0) {
      const uint16_t hm = lane < 2 * w2
                              ? input_flags(th[r], sh[r], hx, y, w, h, border, hi, lo)
                              : (uint16_t)0;
      const unsigned long long hb = __ballot((hm & XB_MASK_BAD_INPUT) != 0);
      // ext bit p: column x0 - w2 + p (p < 64 + 2 w2), as (lo, hi) words
      const unsigned long long eLo = (mid << w2) | (hb & ((1ull << w2) - 1ull));
      const unsigned long long eHi = (mid >> (64 - w2)) | ((hb >> w2) << w2);
      // bit c of the result: any of ext bits c .. c + 2 w2
      hs = eLo;
      for (int k = 1; k <= 2 * w2; k++) hs |= (eLo >> k) | (eHi << (64 - k));
    }
    if (lane == 0) hor[ly] = hs;
  }
  __syncthreads();
  for (int oy = wave; oy < MT_Y; oy += MT_WAVES) {
    const int y = y0 + oy, x = x0 + lane;
    if (y >= h) break;
    unsigned long long v = 0;
    for (int k = 0; k <= 2 * w2; k++) v |= hor[oy + k];
    uint16_t m = own[oy][lane];
    if ((m & XB_MASK_BAD_INPUT) == 0 && ((v >> lane) & 1ull)) m |= XB_MASK_OK_CONV;
    if (x < w) mask[(size_t)y * w + x] = m;
  }
}

hipError_t launch_mask_input(Ctx& c) {
  const int border = c.p.hSStampWidth + c.p.hKernelWidth;
  const int spreadWidth = (int)(c.p.hKernelWidth * c.p.inSpreadMaskFactor);
  const int w2 = spreadWidth / 2;
  if (w2 < 0 || w2 > MAX_W2) return hipErrorInvalidValue;
  static_assert(sizeof(DevScalars) % 4 == 0, "device scalars are whole words");
  const int rows = c.shapes.mask_rows;  // 16, 32 or 64 (checked when set)
  dim3 grid((c.w + MT_X - 1) / MT_X, (c.h + rows - 1) / rows);
  auto* kern = rows == 16 ? k_mask_input<16> : (rows == 32 ? k_mask_input<32> : k_mask_input<64>);
  prof_begin(c, XB_PROF_MASK);
  hipLaunchKernelGGL(kern, grid, dim3(WAVE * MT_WAVES), 0, c.cur, c.img[0], c.img[1], c.mask,
                     c.w, c.h, border, c.m.threshHigh, c.m.threshLow, w2,
                     reinterpret_cast<uint32_t*>(c.sc), (int)(sizeof(DevScalars) / 4));
  prof_end(c, XB_PROF_MASK);
  return hipGetLastError();
}

}  // namespace xb
