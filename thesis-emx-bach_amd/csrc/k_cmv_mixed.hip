// k_cmv_mixed.hip — the mixed-precision form of cmv on MFMA.
//
// The plain form (k_cmv.hip) follows the reference's arithmetic: separable
// Y/X passes (cl_kern/cmv.cl:70-135) and the normal-matrix sums in FP64
// (cmv.cl:195-248).  This form computes the same quantities as two dense
// products on the matrix cores, in BF16 split precision with FP32
// accumulation: x = h + m + l (three BF16 parts, 24 significand bits — every
// FP32 value exactly) and a*b ~ h*h + h*m + m*h + h*l + l*h + m*m (the
// dropped terms are below 2^-24 of the product).  A two-part split (16 bits)
// is not enough: the normal equations are ill-conditioned, and its 2e-5
// errors in W and Q grew into chi2 changes that sent the clip-refit loop
// through many more passes than the plain form's.  The products are:
//
//   W  (961 x 49)  = P (961 x 504) . V (504 x 49)
//       P: the 51x51 window around the substamp, unrolled per pixel (tap (u,v)
//       of pixel p reads window[p/31 + u][p%31 + v]; taps padded 21 -> 24 per
//       row so an MFMA's 8 taps lie in one window row); V: the 2-D basis
//       fy_n[20-u] * fx_n[20-v] (the separable pair of cmv.cl:70-135 as one
//       21x21 filter), split once on the host.
//   G' (53 x 53)   = X^T X,  X = [W'_0 .. W'_51 | target]  (961 x 53)
//       W'_n = W_n - W_0 for the even-even rows (cmv.cl:137-154), rows 49-51
//       the background monomials (cmv.cl:156-193); G = G'[0:52, 0:52] is the
//       Gram matrix of createQ and G'[52, :] the target products of createB
//       (cmv.cl:195-248) and of the cd background right-hand side
//       (cl_kern/cd.cl:318-344).
//
// One workgroup (8 waves, two per SIMD) per stamp.  Each wave owns 4 of the
// 31 pixel tiles (two passes of 2), keeps 2x2 32x32 accumulators, and after
// each pass stages its 64-pixel slab of W in LDS: the slab is written to HBM
// in FP64 (the downstream fit reads W as the plain form does) and summed into
// the wave's partial G' straight from LDS.  The eight partial G' are added in
// wave order, so the result is deterministic.  Everything after cmv (stamp LUs, global
// system, solve, chi2, convolution) is the plain FP64 code.
#include "xb_dev.h"

namespace xb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int MX_WAVES = 8;            // two waves per SIMD: one unpacks while the other multiplies
constexpr int MX_THREADS = MX_WAVES * WAVE;
constexpr int MX_TAPW = 24;            // taps per row, 21 padded to 24
constexpr int MX_KSTEPS = 32;          // 16 taps per MFMA (504 taps -> 512)
constexpr int MX_MT = 31;              // 32-pixel tiles over 961 pixels
constexpr int MX_TPP = 2;              // pixel tiles per pass
constexpr int MX_PASSES = 2;           // passes per wave (8 waves x 2 x 2 = 32 tiles)
constexpr int MX_SLAB = MX_TPP * 32;   // pixels per pass
constexpr int MX_WP = 56;              // window row pitch (words; columns >= 51 are 0)
constexpr int MX_SP = MX_SLAB + 1;     // stage row pitch (floats)
constexpr int MX_SROWS = WROWS + 1;    // staged X columns: W rows 0..48, background rows, target
static_assert(MX_WAVES * MX_PASSES * MX_TPP >= MX_MT, "tiles cover the substamp");

// dynamic LDS carve (bytes)
constexpr size_t MX_WIN_B = sizeof(uint2) * TW * MX_WP;                      // 22,848
constexpr size_t MX_STAGE_B = sizeof(float) * MX_WAVES * MX_SROWS * MX_SP;   // 110,240
constexpr size_t MX_TGT_B = sizeof(float) * 976;                             // 3,904
constexpr size_t MX_MONO_B = sizeof(float) * 2 * 32;                         // 256
constexpr size_t MX_LDS = MX_WIN_B + MX_STAGE_B + MX_TGT_B + MX_MONO_B;
static_assert(MX_LDS <= 160 * 1024, "mixed fill LDS");
static_assert(MX_WAVES * 3 * 16 * WAVE * sizeof(float) <= MX_STAGE_B, "partial G' fits the stage");

// x = h + m + l, each part BF16 rounded to nearest even (gfx950 converts
// FP32 to BF16 in one instruction; the two differences are exact in FP32).
// Packed as .x = h | m << 16, .y = l.
__device__ __forceinline__ uint2 split3(float x) {
  const __bf16 h = (__bf16)x;
  const float r1 = x - (float)h;
  const __bf16 m = (__bf16)r1;
  const __bf16 l = (__bf16)(r1 - (float)m);
  return make_uint2((uint32_t)__builtin_bit_cast(uint16_t, h) |
                        ((uint32_t)__builtin_bit_cast(uint16_t, m) << 16),
                    (uint32_t)__builtin_bit_cast(uint16_t, l));
}

struct Split8 {  // 8 values as the three MFMA operands
  bf16x8 h, m, l;
};

__device__ __forceinline__ Split8 unpack8(const uint2 w[8]) {
  uint4 h, m, l;
  h.x = __builtin_amdgcn_perm(w[1].x, w[0].x, 0x05040100u);
  h.y = __builtin_amdgcn_perm(w[3].x, w[2].x, 0x05040100u);
  h.z = __builtin_amdgcn_perm(w[5].x, w[4].x, 0x05040100u);
  h.w = __builtin_amdgcn_perm(w[7].x, w[6].x, 0x05040100u);
  m.x = __builtin_amdgcn_perm(w[1].x, w[0].x, 0x07060302u);
  m.y = __builtin_amdgcn_perm(w[3].x, w[2].x, 0x07060302u);
  m.z = __builtin_amdgcn_perm(w[5].x, w[4].x, 0x07060302u);
  m.w = __builtin_amdgcn_perm(w[7].x, w[6].x, 0x07060302u);
  l.x = __builtin_amdgcn_perm(w[1].y, w[0].y, 0x05040100u);
  l.y = __builtin_amdgcn_perm(w[3].y, w[2].y, 0x05040100u);
  l.z = __builtin_amdgcn_perm(w[5].y, w[4].y, 0x05040100u);
  l.w = __builtin_amdgcn_perm(w[7].y, w[6].y, 0x05040100u);
  return {__builtin_bit_cast(bf16x8, h), __builtin_bit_cast(bf16x8, m),
          __builtin_bit_cast(bf16x8, l)};
}

// c += a * b over the six leading part products, smallest first
__device__ __forceinline__ f32x16 mfma6(const Split8& a, const Split8& b, f32x16 c) {
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.m, b.m, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.l, b.h, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.h, b.l, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.m, b.h, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.h, b.m, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.h, b.h, c, 0, 0, 0);
  return c;
}

// vtab: [MX_KSTEPS][2 n-tiles][h, m, l][64 lanes] x 8 BF16 (uint4)
__global__ __launch_bounds__(MX_THREADS) void k_fill_mixed(FillArgs A,
                                                           const uint4* __restrict__ vtab) {
  int s, ssx, ssy;
  if (!fill_stamp(A, blockIdx.x, &s, &ssx, &ssy)) return;
  const int g = A.kept[s];
  const bool valid = A.cur[s] < A.count[g];
  const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
  double* Ws = A.W + (size_t)s * WROWS * NPIX;
  if (!valid) {  // as the plain form: zero rows, Gram and targets
    for (int i = tid; i < WROWS * NPIX; i += MX_THREADS) Ws[i] = 0.0;
    for (int i = tid; i < WROWS * WROWS; i += MX_THREADS) {
      const int r = i / WROWS, q = i % WROWS;
      if (q <= r) A.G[(size_t)s * WROWS * WROWS + i] = 0.0;
    }
    if (tid < WROWS) A.Bc[(size_t)s * WROWS + tid] = 0.0;
    if (tid < NBG) A.Br[(size_t)s * NBG + tid] = 0.0;
    return;
  }
  extern __shared__ __align__(16) unsigned char mx_lds[];
  uint2* win = reinterpret_cast<uint2*>(mx_lds);
  float* stage = reinterpret_cast<float*>(mx_lds + MX_WIN_B);
  float* tgt = reinterpret_cast<float*>(mx_lds + MX_WIN_B + MX_STAGE_B);
  float* mono = reinterpret_cast<float*>(mx_lds + MX_WIN_B + MX_STAGE_B + MX_TGT_B);

  const int x0 = ssx - (HSS + HKW), y0 = ssy - (HSS + HKW);
  for (int i = tid; i < TW * MX_WP; i += MX_THREADS) {
    const int r = i / MX_WP, c = i % MX_WP;
    float v = 0.0f;
    if (c < TW) {
      const int xr = min(max(x0 + c, 0), A.w - 1), yr = min(max(y0 + r, 0), A.h - 1);
      v = A.img[(size_t)xr + (size_t)yr * A.w];
    }
    win[i] = split3(v);
  }
  for (int i = tid; i < 976; i += MX_THREADS) {
    float v = 0.0f;
    if (i < NPIX) {
      const int xr = min(max(ssx - HSS + i % SSW, 0), A.w - 1);
      const int yr = min(max(ssy - HSS + i / SSW, 0), A.h - 1);
      v = A.img2[(size_t)xr + (size_t)yr * A.w];
    }
    tgt[i] = v;
  }
  if (tid < 2 * SSW) {  // background monomials (FP32, as cmv.cl:156-193)
    const int b = tid / SSW, i = tid % SSW;
    const float hw = (float)A.w * 0.5f, hh = (float)A.h * 0.5f;
    mono[b * 32 + i] = b == 0 ? ((float)(ssx - HSS + i) - hw) / hw
                              : ((float)(ssy - HSS + i) - hh) / hh;
  }
  // the even-even flags of this lane's two X columns, loaded before any store
  const int half = lane / 32, col = lane % 32;
  const bool sub0 = A.basis->evenSub[col] != 0;
  const bool sub1 = 32 + col < NPSF && A.basis->evenSub[32 + col] != 0;
  __syncthreads();

  // k step ks: this lane's 8 taps start at tap row u, column v0; taps past
  // column 20 or row 20 are padding (their window words are masked to 0, so a
  // NaN next to the kernel's support stays out of W, as in the plain form)
  auto tapRow = [&](int ks, int& rowOff, uint32_t& m45, uint32_t& m67, uint32_t& m03) {
    const int k0 = ks * 16 + 8 * half;
    const int u = k0 / MX_TAPW, v0 = k0 % MX_TAPW;
    const int nreal = u < KW ? min(8, KW - v0) : 0;  // 0, 5 or 8 real taps
    rowOff = min(u, KW - 1) * MX_WP + v0;
    m03 = nreal > 0 ? 0xffffffffu : 0u;
    m45 = nreal == 8 ? 0xffffffffu : nreal == 5 ? 0x0000ffffu : 0u;
    m67 = nreal == 8 ? 0xffffffffu : 0u;
  };
  float* st = stage + (size_t)wv * MX_SROWS * MX_SP;
  f32x16 g00 = {0}, g10 = {0}, g11 = {0};
  for (int ps = 0; ps < MX_PASSES; ps++) {
    const int mt0 = MX_PASSES * MX_TPP * wv + MX_TPP * ps;
    // ---- W slab: MX_TPP pixel tiles x 2 basis tiles (tile 31 is past the
    // substamp: its pixels are clamped to 960 and never written)
    f32x16 acc[MX_TPP][2];
    int base[MX_TPP];
#pragma unroll
    for (int gi = 0; gi < MX_TPP; gi++) {
      acc[gi][0] = f32x16{0};
      acc[gi][1] = f32x16{0};
      const int p = min((mt0 + gi) * 32 + col, NPIX - 1);
      base[gi] = (p / SSW) * MX_WP + p % SSW;
    }
    // (the other wave on the SIMD covers the B loads' latency)
    const uint4* vt = vtab + lane;
    for (int ks = 0; ks < MX_KSTEPS; ks++) {
      const uint4* vn = vt + (size_t)ks * 6 * WAVE;
      const Split8 b0 = {__builtin_bit_cast(bf16x8, vn[0]), __builtin_bit_cast(bf16x8, vn[WAVE]),
                         __builtin_bit_cast(bf16x8, vn[2 * WAVE])};
      const Split8 b1 = {__builtin_bit_cast(bf16x8, vn[3 * WAVE]),
                         __builtin_bit_cast(bf16x8, vn[4 * WAVE]),
                         __builtin_bit_cast(bf16x8, vn[5 * WAVE])};
      int rowOff;
      uint32_t m45, m67, m03;
      tapRow(ks, rowOff, m45, m67, m03);
      Split8 a[MX_TPP];
#pragma unroll
      for (int gi = 0; gi < MX_TPP; gi++) {
        uint2 w[8];
#pragma unroll
        for (int t = 0; t < 8; t++) w[t] = win[base[gi] + rowOff + t];  // inside the window
        const uint32_t mk[8] = {m03, m03, m03, m03, m03, m45, m67, m67};
#pragma unroll
        for (int t = 0; t < 8; t++) {
          w[t].x &= mk[t];
          w[t].y &= mk[t];
        }
        a[gi] = unpack8(w);
      }
#pragma unroll
      for (int gi = 0; gi < MX_TPP; gi++) {
        acc[gi][0] = mfma6(a[gi], b0, acc[gi][0]);
        acc[gi][1] = mfma6(a[gi], b1, acc[gi][1]);
      }
    }
    // ---- stage the slab (rows 0..48) and X columns 49..52
#pragma unroll
    for (int gi = 0; gi < MX_TPP; gi++)
#pragma unroll
      for (int nt = 0; nt < 2; nt++) {
        const int n = nt * 32 + col;
        if (n < NPSF)
#pragma unroll
          for (int v = 0; v < 16; v++)
            st[n * MX_SP + gi * 32 + 8 * (v / 4) + 4 * half + (v % 4)] = acc[gi][nt][v];
      }
    const int pBase = mt0 * 32;
    // X columns 49..52 over the slab: the background monomials 1, yf, xf
    // (W rows 49..51) and the target
    for (int i = lane; i < 4 * MX_SLAB; i += WAVE) {
      const int r = i / MX_SLAB, j = i % MX_SLAB, p = min(pBase + j, NPIX - 1);
      const float v = r == 0 ? 1.0f : r == 1 ? mono[32 + p / SSW] : r == 2 ? mono[p % SSW] : tgt[p];
      st[(NPSF + r) * MX_SP + j] = v;
    }
    __syncthreads();
    // ---- W rows 0..48 of the slab in FP64 (one pixel per lane and row)
    const bool inW = pBase + lane < NPIX;
#pragma unroll 7
    for (int n = 0; n < NPSF; n++) {
      const float v = st[n * MX_SP + lane];
      if (inW) Ws[(size_t)n * NPIX + pBase + lane] = (double)v;
    }
    // ---- partial G' over the slab: X columns col (F0) and 32 + col (F1);
    // columns past 52 are zero (they read row 52 and drop it)
    const int a0 = col, a1 = 32 + col;
    const bool use1 = a1 <= WROWS;
    const float* s0 = st + a0 * MX_SP;
    const float* s1 = st + min(a1, WROWS) * MX_SP;
#pragma unroll
    for (int kk = 0; kk < MX_SLAB / 16; kk++) {
      uint2 w0[8], w1[8];
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const int j = kk * 16 + 8 * half + t;
        const bool in = pBase + j < NPIX;
        const float r0 = st[j];
        float f0 = s0[j], f1 = s1[j];
        f0 = sub0 ? f0 - r0 : f0;
        f1 = sub1 ? f1 - r0 : f1;
        w0[t] = in ? split3(f0) : make_uint2(0u, 0u);
        w1[t] = in && use1 ? split3(f1) : make_uint2(0u, 0u);
      }
      const Split8 x0 = unpack8(w0), x1 = unpack8(w1);
      g00 = mfma6(x0, x0, g00);
      g10 = mfma6(x1, x0, g10);
      g11 = mfma6(x1, x1, g11);
    }
    __syncthreads();  // the stage is rewritten by the next pass
  }
  // ---- add the partial G' in wave order
  float* red = stage;  // [wave][tile][v][lane]
#pragma unroll
  for (int v = 0; v < 16; v++) {
    red[((wv * 3 + 0) * 16 + v) * WAVE + lane] = g00[v];
    red[((wv * 3 + 1) * 16 + v) * WAVE + lane] = g10[v];
    red[((wv * 3 + 2) * 16 + v) * WAVE + lane] = g11[v];
  }
  __syncthreads();
  for (int i = tid; i < NBG * NPIX; i += MX_THREADS) {  // W rows 49..51 (the plain form's values)
    const int b = i / NPIX, p = i % NPIX;
    const double xf = (double)mono[p % SSW], yf = (double)mono[32 + p / SSW];
    Ws[(size_t)(NPSF + b) * NPIX + p] = (b == 2 ? xf : 1.0) * (b == 1 ? yf : 1.0);
  }
  double* G = A.G + (size_t)s * WROWS * WROWS;
  for (int e = tid; e < 3 * 16 * WAVE; e += MX_THREADS) {
    const int tile = e / (16 * WAVE), v = (e / WAVE) % 16, l = e % WAVE;
    float sum = red[((0 * 3 + tile) * 16 + v) * WAVE + l];
    for (int q = 1; q < MX_WAVES; q++) sum += red[((q * 3 + tile) * 16 + v) * WAVE + l];
    const int row = (tile == 0 ? 0 : 32) + 8 * (v / 4) + 4 * (l / 32) + (v % 4);
    const int cc = (tile == 2 ? 32 : 0) + l % 32;
    if (row < WROWS && cc < WROWS && cc <= row) {
      G[row * WROWS + cc] = (double)sum;
    } else if (row == WROWS && cc < WROWS) {
      A.Bc[(size_t)s * WROWS + cc] = (double)sum;
      if (cc >= NPSF) A.Br[(size_t)s * NBG + cc - NPSF] = (double)sum;
    }
  }
}

hipError_t launch_fill_mixed(Ctx& c, const FillArgs& A, int nMax) {
  static const hipError_t attrRc = hipFuncSetAttribute(
      (const void*)k_fill_mixed, hipFuncAttributeMaxDynamicSharedMemorySize, (int)MX_LDS);
  if (attrRc != hipSuccess) return attrRc;
  hipLaunchKernelGGL(k_fill_mixed, dim3(nMax), dim3(MX_THREADS), MX_LDS, c.cur, A,
                     reinterpret_cast<const uint4*>(c.mixTab));
  return hipGetLastError();
}

}  // namespace xb
