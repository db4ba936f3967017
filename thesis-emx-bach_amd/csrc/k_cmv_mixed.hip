// k_cmv_mixed.hip — the mixed-precision form of cmv on the matrix cores.
//
// The plain form (k_cmv.hip) follows the reference's arithmetic: separable
// Y/X passes (cl_kern/cmv.cl:70-135) and the normal-matrix sums in FP64
// (cmv.cl:195-248).  This form computes the same quantities as two dense
// products on the matrix cores:
//
//   W  (961 x 49) = P (961 x 504) . V (504 x 49), in BF16 split precision:
//       P is the 51x51 window around the substamp unrolled per pixel (tap
//       (u,v) of pixel p reads window[p/31 + u][p%31 + v]; taps padded
//       21 -> 24 per row so an MFMA's 8 taps lie in one window row), V the
//       2-D basis fy_n[20-u] * fx_n[20-v] (the separable pair of
//       cmv.cl:70-135 as one 21x21 filter), split once on the host.  Every
//       operand is x = h + m + l, three BF16 parts (24 significand bits: every
//       FP32 value exactly), and a*b ~ h*h + h*m + m*h + h*l + l*h + m*m (the
//       dropped terms are below 2^-24 of the product), accumulated in FP32
//       (v_mfma_f32_32x32x16_bf16).  A two-part split (16 bits) left W 2e-5
//       from FP64, which moved chi2 enough to send the clip-refit loop through
//       many more passes than the plain form's.
//   G' (53 x 53) = X^T X, X = [W'_0 .. W'_51 | target], on the FP64 matrix
//       cores (v_mfma_f64_16x16x4f64): W'_n = W_n - W_0 for the even-even rows
//       (cmv.cl:137-154), rows 49-51 the background monomials (cmv.cl:156-193);
//       G = G'[0:52, 0:52] is the Gram matrix of createQ, G'[52, :] the target
//       products of createB (cmv.cl:195-248) and of the cd background
//       right-hand side (cl_kern/cd.cl:318-344).  FP32-accumulated products
//       are not enough here: near a bright contaminant (a cosmic ray, a star
//       clipped at saturation) the FP32 rounding of partial sums makes Q more
//       than a small perturbation of an exact Gram matrix, and the
//       ill-conditioned fit lost a quarter of its stamps on a crowded field.
//       The W values are FP32, so their products are exact in FP64.
//
//   With XB_FORM_FP16 the W product uses FP16 instead: x = h + l, two FP16
//       parts (22 significand bits), after scaling by a power of two so the
//       values fit FP16's range: the stamp's window by 2^(14 - e) with 2^e above
//       its largest magnitude, the basis table likewise on the host; a*b ~
//       h*h + h*l + l*h (v_mfma_f32_32x32x16_f16, FP32 accumulation), then W
//       is scaled back (exact: powers of two).
//
// k_fill_mixed: one workgroup (8 waves, two per SIMD) per stamp.  Each wave
// owns 4 of the 31 pixel tiles (two passes of 2) and keeps 2x2 32x32
// accumulators; after each pass it stages its 64-pixel slab in LDS and writes
// it to HBM in FP64 (the downstream fit reads W as in the plain form).
// k_fill_gram64: one workgroup (16 waves) per stamp stages X in 64-pixel
// chunks in LDS; wave (J, Q) accumulates the lower tiles (I >= J, 16x16) of
// column block J over pixel quarter Q, and the quarters are added in order,
// so the result is deterministic.  Everything after cmv (stamp LUs, global
// system, solve, chi2, convolution) is the plain FP64 code.
#include "xb_dev.h"

namespace xb {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
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
static_assert(MX_WAVES * MX_PASSES * MX_TPP >= MX_MT, "tiles cover the substamp");

// dynamic LDS carve (bytes)
constexpr size_t MX_WIN_B = sizeof(uint2) * TW * MX_WP;                      // 22,848
constexpr size_t MX_STAGE_B = sizeof(float) * MX_WAVES * NPSF * MX_SP;      // 101,920
constexpr size_t MX_MONO_B = sizeof(float) * 2 * 32;                         // 256
constexpr size_t MX_LDS = MX_WIN_B + MX_STAGE_B + MX_MONO_B;
static_assert(MX_LDS <= 160 * 1024, "mixed fill LDS");

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

// x = h + l, two FP16 parts (x already scaled into FP16's range), packed as
// h | l << 16 in .x
__device__ __forceinline__ uint2 split2h(float x) {
  const _Float16 h = (_Float16)x;
  const _Float16 l = (_Float16)(x - (float)h);
  return make_uint2((uint32_t)__builtin_bit_cast(uint16_t, h) |
                        ((uint32_t)__builtin_bit_cast(uint16_t, l) << 16),
                    0u);
}

struct Split8h {  // 8 values as the two FP16 MFMA operands
  f16x8 h, l;
};

__device__ __forceinline__ Split8h unpack8h(const uint2 w[8]) {
  uint4 h, l;
  h.x = __builtin_amdgcn_perm(w[1].x, w[0].x, 0x05040100u);
  h.y = __builtin_amdgcn_perm(w[3].x, w[2].x, 0x05040100u);
  h.z = __builtin_amdgcn_perm(w[5].x, w[4].x, 0x05040100u);
  h.w = __builtin_amdgcn_perm(w[7].x, w[6].x, 0x05040100u);
  l.x = __builtin_amdgcn_perm(w[1].x, w[0].x, 0x07060302u);
  l.y = __builtin_amdgcn_perm(w[3].x, w[2].x, 0x07060302u);
  l.z = __builtin_amdgcn_perm(w[5].x, w[4].x, 0x07060302u);
  l.w = __builtin_amdgcn_perm(w[7].x, w[6].x, 0x07060302u);
  return {__builtin_bit_cast(f16x8, h), __builtin_bit_cast(f16x8, l)};
}

// c += a * b over the three leading FP16 part products, smallest first
__device__ __forceinline__ f32x16 mfma3h(const Split8h& a, const Split8h& b, f32x16 c) {
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.l, b.h, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.h, b.l, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a.h, b.h, c, 0, 0, 0);
  return c;
}


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

// the split type's operands and products
template <bool F16>
struct MxOps {
  using A = Split8;
  static constexpr int PARTS = 3;
  static __device__ __forceinline__ A unpack(const uint2 w[8]) { return unpack8(w); }
  static __device__ __forceinline__ A load(const uint4* vn) {
    return {__builtin_bit_cast(bf16x8, vn[0]), __builtin_bit_cast(bf16x8, vn[WAVE]),
            __builtin_bit_cast(bf16x8, vn[2 * WAVE])};
  }
  static __device__ __forceinline__ f32x16 mma(const A& a, const A& b, f32x16 c) {
    return mfma6(a, b, c);
  }
};
template <>
struct MxOps<true> {
  using A = Split8h;
  static constexpr int PARTS = 2;
  static __device__ __forceinline__ A unpack(const uint2 w[8]) { return unpack8h(w); }
  static __device__ __forceinline__ A load(const uint4* vn) {
    return {__builtin_bit_cast(f16x8, vn[0]), __builtin_bit_cast(f16x8, vn[WAVE])};
  }
  static __device__ __forceinline__ f32x16 mma(const A& a, const A& b, f32x16 c) {
    return mfma3h(a, b, c);
  }
};

// vtab: [MX_KSTEPS][2 n-tiles][parts][64 lanes] x 8 values (uint4): BF16
// h, m, l, or FP16 h, l scaled by 2^vscale
template <bool F16, int NI>
__global__ __launch_bounds__(MX_THREADS) void k_fill_mixed(ArgsN<FillArgs, NI> P,
                                                           const uint4* __restrict__ vtab,
                                                           int vscale) {
  const FillArgs& A = args_of(P);
  using Ops = MxOps<F16>;
  int s, ssx, ssy;
  if (!fill_stamp(A, blockIdx.x, &s, &ssx, &ssy)) return;
  const int g = A.kept[s];
  const bool valid = A.cur[s] < A.count[g];
  const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
  double* Ws = A.W + (size_t)s * WROWS * NPIX;
  if (!valid) {  // as the plain form: zero rows
    for (int i = tid; i < WROWS * NPIX; i += MX_THREADS) Ws[i] = 0.0;
    return;
  }
  extern __shared__ __align__(16) unsigned char mx_lds[];
  uint2* win = reinterpret_cast<uint2*>(mx_lds);
  float* stage = reinterpret_cast<float*>(mx_lds + MX_WIN_B);
  float* mono = reinterpret_cast<float*>(mx_lds + MX_WIN_B + MX_STAGE_B);

  const int x0 = ssx - (HSS + HKW), y0 = ssy - (HSS + HKW);
  __shared__ float wmax[MX_WAVES];
  float vmax = 0.0f;
  for (int i = tid; i < TW * MX_WP; i += MX_THREADS) {
    const int r = i / MX_WP, c = i % MX_WP;
    float v = 0.0f;
    if (c < TW) {
      const int xr = min(max(x0 + c, 0), A.w - 1), yr = min(max(y0 + r, 0), A.h - 1);
      v = A.img[(size_t)xr + (size_t)yr * A.w];
    }
    if (F16) {
      if (fabsf(v) <= 3.0e38f) vmax = fmaxf(vmax, fabsf(v));  // finite values
      win[i] = make_uint2(__float_as_uint(v), 0u);
    } else {
      win[i] = split3(v);
    }
  }
  int wexp = 0;  // FP16: the window is scaled by 2^(14 - wexp)
  if (F16) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, o));
    if (lane == 0) wmax[wv] = vmax;
    __syncthreads();
    float m = 0.0f;
#pragma unroll
    for (int w = 0; w < MX_WAVES; w++) m = fmaxf(m, wmax[w]);
    if (m > 0.0f) (void)frexpf(m, &wexp);  // m < 2^wexp
    for (int i = tid; i < TW * MX_WP; i += MX_THREADS)
      win[i] = split2h(ldexpf(__uint_as_float(win[i].x), 14 - wexp));
  }
  if (tid < 2 * SSW) {  // background monomials (FP32, as cmv.cl:156-193)
    const int b = tid / SSW, i = tid % SSW;
    const float hw = (float)A.w * 0.5f, hh = (float)A.h * 0.5f;
    mono[b * 32 + i] = b == 0 ? ((float)(ssx - HSS + i) - hw) / hw
                              : ((float)(ssy - HSS + i) - hh) / hh;
  }
  const int half = lane / 32, col = lane % 32;
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
  float* st = stage + (size_t)wv * NPSF * MX_SP;
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
      const uint4* vn = vt + (size_t)ks * 2 * Ops::PARTS * WAVE;
      const typename Ops::A b0 = Ops::load(vn);
      const typename Ops::A b1 = Ops::load(vn + Ops::PARTS * WAVE);
      int rowOff;
      uint32_t m45, m67, m03;
      tapRow(ks, rowOff, m45, m67, m03);
      typename Ops::A a[MX_TPP];
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
        a[gi] = Ops::unpack(w);
      }
#pragma unroll
      for (int gi = 0; gi < MX_TPP; gi++) {
        acc[gi][0] = Ops::mma(a[gi], b0, acc[gi][0]);
        acc[gi][1] = Ops::mma(a[gi], b1, acc[gi][1]);
      }
    }
    // ---- stage the slab (rows 0..48)
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
    __syncthreads();
    // ---- W rows 0..48 of the slab in FP64 (one pixel per lane and row);
    // FP16: the two scales undone (powers of two, exact in FP64)
    const bool inW = pBase + lane < NPIX;
    const int unscale = F16 ? wexp - 14 - vscale : 0;
#pragma unroll 7
    for (int n = 0; n < NPSF; n++) {
      const float v = st[n * MX_SP + lane];
      if (inW) Ws[(size_t)n * NPIX + pBase + lane] = ldexp((double)v, unscale);
    }
    __syncthreads();  // the stage is rewritten by the next pass
  }
  for (int i = tid; i < NBG * NPIX; i += MX_THREADS) {  // W rows 49..51 (the plain form's values)
    const int b = i / NPIX, p = i % NPIX;
    const double xf = (double)mono[p % SSW], yf = (double)mono[32 + p / SSW];
    Ws[(size_t)(NPSF + b) * NPIX + p] = (b == 2 ? xf : 1.0) * (b == 1 ? yf : 1.0);
  }
}

// ---------------------------------------------------------------------------
// G' = X^T X on the FP64 matrix cores
// ---------------------------------------------------------------------------
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int G64_WAVES = 16;
constexpr int G64_THREADS = G64_WAVES * WAVE;
constexpr int G64_CH = 64;                        // pixels per LDS chunk
constexpr int G64_NCH = (NPIX + G64_CH - 1) / G64_CH;  // 16 chunks (1024 px, zero past 961)
constexpr int G64_XP = 64 + 1;                    // LDS pitch of a staged pixel (64 columns)
constexpr int G64_TILES = 10;                     // lower 16x16 tiles of the 64x64 G'

__device__ __forceinline__ int g64_tile(int I, int J) { return I * (I + 1) / 2 + J; }

template <int NI>
__global__ __launch_bounds__(G64_THREADS) void k_fill_gram64(ArgsN<FillArgs, NI> P) {
  const FillArgs& A = args_of(P);
  int s, ssx, ssy;
  if (!fill_stamp(A, blockIdx.x, &s, &ssx, &ssy)) return;
  const int g = A.kept[s];
  const bool valid = A.cur[s] < A.count[g];
  const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
  double* G = A.G + (size_t)s * WROWS * WROWS;
  if (!valid) {  // as the plain form: zero Gram and targets
    for (int i = tid; i < WROWS * WROWS; i += G64_THREADS) {
      const int r = i / WROWS, q = i % WROWS;
      if (q <= r) G[i] = 0.0;
    }
    if (tid < WROWS) A.Bc[(size_t)s * WROWS + tid] = 0.0;
    if (tid < NBG) A.Br[(size_t)s * NBG + tid] = 0.0;
    return;
  }
  const double* Ws = A.W + (size_t)s * WROWS * NPIX;
  const int* evenSub = A.basis->evenSub;
  __shared__ double X[G64_CH][G64_XP];
  __shared__ double red[G64_TILES][4][WAVE];
  const int J = wv % 4, Q = wv / 4;  // column block, pixel quarter of each chunk
  f64x4 acc[4];
#pragma unroll
  for (int I = 0; I < 4; I++) acc[I] = f64x4{0.0, 0.0, 0.0, 0.0};
  // X for pixels ch*64 .. ch*64+63: column c = W'_c (c < 52), the target
  // (c = 52), zero (c > 52 or past the last pixel).  Thread t holds the
  // entries e = t + 1024 r (r < 4) of a chunk; the next chunk's entries are
  // loaded into registers while the current chunk's products run.
  constexpr int PER_T = 64 * G64_CH / G64_THREADS;  // 4
  auto xval = [&](int e, int ch) -> double {
    const int c = e / G64_CH, pp = e % G64_CH, p = ch * G64_CH + pp;
    double v = 0.0;
    if (p < NPIX) {
      if (c < WROWS) {
        v = Ws[(size_t)c * NPIX + p];
        if (c < NPSF && evenSub[c]) v = v - Ws[p];
      } else if (c == WROWS) {
        const int xr = min(max(ssx - HSS + p % SSW, 0), A.w - 1);
        const int yr = min(max(ssy - HSS + p / SSW, 0), A.h - 1);
        v = (double)A.img2[(size_t)xr + (size_t)yr * A.w];
      }
    }
    return v;
  };
  double nx[PER_T];
#pragma unroll
  for (int r = 0; r < PER_T; r++) nx[r] = xval(tid + G64_THREADS * r, 0);
  for (int ch = 0; ch < G64_NCH; ch++) {
#pragma unroll
    for (int r = 0; r < PER_T; r++) {
      const int e = tid + G64_THREADS * r;
      X[e % G64_CH][e / G64_CH] = nx[r];
    }
    __syncthreads();
    if (ch + 1 < G64_NCH)
#pragma unroll
      for (int r = 0; r < PER_T; r++) nx[r] = xval(tid + G64_THREADS * r, ch + 1);
    // wave (J, Q): pixels 16 Q .. 16 Q + 15 of the chunk, 4 per MFMA;
    // lane: pixel k = l / 16 of the step, column l % 16 of the block
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int pp = 16 * Q + 4 * kk + lane / 16;
      const double* xr = &X[pp][lane % 16];
      const double xb = xr[16 * J];
      // tiles I >= J (J is wave-uniform); the products of the other tiles are
      // formed with a zero operand and never read
      const double x0 = J <= 0 ? xr[0] : 0.0, x1 = J <= 1 ? xr[16] : 0.0;
      const double x2 = J <= 2 ? xr[32] : 0.0, x3 = xr[48];
      acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, xb, acc[0], 0, 0, 0);
      acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, xb, acc[1], 0, 0, 0);
      acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(x2, xb, acc[2], 0, 0, 0);
      acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(x3, xb, acc[3], 0, 0, 0);
    }
    __syncthreads();
  }
  // add the four pixel quarters in order
  for (int q = 0; q < 4; q++) {
    if (Q == q) {
#pragma unroll
      for (int I = 0; I < 4; I++) {
        if (I >= J) {
#pragma unroll
          for (int v = 0; v < 4; v++) {
            double* r = &red[g64_tile(I, J)][v][lane];
            *r = q == 0 ? acc[I][v] : *r + acc[I][v];
          }
        }
      }
    }
    __syncthreads();
  }
  // tile (I, J), lane l, value v: G'[16 I + l / 16 + 4 v][16 J + l % 16] (the
  // accumulator layout of v_mfma_f64_16x16x4f64)
  for (int e = tid; e < G64_TILES * 4 * WAVE; e += G64_THREADS) {
    const int t = e / (4 * WAVE), v = (e / WAVE) % 4, l = e % WAVE;
    int I = 0;
    while (g64_tile(I + 1, 0) <= t) I++;
    const int Jt = t - g64_tile(I, 0);
    const int row = 16 * I + l / 16 + 4 * v, cc = 16 * Jt + l % 16;
    const double val = red[t][v][l];
    if (row < WROWS && cc < WROWS && cc <= row) {
      G[row * WROWS + cc] = val;
    } else if (row == WROWS && cc < WROWS) {
      A.Bc[(size_t)s * WROWS + cc] = val;
      if (cc >= NPSF) A.Br[(size_t)s * NBG + cc - NPSF] = val;
    }
  }
}

template <int NI>
static hipError_t launch_fill_mixed_n(Ctx& c, const ArgsN<FillArgs, NI>& P, int nMax) {
  static const hipError_t attrRc = hipFuncSetAttribute(
      (const void*)k_fill_mixed<false, NI>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)MX_LDS);
  static const hipError_t attrRc16 = hipFuncSetAttribute(
      (const void*)k_fill_mixed<true, NI>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)MX_LDS);
  if (attrRc != hipSuccess) return attrRc;
  if (attrRc16 != hipSuccess) return attrRc16;
  const dim3 grid(nMax, 1, NI);
  if (c.form & XB_FORM_FP16)
    hipLaunchKernelGGL((k_fill_mixed<true, NI>), grid, dim3(MX_THREADS), MX_LDS, c.cur, P,
                       reinterpret_cast<const uint4*>(c.mixTab16), c.mixScale16);
  else
    hipLaunchKernelGGL((k_fill_mixed<false, NI>), grid, dim3(MX_THREADS), MX_LDS, c.cur, P,
                       reinterpret_cast<const uint4*>(c.mixTab), 0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill_gram64<NI>, grid, dim3(G64_THREADS), 0, c.cur, P);
  return hipGetLastError();
}

hipError_t launch_fill_mixed(Ctx& c, const FillArgs& A, int nMax) {
  return launch_fill_mixed_n<1>(c, ArgsN<FillArgs, 1>{{A}}, nMax);
}

hipError_t launch_fill_mixed_pair(Ctx& c, const FillArgs& A0, const FillArgs& A1, int nMax) {
  return launch_fill_mixed_n<2>(c, ArgsN<FillArgs, 2>{{A0, A1}}, nMax);
}

}  // namespace xb
