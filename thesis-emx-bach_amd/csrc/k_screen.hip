// k_screen.hip — the learned substamp screen: features, a 32-64-1 MLP with
// FP16 weights (forward, backward, Adam), and its use in the clip-refit loop.
//
// The reference's clip-refit loop (src/cdkscUtil.cpp:585-684, ksc.cl:1-72)
// rejects substamps one pass at a time: a pass fits all stamps, sigma-clips
// their chi2 and replaces the outliers (chi2 - mean > 2 sd), and the loop ends
// when a pass replaces nothing.  After each pass the screen looks at every
// substamp the pass kept, predicts from 32 features which ones a later pass
// would reject, and replaces those at once.  The loop itself runs unchanged
// and is the judge: every substamp the fit finally uses has passed its chi2
// test.  The first SCR_PASSES passes are screened (and recorded).
//
// Features of a stamp (its substamp of the image being convolved, T, and of
// the target, S, both 31x31 around the substamp centre, minus the stamp's sky
// and in units of its sky sigma from sss):
//   0 this pass's chi2 z-score (chi2 - mean) / sd, clamped to [-10, 50] / 10
//   1 asinh(chi2 / mean)
//   2,3 asinh(flux / 100)         T, S
//   4,5 asinh(peak / 10)          T, S
//   6,7 peak / positive flux      T, S
//   8..11 centroid x, y / 15      T, S
//   12,13 sqrt(second moment) / 15 T, S
//   14,15 (xx - yy), 2 xy over (xx + yy + 1) of S
//   16,17 asinh(max |Laplacian| / 10) T, S
//   18,19 10 * fraction of pixels above 5 sigma  T, S
//   20,21 asinh(min)              T, S
//   22,23 raw maximum / threshHigh T, S
//   24 clip spread of this pass, sd / mean of the clipped chi2
//   25 substamp index / 3, 26 substamps found / 3
//   27 log(sigma S / sigma T)
//   28 L1 distance of the flux-normalised profiles, clamped to [0, 4] / 2
//   29 asinh((flux S - flux T) / (|flux T| + |flux S| + 1))
//   30 pass index / 10
//   31 the same substamp's z-score in the previous pass (feature 0 then; 0 when
//      it was not in use or this is the first pass)
//
// Model (MLP_NPARAM floats): feature mean[32], sd[32], W1[64][32], b1[64],
// w2[64], b2, threshold.  W1 and w2 are used rounded to FP16; the rest FP32.
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "xb_dev.h"

namespace xb {

constexpr int MLP_IN = 32, MLP_HID = 64;
constexpr int MP_MU = 0, MP_SD = 32, MP_W1 = 64, MP_B1 = MP_W1 + MLP_HID * MLP_IN,
              MP_W2 = MP_B1 + MLP_HID, MP_B2 = MP_W2 + MLP_HID, MP_TAU = MP_B2 + 1;
constexpr int MLP_NPARAM = MP_TAU + 1;                          // 2242
constexpr int MLP_NGRAD = MLP_HID * MLP_IN + 2 * MLP_HID + 1;   // 2177 trainable
constexpr int MLP_CHUNK = 64;                                   // samples per gradient block
static_assert(MLP_NPARAM == XB_MLP_NPARAM, "model size");

__device__ __forceinline__ float h16(float x) { return __half2float(__float2half(x)); }

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ float wave_max_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_min_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}

// One 64-thread wave: hidden unit j = lane.  x: normalised inputs (LDS).
// Returns z (all lanes) and this lane's hidden value.
__device__ __forceinline__ float mlp_hidden(const float* P, const float* x, int j) {
  float a = P[MP_B1 + j];
  const float* w = P + MP_W1 + j * MLP_IN;
#pragma unroll 8
  for (int i = 0; i < MLP_IN; i++) a += h16(w[i]) * x[i];
  return fmaxf(a, 0.0f);
}

// ---------------------------------------------------------------------------
// features of kept stamp s (one 256-thread block per stamp)
// ---------------------------------------------------------------------------
struct ScreenArgs {
  const float* T;        // image being convolved
  const float* S;        // target image
  int w, h;
  const int* nKept;
  const int* kept;
  const int* cur;        // substamp each stamp had in this pass
  const int* count;
  const int* ssx;
  const int* ssy;
  const double* skyT;
  const double* skyS;
  const double* sdT;     // sky sigma of sss (the reference's "fwhm" arrays)
  const double* sdS;
  const double* sig;     // this pass's chi2
  const KscScalars* ks;  // clip mean / sd of this pass
  double threshHigh;
  int pass;
  const int* prevCur;    // substamp of each stamp in the previous pass (or null)
  float* zPrev;          // [N] feature 0 of the previous pass, updated here
  float* feat;           // [N][32] of this pass
  // the decision (apply): p(reject) of each stamp's features by the model, and
  // the flag of an eligible stamp (its substamp kept by the pass's check)
  int apply;
  const float* model;
  const int* curNow;     // substamp of each stamp after the check
  float* prob;
  uint8_t* flag;
};

constexpr int SCR_THREADS = 256;

__device__ __forceinline__ void screen_features_body(const ScreenArgs& A, int s, int g, int c,
                                                     int tid, int lane, int wid, float* out,
                                                     float* fs) {
  const int cx = A.ssx[g * MAXSS + c], cy = A.ssy[g * MAXSS + c];
  const float skT = (float)A.skyT[g], skS = (float)A.skyS[g];
  const float sgT = fmaxf((float)A.sdT[g], 1e-3f), sgS = fmaxf((float)A.sdS[g], 1e-3f);
  __shared__ float wT[NPIX], wS[NPIX];
  __shared__ float red[4][24];
  float rawT = -INFINITY, rawS = -INFINITY;
  for (int p = tid; p < NPIX; p += SCR_THREADS) {
    const int x = min(max(cx - HSS + p % SSW, 0), A.w - 1);
    const int y = min(max(cy - HSS + p / SSW, 0), A.h - 1);
    const float t = A.T[(size_t)y * A.w + x], v = A.S[(size_t)y * A.w + x];
    rawT = fmaxf(rawT, t);
    rawS = fmaxf(rawS, v);
    wT[p] = (t - skT) / sgT;
    wS[p] = (v - skS) / sgS;
  }
  __syncthreads();
  // pass 1: per image flux, positive flux, peak, min, moments, Laplacian, >5 sigma
  float q[24];
#pragma unroll
  for (int k = 0; k < 24; k++) q[k] = 0.0f;
  float pkT = -INFINITY, pkS = -INFINITY, mnT = INFINITY, mnS = INFINITY, lpT = 0.0f, lpS = 0.0f;
  for (int p = tid; p < NPIX; p += SCR_THREADS) {
    const float dx = (float)(p % SSW - HSS), dy = (float)(p / SSW - HSS);
    const float t = wT[p], v = wS[p];
    const float tp = fmaxf(t, 0.0f), vp = fmaxf(v, 0.0f);
    q[0] += t; q[1] += tp; q[2] += dx * tp; q[3] += dy * tp; q[4] += dx * dx * tp;
    q[5] += dy * dy * tp; q[6] += t > 5.0f ? 1.0f : 0.0f;
    q[7] += v; q[8] += vp; q[9] += dx * vp; q[10] += dy * vp; q[11] += dx * dx * vp;
    q[12] += dy * dy * vp; q[13] += dx * dy * vp; q[14] += v > 5.0f ? 1.0f : 0.0f;
    pkT = fmaxf(pkT, t); pkS = fmaxf(pkS, v); mnT = fminf(mnT, t); mnS = fminf(mnS, v);
    const int i = p % SSW, j = p / SSW;
    if (i > 0 && i < SSW - 1 && j > 0 && j < SSW - 1) {
      lpT = fmaxf(lpT, fabsf(4.0f * t - wT[p - 1] - wT[p + 1] - wT[p - SSW] - wT[p + SSW]));
      lpS = fmaxf(lpS, fabsf(4.0f * v - wS[p - 1] - wS[p + 1] - wS[p - SSW] - wS[p + SSW]));
    }
  }
  q[15] = pkT; q[16] = pkS; q[17] = mnT; q[18] = mnS; q[19] = lpT; q[20] = lpS;
  q[21] = rawT; q[22] = rawS;
#pragma unroll
  for (int k = 0; k < 15; k++) q[k] = wave_sum_f(q[k]);
  q[15] = wave_max_f(q[15]); q[16] = wave_max_f(q[16]);
  q[17] = wave_min_f(q[17]); q[18] = wave_min_f(q[18]);
  q[19] = wave_max_f(q[19]); q[20] = wave_max_f(q[20]);
  q[21] = wave_max_f(q[21]); q[22] = wave_max_f(q[22]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 23; k++) red[wid][k] = q[k];
  __syncthreads();
  __shared__ float tot[24];
  if (tid < 23) {
    float a = red[0][tid];
    for (int w2 = 1; w2 < 4; w2++) {
      const float b = red[w2][tid];
      a = tid < 15 ? a + b : (tid == 17 || tid == 18) ? fminf(a, b) : fmaxf(a, b);
    }
    tot[tid] = a;
  }
  __syncthreads();
  const float FT = tot[0], FS = tot[7];
  // pass 2: L1 distance of the flux-normalised profiles
  float d = 0.0f;
  const float iT = 1.0f / fmaxf(fabsf(FT), 1.0f), iS = 1.0f / fmaxf(fabsf(FS), 1.0f);
  for (int p = tid; p < NPIX; p += SCR_THREADS) d += fabsf(wS[p] * iS - wT[p] * iT);
  d = wave_sum_f(d);
  __syncthreads();
  if (lane == 0) red[wid][23] = d;
  __syncthreads();
  if (tid == 0) {
    const float D = red[0][23] + red[1][23] + red[2][23] + red[3][23];
    const float chi = (float)A.sig[s];
    const float mean = (float)A.ks->clipMean, sd = (float)A.ks->clipStd;
    float f[MLP_IN];
    const float z = sd > 0.0f ? (chi - mean) / sd : 0.0f;
    f[0] = fminf(fmaxf(z, -10.0f), 50.0f) / 10.0f;
    f[1] = asinhf(mean > 0.0f ? chi / mean : 0.0f);
    f[2] = asinhf(FT / 100.0f);
    f[3] = asinhf(FS / 100.0f);
    f[4] = asinhf(tot[15] / 10.0f);
    f[5] = asinhf(tot[16] / 10.0f);
    const float fpT = fmaxf(tot[1], 1.0f), fpS = fmaxf(tot[8], 1.0f);
    f[6] = tot[15] / fpT;
    f[7] = tot[16] / fpS;
    const float cxT = tot[2] / fpT, cyT = tot[3] / fpT, cxS = tot[9] / fpS, cyS = tot[10] / fpS;
    f[8] = cxT / 15.0f;
    f[9] = cyT / 15.0f;
    f[10] = cxS / 15.0f;
    f[11] = cyS / 15.0f;
    const float vxT = tot[4] / fpT - cxT * cxT, vyT = tot[5] / fpT - cyT * cyT;
    const float vxS = tot[11] / fpS - cxS * cxS, vyS = tot[12] / fpS - cyS * cyS;
    const float cvS = tot[13] / fpS - cxS * cyS;
    f[12] = sqrtf(fmaxf(vxT + vyT, 0.0f)) / 15.0f;
    f[13] = sqrtf(fmaxf(vxS + vyS, 0.0f)) / 15.0f;
    f[14] = (vxS - vyS) / (fabsf(vxS + vyS) + 1.0f);
    f[15] = 2.0f * cvS / (fabsf(vxS + vyS) + 1.0f);
    f[16] = asinhf(tot[19] / 10.0f);
    f[17] = asinhf(tot[20] / 10.0f);
    f[18] = 10.0f * tot[6] / (float)NPIX;
    f[19] = 10.0f * tot[14] / (float)NPIX;
    f[20] = asinhf(tot[17]);
    f[21] = asinhf(tot[18]);
    f[22] = tot[21] / (float)A.threshHigh;
    f[23] = tot[22] / (float)A.threshHigh;
    f[24] = mean > 0.0f ? sd / mean : 0.0f;
    f[25] = (float)c / 3.0f;
    f[26] = (float)A.count[g] / 3.0f;
    f[27] = logf(sgS / sgT);
    f[28] = fminf(fmaxf(D, 0.0f), 4.0f) / 2.0f;
    f[29] = asinhf((FS - FT) / (fabsf(FT) + fabsf(FS) + 1.0f));
    f[30] = (float)A.pass / 10.0f;
    f[31] = (A.prevCur && A.prevCur[s] == c) ? A.zPrev[s] : 0.0f;
    A.zPrev[s] = f[0];
#pragma unroll
    for (int k = 0; k < MLP_IN; k++) {
      const float v = isfinite(f[k]) ? f[k] : 0.0f;
      out[k] = v;
      fs[k] = v;
    }
  }
}

// The features of kept stamp s (one 256-thread block per stamp) and, with
// apply, the screen's decision for it, by the block's first wave from the
// features in LDS: flag = p(reject) > threshold for a stamp whose substamp the
// pass's check kept.
__global__ __launch_bounds__(SCR_THREADS) void k_screen_features(ScreenArgs A) {
  const int s = blockIdx.x;
  if (s >= *A.nKept) return;
  const int tid = threadIdx.x, lane = tid % WAVE, wid = tid / WAVE;
  float* out = A.feat + (size_t)s * MLP_IN;
  const int g = A.kept[s], c = A.cur[s];
  __shared__ float fs[MLP_IN];  // this stamp's features
  if (c >= A.count[g]) {  // no substamp: nothing to screen
    if (tid < MLP_IN) {
      out[tid] = 0.0f;
      fs[tid] = 0.0f;
    }
  } else {
    screen_features_body(A, s, g, c, tid, lane, wid, out, fs);
  }
  if (!A.apply) return;
  __syncthreads();
  if (wid == 0) {
    __shared__ float xs[MLP_IN];
    const float* P = A.model;
    if (lane < MLP_IN) xs[lane] = (fs[lane] - P[MP_MU + lane]) / P[MP_SD + lane];
    __builtin_amdgcn_wave_barrier();  // one wave: its LDS accesses run in order
    const float hj = mlp_hidden(P, xs, lane);
    const float z = P[MP_B2] + wave_sum_f(h16(P[MP_W2 + lane]) * hj);
    const float p = 1.0f / (1.0f + expf(-z));
    if (lane == 0) {
      const bool eligible = c < A.count[g] && A.curNow[s] == c;
      A.prob[s] = p;
      A.flag[s] = eligible && p > P[MP_TAU] ? 1 : 0;
    }
  }
}


// ---------------------------------------------------------------------------
// inference: p(reject) per sample (one wave per sample)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float mlp_prob(const float* P, const float* f, float* xs, int lane) {
  if (lane < MLP_IN) xs[lane] = (f[lane] - P[MP_MU + lane]) / P[MP_SD + lane];
  __syncthreads();
  const float hj = mlp_hidden(P, xs, lane);
  const float z = P[MP_B2] + wave_sum_f(h16(P[MP_W2 + lane]) * hj);
  return 1.0f / (1.0f + expf(-z));
}

__global__ __launch_bounds__(WAVE) void k_mlp_forward(const float* __restrict__ P,
                                                      const float* __restrict__ X, int n,
                                                      float* __restrict__ prob) {
  const int k = blockIdx.x;
  if (k >= n) return;
  __shared__ float xs[MLP_IN];
  const float p = mlp_prob(P, X + (size_t)k * MLP_IN, xs, threadIdx.x);
  if (threadIdx.x == 0) prob[k] = p;
}

// apply the screen: flagged stamps move to their next substamp; the refill
// list is rebuilt in ascending stamp order from the pass's own replacements
// (cur != passCur) and the screened ones
__global__ __launch_bounds__(1024) void k_screen_merge(const int* passCur, int* cur,
                                                       const uint8_t* flag, uint8_t* bits,
                                                       int* list, const int* nKept,
                                                       KscScalars* ks, KscScalars* hks) {
  __shared__ int isc[20];
  __shared__ int s_n, s_scr;
  const int tid = threadIdx.x;
  const int n = *nKept;
  if (tid == 0) {
    s_n = 0;
    s_scr = 0;
  }
  __syncthreads();
  for (int base = 0; base < n; base += blockDim.x) {
    const int s = base + tid;
    bool r = false;
    if (s < n) {
      if (flag[s]) {  // screened out: remembered for the chi2 re-check
        bits[s] = (uint8_t)(bits[s] | (1u << cur[s]));
        cur[s] = cur[s] + 1;
        atomicAdd(&s_scr, 1);
      }
      r = cur[s] != passCur[s];
    }
    int tot;
    const int pos = block_prefix_count(r, isc, &tot);
    const int c0 = s_n;
    if (r) list[c0 + pos] = s;
    __syncthreads();
    if (tid == 0) s_n = c0 + tot;
    __syncthreads();
  }
  if (tid == 0) {
    ks->nList = s_n;
    ks->check = s_n > 0 ? 1 : 0;
    ks->nScreened += s_scr;  // over the screened passes (zeroed by the first check)
    publish_pass(*ks, hks);  // the pass's final scalars (after its check's)
  }
}

// ---------------------------------------------------------------------------
// chi2 re-check of the screened substamps, after the loop converged
// ---------------------------------------------------------------------------
// Each stamp with screened substamps left probes its lowest one: its current
// substamp is kept in vOrig and the screened one made current (the fill and
// the chi2 pass that follow see it); the probed stamps are listed in stamp
// order.
__global__ __launch_bounds__(1024) void k_verify_select(const int* nKept, int* cur,
                                                        const uint8_t* bits, int* vOrig,
                                                        int* vList, KscScalars* ks) {
  __shared__ int isc[20];
  __shared__ int s_n;
  const int tid = threadIdx.x;
  const int n = *nKept;
  if (tid == 0) s_n = 0;
  __syncthreads();
  for (int base = 0; base < n; base += blockDim.x) {
    const int s = base + tid;
    const unsigned b = s < n ? bits[s] : 0u;
    if (b) {
      vOrig[s] = cur[s];
      cur[s] = __builtin_ctz(b);
    }
    int tot;
    const int pos = block_prefix_count(b != 0u, isc, &tot);
    const int c0 = s_n;
    if (b) vList[c0 + pos] = s;
    __syncthreads();
    if (tid == 0) s_n = c0 + tot;
    __syncthreads();
  }
  if (tid == 0) {
    ks->nVerify = s_n;
    ks->nRestoredRound = 0;
  }
}

// The reference's rejection tests of a pass (ksc.cl:1-72,
// src/cdkscUtil.cpp:686-770) applied to each probe with the converged fit and
// the last pass's clip: chi2 -1 (no valid pixel) or chi2 - mean > sigKernFit
// sd rejects.  A kept probe stays the stamp's substamp (restored: the loop
// then runs again without the screen); a rejected one is confirmed and the
// stamp returns to its substamp, listed for a refill.  Either way the probed
// bit is cleared.
__global__ __launch_bounds__(1024) void k_verify_decide(const int* vList, int* cur,
                                                        const int* vOrig, uint8_t* bits,
                                                        const double* vSig, int* list,
                                                        KscScalars* ks, double sigKernFit) {
  __shared__ int isc[20];
  __shared__ int s_n, s_kept;
  const int tid = threadIdx.x;
  const int n = ks->nVerify;
  const double mean = ks->clipMean, sd = ks->clipStd;
  if (tid == 0) {
    s_n = 0;
    s_kept = 0;
  }
  __syncthreads();
  for (int base = 0; base < n; base += blockDim.x) {
    const int u = base + tid;
    bool back = false;
    if (u < n) {
      const int s = vList[u];
      const double v = vSig[u];
      const bool reject = v == -1.0 || (v - mean) > sigKernFit * sd;
      bits[s] = (uint8_t)(bits[s] & ~(1u << cur[s]));
      if (reject) {
        cur[s] = vOrig[s];
        back = true;
      } else {
        atomicAdd(&s_kept, 1);
      }
    }
    int tot;
    const int pos = block_prefix_count(back, isc, &tot);
    const int c0 = s_n;
    if (back) list[c0 + pos] = vList[u];
    __syncthreads();
    if (tid == 0) s_n = c0 + tot;
    __syncthreads();
  }
  if (tid == 0) {
    ks->nList = s_n;
    ks->nRestoredRound = s_kept;
    ks->nRestored += s_kept;
    ks->nConfirmed += n - s_kept;
  }
}

// after the loop, over nRec passes of n stamps: a substamp kept by its pass
// that the final fit does not use is 1 when the chi2 test under the converged
// fit rejects it (finalRej, from the recording's probes; all of them when
// finalRej is null) and 3 when it passes that test; 0 = the final fit uses it
__global__ void k_screen_labels(int n, int nRec, const int* passCur, const int* cur,
                                const uint8_t* finalRej, int* label) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nRec) return;
  const int s = e % n;
  if (label[e] == 0 && cur[s] != passCur[e])
    label[e] = (!finalRej || finalRej[s * MAXSS + passCur[e]]) ? 1 : 3;
}

// The substamps a recorded pass used that the converged fit does not: bit c of
// bits[s] for each such substamp c of stamp s (the recording probes them)
__global__ void k_label_bits(int n, int nRec, const int* passCur, const int* cur,
                             const int* kept, const int* count, uint8_t* bits) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  unsigned b = 0;
  const int cnt = count[kept[s]];
  for (int p = 0; p < nRec; p++) {
    const int c = passCur[(size_t)p * n + s];
    if (c != cur[s] && c < cnt) b |= 1u << c;
  }
  bits[s] = (uint8_t)b;
}

// A recording's probe round: the chi2 test of each probe under the converged
// fit and the last pass's clip (the rule of k_verify_decide) is stored in
// finalRej[s][c]; every stamp returns to its own substamp and is listed for a
// refill.
__global__ __launch_bounds__(1024) void k_label_decide(const int* vList, int* cur,
                                                       const int* vOrig, uint8_t* bits,
                                                       const double* vSig, int* list,
                                                       KscScalars* ks, double sigKernFit,
                                                       uint8_t* finalRej) {
  const int n = ks->nVerify;
  const double mean = ks->clipMean, sd = ks->clipStd;
  for (int u = threadIdx.x; u < n; u += blockDim.x) {
    const int s = vList[u], c = cur[s];
    const double v = vSig[u];
    finalRej[s * MAXSS + c] = (v == -1.0 || (v - mean) > sigKernFit * sd) ? 1 : 0;
    bits[s] = (uint8_t)(bits[s] & ~(1u << c));
    cur[s] = vOrig[s];
    list[u] = s;
  }
  if (threadIdx.x == 0) ks->nList = n;
}

// ---------------------------------------------------------------------------
// training: gradient of the weighted mean binary cross-entropy
// ---------------------------------------------------------------------------
// Block b handles samples [b*CHUNK, (b+1)*CHUNK); lane j = hidden unit j.
// Partial sums per block: dW1[64][32], db1[64], dw2[64], db2, loss.
__global__ __launch_bounds__(WAVE) void k_mlp_grad(const float* __restrict__ P,
                                                   const float* __restrict__ X,
                                                   const float* __restrict__ y,
                                                   const float* __restrict__ wt, int n,
                                                   float invN, float* __restrict__ part) {
  const int j = threadIdx.x;
  const int k0 = blockIdx.x * MLP_CHUNK;
  __shared__ float xs[MLP_CHUNK][MLP_IN + 1];
  for (int e = j; e < MLP_CHUNK * MLP_IN; e += WAVE) {
    const int k = e / MLP_IN, i = e % MLP_IN;
    float v = 0.0f;
    if (k0 + k < n) v = (X[(size_t)(k0 + k) * MLP_IN + i] - P[MP_MU + i]) / P[MP_SD + i];
    xs[k][i] = v;
  }
  __syncthreads();
  float w1[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; i++) w1[i] = h16(P[MP_W1 + j * MLP_IN + i]);
  const float b1 = P[MP_B1 + j], w2 = h16(P[MP_W2 + j]), b2 = P[MP_B2];
  float gw1[MLP_IN];
#pragma unroll
  for (int i = 0; i < MLP_IN; i++) gw1[i] = 0.0f;
  float gb1 = 0.0f, gw2 = 0.0f, gb2 = 0.0f, loss = 0.0f;
  const int kn = min(MLP_CHUNK, n - k0);
  for (int k = 0; k < kn; k++) {
    float a = b1;
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) a += w1[i] * xs[k][i];
    const float hj = fmaxf(a, 0.0f);
    const float z = b2 + wave_sum_f(w2 * hj);
    const float p = 1.0f / (1.0f + expf(-z));
    const float yk = y[k0 + k], wk = wt[k0 + k] * invN;
    const float dz = wk * (p - yk);
    // -[y log p + (1-y) log(1-p)] from z without cancellation
    const float l = fmaxf(z, 0.0f) - z * yk + log1pf(expf(-fabsf(z)));
    loss += wk * l;
    gb2 += dz;
    gw2 += dz * hj;
    const float dp = a > 0.0f ? dz * w2 : 0.0f;
    gb1 += dp;
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) gw1[i] += dp * xs[k][i];
  }
  float* o = part + (size_t)blockIdx.x * (MLP_NGRAD + 1);
#pragma unroll
  for (int i = 0; i < MLP_IN; i++) o[j * MLP_IN + i] = gw1[i];
  o[MLP_HID * MLP_IN + j] = gb1;
  o[MLP_HID * MLP_IN + MLP_HID + j] = gw2;
  if (j == 0) {
    o[MLP_NGRAD - 1] = gb2;
    o[MLP_NGRAD] = loss;
  }
}

// Block partials of the allocation-free gradient (xb_mlp_grad_noalloc): one
// row of MLP_NGRAD + 1 floats per gradient block, so at most
// XB_MLP_NOALLOC_MAX / MLP_CHUNK blocks (2.2 MB per device, held until the
// process ends).  Host code never passes this name to a kernel: it uses the
// per-device address from hipGetSymbolAddress (scratch_address below).
constexpr int MLP_SCRATCH_BLOCKS = XB_MLP_NOALLOC_MAX / MLP_CHUNK;
static_assert(XB_MLP_NOALLOC_MAX % MLP_CHUNK == 0, "whole blocks");
static __device__ __align__(256) float scratch[(size_t)MLP_SCRATCH_BLOCKS * (MLP_NGRAD + 1)];

// sum of the block partials in block order (deterministic); out[MLP_NGRAD]
// receives the loss
__global__ void k_mlp_reduce(const float* __restrict__ part, int nb, float* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > MLP_NGRAD) return;
  float a = 0.0f;
  for (int b = 0; b < nb; b++) a += part[(size_t)b * (MLP_NGRAD + 1) + q];
  out[q] = a;
}

// Adam on the trainable parameters (W1, b1, w2, b2 are contiguous in P)
__global__ void k_mlp_adam(float* __restrict__ P, const float* __restrict__ g,
                           float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                           float b2, float c1, float c2, float eps) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= MLP_NGRAD) return;
  const float gq = g[q];
  const float mq = b1 * m[q] + (1.0f - b1) * gq;
  const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  P[MP_W1 + q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

// ---------------------------------------------------------------------------
// launchers used by the pipeline
// ---------------------------------------------------------------------------
hipError_t launch_screen(Ctx& c, int im, int pass, bool apply) {
  const int N = c.nKept[im];
  if (N <= 0 || pass >= SCR_PASSES) return hipSuccess;
  int* passCur = c.scr.passCur + (size_t)pass * N;
  ScreenArgs A;
  A.T = c.img[im];
  A.S = c.img[1 - im];
  A.w = c.w;
  A.h = c.h;
  A.nKept = &c.sc->nKept[im];
  A.kept = c.st[im].kept;
  A.cur = passCur;
  A.count = c.st[im].count;
  A.ssx = c.st[im].ssx;
  A.ssy = c.st[im].ssy;
  A.skyT = c.st[im].sky;
  A.skyS = c.st[1 - im].sky;
  A.sdT = c.st[im].fwhm;
  A.sdS = c.st[1 - im].fwhm;
  A.sig = c.st[im].sigK;
  A.ks = &c.sc->k[im];
  A.threshHigh = c.m.threshHigh;
  A.pass = pass;
  A.prevCur = pass > 0 ? passCur - N : nullptr;
  A.zPrev = c.scr.zPrev;
  A.feat = c.scr.feat + (size_t)pass * N * MLP_IN;
  // (the pass's labels come from its check, k_ksc_check)
  A.apply = apply ? 1 : 0;
  A.model = c.scr.model;
  A.curNow = c.st[im].cur;
  A.prob = c.scr.prob + (size_t)pass * N;
  A.flag = c.scr.flag;
  hipLaunchKernelGGL(k_screen_features, dim3(N), dim3(SCR_THREADS), 0, c.cur, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !apply) return e;
  hipLaunchKernelGGL(k_screen_merge, dim3(1), dim3(1024), 0, c.cur, passCur, c.st[im].cur,
                     c.scr.flag, c.scr.bits, c.st[im].listK, &c.sc->nKept[im], &c.sc->k[im],
                     &c.hscDev->k[im]);
  return hipGetLastError();
}

hipError_t launch_screen_verify_select(Ctx& c, int im) {
  hipLaunchKernelGGL(k_verify_select, dim3(1), dim3(1024), 0, c.cur, &c.sc->nKept[im],
                     c.st[im].cur, c.scr.bits, c.scr.vOrig, c.scr.vList, &c.sc->k[im]);
  return hipGetLastError();
}

hipError_t launch_screen_verify_decide(Ctx& c, int im) {
  hipLaunchKernelGGL(k_verify_decide, dim3(1), dim3(1024), 0, c.cur, c.scr.vList,
                     c.st[im].cur, c.scr.vOrig, c.scr.bits, c.scr.vSig, c.st[im].listK,
                     &c.sc->k[im], c.m.sigKernFit);
  return hipGetLastError();
}

hipError_t launch_screen_labels(Ctx& c, int im, int nRec, bool finalFit) {
  const int N = c.nKept[im];
  if (N <= 0 || nRec <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_screen_labels, dim3((N * nRec + 255) / 256), dim3(256), 0, c.cur, N,
                     nRec, c.scr.passCur, c.st[im].cur, finalFit ? c.scr.finalRej : nullptr,
                     c.scr.label);
  return hipGetLastError();
}

hipError_t launch_label_bits(Ctx& c, int im, int nRec) {
  const int N = c.nKept[im];
  if (N <= 0 || nRec <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_label_bits, dim3((N + 255) / 256), dim3(256), 0, c.cur, N, nRec,
                     c.scr.passCur, c.st[im].cur, c.st[im].kept, c.st[im].count, c.scr.bits);
  return hipGetLastError();
}

hipError_t launch_label_decide(Ctx& c, int im) {
  hipLaunchKernelGGL(k_label_decide, dim3(1), dim3(1024), 0, c.cur, c.scr.vList, c.st[im].cur,
                     c.scr.vOrig, c.scr.bits, c.scr.vSig, c.st[im].listK, &c.sc->k[im],
                     c.m.sigKernFit, c.scr.finalRej);
  return hipGetLastError();
}

// Device address of `scratch` on `device`, looked up once per device with that
// device current (each GPU holds its own copy of the array at its own
// address).  Called by xb_ctx_create, so the allocation-free entry point only
// reads the cached value.
namespace {
std::mutex g_scratchMu;
float* g_scratch[XB_MAX_DEVICES] = {};
}  // namespace

hipError_t scratch_address(int device, float** out) {
  *out = nullptr;
  if (device < 0 || device >= XB_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_scratchMu);
  if (!g_scratch[device]) {
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e != hipSuccess) return e;
    if (prev != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<float*>(p);
  }
  *out = g_scratch[device];
  return hipSuccess;
}

}  // namespace xb

// ===========================================================================
// C ABI of the network (training and checks)
// ===========================================================================
using namespace xb;

namespace {

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)locked_free(p);
  }
  hipError_t alloc(size_t bytes) { return locked_malloc(&p, bytes); }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

#define MLP_HIP(expr)                                    \
  do {                                                   \
    hipError_t _e = (expr);                              \
    if (_e != hipSuccess) {                              \
      c.err = std::string(#expr) + ": " + hipGetErrorString(_e); \
      return XB_ERR_HIP;                                 \
    }                                                    \
  } while (0)

}  // namespace

extern "C" {

int xb_mlp_forward(xb_ctx* x, const float* params, const float* X, int n, float* prob) {
  if (!x || !params || n < 0 || (n > 0 && (!X || !prob))) return XB_ERR_ARG;
  Ctx& c = x->c;
  if (n == 0) return XB_OK;
  MLP_HIP(hipSetDevice(c.device));
  DevBuf dP, dX, dp;
  MLP_HIP(dP.alloc(sizeof(float) * MLP_NPARAM));
  MLP_HIP(dX.alloc(sizeof(float) * (size_t)n * MLP_IN));
  MLP_HIP(dp.alloc(sizeof(float) * (size_t)n));
  MLP_HIP(hipMemcpyAsync(dP.p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice,
                         c.stream));
  MLP_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice,
                         c.stream));
  hipLaunchKernelGGL(k_mlp_forward, dim3(n), dim3(WAVE), 0, c.stream, dP.as<float>(),
                     dX.as<float>(), n, dp.as<float>());
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipMemcpyAsync(prob, dp.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost,
                         c.stream));
  MLP_HIP(hipStreamSynchronize(c.stream));
  return XB_OK;
}

int xb_mlp_grad(xb_ctx* x, const float* params, const float* X, const float* y,
                const float* wt, int n, float* grad) {
  if (!x || !params || !grad || n <= 0 || !X || !y || !wt) return XB_ERR_ARG;
  Ctx& c = x->c;
  MLP_HIP(hipSetDevice(c.device));
  const int nb = (n + MLP_CHUNK - 1) / MLP_CHUNK;
  DevBuf dP, dX, dy, dw, dpart, dg;
  MLP_HIP(dP.alloc(sizeof(float) * MLP_NPARAM));
  MLP_HIP(dX.alloc(sizeof(float) * (size_t)n * MLP_IN));
  MLP_HIP(dy.alloc(sizeof(float) * (size_t)n));
  MLP_HIP(dw.alloc(sizeof(float) * (size_t)n));
  MLP_HIP(dpart.alloc(sizeof(float) * (size_t)nb * (MLP_NGRAD + 1)));
  MLP_HIP(dg.alloc(sizeof(float) * (MLP_NGRAD + 1)));
  MLP_HIP(hipMemcpyAsync(dP.p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice,
                         c.stream));
  MLP_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice,
                         c.stream));
  MLP_HIP(hipMemcpyAsync(dy.p, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c.stream));
  MLP_HIP(hipMemcpyAsync(dw.p, wt, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(WAVE), 0, c.stream, dP.as<float>(),
                     dX.as<float>(), dy.as<float>(), dw.as<float>(), n, 1.0f / n,
                     dpart.as<float>());
  MLP_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NGRAD + 256) / 256), dim3(256), 0, c.stream,
                     dpart.as<float>(), nb, dg.as<float>());
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipMemcpyAsync(grad, dg.p, sizeof(float) * (MLP_NGRAD + 1), hipMemcpyDeviceToHost,
                         c.stream));
  MLP_HIP(hipStreamSynchronize(c.stream));
  return XB_OK;
}

// xb_mlp_grad on device-resident arrays, with the block partials in the
// library's static per-device `scratch` instead of an allocation: no hipMalloc,
// hipFree or other allocating call.  Same kernels, same block order, so the
// result equals xb_mlp_grad's bit for bit.  Enqueued on the context's primary
// stream; the caller waits with xb_sync.
int xb_mlp_grad_noalloc(xb_ctx* x, const float* d_params, const float* d_X, const float* d_y,
                        const float* d_wt, int n, float* d_grad) {
  if (!x || !d_params || !d_X || !d_y || !d_wt || !d_grad || n <= 0) return XB_ERR_ARG;
  Ctx& c = x->c;
  if (n > XB_MLP_NOALLOC_MAX) {
    c.err = "xb_mlp_grad_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
            std::to_string(XB_MLP_NOALLOC_MAX) + " samples";
    return XB_ERR_CAPACITY;
  }
  if (!c.mlpScratch) {
    c.err = "xb_mlp_grad_noalloc: no scratch address for this device";
    return XB_ERR_HIP;
  }
  MLP_HIP(hipSetDevice(c.device));
  const int nb = (n + MLP_CHUNK - 1) / MLP_CHUNK;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(WAVE), 0, c.stream, d_params, d_X, d_y, d_wt,
                     n, 1.0f / n, c.mlpScratch);
  MLP_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NGRAD + 256) / 256), dim3(256), 0, c.stream,
                     c.mlpScratch, nb, d_grad);
  MLP_HIP(hipGetLastError());
  return XB_OK;
}

int xb_mlp_train(xb_ctx* x, float* params, const float* X, const float* y, const float* wt,
                 int n, int epochs, int batch, float lr, uint64_t seed, float* loss) {
  if (!x || !params || n <= 0 || !X || !y || !wt || epochs <= 0 || batch <= 0) return XB_ERR_ARG;
  Ctx& c = x->c;
  MLP_HIP(hipSetDevice(c.device));
  const int maxNb = (batch + MLP_CHUNK - 1) / MLP_CHUNK;
  DevBuf dP, dX, dy, dw, dpart, dg, dm, dv;
  MLP_HIP(dP.alloc(sizeof(float) * MLP_NPARAM));
  MLP_HIP(dX.alloc(sizeof(float) * (size_t)n * MLP_IN));
  MLP_HIP(dy.alloc(sizeof(float) * (size_t)n));
  MLP_HIP(dw.alloc(sizeof(float) * (size_t)n));
  MLP_HIP(dpart.alloc(sizeof(float) * (size_t)maxNb * (MLP_NGRAD + 1)));
  MLP_HIP(dg.alloc(sizeof(float) * (MLP_NGRAD + 1)));
  MLP_HIP(dm.alloc(sizeof(float) * MLP_NGRAD));
  MLP_HIP(dv.alloc(sizeof(float) * MLP_NGRAD));
  std::vector<float> zero(MLP_NGRAD, 0.0f);
  MLP_HIP(hipMemcpyAsync(dm.p, zero.data(), sizeof(float) * MLP_NGRAD, hipMemcpyHostToDevice,
                         c.stream));
  MLP_HIP(hipMemcpyAsync(dv.p, zero.data(), sizeof(float) * MLP_NGRAD, hipMemcpyHostToDevice,
                         c.stream));
  MLP_HIP(hipMemcpyAsync(dP.p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice,
                         c.stream));
  // each epoch visits the samples in a fresh seeded order (host shuffle)
  std::mt19937_64 rng(seed);
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::vector<float> hX((size_t)n * MLP_IN), hy(n), hw(n);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int t = 0;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(idx.begin(), idx.end(), rng);
    for (int i = 0; i < n; i++) {
      std::memcpy(&hX[(size_t)i * MLP_IN], &X[(size_t)idx[i] * MLP_IN], sizeof(float) * MLP_IN);
      hy[i] = y[idx[i]];
      hw[i] = wt[idx[i]];
    }
    MLP_HIP(hipMemcpyAsync(dX.p, hX.data(), sizeof(float) * hX.size(), hipMemcpyHostToDevice,
                           c.stream));
    MLP_HIP(hipMemcpyAsync(dy.p, hy.data(), sizeof(float) * n, hipMemcpyHostToDevice, c.stream));
    MLP_HIP(hipMemcpyAsync(dw.p, hw.data(), sizeof(float) * n, hipMemcpyHostToDevice, c.stream));
    double epLoss = 0.0;
    for (int k0 = 0; k0 < n; k0 += batch) {
      const int bn = std::min(batch, n - k0);
      const int nb = (bn + MLP_CHUNK - 1) / MLP_CHUNK;
      hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(WAVE), 0, c.stream, dP.as<float>(),
                         dX.as<float>() + (size_t)k0 * MLP_IN, dy.as<float>() + k0,
                         dw.as<float>() + k0, bn, 1.0f / bn, dpart.as<float>());
      MLP_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NGRAD + 256) / 256), dim3(256), 0, c.stream,
                         dpart.as<float>(), nb, dg.as<float>());
      MLP_HIP(hipGetLastError());
      t++;
      const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
      hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NGRAD + 255) / 256), dim3(256), 0, c.stream,
                         dP.as<float>(), dg.as<float>(), dm.as<float>(), dv.as<float>(), lr, b1,
                         b2, c1, c2, eps);
      MLP_HIP(hipGetLastError());
      if (loss) {
        float l = 0.0f;
        MLP_HIP(hipMemcpyAsync(&l, dg.as<float>() + MLP_NGRAD, sizeof(float),
                               hipMemcpyDeviceToHost, c.stream));
        MLP_HIP(hipStreamSynchronize(c.stream));
        epLoss += (double)l * bn;
      }
    }
    if (loss) loss[ep] = (float)(epLoss / n);
  }
  MLP_HIP(hipMemcpyAsync(params, dP.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost,
                         c.stream));
  MLP_HIP(hipStreamSynchronize(c.stream));
  return XB_OK;
}

}  // extern "C"
