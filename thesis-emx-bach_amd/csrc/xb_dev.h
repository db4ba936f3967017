// xb_dev.h — device helpers for libxbach kernels (gfx950, wave64).
//
// All fit kernels are compiled with -ffp-contract=off: every FP64/FP32
// expression rounds exactly where the reference's C expression does, which is
// what makes the fit stages bit-identical to the CPU restatement.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xb_internal.h"

namespace xb {

// The arguments of a launch that covers N images at once: image z's work is
// blockIdx.z's (both images of a pair in one dispatch, so a one-stream
// context runs them side by side and a pair makes half the dispatches).
template <typename T, int N>
struct ArgsN {
  T a[N];
};
template <typename T, int N>
__device__ __forceinline__ const T& args_of(const ArgsN<T, N>& P) {
  return P.a[N == 1 ? 0 : blockIdx.z];
}

// A pass's scalars (check, the refill list's length, chi2 count, screened
// count) written into the host's pinned mirror (host memory the device
// addresses) by the kernel that set them last, followed by a system-scope
// fence: the host reads them once the stream's event has completed, without a
// copy dispatch.  One thread calls this.
__device__ __forceinline__ void publish_pass(const KscScalars& ks, KscScalars* host) {
  if (!host) return;
  host->check = ks.check;
  host->nList = ks.nList;
  host->nChi2 = ks.nChi2;
  host->nScreened = ks.nScreened;
  __threadfence_system();
}

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// std::max(a, b) semantics: returns b when a < b, otherwise a (NaN keeps a).
__device__ __forceinline__ double std_max(double a, double b) { return (a < b) ? b : a; }

// Exclusive prefix count of `flag` over the block (blockDim.x multiple of 64,
// at most 1024).  Returns the prefix; *total receives the block total.
// scratch: >= 17 ints of LDS.
__device__ __forceinline__ int block_prefix_count(bool flag, int* scratch, int* total) {
  const int tid = threadIdx.x;
  const int wid = tid / WAVE;
  const int nw = blockDim.x / WAVE;
  unsigned long long b = __ballot(flag);
  int lane = lane_id();
  unsigned long long below = lane == 0 ? 0ull : (b & ((~0ull) >> (64 - lane)));
  int pre = __popcll(below);
  if (lane == 0) scratch[wid] = __popcll(b);
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int i = 0; i < nw; i++) {
      int v = scratch[i];
      scratch[i] = acc;
      acc += v;
    }
    scratch[16] = acc;
  }
  __syncthreads();
  int res = scratch[wid] + pre;
  *total = scratch[16];
  __syncthreads();
  return res;
}

// Exclusive prefix sum of a non-negative int over the block (thread order).
// *total receives the block total.  scratch >= 17 ints.
__device__ __forceinline__ int block_prefix_sum(int v, int* scratch, int* total) {
  const int tid = threadIdx.x;
  const int wid = tid / WAVE;
  const int nw = blockDim.x / WAVE;
  const int lane = lane_id();
  int incl = v;
  for (int o = 1; o < WAVE; o <<= 1) {
    int t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == WAVE - 1) scratch[wid] = incl;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int i = 0; i < nw; i++) {
      int t = scratch[i];
      scratch[i] = acc;
      acc += t;
    }
    scratch[16] = acc;
  }
  __syncthreads();
  int res = scratch[wid] + incl - v;
  *total = scratch[16];
  __syncthreads();
  return res;
}

// Block-wide sum of an int (order-independent).  scratch >= 17 ints.
__device__ __forceinline__ int block_sum_int(int v, int* scratch) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int wid = threadIdx.x / WAVE;
  if (lane_id() == 0) scratch[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int i = 0; i < (int)(blockDim.x / WAVE); i++) acc += scratch[i];
    scratch[16] = acc;
  }
  __syncthreads();
  int r = scratch[16];
  __syncthreads();
  return r;
}

// Block-wide max / min of an int.
__device__ __forceinline__ int block_max_int(int v, int* scratch) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  const int wid = threadIdx.x / WAVE;
  if (lane_id() == 0) scratch[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = scratch[0];
    for (int i = 1; i < (int)(blockDim.x / WAVE); i++) acc = max(acc, scratch[i]);
    scratch[16] = acc;
  }
  __syncthreads();
  int r = scratch[16];
  __syncthreads();
  return r;
}

__device__ __forceinline__ int block_min_int(int v, int* scratch) {
  return -block_max_int(-v, scratch);
}

// Sigma clip of a short array held in LDS (n <= 1024): the device form of
// src/bachUtil.cpp:34-128 with cl_kern/bach.cl:196-251 — 32-element block
// partial sums in list order, then their sequential sum.  All threads call it;
// results are returned to every thread.  clip: n bytes of LDS scratch,
// part: 2*32 doubles of LDS (n <= 1024 -> <= 32 blocks).
__device__ void sigma_clip_small(const double* data, int n, int maxIter, double alpha,
                                 uint8_t* clip, double* part, double* out /*2*/,
                                 int* iscratch) {
  const int tid = threadIdx.x;
  if (n == 0) {
    if (tid == 0) { out[0] = 0.0; out[1] = 1e10; }
    __syncthreads();
    return;
  }
  for (int i = tid; i < n; i += blockDim.x) clip[i] = 0;
  __shared__ long long s_cur, s_prev;
  __shared__ double s_mean, s_inv;
  if (tid == 0) { s_cur = 0; s_prev = n; }
  __syncthreads();
  const int nb = (n + 31) / 32;
  for (int it = 0; it < maxIter; it++) {
    if (s_cur == s_prev) break;
    if (s_prev <= 1) {
      if (tid == 0) { out[0] = 0.0; out[1] = 1e10; }
      __syncthreads();
      return;
    }
    if (tid < nb) {
      double s = 0.0, s2 = 0.0;
      int cnt = min(32, n - tid * 32);
      for (int i = 0; i < cnt; i++) {
        int k = tid * 32 + i;
        double d = clip[k] == 0 ? data[k] : 0.0;
        s += d;
        s2 += d * d;
      }
      part[tid] = s;
      part[32 + tid] = s2;
    }
    __syncthreads();
    if (tid == 0) {
      s_cur = s_prev;
      double sum = 0.0, sum2 = 0.0;
      for (int g = 0; g < nb; g++) sum += part[g];
      for (int g = 0; g < nb; g++) sum2 += part[32 + g];
      double np = (double)s_prev;
      double m = sum / np;
      double sd = sqrt((sum2 - np * m * m) / (double)(s_prev - 1));
      s_mean = m;
      s_inv = 1.0 / sd;
      out[0] = m;
      out[1] = sd;
    }
    __syncthreads();
    const double m = s_mean, inv = s_inv;
    int c = 0;
    for (int k = tid; k < n; k += blockDim.x)
      if (clip[k] == 0 && fabs(data[k] - m) * inv > alpha) {
        clip[k] = 1;
        c++;
      }
    int tot = block_sum_int(c, iscratch);
    if (tid == 0) s_prev = s_cur - tot;
    __syncthreads();
  }
  __syncthreads();
}

// Mask bit OR into the 16-bit mask with a 32-bit atomic (pixels of two
// stamps' substamp windows can share a word).
__device__ __forceinline__ void mask_or_atomic(uint16_t* mask, size_t idx, uint16_t bits) {
  uintptr_t a = (uintptr_t)(mask + idx);
  unsigned* w = (unsigned*)(a & ~(uintptr_t)3);
  unsigned sh = (a & 2) ? 16u : 0u;
  atomicOr(w, ((unsigned)bits) << sh);
}

// ---- cmv: the stamps one fill launch covers (plain and mixed forms) ----
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
  double* Wc;  // W in column-major pixel order (plain form's target products)
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

// the mixed-precision form of the fill (k_cmv_mixed.hip)
hipError_t launch_fill_mixed(Ctx& c, const FillArgs& A, int nMax);
// both images' full fills in one dispatch per kernel (image z = blockIdx.z)
hipError_t launch_fill_mixed_pair(Ctx& c, const FillArgs& A0, const FillArgs& A1, int nMax);

}  // namespace xb
