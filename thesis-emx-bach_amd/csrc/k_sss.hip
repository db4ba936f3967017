// k_sss.hip — stamp statistics and substamp selection (the reference's sss).
//
// One 256-thread workgroup per stamp.  The reference runs these steps as one
// serial work-item per stamp (cl_kern/sss.cl:81-505) with ~1600 host round
// trips per image for the per-stamp sigma clip (src/bachUtil.cpp:255-266).
// Here each stamp's whole chain runs inside its workgroup:
//   * ran1 draws are identical for every stamp (seed -666 reset per stamp,
//     sss.cl:92-95), so they come from a table; candidate draws are tested in
//     parallel and the first 100 valid ones are kept in draw order;
//   * good pixels are gathered in raster order (the canonical order of the
//     reference's atomic_inc list, sss.cl:184) with a block prefix count;
//   * the sigma clip keeps the reference's 32-element block sums and their
//     sequential sum, so mean/stddev are bit-identical;
//   * histogram bins are integer counts (order-free, LDS atomics); the mode
//     and quartile walk is the reference's serial code on one thread;
//   * findSubStamps keeps the sequential scan: each threshold level's pixels
//     are classified by all waves (coalesced, raster-ordered candidate lists)
//     and one wave then processes the candidates in raster order, each with a
//     wave-parallel window search and check.
#include "xb_dev.h"

namespace xb {

struct StampGeom {
  int ox, oy, sw, sh;
};

__device__ __forceinline__ StampGeom stamp_geom(int s, int nx, int ny, int w, int h, int fsw) {
  StampGeom g;
  int sx = s % nx, sy = s / nx;
  g.ox = sx * w / nx;
  g.oy = sy * h / ny;
  g.sw = min(g.ox + fsw, w) - g.ox;
  g.sh = min(g.oy + fsw, h) - g.oy;
  return g;
}

// ---------------------------------------------------------------------------
// calcStats (src/bachUtil.cpp:130-280)
// ---------------------------------------------------------------------------
// n / d for n < 2^20 and 2 <= d <= 1024 by a reciprocal multiply (exact over
// that range: every pair was checked against integer division).
__device__ __forceinline__ int fdiv(int n, unsigned m) { return (int)__umulhi((unsigned)n, m); }
__host__ __device__ __forceinline__ unsigned fdiv_magic(int d) { return 0xFFFFFFFFu / (unsigned)d + 1u; }

constexpr int MAX_CLIP_BLOCKS = 2560;  // ceil(fsw^2 / 32) for fsw <= 286
constexpr int MAX_LEVELS = 32;         // findSubStamps threshold levels


// Block-wide helpers of k_sss_stats (256 threads).
// Inclusive prefix maximum over thread order.  scratch: >= 17 ints.
__device__ __forceinline__ int block_prefix_max_incl(int v, int* scratch) {
  const int lane = lane_id(), wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const int t = __shfl_up(v, o);
    if (lane >= o) v = max(v, t);
  }
  if (lane == WAVE - 1) scratch[wid] = v;
  __syncthreads();
  int before = INT_MIN;
  for (int i = 0; i < wid && i < nw; i++) before = max(before, scratch[i]);
  __syncthreads();
  return max(v, before);
}
// (value, index) maximum, ties to the smaller index; every thread gets the
// index.  dsc: >= 17 doubles, isc: >= 17 ints.
__device__ __forceinline__ int block_argmax_first(double v, int idx, double* dsc, int* isc) {
  const int lane = lane_id(), wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(v, o);
    const int oi = __shfl_xor(idx, o);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  if (lane == 0) {
    dsc[wid] = v;
    isc[wid] = idx;
  }
  __syncthreads();
  double bv = dsc[0];
  int bi = isc[0];
  for (int i = 1; i < nw; i++)
    if (dsc[i] > bv || (dsc[i] == bv && isc[i] < bi)) {
      bv = dsc[i];
      bi = isc[i];
    }
  __syncthreads();
  return bi;
}
// Sum of 64-bit integers (exact in any order).
__device__ __forceinline__ long long block_sum_ll(long long v, long long* scratch) {
  const int lane = lane_id(), wid = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  long long t = 0;
  for (int i = 0; i < nw; i++) t += scratch[i];
  __syncthreads();
  return t;
}
// Sequential sum of a[0..n-1] in index order (one thread), its operands loaded
// one 8-block ahead; a[n .. n+15] must hold +0.0 (adding +0.0 to a partial sum
// that is never -0 changes nothing).
__device__ __forceinline__ double ordered_sum(const double* a, int n) {
  double s = 0.0;
  double c[8], d[8];
#pragma unroll
  for (int t = 0; t < 8; t++) c[t] = a[t];
  for (int b = 0; b < n; b += 16) {
#pragma unroll
    for (int t = 0; t < 8; t++) d[t] = a[b + 8 + t];
#pragma unroll
    for (int t = 0; t < 8; t++) s += c[t];
    if (b + 8 >= n) break;
#pragma unroll
    for (int t = 0; t < 8; t++) c[t] = a[b + 16 + t];
#pragma unroll
    for (int t = 0; t < 8; t++) s += d[t];
  }
  return s;
}

template <int NT>
__global__ __launch_bounds__(NT) void k_sss_stats(const float* __restrict__ img,
                                                   uint16_t* __restrict__ mask, int w, int h,
                                                   int nx, int ny, int fsw,
                                                   const double* __restrict__ ran1,
                                                   float* __restrict__ goodAll, double alpha,
                                                   double iqRange, double* __restrict__ skyOut,
                                                   double* __restrict__ fwhmOut) {
  const int s = blockIdx.x;
  const int tid = threadIdx.x;
  const StampGeom g = stamp_geom(s, nx, ny, w, h, fsw);
  __shared__ double smp[NSAMP_PAD];
  __shared__ int isc[20];
  __shared__ int s_cnt, s_nGood;
  __shared__ double bs[MAX_CLIP_BLOCKS + 16], bs2[MAX_CLIP_BLOCKS + 16];
  __shared__ double s_mean, s_sd, s_inv, s_sum, s_sum2;
  __shared__ long long s_cur, s_prev;
  __shared__ int bins[256];

  // ---- sampleStamp (sss.cl:81-116) + pad (sss.cl:118-125)
  for (int i = tid; i < NSAMP_PAD; i += blockDim.x) smp[i] = i < NSAMP ? 0.0 : INFINITY;
  if (tid == 0) s_cnt = 0;
  __syncthreads();
  const int npix = g.sw * g.sh;
  for (int base = 0; base < npix; base += blockDim.x) {
    if (s_cnt >= NSAMP) break;
    int t = base + tid;
    bool valid = false;
    double v = 0.0;
    if (t < npix) {
      int rx = (int)floor(ran1[2 * t] * (int)g.sw);
      int ry = (int)floor(ran1[2 * t + 1] * (int)g.sh);
      size_t idx = (size_t)(rx + g.ox) + (size_t)(ry + g.oy) * w;
      v = (double)img[idx];
      valid = !(mask[idx] > 0 || fabs(v) <= 1e-10);
    }
    int tot;
    int pos = block_prefix_count(valid, isc, &tot);
    int c0 = s_cnt;
    if (valid && c0 + pos < NSAMP) smp[c0 + pos] = v;
    __syncthreads();
    if (tid == 0) s_cnt = min(NSAMP, c0 + tot);
    __syncthreads();
  }
  const int sampleCount = s_cnt;
  // ---- bitonic sort network (sss.cl:135-150, bachUtil.cpp:229-240)
  for (int k = 2; k <= NSAMP_PAD; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (tid < NSAMP_PAD) {
        int i = tid, ixj = i ^ j;
        if (ixj > i) {
          double a = smp[i], b = smp[ixj];
          if ((i & k) == 0 && a > b) { smp[i] = b; smp[ixj] = a; }
          if ((i & k) != 0 && a < b) { smp[i] = b; smp[ixj] = a; }
        }
      }
      __syncthreads();
    }

  // ---- maskStamp (sss.cl:157-188): good pixels in raster order.  Each wave
  // owns a raster segment of the stamp: it counts its good pixels (flagging
  // NaN pixels in the mask), the segment offsets follow from the counts, and
  // a second scan writes the values at the wave's offset.  The second scan
  // sees a NaN pixel either flagged or NaN: rejected either way.
  float* good = goodAll + (size_t)s * fsw * fsw;
  const int nAll = fsw * fsw;
  const unsigned mf = fdiv_magic(fsw);
  const int lane = lane_id(), wid = tid / WAVE;
  constexpr int NWS = NT / WAVE;
  const int seg = ((nAll + NWS - 1) / NWS + WAVE - 1) / WAVE * WAVE;
  const int k0 = wid * seg, k1 = min(nAll, k0 + seg);
  __shared__ int wcnt[NWS];
  auto scan = [&](bool write, int off) -> int {
    int cnt = 0;
    for (int base = k0; base < k1; base += 8 * WAVE) {
      float pv[8];
      uint16_t m[8];
      size_t ix[8];
      bool in[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int k = base + u * WAVE + lane;
        const int kc = min(k, k1 - 1);
        const int y = fdiv(kc, mf), x = kc - y * fsw;
        const int xI = x + g.ox, yI = y + g.oy;
        in[u] = k < k1 && xI < w && yI < h;
        ix[u] = (size_t)min(xI, w - 1) + (size_t)min(yI, h - 1) * w;
        pv[u] = img[ix[u]];
        m[u] = mask[ix[u]];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const double pixel = (double)pv[u];
        bool ok = false;
        if (in[u] && !(m[u] > 0 || pixel <= 1e-10)) {
          if (isnan(pixel)) {
            if (!write) mask[ix[u]] = m[u] | (XB_MASK_NAN_PIXEL | XB_MASK_BAD_INPUT);
          } else {
            ok = true;
          }
        }
        const unsigned long long bal = __ballot(ok);
        if (write && ok) good[off + cnt + __popcll(bal & ((1ull << lane) - 1ull))] = pv[u];
        cnt += __popcll(bal);
      }
    }
    return cnt;
  };
  {
    const int cnt = scan(false, 0);
    if (lane == 0) wcnt[wid] = cnt;
    __syncthreads();
    int off = 0, tot = 0;
    for (int i = 0; i < NWS; i++) {
      if (i < wid) off += wcnt[i];
      tot += wcnt[i];
    }
    scan(true, off);
    if (tid == 0) s_nGood = tot;
    __syncthreads();
  }
  const int n = s_nGood;

  // ---- sigmaClip(maxIter 3) (bachUtil.cpp:34-128, bach.cl:196-251).
  // Clipped entries are marked by negating the (positive) stored value.
  if (tid == 0) {
    s_cur = 0;
    s_prev = n;
    s_mean = 0.0;
    s_sd = 1e10;
  }
  __syncthreads();
  if (n > 0) {
    const int nb = (n + 31) / 32;
    for (int it = 0; it < 3; it++) {
      if (s_cur == s_prev) break;
      if (s_prev <= 1) {
        if (tid == 0) { s_mean = 0.0; s_sd = 1e10; }
        __syncthreads();
        break;
      }
      for (int b = tid; b < nb; b += blockDim.x) {
        // the block's 32 loads in flight together, summed in list order
        const int cnt = min(32, n - b * 32);
        float vb[32];
#pragma unroll
        for (int i = 0; i < 32; i++) vb[i] = good[b * 32 + min(i, cnt - 1)];
        double sm = 0.0, sm2 = 0.0;
#pragma unroll
        for (int i = 0; i < 32; i++) {
          if (i < cnt) {
            const float v = vb[i];
            const double d = v > 0.0f ? (double)v : 0.0;
            sm += d;
            sm2 += d * d;
          }
        }
        bs[b] = sm;
        bs2[b] = sm2;
      }
      if (tid < 16) {
        bs[nb + tid] = 0.0;
        bs2[nb + tid] = 0.0;
      }
      __syncthreads();
      // the block sums added in block order (the reference's std::accumulate)
      if (tid == 0) s_sum = ordered_sum(bs, nb);
      else if (tid == 64) s_sum2 = ordered_sum(bs2, nb);
      __syncthreads();
      if (tid == 0) {
        s_cur = s_prev;
        double np = (double)s_prev;
        double m = s_sum / np;
        double sd = sqrt((s_sum2 - np * m * m) / (double)(s_prev - 1));
        s_mean = m;
        s_sd = sd;
        s_inv = 1.0 / sd;
      }
      __syncthreads();
      const double m = s_mean, inv = s_inv;
      int c = 0;
      for (int kb = tid; kb < n; kb += 8 * (int)blockDim.x) {
        float vb[8];
#pragma unroll
        for (int u = 0; u < 8; u++) vb[u] = good[min(kb + u * (int)blockDim.x, n - 1)];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int k = kb + u * (int)blockDim.x;
          const float v = vb[u];
          if (k < n && v > 0.0f && fabs((double)v - m) * inv > alpha) {
            good[k] = -v;
            c++;
          }
        }
      }
      int tot = block_sum_int(c, isc);
      if (tid == 0) s_prev = s_cur - tot;
      __syncthreads();
    }
  }
  const double mean = s_mean;
  const double invStdDev = 1.0 / s_sd;

  // ---- createHistogram (sss.cl:190-324).  The good list is exactly the
  // histogram's pixel set (mask == 0 and value > 1e-10 over the stamp).
  __shared__ double h_binSize, h_lowerBinVal, h_lower, h_upper, h_sky;
  __shared__ int h_state;  // 0 continue, 1 done (setFwhm), 2 done (no fwhm)
  __shared__ int hC[257], hU[256];  // bin prefix sums, window upper ends
  __shared__ double hdv[20];
  __shared__ long long hll[20];
  if (tid == 0) {
    double up = smp[(int)(0.9 * sampleCount)];
    double mid = smp[(int)(0.5 * sampleCount)];
    double binSize = (up - mid) / (double)NSAMP;
    h_binSize = binSize;
    h_lowerBinVal = mid - (128.0 * binSize);
    h_lower = h_upper = h_sky = 0.0;
    h_state = 0;
  }
  __syncthreads();
  for (int attempts = 0;; attempts++) {
    if (attempts >= 5) {
      if (tid == 0) h_state = 2;
      __syncthreads();
      break;
    }
    for (int i = tid; i < 256; i += blockDim.x) bins[i] = 0;
    __syncthreads();
    const double lbv = h_lowerBinVal, bsz = h_binSize;
    int ok = 0;
    for (int kb = tid; kb < n; kb += 8 * (int)blockDim.x) {
      float vb[8];
#pragma unroll
      for (int u = 0; u < 8; u++) vb[u] = good[min(kb + u * (int)blockDim.x, n - 1)];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        if (kb + u * (int)blockDim.x >= n) continue;
        const double v = (double)fabsf(vb[u]);
        if (fabs(v - mean) * invStdDev > alpha) continue;
        const double q = floor((v - lbv) / bsz) + 1.0;
        const int idx = (q != q) ? 0 : (q <= 0.0 ? 0 : (q >= 255.0 ? 255 : (int)q));
        atomicAdd(&bins[idx], 1);
        ok++;
      }
    }
    int okCount = block_sum_int(ok, isc);
    // The reference walks the bins serially, accumulating integer counts in
    // doubles: every partial sum is an exact integer, so each walk is a
    // difference of exact prefix sums C and each stopping point a first
    // crossing of C (binary search).  The sliding window's upper end is the
    // running maximum of the per-start first crossings.
    const double binSize = h_binSize, lowerBinVal = h_lowerBinVal;
    if (okCount == 0 || binSize == 0.0) {
      if (tid == 0) h_state = 2;
      __syncthreads();
      break;
    }
    {
      int tot;
      const int pre = block_prefix_sum(tid < 256 ? bins[tid] : 0, isc, &tot);
      if (tid < 256) hC[tid] = pre;
      if (tid == 0) hC[256] = tot;
      __syncthreads();
    }
    const double thr = okCount / 10.0;
    // window start L = tid (1..254): first V in [L+1, 255] with C[V]-C[L] >= thr
    int V = 0;
    if (tid >= 1 && tid <= 254) {
      const int cl = hC[tid];
      int lo = tid + 1, hi = 255;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((double)(hC[mid] - cl) >= thr) hi = mid;
        else lo = mid + 1;
      }
      V = lo;
    }
    const int Uincl = max(1, block_prefix_max_incl(V, isc));
    if (tid < 256) hU[tid] = Uincl;
    __syncthreads();
    double dens = -INFINITY;
    if (tid >= 1 && tid <= 254) {
      const int prevU = tid == 1 ? 1 : hU[tid - 1];
      if (prevU < 255) {  // the walk reaches start tid
        const double d = (double)(hC[Uincl] - hC[tid]) / (double)(Uincl - tid);
        if (d > 0.0) dens = d;
      }
    }
    int maxIndex = block_argmax_first(dens, tid, hdv, isc);
    {
      // every thread took part, so hdv holds the per-wave maxima: no start
      // had a positive density when all of them are -inf
      bool any = false;
      for (int i = 0; i < NT / WAVE; i++) any = any || hdv[i] > -INFINITY;
      if (!any) maxIndex = -1;
    }
    if (maxIndex < 0 || maxIndex > 255) maxIndex = 0;
    // mode window [maxIndex, iEnd)
    int iEnd;
    {
      const int cm = hC[maxIndex];
      int lo = maxIndex, hi = 255;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((double)(hC[mid] - cm) >= thr) hi = mid;
        else lo = mid + 1;
      }
      iEnd = lo;
    }
    const long long se = block_sum_ll(tid >= maxIndex && tid < iEnd ? (long long)tid * bins[tid] : 0ll,
                                      hll);
    if (tid == 0) {
      const double sumBins = (double)(hC[iEnd] - hC[maxIndex]);
      const double sumExpect = (double)se;
      double modeBin = sumExpect / sumBins + 0.5;
      double skyEst = lowerBinVal + binSize * (modeBin - 1.0);
      double lower = okCount * 0.25, upper = okCount * 0.75;
      auto firstAtLeast = [&](int from, double v) {
        int lo = from, hi = 256;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if ((double)hC[mid] >= v) hi = mid;
          else lo = mid + 1;
        }
        return lo;
      };
      int i = firstAtLeast(0, lower);
      lower = i - ((double)hC[i] - lower) / bins[i - 1];
      i = firstAtLeast(i, upper);
      upper = i - ((double)hC[i] - upper) / bins[i - 1];
      h_sky = skyEst;
      h_lower = lower;
      h_upper = upper;
      if (lower < 1.0 || upper > 255.0) {
        h_lowerBinVal = lowerBinVal - 128.0 * binSize;
        h_binSize = binSize * 2;
      } else if (upper - lower < 40.0) {
        h_binSize = binSize / 3.0;
        h_lowerBinVal = skyEst - 128.0 * h_binSize;
      } else {
        h_state = 1;
      }
    }
    __syncthreads();
    if (h_state != 0) break;
  }
  if (tid == 0) {
    fwhmOut[s] = h_state == 1 ? h_binSize * (h_upper - h_lower) / iqRange : 0.0;
    skyOut[s] = h_sky;
  }
}

// ---------------------------------------------------------------------------
// findSubStamps (sss.cl:326-505, sssUtil.cpp:40-104)
// ---------------------------------------------------------------------------
struct FindArgs {
  const float* img;
  uint16_t* mask;
  int w, h, nx, ny, fsw;
  const double* sky;
  const double* fwhm;
  int* found;
  int* count;
  int* ssx;
  int* ssy;
  double* ssv;
  double hi, tkf;
  uint16_t badMask, badPixelMask, skipMask;
  int* candAll;  // [N][fsw*fsw] candidate-list scratch
};

// Wave-wide reductions by DPP, result wave-uniform: rotations by 1, 2, 4, 8
// within each row of 16 lanes give every lane its row's result; the row
// broadcasts of lanes 15 and 31 carry them into the last row, and lane 63
// holds the wave's.  Lanes a broadcast does not write keep their own value
// (min, max and argmax are idempotent) or take 0 (sums).  Every reduction
// here is exact in any combining order.
template <int CTRL, int RM>
__device__ __forceinline__ int dpp_keep(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, RM, 0xF, false);
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_keep_f64(double v) {
  const int lo = dpp_keep<CTRL, RM>(__double2loint(v)), hi = dpp_keep<CTRL, RM>(__double2hiint(v));
  return __hiloint2double(hi, lo);
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_zero_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, RM, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, RM, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane63_f64(double v) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                          __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// (value, index) maximum: larger value wins, ties -> smaller index; NaN never
// wins (callers pass -inf for lanes without a candidate).
template <int CTRL, int RM>
__device__ __forceinline__ void argmax_first_step(double& v, int& idx) {
  const double ov = dpp_keep_f64<CTRL, RM>(v);
  const int oi = dpp_keep<CTRL, RM>(idx);
  if (ov > v || (ov == v && oi < idx)) {
    v = ov;
    idx = oi;
  }
}
__device__ __forceinline__ int wave_argmax_first(double v, int idx) {
  argmax_first_step<0x121, 0xF>(v, idx);
  argmax_first_step<0x122, 0xF>(v, idx);
  argmax_first_step<0x124, 0xF>(v, idx);
  argmax_first_step<0x128, 0xF>(v, idx);
  argmax_first_step<0x142, 0xA>(v, idx);
  argmax_first_step<0x143, 0xC>(v, idx);
  return __builtin_amdgcn_readlane(idx, 63);
}

// The walk's workgroup barrier.  Staged, every read of the walk is of LDS
// (the stamp's values and bad bits), and the global mask's atomic ORs are
// output only, so the barrier waits for LDS traffic alone; unstaged, the
// walk reads the global mask, and the full barrier with its fences is needed.
template <bool STAGED>
__device__ __forceinline__ void walk_barrier() {
  if (STAGED) {
    __asm__ __volatile__("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  } else {
    __syncthreads();
  }
}

// The winning value of wave_argmax_first (the lane holding index wi, or
// -inf when no lane has a candidate).
__device__ __forceinline__ double readlane63_max_f64(double v, int idx, int wi) {
  const unsigned long long own = __ballot(idx == wi && wi != INT_MAX);
  if (own == 0ull) return -INFINITY;
  const int src = __builtin_ctzll(own);
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), src),
                          __builtin_amdgcn_readlane(__double2loint(v), src));
}

// The check's reductions in one pass: first bad position (with its type),
// exponent range, partial sum.
template <int CTRL, int RM>
__device__ __forceinline__ void check_step(int& fb, int& bt, int& emn, int& emx, double& part) {
  const int ofb = dpp_keep<CTRL, RM>(fb), obt = dpp_keep<CTRL, RM>(bt);
  const int omn = dpp_keep<CTRL, RM>(emn), omx = dpp_keep<CTRL, RM>(emx);
  const double op = dpp_zero_f64<CTRL, RM>(part);
  if (ofb < fb) {
    fb = ofb;
    bt = obt;
  }
  emn = min(emn, omn);
  emx = max(emx, omx);
  part += op;
}
__device__ __forceinline__ void wave_check_reduce(int& fb, int& bt, int& emn, int& emx,
                                                  double& part) {
  check_step<0x121, 0xF>(fb, bt, emn, emx, part);
  check_step<0x122, 0xF>(fb, bt, emn, emx, part);
  check_step<0x124, 0xF>(fb, bt, emn, emx, part);
  check_step<0x128, 0xF>(fb, bt, emn, emx, part);
  check_step<0x142, 0xA>(fb, bt, emn, emx, part);
  check_step<0x143, 0xC>(fb, bt, emn, emx, part);
  fb = __builtin_amdgcn_readlane(fb, 63);
  bt = __builtin_amdgcn_readlane(bt, 63);
  emn = __builtin_amdgcn_readlane(emn, 63);
  emx = __builtin_amdgcn_readlane(emx, 63);
  part = readlane63_f64(part);
}

__device__ __forceinline__ int float_exp_field(float f) {
  int e = (int)((__float_as_uint(f) >> 23) & 0xff);
  return e == 0 ? 1 : e;
}

constexpr int FIND_THREADS = 1024;
// dynamic LDS of the staged search (gfx950: 160 KiB per workgroup), leaving
// room for the kernel's static arrays
constexpr size_t FIND_LDS_MAX = 160 * 1024 - 2048;
constexpr int FIND_WAVES = FIND_THREADS / WAVE;
// Windows (at most 31x31) are walked in a 32-wide layout, slot pp -> column
// pp & 31, row pp >> 5, so no lane divides by the window width; the raster
// index row * width + column keeps the reference's scan order for ties.
constexpr int WIN_PER_LANE = 32 * 32 / WAVE;
// The candidate walk's window work: four waves (one per SIMD), four slots each
constexpr int WALK_WAVES = 4;
constexpr int SLOTS_W = WIN_PER_LANE / WALK_WAVES;

// Pixel access of one stamp's search.  With STAGED, the stamp's fsw x fsw
// region lives in LDS: the pixel values (float) and one bit per pixel telling
// whether the mask is bad for this search ((mask & badMask) != 0).  Every
// access of the search stays inside that region, and only this search sets
// bits of its own badMask (the other image's search sets the other image's
// bits), so the bitmap is an exact view of the mask; mask bits are still ORed
// into the global mask, which nothing reads back here.  Without STAGED the
// global image and mask are read directly and each mask write is fenced
// before later reads.
template <bool STAGED>
struct StampPix {
  const float* img;
  uint16_t* mask;
  long imgW;
  int ox, oy, fsw;
  uint16_t badMask;
  float* simg;     // STAGED: [fsw*fsw]
  uint32_t* sbad;  // STAGED: [(fsw*fsw + 63) / 64 * 2]
  __device__ __forceinline__ long addr(int x, int y) const { return x + (long)y * imgW; }
  __device__ __forceinline__ int local(int x, int y) const { return (y - oy) * fsw + (x - ox); }
  __device__ __forceinline__ float value(int x, int y) const {
    return STAGED ? simg[local(x, y)] : img[addr(x, y)];
  }
  __device__ __forceinline__ bool bad(int x, int y) const {
    if (STAGED) {
      const int k = local(x, y);
      return (sbad[k >> 5] >> (k & 31)) & 1u;
    }
    return (mask[addr(x, y)] & badMask) != 0;
  }
  // bits always belong to badMask (bad pixel or skip bit of this image)
  __device__ __forceinline__ void mark(int x, int y, uint16_t bits) const {
    mask_or_atomic(mask, (size_t)addr(x, y), bits);
    if (STAGED) {
      const int k = local(x, y);
      atomicOr(&sbad[k >> 5], 1u << (k & 31));
    }
  }
};

// One 1024-thread workgroup per stamp.  Per threshold level, all 16 waves
// classify the stamp's pixels against the mask as it is now: wave w scans the
// raster segment [w*seg, (w+1)*seg) and appends its candidates, in raster
// order, to its own slice of the scratch list (ballot + popcount).  Mask bits
// are only ever added, so a pixel rejected here stays rejected; flagging
// pixels above threshHigh commutes with everything else.  The workgroup then
// processes the candidates one by one in raster order, re-checking each
// against the current mask: every wave makes the same decisions, four of
// them evaluate the 32 x 32 windows, and each step meets at a barrier (with
// the stamp staged in LDS, one that waits for LDS traffic only).  Mask bits are
// set with 32-bit atomic ORs: the template and science searches run at the
// same time, each reading only its own bits (badMask) but sharing mask words.
template <bool STAGED, int NI>
__global__ __launch_bounds__(FIND_THREADS) void k_sss_find(ArgsN<FindArgs, NI> PA) {
  const FindArgs& A = args_of(PA);
  extern __shared__ uint32_t find_lds[];
  const int s = blockIdx.x;
  const int tid = threadIdx.x, lane = lane_id(), wid = tid / WAVE;
  const StampGeom g = stamp_geom(s, A.nx, A.ny, A.w, A.h, A.fsw);
  const double sky = A.sky[s], fwhm = A.fwhm[s];
  const double floorV = sky + A.tkf * fwhm;
  const int fsw = A.fsw;
  const int nAll = fsw * fsw;
  const unsigned mf = fdiv_magic(fsw);  // stamp-relative raster index -> row
  StampPix<STAGED> P;
  P.img = A.img;
  P.mask = A.mask;
  P.imgW = A.w;
  P.ox = g.ox;
  P.oy = g.oy;
  P.fsw = fsw;
  P.badMask = A.badMask;
  P.simg = reinterpret_cast<float*>(find_lds);
  P.sbad = find_lds + (nAll + 63) / 64 * 64;
  __shared__ int pre[FIND_WAVES + 1];
  __shared__ int lx[MAXSS], ly[MAXSS];
  __shared__ double lv[MAXSS];
  __shared__ int s_counter;
  __shared__ double wWinV[2][WALK_WAVES], wPart[2][WALK_WAVES], s_total;
  __shared__ int wWinI[2][WALK_WAVES], wFb[2][WALK_WAVES], wBt[2][WALK_WAVES];
  __shared__ int wEmn[2][WALK_WAVES], wEmx[2][WALK_WAVES];
  static_assert(WALK_WAVES * SLOTS_W == WIN_PER_LANE, "the walkers cover the window");
  if (tid == 0) s_counter = 0;
  if (STAGED) {
    // stage the stamp: values, and the bad bits 64 pixels per ballot
    for (int base = wid * WAVE; base < nAll; base += FIND_THREADS) {
      const int k = base + lane;
      bool b = false;
      if (k < nAll) {
        const int ky = fdiv(k, mf), kx = k - ky * fsw;
        const long a = (kx + g.ox) + (long)(ky + g.oy) * A.w;
        P.simg[k] = A.img[a];
        b = (A.mask[a] & A.badMask) != 0;
      }
      const unsigned long long bal = __ballot(b);
      if (lane == 0) {
        P.sbad[base >> 5] = (uint32_t)bal;
        P.sbad[(base >> 5) + 1] = (uint32_t)(bal >> 32);
      }
    }
  }
  __syncthreads();
  double dfrac = 0.9;
  const int seg = ((nAll + FIND_WAVES - 1) / FIND_WAVES + WAVE - 1) / WAVE * WAVE;
  int* cand = A.candAll + (size_t)s * nAll;
  // The reference loops until the threshold reaches the floor; with a
  // non-finite sky or sigma that never happens, so the levels are capped
  // (finite statistics end within 6 levels).
  for (int level = 0; level < MAX_LEVELS && s_counter < MAXSS; level++) {
    const double lim = std_max(floorV, sky + (A.hi - sky) * dfrac);
    {
      const int k0 = wid * seg, k1 = min(nAll, k0 + seg);
      int cnt = 0;
      for (int base = k0; base < k1; base += 4 * WAVE) {
        bool bd[4];
        float v[4];
        int xs[4], ys[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int k = base + u * WAVE + lane;
          const int kc = min(k, k1 - 1);  // in bounds; past the segment counts as bad
          const int ky = fdiv(kc, mf);
          xs[u] = kc - ky * fsw + g.ox;
          ys[u] = ky + g.oy;
          const bool b = P.bad(xs[u], ys[u]);
          v[u] = P.value(xs[u], ys[u]);
          bd[u] = k >= k1 || b;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int k = base + u * WAVE + lane;
          bool c = false;
          if (k < k1 && !bd[u]) {
            const double dv = (double)v[u];
            if (dv > A.hi) {
              P.mark(xs[u], ys[u], A.badPixelMask);
            } else if (!((dv - sky) * (1.0 / fwhm) < A.tkf) && dv > lim) {
              c = true;
            }
          }
          const unsigned long long bal = __ballot(c);
          if (c) cand[k0 + cnt + __popcll(bal & ((1ull << lane) - 1ull))] = k;
          cnt += __popcll(bal);
        }
      }
      if (lane == 0) pre[wid + 1] = cnt;
      if (tid == 0) pre[0] = 0;
    }
    __syncthreads();
    if (tid == 0)
      for (int w = 0; w < FIND_WAVES; w++) pre[w + 1] += pre[w];
    __syncthreads();
    {
      // The walk, by all 16 waves in step: each runs the same candidate
      // selection (same data, same decisions); the first WALK_WAVES (one per
      // SIMD) evaluate each 32 x 32 window, eight rows each (slots u = 4w ..
      // 4w + 3 of the one-wave layout), and the others only meet them at the
      // barriers.  Partial results meet in LDS (double-buffered by candidate
      // parity) and are combined in the order-independent way the one-wave
      // reductions use (maximum with the first index among ties, minimum
      // first-bad position, exponent range, sums that are exact when the
      // exponents are close).
      const int nc = pre[FIND_WAVES];
      int counter = s_counter;
      int par = 0;
      // logical list position -> raster index
      auto candAt = [&](int q) {
        int w = 0;
        while (w + 1 < FIND_WAVES && pre[w + 1] <= q) w++;
        return cand[w * seg + (q - pre[w])];
      };
      const bool walker = wid < WALK_WAVES;
      const int px = lane & 31, py0 = (lane >> 5) + 2 * SLOTS_W * wid;  // this wave's first slot
      // candidates in batches of 64 (one per lane, raster order); within a
      // batch the next candidate is the first pending one whose pixel is
      // unmasked now -- the same one a sequential walk reaches next
      for (int b0 = 0; b0 < nc && counter < MAXSS; b0 += WAVE) {
        const int q = b0 + lane;
        const int kq = q < nc ? candAt(q) : 0;
        const int qy = fdiv(kq, mf) + g.oy, qx = kq - (qy - g.oy) * fsw + g.ox;
        unsigned long long pending = __ballot(q < nc);
        while (pending != 0ull && counter < MAXSS) {
          const bool okq = ((pending >> lane) & 1ull) && !P.bad(qx, qy);
          const unsigned long long bal = __ballot(okq);
          if (bal == 0ull) break;
          const int first = __ffsll((unsigned long long)bal) - 1;
          pending &= ~((2ull << first) - 1ull);  // drop this lane and those before it
          const int absx = __builtin_amdgcn_readlane(qx, first);
          const int absy = __builtin_amdgcn_readlane(qy, first);
          // ---- window maximum (sss.cl:427-459), this wave's slot
          const int startX = max(absx - HSS, g.ox), startY = max(absy - HSS, g.oy);
          const int endX = min(absx + HSS, g.ox + fsw - 1), endY = min(absy + HSS, g.oy + fsw - 1);
          const int ww = endX - startX + 1, wh = endY - startY + 1;
          if (walker) {
            double bestV = 0.0;
            int bestI = INT_MAX;
            bool kb[SLOTS_W];
            float kvf[SLOTS_W];
#pragma unroll
            for (int u = 0; u < SLOTS_W; u++) {
              const int py = py0 + 2 * u;
              const bool in = px < ww && py < wh;
              const int x = startX + min(px, ww - 1), y = startY + min(py, wh - 1);
              const bool b = P.bad(x, y);
              kvf[u] = P.value(x, y);
              kb[u] = !in || b;
            }
#pragma unroll
            for (int u = 0; u < SLOTS_W; u++) {
              const int py = py0 + 2 * u;
              if (kb[u]) continue;  // also every slot outside the window
              const double kv = (double)kvf[u];
              if (kv >= A.hi) {
                P.mark(startX + px, startY + py, A.badPixelMask);
                continue;
              }
              if ((kv - sky) * (1.0 / fwhm) < A.tkf) continue;
              if (kv > bestV) {
                bestV = kv;
                bestI = py * ww + px;
              }
            }
            if (bestI == INT_MAX) bestV = -INFINITY;
            const int wi = wave_argmax_first(bestV, bestI);
            const double wv = readlane63_max_f64(bestV, bestI, wi);
            if (lane == 0) {
              wWinV[par][wid] = wv;
              wWinI[par][wid] = wi;
            }
          }
          walk_barrier<STAGED>();  // partials and the window's marks
          int win = INT_MAX;
          {
            double bv = -INFINITY;
#pragma unroll
            for (int w = 0; w < WALK_WAVES; w++) {
              const double v = wWinV[par][w];
              const int i = wWinI[par][w];
              if (i != INT_MAX && (v > bv || (v == bv && i < win))) {
                bv = v;
                win = i;
              }
            }
          }
          int mx = absx, my = absy;
          if (win != INT_MAX) {
            mx = startX + win % ww;
            my = startY + win / ww;
          }
          // ---- checkSStamp (sss.cl:326-361), this wave's slot
          const int cx0 = max(mx - HSS, g.ox), cy0 = max(my - HSS, g.oy);
          const int cx1 = min(mx + HSS, g.ox + g.sw - 1), cy1 = min(my + HSS, g.oy + g.sh - 1);
          const int cw = cx1 - cx0 + 1, ch = cy1 - cy0 + 1;
          if (walker) {
            int firstBad = INT_MAX;
            int badType = 0;
            double part = 0.0;
            int eMin = 1000, eMax = -1000;
            bool cb[SLOTS_W], cin[SLOTS_W];
            float cvf[SLOTS_W];
#pragma unroll
            for (int u = 0; u < SLOTS_W; u++) {
              const int py = py0 + 2 * u;
              cin[u] = px < cw && py < ch;
              const int x = cx0 + min(px, cw - 1), y = cy0 + min(py, ch - 1);
              cb[u] = P.bad(x, y);
              cvf[u] = P.value(x, y);
            }
#pragma unroll
            for (int u = 0; u < SLOTS_W; u++) {
              const int py = py0 + 2 * u;
              if (!cin[u]) continue;
              const int pc = py * cw + px;
              if (cb[u]) {
                if (pc < firstBad) { firstBad = pc; badType = 1; }
                continue;
              }
              const float fv = cvf[u];
              const double v = (double)fv;
              if (v > A.hi) {
                if (pc < firstBad) { firstBad = pc; badType = 2; }
                continue;
              }
              if ((v - sky) / fwhm > A.tkf) {
                part += v;
                if (fv != 0.0f) {
                  const int e = float_exp_field(fv);
                  eMin = min(eMin, e);
                  eMax = max(eMax, e);
                }
              }
            }
            wave_check_reduce(firstBad, badType, eMin, eMax, part);
            if (lane == 0) {
              wFb[par][wid] = firstBad;
              wBt[par][wid] = badType;
              wEmn[par][wid] = eMin;
              wEmx[par][wid] = eMax;
              wPart[par][wid] = part;
            }
          }
          walk_barrier<STAGED>();  // the check's partials
          int fb = INT_MAX, bt = 0, gMin = 1000, gMax = -1000;
          double total = 0.0;
#pragma unroll
          for (int w = 0; w < WALK_WAVES; w++) {
            const int f = wFb[par][w];
            if (f < fb) {
              fb = f;
              bt = wBt[par][w];
            }
            gMin = min(gMin, wEmn[par][w]);
            gMax = max(gMax, wEmx[par][w]);
            total += wPart[par][w];
          }
          par ^= 1;
          if (fb != INT_MAX) {
            // first bad pixel in raster order: checkSStamp returned 0
            if (bt == 2) {
              if (tid == 0) P.mark(cx0 + fb % cw, cy0 + fb / cw, A.badPixelMask);
              walk_barrier<STAGED>();  // the mark before the next candidate's reads
            }
            continue;
          }
          // sum of qualifying values in raster order: exact in any order when
          // the exponent spread is small (all partial sums representable),
          // otherwise summed sequentially by one lane
          if (!(gMax < gMin || gMax - gMin <= 19)) {
            if (tid == 0) {
              double r = 0.0;
              for (int p = 0; p < cw * ch; p++) {
                const double v = (double)P.value(cx0 + p % cw, cy0 + p / cw);
                if ((v - sky) / fwhm > A.tkf) r += v;
              }
              s_total = r;
            }
            walk_barrier<STAGED>();
            total = s_total;
          }
          if (total == 0.0) continue;
          if (tid == 0) {
            lx[counter] = mx;
            ly[counter] = my;
            lv[counter] = total;
          }
          counter++;
          // skip marks (sss.cl:473-482), clamped with the stamp size: this
          // wave's eight rows of the box
          if (walker) {
#pragma unroll
            for (int u = 0; u < SLOTS_W; u++) {
              const int py = py0 + 2 * u;
              if (px < cw && py < ch) P.mark(cx0 + px, cy0 + py, A.skipMask);
            }
          }
          walk_barrier<STAGED>();  // the marks before the next candidate's reads
        }
      }
      if (tid == 0) s_counter = counter;
    }
    __syncthreads();
    if (lim == floorV) break;
    dfrac -= 0.2;
  }
  if (tid == 0) {
    int counter = s_counter;
    for (int i = 1; i < counter; i++)
      for (int j = i; j > 0 && lv[j - 1] < lv[j]; j--) {
        int tx = lx[j - 1]; lx[j - 1] = lx[j]; lx[j] = tx;
        int ty = ly[j - 1]; ly[j - 1] = ly[j]; ly[j] = ty;
        double tv = lv[j - 1]; lv[j - 1] = lv[j]; lv[j] = tv;
      }
    A.found[s] = counter;
    A.count[s] = min(counter, KEEPSS);
    for (int i = 0; i < MAXSS; i++) {
      size_t o = (size_t)s * MAXSS + i;
      A.ssx[o] = i < counter ? lx[i] : INT_MAX;
      A.ssy[o] = i < counter ? ly[i] : INT_MAX;
      A.ssv[o] = i < counter ? lv[i] : -INFINITY;
    }
  }
}

// removeEmptyStamps (sss.cl:507-571): ordered compaction of non-empty stamps,
// both images; retry flag when either keeps less than 10 % (bach.cpp:132-142).
__global__ __launch_bounds__(1024) void k_sss_compact(const int* __restrict__ count0,
                                                      const int* __restrict__ count1,
                                                      int* kept0, int* kept1, int* cur0,
                                                      int* cur1, DevScalars* sc, int nGrid,
                                                      DevScalars* hsc) {
  __shared__ int isc[20];
  __shared__ int s_n;
  for (int im = 0; im < 2; im++) {
    const int* cnt = im == 0 ? count0 : count1;
    int* kept = im == 0 ? kept0 : kept1;
    int* cur = im == 0 ? cur0 : cur1;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    for (int base = 0; base < nGrid; base += blockDim.x) {
      int s = base + threadIdx.x;
      bool keep = s < nGrid && cnt[s] > 0;
      int tot;
      int pos = block_prefix_count(keep, isc, &tot);
      int c0 = s_n;
      if (keep) {
        kept[c0 + pos] = s;
        cur[c0 + pos] = 0;
      }
      __syncthreads();
      if (threadIdx.x == 0) s_n = c0 + tot;
      __syncthreads();
    }
    if (threadIdx.x == 0) sc->nKept[im] = s_n;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double f0 = (double)sc->nKept[0] / nGrid, f1 = (double)sc->nKept[1] / nGrid;
    sc->retry = (f0 < 0.1 || f1 < 0.1) ? 1 : 0;
    if (hsc) {  // the host's pinned mirror: the counts and the retry flag
      hsc->nKept[0] = sc->nKept[0];
      hsc->nKept[1] = sc->nKept[1];
      hsc->retry = sc->retry;
      __threadfence_system();
    }
  }
}

__global__ void k_reset_skip(uint16_t* mask, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mask[i] &= (uint16_t) ~(XB_MASK_SKIP_S | XB_MASK_SKIP_T);
}

hipError_t launch_sss_stats(Ctx& c, int im) {
  if ((c.fsw * c.fsw + 31) / 32 > MAX_CLIP_BLOCKS) return hipErrorInvalidValue;
  prof_begin(c, XB_PROF_SSS);
  // threads per stamp from the context's launch shapes (256, 512 or 1024)
  if (c.shapes.stats_threads == 256)
    hipLaunchKernelGGL(k_sss_stats<256>, dim3(c.nGrid), dim3(256), 0, c.cur, c.img[im], c.mask,
                       c.w, c.h, c.stampsx, c.stampsy, c.fsw, c.ran1, c.goodPix,
                       c.m.sigClipAlpha, c.m.iqRange, c.st[im].sky, c.st[im].fwhm);
  else if (c.shapes.stats_threads == 512)
    hipLaunchKernelGGL(k_sss_stats<512>, dim3(c.nGrid), dim3(512), 0, c.cur, c.img[im], c.mask,
                       c.w, c.h, c.stampsx, c.stampsy, c.fsw, c.ran1, c.goodPix,
                       c.m.sigClipAlpha, c.m.iqRange, c.st[im].sky, c.st[im].fwhm);
  else
    hipLaunchKernelGGL(k_sss_stats<1024>, dim3(c.nGrid), dim3(1024), 0, c.cur, c.img[im], c.mask,
                       c.w, c.h, c.stampsx, c.stampsy, c.fsw, c.ran1, c.goodPix,
                       c.m.sigClipAlpha, c.m.iqRange, c.st[im].sky, c.st[im].fwhm);
  prof_end(c, XB_PROF_SSS);
  return hipGetLastError();
}

static FindArgs find_args(Ctx& c, int im) {
  FindArgs A;
  A.img = c.img[im];
  A.mask = c.mask;
  A.w = c.w; A.h = c.h; A.nx = c.stampsx; A.ny = c.stampsy; A.fsw = c.fsw;
  A.sky = c.st[im].sky; A.fwhm = c.st[im].fwhm;
  A.found = c.st[im].found; A.count = c.st[im].count;
  A.ssx = c.st[im].ssx; A.ssy = c.st[im].ssy; A.ssv = c.st[im].ssv;
  A.hi = c.m.threshHigh; A.tkf = c.m.threshKernFit;
  uint16_t badMask = XB_MASK_ALL & ~XB_MASK_OK_CONV;
  if (im == 0) {
    badMask &= ~(XB_MASK_BAD_PIXEL_S | XB_MASK_SKIP_S);
    A.badPixelMask = XB_MASK_BAD_PIXEL_T;
    A.skipMask = XB_MASK_SKIP_T;
  } else {
    badMask &= ~(XB_MASK_BAD_PIXEL_T | XB_MASK_SKIP_T);
    A.badPixelMask = XB_MASK_BAD_PIXEL_S;
    A.skipMask = XB_MASK_SKIP_S;
  }
  A.badMask = badMask;
  // the good-pixel scratch of k_sss_stats is free again; each image's search
  // uses its own half (the two searches run concurrently)
  A.candAll = reinterpret_cast<int*>(c.goodPix) + (size_t)im * c.capStamps * c.capFsw * c.capFsw;
  return A;
}

// The searches of NI images (image z = blockIdx.z): each reads only its own
// image's mask bits and sets them with atomic ORs, so they may run at once.
template <int NI>
static hipError_t launch_find_n(Ctx& c, const ArgsN<FindArgs, NI>& P) {
  // the stamp region in LDS when it fits (float values + one bad bit per pixel)
  const size_t nAll = (size_t)c.fsw * c.fsw;
  const size_t lds = ((nAll + 63) / 64 * 64) * sizeof(float) + (nAll + 63) / 64 * 8;
  prof_begin(c, XB_PROF_SSS);
  if (c.shapes.find_staged && lds <= FIND_LDS_MAX) {
    // once per process (the attribute belongs to the kernel); the static's
    // initialisation is thread-safe, so batch workers may race here
    static const hipError_t attrRc = hipFuncSetAttribute(
        (const void*)k_sss_find<true, NI>, hipFuncAttributeMaxDynamicSharedMemorySize,
        (int)FIND_LDS_MAX);
    if (attrRc != hipSuccess) return attrRc;
    hipLaunchKernelGGL((k_sss_find<true, NI>), dim3(c.nGrid, 1, NI), dim3(FIND_THREADS), lds,
                       c.cur, P);
  } else {
    hipLaunchKernelGGL((k_sss_find<false, NI>), dim3(c.nGrid, 1, NI), dim3(FIND_THREADS), 0,
                       c.cur, P);
  }
  prof_end(c, XB_PROF_SSS);
  return hipGetLastError();
}

hipError_t launch_sss_find(Ctx& c, int im) {
  return launch_find_n<1>(c, ArgsN<FindArgs, 1>{{find_args(c, im)}});
}

hipError_t launch_sss_find_pair(Ctx& c) {
  return launch_find_n<2>(c, ArgsN<FindArgs, 2>{{find_args(c, 0), find_args(c, 1)}});
}

hipError_t launch_sss_compact(Ctx& c) {
  hipLaunchKernelGGL(k_sss_compact, dim3(1), dim3(1024), 0, c.cur, c.st[0].count,
                     c.st[1].count, c.st[0].kept, c.st[1].kept, c.st[0].cur, c.st[1].cur, c.sc,
                     c.nGrid, c.hscDev);
  return hipGetLastError();
}

hipError_t launch_reset_skip(Ctx& c) {
  size_t n = (size_t)c.w * c.h;
  hipLaunchKernelGGL(k_reset_skip, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.cur,
                     c.mask, n);
  return hipGetLastError();
}

}  // namespace xb
