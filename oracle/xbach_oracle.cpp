/*
 * xbach_oracle.cpp — CPU restatement of X-BACH's init -> sub path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU library, never part of it.
 * See xbach_oracle.h for the contract (canonical order where the reference uses
 * atomics, operand order and types kept, no FMA contraction).
 *
 * Reference files followed (paths relative to the reference repository):
 *   src/bach.cpp, src/bachUtil.cpp, src/sssUtil.cpp, src/cmvUtil.cpp,
 *   src/cdkscUtil.cpp, include/datatypeUtil.h, include/argsUtil.h,
 *   cl_kern/{ini,sss,bach,cmv,cd,ksc,conv,sub}.cl
 */
#include "xbach_oracle.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef uint16_t u16;

/* Optional capture of every global normal matrix + RHS built (pin tests). */
bool g_capture = false;
std::vector<std::vector<double>> g_capMats, g_capRhs;
void capture(const std::vector<double>& M, const std::vector<double>& rhs) {
  if (!g_capture) return;
  g_capMats.push_back(M);
  g_capRhs.push_back(rhs);
}

const double kInf = std::numeric_limits<double>::infinity();

/* ------------------------------------------------------------------------ */
/* Derived sizes (src/cmvUtil.cpp:9-12, src/cdkscUtil.cpp:8-12)              */
/* ------------------------------------------------------------------------ */
struct Cfg {
  xb_params a;
  int mode = XO_MODE_REF;
  int nPSF = 0;
  int nComp1 = 0, nComp2 = 0, nComp = 0, nBG = 0;
  int matSize = 0, nSolComp = 0;
  int wRows = 0, wCols = 0, qCount = 0, bCount = 0;
  int maxSS = 0;
  int kw = 0, hkw = 0, ssw = 0, hssw = 0;
  std::vector<int> kx, ky, kg;        /* basis (x, y, gauss), src/bach.cpp:181-188 */
  std::vector<int> bgx, bgy;          /* background monomials, src/bach.cpp:212-216 */
  std::vector<int> cdx, cdy;          /* spatial monomials, src/bach.cpp:278-282 */
  std::vector<double> fx, fy, vec;    /* basis tables */
};

int tri(int t) { return t * (t + 1) / 2; } /* include/mathUtil.h:3 */

/* Basis filters.  cl_kern/cmv.cl:1-43 (createKernelFilter) and the CPU copy in
 * include/datatypeUtil.h:62-118 compute the same values; bg is float. */
void buildBasis(Cfg& c) {
  const xb_params& a = c.a;
  c.kx.clear(); c.ky.clear(); c.kg.clear();
  for (int g = 0; g < a.ngauss; g++)
    for (int x = 0; x <= a.dg[g]; x++)
      for (int y = 0; y <= a.dg[g] - x; y++) {
        c.kg.push_back(g); c.kx.push_back(x); c.ky.push_back(y);
      }
  c.nPSF = (int)c.kx.size();
  const int kw = a.fKernelWidth;
  c.fx.assign((size_t)c.nPSF * kw, 0.0);
  c.fy.assign((size_t)c.nPSF * kw, 0.0);
  c.vec.assign((size_t)c.nPSF * kw * kw, 0.0);
  for (int n = 0; n < c.nPSF; n++) {
    const double b = (double)a.bg[c.kg[n]];
    double sumX = 0.0, sumY = 0.0;
    for (int i = 0; i < kw; i++) {
      double x = (double)(i - kw / 2);
      double qe = std::exp(-x * x * b);
      double px = 1.0, py = 1.0; /* x^k, exact for these small integers */
      for (int e = 0; e < c.kx[n]; e++) px *= x;
      for (int e = 0; e < c.ky[n]; e++) py *= x;
      double vx = qe * px, vy = qe * py;
      c.fx[(size_t)n * kw + i] = vx;
      c.fy[(size_t)n * kw + i] = vy;
      sumX += vx;
      sumY += vy;
    }
    const bool even = (c.kx[n] % 2 == 0) && (c.ky[n] % 2 == 0);
    if (even) {
      double invX = 1.0 / sumX, invY = 1.0 / sumY;
      for (int i = 0; i < kw; i++) {
        c.fx[(size_t)n * kw + i] *= invX;
        c.fy[(size_t)n * kw + i] *= invY;
      }
    }
  }
  /* cl_kern/cmv.cl:45-68 (createKernelVector): vec[n][v][u] = fx[u]*fy[v],
   * minus basis 0 for even-even n > 0. */
  for (int n = 0; n < c.nPSF; n++) {
    const bool even = (c.kx[n] % 2 == 0) && (c.ky[n] % 2 == 0);
    for (int v = 0; v < kw; v++)
      for (int u = 0; u < kw; u++) {
        double vv = c.fx[(size_t)n * kw + u] * c.fy[(size_t)n * kw + v];
        if (even && n > 0) vv -= c.fx[u] * c.fy[v];
        c.vec[(size_t)n * kw * kw + (size_t)v * kw + u] = vv;
      }
  }
}

bool setupCfg(Cfg& c, const xb_params* p, int mode) {
  c.a = *p;
  c.mode = mode;
  const xb_params& a = c.a;
  if (a.ngauss < 1 || a.ngauss > 3) return false;
  buildBasis(c);
  c.nComp1 = c.nPSF - 1;
  c.nComp2 = tri(a.kernelOrder + 1);
  c.nComp = c.nComp1 * c.nComp2;
  c.nBG = tri(a.backgroundOrder + 1);
  c.matSize = c.nComp + c.nBG + 1;
  c.nSolComp = c.nPSF * c.nComp2 + c.nBG + 1;
  c.wCols = a.fSStampWidth * a.fSStampWidth;
  c.wRows = c.nPSF + c.nBG;
  c.qCount = c.nPSF + 2;
  c.bCount = c.nPSF + 2;
  c.maxSS = 2 * a.maxKSStamps;
  c.kw = a.fKernelWidth;
  c.hkw = a.hKernelWidth;
  c.ssw = a.fSStampWidth;
  c.hssw = a.hSStampWidth;
  c.bgx.clear(); c.bgy.clear();
  for (int x = 0; x <= a.backgroundOrder; x++)
    for (int y = 0; y <= a.backgroundOrder - x; y++) { c.bgx.push_back(x); c.bgy.push_back(y); }
  c.cdx.clear(); c.cdy.clear();
  for (int i = 0; i <= a.kernelOrder; i++)
    for (int j = 0; j <= a.kernelOrder - i; j++) { c.cdx.push_back(i); c.cdy.push_back(j); }
  if (c.maxSS > XB_MAX_SS) return false;
  if (c.nSolComp > XB_MAX_SOL) return false;
  return true;
}

/* Integer power as OpenCL pown for the small exponents used here. */
double pown(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

/* ------------------------------------------------------------------------ */
/* Image state                                                               */
/* ------------------------------------------------------------------------ */
struct Pair {
  int w = 0, h = 0;
  std::vector<double> img[2]; /* 0 = template, 1 = science (before the swap) */
  std::vector<u16> mask;
};

/* cl_kern/ini.cl:15-35 (maskInput) and 37-66 (spreadMask),
 * src/bachUtil.cpp:8-32. */
void maskInput(const Cfg& c, Pair& P) {
  const int w = P.w, h = P.h;
  const int border = c.a.hSStampWidth + c.a.hKernelWidth;
  const double hi = c.a.threshHigh, lo = c.a.threshLow;
  const double* T = P.img[0].data();
  const double* S = P.img[1].data();
  P.mask.assign((size_t)w * h, 0);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      size_t id = (size_t)y * w + x;
      u16 m = 0;
      double t = T[id], s = S[id];
      if (t == 0.0 || s == 0.0) m |= XB_MASK_BAD_INPUT | XB_MASK_BAD_PIX_VAL;
      if (t >= hi || s >= hi) m |= XB_MASK_BAD_INPUT | XB_MASK_SAT_PIXEL;
      if (t <= lo || s <= lo) m |= XB_MASK_BAD_INPUT | XB_MASK_LOW_PIXEL;
      if (x < border || x >= w - border || y < border || y >= h - border)
        m |= XB_MASK_BAD_PIXEL_S | XB_MASK_BAD_PIXEL_T;
      P.mask[id] = m;
    }
  /* spreadMask reads only BAD_INPUT bits, which it never writes, so a
   * snapshot of the BAD_INPUT plane is equivalent. */
  const int spreadWidth = (int)(c.a.hKernelWidth * c.a.inSpreadMaskFactor);
  const int w2 = spreadWidth / 2;
  std::vector<uint8_t> bad((size_t)w * h);
  for (size_t i = 0; i < bad.size(); i++) bad[i] = (P.mask[i] & XB_MASK_BAD_INPUT) != 0;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      size_t id = (size_t)y * w + x;
      if (bad[id]) continue;
      int sx = std::max(x - w2, 0), sy = std::max(y - w2, 0);
      int ex = std::min(x + w2, w - 1), ey = std::min(y + w2, h - 1);
      bool ok = false;
      for (int y2 = sy; y2 <= ey && !ok; y2++)
        for (int x2 = sx; x2 <= ex; x2++)
          if (bad[(size_t)y2 * w + x2]) { ok = true; break; }
      if (ok) P.mask[id] |= XB_MASK_OK_CONV;
    }
}

/* ------------------------------------------------------------------------ */
/* NR ran1 (cl_kern/sss.cl:41-79): Numerical Recipes in C (1st ed.) ran1.    */
/* ------------------------------------------------------------------------ */
struct Ran1 {
  long ix1 = 0, ix2 = 0, ix3 = 0;
  double r[98];
  int iff = 0;
  int idum = -666;
  double next() {
    const long M1 = 259200, IA1 = 7141, IC1 = 54773;
    const long M2 = 134456, IA2 = 8121, IC2 = 28411;
    const long M3 = 243000, IA3 = 4561, IC3 = 51349;
    const double RM1 = 1.0 / M1, RM2 = 1.0 / M2;
    if (idum < 0 || iff == 0) {
      iff = 1;
      ix1 = (IC1 - idum) % M1;
      ix1 = (IA1 * ix1 + IC1) % M1;
      ix2 = ix1 % M2;
      ix1 = (IA1 * ix1 + IC1) % M1;
      ix3 = ix1 % M3;
      for (int j = 1; j <= 97; j++) {
        ix1 = (IA1 * ix1 + IC1) % M1;
        ix2 = (IA2 * ix2 + IC2) % M2;
        r[j] = (ix1 + ix2 * RM2) * RM1;
      }
      idum = 1;
    }
    ix1 = (IA1 * ix1 + IC1) % M1;
    ix2 = (IA2 * ix2 + IC2) % M2;
    ix3 = (IA3 * ix3 + IC3) % M3;
    int j = (int)(1 + ((97 * ix3) / M3));
    double temp = r[j];
    r[j] = (ix1 + ix2 * RM2) * RM1;
    return temp;
  }
};

/* Bitonic network of cl_kern/sss.cl:135-150 / src/bachUtil.cpp:229-240. */
void bitonic(double* a, int n) {
  for (int k = 2; k <= n; k *= 2)
    for (int j = k >> 1; j > 0; j >>= 1)
      for (int i = 0; i < n; i++) {
        int ixj = i ^ j;
        if (ixj > i) {
          if ((i & k) == 0 && a[i] > a[ixj]) std::swap(a[i], a[ixj]);
          if ((i & k) != 0 && a[i] < a[ixj]) std::swap(a[i], a[ixj]);
        }
      }
}

/* Sigma clip: src/bachUtil.cpp:34-128 with cl_kern/bach.cl:196-251.
 * Block-of-32 partial sums in list order, then a sequential host sum. */
void sigmaClip(const double* data, int n, int maxIter, double alpha, double* mean,
               double* stdDev) {
  if (n == 0) {
    *mean = 0.0;
    *stdDev = 1e10;
    return;
  }
  const int L = 32;
  const int nb = (n + L - 1) / L;
  std::vector<uint8_t> clip((size_t)n, 0);
  std::vector<double> bs((size_t)nb), bs2((size_t)nb);
  size_t cur = 0, prev = (size_t)n;
  for (int it = 0; it < maxIter && cur != prev; it++) {
    if (prev <= 1) {
      *mean = 0.0;
      *stdDev = 1e10;
      return;
    }
    cur = prev;
    for (int g = 0; g < nb; g++) {
      double s = 0.0, s2 = 0.0;
      int cnt = std::min(L, n - g * L);
      for (int i = 0; i < cnt; i++) {
        int k = g * L + i;
        double d = clip[k] == 0 ? data[k] : 0.0;
        s += d;
        s2 += d * d;
      }
      bs[g] = s;
      bs2[g] = s2;
    }
    double sum = 0.0, sum2 = 0.0;
    for (int g = 0; g < nb; g++) sum += bs[g];
    for (int g = 0; g < nb; g++) sum2 += bs2[g];
    double np = (double)prev;
    double tMean = sum / np;
    double tSd = std::sqrt((sum2 - np * tMean * tMean) / (double)(prev - 1));
    double inv = 1.0 / tSd;
    int clipCount = 0;
    for (int k = 0; k < n; k++)
      if (clip[k] == 0 && std::fabs(data[k] - tMean) * inv > alpha) {
        clip[k] = 1;
        clipCount++;
      }
    prev = cur - (size_t)clipCount;
    *mean = tMean;
    *stdDev = tSd;
  }
}

/* Histogram statistics of one stamp: cl_kern/sss.cl:190-324 (createHistogram).
 * `pixOk(i)` enumerates the stamp's pixels in raster order. */
template <class Visit>
void histogramStats(Visit visit, const double* sorted, int sampleCount, int nSamples,
                    double mean, double invStdDev, double iqRange, double alpha,
                    double* skyOut, double* fwhmOut) {
  const double upProc = 0.9, midProc = 0.5;
  double up = sorted[(int)(upProc * sampleCount)];
  double mid = sorted[(int)(midProc * sampleCount)];
  double binSize = (up - mid) / (double)nSamples;
  double lowerBinVal = mid - (128.0 * binSize);
  int bins[256];
  int attempts = 0, okCount = 0;
  double lower = 0.0, upper = 0.0, skyEst = 0.0;
  bool setFwhm = true;
  while (true) {
    if (attempts >= 5) { setFwhm = false; break; }
    for (int i = 0; i < 256; i++) bins[i] = 0;
    okCount = 0;
    visit([&](double v) {
      if (std::fabs(v - mean) * invStdDev > alpha) return;
      double q = std::floor((v - lowerBinVal) / binSize) + 1.0;
      /* saturating float->int as the device conversion, then clamp(0,255) */
      int idx = (q != q) ? 0 : (q <= 0.0 ? 0 : (q >= 255.0 ? 255 : (int)q));
      bins[idx]++;
      okCount++;
    });
    if (okCount == 0 || binSize == 0.0) { setFwhm = false; break; }
    double sumBins = 0.0, maxDens = 0.0;
    int lowerIndex = 1, upperIndex = 1, maxIndex = -1;
    while (upperIndex < 255) {
      while (sumBins < okCount / 10.0 && upperIndex < 255) sumBins += bins[upperIndex++];
      if (sumBins / (upperIndex - lowerIndex) > maxDens) {
        maxDens = sumBins / (upperIndex - lowerIndex);
        maxIndex = lowerIndex;
      }
      sumBins -= bins[lowerIndex++];
    }
    if (maxIndex < 0 || maxIndex > 255) maxIndex = 0;
    sumBins = 0.0;
    double sumExpect = 0.0;
    for (int i = maxIndex; sumBins < okCount / 10.0 && i < 255; i++) {
      sumBins += bins[i];
      sumExpect += i * bins[i];
    }
    double modeBin = sumExpect / sumBins + 0.5;
    skyEst = lowerBinVal + binSize * (modeBin - 1.0);
    lower = okCount * 0.25;
    upper = okCount * 0.75;
    sumBins = 0.0;
    int i = 0;
    while (sumBins < lower) sumBins += bins[i++];
    lower = i - (sumBins - lower) / bins[i - 1];
    while (sumBins < upper) sumBins += bins[i++];
    upper = i - (sumBins - upper) / bins[i - 1];
    if (lower < 1.0 || upper > 255.0) {
      lowerBinVal -= 128.0 * binSize;
      binSize *= 2;
    } else if (upper - lower < 40.0) {
      binSize /= 3.0;
      lowerBinVal = skyEst - 128.0 * binSize;
    } else {
      break;
    }
    attempts++;
  }
  *fwhmOut = setFwhm ? binSize * (upper - lower) / iqRange : 0.0;
  *skyOut = skyEst;
}

/* ------------------------------------------------------------------------ */
/* Stamps of one image                                                       */
/* ------------------------------------------------------------------------ */
struct Stamps {
  int n = 0;
  std::vector<int> ox, oy, sw, sh;
  std::vector<double> sky, fwhm;
  std::vector<int> found, count, cur, grid;
  std::vector<int> ssx, ssy;
  std::vector<double> ssv;
  std::vector<double> W, Q, B;   /* per stamp: wRows*wCols, q*q, q */
};

/* cl_kern/sss.cl:19-38 (createStampBounds), src/sssUtil.cpp:24-37. */
void createStamps(const Cfg& c, int w, int h, int nx, int ny, int fsw, Stamps& st) {
  st.n = nx * ny;
  st.ox.assign(st.n, 0); st.oy.assign(st.n, 0); st.sw.assign(st.n, 0); st.sh.assign(st.n, 0);
  for (int id = 0; id < st.n; id++) {
    int sx = id % nx, sy = id / nx;
    int startX = sx * w / nx, startY = sy * h / ny;
    int stopX = std::min(startX + fsw, w), stopY = std::min(startY + fsw, h);
    st.ox[id] = startX; st.oy[id] = startY;
    st.sw[id] = stopX - startX; st.sh[id] = stopY - startY;
  }
  st.sky.assign(st.n, 0.0); st.fwhm.assign(st.n, 0.0);
  st.found.assign(st.n, 0); st.count.assign(st.n, 0); st.cur.assign(st.n, 0);
  st.grid.resize(st.n);
  for (int i = 0; i < st.n; i++) st.grid[i] = i;
  st.ssx.assign((size_t)st.n * c.maxSS, INT_MAX);
  st.ssy.assign((size_t)st.n * c.maxSS, INT_MAX);
  st.ssv.assign((size_t)st.n * c.maxSS, -kInf);
}

/* calcStats: src/bachUtil.cpp:130-280 with cl_kern/sss.cl:81-188 and 190-324. */
int calcStats(const Cfg& c, Pair& P, int which, int fsw, Stamps& st) {
  const int w = P.w, h = P.h;
  const double* img = P.img[which].data();
  const int nSamples = 100, padded = 128;
  for (int s = 0; s < st.n; s++)
    if (st.sw[s] * st.sh[s] < nSamples) return XB_ERR_STAMP_PIXELS;
  const int nPix = fsw * fsw;
  std::vector<double> means(st.n), inv(st.n);
  std::vector<std::vector<double>> sortedS(st.n);
  std::vector<int> sampleCounts(st.n);
  /* sampleStamp + pad + bitonic sort: independent per stamp */
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < st.n; s++) {
    Ran1 rng;
    int cnt = 0;
    std::vector<double> smp(padded, kInf);
    for (int i = 0; i < nSamples; i++) smp[i] = 0.0;
    const int sx = st.sw[s], sy = st.sh[s];
    const int npix = sx * sy;
    for (int iter = 0; cnt < nSamples && iter < npix; iter++) {
      int rx = (int)std::floor(rng.next() * (int)sx);
      int ry = (int)std::floor(rng.next() * (int)sy);
      size_t idx = (size_t)(rx + st.ox[s]) + (size_t)(ry + st.oy[s]) * w;
      if (P.mask[idx] > 0 || std::fabs(img[idx]) <= 1e-10) continue;
      smp[cnt++] = img[idx];
    }
    sampleCounts[s] = cnt;
    bitonic(smp.data(), padded);
    sortedS[s] = smp;
  }
  /* maskStamp (canonical raster order of good pixels) + sigma clip */
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < st.n; s++) {
    std::vector<double> good;
    good.reserve(nPix);
    for (int k = 0; k < nPix; k++) {
      int x = k % fsw, y = k / fsw;
      int xI = x + st.ox[s], yI = y + st.oy[s];
      if (xI >= w || yI >= h) continue;
      size_t idx = (size_t)xI + (size_t)yI * w;
      double pixel = img[idx];
      if (P.mask[idx] > 0 || pixel <= 1e-10) continue;
      if (std::isnan(pixel)) {
        P.mask[idx] |= XB_MASK_NAN_PIXEL | XB_MASK_BAD_INPUT;
        continue;
      }
      good.push_back(pixel);
    }
    double m, sd;
    sigmaClip(good.data(), (int)good.size(), 3, c.a.sigClipAlpha, &m, &sd);
    means[s] = m;
    inv[s] = 1.0 / sd;
  }
  /* createHistogram */
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < st.n; s++) {
    const int ox = st.ox[s], oy = st.oy[s], sx = st.sw[s], sy = st.sh[s];
    auto visit = [&](auto&& f) {
      for (int y = 0; y < sy; y++)
        for (int x = 0; x < sx; x++) {
          size_t idx = (size_t)(ox + x) + (size_t)(oy + y) * w;
          double v = img[idx];
          if (P.mask[idx] != 0 || v <= 1e-10) continue;
          f(v);
        }
    };
    histogramStats(visit, sortedS[s].data(), sampleCounts[s], nSamples, means[s], inv[s],
                   c.a.iqRange, c.a.sigClipAlpha, &st.sky[s], &st.fwhm[s]);
  }
  return XB_OK;
}

/* checkSStamp: cl_kern/sss.cl:326-361. */
double checkSStamp(const double* img, u16* mask, double sky, double fwhm, long imgW, int cx,
                   int cy, int hss, int ox, int oy, int sw, int sh, double hi, double tkf,
                   u16 badMask, u16 badPixelMask) {
  double ret = 0.0;
  int startX = std::max(cx - hss, ox), startY = std::max(cy - hss, oy);
  int endX = std::min(cx + hss, ox + sw - 1), endY = std::min(cy + hss, oy + sh - 1);
  for (int y = startY; y <= endY; y++)
    for (int x = startX; x <= endX; x++) {
      long a = x + y * imgW;
      if ((mask[a] & badMask) > 0) return 0.0;
      double v = img[a];
      if (v > hi) {
        mask[a] |= badPixelMask;
        return 0.0;
      }
      if ((v - sky) / fwhm > tkf) ret += v;
    }
  return ret;
}

/* findSubStamps: cl_kern/sss.cl:381-505, src/sssUtil.cpp:40-104. */
void findSubStamps(const Cfg& c, Pair& P, int which, int fsw, Stamps& st) {
  const bool isT = which == 0;
  u16 badMask = XB_MASK_ALL & ~XB_MASK_OK_CONV;
  u16 badPixelMask, skipMask;
  if (isT) {
    badMask &= ~(XB_MASK_BAD_PIXEL_S | XB_MASK_SKIP_S);
    badPixelMask = XB_MASK_BAD_PIXEL_T;
    skipMask = XB_MASK_SKIP_T;
  } else {
    badMask &= ~(XB_MASK_BAD_PIXEL_T | XB_MASK_SKIP_T);
    badPixelMask = XB_MASK_BAD_PIXEL_S;
    skipMask = XB_MASK_SKIP_S;
  }
  const double* img = P.img[which].data();
  u16* mask = P.mask.data();
  const long imgW = P.w;
  const int hss = c.a.hSStampWidth;
  const int maxSS = c.maxSS;
  const double hi = c.a.threshHigh, tkf = c.a.threshKernFit;
  /* stamps are disjoint and every access below stays inside the stamp */
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < st.n; s++) {
    const double sky = st.sky[s], fwhm = st.fwhm[s];
    const double floorV = sky + tkf * fwhm;
    double dfrac = 0.9;
    const int ox = st.ox[s], oy = st.oy[s], sw = st.sw[s], sh = st.sh[s];
    int lx[XB_MAX_SS], ly[XB_MAX_SS];
    double lv[XB_MAX_SS];
    int counter = 0;
    /* The reference loops until the threshold reaches the floor; with a
     * non-finite sky or sigma it never does, so the levels are capped at 32
     * (finite statistics finish within 6 levels). */
    for (int level = 0; level < 32 && counter < maxSS; level++) {
      double lim = std::max(floorV, sky + (hi - sky) * dfrac);
      for (int y = 0; y < fsw; y++) {
        int absy = y + oy;
        for (int x = 0; x < fsw; x++) {
          int absx = x + ox;
          long a = absx + (absy * imgW);
          if ((mask[a] & badMask) > 0) continue;
          double v = img[a];
          if (v > hi) {
            mask[a] |= badPixelMask;
            continue;
          }
          if ((v - sky) * (1.0 / fwhm) < tkf) continue;
          if (v > lim) {
            double maxVal = 0;
            int mx = absx, my = absy;
            int startX = std::max(absx - hss, ox), startY = std::max(absy - hss, oy);
            int endX = std::min(absx + hss, ox + fsw - 1), endY = std::min(absy + hss, oy + fsw - 1);
            for (int ky = startY; ky <= endY; ky++)
              for (int kx = startX; kx <= endX; kx++) {
                long kc = kx + (ky * imgW);
                double kv = img[kc];
                if ((mask[kc] & badMask) > 0) continue;
                if (kv >= hi) {
                  mask[kc] |= badPixelMask;
                  continue;
                }
                if ((kv - sky) * (1.0 / fwhm) < tkf) continue;
                if (kv > maxVal) {
                  maxVal = kv;
                  mx = kx;
                  my = ky;
                }
              }
            maxVal = checkSStamp(img, mask, sky, fwhm, imgW, mx, my, hss, ox, oy, sw, sh, hi,
                                 tkf, badMask, badPixelMask);
            if (maxVal == 0.0) continue;
            lx[counter] = mx;
            ly[counter] = my;
            lv[counter] = maxVal;
            counter++;
            int sx2 = std::max(mx - hss, ox), sy2 = std::max(my - hss, oy);
            int ex2 = std::min(mx + hss, ox + sw - 1), ey2 = std::min(my + hss, oy + sh - 1);
            for (int yy = sy2; yy <= ey2; yy++)
              for (int xx = sx2; xx <= ex2; xx++) mask[xx + yy * imgW] |= skipMask;
          }
          if (counter >= maxSS) break;
        }
        if (counter >= maxSS) break;
      }
      if (lim == floorV) break;
      dfrac -= 0.2;
    }
    /* sortSubStamps: insertion sort, descending by value (sss.cl:363-379) */
    for (int i = 1; i < counter; i++)
      for (int j = i; j > 0 && lv[j - 1] < lv[j]; j--) {
        std::swap(lx[j - 1], lx[j]);
        std::swap(ly[j - 1], ly[j]);
        std::swap(lv[j - 1], lv[j]);
      }
    st.found[s] = counter;
    st.count[s] = std::min(counter, maxSS / 2);
    for (int i = 0; i < maxSS; i++) {
      size_t o = (size_t)s * maxSS + i;
      if (i < counter) {
        st.ssx[o] = lx[i]; st.ssy[o] = ly[i]; st.ssv[o] = lv[i];
      } else {
        st.ssx[o] = INT_MAX; st.ssy[o] = INT_MAX; st.ssv[o] = -kInf;
      }
    }
  }
}

/* removeEmptyStamps: cl_kern/sss.cl:507-571, src/sssUtil.cpp:106-197.  The
 * bitonic sort of keep indices restores ascending stamp order. */
void removeEmpty(const Cfg& c, Stamps& st) {
  Stamps o;
  o.n = 0;
  for (int s = 0; s < st.n; s++) {
    if (st.count[s] <= 0) continue;
    o.ox.push_back(st.ox[s]); o.oy.push_back(st.oy[s]);
    o.sw.push_back(st.sw[s]); o.sh.push_back(st.sh[s]);
    o.sky.push_back(st.sky[s]); o.fwhm.push_back(st.fwhm[s]);
    o.found.push_back(st.found[s]); o.count.push_back(st.count[s]);
    o.cur.push_back(0);
    o.grid.push_back(st.grid[s]);
    for (int i = 0; i < c.maxSS; i++) {
      size_t k = (size_t)s * c.maxSS + i;
      bool copied = i < st.count[s];
      o.ssx.push_back(copied ? st.ssx[k] : INT_MAX);
      o.ssy.push_back(copied ? st.ssy[k] : INT_MAX);
      o.ssv.push_back(copied ? st.ssv[k] : -kInf);
    }
    o.n++;
  }
  st = o;
}

/* ------------------------------------------------------------------------ */
/* fillStamps: cl_kern/cmv.cl:70-248, src/cmvUtil.cpp:37-187                 */
/* ------------------------------------------------------------------------ */
void fillOne(const Cfg& c, int w, int h, const double* img, const double* img2, bool valid,
             int ssx, int ssy, double* W, double* Q, double* B) {
  const int kw = c.kw, hkw = c.hkw, ssw = c.ssw, hss = c.hssw;
  const int nP = c.nPSF;
  const int xHalf = hss + hkw, yHalf = hss;
  const int tw = 2 * xHalf + 1, th = 2 * yHalf + 1;
  const int wCols = c.wCols;
  std::vector<double> tmp((size_t)tw * th);
  for (int n = 0; n < nP; n++) {
    /* convStampY (cmv.cl:70-106): float image value, double filter tap,
     * FP32 accumulator (FP64 in full-FP64 mode) */
    for (int p = 0; p < tw * th; p++) {
      if (!valid) { tmp[p] = 0.0; continue; }
      int j = ssy + p / tw - yHalf;
      int i = ssx + p % tw - xHalf;
      if (c.mode == XO_MODE_REF) {
        float v = 0.0f;
        for (int y = -hkw; y <= hkw; y++) {
          float iv = (float)img[(size_t)i + (size_t)(j + y) * w];
          v = (float)((double)v + (double)iv * c.fy[(size_t)n * kw + hkw - y]);
        }
        tmp[p] = (double)v;
      } else {
        double v = 0.0;
        for (int y = -hkw; y <= hkw; y++) {
          double iv = (double)(float)img[(size_t)i + (size_t)(j + y) * w];
          v = v + iv * c.fy[(size_t)n * kw + hkw - y];
        }
        tmp[p] = v;
      }
    }
    /* convStampX (cmv.cl:108-135): FP64 accumulator */
    const int subWidth = kw + ssw - 1;
    for (int p = 0; p < wCols; p++) {
      int j = p / ssw - hss;
      int i = p % ssw - hss;
      double w0 = 0.0;
      for (int x = -hkw; x <= hkw; x++) {
        int idx = (i + x) + subWidth / 2 + (j + hss) * subWidth;
        w0 += tmp[idx] * c.fx[(size_t)n * kw + hkw - x];
      }
      W[(size_t)n * wCols + p] = w0;
    }
  }
  /* convStampOdd (cmv.cl:137-154) */
  for (int n = 1; n < nP; n++) {
    if ((c.kx[n] % 2 == 0) && (c.ky[n] % 2 == 0))
      for (int p = 0; p < wCols; p++) W[(size_t)n * wCols + p] -= W[p];
  }
  /* convStampBg (cmv.cl:156-193): xf, yf are computed in FP32 */
  for (int b = 0; b < c.nBG; b++)
    for (int p = 0; p < wCols; p++) {
      double a = 0.0;
      if (valid) {
        int x = ssx - hss + p % ssw;
        int y = ssy - hss + p / ssw;
        float hw = (float)w * 0.5f, hh = (float)h * 0.5f;
        double xf = (double)(((float)x - hw) / hw);
        double yf = (double)(((float)y - hh) / hh);
        a = pown(xf, c.bgx[b]) * pown(yf, c.bgy[b]);
      }
      W[(size_t)(nP + b) * wCols + p] = a;
    }
  /* createQ (cmv.cl:195-215) */
  const int q = c.qCount;
  for (int i = 0; i < q; i++)
    for (int j = 0; j < q; j++) {
      double q0 = 0.0;
      if (i > 0 && j > 0 && j <= i)
        for (int k = 0; k < ssw * ssw; k++)
          q0 += W[(size_t)(i - 1) * wCols + k] * W[(size_t)(j - 1) * wCols + k];
      Q[(size_t)i * q + j] = q0;
    }
  /* createB (cmv.cl:217-248): x outer, y inner */
  for (int i = 0; i < c.bCount; i++) {
    double p0 = 0.0;
    if (valid && i > 0)
      for (int x = -hss; x <= hss; x++)
        for (int y = -hss; y <= hss; y++) {
          int k = x + hss + ssw * (y + hss);
          size_t ii = (size_t)(x + ssx) + (size_t)(y + ssy) * w;
          p0 += W[(size_t)(i - 1) * wCols + k] * img2[ii];
        }
    B[i] = p0;
  }
}

void fillStamps(const Cfg& c, const Pair& P, const double* img, const double* img2, Stamps& st,
                const std::vector<int>& which) {
  const size_t wsz = (size_t)c.wRows * c.wCols, qsz = (size_t)c.qCount * c.qCount;
  st.W.resize((size_t)st.n * wsz);
  st.Q.resize((size_t)st.n * qsz);
  st.B.resize((size_t)st.n * c.bCount);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < (int)which.size(); k++) {
    int s = which[k];
    bool valid = st.cur[s] < st.count[s];
    int sx = 0, sy = 0;
    if (valid) {
      sx = st.ssx[(size_t)s * c.maxSS + st.cur[s]];
      sy = st.ssy[(size_t)s * c.maxSS + st.cur[s]];
    }
    fillOne(c, P.w, P.h, img, img2, valid, sx, sy, &st.W[s * wsz], &st.Q[s * qsz],
            &st.B[(size_t)s * c.bCount]);
  }
}

/* ------------------------------------------------------------------------ */
/* LU solvers                                                                */
/* ------------------------------------------------------------------------ */
/* CPU NR Crout (src/bachUtil.cpp:319-384), 1-based, a is (n+1)x(n+1). */
int ludcmp(double* a, int n, int* index, double* d) {
  const int N = n + 1;
  std::vector<double> vv((size_t)n + 1, 0.0);
  int maxI = 0;
  *d = 1.0;
  for (int i = 1; i <= n; i++) {
    double big = 0.0;
    for (int j = 1; j <= n; j++) {
      double t = std::fabs(a[(size_t)i * N + j]);
      if (t > big) big = t;
    }
    if (big == 0.0) return 1;
    vv[i] = 1.0 / big;
  }
  for (int j = 1; j <= n; j++) {
    for (int i = 1; i < j; i++) {
      double sum = a[(size_t)i * N + j];
      for (int k = 1; k < i; k++) sum -= a[(size_t)i * N + k] * a[(size_t)k * N + j];
      a[(size_t)i * N + j] = sum;
    }
    double big = 0.0;
    for (int i = j; i <= n; i++) {
      double sum = a[(size_t)i * N + j];
      for (int k = 1; k < j; k++) sum -= a[(size_t)i * N + k] * a[(size_t)k * N + j];
      a[(size_t)i * N + j] = sum;
      double dum = vv[i] * std::fabs(sum);
      if (dum >= big) {
        big = dum;
        maxI = i;
      }
    }
    if (j != maxI) {
      for (int k = 1; k <= n; k++) std::swap(a[(size_t)maxI * N + k], a[(size_t)j * N + k]);
      *d = -*d;
      vv[maxI] = vv[j];
    }
    index[j] = maxI;
    if (a[(size_t)j * N + j] == 0.0) a[(size_t)j * N + j] = 1.0e-20;
    if (j != n) {
      double dum = 1.0 / a[(size_t)j * N + j];
      for (int i = j + 1; i <= n; i++) a[(size_t)i * N + j] *= dum;
    }
  }
  return 0;
}

/* src/bachUtil.cpp:386-411 */
void lubksb(const double* a, int n, const int* index, double* b) {
  const int N = n + 1;
  int ii = 0;
  for (int i = 1; i <= n; i++) {
    int ip = index[i];
    double sum = b[ip];
    b[ip] = b[i];
    if (ii) {
      for (int j = ii; j <= i - 1; j++) sum -= a[(size_t)i * N + j] * b[j];
    } else if (sum) {
      ii = i;
    }
    b[i] = sum;
  }
  for (int i = n; i >= 1; i--) {
    double sum = b[i];
    for (int j = i + 1; j <= n; j++) sum -= a[(size_t)i * N + j] * b[j];
    b[i] = sum / a[(size_t)i * N + i];
  }
}

/* Per-stamp device LU (cl_kern/bach.cl:68-194): matrix m x m, rows 1..m-1.
 * ludcmpBig has no zero-row exit (1/0 = inf). */
void luStamp(double* a, int m, double* b) {
  std::vector<double> vv((size_t)m, 0.0);
  std::vector<int> index((size_t)m, 0);
  for (int i = 1; i < m; i++) {
    double big = 0.0;
    for (int j = 1; j < m; j++) {
      double v = std::fabs(a[(size_t)i * m + j]);
      if (v > big) big = v;
    }
    vv[i] = 1.0 / big;
  }
  int maxI = 0;
  for (int j = 1; j < m; j++) {
    for (int i = 1; i < j; i++) {
      double sum = a[(size_t)i * m + j];
      for (int k = 1; k < i; k++) sum -= a[(size_t)i * m + k] * a[(size_t)k * m + j];
      a[(size_t)i * m + j] = sum;
    }
    double big = 0.0;
    for (int i = j; i < m; i++) {
      double sum = a[(size_t)i * m + j];
      for (int k = 1; k < j; k++) sum -= a[(size_t)i * m + k] * a[(size_t)k * m + j];
      a[(size_t)i * m + j] = sum;
      double dum = vv[i] * std::fabs(sum);
      if (dum >= big) {
        big = dum;
        maxI = i;
      }
    }
    if (j != maxI) {
      for (int k = 1; k < m; k++) std::swap(a[(size_t)maxI * m + k], a[(size_t)j * m + k]);
      vv[maxI] = vv[j];
    }
    index[j] = maxI;
    if (a[(size_t)j * m + j] == 0.0) a[(size_t)j * m + j] = 1.0e-20;
    if (j != m - 1) {
      double dum = 1.0 / a[(size_t)j * m + j];
      for (int i = j + 1; i < m; i++) a[(size_t)i * m + j] *= dum;
    }
  }
  int ii = 0;
  for (int i = 1; i < m; i++) {
    int ip = index[i];
    double sum = b[ip];
    b[ip] = b[i];
    if (ii != 0) {
      for (int j = ii; j <= i - 1; j++) sum -= a[(size_t)i * m + j] * b[j];
    } else if (sum != 0.0) {
      ii = i;
    }
    b[i] = sum;
  }
  for (int i = m - 1; i >= 1; i--) {
    double sum = b[i];
    for (int j = i + 1; j < m; j++) sum -= a[(size_t)i * m + j] * b[j];
    b[i] = sum / a[(size_t)i * m + i];
  }
}

/* ------------------------------------------------------------------------ */
/* Global normal matrix                                                      */
/* ------------------------------------------------------------------------ */
/* Spatial weights of one stamp from its current substamp; xf, yf in FP32 as
 * cl_kern/cd.cl:170-171 and src/cdkscUtil.cpp:317-332. */
void stampWeights(const Cfg& c, int w, int h, int x, int y, double* wt) {
  float hw = (float)(0.5 * w), hh = (float)(0.5 * h);
  double fx = (double)(((float)x - hw) / hw);
  double fy = (double)(((float)y - hh) / hh);
  for (int k = 0; k < c.nComp2; k++) wt[k] = pown(fx, c.cdx[k]) * pown(fy, c.cdy[k]);
}

/* Global matrix, both reference forms (the element sums are identical:
 * cl_kern/cd.cl:182-295 GPU form used by cd, src/cdkscUtil.cpp:307-402 CPU
 * form used by ksc).  `use` lists stamps in order; stamps whose current
 * substamp is exhausted are skipped (CPU form) or contribute zeros (GPU form).
 * M is (matSize+1)^2 row-major, 1-based. */
void globalMatrix(const Cfg& c, int w, int h, const Stamps& st, const std::vector<int>& use,
                  std::vector<double>& M, std::vector<double>& wts) {
  const int N = c.matSize + 1, nC2 = c.nComp2, nComp = c.nComp, nC1 = c.nComp1;
  const size_t wsz = (size_t)c.wRows * c.wCols, qsz = (size_t)c.qCount * c.qCount;
  const int q = c.qCount, pix = c.wCols;
  M.assign((size_t)N * N, 0.0);
  wts.assign(use.size() * nC2, 0.0);
  for (size_t u = 0; u < use.size(); u++) {
    int s = use[u];
    if (st.cur[s] < st.count[s]) {
      int x = st.ssx[(size_t)s * c.maxSS + st.cur[s]];
      int y = st.ssy[(size_t)s * c.maxSS + st.cur[s]];
      stampWeights(c, w, h, x, y, &wts[u * nC2]);
    }
  }
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < nComp; i++) {
    int i1 = i / nC2, i2 = i - i1 * nC2;
    for (size_t u = 0; u < use.size(); u++) {
      int s = use[u];
      if (st.cur[s] >= st.count[s]) continue;
      const double* wt = &wts[u * nC2];
      const double* Q = &st.Q[s * qsz];
      for (int j = 0; j <= i; j++) {
        int j1 = j / nC2, j2 = j - j1 * nC2;
        M[(size_t)(i + 2) * N + (j + 2)] += wt[i2] * wt[j2] * Q[(size_t)(i1 + 2) * q + (j1 + 2)];
      }
    }
  }
  for (size_t u = 0; u < use.size(); u++) {
    int s = use[u];
    if (st.cur[s] >= st.count[s]) continue;
    const double* wt = &wts[u * nC2];
    const double* Q = &st.Q[s * qsz];
    const double* W = &st.W[s * wsz];
    M[(size_t)1 * N + 1] += Q[(size_t)1 * q + 1];
    for (int i = 0; i < nComp; i++) {
      int i1 = i / nC2, i2 = i - i1 * nC2;
      M[(size_t)(i + 2) * N + 1] += wt[i2] * Q[(size_t)(i1 + 2) * q + 1];
    }
    for (int iBG = 0; iBG < c.nBG; iBG++) {
      int i = nComp + iBG + 1;
      int iv = nC1 + iBG + 1;
      for (int i1 = 1; i1 < nC1 + 1; i1++) {
        double p0 = 0.0;
        for (int k = 0; k < pix; k++) p0 += W[(size_t)i1 * pix + k] * W[(size_t)iv * pix + k];
        for (int i2 = 0; i2 < nC2; i2++) {
          int jj = (i1 - 1) * nC2 + i2 + 1;
          M[(size_t)(i + 1) * N + (jj + 1)] += p0 * wt[i2];
        }
      }
      double p0 = 0.0;
      for (int k = 0; k < pix; k++) p0 += W[k] * W[(size_t)iv * pix + k];
      M[(size_t)(i + 1) * N + 1] += p0;
      for (int jBG = 0; jBG <= iBG; jBG++) {
        double qq = 0.0;
        for (int k = 0; k < pix; k++)
          qq += W[(size_t)iv * pix + k] * W[(size_t)(nC1 + jBG + 1) * pix + k];
        M[(size_t)(i + 1) * N + (nComp + jBG + 2)] += qq;
      }
    }
  }
  for (int i = 0; i < c.matSize; i++)
    for (int j = 0; j <= i; j++) M[(size_t)(j + 1) * N + (i + 1)] = M[(size_t)(i + 1) * N + (j + 1)];
}

/* Right-hand side.  rowMajorBg selects the GPU form's (cl_kern/cd.cl:297-352,
 * y outer) background sums; otherwise the CPU form (src/cdkscUtil.cpp:429-475,
 * x outer). */
void globalRhs(const Cfg& c, int w, const Stamps& st, const std::vector<int>& use,
               const std::vector<double>& wts, const double* img, bool rowMajorBg,
               std::vector<double>& res) {
  const int nC1 = c.nComp1, nC2 = c.nComp2, hss = c.hssw, ssw = c.ssw;
  const size_t wsz = (size_t)c.wRows * c.wCols;
  res.assign((size_t)c.nSolComp, 0.0);
  for (size_t u = 0; u < use.size(); u++) {
    int s = use[u];
    if (st.cur[s] >= st.count[s]) continue;
    const double* B = &st.B[(size_t)s * c.bCount];
    const double* W = &st.W[s * wsz];
    const double* wt = &wts[u * nC2];
    int ssx = st.ssx[(size_t)s * c.maxSS + st.cur[s]];
    int ssy = st.ssy[(size_t)s * c.maxSS + st.cur[s]];
    res[1] += B[1];
    for (int i = 1; i < nC1 + 1; i++) {
      double p0 = B[i + 1];
      for (int j = 0; j < nC2; j++) {
        int indx = (i - 1) * nC2 + j + 1;
        res[indx + 1] += p0 * wt[j];
      }
    }
    for (int b = 0; b < c.nBG; b++) {
      double q = 0.0;
      const double* Wb = &W[(size_t)(nC1 + b + 1) * c.wCols];
      if (rowMajorBg) {
        for (int y = -hss; y <= hss; y++)
          for (int x = -hss; x <= hss; x++) {
            int index = x + hss + ssw * (y + hss);
            q += Wb[index] * img[(size_t)(x + ssx) + (size_t)(y + ssy) * w];
          }
      } else {
        for (int x = -hss; x <= hss; x++)
          for (int y = -hss; y <= hss; y++) {
            int index = x + hss + ssw * (y + hss);
            q += Wb[index] * (double)(float)img[(size_t)(x + ssx) + (size_t)(y + ssy) * w];
          }
      }
      res[(size_t)c.nComp + b + 2] += q;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Kernel realisation                                                        */
/* ------------------------------------------------------------------------ */
/* Coefficients at (xf, yf): cl_kern/cd.cl:354-383, src/bachUtil.cpp:508-526. */
void kernelCoeffs(const Cfg& c, const double* sol, double xf, double yf, double* co) {
  int k = 2;
  for (int i = 1; i < c.nPSF; i++) {
    double ci = 0.0, aX = 1.0;
    for (int ix = 0; ix <= c.a.kernelOrder; ix++) {
      double aY = 1.0;
      for (int iy = 0; iy <= c.a.kernelOrder - ix; iy++) {
        ci += sol[k++] * aX * aY;
        aY *= yf;
      }
      aX *= xf;
    }
    co[i] = ci;
  }
  co[0] = sol[1];
}

/* GPU makeKernel + sumKernel (cl_kern/cd.cl:385-442, src/bachUtil.cpp:413-500):
 * kernel sum as block-of-32 partial sums then their sum. */
double kernelSumDevice(const Cfg& c, const double* sol, int w, int h, int x, int y) {
  double hw = 0.5 * w, hh = 0.5 * h;
  double xf = (x - hw) / hw, yf = (y - hh) / hh;
  std::vector<double> co((size_t)c.nPSF);
  kernelCoeffs(c, sol, xf, yf, co.data());
  const int cnt = c.kw * c.kw;
  std::vector<double> kern((size_t)cnt);
  for (int i = 0; i < cnt; i++) {
    double k0 = 0.0;
    for (int p = 0; p < c.nPSF; p++) k0 += co[p] * c.vec[(size_t)p * cnt + i];
    kern[i] = k0;
  }
  int count = cnt;
  std::vector<double> a = kern, b;
  while (count > 1) {
    int nb = (count + 31) / 32;
    b.assign((size_t)nb, 0.0);
    for (int g = 0; g < nb; g++) {
      double s = 0.0;
      int lc = std::min(32, count - g * 32);
      for (int i = 0; i < lc; i++) s += a[(size_t)g * 32 + i];
      b[g] = s;
    }
    a.swap(b);
    count = nb;
  }
  return a[0];
}

/* CPU makeKernel (src/bachUtil.cpp:502-541): kernel at (x, y) and its sum. */
double makeKernelCpu(const Cfg& c, const double* sol, int w, int h, int x, int y, double* kern) {
  double hw = 0.5 * w, hh = 0.5 * h;
  double xf = (x - hw) / hw, yf = (y - hh) / hh;
  double co[64];
  kernelCoeffs(c, sol, xf, yf, co);
  const int cnt = c.kw * c.kw;
  double sum = 0.0;
  for (int i = 0; i < cnt; i++) {
    double k0 = 0.0;
    for (int p = 0; p < c.nPSF; p++) k0 += co[p] * c.vec[(size_t)p * cnt + i];
    kern[i] = k0;
    sum += k0;
  }
  return sum;
}

/* getBackground (cl_kern/cd.cl:444-467). */
double background(const Cfg& c, int x, int y, const double* sol, int w, int h) {
  double xf = (x - 0.5 * w) / (0.5 * w);
  double yf = (y - 0.5 * h) / (0.5 * h);
  double bg = 0.0, ax = 1.0;
  int k = c.nComp + 2;
  for (int i = 0; i <= c.a.backgroundOrder; i++) {
    double ay = 1.0;
    for (int j = 0; j <= c.a.backgroundOrder - i; j++) {
      bg += sol[k++] * ax * ay;
      ay *= yf;
    }
    ax *= xf;
  }
  return bg;
}

/* calcSigs (src/cdkscUtil.cpp:477-583, cl_kern/cd.cl:469-655): per stamp mean
 * chi2 over good pixels of the current substamp, or -1.  NaN flags are
 * collected and applied after the parallel loop (same final mask). */
void calcSigs(const Cfg& c, Pair& P, const double* tImg, const double* sImg,
              const Stamps& st, const std::vector<int>& use, const double* sol,
              std::vector<double>& sig) {
  const int w = P.w, h = P.h, ssw = c.ssw, pix = c.wCols;
  const size_t wsz = (size_t)c.wRows * c.wCols;
  sig.assign(use.size(), 0.0);
  std::vector<std::vector<size_t>> nanFlags(use.size());
#pragma omp parallel for schedule(dynamic)
  for (int u = 0; u < (int)use.size(); u++) {
    int s = use[u];
    bool valid = st.cur[s] < st.count[s];
    int ssx = 0, ssy = 0;
    double bg = 0.0;
    if (valid) {
      ssx = st.ssx[(size_t)s * c.maxSS + st.cur[s]];
      ssy = st.ssy[(size_t)s * c.maxSS + st.cur[s]];
      bg = background(c, ssx, ssy, sol, w, h);
    }
    /* makeModel (cd.cl:489-531) */
    std::vector<double> model((size_t)pix, 0.0);
    if (valid) {
      const double* W = &st.W[s * wsz];
      double xf = (ssx - w * 0.5) / (w * 0.5);
      double yf = (ssy - h * 0.5) / (h * 0.5);
      std::vector<double> co((size_t)c.nPSF);
      kernelCoeffs(c, sol, xf, yf, co.data());
      for (int j = 0; j < pix; j++) {
        if (c.mode == XO_MODE_REF) {
          float m0 = (float)(sol[1] * W[j]);
          for (int i = 1; i < c.nPSF; i++)
            m0 = (float)((double)m0 + co[i] * W[(size_t)i * pix + j]);
          model[j] = (double)m0;
        } else {
          double m0 = sol[1] * W[j];
          for (int i = 1; i < c.nPSF; i++) m0 = m0 + co[i] * W[(size_t)i * pix + j];
          model[j] = m0;
        }
      }
    }
    /* calcSig + reduceSig: 32-lane blocks over the padded pixel range */
    const int L = 32;
    const int reduceCount = (pix + L - 1) / L;
    std::vector<double> part((size_t)reduceCount);
    std::vector<int> cnt((size_t)reduceCount);
    for (int g = 0; g < reduceCount; g++) {
      double ls = 0.0;
      int lc = 0;
      for (int li = 0; li < L; li++) {
        int gi = g * L + li;
        int gx = gi % ssw, gy = gi / ssw;
        double s0 = 0.0;
        if (valid && gx < ssw && gy < ssw) {
          int absX = gx - ssw / 2 + ssx, absY = gy - ssw / 2 + ssy;
          size_t ai = (size_t)absX + (size_t)absY * w;
          double tDat = model[(size_t)gi];
          double sDat = sImg[ai];
          double diff = tDat - sDat + bg;
          if ((P.mask[ai] & XB_MASK_BAD_INPUT) == 0 && std::fabs(sDat) > 1e-10) {
            if (std::isnan(tDat) || std::isnan(sDat)) {
              nanFlags[u].push_back(ai);
            } else {
              lc++;
              s0 = diff * diff / (std::fabs(tImg[ai]) + std::fabs(sDat));
            }
          }
        }
        ls += s0;
      }
      part[g] = ls;
      cnt[g] = lc;
    }
    /* reduceSig levels */
    int count = reduceCount;
    bool first = true;
    double result = 0.0;
    std::vector<double> pin = part, pout;
    std::vector<int> cin = cnt, cout;
    while (count > 1 || first) {
      int next = (count + L - 1) / L;
      pout.assign((size_t)next, 0.0);
      cout.assign((size_t)next, 0);
      for (int g = 0; g < next; g++) {
        double sum = 0.0;
        int lc = 0;
        int c2 = std::min(L, count - g * L);
        for (int i = 0; i < c2; i++) sum += pin[(size_t)g * L + i];
        for (int i = 0; i < c2; i++) lc += cin[(size_t)g * L + i];
        if (next == 1) {
          if (lc == 0) sum = -1.0;
          else {
            sum /= lc;
            if (sum >= 1e10) sum = -1.0;
          }
          result = sum;
        }
        pout[g] = sum;
        cout[g] = lc;
      }
      pin.swap(pout);
      cin.swap(cout);
      count = next;
      first = false;
    }
    sig[u] = result;
  }
  for (auto& v : nanFlags)
    for (size_t ai : v) P.mask[ai] |= XB_MASK_NAN_PIXEL;
}

/* ------------------------------------------------------------------------ */
/* cd: testFit (src/cdkscUtil.cpp:4-266)                                     */
/* ------------------------------------------------------------------------ */
double testFit(const Cfg& c, Pair& P, int tIdx, Stamps& st, int* nTest, double* ksumOut,
               double* solOut) {
  const double* tImg = P.img[tIdx].data();
  const double* sImg = P.img[1 - tIdx].data();
  const int n = st.n;
  const int q = c.qCount;
  const size_t qsz = (size_t)q * q;
  *nTest = 0;
  std::vector<double> ksum((size_t)n);
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < n; s++) {
    std::vector<double> m(qsz), v((size_t)q);
    const double* Q = &st.Q[s * qsz];
    for (int i = 0; i < q; i++)
      for (int j = 0; j < q; j++)
        m[(size_t)i * q + j] = (i > 0 && j > 0) ? Q[(size_t)std::max(i, j) * q + std::min(i, j)] : 0.0;
    for (int i = 0; i < q; i++) v[i] = i > 0 ? st.B[(size_t)s * c.bCount + i] : 0.0;
    luStamp(m.data(), q, v.data());
    ksum[s] = v[1];
  }
  double kMean, kStd;
  sigmaClip(ksum.data(), n, 10, c.a.sigClipAlpha, &kMean, &kStd);
  std::vector<int> test;
  for (int s = 0; s < n; s++) {
    double diff = std::fabs((ksum[s] - kMean) / kStd);
    if (diff < c.a.sigKernFit) test.push_back(s);
  }
  *nTest = (int)test.size();
  if (test.empty()) return 666;
  std::vector<double> M, wts, rhs;
  globalMatrix(c, P.w, P.h, st, test, M, wts);
  globalRhs(c, P.w, st, test, wts, sImg, true, rhs);
  capture(M, rhs);
  std::vector<int> index((size_t)c.nSolComp + 1, 0);
  double d;
  ludcmp(M.data(), c.matSize, index.data(), &d);
  lubksb(M.data(), c.matSize, index.data(), rhs.data());
  for (int i = 0; i < c.nSolComp; i++) solOut[i] = rhs[i];
  double kernelSum = kernelSumDevice(c, rhs.data(), P.w, P.h, 0, 0);
  *ksumOut = kernelSum;
  std::vector<double> merits;
  calcSigs(c, P, tImg, sImg, st, test, rhs.data(), merits);
  std::vector<double> clean;
  for (double m : merits)
    if (m >= 0.0) clean.push_back(m);
  if (clean.empty()) return 666;
  double mMean, mStd;
  sigmaClip(clean.data(), (int)clean.size(), 10, c.a.sigClipAlpha, &mMean, &mStd);
  return mMean / kernelSum;
}

/* ------------------------------------------------------------------------ */
/* ksc: fitKernel + checkFitSolution (src/cdkscUtil.cpp:585-802)             */
/* ------------------------------------------------------------------------ */
void fitKernel(const Cfg& c, Pair& P, int tIdx, Stamps& st, std::vector<double>& sol,
               int* iterations) {
  const double* tImg = P.img[tIdx].data();
  const double* sImg = P.img[1 - tIdx].data();
  std::vector<int> all((size_t)st.n);
  for (int s = 0; s < st.n; s++) all[s] = s;
  int it = 0;
  bool check;
  do {
    std::vector<double> M, wts, rhs;
    globalMatrix(c, P.w, P.h, st, all, M, wts);
    globalRhs(c, P.w, st, all, wts, sImg, false, rhs);
    capture(M, rhs);
    std::vector<int> index((size_t)c.matSize + 1, 0);
    double d;
    ludcmp(M.data(), c.matSize, index.data(), &d);
    lubksb(M.data(), c.matSize, index.data(), rhs.data());
    sol = rhs;
    check = false;
    std::vector<double> sig;
    calcSigs(c, P, tImg, sImg, st, all, sol.data(), sig);
    /* checkBadSubStamps (ksc.cl:1-49) */
    std::vector<double> chi2;
    std::vector<int> refill;
    for (int s = 0; s < st.n; s++) {
      if (st.cur[s] < st.count[s]) {
        if (sig[s] == -1.0) {
          st.cur[s]++;
          refill.push_back(s);
        } else {
          chi2.push_back(sig[s]);
        }
      }
    }
    if (!refill.empty()) {
      fillStamps(c, P, tImg, sImg, st, refill);
      check = true;
    }
    double mean = 0.0, sd = 0.0;
    sigmaClip(chi2.data(), (int)chi2.size(), 10, c.a.sigClipAlpha, &mean, &sd);
    /* checkBadSubStampsFromSigmaClip (ksc.cl:51-72) */
    refill.clear();
    for (int s = 0; s < st.n; s++) {
      if (st.cur[s] < st.count[s]) {
        if ((sig[s] - mean) > c.a.sigKernFit * sd) {
          st.cur[s]++;
          refill.push_back(s);
        }
      }
    }
    if (!refill.empty()) {
      fillStamps(c, P, tImg, sImg, st, refill);
      check = true;
    }
    it++;
  } while (check);
  *iterations = it;
}

/* ------------------------------------------------------------------------ */
/* conv + sub (src/bach.cpp:325-440, cl_kern/conv.cl, cl_kern/sub.cl)         */
/* ------------------------------------------------------------------------ */
void convSub(const Cfg& c, Pair& P, int tIdx, bool convTemplate, const std::vector<double>& sol,
             std::vector<double>& convImg, std::vector<double>& diffImg, double* kernSumOut) {
  const int w = P.w, h = P.h, kw = c.kw, hk = kw / 2;
  const double* T = P.img[tIdx].data();
  const double* S = P.img[1 - tIdx].data();
  const bool scaleConv = (c.a.normalizeTemplate && convTemplate) || (!c.a.normalizeTemplate && !convTemplate);
  const int xSteps = (int)std::ceil(w / (double)kw);
  const int ySteps = (int)std::ceil(h / (double)kw);
  const int kk = kw * kw;
  std::vector<double> kern((size_t)xSteps * ySteps * kk);
#pragma omp parallel for schedule(static)
  for (int t = 0; t < xSteps * ySteps; t++) {
    int xs = t % xSteps, ys = t / xSteps;
    makeKernelCpu(c, sol.data(), w, h, xs * kw + c.hkw + c.hkw, ys * kw + c.hkw + c.hkw,
                  &kern[(size_t)t * kk]);
  }
  std::vector<double> tmpK((size_t)kk);
  double kernSum = makeKernelCpu(c, sol.data(), w, h, w / 2, h / 2, tmpK.data());
  *kernSumOut = kernSum;
  const double invKernMult = scaleConv ? 1.0 / kernSum : 1.0;
  const double hi = c.a.threshHigh, lo = c.a.threshLow;
  std::vector<u16> cm((size_t)w * h);
  for (size_t i = 0; i < cm.size(); i++) {
    double t = T[i];
    u16 m = 0;
    if (t == 0.0) m |= XB_MASK_BAD_INPUT | XB_MASK_BAD_PIX_VAL;
    if (t >= hi) m |= XB_MASK_BAD_INPUT | XB_MASK_SAT_PIXEL;
    if (t <= lo) m |= XB_MASK_BAD_INPUT | XB_MASK_LOW_PIXEL;
    cm[i] = m;
  }
  convImg.assign((size_t)w * h, 0.0);
  const double thr = (double)0.99f;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      size_t id = (size_t)y * w + x;
      if (x >= hk && x < w - hk && y >= hk && y < h - hk) {
        int xS = (x - hk) / kw, yS = (y - hk) / kw;
        const double* K = &kern[(size_t)(xS + yS * xSteps) * kk];
        double acc = 0.0, aks = 0.0, uks = 0.0;
        int maskAcc = 0;
        for (int j = y - hk; j <= y + hk; j++) {
          int jk = y - j + hk;
          for (int i = x - hk; i <= x + hk; i++) {
            int ik = x - i + hk;
            double k = K[ik + jk * kw];
            size_t ii = (size_t)i + (size_t)w * j;
            acc += k * T[ii];
            maskAcc |= cm[ii];
            aks += std::fabs(k);
            if ((cm[ii] & XB_MASK_BAD_INPUT) == 0) uks += std::fabs(k);
          }
        }
        acc += background(c, x, y, sol.data(), w, h);
        acc *= invKernMult;
        convImg[id] = acc;
        u16 nm = cm[id];
        if (cm[id] & XB_MASK_BAD_INPUT) nm |= XB_MASK_BAD_OUTPUT;
        if (maskAcc != 0) {
          if ((uks / aks) < thr) nm |= XB_MASK_BAD_OUTPUT | XB_MASK_BAD_CONV;
          else nm |= XB_MASK_OK_CONV;
        }
        P.mask[id] = nm;
      } else {
        convImg[id] = 1e-30;
      }
    }
  /* maskAfterConv (conv.cl:99-113) */
  for (size_t i = 0; i < (size_t)w * h; i++) {
    double t = S[i];
    u16 m = 0;
    if (t == 0.0) m |= XB_MASK_BAD_OUTPUT | XB_MASK_BAD_INPUT | XB_MASK_BAD_PIX_VAL;
    if (t >= hi) m |= XB_MASK_BAD_OUTPUT | XB_MASK_BAD_INPUT | XB_MASK_SAT_PIXEL;
    if (t <= lo) m |= XB_MASK_BAD_OUTPUT | XB_MASK_BAD_INPUT | XB_MASK_LOW_PIXEL;
    P.mask[i] |= m;
  }
  /* sub (sub.cl:15-32) */
  const double convFactor = scaleConv ? kernSum : 1.0;
  const double finalFactor = scaleConv ? -(1.0 / kernSum) : 1.0;
  diffImg.assign((size_t)w * h, 0.0);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      size_t id = (size_t)y * w + x;
      double d = 1e-30;
      if (x >= hk && x < w - hk && y >= hk && y < h - hk)
        if ((P.mask[id] & XB_MASK_BAD_OUTPUT) == 0) d = (convImg[id] * convFactor - S[id]) * finalFactor;
      diffImg[id] = d;
    }
}

double msSince(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void recordStamps(const Cfg& c, const Stamps& st, int im, xb_trace* tr) {
  for (int s = 0; s < st.n && s < XB_MAX_STAMPS; s++) {
    tr->sky[im][s] = st.sky[s];
    tr->fwhm[im][s] = st.fwhm[s];
    tr->ssCount[im][s] = st.count[s];
    for (int i = 0; i < c.maxSS; i++) {
      tr->ssX[im][s * XB_MAX_SS + i] = st.ssx[(size_t)s * c.maxSS + i];
      tr->ssY[im][s * XB_MAX_SS + i] = st.ssy[(size_t)s * c.maxSS + i];
      tr->ssVal[im][s * XB_MAX_SS + i] = st.ssv[(size_t)s * c.maxSS + i];
    }
  }
}

}  // namespace

/* ------------------------------------------------------------------------ */
/* Public entry points                                                       */
/* ------------------------------------------------------------------------ */
extern "C" int xo_run(const xb_params* pin, int w, int h, const float* tmpl, const float* sci,
                      int mode, int threads, float* convOut, float* diffOut, uint16_t* maskOut,
                      xb_trace* tr, double* stage_ms) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  Cfg c;
  if (!pin || !tmpl || !sci || w <= 0 || h <= 0) return XB_ERR_ARG;
  if (!setupCfg(c, pin, mode)) return XB_ERR_ARG;
  if (tr) std::memset(tr, 0, sizeof(*tr));
  double ms[9] = {0};
  auto tAll = std::chrono::steady_clock::now();
  auto t0 = tAll;
  /* init (src/bach.cpp:14-42) */
  Pair P;
  P.w = w;
  P.h = h;
  P.img[0].resize((size_t)w * h);
  P.img[1].resize((size_t)w * h);
  for (size_t i = 0; i < (size_t)w * h; i++) {
    P.img[0][i] = (double)tmpl[i];
    P.img[1][i] = (double)sci[i];
  }
  maskInput(c, P);
  ms[0] = msSince(t0);
  /* sss (src/bach.cpp:44-170) */
  t0 = std::chrono::steady_clock::now();
  xb_params& a = c.a;
  int fsw = std::min(w / a.stampsx, h / a.stampsy);
  fsw -= a.fKernelWidth;
  fsw -= fsw % 2 == 0 ? 1 : 0;
  if (fsw < a.fSStampWidth) {
    fsw = a.fSStampWidth + a.fKernelWidth;
    fsw -= fsw % 2 == 0 ? 1 : 0;
    a.stampsx = w / fsw;
    a.stampsy = h / fsw;
  }
  if (a.stampsx * a.stampsy > XB_MAX_STAMPS || a.stampsx <= 0 || a.stampsy <= 0) return XB_ERR_CAPACITY;
  Stamps st[2];
  auto identify = [&]() -> int {
    createStamps(c, w, h, a.stampsx, a.stampsy, fsw, st[0]);
    createStamps(c, w, h, a.stampsx, a.stampsy, fsw, st[1]);
    int rc = calcStats(c, P, 0, fsw, st[0]);
    if (rc) return rc;
    rc = calcStats(c, P, 1, fsw, st[1]);
    if (rc) return rc;
    findSubStamps(c, P, 0, fsw, st[0]);
    findSubStamps(c, P, 1, fsw, st[1]);
    return XB_OK;
  };
  int rc = identify();
  if (rc) return rc;
  Stamps full[2] = {st[0], st[1]};
  const int oldCount = a.stampsx * a.stampsy;
  removeEmpty(c, st[0]);
  removeEmpty(c, st[1]);
  int retried = 0;
  if ((double)st[0].n / oldCount < 0.1 || (double)st[1].n / oldCount < 0.1) {
    retried = 1;
    for (auto& m : P.mask) m &= (u16) ~(XB_MASK_SKIP_S | XB_MASK_SKIP_T);
    rc = identify();
    if (rc) return rc;
    full[0] = st[0];
    full[1] = st[1];
    removeEmpty(c, st[0]);
    removeEmpty(c, st[1]);
  }
  if (tr) {
    tr->w = w; tr->h = h;
    tr->stampsx = a.stampsx; tr->stampsy = a.stampsy; tr->fStampWidth = fsw;
    tr->nGrid = oldCount; tr->retried = retried;
    recordStamps(c, full[0], 0, tr);
    recordStamps(c, full[1], 1, tr);
    for (int im = 0; im < 2; im++) {
      tr->stampCount[im] = st[im].n;
      for (int s = 0; s < st[im].n; s++) tr->stampIndex[im][s] = st[im].grid[s];
    }
  }
  ms[1] = msSince(t0);
  if (st[0].n == 0 && st[1].n == 0) return XB_ERR_NO_SUBSTAMPS;
  /* cmv (src/bach.cpp:172-268) */
  t0 = std::chrono::steady_clock::now();
  for (int im = 0; im < 2; im++) {
    std::vector<int> all((size_t)st[im].n);
    for (int s = 0; s < st[im].n; s++) all[s] = s;
    fillStamps(c, P, P.img[im].data(), P.img[1 - im].data(), st[im], all);
  }
  ms[2] = msSince(t0);
  /* cd (src/bach.cpp:270-313) */
  t0 = std::chrono::steady_clock::now();
  double merit[2], ksumCd[2];
  int nTest[2];
  std::vector<double> cdSol[2];
  for (int im = 0; im < 2; im++) {
    cdSol[im].assign((size_t)c.nSolComp, 0.0);
    ksumCd[im] = 0.0;
    if (st[im].n == 0) { merit[im] = 666; nTest[im] = 0; continue; }
    merit[im] = testFit(c, P, im, st[im], &nTest[im], &ksumCd[im], cdSol[im].data());
  }
  const bool convTemplate = merit[1] > merit[0];
  const int tIdx = convTemplate ? 0 : 1;
  ms[3] = msSince(t0);
  /* ksc (src/bach.cpp:315-323) */
  t0 = std::chrono::steady_clock::now();
  std::vector<double> sol;
  int iters = 0;
  fitKernel(c, P, tIdx, st[tIdx], sol, &iters);
  ms[4] = msSince(t0);
  /* conv + sub */
  t0 = std::chrono::steady_clock::now();
  std::vector<double> convImg, diffImg;
  double kernSum = 0.0;
  convSub(c, P, tIdx, convTemplate, sol, convImg, diffImg, &kernSum);
  ms[5] = msSince(t0);
  ms[6] = 0.0;
  ms[7] = 0.0;
  ms[8] = msSince(tAll);
  if (convOut)
    for (size_t i = 0; i < convImg.size(); i++) convOut[i] = (float)convImg[i];
  if (diffOut)
    for (size_t i = 0; i < diffImg.size(); i++) diffOut[i] = (float)diffImg[i];
  if (maskOut) std::memcpy(maskOut, P.mask.data(), P.mask.size() * sizeof(u16));
  if (tr) {
    for (int im = 0; im < 2; im++) {
      tr->nTest[im] = nTest[im];
      tr->merit[im] = merit[im];
      tr->kernelSumCd[im] = ksumCd[im];
      for (int i = 0; i < c.nSolComp; i++) tr->cdSolution[im][i] = cdSol[im][i];
    }
    tr->convTemplate = convTemplate ? 1 : 0;
    tr->kscIterations = iters;
    tr->nSol = c.nSolComp;
    for (int i = 0; i < c.nSolComp; i++) tr->solution[i] = sol[i];
    for (int s = 0; s < st[tIdx].n; s++) tr->finalCurrent[s] = st[tIdx].cur[s];
    tr->kernSum = kernSum;
  }
  if (stage_ms)
    for (int i = 0; i < 9; i++) stage_ms[i] = ms[i];
  return XB_OK;
}

extern "C" void xo_ran1_seq(int n, double* out) {
  Ran1 r;
  for (int i = 0; i < n; i++) out[i] = r.next();
}

extern "C" void xo_sigma_clip(const double* data, int n, int maxIter, double alpha, double* mean,
                              double* stdDev) {
  sigmaClip(data, n, maxIter, alpha, mean, stdDev);
}

extern "C" int xo_lu_solve(double* a, int n, double* b, int* index) {
  double d;
  for (int i = 0; i <= n; i++) index[i] = 0;
  int rc = ludcmp(a, n, index, &d);
  lubksb(a, n, index, b);
  return rc;
}

extern "C" void xo_lu_solve_stamp(double* a, int m, double* b) { luStamp(a, m, b); }

extern "C" int xo_basis(const xb_params* p, double* fx, double* fy, double* vec) {
  Cfg c;
  if (!setupCfg(c, p, XO_MODE_REF)) return -1;
  if (fx) std::memcpy(fx, c.fx.data(), c.fx.size() * sizeof(double));
  if (fy) std::memcpy(fy, c.fy.data(), c.fy.size() * sizeof(double));
  if (vec) std::memcpy(vec, c.vec.data(), c.vec.size() * sizeof(double));
  return c.nPSF;
}

extern "C" void xo_bitonic(double* v, int n) { bitonic(v, n); }

extern "C" void xo_histogram(const double* pix, int npix, const double* sorted, int sampleCount,
                             double mean, double invStdDev, double iqRange, double alpha,
                             double* sky, double* fwhm) {
  auto visit = [&](auto&& f) {
    for (int i = 0; i < npix; i++)
      if (!(pix[i] <= 1e-10)) f(pix[i]);
  };
  histogramStats(visit, sorted, sampleCount, 100, mean, invStdDev, iqRange, alpha, sky, fwhm);
}

extern "C" int xo_fill_one(const xb_params* p, int mode, int w, int h, const double* img,
                           const double* img2, int ssx, int ssy, double* W, double* Q, double* B) {
  Cfg c;
  if (!setupCfg(c, p, mode)) return -1;
  fillOne(c, w, h, img, img2, true, ssx, ssy, W, Q, B);
  return 0;
}

extern "C" void xo_capture(int enable) {
  g_capture = enable != 0;
  g_capMats.clear();
  g_capRhs.clear();
}

extern "C" int xo_captured_count(void) { return (int)g_capMats.size(); }

/* Copies captured matrix i ((matSize+1)^2) and rhs (nSolComp); returns the
 * matrix dimension matSize+1 or -1. */
extern "C" int xo_captured_get(int i, double* M, double* rhs) {
  if (i < 0 || i >= (int)g_capMats.size()) return -1;
  if (M) std::memcpy(M, g_capMats[i].data(), g_capMats[i].size() * sizeof(double));
  if (rhs) std::memcpy(rhs, g_capRhs[i].data(), g_capRhs[i].size() * sizeof(double));
  return (int)std::lround(std::sqrt((double)g_capMats[i].size()));
}
