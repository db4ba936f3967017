// xb_api.cpp — host side of libxbach: context, resident buffers, the stage
// functions of the reference pipeline (src/bach.cpp:14-440) and the C ABI.
//
// Control flow stays on the host as in the reference (stage order, the <10 %
// retry, the cd direction choice, the ksc do-while), but every stage is a few
// stream-ordered launches and the host reads back only the handful of scalars
// that steer it (kept-stamp counts, merits, the ksc `check` flag).
#include <hip/hip_runtime.h>

#include <cstddef>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "xb_internal.h"

using namespace xb;

std::shared_mutex& xb::alloc_lock() {
  static std::shared_mutex m;
  return m;
}

namespace {

#define XB_HIP(expr)                                                              \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      c.err = std::string(#expr) + ": " + hipGetErrorString(_e);                  \
      return XB_ERR_HIP;                                                          \
    }                                                                             \
  } while (0)

template <class T>
hipError_t dalloc(T** p, size_t n) {
  if (*p) {
    hipError_t e = locked_free(*p);
    if (e != hipSuccess) return e;
    *p = nullptr;
  }
  if (n == 0) n = 1;
  return locked_malloc(p, n * sizeof(T));
}

template <class T>
void dfree(T** p) {
  if (*p) {
    (void)locked_free(*p);
    *p = nullptr;
  }
}

// ---- basis tables (cl_kern/cmv.cl:1-68, include/datatypeUtil.h:37-118) ----
void buildBasis(const xb_params& a, Basis& B) {
  int n = 0;
  for (int g = 0; g < 3; g++)
    for (int x = 0; x <= a.dg[g]; x++)
      for (int y = 0; y <= a.dg[g] - x; y++) {
        const double b = (double)a.bg[g];
        double sumX = 0.0, sumY = 0.0;
        for (int i = 0; i < KW; i++) {
          double xv = (double)(i - KW / 2);
          double qe = std::exp(-xv * xv * b);
          double px = 1.0, py = 1.0;
          for (int e = 0; e < x; e++) px *= xv;
          for (int e = 0; e < y; e++) py *= xv;
          B.fx[n * KW + i] = qe * px;
          B.fy[n * KW + i] = qe * py;
          sumX += B.fx[n * KW + i];
          sumY += B.fy[n * KW + i];
        }
        const bool even = (x % 2 == 0) && (y % 2 == 0);
        if (even) {
          double ix = 1.0 / sumX, iy = 1.0 / sumY;
          for (int i = 0; i < KW; i++) {
            B.fx[n * KW + i] *= ix;
            B.fy[n * KW + i] *= iy;
          }
        }
        B.kx[n] = x;
        B.ky[n] = y;
        B.evenSub[n] = (even && n > 0) ? 1 : 0;
        n++;
      }
  for (int k = 0; k < NPSF; k++)
    for (int v = 0; v < KW; v++)
      for (int u = 0; u < KW; u++) {
        double vv = B.fx[k * KW + u] * B.fy[k * KW + v];
        if (B.evenSub[k]) vv -= B.fx[u] * B.fy[v];
        B.vec[k * KK + v * KW + u] = vv;
      }
}

// ---- NR ran1 draws, seed -666 (cl_kern/sss.cl:41-79) ----------------------
// BF16 bits of x, round to nearest even (NaN stays a quiet NaN).
uint16_t bf16Bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// The mixed form's 2-D basis as MFMA B operands (k_cmv_mixed.hip): for MFMA
// step ks, basis tile nt, part (BF16 h, m, l) and lane l, 8 taps
// k = 16 ks + 8 (l / 32) + t of basis n = 32 nt + l % 32, tap k = (u, v) =
// (k / 24, k % 24); value fy_n[20-u] * fx_n[20-v] for u, v < 21 and n < 49,
// else 0, rounded to FP32 and split x = h + m + l (each part BF16 of what the
// previous parts leave).
float bf16Float(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

void mixedTable(const Basis& B, std::vector<uint16_t>& out) {
  constexpr int KSTEPS = 32, TAPW = 24, PARTS = 3;
  out.assign((size_t)KSTEPS * 2 * PARTS * 64 * 8, 0);
  for (int ks = 0; ks < KSTEPS; ks++)
    for (int nt = 0; nt < 2; nt++)
      for (int l = 0; l < 64; l++)
        for (int t = 0; t < 8; t++) {
          const int k = 16 * ks + 8 * (l / 32) + t, u = k / TAPW, v = k % TAPW;
          const int n = 32 * nt + l % 32;
          double x = 0.0;
          if (u < KW && v < KW && n < NPSF) x = B.fy[n * KW + 20 - u] * B.fx[n * KW + 20 - v];
          const float xf = (float)x;
          const uint16_t h = bf16Bits(xf);
          const float r1 = xf - bf16Float(h);
          const uint16_t m = bf16Bits(r1);
          const uint16_t lo = bf16Bits(r1 - bf16Float(m));
          const size_t at = ((((size_t)ks * 2 + nt) * PARTS) * 64 + l) * 8 + t;
          out[at] = h;
          out[at + 64 * 8] = m;
          out[at + 2 * 64 * 8] = lo;
        }
}

// The same table for the FP16 split (XB_FORM_FP16): [step][basis tile][h, l]
// [lane][8], value * 2^scale rounded to FP32, then x = h + l in FP16 (round to
// nearest even); the kernel scales W back.  scale puts the largest magnitude
// just below 2^14 (the odd terms, x^a y^b times a Gaussian, are not normalised
// and exceed 1).  Returns scale.
int mixedTable16(const Basis& B, std::vector<uint16_t>& out) {
  constexpr int KSTEPS = 32, TAPW = 24, PARTS = 2;
  out.assign((size_t)KSTEPS * 2 * PARTS * 64 * 8, 0);
  double big = 0.0;
  for (int n = 0; n < NPSF; n++)
    for (int u = 0; u < KW; u++)
      for (int v = 0; v < KW; v++)
        big = std::max(big, std::fabs((double)(float)(B.fy[n * KW + u] * B.fx[n * KW + v])));
  int e = 0;
  if (big > 0.0) (void)std::frexp(big, &e);  // big < 2^e
  const int scale = 14 - e;
  for (int ks = 0; ks < KSTEPS; ks++)
    for (int nt = 0; nt < 2; nt++)
      for (int l = 0; l < 64; l++)
        for (int t = 0; t < 8; t++) {
          const int k = 16 * ks + 8 * (l / 32) + t, u = k / TAPW, v = k % TAPW;
          const int n = 32 * nt + l % 32;
          double x = 0.0;
          if (u < KW && v < KW && n < NPSF) x = B.fy[n * KW + 20 - u] * B.fx[n * KW + 20 - v];
          const float xf = std::ldexp((float)x, scale);
          const _Float16 h = (_Float16)xf;
          const _Float16 lo = (_Float16)(xf - (float)h);
          const size_t at = ((((size_t)ks * 2 + nt) * PARTS) * 64 + l) * 8 + t;
          std::memcpy(&out[at], &h, 2);
          std::memcpy(&out[at + 64 * 8], &lo, 2);
        }
  return scale;
}

void ran1Table(int n, std::vector<double>& out) {
  const long M1 = 259200, IA1 = 7141, IC1 = 54773;
  const long M2 = 134456, IA2 = 8121, IC2 = 28411;
  const long M3 = 243000, IA3 = 4561, IC3 = 51349;
  const double RM1 = 1.0 / M1, RM2 = 1.0 / M2;
  long ix1, ix2, ix3;
  double r[98];
  int idum = -666;
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
  out.resize((size_t)n);
  for (int k = 0; k < n; k++) {
    ix1 = (IA1 * ix1 + IC1) % M1;
    ix2 = (IA2 * ix2 + IC2) % M2;
    ix3 = (IA3 * ix3 + IC3) % M3;
    int j = (int)(1 + ((97 * ix3) / M3));
    out[k] = r[j];
    r[j] = (ix1 + ix2 * RM2) * RM1;
  }
}

int checkParams(const xb_params& p, std::string& err) {
  bool ok = p.fKernelWidth == KW && p.hKernelWidth == HKW && p.fSStampWidth == SSW &&
            p.hSStampWidth == HSS && p.maxKSStamps == KEEPSS && p.kernelOrder == 2 &&
            p.backgroundOrder == 1 && p.ngauss == 3 && p.dg[0] == 6 && p.dg[1] == 4 &&
            p.dg[2] == 2;
  if (!ok) {
    err = "libxbach is built for the reference's fixed method constants "
          "(21x21 kernel, 31x31 substamps, dg 6/4/2, kernel order 2, background order 1, "
          "maxKSStamps 3)";
    return XB_ERR_ARG;
  }
  if (p.stampsx <= 0 || p.stampsy <= 0) {
    err = "stamp grid must be positive";
    return XB_ERR_ARG;
  }
  return XB_OK;
}

int ensureImage(Ctx& c, int w, int h) {
  size_t n = (size_t)w * h;
  if (n <= c.capPix) return XB_OK;
  XB_HIP(dalloc(&c.img[0], n));
  XB_HIP(dalloc(&c.img[1], n));
  XB_HIP(dalloc(&c.mask, n));
  XB_HIP(dalloc(&c.conv, n));
  XB_HIP(dalloc(&c.diff, n));
  const int xSteps = (w + KW - 1) / KW, ySteps = (h + KW - 1) / KW;
  const int tiles = xSteps * ySteps;
  XB_HIP(dalloc(&c.tileK64, (size_t)tiles * KK));
  XB_HIP(dalloc(&c.tileAks, (size_t)tiles));
  c.capTiles = tiles;
  c.capPix = n;
  return XB_OK;
}

int ensureStamps(Ctx& c, int N, int fsw) {
  if (N > c.capStamps) {
    for (int im = 0; im < 2; im++) {
      DevStamps& s = c.st[im];
      XB_HIP(dalloc(&s.sky, N));
      XB_HIP(dalloc(&s.fwhm, N));
      XB_HIP(dalloc(&s.found, N));
      XB_HIP(dalloc(&s.count, N));
      XB_HIP(dalloc(&s.ssx, (size_t)N * MAXSS));
      XB_HIP(dalloc(&s.ssy, (size_t)N * MAXSS));
      XB_HIP(dalloc(&s.ssv, (size_t)N * MAXSS));
      XB_HIP(dalloc(&s.kept, N));
      XB_HIP(dalloc(&s.cur, N));
      XB_HIP(dalloc(&s.W, (size_t)N * WROWS * NPIX));
      XB_HIP(dalloc(&s.Wc, (size_t)N * WROWS * NPIX));
      XB_HIP(dalloc(&s.G, (size_t)N * WROWS * WROWS));
      XB_HIP(dalloc(&s.Bc, (size_t)N * WROWS));
      XB_HIP(dalloc(&s.Br, (size_t)N * NBG));
      XB_HIP(dalloc(&s.sig, N));
      XB_HIP(dalloc(&s.ksum, N));
      XB_HIP(dalloc(&s.list, N));
      XB_HIP(dalloc(&s.sigK, N));
      XB_HIP(dalloc(&s.listK, N));
    }
    c.capStamps = N;
    c.capFsw = 0;  // good-pixel scratch depends on N too
  }
  if (fsw > c.capFsw) {
    // two halves: the substamp searches of the two images run at the same
    // time and each keeps its candidate lists in its own half
    XB_HIP(dalloc(&c.goodPix, (size_t)2 * c.capStamps * fsw * fsw));
    c.capFsw = fsw;
  }
  const int draws = 2 * fsw * fsw;
  if (draws > c.ran1Count) {
    std::vector<double> t;
    ran1Table(draws, t);
    XB_HIP(dalloc(&c.ran1, (size_t)draws));
    XB_HIP(hipMemcpyAsync(c.ran1, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice,
                          c.stream));
    XB_HIP(hipStreamSynchronize(c.stream));
    c.ran1Count = draws;
  }
  return XB_OK;
}

// Waits for an event by polling: the scalar read-backs between passes are on
// the critical path of every pair, and a blocking wait adds ~40 us of wake-up.
// The polling thread yields its core between polls after the first few, so
// many contexts waiting at once (batch workers) leave the host's cores to the
// threads that have work.
hipError_t spinEvent(hipEvent_t ev) {
  hipError_t e;
  for (int n = 0; (e = hipEventQuery(ev)) == hipErrorNotReady; n++)
    if (n >= 32) std::this_thread::yield();
  return e;
}

// Waits for the work queued on the primary stream by polling an event: the
// scalar read-backs between stages are on the critical path of every pair, and
// a blocking wait added ~40 us of wake-up to each.
hipError_t spinSync(Ctx& c) {
  hipError_t e = hipEventRecord(c.evSync, c.stream);
  if (e != hipSuccess) return e;
  return spinEvent(c.evSync);
}

// Copies `bytes` of the device scalars at byte offset `off` into the pinned
// mirror and waits for them.
hipError_t readBack(Ctx& c, size_t off, size_t bytes) {
  hipError_t e = hipMemcpyAsync(reinterpret_cast<char*>(c.hsc) + off,
                                reinterpret_cast<const char*>(c.sc) + off, bytes,
                                hipMemcpyDeviceToHost, c.stream);
  return e != hipSuccess ? e : spinSync(c);
}

int readScalars(Ctx& c, DevScalars& out) {
  XB_HIP(readBack(c, 0, sizeof(DevScalars)));
  out = *c.hsc;
  return XB_OK;
}

double nowMs() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// ---- stages -----------------------------------------------------------------
// The two global systems (one per direction) and the solution buffers.
int ensureSolver(Ctx& c) {
  if (c.sol) return XB_OK;
  for (int i = 0; i < 4; i++) {
    LuSet& L = i < 2 ? c.lu[i] : c.ksLu[i - 2];
    XB_HIP(locked_malloc((void**)&L.M, sizeof(double) * MATN * MATN));
    XB_HIP(locked_malloc((void**)&L.rhs, sizeof(double) * NSOL));
    XB_HIP(locked_malloc((void**)&L.M0, sizeof(double) * MATN * MATN));
    XB_HIP(locked_malloc((void**)&L.rhs0, sizeof(double) * NSOL));
    XB_HIP(locked_malloc((void**)&L.luIndex, sizeof(int) * MATN));
    XB_HIP(locked_malloc((void**)&L.luVV, sizeof(double) * MATN));
    XB_HIP(locked_malloc((void**)&L.luState, LU_STATE_BYTES));
    XB_HIP(locked_malloc((void**)&L.rr, sizeof(double) * NSOL));
    XB_HIP(locked_malloc((void**)&L.F, sizeof(double) * LDLT_WORK_DOUBLES));
  }
  XB_HIP(locked_malloc((void**)&c.cdSol, sizeof(double) * 2 * NSOL));
  XB_HIP(locked_malloc((void**)&c.ksSol, sizeof(double) * 2 * NSOL));
  XB_HIP(locked_malloc((void**)&c.sol, sizeof(double) * NSOL));
  return XB_OK;
}

// Runs f(0) on the primary stream and f(1) on the side stream, both after the
// work already queued on the primary stream; the primary stream then waits for
// both.  Used where the two images' work is independent (the reference runs it
// one after the other; the results do not depend on the order).
template <class F>
int forkBoth(Ctx& c, F f) {
  if (c.side == c.stream) {  // one-stream context: the two parts run in order
    int rc = XB_OK;
    for (int im = 0; im < 2 && rc == XB_OK; im++) rc = f(im);
    return rc;
  }
  XB_HIP(hipEventRecord(c.evFork, c.stream));
  XB_HIP(hipStreamWaitEvent(c.side, c.evFork, 0));
  int rc = XB_OK;
  for (int im = 0; im < 2 && rc == XB_OK; im++) {
    c.cur = im == 0 ? c.stream : c.side;
    rc = f(im);
  }
  c.cur = c.stream;
  XB_HIP(hipEventRecord(c.evJoin, c.side));
  XB_HIP(hipStreamWaitEvent(c.stream, c.evJoin, 0));
  return rc;
}

int syncAll(Ctx& c);

int stageInitCommon(Ctx& c, const xb_params* p, int w, int h) {
  if (!p || w <= 0 || h <= 0) {
    c.err = "bad arguments";
    return XB_ERR_ARG;
  }
  if (c.specRestore >= 0) {  // a restore of the last run is still queued
    const int rc = syncAll(c);
    if (rc) return rc;
  }
  int rc = checkParams(*p, c.err);
  if (rc) return rc;
  c.p = *p;
  c.m.threshLow = p->threshLow;
  c.m.threshHigh = p->threshHigh;
  c.m.threshKernFit = p->threshKernFit;
  c.m.sigKernFit = p->sigKernFit;
  c.m.sigClipAlpha = p->sigClipAlpha;
  c.m.iqRange = p->iqRange;
  c.m.normalizeTemplate = p->normalizeTemplate;
  c.w = w;
  c.h = h;
  c.haveResult = false;
  rc = ensureImage(c, w, h);
  if (rc) return rc;
  if (!c.basis) XB_HIP(locked_malloc((void**)&c.basis, sizeof(Basis)));
  if (!c.sc) XB_HIP(locked_malloc((void**)&c.sc, sizeof(DevScalars)));
  if (!c.hsc) {
    XB_HIP(locked_host_malloc((void**)&c.hsc, sizeof(DevScalars)));
    XB_HIP(hipHostGetDevicePointer((void**)&c.hscDev, c.hsc, 0));
  }
  {
    int rc2 = ensureSolver(c);
    if (rc2) return rc2;
  }
  // The basis depends only on the parameters: it is built and uploaded when
  // they change (and the mixed form's table when the basis widths do), so a
  // run with the same parameters starts without a host wait.
  static_assert(sizeof(Basis) < (1 << 20), "basis size");
  const bool newBasis =
      !c.basisValid || std::memcmp(&c.basisParams, p, sizeof(xb_params)) != 0;
  const bool newTab = (c.form & XB_FORM_MIXED) &&
                      (!c.mixTab || !c.mixTab16 ||
                       std::memcmp(c.mixBg, c.p.bg, sizeof(c.mixBg)) != 0);
  std::vector<uint16_t> tab, tab16;
  if (newBasis || newTab) {
    Basis hb;
    std::memset(&hb, 0, sizeof(hb));
    buildBasis(c.p, hb);
    if (newBasis)
      XB_HIP(hipMemcpyAsync(c.basis, &hb, sizeof(Basis), hipMemcpyHostToDevice, c.stream));
    if (newTab) {
      mixedTable(hb, tab);
      c.mixScale16 = mixedTable16(hb, tab16);
      if (!c.mixTab) XB_HIP(locked_malloc(&c.mixTab, tab.size() * sizeof(uint16_t)));
      if (!c.mixTab16) XB_HIP(locked_malloc(&c.mixTab16, tab16.size() * sizeof(uint16_t)));
      XB_HIP(hipMemcpyAsync(c.mixTab, tab.data(), tab.size() * sizeof(uint16_t),
                            hipMemcpyHostToDevice, c.stream));
      XB_HIP(hipMemcpyAsync(c.mixTab16, tab16.data(), tab16.size() * sizeof(uint16_t),
                            hipMemcpyHostToDevice, c.stream));
      std::memcpy(c.mixBg, c.p.bg, sizeof(c.mixBg));
    }
    // the host sources (hb, tab) end with this scope
    XB_HIP(hipStreamSynchronize(c.stream));
    if (newBasis) {
      c.basisParams = *p;
      c.basisValid = true;
    }
  }
  return XB_OK;
}

// the mask kernel also clears the run's device scalars (c.sc)
int stageMask(Ctx& c) {
  c.cur = c.stream;
  XB_HIP(launch_mask_input(c));
  return XB_OK;
}

// sss (src/bach.cpp:44-170)
// sss geometry and buffers (src/bach.cpp:44-66, bachUtil.cpp:144-156)
int sssSetup(Ctx& c) {
  const xb_params& a = c.p;
  int fsw = std::min(c.w / a.stampsx, c.h / a.stampsy);
  int nx = a.stampsx, ny = a.stampsy;
  fsw -= a.fKernelWidth;
  fsw -= fsw % 2 == 0 ? 1 : 0;
  if (fsw < a.fSStampWidth) {
    fsw = a.fSStampWidth + a.fKernelWidth;
    fsw -= fsw % 2 == 0 ? 1 : 0;
    nx = c.w / fsw;
    ny = c.h / fsw;
  }
  if (nx <= 0 || ny <= 0 || nx * ny > XB_MAX_STAMPS) {
    c.err = "stamp grid outside 1..1024 stamps";
    return XB_ERR_CAPACITY;
  }
  if ((fsw * fsw + 31) / 32 > 2560) {
    c.err = "stamps wider than 286 px are not supported";
    return XB_ERR_CAPACITY;
  }
  c.fsw = fsw;
  c.stampsx = nx;
  c.stampsy = ny;
  c.nGrid = nx * ny;
  // the substamp scan covers fsw x fsw from each origin (sss.cl:406-410)
  if ((nx - 1) * c.w / nx + fsw > c.w || (ny - 1) * c.h / ny + fsw > c.h) {
    c.err = "stamp grid exceeds the image";
    return XB_ERR_ARG;
  }
  for (int s = 0; s < c.nGrid; s++) {  // bachUtil.cpp:144-156
    int sx = s % nx, sy = s / nx;
    int ox = sx * c.w / nx, oy = sy * c.h / ny;
    int sw = std::min(ox + fsw, c.w) - ox, sh = std::min(oy + fsw, c.h) - oy;
    if (sw * sh < NSAMP) {
      c.err = "Not enough pixels in a stamp";
      return XB_ERR_STAMP_PIXELS;
    }
  }
  c.retried = 0;
  return ensureStamps(c, c.nGrid, fsw);
}

// calcStats T, calcStats S, then both searches (src/sssUtil.cpp:11-14); the
// searches read only their own image's mask bits, so they overlap
// (a one-stream context launches both searches as one dispatch)
int sssIdentify(Ctx& c) {
  XB_HIP(launch_sss_stats(c, 0));
  XB_HIP(launch_sss_stats(c, 1));
  if (c.side == c.stream) {
    XB_HIP(launch_sss_find_pair(c));
  } else {
    int rc = forkBoth(c, [&](int im) -> int {
      XB_HIP(launch_sss_find(c, im));
      return XB_OK;
    });
    if (rc) return rc;
  }
  XB_HIP(launch_sss_compact(c));
  return XB_OK;
}

int sssDone(Ctx& c, const DevScalars& sc) {
  c.nKept[0] = sc.nKept[0];
  c.nKept[1] = sc.nKept[1];
  if (c.nKept[0] == 0 && c.nKept[1] == 0) {
    c.err = "No substamps found";
    return XB_ERR_NO_SUBSTAMPS;
  }
  return XB_OK;
}

int stageSss(Ctx& c) {
  int rc = sssSetup(c);
  if (rc) return rc;
  rc = sssIdentify(c);
  if (rc) return rc;
  // k_sss_compact published the counts and the retry flag to the pinned mirror
  XB_HIP(spinSync(c));
  if (c.hsc->retry) {  // fewer than 10 % of the stamps filled (src/bach.cpp:142-161)
    c.retried = 1;
    XB_HIP(launch_reset_skip(c));
    rc = sssIdentify(c);
    if (rc) return rc;
    XB_HIP(spinSync(c));
  }
  return sssDone(c, *c.hsc);
}

// cmv (src/bach.cpp:172-268)
int stageCmv(Ctx& c) {
  if (c.side == c.stream) {  // one stream: both images' fills as one dispatch per kernel
    XB_HIP(launch_fill_pair(c));
    return XB_OK;
  }
  return forkBoth(c, [&](int im) -> int {
    XB_HIP(launch_fill(c, im, nullptr, nullptr, c.nKept[im]));
    return XB_OK;
  });
}

// cd (src/bach.cpp:270-313, testFit src/cdkscUtil.cpp:4-266)
int stageCd(Ctx& c, int* convTemplate) {
  // the two trial fits are independent: each direction has its own system.  A
  // one-stream context with stamps in both images runs them side by side, as
  // one dispatch per kernel and one batched solve of the two systems.
  const bool pair = c.side == c.stream && c.nKept[0] > 0 && c.nKept[1] > 0;
  int rc = XB_OK;
  if (pair) {
    XB_HIP(launch_cd_begin_pair(c));
    XB_HIP(launch_global_solve_pair(c, c.lu[0], c.cdSol, &c.sc->nTest[0], c.lu[1],
                                    c.cdSol + NSOL, &c.sc->nTest[1]));
    XB_HIP(launch_cd_end_pair(c));
  } else rc = forkBoth(c, [&](int im) -> int {
    if (c.nKept[im] == 0) return XB_OK;
    double* cdSol = c.cdSol + (size_t)im * NSOL;
    XB_HIP(launch_stamp_lu(c, im));
    XB_HIP(launch_cd_select(c, im));
    XB_HIP(launch_global_build(c, im, true));
    XB_HIP(launch_global_solve(c, c.lu[im], cdSol, &c.sc->nTest[im]));
    XB_HIP(launch_calc_sigs(c, im, true, cdSol));
    XB_HIP(launch_cd_finish(c, im));
    return XB_OK;
  });
  if (rc) return rc;
  // k_cd_finish published the merits to the pinned mirror
  XB_HIP(spinSync(c));
  double merit[2];
  for (int im = 0; im < 2; im++) merit[im] = c.nKept[im] == 0 ? 666 : c.hsc->merit[im];
  c.convTemplate = merit[1] > merit[0] ? 1 : 0;
  c.tIdx = c.convTemplate ? 0 : 1;
  if (convTemplate) *convTemplate = c.convTemplate;
  return XB_OK;
}

int ensureScreen(Ctx& c) {
  ScreenBufs& S = c.scr;
  if (S.cap >= c.capStamps) return XB_OK;
  dfree(&S.feat);
  dfree(&S.passCur);
  dfree(&S.label);
  dfree(&S.prob);
  dfree(&S.flag);
  dfree(&S.zPrev);
  dfree(&S.bits);
  dfree(&S.vOrig);
  dfree(&S.vList);
  dfree(&S.vSig);
  dfree(&S.finalRej);
  const size_t np = (size_t)SCR_PASSES * c.capStamps;
  XB_HIP(dalloc(&S.feat, np * XB_MLP_NFEAT));
  XB_HIP(dalloc(&S.passCur, np));
  XB_HIP(dalloc(&S.label, np));
  XB_HIP(dalloc(&S.prob, np));
  XB_HIP(dalloc(&S.flag, (size_t)c.capStamps));
  XB_HIP(dalloc(&S.zPrev, (size_t)c.capStamps));
  XB_HIP(dalloc(&S.bits, (size_t)c.capStamps));
  XB_HIP(dalloc(&S.vOrig, (size_t)c.capStamps));
  XB_HIP(dalloc(&S.vList, (size_t)c.capStamps));
  XB_HIP(dalloc(&S.vSig, (size_t)c.capStamps));
  XB_HIP(dalloc(&S.finalRej, (size_t)c.capStamps * MAXSS));
  S.cap = c.capStamps;
  return XB_OK;
}


constexpr size_t kscOffset(int im) { return offsetof(DevScalars, k) + im * sizeof(KscScalars); }

// One pass of the clip-refit loop that convolves image im (fitKernel +
// checkFitSolution, src/cdkscUtil.cpp:585-684) on c.cur: the global fit, every
// stamp's chi2, the clip check (which moves rejected stamps to their next
// substamp and lists them), the screen; then the pass's check flag (with the
// refill list's length, chi2 and screened counts beside it) is copied to the
// pinned mirror and evK[im] recorded.  The refill of the listed stamps
// (kscRefill) is launched by the caller once it has read that length, before
// the next pass, so its grid covers exactly the listed stamps and a converged
// pass (empty list) launches none.  waitCd (may be null): cd's last read of
// image im's fit data, which the check and the refill must follow.
int kscPass(Ctx& c, int im, int it, bool keep, bool screen, uint16_t nanMark,
            hipEvent_t waitCd) {
  double* sol = c.ksSol + (size_t)im * NSOL;
  XB_HIP(launch_global_build(c, im, false));
  XB_HIP(launch_global_solve(c, c.ksLu[im], sol, nullptr));
  XB_HIP(launch_calc_sigs(c, im, false, sol, nanMark));
  if (waitCd) XB_HIP(hipStreamWaitEvent(c.cur, waitCd, 0));
  const bool rec = keep && it < SCR_PASSES;
  // the check records the substamps of this pass; the screen's flags start
  // cleared at a screened loop's first pass
  // the check (and the screen's merge after it, which has the last word)
  // publishes the pass's scalars to the host mirror
  XB_HIP(launch_ksc_check(c, im, it == 0,
                          rec ? c.scr.passCur + (size_t)it * c.nKept[im] : nullptr,
                          (screen && it == 0) ? c.scr.bits : nullptr, true,
                          rec ? c.scr.label + (size_t)it * c.nKept[im] : nullptr));
  if (rec) XB_HIP(launch_screen(c, im, it, screen));
  XB_HIP(hipEventRecord(c.evK[im], c.cur));
  return XB_OK;
}

// The refill of the last pass of image im's loop (its list's length read
// back with its check flag), on c.cur before the next pass.
int kscRefill(Ctx& c, int im) {
  const int n = c.hsc->k[im].nList;
  if (n > 0) XB_HIP(launch_fill(c, im, c.st[im].listK, &c.sc->k[im].nList, n));
  return XB_OK;
}

constexpr int KSC_MAX_PASSES = 3 * XB_MAX_STAMPS + 2;

// One chi2 re-check round of the screened substamps (stageKsc): each stamp
// with screened substamps left probes its lowest one (filled, then its chi2
// under the converged fit), the reference's reject rule decides, and stamps
// whose probe is rejected get their own substamp's fit data back.  *probed:
// stamps probed; *kept: probes the chi2 test kept (restored).
int screenRecheck(Ctx& c, int im, int* probed, int* kept) {
  const double* sol = c.ksSol + (size_t)im * NSOL;
  KscScalars* ks = &c.sc->k[im];
  XB_HIP(launch_screen_verify_select(c, im));
  XB_HIP(launch_fill(c, im, c.scr.vList, &ks->nVerify, c.nKept[im]));
  XB_HIP(launch_calc_sigs_list(c, im, c.scr.vList, &ks->nVerify, sol, c.scr.vSig));
  XB_HIP(launch_screen_verify_decide(c, im));
  XB_HIP(launch_fill(c, im, c.st[im].listK, &ks->nList, c.nKept[im]));
  XB_HIP(readBack(c, kscOffset(im) + offsetof(KscScalars, nVerify), 2 * sizeof(int)));
  *probed = c.hsc->k[im].nVerify;
  *kept = c.hsc->k[im].nRestoredRound;
  return XB_OK;
}

// ksc (src/cdkscUtil.cpp:585-684): the loop of the chosen direction, after cd.
// With the learned screen, the loop that converged is followed by the chi2
// re-check of every substamp the screen replaced: re-check rounds run until
// none is left or one is kept; a kept one is restored and the loop runs again
// without the screen, then the re-check resumes on what is left.  Every
// screened substamp is thus judged by the true chi2 test under a converged
// fit, and the fit finally uses only substamps that passed it.
int stageKsc(Ctx& c) {
  const int im = c.tIdx;
  const bool screen = (c.form & XB_FORM_SCREEN) != 0;
  const bool keep = screen || (c.form & XB_FORM_RECORD) != 0;
  c.scrSamples = 0;
  c.nScreened = 0;
  c.nRestored = 0;
  c.nConfirmed = 0;
  if (screen && !c.scr.haveModel) {
    c.err = "XB_FORM_SCREEN needs a model (xb_screen_load)";
    return XB_ERR_ARG;
  }
  if (keep) {
    int rc = ensureScreen(c);
    if (rc) return rc;
  }
  c.cur = c.stream;
  bool screenOn = screen;
  int it = 0;
  for (;;) {
    int rc = kscPass(c, im, it, keep, screenOn, XB_MASK_NAN_PIXEL, nullptr);
    if (rc) return rc;
    it++;
    XB_HIP(spinEvent(c.evK[im]));
    if (it > KSC_MAX_PASSES) {
      c.err = "ksc loop did not terminate";
      return XB_ERR_HIP;
    }
    if (c.hsc->k[im].check) {
      rc = kscRefill(c, im);
      if (rc) return rc;
      continue;
    }
    // nothing screened (the count is cumulative): no substamp to re-check
    if (!screen || c.hsc->k[im].nScreened == 0) break;
    bool restored = false;
    for (;;) {
      int probed = 0, kept = 0;
      rc = screenRecheck(c, im, &probed, &kept);
      if (rc) return rc;
      if (probed == 0) break;
      if (kept > 0) {
        restored = true;
        break;
      }
    }
    if (!restored) break;
    screenOn = false;  // the restored substamps are judged by plain passes
  }
  c.kscIterations = it;
  if (keep) {
    const int nRec = std::min(it, SCR_PASSES);
    if (!screen) {
      // recording: every substamp a recorded pass used that the converged fit
      // does not is probed (filled, chi2 under that fit, the reference's
      // reject rule with the last clip), then its stamp gets its own
      // substamp's fit data back; the labels follow those verdicts
      const double* sol = c.ksSol + (size_t)im * NSOL;
      KscScalars* ks = &c.sc->k[im];
      XB_HIP(launch_label_bits(c, im, nRec));
      for (;;) {
        XB_HIP(launch_screen_verify_select(c, im));
        XB_HIP(launch_fill(c, im, c.scr.vList, &ks->nVerify, c.nKept[im]));
        XB_HIP(launch_calc_sigs_list(c, im, c.scr.vList, &ks->nVerify, sol, c.scr.vSig));
        XB_HIP(launch_label_decide(c, im));
        XB_HIP(launch_fill(c, im, c.st[im].listK, &ks->nList, c.nKept[im]));
        XB_HIP(readBack(c, kscOffset(im) + offsetof(KscScalars, nVerify), sizeof(int)));
        if (c.hsc->k[im].nVerify == 0) break;
      }
    }
    XB_HIP(launch_screen_labels(c, im, nRec, !screen));
    // the counts: read back, unless the screen replaced nothing (all zero)
    if (!screen || c.hsc->k[im].nScreened > 0) {
      XB_HIP(readBack(c, kscOffset(im), sizeof(KscScalars)));
      if (screen) {
        c.nScreened = c.hsc->k[im].nScreened;
        c.nRestored = c.hsc->k[im].nRestored;
        c.nConfirmed = c.hsc->k[im].nConfirmed;
      }
    } else {
      c.nScreened = c.nRestored = c.nConfirmed = 0;
    }
    c.scrSamples = nRec * c.nKept[im];
  }
  return XB_OK;
}

// Whether cd and both directions' clip-refit loops can run at once
// (stageCdKsc): a context with its own streams for them, and no screen or
// sample recording (their buffers serve one loop).
bool canSpeculate(const Ctx& c) {
  return c.side != c.stream && c.kstream[0] && c.kstream[1] &&
         (c.form & (XB_FORM_SCREEN | XB_FORM_RECORD)) == 0;
}

// cd and ksc of the whole pipeline at once.  The reference fits both
// directions in cd, then runs the clip-refit loop of the one it chose
// (src/bach.cpp:270-323).  A direction's loop needs only cmv's fit data, not
// the choice, so here each direction's loop starts at the same time as cd, on
// a stream of its own, and the loop of the direction cd rejects is abandoned
// when the merits arrive.  The results are those of cd followed by the chosen
// loop, bit for bit:
//   * the two loops and cd keep separate systems, solutions, chi2 arrays,
//     refill lists and scalars;
//   * a loop's first check (which moves rejected stamps on) and refill wait
//     for cd's last read of that image's substamps and fit data;
//   * the loops' NaN marks use a spare mask bit per direction; the
//     convolution keeps the chosen one's as NAN_PIXEL;
//   * the abandoned loop's image gets its cmv fit data back (the stamps it
//     moved on are refilled on their first substamp) on its own stream, beside
//     the convolution; runStages waits for it before returning.
int stageCdKsc(Ctx& c) {
  c.scrSamples = 0;
  c.nScreened = 0;
  XB_HIP(hipEventRecord(c.evFork, c.stream));
  XB_HIP(hipStreamWaitEvent(c.side, c.evFork, 0));
  for (int d = 0; d < 2; d++) XB_HIP(hipStreamWaitEvent(c.kstream[d], c.evFork, 0));
  // the two directions' trial fits (as stageCd)
  for (int im = 0; im < 2; im++) {
    c.cur = im == 0 ? c.stream : c.side;
    if (c.nKept[im] > 0) {
      double* cdSol = c.cdSol + (size_t)im * NSOL;
      XB_HIP(launch_stamp_lu(c, im));
      XB_HIP(launch_cd_select(c, im));
      XB_HIP(launch_global_build(c, im, true));
      XB_HIP(launch_global_solve(c, c.lu[im], cdSol, &c.sc->nTest[im]));
      XB_HIP(launch_calc_sigs(c, im, true, cdSol));
    }
    XB_HIP(hipEventRecord(c.evCdSig[im], c.cur));
    if (c.nKept[im] > 0) XB_HIP(launch_cd_finish(c, im));
  }
  XB_HIP(hipEventRecord(c.evJoin, c.side));
  XB_HIP(hipStreamWaitEvent(c.stream, c.evJoin, 0));
  c.cur = c.stream;
  {
    const size_t off = offsetof(DevScalars, merit);
    XB_HIP(hipMemcpyAsync(reinterpret_cast<char*>(c.hsc) + off,
                          reinterpret_cast<const char*>(c.sc) + off, sizeof(double) * 2,
                          hipMemcpyDeviceToHost, c.stream));
    XB_HIP(hipEventRecord(c.evMerit, c.stream));
  }
  int chosen = -1;
  auto decide = [&]() {
    double merit[2];
    for (int im = 0; im < 2; im++) merit[im] = c.nKept[im] == 0 ? 666 : c.hsc->merit[im];
    c.convTemplate = merit[1] > merit[0] ? 1 : 0;
    c.tIdx = c.convTemplate ? 0 : 1;
    chosen = c.tIdx;
  };
  bool active[2] = {true, true};
  int it[2] = {0, 0};
  for (int d = 0; d < 2; d++) {
    c.cur = c.kstream[d];
    int rc = kscPass(c, d, 0, false, false, mask_pend(d), c.evCdSig[d]);
    if (rc) return rc;
  }
  for (;;) {
    for (int d = 0; d < 2; d++) {
      if (!active[d]) continue;
      XB_HIP(spinEvent(c.evK[d]));
      it[d]++;
      if (!c.hsc->k[d].check) active[d] = false;
    }
    if (chosen < 0) {
      const hipError_t q = hipEventQuery(c.evMerit);
      if (q == hipSuccess) decide();
      else if (q != hipErrorNotReady) XB_HIP(q);
    }
    if (chosen < 0 && !active[0] && !active[1]) {
      XB_HIP(spinEvent(c.evMerit));
      decide();
    }
    if (chosen >= 0) {
      active[1 - chosen] = false;
      if (!active[chosen]) break;
    }
    for (int d = 0; d < 2; d++) {
      if (!active[d]) continue;
      if (it[d] > KSC_MAX_PASSES) {
        c.cur = c.stream;
        c.err = "ksc loop did not terminate";
        return XB_ERR_HIP;
      }
      c.cur = c.kstream[d];
      int rc = kscRefill(c, d);
      if (!rc) rc = kscPass(c, d, it[d], false, false, mask_pend(d), nullptr);
      if (rc) {
        c.cur = c.stream;
        return rc;
      }
    }
  }
  c.kscIterations = it[chosen];
  // both loops' passes end before the primary stream goes on (the convolution
  // rewrites the mask the abandoned loop's chi2 pass may still mark)
  for (int d = 0; d < 2; d++) {
    XB_HIP(hipEventRecord(c.evKEnd[d], c.kstream[d]));
    XB_HIP(hipStreamWaitEvent(c.stream, c.evKEnd[d], 0));
  }
  const int o = 1 - chosen;
  c.cur = c.kstream[o];
  {
    const hipError_t e = launch_restore_fill(c, o);
    c.cur = c.stream;
    XB_HIP(e);
  }
  c.specRestore = o;
  return XB_OK;
}

// Waits for every stream of the context (after an error inside stageCdKsc,
// its loops may still have work queued).
int syncAll(Ctx& c) {
  c.cur = c.stream;
  c.specRestore = -1;
  XB_HIP(hipStreamSynchronize(c.stream));
  if (c.side != c.stream) XB_HIP(hipStreamSynchronize(c.side));
  for (int d = 0; d < 2; d++)
    if (c.kstream[d]) XB_HIP(hipStreamSynchronize(c.kstream[d]));
  return XB_OK;
}

// The restore of stageCdKsc, joined into the primary stream.
int joinRestore(Ctx& c) {
  if (c.specRestore < 0) return XB_OK;
  const int o = c.specRestore;
  c.specRestore = -1;
  XB_HIP(hipEventRecord(c.evKEnd[o], c.kstream[o]));
  XB_HIP(hipStreamWaitEvent(c.stream, c.evKEnd[o], 0));
  return XB_OK;
}

int stageConv(Ctx& c, double* kernSum, bool readSum) {
  XB_HIP(launch_conv(c));
  if (readSum) {
    XB_HIP(readBack(c, offsetof(DevScalars, kernSum), sizeof(double)));
    c.kernSum = c.hsc->kernSum;
    if (kernSum) *kernSum = c.kernSum;
  }
  return XB_OK;
}

// the difference image is formed in k_conv's epilogue (the difference of each
// interior pixel from its unrounded FP64 convolution, 1e-30 on the border and
// on BAD_OUTPUT); sub only publishes it
int stageSub(Ctx& c) {
  c.haveResult = true;
  return XB_OK;
}

void freeAll(Ctx& c) {
  {
    std::unique_lock<std::shared_mutex> g(alloc_lock());
    for (auto& sg : c.solveGraphs) (void)hipGraphExecDestroy(sg.exec);
  }
  c.solveGraphs.clear();
  dfree(&c.img[0]);
  dfree(&c.img[1]);
  dfree(&c.mask);
  dfree(&c.conv);
  dfree(&c.diff);
  for (int im = 0; im < 2; im++) {
    DevStamps& s = c.st[im];
    dfree(&s.sky); dfree(&s.fwhm); dfree(&s.found); dfree(&s.count);
    dfree(&s.ssx); dfree(&s.ssy); dfree(&s.ssv); dfree(&s.kept); dfree(&s.cur);
    dfree(&s.W); dfree(&s.Wc); dfree(&s.G); dfree(&s.Bc); dfree(&s.Br); dfree(&s.sig);
    dfree(&s.ksum); dfree(&s.list); dfree(&s.sigK); dfree(&s.listK);
  }
  dfree(&c.sc);
  if (c.hsc) {
    (void)locked_host_free(c.hsc);
    c.hsc = nullptr;
  }
  dfree(&c.basis);
  c.basisValid = false;
  dfree(&c.ran1);
  dfree(&c.goodPix);
  for (int i = 0; i < 4; i++) {
    LuSet& L = i < 2 ? c.lu[i] : c.ksLu[i - 2];
    dfree(&L.M); dfree(&L.rhs); dfree(&L.M0); dfree(&L.rhs0);
    dfree(&L.luIndex); dfree(&L.luVV); dfree(&L.luState); dfree(&L.rr);
    dfree(&L.F);
  }
  dfree(&c.sol); dfree(&c.cdSol); dfree(&c.ksSol);
  dfree(&c.tileK64); dfree(&c.tileAks);
  dfree(&c.scr.feat); dfree(&c.scr.passCur); dfree(&c.scr.label); dfree(&c.scr.prob);
  dfree(&c.scr.flag); dfree(&c.scr.zPrev); dfree(&c.scr.model);
  dfree(&c.scr.bits); dfree(&c.scr.vOrig); dfree(&c.scr.vList); dfree(&c.scr.vSig);
  dfree(&c.scr.finalRej);
  c.scr.cap = 0;
  c.scr.haveModel = false;
  if (c.mixTab) (void)locked_free(c.mixTab);
  c.mixTab = nullptr;
  if (c.mixTab16) (void)locked_free(c.mixTab16);
  c.mixTab16 = nullptr;
}

int setDev(Ctx& c) {
  XB_HIP(hipSetDevice(c.device));
  return XB_OK;
}

hipEvent_t takeEvent(Ctx& c) {
  if (!c.evPool.empty()) {
    hipEvent_t e = c.evPool.back();
    c.evPool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Resolve recorded event pairs into the per-slot totals.
int drainProfile(Ctx& c) {
  if (c.evPending.empty()) return XB_OK;
  XB_HIP(hipStreamSynchronize(c.stream));
  for (auto& p : c.evPending) {
    float ms = 0.0f;
    XB_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    c.profMs[p.slot] += ms;
    c.profCnt[p.slot] += 1;
    c.evPool.push_back(p.a);
    c.evPool.push_back(p.b);
  }
  c.evPending.clear();
  return XB_OK;
}

}  // namespace

namespace xb {
// ---- the pipeline in pieces, for a driver that interleaves several contexts
// on one host thread (the batch runtime's lockstep mode).  Each piece
// enqueues on the context's primary stream; a read-back is copied to the
// pinned scalar mirror and signalled by c.evSync (c.evK[image] for a
// clip-refit pass), which the driver waits for before the next piece.
static hipError_t copyScalarsAsync(Ctx& c, size_t off, size_t bytes, hipEvent_t ev) {
  hipError_t e = hipMemcpyAsync(reinterpret_cast<char*>(c.hsc) + off,
                                reinterpret_cast<const char*>(c.sc) + off, bytes,
                                hipMemcpyDeviceToHost, c.stream);
  return e != hipSuccess ? e : hipEventRecord(ev, c.stream);
}

int ls_sss_begin(Ctx& c) {
  c.cur = c.stream;
  int rc = sssSetup(c);
  if (rc) return rc;
  rc = sssIdentify(c);
  if (rc) return rc;
  XB_HIP(copyScalarsAsync(c, 0, sizeof(DevScalars), c.evSync));
  return XB_OK;
}

int ls_sss_next(Ctx& c, bool* more) {
  *more = false;
  XB_HIP(spinEvent(c.evSync));
  if (!c.retried && c.hsc->retry) {
    c.retried = 1;
    XB_HIP(launch_reset_skip(c));
    int rc = sssIdentify(c);
    if (rc) return rc;
    XB_HIP(copyScalarsAsync(c, 0, sizeof(DevScalars), c.evSync));
    *more = true;
    return XB_OK;
  }
  const DevScalars sc = *c.hsc;
  return sssDone(c, sc);
}

int ls_cmv(Ctx& c) { return stageCmv(c); }

int ls_cd_build(Ctx& c, LuBatch& B) {
  for (int im = 0; im < 2; im++) {
    if (c.nKept[im] == 0) continue;
    XB_HIP(launch_stamp_lu(c, im));
    XB_HIP(launch_cd_select(c, im));
    XB_HIP(launch_global_build(c, im, true));
    B.s[B.n++] = lu_sys(c.lu[im], c.cdSol + (size_t)im * NSOL, &c.sc->nTest[im]);
  }
  return XB_OK;
}

int ls_cd_post(Ctx& c) {
  for (int im = 0; im < 2; im++) {
    if (c.nKept[im] == 0) continue;
    XB_HIP(launch_calc_sigs(c, im, true, c.cdSol + (size_t)im * NSOL));
    XB_HIP(launch_cd_finish(c, im));
  }
  XB_HIP(copyScalarsAsync(c, offsetof(DevScalars, merit), 2 * sizeof(double), c.evSync));
  return XB_OK;
}

int ls_cd_decide(Ctx& c) {
  XB_HIP(spinEvent(c.evSync));
  double merit[2];
  for (int im = 0; im < 2; im++) merit[im] = c.nKept[im] == 0 ? 666 : c.hsc->merit[im];
  c.convTemplate = merit[1] > merit[0] ? 1 : 0;
  c.tIdx = c.convTemplate ? 0 : 1;
  c.scrSamples = c.nScreened = c.nRestored = c.nConfirmed = 0;
  return XB_OK;
}

int ls_ksc_build(Ctx& c, LuBatch& B) {
  const int im = c.tIdx;
  XB_HIP(launch_global_build(c, im, false));
  B.s[B.n++] = lu_sys(c.ksLu[im], c.ksSol + (size_t)im * NSOL, nullptr);
  return XB_OK;
}

int ls_ksc_post(Ctx& c, int it) {
  const int im = c.tIdx;
  const double* sol = c.ksSol + (size_t)im * NSOL;
  XB_HIP(launch_calc_sigs(c, im, false, sol));
  XB_HIP(launch_ksc_check(c, im, it == 0));
  XB_HIP(launch_fill(c, im, c.st[im].listK, &c.sc->k[im].nList, c.nKept[im]));
  XB_HIP(copyScalarsAsync(c, kscOffset(im) + offsetof(KscScalars, check), sizeof(int),
                          c.evK[im]));
  return XB_OK;
}

int ls_ksc_check(Ctx& c, int it, bool* more) {
  const int im = c.tIdx;
  XB_HIP(spinEvent(c.evK[im]));
  *more = c.hsc->k[im].check != 0;
  c.kscIterations = it;
  if (*more && it > KSC_MAX_PASSES) {
    c.err = "ksc loop did not terminate";
    return XB_ERR_HIP;
  }
  return XB_OK;
}

int ls_finish(Ctx& c) {
  int rc = stageConv(c, nullptr, false);
  if (rc) return rc;
  return stageSub(c);
}

void prof_begin(Ctx& c, int slot) {
  if (!c.profOn || slot < 0 || slot >= XB_PROF_N) return;
  hipEvent_t e = takeEvent(c);
  if (!e) return;
  if (hipEventRecord(e, c.cur) != hipSuccess) {
    c.evPool.push_back(e);
    return;
  }
  c.evOpen[slot] = e;
}

void prof_end(Ctx& c, int slot) {
  if (!c.profOn || slot < 0 || slot >= XB_PROF_N || !c.evOpen[slot]) return;
  hipEvent_t e = takeEvent(c);
  if (!e) return;
  if (hipEventRecord(e, c.cur) != hipSuccess) {
    c.evPool.push_back(e);
    return;
  }
  c.evPending.push_back({slot, c.evOpen[slot], e});
  c.evOpen[slot] = nullptr;
}
}  // namespace xb

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

void xb_params_default(xb_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->stampsx = p->stampsy = 10;
  p->threshLow = 0.0;
  p->threshHigh = 25000.0;
  p->threshKernFit = 20.0;
  p->sigKernFit = 2.0;
  p->sigClipAlpha = 3.0;
  p->iqRange = 1.35;
  p->maxKSStamps = 3;
  p->hSStampWidth = 15;
  p->fSStampWidth = 31;
  p->hKernelWidth = 10;
  p->fKernelWidth = 21;
  p->inSpreadMaskFactor = 1.0;
  p->normalizeTemplate = 1;
  p->backgroundOrder = 1;
  p->kernelOrder = 2;
  p->ngauss = 3;
  p->dg[0] = 6;
  p->dg[1] = 4;
  p->dg[2] = 2;
  p->bg[0] = (float)(1.0 / (2.0 * 0.7 * 0.7));
  p->bg[1] = (float)(1.0 / (2.0 * 1.5 * 1.5));
  p->bg[2] = (float)(1.0 / (2.0 * 3.0 * 3.0));
  p->verbose = 0;
}

int xb_ctx_create(int device, int flags, xb_ctx** out) {
  if (!out) return XB_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return XB_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return XB_ERR_NO_DEVICE;
  xb_ctx* x = new xb_ctx();
  Ctx& c = x->c;
  c.device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete x;
    return XB_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    std::snprintf(c.name, sizeof(c.name), "%s (%s)", prop.name, prop.gcnArchName);
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      delete x;
      return XB_ERR_NOT_BUILT;
    }
  }
  // two streams of the default priority, and the events that fork and join them
  // (one stream serves both roles under XB_CTX_ONE_STREAM)
  const bool one = (flags & XB_CTX_ONE_STREAM) != 0;
  if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess ||
      (one ? (c.side = c.stream, false)
           : hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking) != hipSuccess) ||
      hipEventCreateWithFlags(&c.evFork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c.evJoin, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c.evSync, hipEventDisableTiming) != hipSuccess) {
    if (c.evFork) (void)hipEventDestroy(c.evFork);
    if (c.evJoin) (void)hipEventDestroy(c.evJoin);
    if (c.evSync) (void)hipEventDestroy(c.evSync);
    if (c.side && c.side != c.stream) (void)hipStreamDestroy(c.side);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete x;
    return XB_ERR_HIP;
  }
  c.cur = c.stream;
  c.eagerSolve = (flags & XB_CTX_EAGER) != 0;
  // the clip-refit loops' streams (speculative cd + ksc, multi-stream contexts
  // only) and the events every context's loop uses
  bool evOk = true;
  for (int d = 0; d < 2; d++) {
    evOk = evOk && hipEventCreateWithFlags(&c.evK[d], hipEventDisableTiming) == hipSuccess &&
           hipEventCreateWithFlags(&c.evKEnd[d], hipEventDisableTiming) == hipSuccess &&
           hipEventCreateWithFlags(&c.evCdSig[d], hipEventDisableTiming) == hipSuccess;
    if (evOk && !one)
      evOk = hipStreamCreateWithFlags(&c.kstream[d], hipStreamNonBlocking) == hipSuccess;
  }
  evOk = evOk && hipEventCreateWithFlags(&c.evMerit, hipEventDisableTiming) == hipSuccess;
  if (!evOk) {
    xb_ctx_destroy(x);
    return XB_ERR_HIP;
  }
  // the allocation-free gradient's scratch: its address on this device is
  // looked up here, once per device, so that entry point makes no HIP call
  // beyond its launches
  if (scratch_address(device, &c.mlpScratch) != hipSuccess) c.mlpScratch = nullptr;
  *out = x;
  return XB_OK;
}

int xb_ctx_destroy(xb_ctx* x) {
  if (!x) return XB_OK;
  Ctx& c = x->c;
  (void)hipSetDevice(c.device);
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  if (c.side && c.side != c.stream) (void)hipStreamSynchronize(c.side);
  for (int d = 0; d < 2; d++)
    if (c.kstream[d]) (void)hipStreamSynchronize(c.kstream[d]);
  for (auto& p : c.evPending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c.evOpen)
    if (e) (void)hipEventDestroy(e);
  for (auto e : c.evPool) (void)hipEventDestroy(e);
  freeAll(c);
  if (c.evFork) (void)hipEventDestroy(c.evFork);
  if (c.evJoin) (void)hipEventDestroy(c.evJoin);
  if (c.evSync) (void)hipEventDestroy(c.evSync);
  if (c.evMerit) (void)hipEventDestroy(c.evMerit);
  for (int d = 0; d < 2; d++) {
    if (c.evK[d]) (void)hipEventDestroy(c.evK[d]);
    if (c.evKEnd[d]) (void)hipEventDestroy(c.evKEnd[d]);
    if (c.evCdSig[d]) (void)hipEventDestroy(c.evCdSig[d]);
    if (c.kstream[d]) (void)hipStreamDestroy(c.kstream[d]);
  }
  if (c.side && c.side != c.stream) (void)hipStreamDestroy(c.side);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete x;
  return XB_OK;
}

const char* xb_last_error(xb_ctx* x) { return x ? x->c.err.c_str() : "no context"; }

int xb_device_name(xb_ctx* x, char* buf, size_t len) {
  if (!x || !buf || len == 0) return XB_ERR_ARG;
  std::snprintf(buf, len, "%s", x->c.name);
  return XB_OK;
}

int xb_init(xb_ctx* x, const xb_params* p, int w, int h, const float* tmpl, const float* sci) {
  if (!x || !tmpl || !sci) return XB_ERR_ARG;
  Ctx& c = x->c;
  int rc = setDev(c);
  if (rc) return rc;
  rc = stageInitCommon(c, p, w, h);
  if (rc) return rc;
  size_t n = (size_t)w * h;
  XB_HIP(hipMemcpyAsync(c.img[0], tmpl, n * sizeof(float), hipMemcpyHostToDevice, c.stream));
  XB_HIP(hipMemcpyAsync(c.img[1], sci, n * sizeof(float), hipMemcpyHostToDevice, c.stream));
  // the caller's arrays are copied before this returns (they may change next)
  XB_HIP(spinSync(c));
  return stageMask(c);
}

int xb_init_device(xb_ctx* x, const xb_params* p, int w, int h, const void* d_tmpl,
                   const void* d_sci) {
  if (!x || !d_tmpl || !d_sci) return XB_ERR_ARG;
  Ctx& c = x->c;
  int rc = setDev(c);
  if (rc) return rc;
  rc = stageInitCommon(c, p, w, h);
  if (rc) return rc;
  size_t n = (size_t)w * h;
  XB_HIP(hipMemcpyAsync(c.img[0], d_tmpl, n * sizeof(float), hipMemcpyDeviceToDevice, c.stream));
  XB_HIP(hipMemcpyAsync(c.img[1], d_sci, n * sizeof(float), hipMemcpyDeviceToDevice, c.stream));
  return stageMask(c);
}

int xb_set_form(xb_ctx* x, int form) {
  if (!x || (form & ~(XB_FORM_MIXED | XB_FORM_SCREEN | XB_FORM_RECORD | XB_FORM_FP16 |
                      XB_FORM_REFINE | XB_FORM_LDLT)))
    return XB_ERR_ARG;
  x->c.form = form;
  return XB_OK;
}

int xb_set_shapes(xb_ctx* x, const xb_shapes* s) {
  if (!x) return XB_ERR_ARG;
  const xb_shapes v = s ? *s : default_shapes();
  if (!shapes_valid(v)) return XB_ERR_ARG;
  x->c.shapes = v;
  x->c.shapes.pad = 0;
  return XB_OK;
}

int xb_get_shapes(xb_ctx* x, xb_shapes* s) {
  if (!x || !s) return XB_ERR_ARG;
  *s = x->c.shapes;
  return XB_OK;
}

int xb_screen_load(xb_ctx* x, const float* model) {
  if (!x || !model) return XB_ERR_ARG;
  Ctx& c = x->c;
  for (int i = 0; i < XB_MLP_NFEAT; i++)
    if (!(model[XB_MLP_NFEAT + i] > 0.0f)) {
      c.err = "screen model: feature sd must be positive";
      return XB_ERR_ARG;
    }
  XB_HIP(hipSetDevice(c.device));
  if (!c.scr.model) XB_HIP(locked_malloc((void**)&c.scr.model, sizeof(float) * XB_MLP_NPARAM));
  XB_HIP(hipMemcpy(c.scr.model, model, sizeof(float) * XB_MLP_NPARAM, hipMemcpyHostToDevice));
  c.scr.haveModel = true;
  return XB_OK;
}

int xb_screen_samples(xb_ctx* x, float* feat, int* label, float* prob, int cap, int* n,
                      int* screened) {
  if (!x || cap < 0) return XB_ERR_ARG;
  Ctx& c = x->c;
  const int N = c.scrSamples;
  if (n) *n = N;
  if (screened) *screened = c.nScreened;
  if (N == 0 || (!feat && !label && !prob)) return XB_OK;  // size query
  if (cap < N) {
    c.err = "sample buffer too small";
    return XB_ERR_CAPACITY;
  }
  XB_HIP(hipStreamSynchronize(c.stream));
  if (feat) XB_HIP(hipMemcpy(feat, c.scr.feat, sizeof(float) * N * XB_MLP_NFEAT,
                             hipMemcpyDeviceToHost));
  if (label) XB_HIP(hipMemcpy(label, c.scr.label, sizeof(int) * N, hipMemcpyDeviceToHost));
  if (prob) {
    if (c.form & XB_FORM_SCREEN) {
      XB_HIP(hipMemcpy(prob, c.scr.prob, sizeof(float) * N, hipMemcpyDeviceToHost));
    } else {
      for (int i = 0; i < N; i++) prob[i] = -1.0f;
    }
  }
  return XB_OK;
}

int xb_sss(xb_ctx* x) { return x ? stageSss(x->c) : XB_ERR_ARG; }
int xb_cmv(xb_ctx* x) { return x ? stageCmv(x->c) : XB_ERR_ARG; }

int xb_sync(xb_ctx* x) {
  if (!x) return XB_ERR_ARG;
  Ctx& c = x->c;
  int rc = setDev(c);
  if (rc) return rc;
  return syncAll(c);
}
int xb_cd(xb_ctx* x, int* convTemplate) { return x ? stageCd(x->c, convTemplate) : XB_ERR_ARG; }
int xb_ksc(xb_ctx* x) { return x ? stageKsc(x->c) : XB_ERR_ARG; }
int xb_conv(xb_ctx* x, double* kernSum) { return x ? stageConv(x->c, kernSum, true) : XB_ERR_ARG; }
int xb_sub(xb_ctx* x) {
  if (!x) return XB_ERR_ARG;
  Ctx& c = x->c;
  int rc = stageSub(c);
  if (rc) return rc;
  XB_HIP(hipStreamSynchronize(c.stream));
  return XB_OK;
}

static int getImage(xb_ctx* x, const void* d, void* out, size_t elem) {
  if (!x || !out) return XB_ERR_ARG;
  Ctx& c = x->c;
  if (!c.haveResult) {
    c.err = "no result yet";
    return XB_ERR_ARG;
  }
  XB_HIP(hipMemcpyAsync(out, d, (size_t)c.w * c.h * elem, hipMemcpyDeviceToHost, c.stream));
  XB_HIP(hipStreamSynchronize(c.stream));
  return XB_OK;
}

int xb_get_conv(xb_ctx* x, float* out) { return x ? getImage(x, x->c.conv, out, 4) : XB_ERR_ARG; }
int xb_get_diff(xb_ctx* x, float* out) { return x ? getImage(x, x->c.diff, out, 4) : XB_ERR_ARG; }
int xb_get_mask(xb_ctx* x, uint16_t* out) { return x ? getImage(x, x->c.mask, out, 2) : XB_ERR_ARG; }

int xb_lu_solve(xb_ctx* x, const double* a, const double* b, double* out) {
  if (!x || !a || !b || !out) return XB_ERR_ARG;
  Ctx& c = x->c;
  int rc = setDev(c);
  if (rc) return rc;
  {
    int rc2 = ensureSolver(c);
    if (rc2) return rc2;
  }
  const size_t mb = sizeof(double) * MATN * MATN, vb = sizeof(double) * NSOL;
  LuSet& L = c.lu[0];
  XB_HIP(hipMemcpyAsync(L.M, a, mb, hipMemcpyHostToDevice, c.stream));
  XB_HIP(hipMemcpyAsync(L.M0, a, mb, hipMemcpyHostToDevice, c.stream));
  XB_HIP(hipMemcpyAsync(L.rhs, b, vb, hipMemcpyHostToDevice, c.stream));
  XB_HIP(hipMemcpyAsync(L.rhs0, b, vb, hipMemcpyHostToDevice, c.stream));
  XB_HIP(launch_global_solve(c, L, c.sol, nullptr));
  XB_HIP(hipMemcpyAsync(out, c.sol, vb, hipMemcpyDeviceToHost, c.stream));
  XB_HIP(hipStreamSynchronize(c.stream));
  return XB_OK;
}

int xb_buf_alloc(xb_ctx* x, size_t bytes, void** d_out) {
  if (!x || !d_out) return XB_ERR_ARG;
  Ctx& c = x->c;
  *d_out = nullptr;
  int rc = setDev(c);
  if (rc) return rc;
  XB_HIP(locked_malloc(d_out, bytes ? bytes : 1));
  return XB_OK;
}

int xb_buf_free(xb_ctx* x, void* d) {
  if (!x) return XB_ERR_ARG;
  Ctx& c = x->c;
  if (!d) return XB_OK;
  XB_HIP(hipStreamSynchronize(c.stream));
  XB_HIP(locked_free(d));
  return XB_OK;
}

int xb_buf_upload(xb_ctx* x, void* d, const void* host, size_t bytes) {
  if (!x || !d || !host) return XB_ERR_ARG;
  Ctx& c = x->c;
  XB_HIP(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, c.stream));
  XB_HIP(hipStreamSynchronize(c.stream));
  return XB_OK;
}

int xb_buf_download(xb_ctx* x, void* host, const void* d, size_t bytes) {
  if (!x || !d || !host) return XB_ERR_ARG;
  Ctx& c = x->c;
  XB_HIP(hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, c.stream));
  XB_HIP(hipStreamSynchronize(c.stream));
  return XB_OK;
}

int xb_profile_enable(xb_ctx* x, int enable) {
  if (!x) return XB_ERR_ARG;
  Ctx& c = x->c;
  int rc = drainProfile(c);
  if (rc) return rc;
  c.profOn = enable != 0;
  for (int i = 0; i < XB_PROF_N; i++) {
    c.profMs[i] = 0.0;
    c.profCnt[i] = 0;
  }
  return XB_OK;
}

int xb_profile_get(xb_ctx* x, double* ms, int* counts, int n) {
  if (!x) return XB_ERR_ARG;
  Ctx& c = x->c;
  int rc = drainProfile(c);
  if (rc) return rc;
  for (int i = 0; i < n && i < XB_PROF_N; i++) {
    if (ms) ms[i] = c.profMs[i];
    if (counts) counts[i] = c.profCnt[i];
  }
  return XB_OK;
}

int xb_get_stamp_fit(xb_ctx* x, int im, int k, double* W, double* G, double* Bc, double* Br,
                     int* ssxy) {
  if (!x || im < 0 || im > 1) return XB_ERR_ARG;
  Ctx& c = x->c;
  if (k < 0 || k >= c.nKept[im]) {
    c.err = "no such kept stamp";
    return XB_ERR_ARG;
  }
  XB_HIP(hipStreamSynchronize(c.stream));
  const DevStamps& S = c.st[im];
  if (W) XB_HIP(hipMemcpy(W, S.W + (size_t)k * WROWS * NPIX, sizeof(double) * WROWS * NPIX,
                          hipMemcpyDeviceToHost));
  if (G) XB_HIP(hipMemcpy(G, S.G + (size_t)k * WROWS * WROWS, sizeof(double) * WROWS * WROWS,
                          hipMemcpyDeviceToHost));
  if (Bc) XB_HIP(hipMemcpy(Bc, S.Bc + (size_t)k * WROWS, sizeof(double) * WROWS,
                           hipMemcpyDeviceToHost));
  if (Br) XB_HIP(hipMemcpy(Br, S.Br + (size_t)k * NBG, sizeof(double) * NBG,
                           hipMemcpyDeviceToHost));
  if (ssxy) {
    int g = 0, cur = 0, cnt = 0;
    XB_HIP(hipMemcpy(&g, S.kept + k, sizeof(int), hipMemcpyDeviceToHost));
    XB_HIP(hipMemcpy(&cur, S.cur + k, sizeof(int), hipMemcpyDeviceToHost));
    XB_HIP(hipMemcpy(&cnt, S.count + g, sizeof(int), hipMemcpyDeviceToHost));
    ssxy[0] = ssxy[1] = -1;
    if (cur < cnt) {
      XB_HIP(hipMemcpy(&ssxy[0], S.ssx + (size_t)g * MAXSS + cur, sizeof(int),
                       hipMemcpyDeviceToHost));
      XB_HIP(hipMemcpy(&ssxy[1], S.ssy + (size_t)g * MAXSS + cur, sizeof(int),
                       hipMemcpyDeviceToHost));
    }
  }
  return XB_OK;
}

int xb_get_diagnostics(xb_ctx* x, xb_diag* out) {
  if (!x || !out) return XB_ERR_ARG;
  Ctx& c = x->c;
  if (!c.haveResult) {
    c.err = "no result yet";
    return XB_ERR_ARG;
  }
  const int rc = setDev(c);
  if (rc) return rc;
  double* d = nullptr;
  XB_HIP(locked_malloc((void**)&d, sizeof(double) * (DIAG_BLOCKS * 10 + 10)));
  hipError_t e = launch_diag(c, d, d + DIAG_BLOCKS * 10);
  double q[10] = {0};
  DevScalars sc;
  if (e == hipSuccess)
    e = hipMemcpyAsync(q, d + DIAG_BLOCKS * 10, sizeof(q), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&sc, c.sc, sizeof(sc), hipMemcpyDeviceToHost, c.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
  (void)locked_free(d);
  XB_HIP(e);
  auto fill = [](const double* s, int64_t& n, double& mean, double& sige, double& sig,
                 double* x2) {
    n = (int64_t)s[0];
    mean = n > 0 ? s[1] / s[0] : 0.0;
    sige = n > 1 ? std::sqrt(std::max(0.0, (s[2] - s[1] * s[1] / s[0]) / (s[0] - 1.0))) : 1e10;
    sig = n > 0 ? s[3] / s[0] : 0.0;
    if (x2) *x2 = n > 0 ? s[4] / s[0] : 0.0;
  };
  std::memset(out, 0, sizeof(*out));
  out->kernSum = sc.kernSum;
  out->convTemplate = c.convTemplate;
  out->normalizeTemplate = c.m.normalizeTemplate;
  const KscScalars& ks = sc.k[c.tIdx];
  out->sssig = ks.firstClip[0];
  out->ssscat = ks.firstClip[1];
  out->fsig = ks.clipMean;
  out->fscat = ks.clipStd;
  fill(q, out->nx2nrm, out->dmean, out->dsige, out->dsig, &out->x2nrm);
  fill(q + 5, out->nok, out->dmeano, out->dsigeo, out->dsigo, nullptr);
  return XB_OK;
}

int xb_device_outputs(xb_ctx* x, const void** d_conv, const void** d_diff) {
  if (!x) return XB_ERR_ARG;
  if (d_conv) *d_conv = x->c.conv;
  if (d_diff) *d_diff = x->c.diff;
  return XB_OK;
}

int xb_get_trace(xb_ctx* x, xb_trace* tr) {
  if (!x || !tr) return XB_ERR_ARG;
  Ctx& c = x->c;
  std::memset(tr, 0, sizeof(*tr));
  tr->w = c.w;
  tr->h = c.h;
  tr->stampsx = c.stampsx;
  tr->stampsy = c.stampsy;
  tr->fStampWidth = c.fsw;
  tr->nGrid = c.nGrid;
  tr->retried = c.retried;
  const int N = c.nGrid;
  std::vector<int> ix((size_t)N * MAXSS), iy((size_t)N * MAXSS), cnt(N), kept(N), cur(N);
  std::vector<double> sv((size_t)N * MAXSS), sky(N), fw(N);
  for (int im = 0; im < 2; im++) {
    DevStamps& s = c.st[im];
    if (N == 0) break;
    XB_HIP(hipMemcpy(sky.data(), s.sky, N * sizeof(double), hipMemcpyDeviceToHost));
    XB_HIP(hipMemcpy(fw.data(), s.fwhm, N * sizeof(double), hipMemcpyDeviceToHost));
    XB_HIP(hipMemcpy(cnt.data(), s.count, N * sizeof(int), hipMemcpyDeviceToHost));
    XB_HIP(hipMemcpy(ix.data(), s.ssx, ix.size() * sizeof(int), hipMemcpyDeviceToHost));
    XB_HIP(hipMemcpy(iy.data(), s.ssy, iy.size() * sizeof(int), hipMemcpyDeviceToHost));
    XB_HIP(hipMemcpy(sv.data(), s.ssv, sv.size() * sizeof(double), hipMemcpyDeviceToHost));
    XB_HIP(hipMemcpy(kept.data(), s.kept, N * sizeof(int), hipMemcpyDeviceToHost));
    for (int g = 0; g < N; g++) {
      tr->sky[im][g] = sky[g];
      tr->fwhm[im][g] = fw[g];
      tr->ssCount[im][g] = cnt[g];
      for (int i = 0; i < MAXSS; i++) {
        tr->ssX[im][g * XB_MAX_SS + i] = ix[(size_t)g * MAXSS + i];
        tr->ssY[im][g * XB_MAX_SS + i] = iy[(size_t)g * MAXSS + i];
        tr->ssVal[im][g * XB_MAX_SS + i] = sv[(size_t)g * MAXSS + i];
      }
    }
    tr->stampCount[im] = c.nKept[im];
    for (int k = 0; k < c.nKept[im]; k++) tr->stampIndex[im][k] = kept[k];
  }
  DevScalars sc;
  int rc = readScalars(c, sc);
  if (rc) return rc;
  for (int im = 0; im < 2; im++) {
    tr->nTest[im] = c.nKept[im] ? sc.nTest[im] : 0;
    tr->merit[im] = c.nKept[im] ? sc.merit[im] : 666;
    tr->kernelSumCd[im] = c.nKept[im] ? sc.kernelSumCd[im] : 0.0;
    // a direction without stamps has no trial fit (cd returns 666 before it)
    if (c.nKept[im])
      XB_HIP(hipMemcpy(tr->cdSolution[im], c.cdSol + (size_t)im * NSOL, NSOL * sizeof(double),
                       hipMemcpyDeviceToHost));
  }
  tr->convTemplate = c.convTemplate;
  tr->kscIterations = c.kscIterations;
  tr->nSol = NSOL;
  XB_HIP(hipMemcpy(tr->solution, c.ksSol + (size_t)c.tIdx * NSOL, NSOL * sizeof(double),
                   hipMemcpyDeviceToHost));
  if (N > 0) {
    XB_HIP(hipMemcpy(cur.data(), c.st[c.tIdx].cur, N * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < c.nKept[c.tIdx]; k++) tr->finalCurrent[k] = cur[k];
  }
  tr->kernSum = sc.kernSum;
  tr->captureFailures = c.captureFailures;
  tr->screened = c.nScreened;
  tr->screenRestored = c.nRestored;
  tr->screenConfirmed = c.nConfirmed;
  return XB_OK;
}

static int runStages(Ctx& c, double* ms, bool timed) {
  double t0 = nowMs(), t = t0;
  auto mark = [&](int i) -> int {
    if (!timed) return XB_OK;
    XB_HIP(hipStreamSynchronize(c.stream));
    double n = nowMs();
    ms[i] = n - t;
    t = n;
    return XB_OK;
  };
  int rc = mark(0);
  if (rc) return rc;
  rc = stageSss(c);
  if (rc) return rc;
  if ((rc = mark(1))) return rc;
  rc = stageCmv(c);
  if (rc) return rc;
  if ((rc = mark(2))) return rc;
  if (canSpeculate(c)) {
    // cd and both directions' clip-refit loops at once; the -vt breakdown
    // books their time under KSC (CD reads 0)
    rc = stageCdKsc(c);
    if (rc) {
      (void)syncAll(c);
      return rc;
    }
    if ((rc = mark(3))) return rc;
    ms[4] = ms[3];
    ms[3] = 0.0;
  } else {
    rc = stageCd(c, nullptr);
    if (rc) return rc;
    if ((rc = mark(3))) return rc;
    rc = stageKsc(c);
    if (rc) return rc;
    if ((rc = mark(4))) return rc;
  }
  rc = stageConv(c, nullptr, false);
  if (rc) return rc;
  if ((rc = mark(5))) return rc;
  rc = stageSub(c);
  if (rc) return rc;
  rc = joinRestore(c);
  if (rc) return rc;
  if ((rc = mark(6))) return rc;
  return XB_OK;
}

}  // extern "C"

namespace xb {

// The whole pipeline on the caller's device buffers: the images are read where
// they lie and the convolved and difference images written where the caller
// wants them (null: the context's own buffers), so no image is copied.  The
// context's pointers are restored before this returns, after its streams
// drained; c.kernSum holds the kernel sum.
int pipelineRunInto(Ctx& c, const xb_params* p, int w, int h, const void* tin, const void* sin,
                    void* cout, void* dout) {
  if (!tin || !sin || !p || w <= 0 || h <= 0) return XB_ERR_ARG;
  // an output written in place must not overlap an input
  const size_t bytes = (size_t)w * h * sizeof(float);
  auto overlap = [bytes](const void* a, const void* b) {
    const char* x = static_cast<const char*>(a);
    const char* y = static_cast<const char*>(b);
    return a && b && x < y + bytes && y < x + bytes;
  };
  if (overlap(cout, tin) || overlap(cout, sin) || overlap(dout, tin) || overlap(dout, sin) ||
      overlap(cout, dout)) {
    c.err = "a pair's output buffer overlaps one of its inputs or the other output";
    return XB_ERR_ARG;
  }
  int rc = setDev(c);
  if (rc) return rc;
  rc = stageInitCommon(c, p, w, h);
  if (rc) return rc;
  float* const own[4] = {c.img[0], c.img[1], c.conv, c.diff};
  c.img[0] = const_cast<float*>(static_cast<const float*>(tin));
  c.img[1] = const_cast<float*>(static_cast<const float*>(sin));
  if (cout) c.conv = static_cast<float*>(cout);
  if (dout) c.diff = static_cast<float*>(dout);
  rc = stageMask(c);
  double ms[9];
  if (!rc) rc = runStages(c, ms, false);
  const int rs = syncAll(c);  // nothing may still use the caller's buffers
  // the tile kernels published the kernel sum into the pinned mirror
  if (!rc && !rs) c.kernSum = c.hsc->kernSum;
  c.img[0] = own[0];
  c.img[1] = own[1];
  c.conv = own[2];
  c.diff = own[3];
  c.haveResult = false;  // the outputs are not in the context's buffers
  return rc ? rc : rs;
}

}  // namespace xb

extern "C" {

int xb_pipeline_run(xb_ctx* x, const xb_params* p, int w, int h, const float* tmpl,
                    const float* sci, float* conv, float* diff, double* stage_ms) {
  if (!x) return XB_ERR_ARG;
  Ctx& c = x->c;
  double ms[9] = {0};
  double t0 = nowMs();
  int rc = xb_init(x, p, w, h, tmpl, sci);
  if (rc) return rc;
  XB_HIP(hipStreamSynchronize(c.stream));
  ms[0] = nowMs() - t0;
  double ms2[9] = {0};
  rc = runStages(c, ms2, true);
  if (rc) return rc;
  for (int i = 1; i <= 6; i++) ms[i] = ms2[i];
  double tf = nowMs();
  size_t n = (size_t)w * h;
  if (conv)
    XB_HIP(hipMemcpyAsync(conv, c.conv, n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  if (diff)
    XB_HIP(hipMemcpyAsync(diff, c.diff, n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  XB_HIP(hipStreamSynchronize(c.stream));
  XB_HIP(hipMemcpy(&c.kernSum, &c.sc->kernSum, sizeof(double), hipMemcpyDeviceToHost));
  ms[7] = nowMs() - tf;
  ms[8] = nowMs() - t0;
  if (stage_ms)
    for (int i = 0; i < 9; i++) stage_ms[i] = ms[i];
  return XB_OK;
}

int xb_pipeline_run_device(xb_ctx* x, const xb_params* p, int w, int h, const void* d_tmpl,
                           const void* d_sci) {
  if (!x) return XB_ERR_ARG;
  Ctx& c = x->c;
  int rc = xb_init_device(x, p, w, h, d_tmpl, d_sci);
  if (rc) return rc;
  double ms[9];
  rc = runStages(c, ms, false);
  if (rc) return rc;
  XB_HIP(hipStreamSynchronize(c.stream));
  return XB_OK;
}

}  // extern "C"
