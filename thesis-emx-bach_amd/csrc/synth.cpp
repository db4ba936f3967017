// synth.cpp — seeded synthetic image pairs for tests and the benchmark.
//
// A star field rendered twice: template and science share star positions and
// fluxes and differ in PSF width and noise.  Generator: xoshiro256** seeded by
// SplitMix64 (Vigna); fluxes by the inverse CDF of dN/dF ~ F^slope on
// [fluxMin, fluxMax]; circular Gaussian PSFs integrated over each pixel with
// erf; sky + linear gradient; Poisson noise by Knuth's product method below
// lambda = 30 and by PTRS (Hoermann 1993, transformed rejection) above.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/xbach.h"

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; i++) {  // SplitMix64
      z += 0x9e3779b97f4a7c15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
      x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
      s[i] = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {  // xoshiro256**
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

long poisson(Rng& r, double mu) {
  if (!(mu > 0.0)) return 0;
  if (mu < 30.0) {
    double L = std::exp(-mu), p = 1.0;
    long k = 0;
    do {
      k++;
      p *= r.uniform();
    } while (p > L);
    return k - 1;
  }
  const double smu = std::sqrt(mu);
  const double b = 0.931 + 2.53 * smu;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  const double logmu = std::log(mu);
  for (;;) {
    double U = r.uniform() - 0.5;
    double V = r.uniform();
    double us = 0.5 - std::fabs(U);
    long k = (long)std::floor((2.0 * a / us + b) * U + mu + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0 || (us < 0.013 && V > us)) continue;
    if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <=
        -mu + k * logmu - std::lgamma((double)k + 1.0))
      return k;
  }
}

struct Star {
  double x, y, f;
};

void render(const std::vector<Star>& stars, int w, int h, double fluxScale,
            double (*fwhmAt)(const xb_synth_cfg&, double, double), const xb_synth_cfg& cfg,
            std::vector<double>& model) {
  const double kFwhm = 2.0 * std::sqrt(2.0 * std::log(2.0));
  for (const Star& s : stars) {
    double sig = fwhmAt(cfg, s.x, s.y) / kFwhm;
    int rad = (int)std::ceil(5.0 * sig) + 1;
    int x0 = std::max(0, (int)std::floor(s.x) - rad), x1 = std::min(w - 1, (int)std::floor(s.x) + rad);
    int y0 = std::max(0, (int)std::floor(s.y) - rad), y1 = std::min(h - 1, (int)std::floor(s.y) + rad);
    if (x0 > x1 || y0 > y1) continue;
    const double inv = 1.0 / (std::sqrt(2.0) * sig);
    std::vector<double> gx(x1 - x0 + 1), gy(y1 - y0 + 1);
    for (int x = x0; x <= x1; x++)
      gx[x - x0] = 0.5 * (std::erf((x + 0.5 - s.x) * inv) - std::erf((x - 0.5 - s.x) * inv));
    for (int y = y0; y <= y1; y++)
      gy[y - y0] = 0.5 * (std::erf((y + 0.5 - s.y) * inv) - std::erf((y - 0.5 - s.y) * inv));
    const double f = s.f * fluxScale;
    for (int y = y0; y <= y1; y++)
      for (int x = x0; x <= x1; x++) model[(size_t)y * w + x] += f * gx[x - x0] * gy[y - y0];
  }
}

double fwhmTemplate(const xb_synth_cfg& c, double, double) { return c.fwhmT; }
double fwhmScience(const xb_synth_cfg& c, double x, double y) {
  double fx = c.w > 1 ? x / (c.w - 1) : 0.0, fy = c.h > 1 ? y / (c.h - 1) : 0.0;
  return c.fwhmS0 + c.dfwhmX * fx + c.dfwhmY * fy;
}

}  // namespace

extern "C" int xb_synth_pair(const xb_synth_cfg* cfg, float* tmpl, float* sci) {
  if (!cfg || !tmpl || !sci || cfg->w <= 0 || cfg->h <= 0 || cfg->nstars < 0) return XB_ERR_ARG;
  if (!(cfg->fluxMin > 0.0) || !(cfg->fluxMax >= cfg->fluxMin)) return XB_ERR_ARG;
  const int w = cfg->w, h = cfg->h;
  Rng rs(cfg->seed);
  std::vector<Star> stars((size_t)cfg->nstars);
  const double a = cfg->fluxSlope + 1.0;
  const double lo = cfg->fluxMin, hi = cfg->fluxMax;
  for (Star& s : stars) {
    s.x = rs.uniform() * w;
    s.y = rs.uniform() * h;
    double u = rs.uniform();
    if (std::fabs(a) < 1e-12) s.f = lo * std::exp(u * std::log(hi / lo));
    else s.f = std::pow(std::pow(lo, a) + u * (std::pow(hi, a) - std::pow(lo, a)), 1.0 / a);
  }
  std::vector<double> mt((size_t)w * h), ms((size_t)w * h);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      double sky = cfg->sky + cfg->skyGx * x + cfg->skyGy * y;
      mt[(size_t)y * w + x] = sky;
      ms[(size_t)y * w + x] = sky;
    }
  render(stars, w, h, 1.0, fwhmTemplate, *cfg, mt);
  render(stars, w, h, cfg->sciFluxScale > 0 ? cfg->sciFluxScale : 1.0, fwhmScience, *cfg, ms);
  Rng nt(cfg->seed ^ 0x5445'4d50'4c41'5445ull), ns(cfg->seed ^ 0x5343'4945'4e43'4545ull);
  for (size_t i = 0; i < mt.size(); i++) {
    tmpl[i] = (float)poisson(nt, mt[i]);
    sci[i] = (float)poisson(ns, ms[i]);
  }
  return XB_OK;
}
