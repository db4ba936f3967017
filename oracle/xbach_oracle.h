/*
 * xbach_oracle.h — CPU restatement of the X-BACH PSF-matching path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU path is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product (libxbach) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Where the reference
 * compacts with atomics (nondeterministic order: cl_kern/sss.cl:184,
 * cl_kern/cd.cl:76,85,679,687, cl_kern/ksc.cl:30,41) the oracle uses the
 * canonical ascending-index order.  Floating-point expressions keep the
 * reference's operand order and types; the library is compiled with
 * -ffp-contract=off so no multiply-add is fused.
 *
 * Modes: XO_MODE_REF keeps the reference's FP32 steps (FP32 Y-pass accumulation,
 * FP32 stamp model); XO_MODE_FP64 runs them in FP64 (the full-FP64 oracle the
 * mixed-precision errors are reported against).
 */
#ifndef XBACH_ORACLE_H
#define XBACH_ORACLE_H

#include <stdint.h>

#include "../include/xbach.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { XO_MODE_REF = 0, XO_MODE_FP64 = 1 };

/* Whole path init -> sub.  conv/diff/mask/trace/stage_ms may be NULL.
 * stage_ms: 9 entries as the -vt lines (Ini SSS CMV CD KSC Conv Sub Fin BACH). */
int xo_run(const xb_params* p, int w, int h, const float* tmpl, const float* sci,
           int mode, int threads, float* conv, float* diff, uint16_t* mask,
           xb_trace* trace, double* stage_ms);

/* ---- unit-level entry points used by the oracle's own pin tests ---------- */
/* First n draws of NR ran1 with idum = -666 (cl_kern/sss.cl:41-79). */
void xo_ran1_seq(int n, double* out);
/* GPU-style sigma clip (src/bachUtil.cpp:34-128, cl_kern/bach.cl:196-251). */
void xo_sigma_clip(const double* data, int n, int maxIter, double alpha,
                   double* mean, double* stdDev);
/* NR Crout LU with implicit pivoting, 1-based, on an (n+1)x(n+1) row-major
 * buffer, then back-substitution in place on b[0..n] (src/bachUtil.cpp:319-411).
 * Returns the ludcmp status (1 = singular row, factorisation skipped). */
int xo_lu_solve(double* a, int n, double* b, int* index);
/* Same LU in the per-stamp device form (cl_kern/bach.cl:68-194), matrix of
 * size m x m with row/column 0 unused. */
void xo_lu_solve_stamp(double* a, int m, double* b);
/* Basis tables (cl_kern/cmv.cl:1-68, include/datatypeUtil.h:37-118):
 * fx, fy: nPSF*fKernelWidth; vec: nPSF*fKernelWidth^2.  Returns nPSF. */
int xo_basis(const xb_params* p, double* fx, double* fy, double* vec);
/* Bitonic sort network of cl_kern/sss.cl:135-150 on n (power of 2) values. */
void xo_bitonic(double* v, int n);
/* Histogram sky / sky-sigma of one stamp (cl_kern/sss.cl:190-324) on a plain
 * pixel list (all pixels unmasked). */
void xo_histogram(const double* pix, int npix, const double* sortedSamples,
                  int sampleCount, double mean, double invStdDev, double iqRange,
                  double sigClipAlpha, double* skyEst, double* fwhm);
/* fillStamps for one substamp (cl_kern/cmv.cl:70-248): W (wRows x ss^2),
 * Q (q x q), B (q) with q = nPSF+2.  img = convolved image, img2 = target. */
int xo_fill_one(const xb_params* p, int mode, int w, int h, const double* img,
                const double* img2, int ssx, int ssy, double* W, double* Q,
                double* B);

/* Capture every global normal matrix and RHS built by the next xo_run
 * (cd template, cd science, then each ksc iteration). */
void xo_capture(int enable);
int xo_captured_count(void);
int xo_captured_get(int i, double* M, double* rhs);

#ifdef __cplusplus
}
#endif
#endif
