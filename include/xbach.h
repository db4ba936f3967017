/*
 * xbach.h — C ABI of the MI355X-native X-BACH PSF-matching + subtraction path.
 *
 * The reference (X-BACH, C++20 + OpenCL) reaches this path through its C++ stage
 * API `init/sss/cmv/cd/ksc/conv/sub` (include/bach.h:66-87) over the shared
 * `ClData` state (include/bach.h:7-64) and through the OpenCL plumbing in
 * include/clUtil.h:9-41 (platform/device choice, program build).  This header is
 * the drop-in replacement for that plumbing: plain C, opaque handles, pointers and
 * sizes, integer status codes, no exceptions or exits across the boundary.
 *
 * Images cross the boundary as FP32 row-major (x fastest) arrays of w*h pixels,
 * exactly what the reference reads from BITPIX -32 FITS files
 * (src/fitsUtil.cpp:8-38).  Outputs are FP32 as the reference writes FLOAT_IMG
 * (src/fitsUtil.cpp:40-64).
 */
#ifndef XBACH_H
#define XBACH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference calls std::exit(1) or throws; see
 *      src/bach.cpp:20-24,166-169, src/bachUtil.cpp:149-155) ---------------- */
enum {
  XB_OK = 0,
  XB_ERR_ARG = 1,          /* bad argument / unsupported parameter            */
  XB_ERR_SIZE_MISMATCH = 2,/* template and science sizes differ (bach.cpp:20)  */
  XB_ERR_NO_SUBSTAMPS = 3, /* "No substamps found" (bach.cpp:166-169)          */
  XB_ERR_STAMP_PIXELS = 4, /* "Not enough pixels in a stamp" (bachUtil.cpp:151)*/
  XB_ERR_HIP = 5,          /* HIP runtime error                                */
  XB_ERR_NO_DEVICE = 6,    /* no GPU / bad device index                        */
  XB_ERR_NOT_BUILT = 7,    /* kernels for gfx950 missing                       */
  XB_ERR_CAPACITY = 8,     /* a size exceeds a documented maximum              */
  XB_ERR_IO = 9            /* FITS file cannot be opened, read or written       */
};

/* ---- 16-bit pixel mask bits (include/datatypeUtil.h:172-188) ------------- */
enum {
  XB_MASK_BAD_PIX_VAL = 1 << 0,
  XB_MASK_SAT_PIXEL = 1 << 1,
  XB_MASK_LOW_PIXEL = 1 << 2,
  XB_MASK_NAN_PIXEL = 1 << 3,
  XB_MASK_BAD_CONV = 1 << 4,
  XB_MASK_INPUT_MASK = 1 << 5,
  XB_MASK_OK_CONV = 1 << 6,
  XB_MASK_BAD_INPUT = 1 << 7,
  XB_MASK_BAD_PIXEL_T = 1 << 8,
  XB_MASK_SKIP_T = 1 << 9,
  XB_MASK_BAD_PIXEL_S = 1 << 10,
  XB_MASK_SKIP_S = 1 << 11,
  XB_MASK_BAD_OUTPUT = 1 << 12,
  XB_MASK_ALL = (1 << 13) - 1
};

/* ---- method parameters: the reference's `Arguments` (include/argsUtil.h:7-57)
 *      minus the I/O names.  xb_params_default() fills the reference defaults. */
typedef struct xb_params {
  int32_t stampsx, stampsy;        /* stamp grid (-sx / -sy)                     */
  double threshLow, threshHigh;    /* 0, 25000                                   */
  double threshKernFit;            /* 20                                         */
  double sigKernFit;               /* 2                                          */
  double sigClipAlpha;             /* 3                                          */
  double iqRange;                  /* 1.35                                       */
  int32_t maxKSStamps;             /* 3 (6 candidates kept per stamp)            */
  int32_t hSStampWidth, fSStampWidth; /* 15, 31                                  */
  int32_t hKernelWidth, fKernelWidth; /* 10, 21                                  */
  double inSpreadMaskFactor;       /* 1.0                                        */
  int32_t normalizeTemplate;       /* 1                                          */
  int32_t backgroundOrder;         /* 1                                          */
  int32_t kernelOrder;             /* 2                                          */
  int32_t ngauss;                  /* 3                                          */
  int32_t dg[3];                   /* 6, 4, 2                                    */
  float bg[3];                     /* 1/(2 sigma^2), sigma = 0.7, 1.5, 3.0 (float) */
  int32_t verbose;
} xb_params;

/* ---- capacities of the trace record ------------------------------------- */
#define XB_MAX_STAMPS 1024
#define XB_MAX_SS 16
#define XB_MAX_SOL 512

/* Everything the path decides, recorded so two implementations can be compared
 * field by field.  Image index 0 = template, 1 = science (before the cd swap). */
typedef struct xb_trace {
  int32_t w, h;
  int32_t stampsx, stampsy, fStampWidth, nGrid;
  int32_t retried;                              /* the <10 % refill retry ran   */
  /* sss result (final pass), grid order */
  double sky[2][XB_MAX_STAMPS];
  double fwhm[2][XB_MAX_STAMPS];
  int32_t ssCount[2][XB_MAX_STAMPS];            /* min(found, maxKSStamps)       */
  int32_t ssX[2][XB_MAX_STAMPS * XB_MAX_SS];    /* 2*maxKSStamps slots / stamp   */
  int32_t ssY[2][XB_MAX_STAMPS * XB_MAX_SS];    /* (INT_MAX padding)            */
  double ssVal[2][XB_MAX_STAMPS * XB_MAX_SS];   /* (-INF padding)               */
  int32_t stampCount[2];                        /* non-empty stamps             */
  int32_t stampIndex[2][XB_MAX_STAMPS];         /* grid index of kept stamps    */
  /* cd */
  int32_t nTest[2];
  double kernelSumCd[2];
  double merit[2];
  double cdSolution[2][XB_MAX_SOL];
  int32_t convTemplate;
  /* ksc */
  int32_t kscIterations;
  int32_t nSol;
  double solution[XB_MAX_SOL];
  int32_t finalCurrent[XB_MAX_STAMPS];         /* chosen direction, kept order */
  /* conv */
  double kernSum;
  /* graph captures of the global solve that failed on this context since it
   * was created (the solve then ran its kernels directly; after two failures
   * the context stops capturing).  Zero in a healthy run. */
  int32_t captureFailures;
  /* XB_FORM_SCREEN: substamps the screen replaced, and of those, the ones the
   * chi2 re-check after convergence kept (restored, the loop then ran again
   * without the screen) and the ones it rejected too.  Zero otherwise. */
  int32_t screened, screenRestored, screenConfirmed;
} xb_trace;

void xb_params_default(xb_params* p);

/* ---- context: one per host thread, one primary HIP stream ---------------- */
typedef struct xb_ctx xb_ctx;

/* Replaces getDefaultPlatform/getDefaultDevice/loadBuildPrograms
 * (include/clUtil.h:9-41, src/main.cpp:42-48).  `device` is honoured (the
 * reference ignores -d, src/clUtil.cpp:45).  flags: reserved, pass 0. */
int xb_ctx_create(int device, int flags, xb_ctx** out);
int xb_ctx_destroy(xb_ctx* ctx);
const char* xb_last_error(xb_ctx* ctx);
int xb_device_name(xb_ctx* ctx, char* buf, size_t len);

/* ---- stage API: the reference's stages over the context's resident state.
 * Each mirrors one function of include/bach.h:66-87 and keeps its meaning. */

/* init (src/bach.cpp:14-42): upload the pair, size check, maskInput+spreadMask. */
int xb_init(xb_ctx* ctx, const xb_params* p, int w, int h, const float* tmpl,
            const float* sci);
/* Same as xb_init with device-resident FP32 inputs (no host copy). */
int xb_init_device(xb_ctx* ctx, const xb_params* p, int w, int h,
                   const void* d_tmpl, const void* d_sci);
/* sss (src/bach.cpp:44-170): stamp stats + substamp selection, both images. */
int xb_sss(xb_ctx* ctx);
/* cmv (src/bach.cpp:172-268): basis + per-stamp W, Q, B, both directions. */
int xb_cmv(xb_ctx* ctx);
/* cd (src/bach.cpp:270-313): test fits, merits, direction.  *convTemplate out. */
int xb_cd(xb_ctx* ctx, int* convTemplate);
/* ksc (src/bach.cpp:315-323): global fit with the chi2 clip-refit loop. */
int xb_ksc(xb_ctx* ctx);
/* conv (src/bach.cpp:325-413): spatially varying convolution; *kernSum out.
 * The same kernel forms the difference image (src/bach.cpp:415-440) in its
 * epilogue, from the unrounded FP64 convolution of each pixel. */
int xb_conv(xb_ctx* ctx, double* kernSum);
/* sub (src/bach.cpp:415-440): marks the difference image formed by xb_conv
 * as the pipeline's result (no kernel of its own). */
int xb_sub(xb_ctx* ctx);

/* Results of the last pipeline (host copies). */
int xb_get_conv(xb_ctx* ctx, float* out);
int xb_get_diff(xb_ctx* ctx, float* out);
int xb_get_mask(xb_ctx* ctx, uint16_t* out);
int xb_get_trace(xb_ctx* ctx, xb_trace* out);
/* Fit data of kept stamp k (kept order) of image im (0 template, 1 science)
 * as the last cmv or ksc refill left it: the W rows of fillStamps
 * (cl_kern/cmv.cl:70-193; 52 x 961 row-major, raw: the even-even rows before
 * the W[n] - W[0] of cmv.cl:137-154), G (52 x 52, lower triangle: the Gram
 * matrix of the rows after that subtraction; createQ, cmv.cl:195-215, is G
 * shifted to 1-based), Bc (52: the rows' products with the target, createB
 * cmv.cl:217-248), Br (3: the background rows' products in row-major order,
 * cl_kern/cd.cl:318-344), and the substamp centre ssxy[2] (-1 when the stamp
 * has run out of substamps).  Any output may be NULL. */
int xb_get_stamp_fit(xb_ctx* ctx, int im, int k, double* W, double* G, double* Bc, double* Br,
                     int* ssxy);
/* Quality figures of the last result, the ones HOTPANTS writes to the
 * difference image's header (tests/test01_sub.fits: KSUM00, SSSIG00,
 * SSSCAT00, FSIG00, FSCAT00, X2NRM00, NX2NRM00, DMEAN00, DSIGE00, DSIG00 and
 * the *O00 "OK pixel" variants).  Good pixels: interior (10 px border) and not
 * BAD_OUTPUT; OK pixels: good and OK_CONV.  noise = sqrt(|conv| + |S|), the
 * per-pixel noise of the reference's chi2 (cl_kern/cd.cl:578).  The stamp
 * figures of merit are the clipped mean and sd of the stamps' chi2 in the first
 * and in the last pass of the clip-refit loop.  A standard deviation over fewer
 * than two pixels reads 1e10, as HOTPANTS writes it. */
typedef struct xb_diag {
  double kernSum;                 /* KSUM00                                  */
  double sssig, ssscat;           /* SSSIG00, SSSCAT00 (first pass)          */
  double fsig, fscat;             /* FSIG00, FSCAT00 (final pass)            */
  double x2nrm;                   /* X2NRM00: mean (d / noise)^2, good px    */
  int64_t nx2nrm;                 /* NX2NRM00: good pixels                   */
  double dmean, dsige, dsig;      /* DMEAN00, DSIGE00, DSIG00 (good pixels)  */
  int64_t nok;                    /* OK pixels                               */
  double dmeano, dsigeo, dsigo;   /* DMEANO00, DSIGEO00, DSIGO00             */
  int32_t convTemplate;           /* CONVOL00: 1 template, 0 science         */
  int32_t normalizeTemplate;      /* PHOTNORM: 1 template, 0 science         */
} xb_diag;
int xb_get_diagnostics(xb_ctx* ctx, xb_diag* out);
/* Device pointers of the results (valid until the next xb_init). */
int xb_device_outputs(xb_ctx* ctx, const void** d_conv, const void** d_diff);

/* Whole path init->sub in one call (host buffers in and out).  Any output
 * pointer may be NULL.  stage_ms (9 doubles, may be NULL) receives the -vt
 * breakdown Ini, SSS, CMV, CD, KSC, Conv, Sub, Fin(0), total. */
int xb_pipeline_run(xb_ctx* ctx, const xb_params* p, int w, int h,
                    const float* tmpl, const float* sci, float* conv, float* diff,
                    double* stage_ms);
/* Whole path on device-resident inputs; outputs stay resident
 * (xb_device_outputs).  This is the form the throughput benchmark times. */
int xb_pipeline_run_device(xb_ctx* ctx, const xb_params* p, int w, int h,
                           const void* d_tmpl, const void* d_sci);

/* ---- the global FP64 solve on its own: ludcmp + lubksb for the 292-unknown
 * system (include/bachUtil.h:33-36, src/bachUtil.cpp:319-411), 1-based as
 * the reference: a is 293x293 row-major (row/column 0 unused), b has 298
 * entries (1..292 used), x receives 298 entries.  Bit-identical to the
 * reference's arithmetic, including its zero-row early return. */
int xb_lu_solve(xb_ctx* ctx, const double* a, const double* b, double* x);

/* ---- device buffers (replaces the cl::Buffer handling of the reference,
 *      src/bach.cpp:29-39; owned by the caller, freed with xb_buf_free) ---- */
int xb_buf_alloc(xb_ctx* ctx, size_t bytes, void** d_out);
int xb_buf_free(xb_ctx* ctx, void* d);
int xb_buf_upload(xb_ctx* ctx, void* d, const void* host, size_t bytes);
int xb_buf_download(xb_ctx* ctx, void* host, const void* d, size_t bytes);

/* ---- per-kernel timing with HIP events on the context's stream ---------- */
/* XB_PROF_SUB stays at zero: the difference is formed inside XB_PROF_CONV. */
enum { XB_PROF_CONV = 0, XB_PROF_SUB = 1, XB_PROF_SOLVE = 2, XB_PROF_FILL = 3,
       XB_PROF_SSS = 4, XB_PROF_MASK = 5, XB_PROF_N = 6 };
/* enable != 0 starts recording (and clears the totals). */
int xb_profile_enable(xb_ctx* ctx, int enable);
/* Total milliseconds and launch counts per XB_PROF_* slot since enabling. */
int xb_profile_get(xb_ctx* ctx, double* ms, int* counts, int n);

/* ---- host-side helpers of the library (no GPU needed) -------------------- */
/* Seeded synthetic pair: Gaussian stars (pixel-integrated), power-law fluxes,
 * sky + linear gradient, Poisson noise.  FWHM of the science PSF ramps linearly
 * from fwhmS0 at (0,0) by dfwhmX across x and dfwhmY across y.  xoshiro256**
 * seeded by SplitMix64; Poisson by Knuth below lambda 30, PTRS above. */
typedef struct xb_synth_cfg {
  int32_t w, h;
  int32_t nstars;
  double fluxMin, fluxMax, fluxSlope;  /* dN/dF ~ F^fluxSlope on [min,max]  */
  double fwhmT;                        /* template FWHM (px)                */
  double fwhmS0, dfwhmX, dfwhmY;       /* science FWHM ramp                 */
  double sky, skyGx, skyGy;            /* sky + gx*x + gy*y (both images)   */
  double sciFluxScale;                 /* science/template flux ratio       */
  uint64_t seed;
} xb_synth_cfg;
int xb_synth_pair(const xb_synth_cfg* cfg, float* tmpl, float* sci);

/* ---- FITS I/O (replaces readImage / writeImage, src/fitsUtil.cpp:8-64) ----
 * xb_fits_read: primary HDU of BITPIX -32 or -64 (other types are refused
 * with the reference's message), BSCALE/BZERO applied, pixels narrowed to
 * float as Image::operator[] does (include/datatypeUtil.h:158); plain or
 * gzip-compressed files.  out == NULL queries the size only; cap = number of
 * floats out can hold.  msg (optional) receives the error text.
 * xb_fits_write: FLOAT_IMG primary HDU, overwriting like the reference's "!"
 * prefix (include/datatypeUtil.h:169).  Neither needs a GPU. */
int xb_fits_read(const char* path, int* w, int* h, float* out, size_t cap, char* msg,
                 size_t msglen);
int xb_fits_write(const char* path, int w, int h, const float* data, char* msg, size_t msglen);
/* As xb_fits_write with ncards extra 80-character header cards (cards holds
 * ncards * 80 characters, no terminators needed) before END. */
int xb_fits_write_cards(const char* path, int w, int h, const float* data, const char* cards,
                        int ncards, char* msg, size_t msglen);

/* Waits until the work queued on the context has finished.  The reference's
 * stage functions return only when their device work is done (blocking reads
 * after every launch, e.g. src/bach.cpp:29-39); a caller that times stages
 * the way src/main.cpp:56-142 does calls this after each stage. */
int xb_sync(xb_ctx* ctx);

/* ---- forms of the path ----------------------------------------------------
 * XB_FORM_PLAIN (0): the reference's arithmetic, bit-identical to it.
 * XB_FORM_MIXED: cmv's basis convolution and normal-matrix products run on
 * the matrix cores in BF16 split precision (hi + lo) with FP32 accumulation
 * for every kept stamp (the clip-refit loop's refills of a few stamps take the
 * FP64 fill, faster for a short list); the stamp and global LU solves, chi2,
 * convolution and difference stay FP64, the convolution's terms as fused
 * multiply-adds (in the reference's order) and the tile kernels as one FP64
 * matrix-core product per 16 tiles.
 * Results differ from the plain form at the 1e-5 level of W and Q; the
 * selected substamps (sss) are the same.  Applies from the next xb_init. */
#define XB_FORM_PLAIN 0
#define XB_FORM_MIXED 1
/* XB_FORM_SCREEN: after each of the first 16 passes of the clip-refit loop
 * (ksc), a small network (xb_screen_load) predicts from 32 features per stamp
 * which of the substamps the pass kept a later pass would replace, and
 * replaces them at once; the loop runs as usual and judges every substamp by
 * its chi2, and when it converges every screened substamp is re-judged by the
 * chi2 test under the converged fit (kept ones are restored and the loop runs
 * again without the screen).  XB_FORM_RECORD: keep the 32 features of every substamp each of
 * those passes kept and whether the loop finally replaced it
 * (xb_screen_samples): the training data of the network, labelled by the
 * method itself.  Forms combine (bitwise or). */
#define XB_FORM_SCREEN 2
#define XB_FORM_RECORD 4
/* With XB_FORM_MIXED: XB_FORM_FP16 forms W in FP16 split precision (hi + lo,
 * 22 significand bits, each stamp's window and the basis table scaled by
 * powers of two into FP16's range) instead of the BF16 three-part split;
 * XB_FORM_REFINE adds one step of FP64 iterative refinement to every global
 * solve (residual b - A x of the system as built, with compensated FP64
 * products and sums, solved with the same LU factors and added to x). */
#define XB_FORM_FP16 8
#define XB_FORM_REFINE 16
/* XB_FORM_LDLT: every global solve factors the symmetric positive definite
 * normal matrix as L D L^T without pivoting (FP64; the trailing updates on the
 * FP64 matrix cores) and substitutes, in one launch per system, instead of
 * Numerical Recipes' ludcmp + lubksb in the reference's order.  The solution
 * is not bit-identical to the reference's; on the fit's systems its error
 * against the exact solution is that of ludcmp's (~1e-8 normwise), and with
 * XB_FORM_REFINE (one step of FP64 iterative refinement in the same launch:
 * compensated residual of the system as built, the same factors) ~1e-11.  A
 * system with a zero row takes the reference's path (lubksb on the
 * unfactorised matrix), bit for bit. */
#define XB_FORM_LDLT 32
int xb_set_form(xb_ctx* ctx, int form);

/* ---- launch shapes ----------------------------------------------------------
 * The tiling of the kernels where it is a free choice (no counterpart in the
 * reference, whose kernels have fixed work-group sizes).  Every shape gives
 * the same results bit for bit; the batch autotuner picks them per size
 * bucket (xb_batch_autotune, xb_batch_get_shapes). */
typedef struct xb_shapes {
  int32_t conv_tiles;  /* 21x21 output tiles (one wave each) per convolution
                          workgroup: 1, 2, 4 or 8 (default 4)                 */
  int32_t mask_rows;   /* rows per 64-column tile of the input-mask pass: 16,
                          32 or 64 (default 64)                               */
  int32_t ldlt_group;  /* trailing tiles per load group of the L D L^T solve
                          (XB_FORM_LDLT): 2 or 4 (default 4)                  */
  int32_t find_staged; /* 1: the substamp search stages each stamp in LDS when
                          it fits (up to ~158 KB, one workgroup per CU); 0:
                          it reads the image and mask in place (default 1)    */
  int32_t stats_threads; /* threads per stamp of the stamp statistics: 256,
                            512 or 1024 (default 1024 for a context; 512 for
                            a batch's pairs, which share the GPU)             */
  int32_t pad;
} xb_shapes;
/* s == NULL restores the defaults; other values are refused (XB_ERR_ARG). */
int xb_set_shapes(xb_ctx* ctx, const xb_shapes* s);
int xb_get_shapes(xb_ctx* ctx, xb_shapes* s);

/* ---- the learned substamp screen ------------------------------------------
 * Model: XB_MLP_NPARAM floats = feature mean[32], feature sd[32],
 * W1[64][32] (row = hidden unit), b1[64], w2[64], b2, threshold.  The network
 * is 32 -> 64 (ReLU) -> 1 (logistic) on (feature - mean) / sd; W1 and w2 are
 * used rounded to FP16.  A substamp is screened out when p > threshold. */
#define XB_MLP_NFEAT 32
#define XB_MLP_NPARAM 2242
#define XB_MLP_NGRAD 2177 /* W1, b1, w2, b2 */
int xb_screen_load(xb_ctx* ctx, const float* model);
/* Samples of the last run with XB_FORM_RECORD or XB_FORM_SCREEN: for each
 * recorded pass (first pass first) the kept stamps of the convolved image in
 * kept order, feat[n][32], label[n] (2 replaced by that pass itself; kept by
 * that pass but not used by the final fit: 1 when the chi2 test under the
 * converged fit rejects it, 3 when it passes that test (XB_FORM_RECORD probes
 * each such substamp; under XB_FORM_SCREEN all are 1); 0 used by the final
 * fit, -1 no substamp), prob[n] (screen only, may be NULL).
 * cap = capacity in samples; *n receives the count, *screened the number of
 * substamps the screen replaced.  Any output may be NULL. */
int xb_screen_samples(xb_ctx* ctx, float* feat, int* label, float* prob, int cap, int* n,
                      int* screened);
/* The network on its own (host arrays; used to train and to check it):
 * xb_mlp_forward: prob[n] for features X[n][32].
 * xb_mlp_grad: gradient of L = sum_k wt[k] * BCE(p_k, y_k) / n with respect
 * to W1, b1, w2, b2 (grad[XB_MLP_NGRAD], in that order) and L (grad[NGRAD]).
 * xb_mlp_train: Adam (lr, 0.9, 0.999) over `epochs` passes in batches of
 * `batch` samples, in a seeded order; model is updated in place, loss[epochs]
 * (may be NULL) receives the mean loss per epoch. */
int xb_mlp_forward(xb_ctx* ctx, const float* model, const float* X, int n, float* prob);
int xb_mlp_grad(xb_ctx* ctx, const float* model, const float* X, const float* y,
                const float* wt, int n, float* grad);
int xb_mlp_train(xb_ctx* ctx, float* model, const float* X, const float* y, const float* wt,
                 int n, int epochs, int batch, float lr, uint64_t seed, float* loss);

/* ---- allocation-free gradient ---------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight training loops, harnesses that
 * forbid hipMalloc): xb_mlp_grad on DEVICE-resident arrays (d_model
 * [XB_MLP_NPARAM], d_X [n][32], d_y [n], d_wt [n] in, d_grad [XB_MLP_NGRAD + 1]
 * out, the loss last), with the per-block partial sums in a static per-device
 * array of the library (about 2.2 MB, held by each GPU the library runs on
 * until the process ends) instead of an allocation.  It makes no hipMalloc,
 * hipFree or other allocating call; n above XB_MLP_NOALLOC_MAX is refused with
 * XB_ERR_CAPACITY.  The result equals xb_mlp_grad's bit for bit.  The work is
 * enqueued on the context's primary stream and the call returns without
 * waiting (xb_sync waits).  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (two contexts, two threads) race. */
#define XB_MLP_NOALLOC_MAX 16384
int xb_mlp_grad_noalloc(xb_ctx* ctx, const float* d_model, const float* d_X, const float* d_y,
                        const float* d_wt, int n, float* d_grad);

/* xb_ctx_create flag: run both images' independent work on the one primary
 * stream instead of forking a second stream (used by the batch runtime, where
 * the concurrency comes from several pairs in flight). */
#define XB_CTX_ONE_STREAM 1
/* xb_ctx_create flag: launch the global solve's kernels one by one.  By
 * default each distinct solve of a context (a direction trial, the clip-refit
 * system, xb_lu_solve) is captured once into a hipGraph (its 23 kernels:
 * scaling, first panel, 19 panel steps, replay guard, substitution) and
 * replayed on later calls; the results are identical either way. */
#define XB_CTX_EAGER 2

/* ---- batches of independent pairs ----------------------------------------
 * The reference processes one pair per process (src/main.cpp:16-159).  A batch
 * runs many pairs, of any sizes, through the same pipeline at once: `workers`
 * host threads each own a context and pull pairs (largest first) from a shared
 * queue, so the latency-bound stages of one pair (the FP64 LU, the clip-refit
 * read-backs) overlap other pairs' work on the GPU.  Every pair's outputs are
 * bit-identical to xb_pipeline_run on that pair alone.  One batch handle per
 * host thread; runs on one handle do not overlap. */
typedef struct xb_batch xb_batch;

typedef struct xb_pair_result {
  int32_t status;          /* XB_OK or the pair's error code                  */
  int32_t convTemplate;    /* direction chosen by cd                          */
  int32_t kscIterations;   /* clip-refit passes                               */
  int32_t worker;          /* worker that ran the pair                        */
  double kernSum;          /* kernel sum at the image centre                  */
  double ms;               /* wall time of the pair on its worker             */
  int32_t captureFailures; /* failed graph captures of its worker's context   */
  int32_t screened;        /* XB_FORM_SCREEN: substamps the screen replaced    */
  int32_t screenRestored;  /* ... that the chi2 re-check restored              */
  int32_t pad;
} xb_pair_result;

/* 1 <= workers <= XB_BATCH_MAX_WORKERS (more is refused with
 * XB_ERR_CAPACITY).  Contexts are created up front with the flags of
 * xb_ctx_create plus XB_CTX_ONE_STREAM and XB_CTX_EAGER implied: worker
 * threads launch the global solve's kernels directly and never capture a
 * graph, since a stream capture on one thread can be invalidated by another
 * thread's work on the device. */
#define XB_BATCH_MAX_WORKERS 32
/* xb_batch_create flag: lockstep mode.  One host thread runs groups of pairs
 * together (each on its own worker context and stream), stage by stage, and
 * every global solve of a group — the direction trials of all its pairs, then
 * each clip-refit pass of the pairs still in their loop — is one batched
 * solve: the 23 launches of a single solve, each covering all the group's
 * systems (up to 32, so up to 16 pairs per group).  The group size is the
 * pairs-in-flight count of the pair's size bucket (set_inflight / autotune).
 * Each distinct batched solve (the same systems, form and launch shape) is
 * captured into a hipGraph at its second occurrence and replayed from then on,
 * unless the flags hold XB_CTX_EAGER (then its kernels launch one by one); a
 * failed capture is counted in every pair's captureFailures.  Outputs are
 * bit-identical to single runs.  Runs with XB_FORM_SCREEN or XB_FORM_RECORD
 * use the worker threads instead. */
#define XB_BATCH_LOCKSTEP 4
int xb_batch_create(int device, int workers, int flags, xb_batch** out);
int xb_batch_destroy(xb_batch* b);
const char* xb_batch_last_error(xb_batch* b);
/* Pairs in host memory: tmpl[i], sci[i] of w[i] x h[i] FP32 pixels; conv[i],
 * diff[i] receive the outputs (either array, or any entry, may be NULL).
 * results (n entries, may be NULL).  Returns XB_OK when every pair succeeded,
 * else the first failing pair's code (all pairs are still attempted). */
int xb_batch_run(xb_batch* b, const xb_params* p, int n, const int* w, const int* h,
                 const float* const* tmpl, const float* const* sci, float* const* conv,
                 float* const* diff, xb_pair_result* results);
/* Same with device-resident inputs and outputs (device pointers from
 * xb_buf_alloc or any HIP allocation on the batch's device).  The pipeline
 * reads the inputs and writes the outputs in place (no copies): they must stay
 * allocated and unchanged until the call returns, and a pair's outputs must not
 * overlap its inputs. */
int xb_batch_run_device(xb_batch* b, const xb_params* p, int n, const int* w, const int* h,
                        const void* const* d_tmpl, const void* const* d_sci,
                        void* const* d_conv, void* const* d_diff, xb_pair_result* results);
/* Number of pairs in flight used by the next run (1..workers; more is
 * refused with XB_ERR_CAPACITY); 0 restores "all workers".  Clears the
 * autotuner's per-bucket counts. */
int xb_batch_set_inflight(xb_batch* b, int inflight);
/* Form (XB_FORM_*) of every pair of the next run. */
int xb_batch_set_form(xb_batch* b, int form);
/* Autotuner: for each size bucket present among the given device-resident
 * pairs (max(w, h) <= 1024, <= 2048, <= 3072, larger), times the bucket's
 * pairs at 1, 2, 4, 6, 8, 12, 16, 24 and 32 pairs in flight (up to the
 * workers; best of `reps` runs after a warm-up, a larger count must win by
 * 3 %) and keeps the fastest.  In later runs one pool of workers takes the
 * pairs largest first and starts a pair only while fewer pairs run than its
 * bucket's count.  table[4] (may be NULL) receives the counts (0 for
 * an absent bucket).  The timing runs are ordinary runs: outputs are written. */
int xb_batch_autotune(xb_batch* b, const xb_params* p, int n, const int* w, const int* h,
                      const void* const* d_tmpl, const void* const* d_sci, void* const* d_conv,
                      void* const* d_diff, int reps, int* table);
/* Screen model of every worker (as xb_screen_load). */
int xb_batch_screen_load(xb_batch* b, const float* model);
/* Launch shapes used for pairs of size bucket `bucket` (0..3, as the
 * autotuner's table): the defaults until xb_batch_autotune has tuned that
 * bucket (after its pairs-in-flight count, it tries each shape value in turn
 * at that count, keeping one only when it is faster by 3 %), or
 * xb_batch_set_shapes set them. */
int xb_batch_get_shapes(xb_batch* b, int bucket, xb_shapes* s);
/* The same shapes for every bucket (s == NULL: a batch's defaults, which are a
 * context's but for stats_threads 512). */
int xb_batch_set_shapes(xb_batch* b, const xb_shapes* s);

#ifdef __cplusplus
}
#endif
#endif /* XBACH_H */
