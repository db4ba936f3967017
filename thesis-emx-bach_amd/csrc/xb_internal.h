// xb_internal.h — shared definitions of libxbach (MI355X / gfx950).
//
// The device path is specialised for the reference's fixed method constants
// (include/argsUtil.h:29-56): 21x21 kernel, 31x31 substamps, three Gaussians of
// degrees 6/4/2 (49 basis terms), kernel order 2 (6 spatial terms), background
// order 1 (3 terms), 3 substamps kept of 6 candidates.  xb_init rejects other
// values with XB_ERR_ARG.  Everything the reference decides from data (stamp
// grid, thresholds, sigma-clip constants) stays a runtime parameter.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "../../include/xbach.h"

namespace xb {

// The launch shapes of kernels whose tiling is free (xb_set_shapes): every
// shape gives the same results bit for bit.
inline xb_shapes default_shapes() {
  xb_shapes s;
  s.conv_tiles = 4;
  s.mask_rows = 64;
  s.ldlt_group = 4;
  s.find_staged = 1;
  s.stats_threads = 1024;
  s.pad = 0;
  return s;
}
// A batch's defaults: the same but for the stamp statistics, which run with
// 512 threads per stamp (faster when many pairs share the GPU; 1024 is faster
// for a lone pair).
inline xb_shapes default_batch_shapes() {
  xb_shapes s = default_shapes();
  s.stats_threads = 512;
  return s;
}
inline bool shapes_valid(const xb_shapes& s) {
  const bool ct = s.conv_tiles == 1 || s.conv_tiles == 2 || s.conv_tiles == 4 ||
                  s.conv_tiles == 8;
  const bool mr = s.mask_rows == 16 || s.mask_rows == 32 || s.mask_rows == 64;
  const bool lg = s.ldlt_group == 2 || s.ldlt_group == 4;
  const bool fs = s.find_staged == 0 || s.find_staged == 1;
  const bool st = s.stats_threads == 256 || s.stats_threads == 512 || s.stats_threads == 1024;
  return ct && mr && lg && fs && st;
}


// A HIP allocation or release made while another thread has a stream capture
// open can invalidate that capture and leave its stream failing every later
// call (hipFree synchronizes the device).  Every allocation, release and
// graph instantiation takes this lock exclusively; an open capture holds it
// shared, so the captures of different contexts still overlap.
std::shared_mutex& alloc_lock();
template <class T>
inline hipError_t locked_malloc(T** p, size_t bytes) {
  std::unique_lock<std::shared_mutex> g(alloc_lock());
  return hipMalloc((void**)p, bytes);
}
inline hipError_t locked_free(void* p) {
  std::unique_lock<std::shared_mutex> g(alloc_lock());
  return hipFree(p);
}
inline hipError_t locked_host_malloc(void** p, size_t bytes) {
  std::unique_lock<std::shared_mutex> g(alloc_lock());
  return hipHostMalloc(p, bytes, hipHostMallocDefault);
}
inline hipError_t locked_host_free(void* p) {
  std::unique_lock<std::shared_mutex> g(alloc_lock());
  return hipHostFree(p);
}

constexpr int KW = 21;             // fKernelWidth
constexpr int HKW = 10;            // hKernelWidth
constexpr int SSW = 31;            // fSStampWidth
constexpr int HSS = 15;            // hSStampWidth
constexpr int NPIX = SSW * SSW;    // 961 substamp pixels
constexpr int NPSF = 49;           // basis terms
constexpr int NBG = 3;             // background terms
constexpr int WROWS = NPSF + NBG;  // 52 rows of W
constexpr int NC1 = NPSF - 1;      // 48
constexpr int NC2 = 6;             // spatial terms of kernel order 2
constexpr int NCOMP = NC1 * NC2;   // 288
constexpr int MATSIZE = NCOMP + NBG + 1;  // 292 unknowns (1-based)
constexpr int MATN = MATSIZE + 1;         // 293 rows/cols stored
constexpr int NSOL = NPSF * NC2 + NBG + 1;  // 298 solution slots
// LU bookkeeping block: a 64-byte state record, then one zero-row flag per row
// a 64-byte state record, the zero-row flags, and one row map per 16-column
// panel (19 for 292 unknowns)
constexpr int LU_PANELS = (MATSIZE + 15) / 16;
constexpr size_t LU_STATE_BYTES = 64 + sizeof(int) * MATN * (1 + LU_PANELS);
// XB_FORM_LDLT's workspace: L and D of the system padded to 19 blocks of 16
// (0-based, L also transposed above the diagonal), then the 190 lower 16 x 16
// tiles of the trailing matrix in the factorisation's lane order
constexpr int LDLT_WORK_DOUBLES =
    (LU_PANELS * 16) * (LU_PANELS * 16) + LU_PANELS * (LU_PANELS + 1) / 2 * 256;
constexpr int QN = NPSF + 2;              // 51: per-stamp test system (1-based)
constexpr int MAXSS = 6;                  // 2 * maxKSStamps
constexpr int KEEPSS = 3;                 // maxKSStamps
constexpr int TW = 2 * (HSS + HKW) + 1;   // 51: Y-pass window width
constexpr int TH = SSW;                   // 31: Y-pass rows
constexpr int KK = KW * KW;               // 441
constexpr int NSAMP = 100, NSAMP_PAD = 128;

// Method parameters the kernels read (derived from xb_params).
struct Method {
  double threshLow, threshHigh, threshKernFit, sigKernFit, sigClipAlpha, iqRange;
  int normalizeTemplate;
};

// Basis tables built once on the host in FP64 (glibc exp; bg rounded to float
// as include/argsUtil.h:49-53) and uploaded, so the fit (cl_kern/cmv.cl:1-68)
// and the convolution (include/datatypeUtil.h:62-118) use identical values.
struct Basis {
  double fx[NPSF * KW];
  double fy[NPSF * KW];
  double vec[NPSF * KK];
  int kx[NPSF], ky[NPSF];
  int evenSub[NPSF];  // 1 when W[n] -= W[0] (even-even, n > 0; cmv.cl:137-154)
};

// Per-image stamp arrays on the device.  Grid order for the sss results,
// "kept" (compacted, ascending grid index) order for the fit data.
struct DevStamps {
  double* sky = nullptr;   // [N]
  double* fwhm = nullptr;  // [N]
  int* found = nullptr;    // [N]
  int* count = nullptr;    // [N] min(found, 3)
  int* ssx = nullptr;      // [N*MAXSS]
  int* ssy = nullptr;
  double* ssv = nullptr;
  int* kept = nullptr;     // [N] grid index of kept stamp k
  int* cur = nullptr;      // [N] current substamp of kept stamp k
  double* W = nullptr;     // [N][WROWS][NPIX] (raw rows; even rows minus row 0 on use)
  double* Wc = nullptr;    // the same rows in column-major pixel order (target products)
  double* G = nullptr;     // [N][WROWS][WROWS] Gram of final W rows (lower part)
  double* Bc = nullptr;    // [N][WROWS] W_r . target, column-major pixel order
  double* Br = nullptr;    // [N][NBG]   bg rows . target, row-major pixel order
  double* sig = nullptr;   // [N] cd: chi2 of the test stamps (list order)
  double* ksum = nullptr;  // [N] per-stamp test-fit kernel sums
  int* list = nullptr;     // [N] cd: test-stamp list
  double* sigK = nullptr;  // [N] ksc: chi2 per kept stamp of the current pass
  int* listK = nullptr;    // [N] ksc: refill list of the current pass
};

// Scalars of one direction's clip-refit loop (ksc), indexed by the image it
// convolves: each direction has its own, so both loops can run at once.
struct KscScalars {
  int check;            // the pass replaced a substamp
  int nList;            // entries of the refill list
  int nChi2;            // stamps with a valid chi2 in the pass
  int nScreened;        // substamps the learned screen replaced
  int nVerify;          // screened substamps probed by this verification round
  int nRestoredRound;   // ... of which the chi2 test kept (restored)
  int nRestored;        // restored screened substamps, whole loop
  int nConfirmed;       // screened substamps the chi2 test also rejects
  double clipMean, clipStd;
  double firstClip[2];  // clip mean and sd of the first pass
};

// Small device-side scalars of one pass (one struct, one allocation).
struct DevScalars {
  int nKept[2];
  int nList[2];         // cd: entries of the test-stamp list, per image
  int nTest[2];
  int retry;            // sss: refill ratio below 10 %
  int status;           // device-detected error
  double merit[2];
  double kernelSumCd[2];
  double kernSum;       // kernel sum at the image centre (conv)
  KscScalars k[2];      // ksc, per convolved image
};

// The global FP64 system of one fit and its LU work buffers.
struct LuSet {
  double* M = nullptr;          // [MATN*MATN] global matrix
  double* rhs = nullptr;        // [NSOL] right-hand side, then the solution
  double* M0 = nullptr;         // copies of the system as built (LU replay paths)
  double* rhs0 = nullptr;
  int* luIndex = nullptr;       // [MATN] pivot rows
  void* luState = nullptr;      // LU bookkeeping (zero row, last pivot, ii, zero flags)
  double* luVV = nullptr;       // [MATN]
  double* rr = nullptr;         // [NSOL] iterative refinement: residual, then correction
  double* F = nullptr;          // [LDLT_WORK_DOUBLES] L D L^T factors (XB_FORM_LDLT)
};

// The learned substamp screen's per-pass buffers and model (the first
// SCR_PASSES passes of the clip-refit loop are screened and recorded).
constexpr int SCR_PASSES = 16;
struct ScreenBufs {
  float* feat = nullptr;    // [pass][N][32] features of each pass's substamps
  int* passCur = nullptr;   // [pass][N] substamp of each stamp in that pass
  int* label = nullptr;     // [pass][N] 2 replaced by the pass, 1 later, 0 kept, -1 none
  float* prob = nullptr;    // [pass][N] p(reject)
  uint8_t* flag = nullptr;  // [N] screened out in the current pass
  uint8_t* bits = nullptr;  // [N] bit c: substamp c was screened out, not yet re-checked
  int* vOrig = nullptr;     // [N] substamp before a re-check probe
  int* vList = nullptr;     // [N] stamps probed by a re-check round
  double* vSig = nullptr;   // [N] their chi2 under the final fit (list order)
  uint8_t* finalRej = nullptr;  // [N][MAXSS] recording: the converged fit rejects substamp c
  float* zPrev = nullptr;   // [N] chi2 z-score of the previous pass
  float* model = nullptr;   // [XB_MLP_NPARAM]
  int cap = 0;
  bool haveModel = false;
};

// NaN marks of a clip-refit loop that runs before the direction is known
// (one spare mask bit per convolved image); the convolution keeps the chosen
// direction's as XB_MASK_NAN_PIXEL and drops both.
constexpr uint16_t MASK_PEND0 = 1u << 13;
__host__ __device__ constexpr uint16_t mask_pend(int im) { return (uint16_t)(MASK_PEND0 << im); }
static_assert((XB_MASK_ALL & (MASK_PEND0 | (MASK_PEND0 << 1))) == 0, "spare mask bits");

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;  // primary stream: every host read-back syncs it
  hipStream_t side = nullptr;    // second stream for the other image's independent work
  hipStream_t cur = nullptr;     // stream the launchers use (stream or side)
  // the two directions' clip-refit loops, run beside cd when the pipeline
  // runs whole (multi-stream contexts; see stageCdKsc)
  hipStream_t kstream[2] = {nullptr, nullptr};
  hipEvent_t evCdSig[2] = {nullptr, nullptr};  // cd's last read of a direction's fit data
  hipEvent_t evK[2] = {nullptr, nullptr};      // a ksc pass's check flag has arrived
  hipEvent_t evKEnd[2] = {nullptr, nullptr};   // a direction's loop has ended
  hipEvent_t evMerit = nullptr;                // both cd merits have arrived
  int specRestore = -1;         // image whose fit data a restore is refilling (-1: none)
  // captured global solves, one per (system, output, skip flag); buffers are
  // allocated once per context, so an entry stays valid until freeAll
  struct SolveGraph {
    const double* M;
    const double* out;
    const int* skip;
    hipGraphExec_t exec;
  };
  std::vector<SolveGraph> solveGraphs;
  bool eagerSolve = false;       // XB_CTX_EAGER: no capture
  int captureFailures = 0;       // failed captures of a solve (two stop capturing)
  hipEvent_t evFork = nullptr, evJoin = nullptr;
  std::string err;
  char name[256] = {0};
  // problem
  xb_params p{};
  Method m{};
  int w = 0, h = 0;
  int stampsx = 0, stampsy = 0, fsw = 0, nGrid = 0;
  size_t capPix = 0;
  int capStamps = 0, capFsw = 0;
  // device memory
  float* img[2] = {nullptr, nullptr};  // 0 template, 1 science (FP32)
  uint16_t* mask = nullptr;
  float* conv = nullptr;
  float* diff = nullptr;
  DevStamps st[2];
  DevScalars* sc = nullptr;
  DevScalars* hsc = nullptr;     // pinned host mirror for the scalar read-backs
  DevScalars* hscDev = nullptr;  // the same memory as the device addresses it
                                 // (kernels publish a pass's scalars there)
  hipEvent_t evSync = nullptr;   // completion of the read-backs (polled, no sleep)
  Basis* basis = nullptr;
  double* ran1 = nullptr;       // ran1 draws (identical per stamp: seed -666)
  int ran1Count = 0;
  float* goodPix = nullptr;     // [2][N][fsw*fsw] good-pixel / candidate scratch of sss
  LuSet lu[2];                  // cd: one global system per direction (both run at once)
  LuSet ksLu[2];                // ksc: the clip-refit system of each direction
  double* ksSol = nullptr;      // [2][NSOL] ksc solution of each direction
  double* sol = nullptr;        // [NSOL] xb_lu_solve's output
  double* cdSol = nullptr;      // [2][NSOL]
  int form = 0;                 // XB_FORM_* of the next run
  xb_shapes shapes = default_shapes();  // launch shapes (xb_set_shapes)
  ScreenBufs scr;
  int nScreened = 0;            // substamps the screen replaced (last run)
  int nRestored = 0;            // ... that the chi2 re-check kept (restored)
  int nConfirmed = 0;           // ... that the chi2 re-check also rejected
  int scrSamples = 0;           // samples the last ksc recorded
  void* mixTab = nullptr;       // mixed form: split BF16 2-D basis, MFMA fragment order
  void* mixTab16 = nullptr;     // ... split FP16, scaled by 2^mixScale16 (XB_FORM_FP16)
  int mixScale16 = 0;
  float mixBg[3] = {0, 0, 0};   // basis widths the table was built for
  xb_params basisParams{};      // parameters the device basis was built from
  bool basisValid = false;
  double* tileK64 = nullptr;    // [tiles][KK] FP64 tile kernels
  double* tileAks = nullptr;    // [tiles] sum |k| in the conv loop order
  int capTiles = 0;
  // host mirrors
  int nKept[2] = {0, 0};
  int convTemplate = 1;
  int tIdx = 0;                 // image being convolved
  double kernSum = 0.0;
  int kscIterations = 0;
  int retried = 0;
  bool haveResult = false;
  double stageMs[9] = {0};
  // per-kernel HIP-event timing (xb_profile_*)
  bool profOn = false;
  struct PendingEv {
    int slot;
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> evPool;
  std::vector<PendingEv> evPending;
  hipEvent_t evOpen[XB_PROF_N] = {};
  double profMs[XB_PROF_N] = {0};
  int profCnt[XB_PROF_N] = {0};
  // device address of the static gradient scratch of xb_mlp_grad_noalloc on
  // this context's device (looked up by xb_ctx_create)
  float* mlpScratch = nullptr;
};

// Bracket a launch group with HIP events on the context's stream (no-ops
// unless profiling is enabled).
void prof_begin(Ctx& c, int slot);
void prof_end(Ctx& c, int slot);

// ---- launchers (defined in the .hip files) --------------------------------
hipError_t launch_mask_input(Ctx& c);
hipError_t launch_sss_stats(Ctx& c, int im);
hipError_t launch_sss_find(Ctx& c, int im);
// both images' searches in one dispatch (image z = blockIdx.z)
hipError_t launch_sss_find_pair(Ctx& c);
hipError_t launch_sss_compact(Ctx& c);
hipError_t launch_reset_skip(Ctx& c);
// Fills the stamps of image im listed in d_list (d_nList entries on the
// device), or every kept stamp when d_list is null; nMax bounds the count.
hipError_t launch_fill(Ctx& c, int im, const int* d_list, const int* d_nList, int nMax);
// both images' full fills in one dispatch per kernel (image z = blockIdx.z)
hipError_t launch_fill_pair(Ctx& c);
hipError_t launch_stamp_lu(Ctx& c, int im);
hipError_t launch_cd_select(Ctx& c, int im);
hipError_t launch_global_build(Ctx& c, int im, bool cdForm);
// A batch of global systems solved together (launch_global_solve_batch): each
// system's buffers, passed by value in the kernel arguments.
struct LuState;
struct LuSys {
  double* M;
  double* vv;
  int* piv;
  double* rhs;
  LuState* st;
  const double* M0;
  const double* rhs0;
  double* out;
  const int* skip;
  double* F;
};
constexpr int LU_BATCH_MAX = 32;  // systems per batched launch (kernel argument space)
struct LuBatch {
  LuSys s[LU_BATCH_MAX];
  int n;
};
LuSys lu_sys(const LuSet& L, double* d_out, const int* d_skipIfZero);
// ldlt: the systems are solved by launch_ldlt_solve_batch (XB_FORM_LDLT),
// refined once when refine is set
// (ldltGroup: the L D L^T solve's load group, 2 or 4)
hipError_t launch_global_solve_batch(hipStream_t s, const LuBatch& B, bool ldlt = false,
                                     bool refine = false, int ldltGroup = 4);
// XB_FORM_LDLT (k_ldlt.hip): L D L^T without pivoting, with refine one step of
// FP64 iterative refinement; one workgroup and one launch per system; grp:
// trailing tiles per load group (2 or 4)
hipError_t launch_ldlt_solve(hipStream_t s, const LuSys& S, bool refine, int grp);
hipError_t launch_ldlt_solve_batch(hipStream_t s, const LuBatch& B, bool refine, int grp);
// Solves system set L.  d_skipIfZero: when it points at a zero count the solve
// is skipped and d_out is zeroed (the reference returns before solving an
// empty fit).
hipError_t launch_global_solve(Ctx& c, const LuSet& L, double* d_out, const int* d_skipIfZero);
hipError_t launch_cd_finish(Ctx& c, int im);
// Both images in one dispatch per kernel (one-stream contexts; both images
// have stamps): cd's stamp fits, test-stamp selection and global build; then,
// after the two solves, their chi2 and merits.
hipError_t launch_cd_begin_pair(Ctx& c);
hipError_t launch_cd_end_pair(Ctx& c);
// The two systems' solves as one batched launch sequence.
hipError_t launch_global_solve_pair(Ctx& c, const LuSet& L0, double* out0, const int* skip0,
                                    const LuSet& L1, double* out1, const int* skip1);
// cd form (useList): the test stamps' chi2 into st.sig; ksc form: every kept
// stamp's into st.sigK.  nanMark is the mask bit a NaN model pixel sets.
hipError_t launch_calc_sigs(Ctx& c, int im, bool useList, const double* d_sol,
                            uint16_t nanMark = XB_MASK_NAN_PIXEL);
// chi2 of the listed stamps (d_list, d_nList entries) into out (list order)
hipError_t launch_calc_sigs_list(Ctx& c, int im, const int* d_list, const int* d_nList,
                                 const double* d_sol, double* out);
// publish: also write the pass's check .. nScreened into the host mirror;
// label (with rec): the screen's per-stamp labels of the pass
hipError_t launch_ksc_check(Ctx& c, int im, bool first, int* rec = nullptr,
                            uint8_t* zeroBits = nullptr, bool publish = false,
                            int* label = nullptr);
hipError_t launch_conv(Ctx& c);
// Restores the fit data of image im to what cmv left (every kept stamp on its
// first substamp): stamps a clip-refit loop moved on are listed and refilled.
hipError_t launch_restore_fill(Ctx& c, int im);
// difference-image diagnostics: out[10] = good pixels (n, sum d, sum d^2,
// sum noise, sum (d/noise)^2), then the same over OK_CONV pixels
hipError_t launch_diag(Ctx& c, double* d_part, double* d_out);
constexpr int DIAG_BLOCKS = 512;
// learned screen (k_screen.hip): features of a pass's substamps, and with
// apply, the network's decision merged into the pass's refill list
hipError_t launch_screen(Ctx& c, int im, int pass, bool apply);
// Labels of the recorded samples; finalFit: by the recording's probes under
// the converged fit (launch_label_bits / launch_label_decide rounds)
hipError_t launch_screen_labels(Ctx& c, int im, int nRec, bool finalFit);
hipError_t launch_label_bits(Ctx& c, int im, int nRec);
hipError_t launch_label_decide(Ctx& c, int im);
// The chi2 re-check of screened substamps after the loop converged: select
// (each stamp's lowest screened substamp becomes current: probe), then, after
// the probes are filled and their chi2 formed, decide (the reference's reject
// rule with the last pass's clip: a kept probe stays, a rejected one returns
// its stamp to the substamp it had and lists it for a refill).
hipError_t launch_screen_verify_select(Ctx& c, int im);
hipError_t launch_screen_verify_decide(Ctx& c, int im);
// Device address of the static gradient scratch on `device` (k_screen.hip);
// looked up with hipGetSymbolAddress once per device and cached.
constexpr int XB_MAX_DEVICES = 64;
hipError_t scratch_address(int device, float** out);

// The pipeline in pieces for the batch runtime's lockstep mode (xb_api.cpp):
// sss (begin, then next until *more is false), cmv, cd (build: the trial
// systems appended to B; post after the batched solve; decide after its
// read-back), ksc passes (build: the system appended to B; post after the
// batched solve; check after its read-back), then conv and sub.
int ls_sss_begin(Ctx& c);
int ls_sss_next(Ctx& c, bool* more);
int ls_cmv(Ctx& c);
int ls_cd_build(Ctx& c, LuBatch& B);
int ls_cd_post(Ctx& c);
int ls_cd_decide(Ctx& c);
int ls_ksc_build(Ctx& c, LuBatch& B);
int ls_ksc_post(Ctx& c, int it);
int ls_ksc_check(Ctx& c, int it, bool* more);
int ls_finish(Ctx& c);
// the whole pipeline reading and writing the caller's device buffers
int pipelineRunInto(Ctx& c, const xb_params* p, int w, int h, const void* tin, const void* sin,
                    void* cout, void* dout);
// FP64 iterative refinement of system L's solution (XB_FORM_REFINE) on stream s
hipError_t launch_refine(hipStream_t s, const LuSet& L, double* d_out);

}  // namespace xb

// the opaque handle of the C ABI
struct xb_ctx {
  xb::Ctx c;
};
