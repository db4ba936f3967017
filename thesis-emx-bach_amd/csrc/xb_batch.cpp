// xb_batch.cpp — many independent pairs through the pipeline at once.
//
// The reference runs one pair per process (src/main.cpp:16-159), and a single
// pair cannot fill an MI355X: the global FP64 LU (src/bachUtil.cpp:319-411) is
// a chain of ~300 dependent pivot steps on a handful of CUs, and the clip-refit
// loop (src/cdkscUtil.cpp:585-684) stops at a host read-back every pass.  The
// batch runtime keeps several pairs in flight instead: each worker is a host
// thread that owns one context (one HIP stream, its own resident buffers) and
// pulls pairs from a shared queue, largest first, so one pair's latency-bound
// stages overlap the other pairs' work.  A pair's arithmetic is exactly the
// single-pair pipeline's, so every output is bit-identical to xb_pipeline_run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <condition_variable>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "xb_internal.h"

constexpr int XB_NBUCKET = 4;  // pair sizes: max(w, h) <= 1024, 2048, 3072, larger

struct xb_batch {
  int device = 0;
  int inflight = 0;                    // 0: all workers
  int tuned[XB_NBUCKET] = {0, 0, 0, 0};  // pairs in flight per size bucket (0: untuned)
  xb_shapes shapes[XB_NBUCKET] = {xb::default_batch_shapes(), xb::default_batch_shapes(),
                                  xb::default_batch_shapes(), xb::default_batch_shapes()};
  std::vector<xb_ctx*> ctx;
  std::string err;
  // lockstep mode (XB_BATCH_LOCKSTEP): the batched solves' stream and events
  bool lockstep = false;
  hipStream_t sstream = nullptr;
  hipEvent_t evSolve = nullptr;
  std::vector<hipEvent_t> evBuild;     // per context
  int form = 0;
  // lockstep: each batched solve (its systems, form and shape) is captured
  // into a hipGraph at its second occurrence and replayed from then on
  struct SolveGraph {
    xb::LuBatch key;
    bool ldlt, refine;
    int grp;
    hipGraphExec_t exec;  // null while seen once
  };
  std::vector<SolveGraph> graphs;
  int captureFailures = 0;  // failed captures (after two, solves launch directly)
  bool eagerSolves = false; // XB_CTX_EAGER among the flags: no graphs
};

namespace {

int sizeBucket(int w, int h) {
  const int m = std::max(w, h);
  return m <= 1024 ? 0 : m <= 2048 ? 1 : m <= 3072 ? 2 : 3;
}

double nowMs() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Job {
  const xb_params* p;
  int n;
  const int* w;
  const int* h;
  bool device;
  const void* const* tin;
  const void* const* sin;
  void* const* cout;
  void* const* dout;
  xb_pair_result* res;
  const xb_shapes* shapes = nullptr;  // launch shapes per size bucket
  std::vector<int> order;   // pair indices, largest first
  std::vector<int> status;  // per pair
  std::vector<char> ran;    // per pair: a worker took it
  std::vector<int> cap;     // per order position: pairs that may run at once
  int next = 0;             // next order position to hand out (under gate)
  int running = 0;          // pairs running now (under gate)
  std::mutex gate;
  std::condition_variable cv;
  std::mutex mu;
  int firstBad = -1;        // lowest failing pair index
  std::string firstMsg;
};

// One pair on one worker's context: the single-pair pipeline.  Device pairs
// are read and written in the caller's buffers directly.
int runPair(xb_ctx* x, const Job& J, int i, xb_pair_result* r) {
  xb::Ctx& c = x->c;
  c.shapes = J.shapes[sizeBucket(J.w[i], J.h[i])];
  if (J.device) {
    int rc = xb::pipelineRunInto(c, J.p, J.w[i], J.h[i], J.tin[i], J.sin[i],
                                 J.cout ? J.cout[i] : nullptr, J.dout ? J.dout[i] : nullptr);
    if (rc) return rc;
  } else {
    int rc = xb_pipeline_run(x, J.p, J.w[i], J.h[i], static_cast<const float*>(J.tin[i]),
                             static_cast<const float*>(J.sin[i]),
                             J.cout ? static_cast<float*>(J.cout[i]) : nullptr,
                             J.dout ? static_cast<float*>(J.dout[i]) : nullptr, nullptr);
    if (rc) return rc;
  }
  if (r) {
    r->convTemplate = c.convTemplate;
    r->kscIterations = c.kscIterations;
    r->kernSum = c.kernSum;
    r->captureFailures = c.captureFailures;
    r->screened = c.nScreened;
    r->screenRestored = c.nRestored;
  }
  return XB_OK;
}

void worker(xb_batch* b, int wi, Job* J) {
  xb_ctx* x = b->ctx[wi];
  // a worker that cannot select the device takes no pair; runBatch reports
  // the pairs no worker ran
  if (hipSetDevice(b->device) != hipSuccess) return;
  for (;;) {
    // the next pair starts once fewer pairs run than its size bucket allows
    int k;
    {
      std::unique_lock<std::mutex> g(J->gate);
      J->cv.wait(g, [J] { return J->next >= J->n || J->running < J->cap[J->next]; });
      if (J->next >= J->n) break;
      k = J->next++;
      J->running++;
    }
    const int i = J->order[k];
    xb_pair_result local{};
    xb_pair_result* r = J->res ? &J->res[i] : &local;
    const double t0 = nowMs();
    J->ran[i] = 1;
    const int rc = runPair(x, *J, i, r);
    r->status = rc;
    r->worker = wi;
    r->ms = nowMs() - t0;
    J->status[i] = rc;
    if (rc) {
      std::lock_guard<std::mutex> g(J->mu);
      if (J->firstBad < 0 || i < J->firstBad) {
        J->firstBad = i;
        J->firstMsg = "pair " + std::to_string(i) + ": " + x->c.err;
      }
      // a failed pair may leave work queued; start the next from a clean stream
      (void)hipStreamSynchronize(x->c.stream);
    }
    {
      std::lock_guard<std::mutex> g(J->gate);
      J->running--;
    }
    J->cv.notify_all();
  }
}

// ---- lockstep mode -----------------------------------------------------------
// One host thread runs a group of pairs together, each on its own context and
// stream, stage by stage: every global solve of the group (the direction
// trials of all pairs, then each clip-refit pass of the pairs still in their
// loop) is one batched solve (launch_global_solve_batch: 23 launches for all
// the group's systems) on the batch's solve stream, which waits for each
// pair's build and is waited for by each pair's next stage.  A pair's
// arithmetic is the single-pair pipeline's, so its outputs are bit-identical.
// Groups are consecutive pairs of the largest-first order, up to the size
// bucket's count (the autotuner's) and LU_BATCH_MAX / 2 pairs.
int lsSolve(xb_batch* b, const std::vector<int>& slots, const xb::LuBatch& B,
            std::string& err) {
  hipError_t e;
  for (int j : slots) {
    xb::Ctx& c = b->ctx[j]->c;
    if ((e = hipEventRecord(b->evBuild[j], c.stream)) != hipSuccess ||
        (e = hipStreamWaitEvent(b->sstream, b->evBuild[j], 0)) != hipSuccess) {
      err = std::string("lockstep solve: ") + hipGetErrorString(e);
      return XB_ERR_HIP;
    }
  }
  const bool ldlt = (b->form & XB_FORM_LDLT) != 0;
  const bool refine = (b->form & XB_FORM_REFINE) != 0;
  // the group's pairs share a size bucket's shapes (the first pair's)
  const int grp = slots.empty() ? 4 : b->ctx[slots[0]]->c.shapes.ldlt_group;
  // the batched solve's kernels (and each system's refinement on the solve
  // stream, its system set found among its pair's four: two trials, two
  // clip-refit directions)
  auto enqueue = [&]() -> hipError_t {
    hipError_t q = xb::launch_global_solve_batch(b->sstream, B, ldlt, refine, grp);
    if (q != hipSuccess || !refine || ldlt) return q;
    for (int k = 0; k < B.n; k++) {
      const xb::LuSet* L = nullptr;
      for (int j : slots) {
        xb::Ctx& c = b->ctx[j]->c;
        for (const xb::LuSet* r : {&c.lu[0], &c.lu[1], &c.ksLu[0], &c.ksLu[1]})
          if (r->M == B.s[k].M) L = r;
      }
      if (L && (q = xb::launch_refine(b->sstream, *L, B.s[k].out)) != hipSuccess) return q;
    }
    return hipSuccess;
  };
  // a solve seen before is replayed from its graph (captured at its second
  // occurrence: the first is launched directly, which also sets the kernels'
  // attributes outside any capture)
  xb_batch::SolveGraph* sg = nullptr;
  for (auto& g : b->graphs)
    if (g.key.n == B.n && g.ldlt == ldlt && g.refine == refine && g.grp == grp &&
        std::memcmp(g.key.s, B.s, sizeof(xb::LuSys) * B.n) == 0) {
      sg = &g;
      break;
    }
  constexpr size_t kMaxGraphs = 512;
  if (b->eagerSolves) {
    e = enqueue();
  } else if (!sg && b->graphs.size() < kMaxGraphs) {
    b->graphs.push_back({B, ldlt, refine, grp, nullptr});
    e = enqueue();
  } else if (sg && !sg->exec && b->captureFailures < 2) {
    hipGraph_t graph = nullptr;
    {
      // no thread allocates or frees while the capture is open
      std::shared_lock<std::shared_mutex> lk(xb::alloc_lock());
      e = hipStreamBeginCapture(b->sstream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        const hipError_t e1 = enqueue();
        const hipError_t e2 = hipStreamEndCapture(b->sstream, &graph);
        e = e1 != hipSuccess ? e1 : e2;
      }
    }
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess) {
      std::unique_lock<std::shared_mutex> lk(xb::alloc_lock());
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    }
    if (graph) (void)hipGraphDestroy(graph);
    if (e == hipSuccess) {
      sg->exec = exec;
      e = hipGraphLaunch(exec, b->sstream);
    } else {
      // nothing of the capture ran: end a capture left open, count the
      // failure, and launch the kernels directly
      b->captureFailures++;
      (void)hipGetLastError();
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(b->sstream, &cs) == hipSuccess &&
          cs != hipStreamCaptureStatusNone) {
        hipGraph_t left = nullptr;
        (void)hipStreamEndCapture(b->sstream, &left);
        if (left) (void)hipGraphDestroy(left);
        (void)hipGetLastError();
      }
      e = enqueue();
    }
  } else if (sg && sg->exec) {
    e = hipGraphLaunch(sg->exec, b->sstream);
  } else {
    e = enqueue();
  }
  if (e != hipSuccess) {
    err = std::string("batched solve: ") + hipGetErrorString(e);
    return XB_ERR_HIP;
  }
  if ((e = hipEventRecord(b->evSolve, b->sstream)) != hipSuccess) {
    err = std::string("lockstep solve: ") + hipGetErrorString(e);
    return XB_ERR_HIP;
  }
  for (int j : slots)
    if ((e = hipStreamWaitEvent(b->ctx[j]->c.stream, b->evSolve, 0)) != hipSuccess) {
      err = std::string("lockstep solve: ") + hipGetErrorString(e);
      return XB_ERR_HIP;
    }
  return XB_OK;
}

void lsGroup(xb_batch* b, Job& J, const std::vector<int>& pairs) {
  const int g = (int)pairs.size();
  std::vector<int> rc(g, XB_OK);
  std::vector<char> live(g, 1);
  const double t0 = nowMs();
  auto fail = [&](int j, int code) {
    rc[j] = code;
    live[j] = 0;
    (void)hipStreamSynchronize(b->ctx[j]->c.stream);
  };
  auto liveSlots = [&]() {
    std::vector<int> v;
    for (int j = 0; j < g; j++)
      if (live[j]) v.push_back(j);
    return v;
  };
  // init + sss
  for (int j = 0; j < g; j++) {
    const int i = pairs[j];
    xb_ctx* x = b->ctx[j];
    x->c.shapes = J.shapes[sizeBucket(J.w[i], J.h[i])];
    int r = J.device ? xb_init_device(x, J.p, J.w[i], J.h[i], J.tin[i], J.sin[i])
                     : xb_init(x, J.p, J.w[i], J.h[i], static_cast<const float*>(J.tin[i]),
                               static_cast<const float*>(J.sin[i]));
    if (!r) r = xb::ls_sss_begin(x->c);
    if (r) fail(j, r);
  }
  for (bool again = true; again;) {
    again = false;
    for (int j : liveSlots()) {
      bool more = false;
      const int r = xb::ls_sss_next(b->ctx[j]->c, &more);
      if (r) fail(j, r);
      again = again || more;
    }
  }
  // cmv, then the direction trials as one batched solve
  xb::LuBatch B;
  B.n = 0;
  std::vector<int> slots;
  for (int j : liveSlots()) {
    int r = xb::ls_cmv(b->ctx[j]->c);
    if (!r) r = xb::ls_cd_build(b->ctx[j]->c, B);
    if (r) fail(j, r);
    else slots.push_back(j);
  }
  std::string err;
  if (!slots.empty() && lsSolve(b, slots, B, err) != XB_OK) {
    for (int j : slots) {
      b->ctx[j]->c.err = err;
      fail(j, XB_ERR_HIP);
    }
  }
  for (int j : liveSlots()) {
    int r = xb::ls_cd_post(b->ctx[j]->c);
    if (r) fail(j, r);
  }
  for (int j : liveSlots()) {
    int r = xb::ls_cd_decide(b->ctx[j]->c);
    if (r) fail(j, r);
  }
  // the clip-refit passes of the pairs still in their loop, one batched solve each
  std::vector<char> inLoop(live.begin(), live.end());
  for (int it = 0;; it++) {
    B.n = 0;
    slots.clear();
    for (int j = 0; j < g; j++) {
      if (!live[j] || !inLoop[j]) continue;
      const int r = xb::ls_ksc_build(b->ctx[j]->c, B);
      if (r) fail(j, r);
      else slots.push_back(j);
    }
    if (slots.empty()) break;
    if (lsSolve(b, slots, B, err) != XB_OK) {
      for (int j : slots) {
        b->ctx[j]->c.err = err;
        fail(j, XB_ERR_HIP);
      }
      break;
    }
    for (int j : slots) {
      const int r = xb::ls_ksc_post(b->ctx[j]->c, it);
      if (r) fail(j, r);
    }
    for (int j : slots) {
      if (!live[j]) continue;
      bool more = false;
      const int r = xb::ls_ksc_check(b->ctx[j]->c, it + 1, &more);
      if (r) fail(j, r);
      else if (!more) inLoop[j] = 0;
    }
  }
  // conv, sub and the outputs
  for (int j : liveSlots()) {
    const int r = xb::ls_finish(b->ctx[j]->c);
    if (r) fail(j, r);
  }
  for (int j : liveSlots()) {
    const int i = pairs[j];
    xb::Ctx& c = b->ctx[j]->c;
    const size_t bytes = (size_t)J.w[i] * J.h[i] * sizeof(float);
    const hipMemcpyKind kind = J.device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    hipError_t e = hipSuccess;
    if (J.cout && J.cout[i]) e = hipMemcpyAsync(J.cout[i], c.conv, bytes, kind, c.stream);
    if (e == hipSuccess && J.dout && J.dout[i])
      e = hipMemcpyAsync(J.dout[i], c.diff, bytes, kind, c.stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(&c.kernSum, &c.sc->kernSum, sizeof(double), hipMemcpyDeviceToHost,
                         c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) {
      c.err = std::string("lockstep outputs: ") + hipGetErrorString(e);
      fail(j, XB_ERR_HIP);
    }
  }
  const double ms = nowMs() - t0;
  for (int j = 0; j < g; j++) {
    const int i = pairs[j];
    xb::Ctx& c = b->ctx[j]->c;
    J.ran[i] = 1;
    J.status[i] = rc[j];
    if (J.res) {
      xb_pair_result& r = J.res[i];
      r = xb_pair_result{};
      r.status = rc[j];
      r.worker = j;
      r.ms = ms;
      if (!rc[j]) {
        r.convTemplate = c.convTemplate;
        r.kscIterations = c.kscIterations;
        r.kernSum = c.kernSum;
        r.captureFailures = c.captureFailures + b->captureFailures;
      }
    }
    if (rc[j] && (J.firstBad < 0 || i < J.firstBad)) {
      J.firstBad = i;
      J.firstMsg = "pair " + std::to_string(i) + ": " + c.err;
    }
  }
}

void runLockstep(xb_batch* b, Job& J) {
  if (hipSetDevice(b->device) != hipSuccess) return;  // pairs reported as not run
  const int maxGroup = std::min<int>((int)b->ctx.size(), xb::LU_BATCH_MAX / 2);
  for (int k = 0; k < J.n;) {
    const int cap = std::max(1, std::min(J.cap[k], maxGroup));
    std::vector<int> pairs;
    for (; k < J.n && (int)pairs.size() < cap; k++) pairs.push_back(J.order[k]);
    lsGroup(b, J, pairs);
  }
}

int runBatch(xb_batch* b, Job& J) {
  if (!b) return XB_ERR_ARG;
  if (!J.p || J.n < 0 || (J.n > 0 && (!J.w || !J.h || !J.tin || !J.sin))) {
    b->err = "bad arguments";
    return XB_ERR_ARG;
  }
  for (int i = 0; i < J.n; i++) {
    if (J.w[i] <= 0 || J.h[i] <= 0 || !J.tin[i] || !J.sin[i]) {
      b->err = "pair " + std::to_string(i) + ": bad size or input pointer";
      return XB_ERR_ARG;
    }
  }
  if (J.n == 0) return XB_OK;
  J.shapes = b->shapes;
  J.order.resize(J.n);
  std::iota(J.order.begin(), J.order.end(), 0);
  // largest pairs first (longest-processing-time order balances the workers)
  std::stable_sort(J.order.begin(), J.order.end(), [&](int a, int c) {
    return (long long)J.w[a] * J.h[a] > (long long)J.w[c] * J.h[c];
  });
  J.status.assign(J.n, XB_ERR_HIP);
  J.ran.assign(J.n, 0);
  const int all = b->inflight > 0 ? std::min<int>(b->inflight, (int)b->ctx.size())
                                  : (int)b->ctx.size();
  // pairs in flight: the autotuner's count for the pair's size bucket, else
  // the batch-wide setting; one pool of workers serves every bucket, so the
  // count changes as the queue moves from large pairs to small ones
  J.cap.resize(J.n);
  int nw = 1;
  for (int k = 0; k < J.n; k++) {
    const int t = b->tuned[sizeBucket(J.w[J.order[k]], J.h[J.order[k]])];
    J.cap[k] = t > 0 ? std::min<int>(t, (int)b->ctx.size()) : all;
    nw = std::max(nw, J.cap[k]);
  }
  nw = std::min(nw, J.n);
  if (b->lockstep && (b->form & (XB_FORM_SCREEN | XB_FORM_RECORD)) == 0) {
    runLockstep(b, J);
  } else {
    std::vector<std::thread> th;
    th.reserve(nw);
    for (int wi = 0; wi < nw; wi++) th.emplace_back(worker, b, wi, &J);
    for (auto& t : th) t.join();
  }
  // pairs no worker took (every worker failed to select the device)
  for (int i = 0; i < J.n; i++) {
    if (J.ran[i]) continue;
    if (J.res) {
      J.res[i] = xb_pair_result{};
      J.res[i].status = XB_ERR_HIP;
      J.res[i].worker = -1;
    }
    if (J.firstBad < 0 || i < J.firstBad) {
      J.firstBad = i;
      J.firstMsg = "pair " + std::to_string(i) + ": no worker could select device " +
                   std::to_string(b->device);
    }
  }
  if (J.firstBad >= 0) {
    b->err = J.firstMsg;
    return J.status[J.firstBad];
  }
  b->err.clear();
  return XB_OK;
}

}  // namespace

extern "C" {

int xb_batch_create(int device, int workers, int flags, xb_batch** out) {
  if (!out || workers < 1) return XB_ERR_ARG;
  *out = nullptr;
  if (workers > XB_BATCH_MAX_WORKERS) return XB_ERR_CAPACITY;
  xb_batch* b = new xb_batch();
  b->device = device;
  for (int i = 0; i < workers; i++) {
    xb_ctx* x = nullptr;
    // worker threads never capture graphs: a capture open on one thread's
    // stream has been invalidated by other threads' work on the device, and
    // the graphs measured no faster for a batch than direct launches
    const int rc = xb_ctx_create(device, flags | XB_CTX_ONE_STREAM | XB_CTX_EAGER, &x);
    if (rc) {
      for (auto* y : b->ctx) xb_ctx_destroy(y);
      delete b;
      return rc;
    }
    b->ctx.push_back(x);
  }
  b->eagerSolves = (flags & XB_CTX_EAGER) != 0;
  if (flags & XB_BATCH_LOCKSTEP) {
    b->lockstep = true;
    b->evBuild.assign(workers, nullptr);
    bool ok = hipSetDevice(device) == hipSuccess &&
              hipStreamCreateWithFlags(&b->sstream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->evSolve, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < workers; i++)
      ok = hipEventCreateWithFlags(&b->evBuild[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      xb_batch_destroy(b);
      return XB_ERR_HIP;
    }
  }
  *out = b;
  return XB_OK;
}

int xb_batch_destroy(xb_batch* b) {
  if (!b) return XB_OK;
  if (b->sstream) (void)hipStreamSynchronize(b->sstream);
  for (auto& g : b->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  for (auto* x : b->ctx) xb_ctx_destroy(x);
  for (auto e : b->evBuild)
    if (e) (void)hipEventDestroy(e);
  if (b->evSolve) (void)hipEventDestroy(b->evSolve);
  if (b->sstream) (void)hipStreamDestroy(b->sstream);
  delete b;
  return XB_OK;
}

const char* xb_batch_last_error(xb_batch* b) { return b ? b->err.c_str() : "no batch"; }

int xb_batch_set_inflight(xb_batch* b, int inflight) {
  if (!b || inflight < 0) return XB_ERR_ARG;
  if (inflight > (int)b->ctx.size()) {
    b->err = "pairs in flight exceed the batch's " + std::to_string(b->ctx.size()) + " workers";
    return XB_ERR_CAPACITY;
  }
  b->inflight = inflight;
  std::fill(b->tuned, b->tuned + XB_NBUCKET, 0);
  return XB_OK;
}

int xb_batch_autotune(xb_batch* b, const xb_params* p, int n, const int* w, const int* h,
                      const void* const* d_tmpl, const void* const* d_sci, void* const* d_conv,
                      void* const* d_diff, int reps, int* table) {
  if (!b || reps < 1) return XB_ERR_ARG;
  std::fill(b->tuned, b->tuned + XB_NBUCKET, 0);
  const int workers = (int)b->ctx.size();
  const int cands[] = {1, 2, 4, 6, 8, 12, 16, 24, 32};
  for (int bk = 0; bk < XB_NBUCKET; bk++) {
    std::vector<int> idx;
    for (int i = 0; i < n; i++)
      if (sizeBucket(w[i], h[i]) == bk) idx.push_back(i);
    if (idx.empty()) continue;
    std::vector<int> bw, bh;
    std::vector<const void*> bt, bs;
    std::vector<void*> bc, bd;
    for (int i : idx) {
      bw.push_back(w[i]);
      bh.push_back(h[i]);
      bt.push_back(d_tmpl[i]);
      bs.push_back(d_sci[i]);
      bc.push_back(d_conv ? d_conv[i] : nullptr);
      bd.push_back(d_diff ? d_diff[i] : nullptr);
    }
    double best = 1e300;
    int bestC = 1;
    const int saved = b->inflight;
    // best of reps timed runs after a warm-up run, at b's current settings
    auto timeRuns = [&](double& t) -> int {
      t = 1e300;
      for (int r = 0; r < reps + 1; r++) {
        const double t0 = nowMs();
        const int rc = xb_batch_run_device(b, p, (int)idx.size(), bw.data(), bh.data(), bt.data(),
                                           bs.data(), bc.data(), bd.data(), nullptr);
        if (rc) return rc;
        if (r > 0) t = std::min(t, nowMs() - t0);
      }
      return XB_OK;
    };
    b->shapes[bk] = xb::default_batch_shapes();
    for (int c : cands) {
      if (c > workers) break;
      b->inflight = c;
      double t = 0.0;
      const int rc = timeRuns(t);
      if (rc) {
        b->inflight = saved;
        return rc;
      }
      if (t < best * 0.97) {  // a larger count must win by 3 %
        best = t;
        bestC = c;
      }
      if (c >= (int)idx.size()) break;
    }
    // launch shapes at that count, one knob at a time from the defaults: a
    // value other than the current one must win by 3 %
    b->inflight = bestC;
    b->shapes[bk] = xb::default_batch_shapes();
    struct Knob {
      int32_t xb_shapes::*field;
      std::vector<int> values;
    };
    std::vector<Knob> knobs = {{&xb_shapes::conv_tiles, {1, 2, 4, 8}},
                               {&xb_shapes::mask_rows, {16, 32, 64}},
                               {&xb_shapes::find_staged, {0, 1}},
                               {&xb_shapes::stats_threads, {256, 512, 1024}}};
    if (b->form & XB_FORM_LDLT) knobs.push_back({&xb_shapes::ldlt_group, {2, 4}});
    for (const Knob& k : knobs) {
      const int cur = b->shapes[bk].*k.field;
      for (int v : k.values) {
        if (v == cur) continue;
        const xb_shapes keep = b->shapes[bk];
        b->shapes[bk].*k.field = v;
        double t = 0.0;
        const int rc = timeRuns(t);
        if (rc) {
          b->inflight = saved;
          b->shapes[bk] = keep;
          return rc;
        }
        if (t < best * 0.97) best = t;
        else b->shapes[bk] = keep;
      }
    }
    b->inflight = saved;
    b->tuned[bk] = bestC;
  }
  if (table)
    for (int k = 0; k < XB_NBUCKET; k++) table[k] = b->tuned[k];
  return XB_OK;
}

int xb_batch_get_shapes(xb_batch* b, int bucket, xb_shapes* s) {
  if (!b || !s || bucket < 0 || bucket >= XB_NBUCKET) return XB_ERR_ARG;
  *s = b->shapes[bucket];
  return XB_OK;
}

int xb_batch_set_shapes(xb_batch* b, const xb_shapes* s) {
  if (!b) return XB_ERR_ARG;
  xb_shapes v = s ? *s : xb::default_batch_shapes();
  if (!xb::shapes_valid(v)) return XB_ERR_ARG;
  v.pad = 0;
  for (auto& q : b->shapes) q = v;
  return XB_OK;
}

int xb_batch_set_form(xb_batch* b, int form) {
  if (!b) return XB_ERR_ARG;
  b->form = form;
  for (auto* x : b->ctx) {
    const int rc = xb_set_form(x, form);
    if (rc) return rc;
  }
  return XB_OK;
}

int xb_batch_screen_load(xb_batch* b, const float* model) {
  if (!b) return XB_ERR_ARG;
  for (auto* x : b->ctx) {
    const int rc = xb_screen_load(x, model);
    if (rc) return rc;
  }
  return XB_OK;
}

int xb_batch_run(xb_batch* b, const xb_params* p, int n, const int* w, const int* h,
                 const float* const* tmpl, const float* const* sci, float* const* conv,
                 float* const* diff, xb_pair_result* results) {
  Job J;
  J.p = p;
  J.n = n;
  J.w = w;
  J.h = h;
  J.device = false;
  J.tin = reinterpret_cast<const void* const*>(tmpl);
  J.sin = reinterpret_cast<const void* const*>(sci);
  J.cout = reinterpret_cast<void* const*>(conv);
  J.dout = reinterpret_cast<void* const*>(diff);
  J.res = results;
  return runBatch(b, J);
}

int xb_batch_run_device(xb_batch* b, const xb_params* p, int n, const int* w, const int* h,
                        const void* const* d_tmpl, const void* const* d_sci,
                        void* const* d_conv, void* const* d_diff, xb_pair_result* results) {
  Job J;
  J.p = p;
  J.n = n;
  J.w = w;
  J.h = h;
  J.device = true;
  J.tin = d_tmpl;
  J.sin = d_sci;
  J.cout = d_conv;
  J.dout = d_diff;
  J.res = results;
  return runBatch(b, J);
}

}  // extern "C"
