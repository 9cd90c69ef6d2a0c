// extern "C" entry points of libbsmpo (declared in include/bsmpo.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"

namespace bsmpo {
int launch_observables(int M, int d, int chi, const double *Gamma, const double *Lambda,
                       const int32_t *charge, double *R, double *L, double *marg, double *trace,
                       hipStream_t stream);
int launch_entropy(int M, int chi, const double *Lambda, double *S, hipStream_t stream);
}

namespace bsmpo {
namespace {
struct TimingRec {
  int stage;
  hipEvent_t a, b;
  double flops, bytes;
  bool closed;
};
std::mutex g_tmu;
std::atomic<bool> g_timing{false};
std::vector<TimingRec> g_recs;
double g_ms[ST_COUNT], g_flops[ST_COUNT], g_bytes[ST_COUNT];
int64_t g_cnt[ST_COUNT];
}  // namespace

bool timing_on() { return g_timing; }

// Raises a kernel's dynamic-LDS limit once per (kernel, device), under a lock.
int ensure_lds_limit(const void *kernel, int bytes) {
  static std::mutex mu;
  struct Entry {
    const void *kernel;
    int dev, bytes;
  };
  static std::vector<Entry> done;
  int dev = 0;
  BSMPO_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  for (const Entry &e : done)
    if (e.kernel == kernel && e.dev == dev && e.bytes >= bytes) return 0;
  BSMPO_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  done.push_back(Entry{kernel, dev, bytes});
  return 0;
}

int timing_begin(int stage, hipStream_t stream) {
  if (!g_timing) return -1;
  TimingRec r{stage, nullptr, nullptr, 0.0, 0.0, false};
  if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return -1;
  hipEventRecord(r.a, stream);
  std::lock_guard<std::mutex> lk(g_tmu);
  g_recs.push_back(r);
  return (int)g_recs.size() - 1;
}

void timing_end(int token, hipStream_t stream, double flops, double bytes) {
  if (token < 0) return;
  std::lock_guard<std::mutex> lk(g_tmu);
  if (token >= (int)g_recs.size()) return;
  TimingRec &r = g_recs[token];
  hipEventRecord(r.b, stream);
  r.flops = flops;
  r.bytes = bytes;
  r.closed = true;
}
}  // namespace bsmpo

using namespace bsmpo;

namespace {

int run_group(const void *blob, int precision, void *d_work, size_t work_bytes, hipStream_t st,
              bool contract_only) {
  const GroupHeader *g = group_header(blob);
  if (!g || g->magic != kGroupMagic) return fail(-1, "not a bsmpo group");
  if (!d_work) return fail(-1, "workspace is NULL");
  if ((int64_t)work_bytes < g->work_bytes) return fail(-1, "workspace too small for this group");
  if (precision != BSMPO_C128 && precision != BSMPO_C64 && precision != BSMPO_BF16X3)
    return fail(-1, "unknown precision");
  BSMPO_HIP(hipMemcpyAsync(d_work, reinterpret_cast<const char *>(blob) + kGroupHostHead,
                           (size_t)g->staging_bytes, hipMemcpyHostToDevice, st));
  const char *w = reinterpret_cast<const char *>(d_work);
  int rc = launch_contract(*g, w, precision, st);
  if (rc || contract_only) return rc;
  if ((rc = launch_svd(*g, w, st))) return rc;
  return launch_truncate(*g, w, st);
}

// Single-update entry points run as a group of one.  The host image goes up
// with an asynchronous copy, so an image buffer is reused only after the copy
// that read it has completed: each thread keeps two pinned buffers, each with
// the event recorded after its last upload.
struct ImageSlot {
  char *buf = nullptr;
  int64_t cap = 0;
  hipEvent_t done = nullptr;
  bool pending = false;
};
struct ThreadImages {
  ImageSlot slot[2];
  int turn = 0;
  ~ThreadImages() {
    for (ImageSlot &s : slot) {
      if (s.pending) hipEventSynchronize(s.done);
      if (s.done) hipEventDestroy(s.done);
      if (s.buf) hipHostFree(s.buf);
    }
  }
};
thread_local ThreadImages t_images;

int image_for(int64_t bytes, ImageSlot **out) {
  ImageSlot &s = t_images.slot[t_images.turn];
  t_images.turn ^= 1;
  if (s.pending) {
    BSMPO_HIP(hipEventSynchronize(s.done));  // the last upload from this buffer has run
    s.pending = false;
  }
  if (!s.done) BSMPO_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
  if (s.cap < bytes) {
    if (s.buf) BSMPO_HIP(hipHostFree(s.buf));
    s.buf = nullptr;
    s.cap = 0;
    const int64_t cap = bytes + bytes / 4;
    BSMPO_HIP(hipHostMalloc((void **)&s.buf, (size_t)cap, hipHostMallocDefault));
    s.cap = cap;
  }
  *out = &s;
  return 0;
}

int single(const void *h_plan, const UpdateIO &io, void *d_work, size_t work_bytes, int precision,
           hipStream_t st, bool contract_only) {
  const void *plans[1] = {h_plan};
  int64_t hb = 0, wb = 0;
  int rc = group_query(1, plans, &hb, &wb);
  if (rc) return rc;
  if ((int64_t)work_bytes < wb) return fail(-1, "workspace too small for this plan");
  ImageSlot *img = nullptr;
  if ((rc = image_for(hb, &img))) return rc;
  if ((rc = group_build(1, plans, &io, reinterpret_cast<char *>(d_work), img->buf, hb,
                        contract_only)))
    return rc;
  rc = run_group(img->buf, precision, d_work, work_bytes, st, contract_only);
  // the upload was enqueued first in run_group: the event marks its completion
  if (hipEventRecord(img->done, st) == hipSuccess) img->pending = true;
  return rc;
}

// ------------------------------------------------ allocation-free variant --
// bsmpo_contract_noalloc keeps its staging (update descriptor, work lists and
// the plan's device part) in this file-scope array instead of a caller
// workspace.  Host code never passes `scratch` itself to a kernel or copy:
// each device's own copy is reached through the address hipGetSymbolAddress
// returns for it, looked up once per device and kept in g_scratch_ptr.
constexpr int64_t kNoAllocBytes = 8 << 20;  // plans up to chi = 2048 at d = 13 fit
constexpr int kMaxDevices = 64;
__device__ __align__(256) char scratch[kNoAllocBytes];
std::mutex g_noalloc_mu;                   // guards the two arrays below
void *g_scratch_ptr[kMaxDevices];          // device address of `scratch`, per device
alignas(256) char g_noalloc_image[kGroupHostHead + kNoAllocBytes];  // host image

int scratch_address(void **out) {
  int dev = 0;
  BSMPO_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(-1, "contract_noalloc: device index out of range");
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    BSMPO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

int bsmpo_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timing = on != 0;
  return 0;
}

int bsmpo_timing_read(double *ms, double *flops, double *bytes, int64_t *count, int reset) {
  std::vector<TimingRec> recs;
  {
    std::lock_guard<std::mutex> lk(g_tmu);
    recs.swap(g_recs);
  }
  for (TimingRec &r : recs) {
    if (r.closed) {
      BSMPO_HIP(hipEventSynchronize(r.b));
      float t = 0.f;
      BSMPO_HIP(hipEventElapsedTime(&t, r.a, r.b));
      g_ms[r.stage] += t;
      g_flops[r.stage] += r.flops;
      g_bytes[r.stage] += r.bytes;
      g_cnt[r.stage] += 1;
    }
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  // the kept-pair stages: flops counted on the device by the kept counts
  double kept[3] = {0.0, 0.0, 0.0};
  if (int rc = kept_flops_read(kept, true)) return rc;
  g_flops[ST_WMAT] += kept[0];
  g_flops[ST_VECTORS] += kept[1];
  g_flops[ST_BACKTRANS] += kept[2];
  for (int i = 0; i < ST_COUNT; ++i) {
    if (ms) ms[i] = g_ms[i];
    if (flops) flops[i] = g_flops[i];
    if (bytes) bytes[i] = g_bytes[i];
    if (count) count[i] = g_cnt[i];
    if (reset) g_ms[i] = g_flops[i] = g_bytes[i] = 0.0, g_cnt[i] = 0;
  }
  return 0;
}

int bsmpo_bs_table(int d, int nb, const double *d_B, double *d_U3, void *stream) {
  if (nb > 0 && (!d_B || !d_U3)) return fail(-1, "bs_table: NULL pointer");
  return launch_bs_table(d, nb, d_B, d_U3, (hipStream_t)stream);
}

int bsmpo_contract(int d, int chi, const void *h_plan, const double *d_Glc, const double *d_Gcr,
                   const double *d_LR, const double *d_U3, double *d_C, double *d_T,
                   const double *d_LL, const double *d_LC, int precision, void *d_work,
                   size_t work_bytes, void *stream) {
  int rc = validate_plan(h_plan, d, chi);
  if (rc) return rc;
  if (!d_Glc || !d_Gcr || !d_LR || !d_U3 || !d_C || !d_T) return fail(-1, "contract: NULL pointer");
  UpdateIO io{};
  io.Glc = const_cast<double *>(d_Glc);
  io.Gcr = const_cast<double *>(d_Gcr);
  io.LR = d_LR;
  io.U3 = d_U3;
  io.LL = d_LL;
  io.LC = d_LC;
  io.C_override = d_C;
  io.T_override = d_T;
  return single(h_plan, io, d_work, work_bytes, precision, (hipStream_t)stream, true);
}

int bsmpo_contract_noalloc(int d, int chi, const void *h_plan, const double *d_Glc,
                           const double *d_Gcr, const double *d_LR, const double *d_U3,
                           double *d_C, double *d_T, const double *d_LL, const double *d_LC,
                           int precision, void *stream) {
  int rc = validate_plan(h_plan, d, chi);
  if (rc) return rc;
  if (!d_Glc || !d_Gcr || !d_LR || !d_U3 || !d_C || !d_T)
    return fail(-1, "contract_noalloc: NULL pointer");
  if (precision != BSMPO_C128 && precision != BSMPO_C64 && precision != BSMPO_BF16X3)
    return fail(-1, "contract_noalloc: unknown precision");
  if (timing_on()) return fail(-1, "contract_noalloc: disable stage timing first (it creates events)");
  const void *plans[1] = {h_plan};
  int64_t hb = 0, wb = 0;
  if ((rc = group_query(1, plans, &hb, &wb))) return rc;
  if (hb - kGroupHostHead > kNoAllocBytes)
    return fail(-1, "contract_noalloc: this plan needs " + std::to_string(hb - kGroupHostHead) +
                        " bytes of staging, more than the library's 8 MiB scratch (use "
                        "bsmpo_contract with a workspace)");
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  void *dev_scratch = nullptr;
  if ((rc = scratch_address(&dev_scratch))) return rc;
  UpdateIO io{};
  io.Glc = const_cast<double *>(d_Glc);
  io.Gcr = const_cast<double *>(d_Gcr);
  io.LR = d_LR;
  io.U3 = d_U3;
  io.LL = d_LL;
  io.LC = d_LC;
  io.C_override = d_C;
  io.T_override = d_T;
  if ((rc = group_build(1, plans, &io, reinterpret_cast<char *>(dev_scratch), g_noalloc_image, hb,
                        true)))
    return rc;
  const GroupHeader *g = group_header(g_noalloc_image);
  BSMPO_HIP(hipMemcpyAsync(dev_scratch, g_noalloc_image + kGroupHostHead, (size_t)g->staging_bytes,
                           hipMemcpyHostToDevice, st));
  // the host image is shared by every call: it may be rewritten once the copy has read it
  BSMPO_HIP(hipStreamSynchronize(st));
  return launch_contract(*g, reinterpret_cast<const char *>(dev_scratch), precision, st);
}

int bsmpo_site_update(int d, int chi, const void *h_plan, double *d_Glc, double *d_Gcr,
                      const double *d_LR, const double *d_U3, int32_t *d_charge_c,
                      double *d_lambda_c, int32_t *d_num_lambda, int edge, int precision,
                      void *d_work, size_t work_bytes, void *stream) {
  int rc = validate_plan(h_plan, d, chi);
  if (rc) return rc;
  if (!d_Glc || !d_Gcr || !d_LR || !d_U3 || !d_charge_c || !d_lambda_c || !d_num_lambda)
    return fail(-1, "site_update: NULL pointer");
  UpdateIO io{};
  io.Glc = d_Glc;
  io.Gcr = d_Gcr;
  io.LR = d_LR;
  io.U3 = d_U3;
  io.charge_c = d_charge_c;
  io.lambda_c = d_lambda_c;
  io.num_lambda = d_num_lambda;
  io.edge = edge;
  return single(h_plan, io, d_work, work_bytes, precision, (hipStream_t)stream, false);
}

int bsmpo_group_query(int n_upd, const void *const *h_plans, size_t *host_bytes,
                      size_t *work_bytes) {
  if (!h_plans || !host_bytes || !work_bytes) return fail(-1, "group_query: NULL argument");
  int64_t hb = 0, wb = 0;
  int rc = group_query(n_upd, h_plans, &hb, &wb);
  if (rc) return rc;
  *host_bytes = (size_t)hb;
  *work_bytes = (size_t)wb;
  return 0;
}

int bsmpo_group_build(int n_upd, const void *const *h_plans, const bsmpo_update_io *io,
                      void *d_work, void *h_group, size_t host_bytes) {
  if (!h_plans || !io || !d_work || !h_group) return fail(-1, "group_build: NULL argument");
  std::vector<UpdateIO> v((size_t)std::max(n_upd, 0));
  for (int u = 0; u < n_upd; ++u) {
    if (!io[u].Glc || !io[u].Gcr || !io[u].LR || !io[u].U3 || !io[u].charge_c ||
        !io[u].lambda_c || !io[u].num_lambda)
      return fail(-1, "group_build: NULL pointer in update " + std::to_string(u));
    v[u] = UpdateIO{};
    v[u].Glc = io[u].Glc;
    v[u].Gcr = io[u].Gcr;
    v[u].LR = io[u].LR;
    v[u].U3 = io[u].U3;
    v[u].charge_c = io[u].charge_c;
    v[u].lambda_c = io[u].lambda_c;
    v[u].num_lambda = io[u].num_lambda;
    v[u].edge = io[u].edge;
  }
  return group_build(n_upd, h_plans, v.data(), reinterpret_cast<char *>(d_work),
                     reinterpret_cast<char *>(h_group), (int64_t)host_bytes);
}

int bsmpo_group_run(const void *h_group, int precision, void *d_work, size_t work_bytes,
                    void *stream) {
  return run_group(h_group, precision, d_work, work_bytes, (hipStream_t)stream, false);
}

int bsmpo_group_graph_create(const void *h_group, int precision, void *d_work, size_t work_bytes,
                             void *stream, void **graph_exec) {
  if (!graph_exec) return fail(-1, "group_graph_create: NULL graph_exec");
  *graph_exec = nullptr;
  hipStream_t st = (hipStream_t)stream;
  if (!st) return fail(-1, "group_graph_create: capture needs a non-default stream");
  if (timing_on()) return fail(-1, "group_graph_create: disable timing before capturing");
  BSMPO_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = run_group(h_group, precision, d_work, work_bytes, st, false);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &graph);
  if (rc) {
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail((int)e, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e2 != hipSuccess)
    return fail((int)e2, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
  *graph_exec = exec;
  return 0;
}

int bsmpo_graph_launch(void *graph_exec, void *stream) {
  if (!graph_exec) return fail(-1, "graph_launch: NULL graph");
  BSMPO_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
  return 0;
}

int bsmpo_graph_destroy(void *graph_exec) {
  if (graph_exec) BSMPO_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
  return 0;
}

int bsmpo_observables(int M, int d, int chi, const double *d_Gamma, const double *d_Lambda,
                      const int32_t *d_charge, double *d_R, double *d_L, double *d_marg,
                      double *d_trace, void *stream) {
  if (!d_Gamma || !d_Lambda || !d_charge || !d_R || !d_L || !d_marg || !d_trace)
    return fail(-1, "observables: NULL pointer");
  return launch_observables(M, d, chi, d_Gamma, d_Lambda, d_charge, d_R, d_L, d_marg, d_trace,
                            (hipStream_t)stream);
}

int bsmpo_entanglement_entropy(int M, int chi, const double *d_Lambda, double *d_S,
                               void *stream) {
  if (!d_Lambda || !d_S) return fail(-1, "entanglement_entropy: NULL pointer");
  return launch_entropy(M, chi, d_Lambda, d_S, (hipStream_t)stream);
}

}  // extern "C"
