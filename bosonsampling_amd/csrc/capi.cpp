// extern "C" entry points of libbsmpo (declared in include/bsmpo.h).
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "internal.h"

namespace bsmpo {
int launch_observables(int M, int d, int chi, const double *Gamma, const double *Lambda,
                       const int32_t *charge, double *R, double *L, double *marg, double *trace,
                       hipStream_t stream);
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
bool g_timing = false;
std::vector<TimingRec> g_recs;
double g_ms[ST_COUNT], g_flops[ST_COUNT], g_bytes[ST_COUNT];
int64_t g_cnt[ST_COUNT];
}  // namespace

bool timing_on() { return g_timing; }

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

int upload_plan(const void *h_plan, void *d_work, size_t work_bytes, hipStream_t stream) {
  const PlanHeader *h = header(h_plan);
  if (!d_work) return fail(-1, "workspace is NULL");
  if ((int64_t)work_bytes < h->w_total) return fail(-1, "workspace too small for this plan");
  if (h->dev_bytes > 0)
    BSMPO_HIP(hipMemcpyAsync(d_work, reinterpret_cast<const char *>(h_plan) + h->dev_begin,
                             (size_t)h->dev_bytes, hipMemcpyHostToDevice, stream));
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
  hipStream_t st = (hipStream_t)stream;
  if ((rc = upload_plan(h_plan, d_work, work_bytes, st))) return rc;
  return launch_contract(*header(h_plan), d_work, d_Glc, d_Gcr, d_LR, d_U3, d_C, d_T, d_LL, d_LC,
                         precision, st);
}

int bsmpo_site_update(int d, int chi, const void *h_plan, double *d_Glc, double *d_Gcr,
                      const double *d_LR, const double *d_U3, int32_t *d_charge_c,
                      double *d_lambda_c, int32_t *d_num_lambda, int edge, int precision,
                      void *d_work, size_t work_bytes, void *stream) {
  int rc = validate_plan(h_plan, d, chi);
  if (rc) return rc;
  if (!d_Glc || !d_Gcr || !d_LR || !d_U3 || !d_charge_c || !d_lambda_c || !d_num_lambda)
    return fail(-1, "site_update: NULL pointer");
  const PlanHeader &h = *header(h_plan);
  hipStream_t st = (hipStream_t)stream;
  if ((rc = upload_plan(h_plan, d_work, work_bytes, st))) return rc;
  char *w = reinterpret_cast<char *>(d_work);
  if ((rc = launch_contract(h, d_work, d_Glc, d_Gcr, d_LR, d_U3,
                            reinterpret_cast<double *>(w + h.w_C),
                            reinterpret_cast<double *>(w + h.w_T), nullptr, nullptr, precision,
                            st)))
    return rc;
  if ((rc = launch_svd(h, d_work, st))) return rc;
  return launch_truncate(h, d_work, d_Glc, d_Gcr, d_charge_c, d_lambda_c, d_num_lambda, edge, st);
}

int bsmpo_observables(int M, int d, int chi, const double *d_Gamma, const double *d_Lambda,
                      const int32_t *d_charge, double *d_R, double *d_L, double *d_marg,
                      double *d_trace, void *stream) {
  if (!d_Gamma || !d_Lambda || !d_charge || !d_R || !d_L || !d_marg || !d_trace)
    return fail(-1, "observables: NULL pointer");
  return launch_observables(M, d, chi, d_Gamma, d_Lambda, d_charge, d_R, d_L, d_marg, d_trace,
                            (hipStream_t)stream);
}

}  // extern "C"
