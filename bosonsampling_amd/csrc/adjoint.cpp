// Host side of the reverse-mode path: the backward workspace layout, the
// per-sector tables, and the extern "C" entry points (include/bsmpo.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "adjoint.h"

namespace bsmpo {
namespace {

int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

struct VjpLayout {
  int64_t contract_work = 0;  // bytes at the front: the forward contraction's workspace
  int64_t o_C = 0, o_T = 0, o_gC = 0, o_tab = 0, o_rowpos = 0, o_colpos = 0, o_scr = 0,
          o_dscr = 0, total = 0;
  int64_t t_sec = 0, t_pairs = 0, t_rows = 0, t_cols = 0, t_ch = 0, t_lblk = 0, t_cblk = 0,
          t_rblk = 0, tab_bytes = 0;
  int n_sectors = 0, n_pairs = 0, n_cblk = 0;
  std::vector<char> tab;  // host image of the tables
};

thread_local VjpLayout t_lay;

// Kept slots of every sector from the new centre charges: the forward keeps
// a prefix (descending singular values) of each sector and writes its slots
// contiguously in charge order, so sector tau owns the run of slots whose
// new charge equals tau.
int vjp_layout(int d, int chi, const void *h_plan, const int32_t *h_charge_new, VjpLayout &Lo) {
  int rc = validate_plan(h_plan, d, chi);
  if (rc) return rc;
  if (!h_charge_new) return fail(-1, "site_update_vjp: NULL new charges");
  const PlanHeader *h = header(h_plan);
  const Sector *sec = host_array<Sector>(h_plan, h->o_sectors);
  const Block *lblk = host_array<Block>(h_plan, h->o_lblk);
  const Block *cblk = host_array<Block>(h_plan, h->o_cblk);
  const Block *rblk = host_array<Block>(h_plan, h->o_rblk);
  const int ns = h->n_sectors;
  const void *plans[1] = {h_plan};
  int64_t hb = 0, wb = 0;
  if ((rc = group_query(1, plans, &hb, &wb))) return rc;

  std::vector<VjpSector> vs((size_t)ns);
  int64_t scr = 0, dscr = 0;
  int j = 0;
  for (int s = 0; s < ns; ++s) {
    const Sector &x = sec[s];
    VjpSector v{};
    v.t0 = x.t0;
    v.t1 = x.t1;
    v.L = x.L;
    v.R = x.R;
    v.n = x.n;
    v.left_small = x.left_small;
    v.row_off = x.row_off;
    v.col_off = x.col_off;
    v.c_off = x.c_off;
    // skip slots of charges that sort before this sector (they cannot occur)
    if (j < chi && h_charge_new[2 * j] < d &&
           (h_charge_new[2 * j] < x.t0 ||
            (h_charge_new[2 * j] == x.t0 && h_charge_new[2 * j + 1] < x.t1)))
      return fail(-1, "site_update_vjp: new charge (" + std::to_string(h_charge_new[2 * j]) +
                          "," + std::to_string(h_charge_new[2 * j + 1]) +
                          ") has no sector in this plan");
    v.slot0 = j;
    while (j < chi && h_charge_new[2 * j] == x.t0 && h_charge_new[2 * j + 1] == x.t1) ++j;
    v.k = j - v.slot0;
    if (v.k > v.n) return fail(-1, "site_update_vjp: more kept slots than the sector's rank");
    const int64_t L = x.L, R = x.R, n = x.n, k = v.k;
    const int64_t pmax = std::max(n * R, L * n), xmax = std::max(k * R, L * k);
    v.scr = scr;
    v.dscr = dscr;
    if (k > 0) {
      scr += 3 * n * n + 2 * L * k + 3 * n * k + 2 * pmax + xmax + L * R + k * R;
      dscr += 3 * n + 2 * k;
    }
    vs[s] = v;
  }
  for (; j < chi; ++j)
    if (h_charge_new[2 * j] < d || h_charge_new[2 * j + 1] < d)
      return fail(-1, "site_update_vjp: new charges are not sorted or have no sector");

  std::vector<VjpPair> pairs;
  for (int s = 0; s < ns; ++s) {
    const Sector &x = sec[s];
    int ro = 0;
    for (int a = 0; a < h->n_lblk; ++a) {
      const Block &la = lblk[a];
      if (!(la.c0 >= x.t0 && la.c1 >= x.t1)) continue;
      int co = 0;
      for (int b = 0; b < h->n_rblk; ++b) {
        const Block &rb = rblk[b];
        if (!(rb.c0 <= x.t0 && rb.c1 <= x.t1)) continue;
        if (la.c0 >= rb.c0 && la.c1 >= rb.c1) pairs.push_back(VjpPair{s, a, b, ro, co, 0});
        co += rb.end - rb.start;
      }
      ro += la.end - la.start;
    }
  }

  // host image of the tables
  Lo.n_sectors = ns;
  Lo.n_pairs = (int)pairs.size();
  Lo.n_cblk = h->n_cblk;
  int64_t o = 0;
  auto put = [&](int64_t bytes) {
    const int64_t at = o;
    o = align256(o + bytes);
    return at;
  };
  Lo.t_sec = put((int64_t)sizeof(VjpSector) * ns);
  Lo.t_pairs = put((int64_t)sizeof(VjpPair) * pairs.size());
  Lo.t_rows = put(4 * h->rows_total);
  Lo.t_cols = put(4 * h->cols_total);
  Lo.t_ch = put(4 * 6 * (int64_t)chi);
  Lo.t_lblk = put((int64_t)sizeof(Block) * h->n_lblk);
  Lo.t_cblk = put((int64_t)sizeof(Block) * h->n_cblk);
  Lo.t_rblk = put((int64_t)sizeof(Block) * h->n_rblk);
  Lo.tab_bytes = o;
  Lo.tab.assign((size_t)o, 0);
  char *t = Lo.tab.data();
  if (ns) std::memcpy(t + Lo.t_sec, vs.data(), sizeof(VjpSector) * ns);
  if (!pairs.empty()) std::memcpy(t + Lo.t_pairs, pairs.data(), sizeof(VjpPair) * pairs.size());
  std::memcpy(t + Lo.t_rows, host_array<int32_t>(h_plan, h->o_rows), 4 * h->rows_total);
  std::memcpy(t + Lo.t_cols, host_array<int32_t>(h_plan, h->o_cols), 4 * h->cols_total);
  std::memcpy(t + Lo.t_ch, host_array<int32_t>(h_plan, h->o_charges), 4 * 6 * (int64_t)chi);
  std::memcpy(t + Lo.t_lblk, lblk, sizeof(Block) * h->n_lblk);
  std::memcpy(t + Lo.t_cblk, cblk, sizeof(Block) * h->n_cblk);
  std::memcpy(t + Lo.t_rblk, rblk, sizeof(Block) * h->n_rblk);

  // device workspace
  o = 0;
  Lo.contract_work = wb;
  o = align256(wb);
  Lo.o_C = o;
  o = align256(o + 16 * std::max<int64_t>(h->c_elems, 1));
  Lo.o_T = o;
  o = align256(o + 16 * std::max<int64_t>(h->c_elems, 1));
  Lo.o_gC = o;
  o = align256(o + 16 * std::max<int64_t>(h->c_elems, 1));
  Lo.o_tab = o;
  o = align256(o + Lo.tab_bytes);
  Lo.o_rowpos = o;
  o = align256(o + 4 * (int64_t)ns * chi);
  Lo.o_colpos = o;
  o = align256(o + 4 * (int64_t)ns * chi);
  Lo.o_scr = o;
  o = align256(o + 16 * std::max<int64_t>(scr, 1));
  Lo.o_dscr = o;
  o = align256(o + 8 * std::max<int64_t>(dscr, 1));
  Lo.total = o;
  return 0;
}

}  // namespace
}  // namespace bsmpo

using namespace bsmpo;

extern "C" {

int bsmpo_site_update_vjp_query(int d, int chi, const void *h_plan, const int32_t *h_charge_new,
                                size_t *work_bytes) {
  if (!work_bytes) return fail(-1, "site_update_vjp_query: NULL work_bytes");
  int rc = vjp_layout(d, chi, h_plan, h_charge_new, t_lay);
  if (rc) return rc;
  *work_bytes = (size_t)t_lay.total;
  return 0;
}

int bsmpo_site_update_vjp(int d, int chi, const void *h_plan, const int32_t *h_charge_new,
                          const double *d_Glc, const double *d_Gcr, const double *d_LR,
                          const double *d_U3, const double *d_G0new, const double *d_G1new,
                          const double *d_lambda_new, const double *d_gG0, const double *d_gG1,
                          const double *d_gLam, double *d_gGlc, double *d_gGcr, double *d_gLR,
                          double *d_gU3, int precision, void *d_work, size_t work_bytes,
                          void *stream) {
  if (!d_Glc || !d_Gcr || !d_LR || !d_U3 || !d_G0new || !d_G1new || !d_lambda_new || !d_gG0 ||
      !d_gG1 || !d_gLam || !d_gGlc || !d_gGcr || !d_gLR || !d_gU3 || !d_work)
    return fail(-1, "site_update_vjp: NULL pointer");
  if (d_G0new == d_Glc || d_G1new == d_Gcr)
    return fail(-1, "site_update_vjp: the update's inputs and outputs must be separate buffers");
  VjpLayout &Lo = t_lay;
  int rc = vjp_layout(d, chi, h_plan, h_charge_new, Lo);
  if (rc) return rc;
  if ((int64_t)work_bytes < Lo.total) return fail(-1, "site_update_vjp: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  char *w = reinterpret_cast<char *>(d_work);
  double *C = reinterpret_cast<double *>(w + Lo.o_C);
  double *T = reinterpret_cast<double *>(w + Lo.o_T);
  // forward contraction again (same precision as the forward), C and T packed
  if ((rc = bsmpo_contract(d, chi, h_plan, d_Glc, d_Gcr, d_LR, d_U3, C, T, nullptr, nullptr,
                           precision, d_work, (size_t)Lo.contract_work, stream)))
    return rc;
  char *tab = w + Lo.o_tab;
  BSMPO_HIP(hipMemcpyAsync(tab, Lo.tab.data(), (size_t)Lo.tab_bytes, hipMemcpyHostToDevice, st));
  BSMPO_HIP(hipStreamSynchronize(st));  // the host image is reused by the next call
  VjpArgs a{};
  a.d = d;
  a.chi = chi;
  a.n_sectors = Lo.n_sectors;
  a.n_cblk = Lo.n_cblk;
  a.sec = reinterpret_cast<const VjpSector *>(tab + Lo.t_sec);
  a.pairs = reinterpret_cast<const VjpPair *>(tab + Lo.t_pairs);
  a.rows = reinterpret_cast<const int32_t *>(tab + Lo.t_rows);
  a.cols = reinterpret_cast<const int32_t *>(tab + Lo.t_cols);
  const int32_t *ch = reinterpret_cast<const int32_t *>(tab + Lo.t_ch);
  a.CL = ch;
  a.CC = ch + 2 * chi;
  a.CR = ch + 4 * chi;
  a.lblk = reinterpret_cast<const Block *>(tab + Lo.t_lblk);
  a.cblk = reinterpret_cast<const Block *>(tab + Lo.t_cblk);
  a.rblk = reinterpret_cast<const Block *>(tab + Lo.t_rblk);
  a.rowpos = reinterpret_cast<int32_t *>(w + Lo.o_rowpos);
  a.colpos = reinterpret_cast<int32_t *>(w + Lo.o_colpos);
  a.C = C;
  a.T = T;
  a.gC = reinterpret_cast<double *>(w + Lo.o_gC);
  a.Glc = d_Glc;
  a.Gcr = d_Gcr;
  a.LR = d_LR;
  a.U3 = d_U3;
  a.G0 = d_G0new;
  a.G1 = d_G1new;
  a.lam_new = d_lambda_new;
  a.gG0 = d_gG0;
  a.gG1 = d_gG1;
  a.gLam = d_gLam;
  a.gGlc = d_gGlc;
  a.gGcr = d_gGcr;
  a.gLR = d_gLR;
  a.gU3 = d_gU3;
  a.scr = reinterpret_cast<double *>(w + Lo.o_scr);
  a.dscr = reinterpret_cast<double *>(w + Lo.o_dscr);
  a.drop_rel2 = drop_rel() * drop_rel();
  return launch_site_vjp(a, Lo.n_pairs, st);
}

int bsmpo_bs_table_vjp(int d, int nb, const double *d_B, const double *d_gU3, double *d_gB,
                       void *stream) {
  if (nb > 0 && (!d_B || !d_gU3 || !d_gB)) return fail(-1, "bs_table_vjp: NULL pointer");
  return launch_bs_table_vjp(d, nb, d_B, d_gU3, d_gB, (hipStream_t)stream);
}

int bsmpo_observable_vjp(int M, int d, int chi, const double *d_Gamma, const double *d_Lambda,
                         const int32_t *d_charge, const int32_t *d_pattern,
                         const double *d_weights, double *d_env, double *d_gGamma, double *d_out,
                         void *stream) {
  if (M < 2 || chi < 1 || d < 1) return fail(-1, "observable_vjp: bad sizes");
  if (!d_Gamma || !d_Lambda || !d_charge || !d_pattern || !d_weights || !d_env || !d_gGamma ||
      !d_out)
    return fail(-1, "observable_vjp: NULL pointer");
  return launch_observable_vjp(M, d, chi, d_Gamma, d_Lambda, d_charge, d_pattern, d_weights,
                               d_env, d_gGamma, d_out, (hipStream_t)stream);
}

}  // extern "C"
