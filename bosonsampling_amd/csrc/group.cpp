// Host construction of a group of planned updates: per-update descriptors,
// flattened (update, item) work lists and the workspace layout.  One launch
// per stage then covers every update of the group (all beam splitters of a
// brick-wall layer, across any number of independent interferometers).
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"

namespace bsmpo {
namespace {

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// The plan copies follow the work lists at o_plans, the per-update scratch
// areas follow the staging region at o_ws, each in update order and aligned
// like the other regions (no per-update offset arrays: building a group
// allocates nothing).
struct Layout {
  GroupHeader g;
};

inline int64_t region(int64_t bytes) { return align_up(std::max<int64_t>(bytes, 8), 256); }

// The device counts of a built group (its image's GroupCounts).  The split of
// the one-stage tridiagonal sectors between the two block shapes follows
// launch_tri_values: with few of them (a single update) all take the larger
// shape.
void write_counts(const GroupHeader &g, char *img) {
  GroupCounts &k = *reinterpret_cast<GroupCounts *>(img + g.o_counts);
  std::memset(&k, 0, sizeof(k));
  k.seg[SG_UPD] = make_int2(0, g.n_upd);
  for (int c = 0; c < kTileClasses; ++c)
    k.seg[SG_TILE + c] = make_int2(g.tile_off[c], g.tile_off[c + 1] - g.tile_off[c]);
  k.seg[SG_GRAM] = make_int2(0, g.n_gram);
  for (int c = 0; c < kTriClass; ++c)
    k.seg[SG_JAC + c] = make_int2(g.jac_off[c], g.jac_off[c + 1] - g.jac_off[c]);
  k.seg[SG_JACALL] = make_int2(0, g.jac_off[kTriClass]);
  k.seg[SG_TRI] = make_int2(0, g.n_tri);
  k.seg[SG_TS] = make_int2(0, g.n_ts);
  const int n1 = g.n_tri - g.n_ts;
  const int nbig = n1 < kTriFewSectors ? n1 : g.tri_nbig;
  k.seg[SG_TRIBIG] = make_int2(g.n_ts, nbig);
  k.seg[SG_TRISMALL] = make_int2(g.n_ts + nbig, n1 - nbig);
  k.seg[SG_SEC] = make_int2(0, g.n_sec);
  k.seg[SG_V] = make_int2(0, g.n_v);
  k.seg[SG_W] = make_int2(0, g.n_w);
  k.seg[SG_WYT] = make_int2(0, g.n_wyt);
  k.seg[SG_WYA] = make_int2(0, g.n_wya);
  k.seg[SG_BC] = make_int2(0, g.n_bc);
}

int layout(int n, const void *const *plans, Layout &L) {
  if (n <= 0) return fail(-1, "group: no updates");
  GroupHeader &g = L.g;
  std::memset(&g, 0, sizeof(g));
  g.magic = kGroupMagic;
  g.n_upd = n;
  int64_t n_tiles = 0, n_gram = 0, n_tri = 0, n_sec = 0, n_v = 0, n_w = 0, n_jac = 0;
  int64_t tile_cnt[kTileClasses] = {};
  for (int u = 0; u < n; ++u) {
    const PlanHeader *h = header(plans[u]);
    if (!h || h->magic != kPlanMagic) return fail(-1, "group: entry is not a bsmpo plan");
    n_tiles += h->n_titems;
    for (int c = 0; c < kTileClasses; ++c) tile_cnt[c] += h->tile_off[c + 1] - h->tile_off[c];
    g.flops_weight += h->flops_weight;
    g.mfma_k4 += h->mfma_k4;
    g.mfma_k32 += h->mfma_k32;
    g.d_max = std::max(g.d_max, h->d);
    n_gram += h->n_gram_items;
    n_v += h->n_v_items;
    n_w += h->n_w_items;
    n_sec += h->n_sectors;
    n_tri += h->jac_off[kTriClass + 1] - h->jac_off[kTriClass];
    n_jac += h->jac_off[kTriClass] - h->jac_off[0];
    for (int c = 0; c < kTriClass; ++c) g.jac_max_n[c] = std::max(g.jac_max_n[c], h->jac_max_n[c]);
    g.tri_max_n = std::max(g.tri_max_n, h->jac_max_n[kTriClass]);
    g.chi = std::max(g.chi, h->chi);
    g.max_sv = std::max<int64_t>(g.max_sv, h->sv_total);
    g.flops_alg += h->flops_alg;
    g.flops_gram += h->flops_gram;
    g.flops_jacobi += h->flops_jacobi;
    g.flops_wmat += h->flops_wmat;
    g.flops_vectors += h->flops_vectors;
    g.bytes_alg16 += h->bytes_alg16;
    g.c_elems += h->c_elems;
    g.g_elems += h->g_elems;
    g.v_elems += h->v_elems;
    g.w_elems += h->w_elems;
    g.sv_total += h->sv_total;
    g.flops_tridiag += h->flops_tridiag;
    g.bytes_tridiag += h->bytes_tridiag;
  }
  if (n_tiles > INT32_MAX || n_gram > INT32_MAX || n_w > INT32_MAX || n_v > INT32_MAX)
    return fail(-1, "group: too many work items");
  g.n_tiles = (int32_t)n_tiles;
  g.tile_off[0] = 0;
  for (int c = 0; c < kTileClasses; ++c) g.tile_off[c + 1] = g.tile_off[c] + (int32_t)tile_cnt[c];
  g.n_gram = (int32_t)n_gram;
  g.n_tri = (int32_t)n_tri;
  g.n_sec = (int32_t)n_sec;
  g.n_v = (int32_t)n_v;
  g.n_w = (int32_t)n_w;
  int64_t o = 0;
  auto place = [&](int64_t bytes) {
    const int64_t at = o;
    o = align_up(o + std::max<int64_t>(bytes, 8), 256);
    return at;
  };
  g.o_counts = place((int64_t)sizeof(GroupCounts));
  g.o_desc = place((int64_t)sizeof(UpdDesc) * n);
  g.o_tiles = place(8 * n_tiles);
  g.o_gram = place(8 * n_gram);
  g.o_jac = place(8 * n_jac);
  g.o_tri = place(8 * n_tri);
  g.o_sec = place(8 * n_sec);
  g.o_v = place(8 * n_v);
  g.o_w = place(8 * n_w);
  // back-transformation items of the tridiagonal-route sectors: one per WY
  // block of reflectors and one per panel of (at most) kWyBlock kept vectors
  int64_t n_wyt = 0, n_wya = 0, n_bc = 0;
  for (int u = 0; u < n; ++u) {
    n_wyt += header(plans[u])->n_wyt;
    n_wya += header(plans[u])->n_wya;
    n_bc += header(plans[u])->n_bc;
  }
  if (n_wyt > INT32_MAX || n_wya > INT32_MAX || n_bc > INT32_MAX)
    return fail(-1, "group: too many work items");
  g.n_wyt = (int32_t)n_wyt;
  g.n_wya = (int32_t)n_wya;
  g.n_bc = (int32_t)n_bc;
  g.o_wyt = place(8 * n_wyt);
  g.o_wya = place(8 * n_wya);
  g.o_bc = place(8 * n_bc);
  g.o_plans = o;
  for (int u = 0; u < n; ++u) o += region(header(plans[u])->dev_bytes);
  g.staging_bytes = o;
  g.o_ws = o;
  for (int u = 0; u < n; ++u) {
    const PlanHeader *h = header(plans[u]);
    o += region(h->w_total - h->w_C);
  }
  g.work_bytes = o;
  g.host_bytes = kGroupHostHead + g.staging_bytes;
  return 0;
}

}  // namespace

int group_query(int n, const void *const *plans, int64_t *host_bytes, int64_t *work_bytes) {
  Layout L;
  int rc = layout(n, plans, L);
  if (rc) return rc;
  *host_bytes = L.g.host_bytes;
  *work_bytes = L.g.work_bytes;
  return 0;
}

int group_build(int n, const void *const *plans, const UpdateIO *io, char *d_work,
                char *host_blob, int64_t host_bytes, bool contract_only) {
  Layout L;
  int rc = layout(n, plans, L);
  if (rc) return rc;
  GroupHeader &g = L.g;
  if (host_bytes < g.host_bytes) return fail(-1, "group: host buffer too small");
  char *img = host_blob + kGroupHostHead;  // image of d_work[0, staging_bytes)
  std::memset(img, 0, (size_t)g.staging_bytes);
  UpdDesc *desc = reinterpret_cast<UpdDesc *>(img + g.o_desc);
  int2 *tiles = reinterpret_cast<int2 *>(img + g.o_tiles);
  int32_t tile_fill[kTileClasses];
  for (int c = 0; c < kTileClasses; ++c) tile_fill[c] = g.tile_off[c];
  int2 *gram = reinterpret_cast<int2 *>(img + g.o_gram);
  int2 *jac = reinterpret_cast<int2 *>(img + g.o_jac);
  int2 *tri = reinterpret_cast<int2 *>(img + g.o_tri);
  int2 *sec = reinterpret_cast<int2 *>(img + g.o_sec);
  int2 *vl = reinterpret_cast<int2 *>(img + g.o_v);
  int2 *wl = reinterpret_cast<int2 *>(img + g.o_w);
  int64_t ng = 0, nt = 0, ns = 0, nv = 0, nw = 0;
  int64_t plan_at = g.o_plans, ws_at = g.o_ws;
  for (int u = 0; u < n; ++u) {
    const PlanHeader *h = header(plans[u]);
    std::memcpy(img + plan_at, reinterpret_cast<const char *>(plans[u]) + h->dev_begin,
                (size_t)h->dev_bytes);
    char *pb = d_work + plan_at;           // device address of the plan copy
    char *wb = d_work + ws_at - h->w_C;    // so that wb + h->w_X is the scratch X
    plan_at += region(h->dev_bytes);
    ws_at += region(h->w_total - h->w_C);
    UpdDesc &D = desc[u];
    D.sectors = reinterpret_cast<const Sector *>(pb + h->o_sectors);
    D.rows = reinterpret_cast<const int32_t *>(pb + h->o_rows);
    D.cols = reinterpret_cast<const int32_t *>(pb + h->o_cols);
    D.lblk = reinterpret_cast<const Block *>(pb + h->o_lblk);
    D.cblk = reinterpret_cast<const Block *>(pb + h->o_cblk);
    D.rblk = reinterpret_cast<const Block *>(pb + h->o_rblk);
    D.gram_items = reinterpret_cast<const Item *>(pb + h->o_gram_items);
    D.w_items = reinterpret_cast<const Item *>(pb + h->o_w_items);
    D.v_items = reinterpret_cast<const Item *>(pb + h->o_v_items);
    D.charges = reinterpret_cast<const int32_t *>(pb + h->o_charges);
    D.sv_sector = reinterpret_cast<const int32_t *>(pb + h->o_sv_sector);
    D.C = io[u].C_override ? io[u].C_override : reinterpret_cast<double *>(wb + h->w_C);
    D.T = io[u].T_override ? io[u].T_override : reinterpret_cast<double *>(wb + h->w_T);
    D.G = reinterpret_cast<double *>(wb + h->w_G);
    D.V = reinterpret_cast<double *>(wb + h->w_V);
    D.W = reinterpret_cast<double *>(wb + h->w_W);
    D.S = reinterpret_cast<double *>(wb + h->w_S);
    D.lam = reinterpret_cast<double *>(wb + h->w_lam);
    D.td = reinterpret_cast<double *>(wb + h->w_td);
    D.te = reinterpret_cast<double *>(wb + h->w_te);
    D.tau = reinterpret_cast<double *>(wb + h->w_tau);
    D.triwork = reinterpret_cast<double *>(wb + h->w_triwork);
    D.flag = reinterpret_cast<int32_t *>(wb + h->w_flag);
    D.slot = reinterpret_cast<int32_t *>(wb + h->w_slot);
    D.kcount = reinterpret_cast<int32_t *>(wb + h->w_kcount);
    D.hist = reinterpret_cast<int32_t *>(wb + h->w_hist);
    D.Glc = io[u].Glc;
    D.Gcr = io[u].Gcr;
    D.LR = io[u].LR;
    D.U3 = io[u].U3;
    D.LL = io[u].LL;
    D.LC = io[u].LC;
    if (io[u].LC) g.has_lc = 1;
    D.charge_c = io[u].charge_c;
    D.lambda_c = io[u].lambda_c;
    D.num_lambda = io[u].num_lambda;
    D.d = h->d;
    D.chi = h->chi;
    D.edge = io[u].edge;
    D.n_sectors = h->n_sectors;
    D.n_cblk = h->n_cblk;
    D.eig_thr = h->eig_threshold;
    D.sv_total = h->sv_total;
    D.tie_rel = h->tie_rel;
    D.titems = reinterpret_cast<const TileItem *>(pb + h->o_titems);
    D.tau_sector = reinterpret_cast<const int32_t *>(pb + h->o_tau_sector);
    D.cblk_of = reinterpret_cast<const int32_t *>(pb + h->o_cblk_of);
    D.rowoff = reinterpret_cast<const int32_t *>(pb + h->o_rowoff);
    D.coloff = reinterpret_cast<const int32_t *>(pb + h->o_coloff);
    D.row_info = reinterpret_cast<const int2 *>(pb + h->o_row_info);
    D.col_info = reinterpret_cast<const int2 *>(pb + h->o_col_info);
    D.n_lblk = h->n_lblk;
    D.n_rblk = h->n_rblk;
    D.ts_min = h->ts_min;
    D.drop_rel2 = h->drop_rel * h->drop_rel;
    for (int c = 0; c < kTileClasses; ++c)
      for (int i = h->tile_off[c]; i < h->tile_off[c + 1]; ++i) tiles[tile_fill[c]++] = make_int2(u, i);
    if (contract_only) continue;  // the SVD-stage lists stay empty (zero)
    for (int i = 0; i < h->n_gram_items; ++i) gram[ng++] = make_int2(u, i);
    for (int i = 0; i < h->n_v_items; ++i) vl[nv++] = make_int2(u, i);
    for (int i = 0; i < h->n_w_items; ++i) wl[nw++] = make_int2(u, i);
    for (int i = 0; i < h->n_sectors; ++i) sec[ns++] = make_int2(u, i);
    const int32_t *jo = host_array<int32_t>(plans[u], h->o_jac_order);
    for (int i = h->jac_off[kTriClass]; i < h->jac_off[kTriClass + 1]; ++i) tri[nt++] = make_int2(u, jo[i]);
  }
  if (contract_only) {  // a group that only runs the contraction stage
    g.n_gram = g.n_tri = g.n_sec = g.n_v = g.n_w = g.n_wyt = g.n_wya = g.n_bc = 0;
    for (int c = 0; c <= kTriClass; ++c) g.jac_off[c] = 0;
    write_counts(g, img);
    std::memcpy(host_blob, &g, sizeof(GroupHeader));
    return 0;
  }
  // Tridiagonal-route sectors: the two-stage ones first, then the one-stage
  // ones, each part largest first across the group: blocks are dispatched in
  // list order, so the longest chains start at once and the short ones fill
  // the slots they leave (longest-processing-time order).
  {
    auto size_of = [&](int2 r) {
      return host_array<Sector>(plans[r.x], header(plans[r.x])->o_sectors)[r.y].n;
    };
    auto ts_of = [&](int2 r) { return two_stage(size_of(r), header(plans[r.x])->ts_min); };
    std::stable_sort(tri, tri + nt, [&](int2 x, int2 y) {
      const bool tx = ts_of(x), ty = ts_of(y);
      return tx != ty ? tx : size_of(x) > size_of(y);
    });
    int n_ts = 0;
    while (n_ts < nt && ts_of(tri[n_ts])) ++n_ts;
    g.n_ts = n_ts;
    g.ts_max_n = n_ts > 0 ? size_of(tri[0]) : 0;
    g.ts_nbig = 0;
    while (g.ts_nbig < n_ts && size_of(tri[g.ts_nbig]) - kBand > kBandThreads) ++g.ts_nbig;
    g.tri_split = tri_split();
    g.tri_nbig = 0;  // one-stage sectors above the block-shape split
    while (n_ts + g.tri_nbig < nt && size_of(tri[n_ts + g.tri_nbig]) > g.tri_split) ++g.tri_nbig;
    // back-transformation items in the same order (largest sectors first)
    int2 *wyt = reinterpret_cast<int2 *>(img + g.o_wyt);
    int2 *wya = reinterpret_cast<int2 *>(img + g.o_wya);
    int2 *bc = reinterpret_cast<int2 *>(img + g.o_bc);
    int64_t a = 0, b = 0, cc = 0;
    for (int64_t t = 0; t < nt; ++t) {
      const int nn = size_of(tri[t]);
      const bool ts = t < n_ts;
      const int off = ts ? kBand : 1;  // row offset of reflector i's unit entry
      for (int m = 0; m < (nn - off + kWyBlock - 1) / kWyBlock; ++m)
        wyt[a++] = make_int2(tri[t].x, tri[t].y | m << 20);
      for (int p = 0; p < (nn + kWyBlock - 1) / kWyBlock; ++p)
        wya[b++] = make_int2(tri[t].x, tri[t].y | p << 20);
      if (ts)
        for (int q = 0; q < (nn + bc_vectors(nn) - 1) / bc_vectors(nn); ++q)
          bc[cc++] = make_int2(tri[t].x, tri[t].y | q << 20);
    }
    g.n_wyt = (int32_t)a;
    g.n_bc = (int32_t)cc;
  }
  // LDS Jacobi list grouped by size class across updates
  int64_t nj = 0;
  for (int c = 0; c < kTriClass; ++c) {
    g.jac_off[c] = (int32_t)nj;
    for (int u = 0; u < n; ++u) {
      const PlanHeader *h = header(plans[u]);
      const int32_t *jo = host_array<int32_t>(plans[u], h->o_jac_order);
      for (int i = h->jac_off[c]; i < h->jac_off[c + 1]; ++i) jac[nj++] = make_int2(u, jo[i]);
    }
  }
  g.jac_off[kTriClass] = (int32_t)nj;
  write_counts(g, img);
  std::memcpy(host_blob, &g, sizeof(GroupHeader));
  return 0;
}

}  // namespace bsmpo
