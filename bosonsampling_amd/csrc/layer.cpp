// Layer programs (bsmpo_layer_*): every site update of a brick-wall layer,
// planned on the device (devplan.hip) and run by the grouped stages, captured
// once as one HIP graph and replayed for every layer with the same update set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.h"

namespace bsmpo {
int launch_contract(const GroupHeader &g, const char *d_work, int precision, hipStream_t stream);
int launch_svd(const GroupHeader &g, const char *d_work, hipStream_t stream);
int launch_truncate(const GroupHeader &g, const char *d_work, hipStream_t stream);

namespace {

int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

// BSMPO_TRACE=1 in the environment: program construction steps on stderr
void trace(const char *what) {
  static const bool on = [] {
    const char *e = std::getenv("BSMPO_TRACE");
    return e && e[0] == '1';
  }();
  if (on) {
    std::fprintf(stderr, "[bsmpo layer] %s\n", what);
    std::fflush(stderr);
  }
}
int64_t region(int64_t bytes) { return align256(std::max<int64_t>(bytes, 8)); }

constexpr int kMaxChains = 4;  // scratch arenas, each a chain of groups

struct LayerProgram {
  int device = 0, precision = 0, n_streams = 1;  // n_streams: scratch arenas (concurrent chains)
  DpParams params{};
  DpUpdate *d_upd = nullptr;
  DpGroup *d_groups = nullptr;
  DpSizes *d_sizes = nullptr;
  std::vector<GroupHeader> cap;  // per group: capacity header (grids, LDS, list offsets)
  std::vector<char *> staging;   // per group: device staging region
  // the layer's captured graphs: the planning, and per group its parts
  // (head, Householder 512-thread shape, 256-thread shape, tail); a part with
  // no launch has no graph
  hipGraphExec_t plan = nullptr;
  std::vector<std::array<hipGraphExec_t, 4>> parts;
  // streams of a replay besides the caller's: the second arena's chain, and
  // per chain the stream of its 256-thread Householder launches; events that
  // order them
  hipStream_t chain[kMaxChains] = {}, side[kMaxChains] = {};  // chain[0]: the caller's
  hipEvent_t ev_plan = nullptr, ev_done[kMaxChains] = {};
  std::vector<hipEvent_t> ev_head, ev_small;
  ~LayerProgram() {
    if (plan) (void)hipGraphExecDestroy(plan);
    for (auto &q : parts)
      for (hipGraphExec_t e : q)
        if (e) (void)hipGraphExecDestroy(e);
    for (hipEvent_t e : ev_head) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_small) (void)hipEventDestroy(e);
    if (ev_plan) (void)hipEventDestroy(ev_plan);
    for (hipEvent_t e : ev_done)
      if (e) (void)hipEventDestroy(e);
    for (int c = 1; c < kMaxChains; ++c)
      if (chain[c]) (void)hipStreamDestroy(chain[c]);
    for (hipStream_t x : side)
      if (x) (void)hipStreamDestroy(x);
  }
};

// Byte layout of a program's workspace and of each group's staging region,
// from the spec alone (query and create agree on it).
struct Layout {
  int64_t o_state, o_upd, o_groups, o_sizes, o_u3, total;
  int64_t staging_bytes, fixed_bytes;
  std::vector<int64_t> o_staging, o_fixed, o_var, o_tri_tmp, o_work;  // per group / per stream
  GroupHeader cap;  // staging offsets (shared by every group), capacity counts
};

int check_spec(const bsmpo_layer_spec *s) {
  if (!s) return fail(-1, "layer: NULL spec");
  if (s->d < 1 || s->d > kMaxD) return fail(-1, "layer: d must be in [1, 16]");
  if (s->chi < 1 || s->chi > kDpMaxChi) return fail(-1, "layer: chi must be in [1, 2048]");
  if (s->precision != BSMPO_C128 && s->precision != BSMPO_C64 && s->precision != BSMPO_BF16X3)
    return fail(-1, "layer: unknown precision");
  if (s->n_sets < 1 || s->n_sets > 2) return fail(-1, "layer: n_sets must be 1 or 2");
  for (int k = 0; k < s->n_sets; ++k) {
    if (s->n_upd[k] < 1 || !s->upd[k]) return fail(-1, "layer: an update set is empty");
    if (s->n_groups < 1 || s->n_groups > s->n_upd[k])
      return fail(-1, "layer: n_groups must be in [1, updates of every set]");
    for (int u = 0; u < s->n_upd[k]; ++u) {
      const bsmpo_layer_update &x = s->upd[k][u];
      if (!x.CL || !x.CC || !x.CR || !x.Glc || !x.Gcr || !x.LR || !x.charge_c || !x.lambda_c ||
          !x.num_lambda)
        return fail(-1, "layer: NULL pointer in update " + std::to_string(u) + " of set " +
                            std::to_string(k));
    }
  }
  if (s->n_streams < 1 || s->n_streams > kMaxChains)
    return fail(-1, "layer: n_streams must be in [1, 4]");
  if (s->n_layers < 1 || !s->d_B) return fail(-1, "layer: no BS matrices");
  const bsmpo_layer_caps &c = s->caps;
  if (c.var_bytes < 1 || c.work_bytes < 1 || c.tiles < 1 || c.gram < 1 || c.jac < 0 || c.tri < 0 ||
      c.sec < 1 || c.v < 0 || c.w < 1 || c.wyt < 0 || c.wya < 0)
    return fail(-1, "layer: capacities must be positive");
  if (two_stage_min() != 0)
    return fail(-1, "layer: the two-stage reduction is not planned on the device (bsmpo_set_two_stage(0))");
  return 0;
}

// updates of group g of a set (contiguous, the first n % n_groups groups one more)
int group_size(const bsmpo_layer_spec *s, int set, int g) {
  const int base = s->n_upd[set] / s->n_groups, extra = s->n_upd[set] % s->n_groups;
  return base + (g < extra ? 1 : 0);
}
int group_first(const bsmpo_layer_spec *s, int set, int g) {
  int f = 0;
  for (int k = 0; k < g; ++k) f += group_size(s, set, k);
  return f;
}
int max_group(const bsmpo_layer_spec *s) {
  int m = 0;
  for (int k = 0; k < s->n_sets; ++k) m = std::max(m, group_size(s, k, 0));
  return m;
}
int max_upd(const bsmpo_layer_spec *s) {
  return s->n_sets == 2 ? std::max(s->n_upd[0], s->n_upd[1]) : s->n_upd[0];
}

void layout(const bsmpo_layer_spec *s, Layout &L) {
  const bsmpo_layer_caps &c = s->caps;
  int64_t o = 0;
  auto place = [&](int64_t bytes) {
    const int64_t at = o;
    o = region(o + std::max<int64_t>(bytes, 8));
    return at;
  };
  // staging region of one group: GroupCounts, descriptors, the work lists
  const int mg = max_group(s), mu = max_upd(s);
  const int n_all = s->n_upd[0] + (s->n_sets == 2 ? s->n_upd[1] : 0);
  GroupHeader &h = L.cap;
  std::memset(&h, 0, sizeof(h));
  h.o_counts = place((int64_t)sizeof(GroupCounts));
  h.o_desc = place((int64_t)sizeof(UpdDesc) * mg);
  h.o_tiles = place(8 * (int64_t)c.tiles);
  h.o_gram = place(8 * (int64_t)c.gram);
  h.o_jac = place(8 * (int64_t)c.jac);
  h.o_tri = place(8 * (int64_t)c.tri);
  h.o_sec = place(8 * (int64_t)c.sec);
  h.o_v = place(8 * (int64_t)c.v);
  h.o_w = place(8 * (int64_t)c.w);
  h.o_wyt = place(8 * (int64_t)c.wyt);
  h.o_wya = place(8 * (int64_t)c.wya);
  h.o_bc = place(8);
  L.staging_bytes = o;
  L.fixed_bytes = fixed_layout(s->d, s->chi).bytes;
  // program workspace
  o = 0;
  L.o_state = place(64);
  L.o_upd = place((int64_t)sizeof(DpUpdate) * n_all);
  L.o_groups = place((int64_t)sizeof(DpGroup) * s->n_groups);
  L.o_sizes = place((int64_t)sizeof(DpSizes) * mu);
  L.o_u3 = place(16 * (int64_t)s->d * s->d * s->d * mu);
  L.o_staging.resize(s->n_groups);
  L.o_fixed.resize(s->n_groups);
  L.o_var.resize(s->n_groups);
  L.o_tri_tmp.resize(s->n_groups);
  for (int g = 0; g < s->n_groups; ++g) {
    L.o_staging[g] = place(L.staging_bytes);
    L.o_fixed[g] = place(L.fixed_bytes * mg);
    L.o_var[g] = place(c.var_bytes);
    L.o_tri_tmp[g] = place(8 * (int64_t)c.tri);
  }
  L.o_work.resize(s->n_streams);
  for (int k = 0; k < s->n_streams; ++k) L.o_work[k] = place(c.work_bytes);
  L.total = o;
}

// The capacity header of a group: counts are capacities (they size grids and
// LDS); the kernels read the planned counts from the group's GroupCounts.
GroupHeader capacity_header(const bsmpo_layer_spec *s, const Layout &L, int n_upd, int thr) {
  const bsmpo_layer_caps &c = s->caps;
  GroupHeader h = L.cap;
  h.magic = kGroupMagic;
  h.n_upd = n_upd;
  h.chi = s->chi;
  h.d_max = s->d;
  h.dev_planned = 1;
  h.n_tiles = c.tiles;
  for (int k = 0; k <= kTileClasses; ++k) h.tile_off[k] = k * c.tiles;
  h.n_gram = c.gram;
  h.n_sec = c.sec;
  int off = 0;
  for (int k = 0; k < kTriClass; ++k) {
    const int lo = k == 0 ? 0 : kJacLimit[k - 1];
    h.jac_off[k] = off;
    const bool reachable = lo < thr && c.jac > 0;
    h.jac_max_n[k] = reachable ? std::min(kJacLimit[k], thr) : 0;
    if (reachable) off += c.jac;
  }
  h.jac_off[kTriClass] = off;
  h.n_tri = c.tri;
  h.n_ts = 0;
  h.tri_nbig = c.tri;
  h.tri_split = tri_split();
  h.tri_max_n = s->chi;
  h.n_v = c.v;
  h.n_w = c.w;
  h.n_wyt = c.wyt;
  h.n_wya = c.wya;
  h.n_bc = 0;
  h.max_sv = 4096;  // k_scatter's grid width (it walks the kept values grid-stride)
  h.staging_bytes = L.staging_bytes;
  return h;
}

// The parts of a layer's work, each launched on one stream: the planning, and
// per group the head of its stages (contraction, Gram, Jacobi route), the two
// Householder launches and the rest (bisection to scatter).
int part_plan(LayerProgram &p, hipStream_t st) {
  if (int rc = launch_bs_table_layer(p.params, p.d_upd, st)) return rc;
  return launch_device_plan(p.params, p.d_upd, p.d_groups, p.d_sizes, st);
}
int part_head(LayerProgram &p, int g, hipStream_t st) {
  if (int rc = launch_contract(p.cap[g], p.staging[g], p.precision, st)) return rc;
  return launch_svd_head(p.cap[g], p.staging[g], st);
}
int part_tridiag(LayerProgram &p, int g, int part, hipStream_t st) {
  return launch_tridiag_part(p.cap[g], p.staging[g], st, part);
}
int part_tail(LayerProgram &p, int g, hipStream_t st) {
  if (int rc = launch_tri_sturm(p.cap[g], p.staging[g], st)) return rc;
  return launch_truncate(p.cap[g], p.staging[g], st);
}

// One part captured on `st` (a capture of one stream) and instantiated; a
// part with no launch gives no graph (*exec stays null).
template <class F>
int capture_part(hipStream_t st, hipGraphExec_t *exec, F launch) {
  *exec = nullptr;
  BSMPO_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = launch();
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail((int)e, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  size_t n = 0;
  hipError_t e2 = hipGraphGetNodes(graph, nullptr, &n);
  if (e2 == hipSuccess && n > 0) e2 = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e2 != hipSuccess) return fail((int)e2, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
  return 0;
}

// The layer's graphs.  HIP's graph executor runs the nodes of one graph one
// after another, so the concurrency of a layer (the two arenas' chains, the
// two Householder block shapes side by side) comes from launching the parts'
// graphs on several streams (layer_launch).
int build_graphs(LayerProgram &p, hipStream_t st) {
  if (int rc = capture_part(st, &p.plan, [&] { return part_plan(p, st); })) return rc;
  p.parts.assign(p.cap.size(), {nullptr, nullptr, nullptr, nullptr});
  for (int g = 0; g < (int)p.cap.size(); ++g) {
    auto &q = p.parts[g];
    if (int rc = capture_part(st, &q[0], [&] { return part_head(p, g, st); })) return rc;
    if (int rc = capture_part(st, &q[1], [&] { return part_tridiag(p, g, 0, st); })) return rc;
    if (int rc = capture_part(st, &q[2], [&] { return part_tridiag(p, g, 1, st); })) return rc;
    if (int rc = capture_part(st, &q[3], [&] { return part_tail(p, g, st); })) return rc;
  }
  return 0;
}

// One replay: the planning graph on `st`, then the groups in turn on their
// arena's chain (group g on chain g % n_streams: chain 0 is the caller's
// stream), each as head -> (512-thread Householder on the chain, 256-thread
// on the chain's side stream) -> tail; the caller's stream joins the chains.
int layer_launch(LayerProgram &p, hipStream_t st) {
  BSMPO_HIP(hipGraphLaunch(p.plan, st));
  p.chain[0] = st;
  if (p.n_streams > 1) {
    BSMPO_HIP(hipEventRecord(p.ev_plan, st));
    for (int c = 1; c < p.n_streams; ++c) BSMPO_HIP(hipStreamWaitEvent(p.chain[c], p.ev_plan, 0));
  }
  for (int g = 0; g < (int)p.parts.size(); ++g) {
    const int c = g % p.n_streams;
    const hipStream_t cs = p.chain[c], ss = p.side[c];
    auto &q = p.parts[g];
    if (q[0]) BSMPO_HIP(hipGraphLaunch(q[0], cs));
    if (q[2]) {
      BSMPO_HIP(hipEventRecord(p.ev_head[g], cs));
      BSMPO_HIP(hipStreamWaitEvent(ss, p.ev_head[g], 0));
      BSMPO_HIP(hipGraphLaunch(q[2], ss));
      BSMPO_HIP(hipEventRecord(p.ev_small[g], ss));
    }
    if (q[1]) BSMPO_HIP(hipGraphLaunch(q[1], cs));
    if (q[2]) BSMPO_HIP(hipStreamWaitEvent(cs, p.ev_small[g], 0));
    if (q[3]) BSMPO_HIP(hipGraphLaunch(q[3], cs));
  }
  for (int c = 1; c < p.n_streams; ++c) {
    BSMPO_HIP(hipEventRecord(p.ev_done[c], p.chain[c]));
    BSMPO_HIP(hipStreamWaitEvent(st, p.ev_done[c], 0));
  }
  return 0;
}

int create(const bsmpo_layer_spec *s, void *d_work, size_t work_bytes, hipStream_t stream,
           LayerProgram **out) {
  Layout L;
  layout(s, L);
  if (!d_work) return fail(-1, "layer: workspace is NULL");
  if ((int64_t)work_bytes < L.total) return fail(-1, "layer: workspace too small");
  if (timing_on()) return fail(-1, "layer: disable stage timing before creating a program");
  char *w = reinterpret_cast<char *>(d_work);
  auto *p = new LayerProgram();
  std::unique_ptr<LayerProgram> guard(p);
  BSMPO_HIP(hipGetDevice(&p->device));
  p->precision = s->precision;
  p->n_streams = s->n_streams;
  DpParams &P = p->params;
  P.d = s->d;
  P.chi = s->chi;
  P.thr = eig_threshold();
  P.tri_split = tri_split();
  P.tile_order = tile_order();
  P.n_groups = s->n_groups;
  P.n_layers = s->n_layers;
  P.n_sets = s->n_sets;
  P.max_upd = max_upd(s);
  P.n_upd[0] = s->n_upd[0];
  P.n_upd[1] = s->n_sets == 2 ? s->n_upd[1] : 0;
  P.off[0] = 0;
  P.off[1] = s->n_upd[0];
  P.tie_rel = tie_rel();
  P.drop_rel = drop_rel();
  P.fixed_bytes = L.fixed_bytes;
  P.state = reinterpret_cast<int32_t *>(w + L.o_state);
  P.B = s->d_B;
  p->d_upd = reinterpret_cast<DpUpdate *>(w + L.o_upd);
  p->d_groups = reinterpret_cast<DpGroup *>(w + L.o_groups);
  p->d_sizes = reinterpret_cast<DpSizes *>(w + L.o_sizes);
  // host images of the update and group records
  std::vector<DpUpdate> hu(P.n_upd[0] + P.n_upd[1]);
  std::vector<DpGroup> hg(s->n_groups);
  const int64_t u3_elems = (int64_t)s->d * s->d * s->d;
  const int mg = max_group(s);
  for (int g = 0; g < s->n_groups; ++g) {
    DpGroup &G = hg[g];
    std::memset(&G, 0, sizeof(G));
    G.staging = w + L.o_staging[g];
    G.fixed = w + L.o_fixed[g];
    G.var = w + L.o_var[g];
    G.work = w + L.o_work[g % s->n_streams];
    G.tri_tmp = reinterpret_cast<int2 *>(w + L.o_tri_tmp[g]);
    G.caps = DpCaps{s->caps.var_bytes, s->caps.work_bytes, s->caps.tiles, s->caps.gram,
                    s->caps.jac, s->caps.tri, s->caps.sec, s->caps.v, s->caps.w,
                    s->caps.wyt, s->caps.wya, 0};
    G.o_desc = L.cap.o_desc;
    G.o_tiles = L.cap.o_tiles;
    G.o_gram = L.cap.o_gram;
    G.o_jac = L.cap.o_jac;
    G.o_tri = L.cap.o_tri;
    G.o_sec = L.cap.o_sec;
    G.o_v = L.cap.o_v;
    G.o_w = L.cap.o_w;
    G.o_wyt = L.cap.o_wyt;
    G.o_wya = L.cap.o_wya;
    for (int k = 0; k < s->n_sets; ++k) {
      G.first[k] = group_first(s, k, g);
      G.n_upd[k] = group_size(s, k, g);
      for (int i = 0; i < G.n_upd[k]; ++i) {
        const int idx = G.first[k] + i;  // index inside the set
        const bsmpo_layer_update &x = s->upd[k][idx];
        DpUpdate &U = hu[P.off[k] + idx];
        U.CL = x.CL;
        U.CC = x.CC;
        U.CR = x.CR;
        U.Glc = x.Glc;
        U.Gcr = x.Gcr;
        U.LR = x.LR;
        // the sets never run together: update idx of either set uses table idx
        U.U3 = reinterpret_cast<double *>(w + L.o_u3) + 2 * u3_elems * idx;
        U.charge_c = x.charge_c;
        U.lambda_c = x.lambda_c;
        U.num_lambda = x.num_lambda;
        U.edge = x.edge;
        U.group = g;
        U.local = i;
        U.first_replay = x.first_replay;
      }
    }
    p->cap.push_back(capacity_header(s, L, mg, P.thr));
    p->staging.push_back(G.staging);
  }
  trace("records built");
  BSMPO_HIP(hipMemcpyAsync(p->d_upd, hu.data(), sizeof(DpUpdate) * hu.size(), hipMemcpyHostToDevice,
                           stream));
  BSMPO_HIP(hipMemcpyAsync(p->d_groups, hg.data(), sizeof(DpGroup) * hg.size(),
                           hipMemcpyHostToDevice, stream));
  if (int rc = launch_layer_seek(P.state, 0, stream)) return rc;
  trace("records uploaded");
  // A dry pass of every part with every count zero (the kernels start and
  // return): raises the kernels' LDS limits outside the capture.
  for (int g = 0; g < s->n_groups; ++g)
    BSMPO_HIP(hipMemsetAsync(p->staging[g], 0, sizeof(GroupCounts), stream));
  trace("dry pass");
  for (int g = 0; g < s->n_groups; ++g) {
    int rc = part_head(*p, g, stream);
    if (!rc) rc = part_tridiag(*p, g, 0, stream);
    if (!rc) rc = part_tridiag(*p, g, 1, stream);
    if (!rc) rc = part_tail(*p, g, stream);
    if (rc) return rc;
  }
  BSMPO_HIP(hipStreamSynchronize(stream));
  trace("dry pass done; capturing the parts");
  if (int rc = build_graphs(*p, stream)) {
    trace(bsmpo_last_error());
    return rc;
  }
  for (int c = 0; c < p->n_streams; ++c) {
    BSMPO_HIP(hipStreamCreateWithFlags(&p->side[c], hipStreamNonBlocking));
    if (c > 0) {
      BSMPO_HIP(hipStreamCreateWithFlags(&p->chain[c], hipStreamNonBlocking));
      BSMPO_HIP(hipEventCreateWithFlags(&p->ev_done[c], hipEventDisableTiming));
    }
  }
  BSMPO_HIP(hipEventCreateWithFlags(&p->ev_plan, hipEventDisableTiming));
  p->ev_head.assign(p->parts.size(), nullptr);
  p->ev_small.assign(p->parts.size(), nullptr);
  for (size_t g = 0; g < p->parts.size(); ++g) {
    BSMPO_HIP(hipEventCreateWithFlags(&p->ev_head[g], hipEventDisableTiming));
    BSMPO_HIP(hipEventCreateWithFlags(&p->ev_small[g], hipEventDisableTiming));
  }
  trace("instantiated");
  *out = guard.release();
  return 0;
}

LayerProgram *as_program(void *prog) { return reinterpret_cast<LayerProgram *>(prog); }

}  // namespace
}  // namespace bsmpo

using namespace bsmpo;

extern "C" {

int bsmpo_layer_query(const bsmpo_layer_spec *spec, size_t *work_bytes) {
  if (int rc = check_spec(spec)) return rc;
  if (!work_bytes) return fail(-1, "layer_query: NULL work_bytes");
  Layout L;
  layout(spec, L);
  *work_bytes = (size_t)L.total;
  return 0;
}

int bsmpo_layer_create(const bsmpo_layer_spec *spec, void *d_work, size_t work_bytes,
                       void *stream, void **prog) {
  if (!prog) return fail(-1, "layer_create: NULL prog");
  *prog = nullptr;
  if (int rc = check_spec(spec)) return rc;
  if (!stream) return fail(-1, "layer_create: capture needs a non-default stream");
  LayerProgram *p = nullptr;
  if (int rc = create(spec, d_work, work_bytes, (hipStream_t)stream, &p)) return rc;
  *prog = p;
  return 0;
}

int bsmpo_layer_seek(void *prog, int replay, void *stream) {
  if (!prog) return fail(-1, "layer_seek: NULL program");
  if (replay < 0) return fail(-1, "layer_seek: negative replay");
  return launch_layer_seek(as_program(prog)->params.state, replay, (hipStream_t)stream);
}

int bsmpo_layer_launch(void *prog, void *stream) {
  if (!prog) return fail(-1, "layer_launch: NULL program");
  return layer_launch(*as_program(prog), (hipStream_t)stream);
}

int bsmpo_layer_state_read(void *prog, bsmpo_layer_state *state, bsmpo_layer_caps *group_needs) {
  if (!prog || !state) return fail(-1, "layer_state_read: NULL argument");
  LayerProgram *p = as_program(prog);
  BSMPO_HIP(hipDeviceSynchronize());
  int32_t st[4];
  BSMPO_HIP(hipMemcpy(st, p->params.state, sizeof(st), hipMemcpyDeviceToHost));
  state->status = st[0];
  state->replays = st[1];
  state->failed_at = st[2];
  state->n_groups = p->params.n_groups;
  if (group_needs) {
    std::vector<DpGroup> hg(p->params.n_groups);
    BSMPO_HIP(hipMemcpy(hg.data(), p->d_groups, sizeof(DpGroup) * hg.size(), hipMemcpyDeviceToHost));
    for (size_t g = 0; g < hg.size(); ++g) {
      const DpNeeds &n = hg[g].needs;
      group_needs[g] = bsmpo_layer_caps{n.var_bytes, n.work_bytes, n.tiles, n.gram, n.jac, n.tri,
                                        n.sec, n.v, n.w, n.wyt, n.wya, n.status};
    }
  }
  return 0;
}

int bsmpo_layer_destroy(void *prog) {
  delete as_program(prog);
  return 0;
}

int bsmpo_layer_needs(int n_upd, const void *const *h_plans, bsmpo_layer_caps *needs) {
  if (n_upd < 1 || !h_plans || !needs) return fail(-1, "layer_needs: bad argument");
  bsmpo_layer_caps c{};
  for (int u = 0; u < n_upd; ++u) {
    const PlanHeader *h = header(h_plans[u]);
    if (!h || h->magic != kPlanMagic) return fail(-1, "layer_needs: entry is not a bsmpo plan");
    c.var_bytes += region(var_layout(h->rows_total, h->cols_total, h->n_gram_items, h->n_w_items,
                                     h->n_v_items, h->sv_total, h->n_titems).bytes);
    c.work_bytes += region(h->w_total - h->w_C);
    c.tiles += h->n_titems;
    c.gram += h->n_gram_items;
    c.jac += h->jac_off[kTriClass] - h->jac_off[0];
    c.tri += h->jac_off[kTriClass + 1] - h->jac_off[kTriClass];
    c.sec += h->n_sectors;
    c.v += h->n_v_items;
    c.w += h->n_w_items;
    c.wyt += h->n_wyt;
    c.wya += h->n_wya;
  }
  *needs = c;
  return 0;
}

}  // extern "C"
