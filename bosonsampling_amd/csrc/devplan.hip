// Device-side planning of the site updates of a layer program.
//
// From the bond charges in HBM, these kernels write what the host planner
// (plan.cpp) and the group builder (group.cpp) write: per update the charge
// blocks, the tau sectors with their packed buffer offsets, the offset tables
// of the contraction, the row / column pools, the Gram / W / V work items and
// the contraction tile items; per group the update descriptors and the work
// lists with their GroupCounts.  The stage kernels then run unchanged, with
// grids sized for the group's capacity.  Nothing is read back by the host, so
// a whole layer (planning and stages) can be captured once as a HIP graph.
//
// The selection rules are the reference's, as in plan.cpp:
//   tau loop bounds   src/beam_splitter_parallel/Rank.py:78-91
//   sector windows    src/mpo_sort.py:424-432 (rows cl >= tau, cols cr <= tau)
//   charge blocks     src/mpo_sort.py:215-259 (start index of each charge)
//
// Kernels, in launch order (one launch each covers every group):
//   k_dp_sizes  one block per update: blocks, sectors, tables, sizes
//   k_dp_scan   one block per group: arena offsets and list positions by a
//               scan over the group's updates, capacity check, GroupCounts
//   k_dp_fill   (update, sector stripe) blocks: pools, items, descriptors,
//               the sector / Gram / W / V / Jacobi lists, the unsorted
//               tridiagonal list
//   k_dp_tiles  one block per update: contraction tile items and their list
//   k_dp_tri    one block per group: tridiagonal list largest first, block
//               shape split, back-transformation items
//   k_dp_gate   one block: any failure (or an earlier one) zeroes every
//               group's counts, so the replay leaves the state untouched
#include "common.h"

namespace bsmpo {
namespace {

constexpr int kSizeThreads = 1024;
constexpr int kKeys = kMaxD * kMaxD;  // charge keys c0 * d + c1, c0, c1 < d
constexpr int kMaxChi = kDpMaxChi;
constexpr int kFillStripes = 16;      // k_dp_fill blocks per update

__device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }
__device__ inline int64_t region(int64_t bytes) { return dp_align(bytes > 8 ? bytes : 8); }

__device__ inline char *fixed_of(const DpParams &P, const DpGroup &G, int local) {
  return G.fixed + (int64_t)local * P.fixed_bytes;
}

// ----------------------------------------------------- contraction tiles --
// The 16 x 16 tiles of (left row, right column) elements of an update, in the
// host planner's order: super-tiles of S x S tiles (S = the tile order) row by
// row, tiles row by row inside each; the linear index maps to the tile origin.
__device__ inline void tile_origin(int idx, int nTR, int nTC, int S, int &r0, int &c0) {
  const int SR = idx / (S * nTC);
  int rem = idx - SR * S * nTC;
  const int rows = min(S, nTR - S * SR);
  const int SC = rem / (rows * S);
  rem -= SC * rows * S;
  const int cols = min(S, nTC - S * SC);
  r0 = 16 * (S * SR + rem / cols);
  c0 = 16 * (S * SC + rem % cols);
}

struct TileCtx {
  const int4 *lb, *rb;        // blocks (c0, c1, start, end) of the left / right bond
  const int16_t *rowblk, *colblk;  // block of each slot, -1 for padding
  int nl, nr, nLv, nRv, order;
};

// One pass over rows [r0 + s0, r0 + s0 + rp) of the tile (r0, c0): whether
// any (left block, right block) pair of it has a non-empty box, its LDS need
// and its item (plan.cpp make_pass).
__device__ inline bool make_pass(const TileCtx &t, int r0, int c0, int tnr, int tnc, int s0,
                                 int rp, int lo0, int lo1, int c_lo, int c_hi, int &need,
                                 TileItem &it) {
  int hi0 = -1, hi1 = -1, a_lo = t.nl, a_hi = -1;
  const int iend = r0 + min(tnr, s0 + rp);
  for (int i = r0 + s0; i < iend; ++i) {
    const int a = t.rowblk[i];
    if (a < 0) continue;
    hi0 = max(hi0, t.lb[a].x);
    hi1 = max(hi1, t.lb[a].y);
    a_lo = min(a_lo, a);
    a_hi = max(a_hi, a);
  }
  bool any = false;
  for (int a = a_lo; a <= a_hi && !any; ++a)
    for (int c = c_lo; c <= c_hi && !any; ++c)
      any = t.lb[a].x >= t.rb[c].x && t.lb[a].y >= t.rb[c].y;
  if (!any) return false;
  const int nbox = (hi0 - lo0 + 1) * (hi1 - lo1 + 1);
  const int na = a_hi - a_lo + 1, nbk = c_hi - c_lo + 1;
  need = nbox * rp * 16 + (nbox * (12 + 4 * (na + nbk)) + 7) / 8;
  it.r0 = r0;
  it.c0 = c0;
  it.nr = (int8_t)tnr;
  it.nc = (int8_t)tnc;
  it.s0 = (int8_t)s0;
  it.rp = (int8_t)rp;
  it.lo0 = (int8_t)lo0;
  it.lo1 = (int8_t)lo1;
  it.hi0 = (int8_t)hi0;
  it.hi1 = (int8_t)hi1;
  it.a_lo = (int16_t)a_lo;
  it.b_lo = (int16_t)c_lo;
  it.na = (int8_t)na;
  it.nb = (int8_t)nbk;
  it.pad0 = it.pad1 = 0;
  return true;
}

__device__ inline int tile_class(int need) {
  int c = 0;
  while (c < kTileClasses - 1 && need > kTileElems[c]) ++c;
  return c;
}

// The passes of tile idx: the largest row pass (16, 8, ..., 1) whose passes
// all fit the largest LDS class, each handed to emit(item, class) in row
// order.  Returns false if even single-row passes do not fit, or the tile
// has no column of a block (then it has no passes, as on the host).
template <class Emit>
__device__ bool tile_passes(const TileCtx &t, int idx, Emit emit) {
  const int nTR = cdiv(t.nLv, 16), nTC = cdiv(t.nRv, 16);
  int r0, c0;
  tile_origin(idx, nTR, nTC, t.order, r0, c0);
  const int tnr = min(16, t.nLv - r0), tnc = min(16, t.nRv - c0);
  int lo0 = kMaxD, lo1 = kMaxD, c_lo = t.nr, c_hi = -1;
  for (int j = c0; j < c0 + tnc; ++j) {
    const int c = t.colblk[j];
    if (c < 0) continue;
    lo0 = min(lo0, t.rb[c].x);
    lo1 = min(lo1, t.rb[c].y);
    c_lo = min(c_lo, c);
    c_hi = max(c_hi, c);
  }
  if (c_hi < 0) return true;
  constexpr int kLast = kTileElems[kTileClasses - 1];
  int rp = 16;
  for (; rp >= 1; rp >>= 1) {
    bool fits = true;
    for (int s0 = 0; s0 < tnr && fits; s0 += rp) {
      int need;
      TileItem it;
      if (make_pass(t, r0, c0, tnr, tnc, s0, rp, lo0, lo1, c_lo, c_hi, need, it) && need > kLast)
        fits = false;
    }
    if (fits) break;
  }
  if (rp < 1) return false;
  for (int s0 = 0; s0 < tnr; s0 += rp) {
    int need;
    TileItem it;
    if (make_pass(t, r0, c0, tnr, tnc, s0, rp, lo0, lo1, c_lo, c_hi, need, it))
      emit(it, tile_class(need));
  }
  return true;
}

// ------------------------------------------------------------ k_dp_sizes --
// One block per update: validates and copies the three bonds' charges, finds
// their blocks, the tau sectors (with their update-local offsets), the
// contraction tables and the per-class tile counts, and records the sizes of
// the update's variable plan part and scratch.
__global__ __launch_bounds__(kSizeThreads) void k_dp_sizes(DpParams P,
                                                           const DpUpdate *__restrict__ Ud,
                                                           const DpGroup *__restrict__ Gs,
                                                           DpSizes *__restrict__ Sz) {
  const int set = dp_set(P), u = blockIdx.x, tid = threadIdx.x;
  if (u >= P.n_upd[set]) return;
  const DpUpdate up = Ud[P.off[set] + u];
  const DpGroup &G = Gs[up.group];
  DpSizes &sz = Sz[u];
  const int d = P.d, chi = P.chi, D1 = d + 1;
  __shared__ int s_start[3][kKeys], s_end[3][kKeys], s_idx[3][kKeys];
  __shared__ int4 s_blk[3][kKeys];
  __shared__ int s_nblk[3];
  __shared__ int16_t s_rowblk[kMaxChi], s_colblk[kMaxChi];
  __shared__ int s_tL[kKeys], s_tR[kKeys];
  __shared__ int s_sect[kKeys];  // sector -> t0 | t1 << 8
  __shared__ int s_err, s_lc0, s_lc1, s_sr0, s_sr1, s_nsec, s_tcnt[kTileClasses];
  if (P.state[0] != DP_OK) {  // an earlier replay failed: nothing is planned any more
    if (tid == 0) sz.status = DP_SKIPPED;
    return;
  }
  if (P.state[1] < up.first_replay) {  // beyond the light cone in this replay
    if (tid == 0) {
      sz.status = DP_OK;
      sz.active = 0;
    }
    return;
  }
  const FixedLayout F = fixed_layout(d, chi);
  char *fx = fixed_of(P, G, up.local);
  Sector *sectors = reinterpret_cast<Sector *>(fx + F.o_sectors);
  Block *blk_out[3] = {reinterpret_cast<Block *>(fx + F.o_lblk),
                       reinterpret_cast<Block *>(fx + F.o_cblk),
                       reinterpret_cast<Block *>(fx + F.o_rblk)};
  int32_t *tau_sector = reinterpret_cast<int32_t *>(fx + F.o_tau_sector);
  int32_t *cblk_of = reinterpret_cast<int32_t *>(fx + F.o_cblk_of);
  int32_t *rowoff = reinterpret_cast<int32_t *>(fx + F.o_rowoff);
  int32_t *coloff = reinterpret_cast<int32_t *>(fx + F.o_coloff);
  int2 *row_info = reinterpret_cast<int2 *>(fx + F.o_row_info);
  int2 *col_info = reinterpret_cast<int2 *>(fx + F.o_col_info);
  int32_t *charges = reinterpret_cast<int32_t *>(fx + F.o_charges);
  int4 *secx = reinterpret_cast<int4 *>(fx + F.o_secx);

  for (int k = tid; k < 3 * kKeys; k += kSizeThreads) {
    (&s_start[0][0])[k] = -1;
    (&s_end[0][0])[k] = -1;
  }
  for (int k = tid; k < D1 * D1; k += kSizeThreads) {
    tau_sector[k] = -1;
    cblk_of[k] = -1;
  }
  if (tid == 0) {
    s_err = DP_OK;
    s_lc0 = s_lc1 = 0;
    s_sr0 = s_sr1 = d;
    s_nsec = 0;
  }
  if (tid < kTileClasses) s_tcnt[tid] = 0;
  __syncthreads();
  // charges: copy (CL, CC, CR, as the plan keeps them), validate, find blocks
  for (int b = 0; b < 3; ++b) {
    const int32_t *ch = b == 0 ? up.CL : (b == 1 ? up.CC : up.CR);
    for (int i = tid; i < chi; i += kSizeThreads) {
      const int c0 = ch[2 * i], c1 = ch[2 * i + 1];
      charges[2 * chi * b + 2 * i] = c0;
      charges[2 * chi * b + 2 * i + 1] = c1;
      if (c0 < 0 || c1 < 0 || c0 > d || c1 > d) {
        s_err = DP_BAD_CHARGES;
        continue;
      }
      const int key = c0 * D1 + c1;
      const int pk = i > 0 ? ch[2 * i - 2] * D1 + ch[2 * i - 1] : -1;
      const int nk = i + 1 < chi ? ch[2 * i + 2] * D1 + ch[2 * i + 3] : -1;
      if (key < pk) s_err = DP_BAD_CHARGES;
      if (b == 0) {
        if (c0 != d) atomicMax(&s_lc0, c0);
        if (c1 != d) atomicMax(&s_lc1, c1);
      }
      if (b == 2) {
        atomicMin(&s_sr0, c0);
        atomicMin(&s_sr1, c1);
      }
      if (c0 >= d || c1 >= d) continue;
      const int k = c0 * d + c1;
      if (pk != key) s_start[b][k] = i;
      if (nk != key) s_end[b][k] = i + 1;
    }
  }
  __syncthreads();
  if (s_err != DP_OK) {
    if (tid == 0) sz.status = s_err;
    return;
  }
  // blocks in key order (= slot order, the charges being sorted)
  if (tid < 3) {
    const int b = tid;
    int n = 0;
    for (int k = 0; k < d * d; ++k) {
      if (s_start[b][k] < 0) {
        s_idx[b][k] = -1;
        continue;
      }
      const int4 bk = make_int4(k / d, k % d, s_start[b][k], s_end[b][k]);
      s_blk[b][n] = bk;
      blk_out[b][n] = Block{bk.x, bk.y, bk.z, bk.w};
      if (b == 1) cblk_of[bk.x * D1 + bk.y] = n;
      s_idx[b][k] = n++;
    }
    s_nblk[b] = n;
  }
  __syncthreads();
  const int nl = s_nblk[0], nc = s_nblk[1], nr = s_nblk[2];
  for (int i = tid; i < chi; i += kSizeThreads) {
    int c0 = up.CL[2 * i], c1 = up.CL[2 * i + 1];
    int a = c0 < d && c1 < d ? s_idx[0][c0 * d + c1] : -1;
    s_rowblk[i] = (int16_t)a;
    row_info[i] = a >= 0 ? make_int2(a | c0 << 16 | c1 << 24, i - s_blk[0][a].z) : make_int2(-1, 0);
    c0 = up.CR[2 * i];
    c1 = up.CR[2 * i + 1];
    a = c0 < d && c1 < d ? s_idx[2][c0 * d + c1] : -1;
    s_colblk[i] = (int16_t)a;
    col_info[i] = a >= 0 ? make_int2(a | c0 << 16 | c1 << 24, i - s_blk[2][a].z) : make_int2(-1, 0);
  }
  // tau candidates: L = rows of left blocks >= tau, R = columns of right blocks <= tau
  const int sr0 = s_sr0, sr1 = s_sr1, nt0 = s_lc0 - sr0 + 1, nt1 = s_lc1 - sr1 + 1;
  const int ntau = nt0 > 0 && nt1 > 0 ? nt0 * nt1 : 0;
  for (int t = tid; t < ntau; t += kSizeThreads) {
    const int t0 = sr0 + t / nt1, t1 = sr1 + t % nt1;
    int L = 0, R = 0;
    for (int a = 0; a < nl; ++a)
      if (s_blk[0][a].x >= t0 && s_blk[0][a].y >= t1) L += s_blk[0][a].w - s_blk[0][a].z;
    for (int c = 0; c < nr; ++c)
      if (s_blk[2][c].x <= t0 && s_blk[2][c].y <= t1) R += s_blk[2][c].w - s_blk[2][c].z;
    s_tL[t] = L;
    s_tR[t] = R;
  }
  __syncthreads();
  // sectors in tau order (t0 outer, t1 inner) with their update-local offsets
  if (tid == 0) {
    int ns = 0, rows = 0, cols = 0, ngram = 0, nw = 0, nv = 0, sv = 0, ntri = 0;
    int jac[kTriClass] = {};
    int64_t c_off = 0, g_off = 0, v_off = 0, w_off = 0, tw = 0;
    int err = DP_OK;
    for (int t = 0; t < ntau; ++t) {
      const int L = s_tL[t], R = s_tR[t];
      if (L == 0 || R == 0) continue;  // Rank.py:108-111: skipped sector
      const int t0 = sr0 + t / nt1, t1 = sr1 + t % nt1;
      Sector s{};
      s.t0 = t0;
      s.t1 = t1;
      s.L = L;
      s.R = R;
      s.n = min(L, R);
      s.row_off = rows;
      s.col_off = cols;
      s.left_small = L <= R ? 1 : 0;
      s.c_off = c_off;
      s.g_off = g_off;
      s.v_off = v_off;
      s.w_off = w_off;
      s.s_off = sv;
      const int n = s.n;
      const bool tri = n > P.thr;
      int rank;
      if (tri) {
        if (n > kTriMaxN) err = DP_SECTOR_TOO_BIG;
        s.tw_off = tw;
        tw += tri_work_doubles(n);
        rank = ntri++;
      } else {
        int c = 0;
        while (n > kJacLimit[c]) ++c;
        rank = jac[c]++;
      }
      const int nb = cdiv(n, kGramBlock);
      secx[ns] = make_int4(ngram, nw, nv, rank);
      ngram += tri ? nb * (nb + 1) / 2 : nb * nb;
      nw += nb * cdiv(R, kGramBlock);
      if (!s.left_small) nv += cdiv(L, kGramBlock) * nb;
      rows += L;
      cols += R;
      c_off += (int64_t)L * R;
      g_off += (int64_t)n * n;
      v_off += (int64_t)L * n;
      w_off += (int64_t)n * R;
      sv += n;
      sectors[ns] = s;
      tau_sector[t0 * D1 + t1] = ns;
      s_sect[ns] = t0 | t1 << 8;
      ++ns;
    }
    s_nsec = ns;
    sz.status = err;
    sz.active = 1;
    sz.n_sec = ns;
    sz.n_lblk = nl;
    sz.n_cblk = nc;
    sz.n_rblk = nr;
    sz.rows_total = rows;
    sz.cols_total = cols;
    sz.n_gram = ngram;
    sz.n_w = nw;
    sz.n_v = nv;
    sz.sv_total = sv;
    sz.n_tri = ntri;
    for (int c = 0; c < kTriClass; ++c) sz.jac_cnt[c] = jac[c];
    sz.c_elems = c_off;
    sz.g_elems = g_off;
    sz.v_elems = v_off;
    sz.w_elems = w_off;
    sz.tw = tw;
  }
  __syncthreads();
  // row / column offset of every block inside every sector (-1: not in it)
  const int ns = s_nsec;
  for (int sid = tid; sid < ns; sid += kSizeThreads) {
    const int t0 = s_sect[sid] & 255, t1 = s_sect[sid] >> 8;
    int o = 0;
    for (int a = 0; a < nl; ++a) {
      const bool in = s_blk[0][a].x >= t0 && s_blk[0][a].y >= t1;
      rowoff[sid * nl + a] = in ? o : -1;
      if (in) o += s_blk[0][a].w - s_blk[0][a].z;
    }
    o = 0;
    for (int c = 0; c < nr; ++c) {
      const bool in = s_blk[2][c].x <= t0 && s_blk[2][c].y <= t1;
      coloff[sid * nr + c] = in ? o : -1;
      if (in) o += s_blk[2][c].w - s_blk[2][c].z;
    }
  }
  // contraction tile passes per LDS class
  TileCtx tc;
  tc.lb = s_blk[0];
  tc.rb = s_blk[2];
  tc.rowblk = s_rowblk;
  tc.colblk = s_colblk;
  tc.nl = nl;
  tc.nr = nr;
  tc.nLv = nl > 0 ? s_blk[0][nl - 1].w : 0;
  tc.nRv = nr > 0 ? s_blk[2][nr - 1].w : 0;
  tc.order = P.tile_order;
  const int ntiles = cdiv(tc.nLv, 16) * cdiv(tc.nRv, 16);
  int cnt[kTileClasses] = {};
  bool fits = true;
  for (int idx = tid; idx < ntiles; idx += kSizeThreads)
    fits &= tile_passes(tc, idx, [&](const TileItem &, int c) { ++cnt[c]; });
  if (!fits) s_err = DP_TILE_TOO_BIG;
  for (int c = 0; c < kTileClasses; ++c)
    if (cnt[c]) atomicAdd(&s_tcnt[c], cnt[c]);
  __syncthreads();
  if (tid == 0) {
    if (s_err != DP_OK && sz.status == DP_OK) sz.status = s_err;
    int nt = 0;
    for (int c = 0; c < kTileClasses; ++c) {
      sz.tile_cnt[c] = s_tcnt[c];
      nt += s_tcnt[c];
    }
    sz.n_titems = nt;
    sz.var_bytes = var_layout(sz.rows_total, sz.cols_total, sz.n_gram, sz.n_w, sz.n_v,
                              sz.sv_total, nt).bytes;
    sz.work_bytes = work_layout(sz.c_elems, sz.g_elems, sz.v_elems, sz.w_elems, sz.sv_total,
                                ns, sz.tw).bytes;
  }
}

// ------------------------------------------------------------- k_dp_scan --
// One block per group (thread 0 walks its updates): arena offsets, list
// positions, the capacity check and the group's counts.
__global__ void k_dp_scan(DpParams P, DpGroup *__restrict__ Gs, DpSizes *__restrict__ Sz) {
  if (threadIdx.x != 0) return;
  const int set = dp_set(P);
  DpGroup &G = Gs[blockIdx.x];
  GroupCounts &K = *reinterpret_cast<GroupCounts *>(G.staging);
  const int first = G.first[set], nupd = G.n_upd[set];
  if (P.state[0] != DP_OK) {  // an earlier replay failed: keep the needs it recorded
    for (int k = 0; k < SG_COUNT; ++k) K.seg[k] = make_int2(0, 0);
    K.status = DP_SKIPPED;
    return;
  }
  int status = DP_OK;
  int64_t var = 0, work = 0;
  int64_t tiles[kTileClasses] = {}, jac[kTriClass] = {};
  int64_t gram = 0, tri = 0, sec = 0, nv = 0, nw = 0;
  int nact = 0;
  for (int i = 0; i < nupd; ++i) {
    DpSizes &s = Sz[first + i];
    if (s.status != DP_OK && status == DP_OK) status = s.status;
    if (s.status != DP_OK || !s.active) continue;
    s.desc = nact++;
    s.var_off = var;
    var += region(s.var_bytes);
    s.work_off = work;
    work += region(s.work_bytes);
    for (int c = 0; c < kTileClasses; ++c) {
      s.l_tile[c] = (int32_t)tiles[c];
      tiles[c] += s.tile_cnt[c];
    }
    for (int c = 0; c < kTriClass; ++c) {
      s.l_jac[c] = (int32_t)jac[c];
      jac[c] += s.jac_cnt[c];
    }
    s.l_gram = (int32_t)gram;
    gram += s.n_gram;
    s.l_tri = (int32_t)tri;
    tri += s.n_tri;
    s.l_sec = (int32_t)sec;
    sec += s.n_sec;
    s.l_v = (int32_t)nv;
    nv += s.n_v;
    s.l_w = (int32_t)nw;
    nw += s.n_w;
  }
  int64_t tile_off[kTileClasses + 1], jac_off[kTriClass + 1];
  tile_off[0] = jac_off[0] = 0;
  for (int c = 0; c < kTileClasses; ++c) tile_off[c + 1] = tile_off[c] + tiles[c];
  for (int c = 0; c < kTriClass; ++c) jac_off[c + 1] = jac_off[c] + jac[c];
  for (int i = 0; i < nupd && status == DP_OK; ++i) {
    DpSizes &s = Sz[first + i];
    if (!s.active) continue;
    for (int c = 0; c < kTileClasses; ++c) s.l_tile[c] += (int32_t)tile_off[c];
    for (int c = 0; c < kTriClass; ++c) s.l_jac[c] += (int32_t)jac_off[c];
  }
  G.needs.var_bytes = var;
  G.needs.work_bytes = work;
  G.needs.tiles = (int32_t)min<int64_t>(tile_off[kTileClasses], INT32_MAX);
  G.needs.jac = (int32_t)min<int64_t>(jac_off[kTriClass], INT32_MAX);
  G.needs.gram = (int32_t)min<int64_t>(gram, INT32_MAX);
  G.needs.tri = (int32_t)min<int64_t>(tri, INT32_MAX);
  G.needs.sec = (int32_t)min<int64_t>(sec, INT32_MAX);
  G.needs.v = (int32_t)min<int64_t>(nv, INT32_MAX);
  G.needs.w = (int32_t)min<int64_t>(nw, INT32_MAX);
  G.needs.wyt = G.needs.wya = 0;
  if (status == DP_OK &&
      (var > G.caps.var_bytes || work > G.caps.work_bytes || tile_off[kTileClasses] > G.caps.tiles ||
       jac_off[kTriClass] > G.caps.jac || gram > G.caps.gram || tri > G.caps.tri ||
       sec > G.caps.sec || nv > G.caps.v || nw > G.caps.w))
    status = DP_OVER_CAPACITY;
  G.needs.status = status;
  for (int k = 0; k < SG_COUNT; ++k) K.seg[k] = make_int2(0, 0);
  K.status = status;
  if (status != DP_OK) return;
  K.seg[SG_UPD] = make_int2(0, nact);
  for (int c = 0; c < kTileClasses; ++c)
    K.seg[SG_TILE + c] = make_int2((int)tile_off[c], (int)tiles[c]);
  K.seg[SG_GRAM] = make_int2(0, (int)gram);
  for (int c = 0; c < kTriClass; ++c) K.seg[SG_JAC + c] = make_int2((int)jac_off[c], (int)jac[c]);
  K.seg[SG_JACALL] = make_int2(0, (int)jac_off[kTriClass]);
  K.seg[SG_TRI] = make_int2(0, (int)tri);
  K.seg[SG_SEC] = make_int2(0, (int)sec);
  K.seg[SG_V] = make_int2(0, (int)nv);
  K.seg[SG_W] = make_int2(0, (int)nw);
}

// ------------------------------------------------------------- k_dp_fill --
// Blocks (update, stripe): the update's descriptor (stripe 0), then for the
// stripe's sectors the row / column pools, the Gram / W / V items, the
// singular-value owners and the list entries.
__global__ __launch_bounds__(256) void k_dp_fill(DpParams P, const DpUpdate *__restrict__ Ud,
                                                 const DpGroup *__restrict__ Gs,
                                                 const DpSizes *__restrict__ Sz) {
  const int set = dp_set(P), u = blockIdx.x, tid = threadIdx.x;
  if (u >= P.n_upd[set]) return;
  const DpUpdate up = Ud[P.off[set] + u];
  const DpGroup &G = Gs[up.group];
  if (reinterpret_cast<const GroupCounts *>(G.staging)->status != DP_OK) return;
  const DpSizes s = Sz[u];
  if (!s.active) return;
  const int d = P.d, chi = P.chi;
  const FixedLayout F = fixed_layout(d, chi);
  char *fx = fixed_of(P, G, up.local);
  const VarLayout V = var_layout(s.rows_total, s.cols_total, s.n_gram, s.n_w, s.n_v, s.sv_total,
                                 s.n_titems);
  char *vb = G.var + s.var_off;
  const Sector *sectors = reinterpret_cast<const Sector *>(fx + F.o_sectors);
  const Block *lblk = reinterpret_cast<const Block *>(fx + F.o_lblk);
  const Block *rblk = reinterpret_cast<const Block *>(fx + F.o_rblk);
  const int32_t *rowoff = reinterpret_cast<const int32_t *>(fx + F.o_rowoff);
  const int32_t *coloff = reinterpret_cast<const int32_t *>(fx + F.o_coloff);
  const int4 *secx = reinterpret_cast<const int4 *>(fx + F.o_secx);
  int32_t *rows = reinterpret_cast<int32_t *>(vb + V.o_rows);
  int32_t *cols = reinterpret_cast<int32_t *>(vb + V.o_cols);
  Item *gram = reinterpret_cast<Item *>(vb + V.o_gram);
  Item *witems = reinterpret_cast<Item *>(vb + V.o_w);
  Item *vitems = reinterpret_cast<Item *>(vb + V.o_v);
  int32_t *sv_sector = reinterpret_cast<int32_t *>(vb + V.o_sv);
  int2 *l_gram = reinterpret_cast<int2 *>(G.staging + G.o_gram);
  int2 *l_w = reinterpret_cast<int2 *>(G.staging + G.o_w);
  int2 *l_v = reinterpret_cast<int2 *>(G.staging + G.o_v);
  int2 *l_sec = reinterpret_cast<int2 *>(G.staging + G.o_sec);
  int2 *l_jac = reinterpret_cast<int2 *>(G.staging + G.o_jac);
  const int me = s.desc;  // descriptor and list index among the group's active updates
  if (blockIdx.y == 0 && tid == 0) {
    const WorkLayout W = work_layout(s.c_elems, s.g_elems, s.v_elems, s.w_elems, s.sv_total,
                                     s.n_sec, s.tw);
    char *wb = G.work + s.work_off;
    UpdDesc D;
    D.sectors = sectors;
    D.rows = rows;
    D.cols = cols;
    D.lblk = lblk;
    D.cblk = reinterpret_cast<const Block *>(fx + F.o_cblk);
    D.rblk = rblk;
    D.gram_items = gram;
    D.w_items = witems;
    D.v_items = vitems;
    D.charges = reinterpret_cast<const int32_t *>(fx + F.o_charges);
    D.sv_sector = sv_sector;
    D.C = reinterpret_cast<double *>(wb + W.C);
    D.T = reinterpret_cast<double *>(wb + W.T);
    D.G = reinterpret_cast<double *>(wb + W.G);
    D.V = reinterpret_cast<double *>(wb + W.V);
    D.W = reinterpret_cast<double *>(wb + W.W);
    D.S = reinterpret_cast<double *>(wb + W.S);
    D.lam = reinterpret_cast<double *>(wb + W.lam);
    D.td = reinterpret_cast<double *>(wb + W.td);
    D.te = reinterpret_cast<double *>(wb + W.te);
    D.tau = reinterpret_cast<double *>(wb + W.tau);
    D.triwork = reinterpret_cast<double *>(wb + W.triwork);
    D.flag = reinterpret_cast<int32_t *>(wb + W.flag);
    D.slot = reinterpret_cast<int32_t *>(wb + W.slot);
    D.kcount = reinterpret_cast<int32_t *>(wb + W.kcount);
    D.hist = reinterpret_cast<int32_t *>(wb + W.hist);
    D.Glc = up.Glc;
    D.Gcr = up.Gcr;
    D.LR = up.LR;
    D.U3 = up.U3;
    D.LL = nullptr;
    D.LC = nullptr;
    D.charge_c = up.charge_c;
    D.lambda_c = up.lambda_c;
    D.num_lambda = up.num_lambda;
    D.d = d;
    D.chi = chi;
    D.edge = up.edge;
    D.n_sectors = s.n_sec;
    D.n_cblk = s.n_cblk;
    D.eig_thr = P.thr;
    D.sv_total = s.sv_total;
    D.tie_rel = P.tie_rel;
    D.titems = reinterpret_cast<const TileItem *>(vb + V.o_titems);
    D.tau_sector = reinterpret_cast<const int32_t *>(fx + F.o_tau_sector);
    D.cblk_of = reinterpret_cast<const int32_t *>(fx + F.o_cblk_of);
    D.rowoff = rowoff;
    D.coloff = coloff;
    D.row_info = reinterpret_cast<const int2 *>(fx + F.o_row_info);
    D.col_info = reinterpret_cast<const int2 *>(fx + F.o_col_info);
    D.n_lblk = s.n_lblk;
    D.n_rblk = s.n_rblk;
    D.ts_min = 0;
    D.pad_ts = 0;
    D.drop_rel2 = P.drop_rel * P.drop_rel;
    reinterpret_cast<UpdDesc *>(G.staging + G.o_desc)[me] = D;
  }
  for (int sid = blockIdx.y; sid < s.n_sec; sid += gridDim.y) {
    const Sector sc = sectors[sid];
    const int4 x = secx[sid];
    const int n = sc.n, L = sc.L, R = sc.R;
    for (int a = 0; a < s.n_lblk; ++a) {
      const int ro = rowoff[sid * s.n_lblk + a];
      if (ro < 0) continue;
      const Block b = lblk[a];
      for (int i = tid; i < b.end - b.start; i += blockDim.x) rows[sc.row_off + ro + i] = b.start + i;
    }
    for (int c = 0; c < s.n_rblk; ++c) {
      const int co = coloff[sid * s.n_rblk + c];
      if (co < 0) continue;
      const Block b = rblk[c];
      for (int i = tid; i < b.end - b.start; i += blockDim.x) cols[sc.col_off + co + i] = b.start + i;
    }
    const bool tri = n > P.thr;
    const int nb = cdiv(n, kGramBlock);
    const int ng = tri ? nb * (nb + 1) / 2 : nb * nb;
    for (int k = tid; k < ng; k += blockDim.x) {
      int br, bc;
      if (tri) {  // lower triangle, row by row: row br holds br + 1 blocks
        br = (int)((sqrt(8.0 * k + 1.0) - 1.0) * 0.5);
        while (br * (br + 1) / 2 > k) --br;
        while ((br + 1) * (br + 2) / 2 <= k) ++br;
        bc = k - br * (br + 1) / 2;
      } else {
        br = k / nb;
        bc = k % nb;
      }
      const int r0 = br * kGramBlock, c0 = bc * kGramBlock;
      gram[x.x + k] = Item{sid, 0, r0, min(kGramBlock, n - r0), c0, min(kGramBlock, n - c0), 0, 0};
      l_gram[s.l_gram + x.x + k] = make_int2(me, x.x + k);
    }
    const int nwc = cdiv(R, kGramBlock);
    for (int k = tid; k < nb * nwc; k += blockDim.x) {
      const int r0 = (k / nwc) * kGramBlock, c0 = (k % nwc) * kGramBlock;
      witems[x.y + k] = Item{sid, 0, r0, min(kGramBlock, n - r0), c0, min(kGramBlock, R - c0), 0, 0};
      l_w[s.l_w + x.y + k] = make_int2(me, x.y + k);
    }
    if (!sc.left_small) {
      const int nvr = cdiv(L, kGramBlock);
      for (int k = tid; k < nvr * nb; k += blockDim.x) {
        const int r0 = (k / nb) * kGramBlock, c0 = (k % nb) * kGramBlock;
        vitems[x.z + k] = Item{sid, 0, r0, min(kGramBlock, L - r0), c0, min(kGramBlock, n - c0), 0, 0};
        l_v[s.l_v + x.z + k] = make_int2(me, x.z + k);
      }
    }
    for (int i = tid; i < n; i += blockDim.x) sv_sector[sc.s_off + i] = sid;
    if (tid == 0) {
      l_sec[s.l_sec + sid] = make_int2(me, sid);
      if (tri) {
        G.tri_tmp[s.l_tri + x.w] = make_int2(me, sid);
      } else {
        int c = 0;
        while (n > kJacLimit[c]) ++c;
        l_jac[s.l_jac[c] + x.w] = make_int2(me, sid);
      }
    }
  }
}

// ------------------------------------------------------------ k_dp_tiles --
// One block per update: the contraction tile items, class by class in tile
// order (a block scan over the tiles, 256 at a time), and their list entries.
__device__ inline uint64_t block_excl_scan64(uint64_t v, uint64_t *wsum, uint64_t &total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  uint64_t before = 0, all = 0;
  const int nw = (int)(blockDim.x >> 6);
  for (int w = 0; w < nw; ++w) {
    if (w < wave) before += wsum[w];
    all += wsum[w];
  }
  __syncthreads();  // wsum is reused by the next call
  total = all;
  return before + x - v;
}

__global__ __launch_bounds__(256) void k_dp_tiles(DpParams P, const DpUpdate *__restrict__ Ud,
                                                  const DpGroup *__restrict__ Gs,
                                                  const DpSizes *__restrict__ Sz) {
  const int set = dp_set(P), u = blockIdx.x, tid = threadIdx.x;
  if (u >= P.n_upd[set]) return;
  const DpUpdate up = Ud[P.off[set] + u];
  const DpGroup &G = Gs[up.group];
  if (reinterpret_cast<const GroupCounts *>(G.staging)->status != DP_OK) return;
  const DpSizes s = Sz[u];
  if (!s.active) return;
  const int d = P.d, chi = P.chi;
  const FixedLayout F = fixed_layout(d, chi);
  const char *fx = fixed_of(P, G, up.local);
  __shared__ int4 s_lb[kKeys], s_rb[kKeys];
  __shared__ int16_t s_rowblk[kMaxChi], s_colblk[kMaxChi];
  __shared__ uint64_t wsum[4];
  const Block *lblk = reinterpret_cast<const Block *>(fx + F.o_lblk);
  const Block *rblk = reinterpret_cast<const Block *>(fx + F.o_rblk);
  const int2 *row_info = reinterpret_cast<const int2 *>(fx + F.o_row_info);
  const int2 *col_info = reinterpret_cast<const int2 *>(fx + F.o_col_info);
  for (int a = tid; a < s.n_lblk; a += blockDim.x)
    s_lb[a] = make_int4(lblk[a].c0, lblk[a].c1, lblk[a].start, lblk[a].end);
  for (int c = tid; c < s.n_rblk; c += blockDim.x)
    s_rb[c] = make_int4(rblk[c].c0, rblk[c].c1, rblk[c].start, rblk[c].end);
  for (int i = tid; i < chi; i += blockDim.x) {
    s_rowblk[i] = (int16_t)(row_info[i].x < 0 ? -1 : (row_info[i].x & 0xffff));
    s_colblk[i] = (int16_t)(col_info[i].x < 0 ? -1 : (col_info[i].x & 0xffff));
  }
  __syncthreads();
  TileCtx tc;
  tc.lb = s_lb;
  tc.rb = s_rb;
  tc.rowblk = s_rowblk;
  tc.colblk = s_colblk;
  tc.nl = s.n_lblk;
  tc.nr = s.n_rblk;
  tc.nLv = s.n_lblk > 0 ? s_lb[s.n_lblk - 1].w : 0;
  tc.nRv = s.n_rblk > 0 ? s_rb[s.n_rblk - 1].w : 0;
  tc.order = P.tile_order;
  const int ntiles = cdiv(tc.nLv, 16) * cdiv(tc.nRv, 16);
  const VarLayout V = var_layout(s.rows_total, s.cols_total, s.n_gram, s.n_w, s.n_v, s.sv_total,
                                 s.n_titems);
  TileItem *titems = reinterpret_cast<TileItem *>(G.var + s.var_off + V.o_titems);
  int2 *list = reinterpret_cast<int2 *>(G.staging + G.o_tiles);
  int base[kTileClasses];  // first item of each class inside the update
  base[0] = 0;
  for (int c = 1; c < kTileClasses; ++c) base[c] = base[c - 1] + s.tile_cnt[c - 1];
  int run[kTileClasses] = {};  // items of each class in the tiles before this chunk
  for (int chunk = 0; chunk < ntiles; chunk += blockDim.x) {
    const int idx = chunk + tid;
    uint64_t mine = 0;  // 16-bit count per class
    if (idx < ntiles)
      tile_passes(tc, idx, [&](const TileItem &, int c) { mine += (uint64_t)1 << (16 * c); });
    uint64_t total;
    const uint64_t before = block_excl_scan64(mine, wsum, total);
    if (idx < ntiles) {
      int k[kTileClasses];
      for (int c = 0; c < kTileClasses; ++c) k[c] = run[c] + (int)((before >> (16 * c)) & 0xffff);
      tile_passes(tc, idx, [&](const TileItem &it, int c) {
        const int pos = base[c] + k[c];
        titems[pos] = it;
        list[s.l_tile[c] + k[c]] = make_int2(s.desc, pos);
        ++k[c];
      });
    }
    for (int c = 0; c < kTileClasses; ++c) run[c] += (int)((total >> (16 * c)) & 0xffff);
  }
}

// -------------------------------------------------------------- k_dp_tri --
// One block per group: the tridiagonal-route list sorted largest first (a
// counting sort on n; the order among equal sizes is free, each sector's
// result does not depend on it), the split between the two Householder block
// shapes (as launch_tri_values splits a host-planned group), then the
// back-transformation items in list order.
__global__ __launch_bounds__(1024) void k_dp_tri(DpParams P, DpGroup *__restrict__ Gs) {
  DpGroup &G = Gs[blockIdx.x];
  GroupCounts &K = *reinterpret_cast<GroupCounts *>(G.staging);
  if (K.status != DP_OK) return;
  const int tid = threadIdx.x, NT = blockDim.x;
  __shared__ int hist[kMaxChi + 1];
  __shared__ uint64_t wsum[16];
  __shared__ int s_over;
  const int ntri = K.seg[SG_TRI].y, chi = P.chi;
  int2 *out = reinterpret_cast<int2 *>(G.staging + G.o_tri);
  const int2 *in = G.tri_tmp;
  const UpdDesc *desc = reinterpret_cast<const UpdDesc *>(G.staging + G.o_desc);
  auto size_of = [&](int2 r) { return desc[r.x].sectors[r.y].n; };
  for (int i = tid; i <= chi; i += NT) hist[i] = 0;
  if (tid == 0) s_over = 0;
  __syncthreads();
  for (int i = tid; i < ntri; i += NT) atomicAdd(&hist[size_of(in[i])], 1);
  __syncthreads();
  if (tid == 0) {  // hist[n] becomes the first position of size n (sizes descending)
    int acc = 0;
    for (int n = chi; n >= 0; --n) {
      const int c = hist[n];
      hist[n] = acc;
      acc += c;
    }
    // sectors with n > split come first; with few of them all take the
    // larger block shape, as in a host-planned group (launch_tri_values)
    const int split = min(P.tri_split, chi);
    const int nbig = ntri < kTriFewSectors ? ntri : hist[split];
    K.seg[SG_TRIBIG] = make_int2(0, nbig);
    K.seg[SG_TRISMALL] = make_int2(nbig, ntri - nbig);
  }
  __syncthreads();
  for (int i = tid; i < ntri; i += NT) {
    const int2 r = in[i];
    out[atomicAdd(&hist[size_of(r)], 1)] = r;
  }
  __syncthreads();
  // back-transformation items: WY blocks of reflectors (n - 1 of them, unit
  // entry at row offset 1) and panels of kept vectors, per sector in list order
  uint64_t tot_t = 0, tot_a = 0;
  for (int chunk = 0; chunk < ntri; chunk += NT) {
    const int i = chunk + tid;
    uint64_t mine = 0;
    if (i < ntri) {
      const int n = size_of(out[i]);
      mine = (uint64_t)cdiv(n - 1, kWyBlock) | (uint64_t)cdiv(n, kWyBlock) << 32;
    }
    uint64_t total;
    block_excl_scan64(mine, wsum, total);
    tot_t += total & 0xffffffffu;
    tot_a += total >> 32;
  }
  G.needs.wyt = (int32_t)min<uint64_t>(tot_t, INT32_MAX);
  G.needs.wya = (int32_t)min<uint64_t>(tot_a, INT32_MAX);
  if (tot_t > (uint64_t)G.caps.wyt || tot_a > (uint64_t)G.caps.wya) {
    if (tid == 0) {
      K.status = DP_OVER_CAPACITY;
      G.needs.status = DP_OVER_CAPACITY;
    }
    return;
  }
  int2 *wyt = reinterpret_cast<int2 *>(G.staging + G.o_wyt);
  int2 *wya = reinterpret_cast<int2 *>(G.staging + G.o_wya);
  uint64_t run_t = 0, run_a = 0;
  for (int chunk = 0; chunk < ntri; chunk += NT) {
    const int i = chunk + tid;
    uint64_t mine = 0;
    int n = 0;
    int2 r = make_int2(0, 0);
    if (i < ntri) {
      r = out[i];
      n = size_of(r);
      mine = (uint64_t)cdiv(n - 1, kWyBlock) | (uint64_t)cdiv(n, kWyBlock) << 32;
    }
    uint64_t total;
    const uint64_t before = block_excl_scan64(mine, wsum, total);
    if (i < ntri) {
      const int64_t a = (int64_t)run_t + (int64_t)(before & 0xffffffffu);
      const int64_t b = (int64_t)run_a + (int64_t)(before >> 32);
      for (int m = 0; m < cdiv(n - 1, kWyBlock); ++m) wyt[a + m] = make_int2(r.x, r.y | m << 20);
      for (int p = 0; p < cdiv(n, kWyBlock); ++p) wya[b + p] = make_int2(r.x, r.y | p << 20);
    }
    run_t += total & 0xffffffffu;
    run_a += total >> 32;
  }
  if (tid == 0) {
    K.seg[SG_WYT] = make_int2(0, (int)tot_t);
    K.seg[SG_WYA] = make_int2(0, (int)tot_a);
  }
}

// ------------------------------------------------------------- k_dp_gate --
// After every group is planned: a failure in any group, a missing BS table or
// an earlier failed replay zeroes every group's counts (the stages then do
// nothing and the MPO is left as it was) and is recorded once in the sticky
// state with the replay it happened in.  Then the replay counter advances.
__global__ void k_dp_gate(DpParams P, DpGroup *__restrict__ Gs) {
  __shared__ int s_st;
  if (threadIdx.x == 0) {
    int st = P.state[0];
    if (st == DP_OK && P.state[3] != DP_OK) st = P.state[3];
    for (int g = 0; g < P.n_groups && st == DP_OK; ++g)
      st = reinterpret_cast<const GroupCounts *>(Gs[g].staging)->status;
    if (st != DP_OK && P.state[0] == DP_OK) {
      P.state[0] = st;
      P.state[2] = P.state[1];
    }
    s_st = st;
  }
  __syncthreads();
  if (s_st != DP_OK)
    for (int g = threadIdx.x; g < P.n_groups; g += blockDim.x) {
      GroupCounts &K = *reinterpret_cast<GroupCounts *>(Gs[g].staging);
      for (int k = 0; k < SG_COUNT; ++k) K.seg[k].y = 0;
    }
  __syncthreads();
  if (threadIdx.x == 0) P.state[1] += 1;
}

__global__ void k_dp_seek(int32_t *state, int replay) {
  state[0] = DP_OK;
  state[1] = replay;
  state[2] = 0;
  state[3] = DP_OK;
}

}  // namespace

int launch_layer_seek(int32_t *d_state, int replay, hipStream_t stream) {
  hipLaunchKernelGGL(k_dp_seek, dim3(1), dim3(1), 0, stream, d_state, replay);
  BSMPO_HIP(hipGetLastError());
  return 0;
}

int launch_device_plan(const DpParams &p, const DpUpdate *d_upd, DpGroup *d_groups,
                       DpSizes *d_sizes, hipStream_t stream) {
  if (p.chi > kMaxChi) return fail(-1, "device planning: chi above 2048");
  if (p.d < 1 || p.d > kMaxD) return fail(-1, "device planning: d must be in [1, 16]");
  hipLaunchKernelGGL(k_dp_sizes, dim3(p.max_upd), dim3(kSizeThreads), 0, stream, p, d_upd,
                     d_groups, d_sizes);
  BSMPO_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dp_scan, dim3(p.n_groups), dim3(64), 0, stream, p, d_groups, d_sizes);
  BSMPO_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dp_fill, dim3(p.max_upd, kFillStripes), dim3(256), 0, stream, p, d_upd,
                     d_groups, d_sizes);
  BSMPO_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dp_tiles, dim3(p.max_upd), dim3(256), 0, stream, p, d_upd, d_groups, d_sizes);
  BSMPO_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dp_tri, dim3(p.n_groups), dim3(1024), 0, stream, p, d_groups);
  BSMPO_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_dp_gate, dim3(1), dim3(256), 0, stream, p, d_groups);
  BSMPO_HIP(hipGetLastError());
  return 0;
}

}  // namespace bsmpo
