// Host-side planning of one site update: charge blocks, tau sectors, packed
// buffer offsets and the work lists of every stage.
//
// Follows the selection rules of the reference:
//   tau loop bounds   src/beam_splitter_parallel/Rank.py:78-91
//   sector windows    src/mpo_sort.py:424-432 (rows cl >= tau, cols cr <= tau)
//   charge blocks     src/mpo_sort.py:215-259 (start index of each charge)
// but keeps the sectors compact (no padding to multiples of 8, which the
// reference needs for its FFMA thread tiles, src/mpo_sort.py:263-384): the
// MFMA kernels mask ragged tiles instead.
#include <algorithm>
#include <cstring>
#include <atomic>
#include <mutex>
#include <vector>

#include "internal.h"

namespace bsmpo {

static thread_local std::string g_last_error;
// process-wide knobs, set and read from any thread
static std::atomic<int> g_eig_threshold{96};
static std::atomic<int> g_tri_split{256};
static std::atomic<double> g_tie_rel{1e-6};
static std::atomic<double> g_drop_rel{1e-6};
static std::atomic<int> g_two_stage{0};
static std::atomic<int> g_tile_order{8};
int eig_threshold() { return g_eig_threshold; }
int two_stage_min() { return g_two_stage; }
int tri_split() { return g_tri_split; }
int tile_order() { return g_tile_order; }
double tie_rel() { return g_tie_rel; }
double drop_rel() { return g_drop_rel; }

void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

namespace {

struct Built {
  PlanHeader h;
  std::vector<Sector> sectors;
  std::vector<int32_t> rows, cols;
  std::vector<Block> lblk, cblk, rblk;
  std::vector<Item> gram_items, w_items, v_items;
  std::vector<int32_t> charges;  // CL, CC, CR: 3 x chi x 2
  std::vector<int32_t> sv_sector;  // sector of every singular-value slot
  std::vector<int32_t> jac_order;  // sector indices grouped by Jacobi class
  std::vector<TileItem> titems;    // contraction tiles, by LDS class
  std::vector<int32_t> tau_sector, cblk_of, rowoff, coloff, row_blk, col_blk;
  std::vector<SlotInfo> row_info, col_info;  // device copies of the slot data
};

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

int make_blocks(int d, int chi, const int32_t *ch, std::vector<Block> &out, const char *name) {
  out.clear();
  int64_t prev = -1;
  for (int i = 0; i < chi; ++i) {
    int32_t c0 = ch[2 * i], c1 = ch[2 * i + 1];
    if (c0 < 0 || c1 < 0 || c0 > d || c1 > d)
      return fail(-1, std::string("charge out of range in ") + name);
    int64_t key = (int64_t)c0 * (d + 1) + c1;
    if (key < prev) return fail(-1, std::string("charges not lexicographically sorted in ") + name);
    prev = key;
    if (c0 >= d || c1 >= d) continue;
    if (!out.empty() && out.back().c0 == c0 && out.back().c1 == c1 && out.back().end == i)
      out.back().end = i + 1;
    else
      out.push_back(Block{c0, c1, i, i + 1});
  }
  return 0;
}

void add_tiles(std::vector<Item> &items, int sector, int blk, int r_begin, int r_end, int c_begin,
               int c_end, int t = 16) {
  for (int r = r_begin; r < r_end; r += t)
    for (int c = c_begin; c < c_end; c += t)
      items.push_back(Item{sector, blk, r, std::min(t, r_end - r), c, std::min(t, c_end - c), 0, 0});
}

int build(int d, int chi, const int32_t *CL, const int32_t *CC, const int32_t *CR, int only_t0,
          int only_t1, Built &b) {
  if (d < 1 || d > kMaxD)
    return fail(-1, "d must be in [1, 16] (the sector contraction keeps one BS box side per "
                    "register array; the reference's kernel has the same bound)");
  if (chi < 1) return fail(-1, "chi must be positive");
  std::memset(&b.h, 0, sizeof(PlanHeader));
  int rc;
  if ((rc = make_blocks(d, chi, CL, b.lblk, "CL"))) return rc;
  if ((rc = make_blocks(d, chi, CC, b.cblk, "CC"))) return rc;
  if ((rc = make_blocks(d, chi, CR, b.rblk, "CR"))) return rc;

  // tau bounds, Rank.py:78-91 (largest valid CL per component, smallest CR)
  int lc0 = 0, lc1 = 0, sr0 = d, sr1 = d;
  bool any0 = false, any1 = false;
  for (int i = 0; i < chi; ++i) {
    if (CL[2 * i] != d) { lc0 = any0 ? std::max(lc0, CL[2 * i]) : CL[2 * i]; any0 = true; }
    if (CL[2 * i + 1] != d) { lc1 = any1 ? std::max(lc1, CL[2 * i + 1]) : CL[2 * i + 1]; any1 = true; }
    sr0 = std::min(sr0, CR[2 * i]);
    sr1 = std::min(sr1, CR[2 * i + 1]);
  }

  const int thr = eig_threshold();  // sectors with n > thr take the tridiagonal route
  int64_t c_off = 0, g_off = 0, v_off = 0, w_off = 0, s_off = 0;
  double theta_elems = 0.0;
  for (int t0 = sr0; t0 <= lc0; ++t0) {
    for (int t1 = sr1; t1 <= lc1; ++t1) {
      if (only_t0 >= 0 && (t0 != only_t0 || t1 != only_t1)) continue;
      Sector s{};
      s.t0 = t0;
      s.t1 = t1;
      s.row_off = (int32_t)b.rows.size();
      s.col_off = (int32_t)b.cols.size();
      std::vector<std::pair<int, int>> lsel;  // (block index, sector-local row start)
      int L = 0, R = 0;
      for (int k = 0; k < (int)b.lblk.size(); ++k) {
        const Block &bl = b.lblk[k];
        if (bl.c0 >= t0 && bl.c1 >= t1) {
          lsel.push_back({k, L});
          for (int i = bl.start; i < bl.end; ++i) b.rows.push_back(i);
          L += bl.end - bl.start;
        }
      }
      std::vector<int> rsel;
      for (int k = 0; k < (int)b.rblk.size(); ++k) {
        const Block &br = b.rblk[k];
        if (br.c0 <= t0 && br.c1 <= t1) {
          rsel.push_back(k);
          for (int i = br.start; i < br.end; ++i) b.cols.push_back(i);
          R += br.end - br.start;
        }
      }
      if (L == 0 || R == 0) {  // Rank.py:108-111: skipped sector
        b.rows.resize(s.row_off);
        b.cols.resize(s.col_off);
        continue;
      }
      s.L = L;
      s.R = R;
      s.n = std::min(L, R);
      s.left_small = L <= R ? 1 : 0;
      s.c_off = c_off;
      s.g_off = g_off;
      s.v_off = v_off;
      s.w_off = w_off;
      s.s_off = s_off;
      c_off += (int64_t)L * R;
      g_off += (int64_t)s.n * s.n;
      v_off += (int64_t)L * s.n;
      w_off += (int64_t)s.n * R;
      s_off += s.n;
      theta_elems += (double)L * R;
      // sectors on the tridiagonal route read only the lower triangle of G
      const bool lower_only = s.n > thr;
      {
        const double n = s.n, m = std::max(L, R);
        b.h.flops_gram += lower_only ? 4.0 * m * n * (n + 1) : 8.0 * m * n * n;
        b.h.flops_wmat += 8.0 * n * L * (double)R;
        if (!s.left_small) b.h.flops_vectors += 8.0 * L * (double)R * n;
        if (!lower_only) b.h.flops_jacobi += 8.0 * n * n * n;
      }
      int sid = (int)b.sectors.size();
      // Gram blocks of kGramBlock x kGramBlock; on the tridiagonal route only
      // those on or below the diagonal
      for (int r = 0; r < s.n; r += kGramBlock)
        for (int c = 0; c < s.n && (!lower_only || c <= r); c += kGramBlock)
          b.gram_items.push_back(Item{sid, 0, r, std::min(kGramBlock, s.n - r), c,
                                      std::min(kGramBlock, s.n - c), 0, 0});
      add_tiles(b.w_items, sid, 0, 0, s.n, 0, R, kGramBlock);
      if (!s.left_small) add_tiles(b.v_items, sid, 0, 0, L, 0, s.n, kGramBlock);
      b.sectors.push_back(s);
    }
  }

  // algorithmic counts (roofline accounting, complex MAC = 8 flops), from
  // 2-D prefix sums of the centre-charge histogram:
  //   box(cl, cr) = sum_{cr <= ch <= cl} n_C(ch)
  //   F_alg = 8 sum_{cl >= cr} n_L n_R box,  F_tau = F_alg weighted by the
  //   number of sectors tau with cl >= tau >= cr
  std::vector<double> pc((size_t)(d + 1) * (d + 1), 0.0);  // pc[(a+1)(d+1)+(b+1)] = sum_{c0<=a,c1<=b}
  for (const Block &bc : b.cblk) pc[(size_t)(bc.c0 + 1) * (d + 1) + bc.c1 + 1] += bc.end - bc.start;
  for (int a = 1; a <= d; ++a)
    for (int c = 1; c <= d; ++c)
      pc[(size_t)a * (d + 1) + c] += pc[(size_t)(a - 1) * (d + 1) + c] +
                                     pc[(size_t)a * (d + 1) + c - 1] -
                                     pc[(size_t)(a - 1) * (d + 1) + c - 1];
  auto rect = [&](int a0, int b0, int a1, int b1) {  // sum over a0..a1, b0..b1 (inclusive)
    if (a1 < a0 || b1 < b0) return 0.0;
    auto P = [&](int a, int c) { return pc[(size_t)(a + 1) * (d + 1) + c + 1]; };
    double r = P(a1, b1);
    if (a0 > 0) r -= P(a0 - 1, b1);
    if (b0 > 0) r -= P(a1, b0 - 1);
    if (a0 > 0 && b0 > 0) r += P(a0 - 1, b0 - 1);
    return r;
  };
  double f_alg = 0.0, f_tau = 0.0, lc = 0.0, cr = 0.0;
  for (const Block &bl : b.lblk) {
    const double nl = bl.end - bl.start;
    lc += nl * rect(0, 0, bl.c0, bl.c1);
    for (const Block &br : b.rblk) {
      if (!(bl.c0 >= br.c0 && bl.c1 >= br.c1)) continue;
      const double box = rect(br.c0, br.c1, bl.c0, bl.c1);
      const double f = 8.0 * nl * (br.end - br.start) * box;
      f_alg += f;
      f_tau += f * (double)(bl.c0 - br.c0 + 1) * (bl.c1 - br.c1 + 1);
    }
  }
  for (const Block &br : b.rblk) cr += (br.end - br.start) * rect(br.c0, br.c1, d - 1, d - 1);
  (void)theta_elems;

  for (int sid = 0; sid < (int)b.sectors.size(); ++sid)
    for (int i = 0; i < b.sectors[sid].n; ++i) b.sv_sector.push_back(sid);
  b.h.eig_threshold = thr;
  b.h.tie_rel = tie_rel();
  b.h.drop_rel = drop_rel();
  const int tsm = two_stage_min();
  b.h.ts_min = tsm;
  int64_t tw = 0;
  for (int c = 0; c < kJacClasses; ++c) {
    b.h.jac_off[c] = (int32_t)b.jac_order.size();
    b.h.jac_max_n[c] = 0;
    int lo = c == 0 ? 0 : kJacLimit[c - 1];
    for (int sid = 0; sid < (int)b.sectors.size(); ++sid) {
      Sector &sc = b.sectors[sid];
      const int n = sc.n;
      const bool tri = n > thr;
      const bool in = c == kTriClass ? tri : (!tri && n > lo && n <= kJacLimit[c]);
      if (in) {
        b.jac_order.push_back(sid);
        b.h.jac_max_n[c] = std::max(b.h.jac_max_n[c], n);
        if (c == kTriClass && two_stage(n, tsm)) {
          // band reduction: the lower triangle read and the full matrix
          // written once (scaling, mirror); per panel of kBand columns the
          // full trailing matrix read for Y = A22 V and read + written by the
          // two-sided update (MFMA, 8 flops per complex MAC); the band copy.
          // Bulge chasing: ~n^2 / (2 kBand) steps of ~4 kBand^2 complex MACs.
          const double nn = n;
          b.h.bytes_tridiag += 8.0 * nn * (nn + 1) + 16.0 * nn * nn + 16.0 * (kBand + 1 + 2 * kBand) * nn;
          for (int j0 = 0; n - j0 - kBand >= 2; j0 += kBand) {
            const double m = n - j0 - kBand;
            b.h.bytes_tridiag += 48.0 * m * m + 2.0 * 16.0 * kBand * m;
            b.h.flops_tridiag += 8.0 * (3.0 * kBand * m * m + 4.0 * kBand * kBand * m);
          }
          b.h.flops_tridiag += 8.0 * 4.0 * kBand * kBand * nn * nn / (2.0 * kBand);
          b.h.n_bc += (n + bc_vectors(n) - 1) / bc_vectors(n);
        } else if (c == kTriClass) {
          const double nn = n;
          b.h.flops_tridiag += 16.0 / 3.0 * nn * nn * nn;
          // lower triangle of the trailing matrix (m (m + 1) / 2 complex entries)
          // read once per column (A22 v), read + written once per panel
          for (int j = 0; j + 1 < n; ++j) {
            const double m = n - 1 - j;
            b.h.bytes_tridiag += 8.0 * m * (m + 1);
            if (j % kTriPanel == kTriPanel - 1 || j + 2 == n) b.h.bytes_tridiag += 16.0 * m * (m + 1);
          }
        }
        if (c == kTriClass) {
          sc.tw_off = tw;
          // per-lane inverse-iteration arrays (6 n doubles for each of the
          // min(n, kTriVecThreads) lanes of the sector's block); the Householder
          // panels V, W (2 n kTriPanel complex) use the same area first
          // then the back-transformation's T blocks (kWyBlock^2 complex per
          // block of reflectors), once the inverse iteration is done
          const int nblk = (n - 1 + kWyBlock - 1) / kWyBlock;
          b.h.n_wyt += nblk;
          b.h.n_wya += (n + kWyBlock - 1) / kWyBlock;
          const int64_t wy = 2 * (int64_t)kWyBlock * kWyBlock * nblk;
          // two-stage: the band (2 kBand n complex) and the panel's V, Y, W
          // (kBand n complex each) before the inverse iteration
          const int64_t band = two_stage(n, tsm) ? 2 * (int64_t)(2 * kBand + 3 * kBand) * n : 0;
          tw += std::max<int64_t>(
              std::max<int64_t>((int64_t)n * std::max(6 * std::min(n, kTriVecThreads), 4 * kTriPanel), wy),
              band);
        }
      }
    }
    if (c == kTriClass)  // largest first: the longest column chains start first
      std::stable_sort(b.jac_order.begin() + b.h.jac_off[c], b.jac_order.end(),
                       [&](int32_t x, int32_t y) { return b.sectors[x].n > b.sectors[y].n; });
  }
  b.h.jac_off[kJacClasses] = (int32_t)b.jac_order.size();

  // contraction tables: tau -> sector, centre charge -> block, block offsets in sectors
  {
    const int D1 = d + 1;
    const int ns = (int)b.sectors.size(), nl = (int)b.lblk.size(), nr = (int)b.rblk.size();
    b.tau_sector.assign((size_t)D1 * D1, -1);
    b.cblk_of.assign((size_t)D1 * D1, -1);
    for (int s = 0; s < ns; ++s) b.tau_sector[(size_t)b.sectors[s].t0 * D1 + b.sectors[s].t1] = s;
    for (int c = 0; c < (int)b.cblk.size(); ++c)
      b.cblk_of[(size_t)b.cblk[c].c0 * D1 + b.cblk[c].c1] = c;
    b.rowoff.assign((size_t)ns * nl, -1);
    b.coloff.assign((size_t)ns * nr, -1);
    for (int s = 0; s < ns; ++s) {
      const Sector &x = b.sectors[s];
      int o = 0;
      for (int a = 0; a < nl; ++a)
        if (b.lblk[a].c0 >= x.t0 && b.lblk[a].c1 >= x.t1) {
          b.rowoff[(size_t)s * nl + a] = o;
          o += b.lblk[a].end - b.lblk[a].start;
        }
      o = 0;
      for (int c = 0; c < nr; ++c)
        if (b.rblk[c].c0 <= x.t0 && b.rblk[c].c1 <= x.t1) {
          b.coloff[(size_t)s * nr + c] = o;
          o += b.rblk[c].end - b.rblk[c].start;
        }
    }
    for (int a = 0; a < nl; ++a) {
      const Block &la = b.lblk[a];
      for (int c = 0; c < nr; ++c) {
        const Block &rb = b.rblk[c];
        if (!(la.c0 >= rb.c0 && la.c1 >= rb.c1)) continue;
        const int B0 = la.c0 - rb.c0 + 1, B1 = la.c1 - rb.c1 + 1;
        b.h.flops_weight += 8.0 * (la.end - la.start) * (rb.end - rb.start) *
                            (double)(B0 * B1) * (B0 + B1);
      }
    }
    // contraction tiles: 16 x 16 (left row, right column) tiles, visited in
    // super-tiles of tile_order() x tile_order() tiles so that neighbouring
    // items share operand panels
    b.row_blk.assign((size_t)chi, -1);
    b.col_blk.assign((size_t)chi, -1);
    int nLv = 0, nRv = 0;
    for (int a = 0; a < nl; ++a) {
      for (int i = b.lblk[a].start; i < b.lblk[a].end; ++i) b.row_blk[i] = a;
      nLv = std::max(nLv, b.lblk[a].end);
    }
    for (int c = 0; c < nr; ++c) {
      for (int i = b.rblk[c].start; i < b.rblk[c].end; ++i) b.col_blk[i] = c;
      nRv = std::max(nRv, b.rblk[c].end);
    }
    b.row_info.assign((size_t)chi, SlotInfo{-1, 0});
    b.col_info.assign((size_t)chi, SlotInfo{-1, 0});
    for (int a = 0; a < nl; ++a)
      for (int i = b.lblk[a].start; i < b.lblk[a].end; ++i)
        b.row_info[i] = SlotInfo{a | b.lblk[a].c0 << 16 | b.lblk[a].c1 << 24, i - b.lblk[a].start};
    for (int c = 0; c < nr; ++c)
      for (int i = b.rblk[c].start; i < b.rblk[c].end; ++i)
        b.col_info[i] = SlotInfo{c | b.rblk[c].c0 << 16 | b.rblk[c].c1 << 24, i - b.rblk[c].start};
    std::vector<TileItem> cls[kTileClasses];
    const int kLast = kTileElems[kTileClasses - 1];
    struct Pass {
      TileItem it;
      int need;
      double k4, k32;
      bool any;
    };
    // one pass over rows [r0 + s0, r0 + s0 + rp) of the tile (r0, c0)
    auto make_pass = [&](int r0, int c0, int tnr, int tnc, int s0, int rp, int lo0, int lo1,
                         int c_lo, int c_hi) {
      Pass ps{};
      int hi0 = -1, hi1 = -1, a_lo = nl, a_hi = -1;
      for (int i = r0 + s0; i < r0 + std::min(tnr, s0 + rp); ++i) {
        const int a = b.row_blk[i];
        if (a < 0) continue;
        hi0 = std::max(hi0, b.lblk[a].c0);
        hi1 = std::max(hi1, b.lblk[a].c1);
        a_lo = std::min(a_lo, a);
        a_hi = std::max(a_hi, a);
      }
      for (int a = a_lo; a <= a_hi && !ps.any; ++a)
        for (int c = c_lo; c <= c_hi && !ps.any; ++c) {
          const Block &la = b.lblk[a], &rb = b.rblk[c];
          if (!(la.c0 >= rb.c0 && la.c1 >= rb.c1)) continue;
          if (only_t0 >= 0 && !(la.c0 >= only_t0 && only_t0 >= rb.c0 && la.c1 >= only_t1 &&
                                only_t1 >= rb.c1))
            continue;
          ps.any = true;
        }
      if (!ps.any) return ps;
      for (int q0 = lo0; q0 <= hi0; ++q0)
        for (int q1 = lo1; q1 <= hi1; ++q1) {
          const int cb = b.cblk_of[(size_t)q0 * D1 + q1];
          if (cb < 0) continue;
          const int n = b.cblk[cb].end - b.cblk[cb].start;
          ps.k4 += (n + 3) / 4 * 4;
          ps.k32 += (n + 31) / 32 * 32;
        }
      const int nbox = (hi0 - lo0 + 1) * (hi1 - lo1 + 1);
      const int na = a_hi - a_lo + 1, nbk = c_hi - c_lo + 1;
      ps.need = nbox * rp * 16 + (nbox * (12 + 4 * (na + nbk)) + 7) / 8;
      ps.it = TileItem{r0, c0, (int8_t)tnr, (int8_t)tnc, (int8_t)s0, (int8_t)rp,
                       (int8_t)lo0, (int8_t)lo1, (int8_t)hi0, (int8_t)hi1,
                       (int16_t)a_lo, (int16_t)c_lo, (int8_t)na, (int8_t)nbk, 0, 0};
      return ps;
    };
    std::vector<Pass> passes;
    const int sup = 16 * tile_order();  // super-tile side in elements
    for (int R0 = 0; R0 < nLv; R0 += sup)
      for (int C0 = 0; C0 < nRv; C0 += sup)
        for (int r0 = R0; r0 < std::min(R0 + sup, nLv); r0 += 16)
          for (int c0 = C0; c0 < std::min(C0 + sup, nRv); c0 += 16) {
            const int tnr = std::min(16, nLv - r0), tnc = std::min(16, nRv - c0);
            int lo0 = d, lo1 = d, c_lo = nr, c_hi = -1;
            for (int j = c0; j < c0 + tnc; ++j) {
              const int c = b.col_blk[j];
              if (c < 0) continue;
              lo0 = std::min(lo0, b.rblk[c].c0);
              lo1 = std::min(lo1, b.rblk[c].c1);
              c_lo = std::min(c_lo, c);
              c_hi = std::max(c_hi, c);
            }
            if (c_hi < 0) continue;
            for (int rp = 16; rp >= 1; rp /= 2) {
              passes.clear();
              bool fits = true;
              for (int s0 = 0; s0 < tnr && fits; s0 += rp) {
                Pass ps = make_pass(r0, c0, tnr, tnc, s0, rp, lo0, lo1, c_lo, c_hi);
                if (!ps.any) continue;
                if (ps.need > kLast) fits = false;
                passes.push_back(ps);
              }
              if (fits) break;
              if (rp == 1) return fail(-1, "contraction tile does not fit the LDS classes");
            }
            for (const Pass &ps : passes) {
              int c = 0;
              while (ps.need > kTileElems[c]) ++c;
              cls[c].push_back(ps.it);
              b.h.mfma_k4 += ps.k4;
              b.h.mfma_k32 += ps.k32;
            }
          }
    for (int c = 0; c < kTileClasses; ++c) {
      b.h.tile_off[c] = (int32_t)b.titems.size();
      b.titems.insert(b.titems.end(), cls[c].begin(), cls[c].end());
    }
    b.h.tile_off[kTileClasses] = (int32_t)b.titems.size();
  }

  b.charges.resize((size_t)6 * chi);
  std::memcpy(b.charges.data(), CL, sizeof(int32_t) * 2 * chi);
  std::memcpy(b.charges.data() + 2 * chi, CC, sizeof(int32_t) * 2 * chi);
  std::memcpy(b.charges.data() + 4 * chi, CR, sizeof(int32_t) * 2 * chi);

  PlanHeader &h = b.h;
  h.magic = kPlanMagic;
  h.d = d;
  h.chi = chi;
  h.n_sectors = (int32_t)b.sectors.size();
  h.n_items = (int32_t)b.titems.size();
  h.n_titems = h.n_items;
  h.n_lblk = (int32_t)b.lblk.size();
  h.n_cblk = (int32_t)b.cblk.size();
  h.n_rblk = (int32_t)b.rblk.size();
  h.n_gram_items = (int32_t)b.gram_items.size();
  h.n_w_items = (int32_t)b.w_items.size();
  h.n_v_items = (int32_t)b.v_items.size();
  h.c_elems = c_off;
  h.g_elems = g_off;
  h.sv_total = s_off;
  h.v_elems = v_off;
  h.w_elems = w_off;
  h.rows_total = (int64_t)b.rows.size();
  h.cols_total = (int64_t)b.cols.size();
  h.flops_alg = f_alg;
  h.flops_tau = f_tau;
  h.bytes_alg16 = 16.0 * (lc + cr + theta_elems);

  h.dev_begin = align_up(sizeof(PlanHeader), 256);
  int64_t o = 0;
  auto place = [&](int64_t bytes) {
    int64_t at = o;
    o = align_up(o + bytes, 256);
    return at;
  };
  h.o_sectors = place(sizeof(Sector) * b.sectors.size());
  h.o_rows = place(sizeof(int32_t) * b.rows.size());
  h.o_cols = place(sizeof(int32_t) * b.cols.size());
  h.o_lblk = place(sizeof(Block) * b.lblk.size());
  h.o_cblk = place(sizeof(Block) * b.cblk.size());
  h.o_rblk = place(sizeof(Block) * b.rblk.size());
  h.o_gram_items = place(sizeof(Item) * b.gram_items.size());
  h.o_w_items = place(sizeof(Item) * b.w_items.size());
  h.o_v_items = place(sizeof(Item) * b.v_items.size());
  h.o_charges = place(sizeof(int32_t) * b.charges.size());
  h.o_sv_sector = place(sizeof(int32_t) * b.sv_sector.size());
  h.o_jac_order = place(sizeof(int32_t) * b.jac_order.size());
  h.o_titems = place(sizeof(TileItem) * b.titems.size());
  h.o_tau_sector = place(sizeof(int32_t) * b.tau_sector.size());
  h.o_cblk_of = place(sizeof(int32_t) * b.cblk_of.size());
  h.o_rowoff = place(sizeof(int32_t) * b.rowoff.size());
  h.o_coloff = place(sizeof(int32_t) * b.coloff.size());
  h.o_row_info = place(sizeof(SlotInfo) * b.row_info.size());
  h.o_col_info = place(sizeof(SlotInfo) * b.col_info.size());
  h.dev_bytes = o;
  h.blob_bytes = h.dev_begin + h.dev_bytes;

  int64_t w = align_up(h.dev_bytes, 256);
  auto wplace = [&](int64_t bytes) {
    int64_t at = w;
    w = align_up(w + std::max<int64_t>(bytes, 1), 256);
    return at;
  };
  h.w_C = wplace(16 * h.c_elems);
  h.w_T = wplace(16 * h.c_elems);
  h.w_G = wplace(16 * h.g_elems);
  h.w_V = wplace(16 * h.v_elems);
  h.w_W = wplace(16 * h.w_elems);
  h.w_S = wplace(8 * h.sv_total);
  h.w_lam = wplace(8 * h.sv_total);
  h.w_flag = wplace(4 * h.sv_total);
  h.w_slot = wplace(4 * h.sv_total);
  h.w_td = wplace(8 * h.sv_total);
  h.w_te = wplace(8 * h.sv_total);
  h.w_tau = wplace(16 * h.sv_total);
  h.w_kcount = wplace(4 * (int64_t)b.sectors.size());
  h.w_triwork = wplace(8 * tw);
  h.w_hist = wplace(4 * kHistBins + 8);
  h.w_total = w;
  return 0;
}

// The last plan built on this thread, so that query + build cost one build.
struct Cached {
  bool valid = false;
  int d = 0, chi = 0, t0 = 0, t1 = 0, thr = 0, tsm = 0, order = 0;
  double tie = 0.0, drop = 0.0;
  std::vector<int32_t> key;
  Built b;
};
thread_local Cached t_cache;

int build_cached(int d, int chi, const int32_t *CL, const int32_t *CC, const int32_t *CR, int t0,
                 int t1, const Built **out) {
  Cached &c = t_cache;
  const size_t n = (size_t)2 * chi;
  bool hit = c.valid && c.d == d && c.chi == chi && c.t0 == t0 && c.t1 == t1 &&
             c.thr == eig_threshold() && c.tsm == two_stage_min() && c.tie == tie_rel() &&
             c.order == tile_order() &&
             c.drop == drop_rel() &&
             c.key.size() == 3 * n &&
             !std::memcmp(c.key.data(), CL, n * 4) && !std::memcmp(c.key.data() + n, CC, n * 4) &&
             !std::memcmp(c.key.data() + 2 * n, CR, n * 4);
  if (!hit) {
    c.valid = false;
    c.b = Built();
    int rc = build(d, chi, CL, CC, CR, t0, t1, c.b);
    if (rc) return rc;
    c.key.resize(3 * n);
    std::memcpy(c.key.data(), CL, n * 4);
    std::memcpy(c.key.data() + n, CC, n * 4);
    std::memcpy(c.key.data() + 2 * n, CR, n * 4);
    c.d = d; c.chi = chi; c.t0 = t0; c.t1 = t1; c.thr = eig_threshold(); c.tie = tie_rel();
    c.drop = drop_rel();
    c.tsm = two_stage_min();
    c.order = tile_order();
    c.valid = true;
  }
  *out = &c.b;
  return 0;
}

template <class T>
void put(char *base, int64_t off, const std::vector<T> &v) {
  if (!v.empty()) std::memcpy(base + off, v.data(), sizeof(T) * v.size());
}

}  // namespace

int validate_plan(const void *plan, int d, int chi) {
  if (!plan) return fail(-1, "plan is NULL");
  const PlanHeader *h = header(plan);
  if (h->magic != kPlanMagic) return fail(-1, "not a bsmpo plan");
  if (h->d != d || h->chi != chi) return fail(-1, "plan was built for another d / chi");
  return 0;
}

}  // namespace bsmpo

using namespace bsmpo;

extern "C" {

const char *bsmpo_last_error(void) { return bsmpo::g_last_error.c_str(); }

int bsmpo_abi_version(void) { return 1; }

int bsmpo_set_drop_rel(double rel) {
  if (!(rel >= 1e-8 && rel <= 1e-3)) return fail(-1, "drop threshold must be in [1e-8, 1e-3]");
  g_drop_rel = rel;
  return 0;
}

double bsmpo_drop_rel(void) { return g_drop_rel; }

int bsmpo_set_tie_rel(double rel) {
  if (!(rel >= 0.0 && rel < 1e-2)) return fail(-1, "tie band must be in [0, 1e-2)");
  g_tie_rel = rel;
  return 0;
}

int bsmpo_set_eig_threshold(int n) {
  if (n < 0 || n > 96) return fail(-1, "eig threshold must be in [0, 96]");
  g_eig_threshold = n;
  return 0;
}

int bsmpo_two_stage_threshold(void) { return g_two_stage; }

int bsmpo_set_two_stage(int n) {
  if (n < 0 || n > kBandMaxN) return fail(-1, "two-stage threshold must be in [0, 528]");
  g_two_stage = n == 0 ? 0 : std::max(n, kTwoStageMinN);
  return 0;
}

int bsmpo_set_tile_order(int n) {
  if (n != 2 && n != 4 && n != 8 && n != 16) return fail(-1, "tile order must be 2, 4, 8 or 16");
  g_tile_order = n;
  return 0;
}

int bsmpo_tile_order(void) { return g_tile_order; }

int bsmpo_tri_split(void) { return g_tri_split; }

int bsmpo_eig_threshold(void) { return g_eig_threshold; }

int bsmpo_set_tri_split(int n) {
  if (n < 1 || n > 512) return fail(-1, "tridiagonal block-shape split must be in [1, 512]");
  g_tri_split = n;
  return 0;
}

int bsmpo_plan_query(int d, int chi, const int32_t *h_CL, const int32_t *h_CC,
                     const int32_t *h_CR, int tau0, int tau1, size_t *plan_bytes,
                     size_t *work_bytes) {
  if (!h_CL || !h_CC || !h_CR || !plan_bytes || !work_bytes) return fail(-1, "NULL argument");
  const Built *b = nullptr;
  int rc = build_cached(d, chi, h_CL, h_CC, h_CR, tau0, tau1, &b);
  if (rc) return rc;
  *plan_bytes = (size_t)b->h.blob_bytes;
  // workspace of a single-update run (a group of one): header-only query
  const void *plans[1] = {&b->h};
  int64_t hb = 0, wb = 0;
  if ((rc = group_query(1, plans, &hb, &wb))) return rc;
  *work_bytes = (size_t)wb;
  return 0;
}

int bsmpo_plan_build(int d, int chi, const int32_t *h_CL, const int32_t *h_CC,
                     const int32_t *h_CR, int tau0, int tau1, void *h_plan, size_t plan_bytes) {
  if (!h_CL || !h_CC || !h_CR || !h_plan) return fail(-1, "NULL argument");
  const Built *bp = nullptr;
  int rc = build_cached(d, chi, h_CL, h_CC, h_CR, tau0, tau1, &bp);
  if (rc) return rc;
  const Built &b = *bp;
  if ((int64_t)plan_bytes < b.h.blob_bytes) return fail(-1, "plan buffer too small");
  char *base = reinterpret_cast<char *>(h_plan);
  std::memset(base, 0, (size_t)b.h.dev_begin);
  std::memcpy(base, &b.h, sizeof(PlanHeader));
  char *dev = base + b.h.dev_begin;
  put(dev, b.h.o_sectors, b.sectors);
  put(dev, b.h.o_rows, b.rows);
  put(dev, b.h.o_cols, b.cols);
  put(dev, b.h.o_lblk, b.lblk);
  put(dev, b.h.o_cblk, b.cblk);
  put(dev, b.h.o_rblk, b.rblk);
  put(dev, b.h.o_gram_items, b.gram_items);
  put(dev, b.h.o_w_items, b.w_items);
  put(dev, b.h.o_v_items, b.v_items);
  put(dev, b.h.o_charges, b.charges);
  put(dev, b.h.o_sv_sector, b.sv_sector);
  put(dev, b.h.o_jac_order, b.jac_order);
  put(dev, b.h.o_titems, b.titems);
  put(dev, b.h.o_tau_sector, b.tau_sector);
  put(dev, b.h.o_cblk_of, b.cblk_of);
  put(dev, b.h.o_rowoff, b.rowoff);
  put(dev, b.h.o_coloff, b.coloff);
  put(dev, b.h.o_row_info, b.row_info);
  put(dev, b.h.o_col_info, b.col_info);
  return 0;
}

int bsmpo_plan_info_get(const void *h_plan, bsmpo_plan_info *info) {
  if (!h_plan || !info) return fail(-1, "NULL argument");
  const PlanHeader *h = header(h_plan);
  if (h->magic != kPlanMagic) return fail(-1, "not a bsmpo plan");
  info->n_sectors = h->n_sectors;
  info->n_items = h->n_items;
  info->c_elems = h->c_elems;
  info->g_elems = h->g_elems;
  info->sv_total = h->sv_total;
  info->flops_alg = h->flops_alg;
  info->flops_tau = h->flops_tau;
  info->bytes_alg16 = h->bytes_alg16;
  return 0;
}

int bsmpo_plan_sectors(const void *h_plan, int32_t *h_out, int64_t *h_coff, int32_t *h_rows,
                       int32_t *h_cols) {
  if (!h_plan) return fail(-1, "NULL argument");
  const PlanHeader *h = header(h_plan);
  if (h->magic != kPlanMagic) return fail(-1, "not a bsmpo plan");
  const Sector *s = host_array<Sector>(h_plan, h->o_sectors);
  for (int i = 0; i < h->n_sectors; ++i) {
    if (h_out) {
      h_out[4 * i] = s[i].t0;
      h_out[4 * i + 1] = s[i].t1;
      h_out[4 * i + 2] = s[i].L;
      h_out[4 * i + 3] = s[i].R;
    }
    if (h_coff) h_coff[i] = s[i].c_off;
  }
  if (h_rows) std::memcpy(h_rows, host_array<int32_t>(h_plan, h->o_rows), 4 * h->rows_total);
  if (h_cols) std::memcpy(h_cols, host_array<int32_t>(h_plan, h->o_cols), 4 * h->cols_total);
  return 0;
}

}  // extern "C"
