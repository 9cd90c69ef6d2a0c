// Internal declarations shared by the bsmpo translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>

#include "../../include/bsmpo.h"

namespace bsmpo {

// ------------------------------------------------------------- errors --
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
#define BSMPO_HIP(call)                                                          \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return ::bsmpo::fail((int)e_, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// --------------------------------------------------------------- plan --
// A plan is one contiguous host blob: PlanHeader, then the arrays below at
// the byte offsets the header records.  The part from `dev_begin` to the end
// is copied verbatim to the front of the device workspace, so every offset
// used by kernels is relative to that copy.
constexpr int32_t kPlanMagic = 0x4253504c;  // "BSPL"
// Eigensolver size classes of the sector Gram matrices: n <= 16, 32, 64, 96
// run one-sided Jacobi with the matrix in LDS; the last class (n above the
// runtime threshold, default 96) runs the tridiagonal route (Householder,
// bisection, inverse iteration for the kept pairs only).
constexpr int kJacClasses = 5;
constexpr int kJacLimit[kJacClasses] = {16, 32, 64, 96, 1 << 30};
constexpr int kTriClass = kJacClasses - 1;
constexpr int kTriPanel = 8;  // column panel of the blocked Householder reduction
constexpr int kTriVecThreads = 128;  // inverse-iteration threads per sector
constexpr int kWyBlock = 32;  // back-transformation: reflectors per WY block, vectors per panel
constexpr int kGramBlock = 32;  // Gram stage: output block of one workgroup (2 x 2 MFMA tiles)
// Two-stage reduction of the large tridiagonal-route sectors: dense -> band of
// width kBand (panels on the FP64 MFMA) -> tridiagonal (bulge chasing).
constexpr int kBand = 16;
constexpr int kBandThreads = 512;                // stage-1 block; panel rows in registers
constexpr int kBandMaxN = kBand + kBandThreads;  // one panel row per thread
constexpr int kTwoStageMinN = 4 * kBand;           // smallest threshold accepted
__host__ __device__ inline bool two_stage(int n, int ts_min) {
  return ts_min > 0 && n >= ts_min && n <= kBandMaxN;
}
// kept vectors per work item of the bulge-chasing back-transformation
__host__ __device__ inline int bc_vectors(int n) {
  int nv = 16;
  while (nv > 1 && nv * n > 2048) nv >>= 1;
  return nv;
}
int eig_threshold();  // sectors with n > threshold take the tridiagonal route
int two_stage_min();  // tridiagonal-route sectors with n >= this take the two-stage reduction
int tri_split();      // many-sector groups: Householder blocks of 512 threads for n > split
int tile_order();     // contraction tiles are visited in super-tiles of order x order tiles
double tie_rel();     // relative width of the tie band at the chi-th singular value
double drop_rel();    // singular values <= drop_rel * the sector's largest are dropped

// Singular-value histogram of an update (counts of surely-kept values in bins
// of a quarter octave, bin b holding values >= 2^((b - kHistZero) / 4)) and
// the cutoff below which no value can be among the chi kept (stored after it)
constexpr int kHistBins = 8192;
constexpr int kHistZero = 4096;

constexpr int kTileClasses = 4;
constexpr int kTileElems[kTileClasses] = {1024, 2048, 4096, 8192};

struct PlanHeader {
  int32_t magic, d, chi, n_sectors;
  int32_t n_items, n_lblk, n_cblk, n_rblk;  // n_items: contraction tile items
  int32_t n_gram_items, n_w_items, n_v_items, pad0;
  int64_t c_elems, g_elems, sv_total, v_elems, w_elems;
  int64_t rows_total, cols_total;
  double flops_alg, flops_tau, bytes_alg16;
  // flops of the SVD stages (complex MAC = 8): Gram 8 m n^2, W 8 n L R,
  // vectors 8 L R n, Jacobi one sweep of 8 n^3 (reported per sweep-equivalent);
  // W and vectors are upper bounds over all n pairs (stage timing counts the
  // kept pairs on the device instead)
  double flops_gram, flops_wmat, flops_vectors, flops_jacobi;
  // tridiagonal route: Householder flops (16/3) n^3 and the blocked reduction's
  // trailing-matrix traffic (16 m^2 bytes read per column, 32 m^2 per panel)
  double flops_tridiag, bytes_tridiag;
  // host byte offsets of the device-copied region and total blob size
  int64_t dev_begin, blob_bytes;
  // byte offsets RELATIVE TO dev_begin (i.e. inside the device copy)
  int64_t o_sectors, o_rows, o_cols, o_lblk, o_cblk, o_rblk;
  int64_t o_gram_items, o_w_items, o_v_items, o_charges, o_sv_sector, o_jac_order;
  int64_t dev_bytes;  // size of the device-copied region
  // sectors grouped by Jacobi size class: jac_order[jac_off[c] .. jac_off[c+1])
  int32_t jac_off[kJacClasses + 1];
  int32_t jac_max_n[kJacClasses];
  // workspace layout (byte offsets from the start of d_work)
  int64_t w_C, w_T, w_G, w_V, w_W, w_S, w_lam, w_flag, w_slot;
  int64_t w_td, w_te, w_tau, w_kcount, w_triwork, w_hist, w_total;
  int32_t eig_threshold, pad_t;
  double tie_rel;
  // MFMA tile formulation of the contraction (see TileItem)
  int32_t n_titems, tile_off[kTileClasses + 1];
  int64_t o_titems, o_tau_sector, o_cblk_of, o_rowoff, o_coloff, o_row_info, o_col_info;
  double flops_weight;  // BS transforms: 8 sum_pairs nL nR B0 B1 (B0 + B1)
  double mfma_k4;       // sum over tile items and their centre blocks of K padded to 4
  double mfma_k32;      // the same with K padded to 32 (BF16x3 steps)
  // back-transformation items of the tridiagonal-route sectors (WY blocks of
  // reflectors, panels of kept vectors); layouts read only header fields
  int32_t n_wyt, n_wya;
  // two-stage threshold in effect, and the bulge-chasing back-transformation
  // items (chunks of kept vectors) of the two-stage sectors
  int32_t ts_min, n_bc;
  double drop_rel;  // drop rule in effect (on sigma, relative to the sector's largest)
};

// Tile formulation of the sector contraction.  A TileItem is one pass over
// a 16 x 16 tile of (left row, right column) elements: rows r0.., columns
// c0.. of the update (nr, nc of them exist); the pass stores rows
// [s0, s0 + rp) of the tile.  [lo, hi] is the union box of the pass's
// elements' boxes cr <= q <= cl (lo = min right charge over the tile's
// columns, hi = max left charge over the pass's rows); the left blocks of
// the pass's rows are a_lo .. a_lo + na - 1, the right blocks of the tile's
// columns b_lo .. b_lo + nb - 1.  The kernel keeps, in LDS, the products P_q
// of every union-box position for the pass's rows (nbox x rp x 16 elements)
// and the output offset tables (nbox x (12 + 4 (na + nb)) bytes); a tile
// whose union box is too large for 16 rows in the largest LDS class is split
// into passes of rp < 16 rows.  Items are sorted into LDS classes: class c
// holds passes whose need (elements plus tables in 8-byte units) is at most
// kTileElems[c].
constexpr int kMaxD = 16;  // the reference's limit too: 256-entry centre-charge tables
                           // (src/update_MPO_kernel.cu:236-238)
// Per bond slot of the left (right) bond, for the contraction: x = block
// index | charge component 0 << 16 | component 1 << 24 (-1 for padding
// slots), y = position of the slot inside its block.  One load gives the
// element's charges without a block-table lookup.
struct SlotInfo {
  int32_t x, y;
};
struct TileItem {
  int32_t r0, c0;
  int8_t nr, nc, s0, rp, lo0, lo1, hi0, hi1;
  int16_t a_lo, b_lo;
  int8_t na, nb, pad0, pad1;
};
static_assert(sizeof(TileItem) == 24, "TileItem layout");

struct Sector {
  int32_t t0, t1, L, R;
  int32_t n;           // min(L, R)
  int32_t row_off;     // into the rows pool
  int32_t col_off;     // into the cols pool
  int32_t left_small;  // 1 if L <= R (Gram = T T^H), else 0 (Gram = T^H T)
  int64_t c_off;       // element offset of C / T (L x R row-major)
  int64_t g_off;       // element offset of G (n x n column-major)
  int64_t v_off;       // element offset of V (L x n column-major)
  int64_t w_off;       // element offset of W (n x R row-major)
  int64_t s_off;       // element offset of S / flags / slots (n)
  int64_t tw_off;      // element offset (doubles) of the tridiagonal route's work arrays
};

struct Block {
  int32_t c0, c1, start, end;
};

// GEMM work item of the Gram / W / V stages: a 16-row x 16-col tile of a sector matrix
struct Item {
  int32_t sector;
  int32_t blk;   // left-block index (contraction) or unused
  int32_t r0;    // first row of the tile (sector-local)
  int32_t nr;    // rows in the tile (<= 16)
  int32_t c0;    // first col of the tile (sector-local)
  int32_t nc;    // cols in the tile (<= 16)
  int32_t pad0, pad1;
};

// A group: any number of planned updates run by one launch per stage.
// Device layout of the group workspace (byte offsets from d_work):
//   [UpdDesc x n_upd][work lists (int2 pairs)][plan device parts] = staging,
//   uploaded from the host image in one copy, then each update's scratch.
constexpr int32_t kGroupMagic = 0x4253475a;  // "BSGZ"

// Work-list segments of a group: (offset into its list, count), kept in
// device memory at the front of the staging region (GroupCounts), so that
// every stage kernel reads how much work there is from the device.  A group
// planned on the device launches with grids sized for its capacity, and the
// counts it planned never pass through the host.
enum GroupSeg {
  SG_UPD = 0,                          // the updates (per-update kernels); 0 for a skipped group
  SG_TILE = 1,                         // + LDS class c: contraction tile items
  SG_GRAM = SG_TILE + kTileClasses,    // Gram blocks
  SG_JAC = SG_GRAM + 1,                // + size class c < kTriClass: one-sided Jacobi sectors
  SG_JACALL = SG_JAC + kTriClass,      // every Jacobi-route sector
  SG_TRI = SG_JACALL + 1,              // every tridiagonal-route sector (two-stage ones first)
  SG_TS = SG_TRI + 1,                  // the two-stage ones
  SG_TRIBIG = SG_TS + 1,               // one-stage sectors on 512-thread blocks
  SG_TRISMALL = SG_TRIBIG + 1,         // one-stage sectors on 256-thread blocks
  SG_SEC = SG_TRISMALL + 1,            // every sector
  SG_V = SG_SEC + 1,                   // V = T Y / sigma blocks
  SG_W = SG_V + 1,                     // W = V^H C blocks
  SG_WYT = SG_W + 1,                   // WY blocks of reflectors (T factors)
  SG_WYA = SG_WYT + 1,                 // panels of kept vectors (WY application)
  SG_BC = SG_WYA + 1,                  // bulge-chasing back-transformation chunks
  SG_COUNT = SG_BC + 1
};
struct GroupCounts {
  int2 seg[SG_COUNT];
  int32_t status;  // device planning: 0, or why the group was skipped (all counts zero)
  int32_t pad[3];
};
// A stage's work: the group's list array and the segment it launches over.
struct WorkSeg {
  const int2 *list;
  const int2 *seg;
};
// below this many one-stage tridiagonal sectors in a launch (a single update):
// one block shape for all of them
constexpr int kTriFewSectors = 256;

struct GroupHeader {
  int32_t magic, n_upd, chi, pad0;
  int32_t n_gram, n_tri, n_sec, n_v, n_w, pad4;
  int32_t jac_off[kTriClass + 1];
  int32_t jac_max_n[kTriClass];
  int32_t tri_max_n, pad1;
  int64_t o_desc, o_gram, o_jac, o_tri, o_sec, o_v, o_w, o_plans, o_ws;
  int64_t staging_bytes, work_bytes, max_sv, host_bytes;
  double flops_alg, flops_gram, flops_jacobi, flops_wmat, flops_vectors, bytes_alg16;
  double c_elems, g_elems, v_elems, w_elems, sv_total;
  double flops_tridiag, bytes_tridiag;
  int32_t n_tiles, d_max;
  int32_t tile_off[kTileClasses + 1], has_lc;  // has_lc: some update weights centre rows (LC)
  int64_t o_tiles;
  double flops_weight, mfma_k4, mfma_k32;
  int32_t tri_split, tri_nbig;  // the block-shape split in effect; sectors with n > tri_split
  int32_t n_wyt, n_wya;         // back-transformation items: (sector, block), (sector, panel)
  int64_t o_wyt, o_wya;
  // two-stage sectors: the head of the tridiagonal list (n_ts of them, largest
  // first; ts_nbig with two panel rows per thread), and their bulge-chasing
  // back-transformation items (sector | chunk << 20)
  int32_t n_ts, ts_nbig, n_bc, ts_max_n;
  int64_t o_bc;
  int64_t o_counts;  // GroupCounts at the front of the staging region
  int32_t dev_planned, pad_dp;  // counts are capacities: the group was planned on the device
};
constexpr int64_t kGroupHostHead = 1024;  // host blob: header, then the staging image
static_assert(sizeof(GroupHeader) <= kGroupHostHead, "group header too large");

inline const PlanHeader *header(const void *plan) {
  return reinterpret_cast<const PlanHeader *>(plan);
}
template <class T>
inline const T *host_array(const void *plan, int64_t dev_off) {
  const PlanHeader *h = header(plan);
  return reinterpret_cast<const T *>(reinterpret_cast<const char *>(plan) + h->dev_begin + dev_off);
}
template <class T>
inline T *dev_array(void *work, int64_t dev_off) {
  return reinterpret_cast<T *>(reinterpret_cast<char *>(work) + dev_off);
}

int validate_plan(const void *plan, int d, int chi);

// ------------------------------------------------------------- timing --
// Optional per-stage timing with HIP events recorded on the launch stream
// (bsmpo_timing_enable / bsmpo_timing_read).  Off by default: no events.
enum Stage {
  ST_BSTABLE = 0, ST_CONTRACT, ST_GRAM, ST_JACOBI, ST_VECTORS, ST_WMAT, ST_TOPK,
  ST_SCATTER, ST_OBSERVABLES, ST_TRIDIAG, ST_BISECT, ST_TRIVEC, ST_BACKTRANS, ST_COUNT
};
bool timing_on();
// raise a kernel's dynamic-LDS limit (once per kernel and device, thread-safe)
int ensure_lds_limit(const void *kernel, int bytes);
// begin returns a token (or -1 when timing is off); end closes it.
int timing_begin(int stage, hipStream_t stream);
void timing_end(int token, hipStream_t stream, double flops, double bytes);
// flops of the kept-pair stages (W, V, back-transformation) counted on the device
int kept_flops_read(double *out, bool reset);

// -------------------------------------------------------------- kernels --
// launchers implemented in the .hip files; all enqueue on `stream`.
int launch_bs_table(int d, int nb, const double *d_B, double *d_U3, hipStream_t stream);
int launch_contract(const GroupHeader &g, const char *d_work, int precision, hipStream_t stream);
int launch_svd(const GroupHeader &g, const char *d_work, hipStream_t stream);
// parts of launch_svd, for graphs built from them (layer programs)
int launch_svd_head(const GroupHeader &g, const char *d_work, hipStream_t stream);
int launch_tridiag_part(const GroupHeader &g, const char *d_work, hipStream_t stream, int part);
int launch_tri_sturm(const GroupHeader &g, const char *d_work, hipStream_t stream);
int launch_truncate(const GroupHeader &g, const char *d_work, hipStream_t stream);
int launch_wy_backtransform(const GroupHeader &g, const char *d_work, hipStream_t stream);
int launch_band(const GroupHeader &g, const char *d_work, hipStream_t stream);
int launch_bulge_apply(const GroupHeader &g, const char *d_work, hipStream_t stream);

// group construction (group.cpp)
struct UpdateIO {
  double *Glc, *Gcr;
  const double *LR, *U3, *LL, *LC;
  int32_t *charge_c;
  double *lambda_c;
  int32_t *num_lambda;
  double *C_override, *T_override;  // bsmpo_contract writes to caller buffers
  int edge;
};
int group_query(int n, const void *const *plans, int64_t *host_bytes, int64_t *work_bytes);
// contract_only: build only what the contraction stage reads (the SVD-stage
// work lists stay empty); no heap allocation in either case
int group_build(int n, const void *const *plans, const UpdateIO *io, char *d_work,
                char *host_blob, int64_t host_bytes, bool contract_only = false);

// ------------------------------------------------------ device planning --
// A layer program (layer.cpp) plans its groups on the device (devplan.hip):
// the same Sector / Block / Item / TileItem arrays and work lists the host
// planner builds, written into arenas of fixed capacity.  Per update, a fixed
// part (tables whose size depends on d and chi only) and a variable part
// (pools and item arrays) placed by a scan over the group's updates.
struct DpUpdate {
  const int32_t *CL, *CC, *CR;  // charges of bonds mode, mode + 1, mode + 2 (chi x 2)
  double *Glc, *Gcr;
  const double *LR;
  double *U3;                   // the program's BS table of this update
  int32_t *charge_c;
  double *lambda_c;
  int32_t *num_lambda;
  int32_t edge, group;          // group: index of the update's group in its set
  int32_t local;                // position inside the group
  int32_t first_replay;         // replays before it skip the update (light cone)
};
// sizes found by the first planning kernel, offsets by the group scan (one
// record per update of the replay's set)
struct DpSizes {
  int32_t status, n_sec, n_lblk, n_cblk, n_rblk, rows_total, cols_total, n_gram;
  int32_t n_w, n_v, sv_total, n_tri, n_titems, active;  // active: runs in this replay
  int32_t desc, pad2;  // its descriptor (and list) index among the group's active updates
  int32_t jac_cnt[kTriClass], tile_cnt[kTileClasses];
  int64_t c_elems, g_elems, v_elems, w_elems, tw;
  int64_t var_bytes, work_bytes, var_off, work_off;
  // positions of the update's entries in the group's work lists
  int32_t l_tile[kTileClasses], l_jac[kTriClass];
  int32_t l_gram, l_tri, l_sec, l_v, l_w, pad1;
};
// device planning status codes (GroupCounts::status, DpSizes::status, and
// the program's sticky status)
enum DpStatus {
  DP_OK = 0,
  DP_OVER_CAPACITY = 1,  // a group needed more than its arenas or lists hold
  DP_BAD_CHARGES = 2,    // charges out of range or not sorted
  DP_TILE_TOO_BIG = 3,   // a contraction tile does not fit the LDS classes
  DP_SECTOR_TOO_BIG = 4, // a tridiagonal-route sector above the eigensolver's limit
  DP_NO_TABLE = 5,       // replayed more often than BS matrices were provided
  DP_SKIPPED = 6         // an earlier replay failed: this one did nothing
};
// capacities of one group's arenas and work lists
struct DpCaps {
  int64_t var_bytes, work_bytes;
  int32_t tiles, gram, jac, tri, sec, v, w, wyt, wya, pad;
};
// what a group needed in its last planning (written even when over capacity)
struct DpNeeds {
  int64_t var_bytes, work_bytes;
  int32_t tiles, gram, jac, tri, sec, v, w, wyt, wya, status;
};
struct DpGroup {
  char *staging;  // GroupCounts, UpdDesc[n_upd], the work lists (at the capacity header's offsets)
  char *fixed;    // fixed plan parts, fixed_bytes per update
  char *var;      // variable plan parts
  char *work;     // scratch arena (shared by the groups launched on one stream)
  int2 *tri_tmp;  // the tridiagonal list before its sort (caps.tri entries)
  DpCaps caps;
  DpNeeds needs;
  int64_t o_desc, o_tiles, o_gram, o_jac, o_tri, o_sec, o_v, o_w, o_wyt, o_wya;
  int32_t first[2], n_upd[2];  // per update set: its updates [first, first + n_upd)
};
struct DpParams {
  int32_t d, chi, thr, tri_split, n_groups, n_layers, tile_order, pad0;
  int32_t n_sets, max_upd;     // replay r runs update set r % n_sets
  int32_t n_upd[2], off[2];    // per set: updates, first record in the update array
  double tie_rel, drop_rel;
  int64_t fixed_bytes;
  int32_t *state;  // [0] sticky status, [1] replay counter, [2] first failed replay,
                   // [3] failure seen by the BS-table kernel
  const double *B; // BS matrices, [n_layers][max_upd][2][2] complex128
};
__device__ inline int dp_set(const DpParams &P) { return P.state[1] % P.n_sets; }
// fixed per-update plan part: byte offsets (multiples of 256)
struct FixedLayout {
  int64_t o_sectors, o_lblk, o_cblk, o_rblk, o_tau_sector, o_cblk_of, o_rowoff, o_coloff;
  int64_t o_row_info, o_col_info, o_charges, o_secx, bytes;
};
__host__ __device__ inline int64_t dp_align(int64_t x) { return (x + 255) / 256 * 256; }
__host__ __device__ inline FixedLayout fixed_layout(int d, int chi) {
  const int64_t nb = (int64_t)d * d, D1 = d + 1;  // blocks per bond, sectors: at most d^2
  FixedLayout f;
  int64_t o = 0;
  f.o_sectors = o; o = dp_align(o + (int64_t)sizeof(Sector) * nb);
  f.o_lblk = o; o = dp_align(o + (int64_t)sizeof(Block) * nb);
  f.o_cblk = o; o = dp_align(o + (int64_t)sizeof(Block) * nb);
  f.o_rblk = o; o = dp_align(o + (int64_t)sizeof(Block) * nb);
  f.o_tau_sector = o; o = dp_align(o + 4 * D1 * D1);
  f.o_cblk_of = o; o = dp_align(o + 4 * D1 * D1);
  f.o_rowoff = o; o = dp_align(o + 4 * nb * nb);
  f.o_coloff = o; o = dp_align(o + 4 * nb * nb);
  f.o_row_info = o; o = dp_align(o + (int64_t)sizeof(SlotInfo) * chi);
  f.o_col_info = o; o = dp_align(o + (int64_t)sizeof(SlotInfo) * chi);
  f.o_charges = o; o = dp_align(o + 4 * 6 * (int64_t)chi);
  f.o_secx = o; o = dp_align(o + 16 * nb);  // per sector: gram, W, V item offsets, class rank
  f.bytes = o;
  return f;
}
// variable per-update plan part: byte offsets of its arrays, and its size
struct VarLayout {
  int64_t o_rows, o_cols, o_gram, o_w, o_v, o_sv, o_titems, bytes;
};
__host__ __device__ inline VarLayout var_layout(int64_t rows, int64_t cols, int64_t gram,
                                                int64_t wi, int64_t vi, int64_t sv,
                                                int64_t titems) {
  VarLayout v;
  int64_t o = 0;
  v.o_rows = o; o = dp_align(o + 4 * rows);
  v.o_cols = o; o = dp_align(o + 4 * cols);
  v.o_gram = o; o = dp_align(o + (int64_t)sizeof(Item) * gram);
  v.o_w = o; o = dp_align(o + (int64_t)sizeof(Item) * wi);
  v.o_v = o; o = dp_align(o + (int64_t)sizeof(Item) * vi);
  v.o_sv = o; o = dp_align(o + 4 * sv);
  v.o_titems = o; o = dp_align(o + (int64_t)sizeof(TileItem) * titems);
  v.bytes = o;
  return v;
}
// per-update scratch: the same arrays and sizes as a host plan's workspace
// (plan.cpp), offsets from the update's scratch base
struct WorkLayout {
  int64_t C, T, G, V, W, S, lam, flag, slot, td, te, tau, kcount, triwork, hist, bytes;
};
__host__ __device__ inline WorkLayout work_layout(int64_t c_elems, int64_t g_elems, int64_t v_elems,
                                                  int64_t w_elems, int64_t sv, int64_t n_sec,
                                                  int64_t tw) {
  WorkLayout L;
  int64_t w = 0;
#define BSMPO_WPLACE(f, b) L.f = w; w = dp_align(w + ((b) > 1 ? (b) : 1));
  BSMPO_WPLACE(C, 16 * c_elems)
  BSMPO_WPLACE(T, 16 * c_elems)
  BSMPO_WPLACE(G, 16 * g_elems)
  BSMPO_WPLACE(V, 16 * v_elems)
  BSMPO_WPLACE(W, 16 * w_elems)
  BSMPO_WPLACE(S, 8 * sv)
  BSMPO_WPLACE(lam, 8 * sv)
  BSMPO_WPLACE(flag, 4 * sv)
  BSMPO_WPLACE(slot, 4 * sv)
  BSMPO_WPLACE(td, 8 * sv)
  BSMPO_WPLACE(te, 8 * sv)
  BSMPO_WPLACE(tau, 16 * sv)
  BSMPO_WPLACE(kcount, 4 * n_sec)
  BSMPO_WPLACE(triwork, 8 * tw)
  BSMPO_WPLACE(hist, 4 * kHistBins + 8)
#undef BSMPO_WPLACE
  L.bytes = w;
  return L;
}
// tridiagonal-route work area of a sector of size n (one-stage reduction):
// the inverse iteration's per-lane arrays, the Householder panels and the
// back-transformation's T blocks share it
__host__ __device__ inline int64_t tri_work_doubles(int n) {
  const int nblk = (n - 1 + kWyBlock - 1) / kWyBlock;
  const int64_t wy = 2 * (int64_t)kWyBlock * kWyBlock * nblk;
  const int lanes = n < kTriVecThreads ? n : kTriVecThreads;
  int64_t a = (int64_t)n * (6 * lanes > 4 * kTriPanel ? 6 * lanes : 4 * kTriPanel);
  return a > wy ? a : wy;
}
// launchers of the planning kernels (devplan.hip)
int launch_device_plan(const DpParams &p, const DpUpdate *d_upd, DpGroup *d_groups,
                       DpSizes *d_sizes, hipStream_t stream);
int launch_bs_table_layer(const DpParams &p, const DpUpdate *d_upd, hipStream_t stream);
int launch_layer_seek(int32_t *d_state, int replay, hipStream_t stream);
constexpr int kTriMaxN = 2048;   // largest sector of the tridiagonal route
constexpr int kDpMaxChi = kTriMaxN;  // largest chi a layer program plans
inline const GroupHeader *group_header(const void *blob) {
  return reinterpret_cast<const GroupHeader *>(blob);
}

}  // namespace bsmpo
