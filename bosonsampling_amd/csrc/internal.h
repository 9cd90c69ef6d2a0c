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
constexpr int kTriBtSplit = 512;  // back-transformation: sectors above take the 16-row class
int eig_threshold();  // sectors with n > threshold take the tridiagonal route
int tri_split();      // many-sector groups: Householder blocks of 512 threads for n > split
double tie_rel();     // relative width of the tie band at the chi-th singular value

struct PlanHeader {
  int32_t magic, d, chi, n_sectors;
  int32_t n_items, n_lblk, n_cblk, n_rblk;
  int32_t n_gram_items, n_w_items, n_v_items, pad0;
  int64_t c_elems, g_elems, sv_total, v_elems, w_elems;
  int64_t rows_total, cols_total;
  double flops_alg, flops_tau, bytes_alg16;
  // flops of the SVD stages (complex MAC = 8): Gram 8 m n^2, W 8 n L R,
  // vectors 8 L R n, Jacobi one sweep of 8 n^3 (reported per sweep-equivalent)
  double flops_gram, flops_wmat, flops_vectors, flops_jacobi;
  // tridiagonal route: Householder flops (16/3) n^3 and the blocked reduction's
  // trailing-matrix traffic (16 m^2 bytes read per column, 32 m^2 per panel)
  double flops_tridiag, bytes_tridiag;
  // host byte offsets of the device-copied region and total blob size
  int64_t dev_begin, blob_bytes;
  // byte offsets RELATIVE TO dev_begin (i.e. inside the device copy)
  int64_t o_sectors, o_rows, o_cols, o_lblk, o_cblk, o_rblk, o_items;
  int64_t o_gram_items, o_w_items, o_v_items, o_charges, o_sv_sector, o_jac_order;
  int64_t dev_bytes;  // size of the device-copied region
  // sectors grouped by Jacobi size class: jac_order[jac_off[c] .. jac_off[c+1])
  int32_t jac_off[kJacClasses + 1];
  int32_t jac_max_n[kJacClasses];
  // workspace layout (byte offsets from the start of d_work)
  int64_t w_C, w_T, w_G, w_V, w_W, w_S, w_lam, w_flag, w_slot;
  int64_t w_td, w_te, w_tau, w_kcount, w_triwork, w_total;
  int32_t eig_threshold, pad_t;
  double tie_rel;
  // pair formulation of the contraction (see PairItem)
  int32_t n_pitems, pair_ok;  // pair_ok = 0: a box side exceeds kPairMaxB, use the tile kernel
  int64_t o_pitems, o_tau_sector, o_cblk_of, o_rowoff, o_coloff;
  double flops_pair_weight;   // 8 sum_pairs nL nR B0 B1 (B0 + B1)
};

// Pair formulation of the sector contraction.  For a left block a (charge cl)
// and a right block b (cr <= cl) every tau and every centre charge cc of the
// box cr <= q <= cl take part, and the BS weight factorises over the two
// charge components:
//   u(cl, cc, cr; tau) = alpha[tau0][cc0] conj(beta[tau1][cc1]),
//   alpha[t][c] = U3[cl0 - t][t - cr0][cl0 - c], beta the same on component 1.
// So C_tau[a, b] = sum_cc u P_cc with P_cc = Glc[a, cc] Gcr[cc, b] is
//   C[t0][t1] = sum_c0 alpha[t0][c0] sum_c1 conj(beta[t1][c1]) P[c0][c1],
// computed once per (a, b) for every tau.  A PairItem is a tile of tl x tr
// elements of one (a, b) pair.
constexpr int kPairMaxB = 16;     // largest box side handled (d <= 16 always fits)
constexpr int kPairLds = 3072;    // B0 B1 tl tr <= kPairLds (LDS elements per item)
struct PairItem {
  int32_t a, b, l0, r0, tl, tr, pad0, pad1;
};

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

// contraction / GEMM work item: a 16-row x 16-col tile of a sector output
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
struct GroupHeader {
  int32_t magic, n_upd, chi, pad0;
  int32_t n_items, n_gram, n_tri, n_sec, n_v, n_w;
  int32_t jac_off[kTriClass + 1];
  int32_t jac_max_n[kTriClass];
  int32_t tri_max_n, pad1;
  int64_t o_desc, o_items, o_gram, o_jac, o_tri, o_sec, o_v, o_w;
  int64_t staging_bytes, work_bytes, max_sv, host_bytes;
  double flops_alg, flops_gram, flops_jacobi, flops_wmat, flops_vectors, bytes_alg16;
  double c_elems, g_elems, v_elems, w_elems, sv_total;
  double flops_tridiag, bytes_tridiag;
  int32_t n_pitems, pair_ok;  // pair_ok: every plan of the group supports the pair kernel
  int64_t o_pitems;
  double flops_pair_weight;
  int32_t tri_split, tri_nbig;  // the block-shape split in effect; sectors with n > tri_split
  int32_t tri_nbt, pad2;        // sectors with n > kTriBtSplit
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
// begin returns a token (or -1 when timing is off); end closes it.
int timing_begin(int stage, hipStream_t stream);
void timing_end(int token, hipStream_t stream, double flops, double bytes);

// -------------------------------------------------------------- kernels --
// launchers implemented in the .hip files; all enqueue on `stream`.
int launch_bs_table(int d, int nb, const double *d_B, double *d_U3, hipStream_t stream);
int launch_contract(const GroupHeader &g, const char *d_work, int precision, hipStream_t stream);
int launch_contract_lowp(const GroupHeader &g, const char *d_work, int precision,
                         hipStream_t stream);
int launch_contract_pair(const GroupHeader &g, const char *d_work, int precision,
                         hipStream_t stream);
int launch_svd(const GroupHeader &g, const char *d_work, hipStream_t stream);
int launch_truncate(const GroupHeader &g, const char *d_work, hipStream_t stream);

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
int group_build(int n, const void *const *plans, const UpdateIO *io, char *d_work,
                char *host_blob, int64_t host_bytes);
inline const GroupHeader *group_header(const void *blob) {
  return reinterpret_cast<const GroupHeader *>(blob);
}

}  // namespace bsmpo
