// Device-side helpers shared by the kernels: complex arithmetic and the
// complex 16x16x4 FP64 MFMA tile step (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "internal.h"

namespace bsmpo {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// Pointer into global memory (address space 1).  Descriptor fields are
// generic pointers, whose loads compile to flat instructions; those count
// against the LDS wait counter too, so every LDS wait would also wait for
// the global loads in flight.  Hot loops cast to this type first.  (The host
// pass of a .hip file parses kernel bodies without generating them; there the
// HIP vector types do not accept qualified address spaces, so it sees a plain
// pointer.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
#else
template <typename T>
using gptr = T *;
#endif
template <typename T>
__device__ __forceinline__ gptr<T> gp(T *p) {
  return (gptr<T>)p;
}
typedef double dbl2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 gload2(gptr<const double2> p, int64_t i) {
  const dbl2v v = ((gptr<const dbl2v>)p)[i];
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void gstore2(gptr<double2> p, int64_t i, double x, double y) {
  ((gptr<dbl2v>)p)[i] = dbl2v{x, y};
}
// Pointer into LDS (address space 3), where code is written once for an
// array in LDS or in global memory: a generic pointer would compile to flat
// loads, which count against both wait counters.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using lptr = __attribute__((address_space(3))) T *;
#else
template <typename T>
using lptr = T *;
#endif

struct cplx {
  double x, y;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return cplx{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cplx cmulc(cplx a, cplx b) {  // a * conj(b)
  return cplx{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
__device__ __forceinline__ cplx cconj(cplx a) { return cplx{a.x, -a.y}; }
__device__ __forceinline__ cplx cload(const double *p, int64_t i) {
  const double2 v = reinterpret_cast<const double2 *>(p)[i];
  return cplx{v.x, v.y};
}
__device__ __forceinline__ void cstore(double *p, int64_t i, cplx v) {
  reinterpret_cast<double2 *>(p)[i] = make_double2(v.x, v.y);
}

// sqrt(n!) for n = 0..32, passed by value to the BS-table kernels
struct BsArgs {
  double sqf[33];
};
inline BsArgs bs_args() {
  BsArgs a;
  long double f = 1.0L;
  for (int n = 0; n <= 32; ++n) {
    if (n > 0) f *= (long double)n;
    a.sqf[n] = (double)sqrtl(f);
  }
  return a;
}

__device__ __forceinline__ double binom_u64(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  if (k > n - k) k = n - k;
  unsigned long long r = 1;
  for (int i = 1; i <= k; ++i) r = r * (unsigned long long)(n - k + i) / (unsigned long long)i;
  return (double)r;
}

__device__ __forceinline__ cplx cpow_int(cplx a, int e) {
  cplx r{1.0, 0.0};
  for (int i = 0; i < e; ++i) r = cmul(r, a);
  return r;
}

// One K=4 step of a complex 16x16 tile on FP64 MFMA.
// Lane layout of v_mfma_f64_16x16x4f64 (64 lanes), measured on gfx950:
//   A[i][k] from lane i + 16 k,  B[k][j] from lane j + 16 k,
//   D[l / 16 + 4 r][l % 16] in element r of lane l's accumulator.
__device__ __forceinline__ int mfma64_row(int kq, int r) { return kq + 4 * r; }
__device__ __forceinline__ void cmfma_f64(dbl4 &accr, dbl4 &acci, cplx a, cplx b) {
  accr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, accr, 0, 0, 0);
  accr = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, accr, 0, 0, 0);
  acci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, acci, 0, 0, 0);
  acci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, acci, 0, 0, 0);
}

// Everything a kernel needs about one update of a group: the update's plan
// arrays (device copies), its scratch arrays and the MPO buffers it reads
// and writes.  Kernels take an array of these plus a work list of
// (update, local index) pairs, so one launch covers any number of updates.
struct UpdDesc {
  const Sector *sectors;
  const int32_t *rows, *cols;
  const Block *lblk, *cblk, *rblk;
  const Item *gram_items, *w_items, *v_items;
  const int32_t *charges, *sv_sector;
  double *C, *T, *G, *V, *W, *S, *lam, *td, *te, *tau, *triwork;
  int32_t *flag, *slot, *kcount;
  int32_t *hist;  // kHistBins counts, then the cutoff (a double)
  double *Glc, *Gcr;
  const double *LR, *U3, *LL, *LC;
  int32_t *charge_c;
  double *lambda_c;
  int32_t *num_lambda;
  int32_t d, chi, edge, n_sectors, n_cblk, eig_thr;  // eig_thr: n above it = tridiagonal route
  int64_t sv_total;
  double tie_rel;
  // tile formulation of the contraction
  const TileItem *titems;
  const int32_t *tau_sector, *cblk_of, *rowoff, *coloff;
  const int2 *row_info, *col_info;  // per bond slot, see SlotInfo
  int32_t n_lblk, n_rblk;
  int32_t ts_min, pad_ts;  // two-stage threshold of the plan (see two_stage)
  double drop_rel2;        // lambda <= drop_rel2 * lambda_max of the sector: dropped
};

// Plan arrays as seen by kernels (device copy at the front of the workspace).
struct DevPlan {
  const Sector *sectors;
  const int32_t *rows;
  const int32_t *cols;
  const Block *lblk;
  const Block *cblk;
  const Block *rblk;
  const Item *gram_items;
  const Item *w_items;
  const Item *v_items;
  const int32_t *charges;  // CL, CC, CR (3 x chi x 2)
  const int32_t *sv_sector;
  const int32_t *jac_order;
};

__host__ __device__ inline DevPlan dev_plan(const PlanHeader &h, const void *work) {
  const char *b = reinterpret_cast<const char *>(work);
  DevPlan p;
  p.sectors = reinterpret_cast<const Sector *>(b + h.o_sectors);
  p.rows = reinterpret_cast<const int32_t *>(b + h.o_rows);
  p.cols = reinterpret_cast<const int32_t *>(b + h.o_cols);
  p.lblk = reinterpret_cast<const Block *>(b + h.o_lblk);
  p.cblk = reinterpret_cast<const Block *>(b + h.o_cblk);
  p.rblk = reinterpret_cast<const Block *>(b + h.o_rblk);
  p.gram_items = reinterpret_cast<const Item *>(b + h.o_gram_items);
  p.w_items = reinterpret_cast<const Item *>(b + h.o_w_items);
  p.v_items = reinterpret_cast<const Item *>(b + h.o_v_items);
  p.charges = reinterpret_cast<const int32_t *>(b + h.o_charges);
  p.sv_sector = reinterpret_cast<const int32_t *>(b + h.o_sv_sector);
  p.jac_order = reinterpret_cast<const int32_t *>(b + h.o_jac_order);
  return p;
}

// XCD-aware order of a grid-stride walk over a work list: walk position p
// runs on XCD p % 8 (workgroups are dispatched round robin over the 8 XCDs,
// and the grid is a multiple of 8), so each XCD takes one contiguous eighth
// of the list, and neighbouring items, which read the same operand panels,
// share its L2.  Returns the item of position p (>= nlist: none).
__device__ __forceinline__ int xcd_item(int p, int nlist) {
  const int chunk = (nlist + 7) >> 3;
  return (p & 7) * chunk + (p >> 3);
}
__device__ __forceinline__ int xcd_positions(int nlist) { return ((nlist + 7) >> 3) << 3; }
// a grid for such a walk: a multiple of 8, at most `cap` workgroups
inline int xcd_grid(int n, int cap = 16384) {
  const int g = (n + 7) / 8 * 8;
  return g < 8 ? 8 : (g > cap ? cap : g);
}

inline int grid_for(int64_t waves_needed) {
  int64_t blocks = (waves_needed + 3) / 4;
  if (blocks < 1) blocks = 1;
  if (blocks > 8192) blocks = 8192;
  return (int)blocks;
}

}  // namespace bsmpo
