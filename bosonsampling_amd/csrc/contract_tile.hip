// Sector contraction on MFMA (gfx950): src/update_MPO_kernel.cu:193-734
// semantics, with the LR weight of src/beam_splitter_parallel/Rank.py:114-118
// (T = C diag(LR)) and the LL / LC weights of the pure-state form
// (src/update_kernel.cu:594-595,706-715).
//
// A work item is a 16 x 16 tile of (left row, right column) elements.  For
// element (i, j) with left charge cl and right charge cr the products
//   P_q[i, j] = sum_{c in centre block q} Glc[i, c] Gcr[c, j],  cr <= q <= cl,
// feed every sector tau of the same box through the BS weight, which
// factorises over the two charge components:
//   C_tau[i, j] = sum_q0 W_A0[t0][q0] sum_q1 conj(W_A1[t1][q1]) P_q[i, j],
//   W_A[t][c] = U3[A - t][t][A - c],  A = cl - cr,  t = tau - cr,  c = q - cr.
// Phase 1 forms P_q for every non-empty centre block q of the tile's union
// box [min cr, max cl] as one 16 x 16 MFMA product (K = rows of block q)
// and stores it in LDS; waves take the blocks round robin.  Phase 2 applies
// the two small transforms per element (vector ALUs) and writes C_tau and
// T_tau into every sector of the element's box.  Every product of the update
// is formed once; tiles whose union box holds many blocks store only some
// of their rows per pass (the planner splits them, see TileItem).
//
// Arithmetic: complex128 on v_mfma_f64_16x16x4f64; complex64 on
// v_mfma_f32_16x16x4f32 with FP32 accumulation; split-BF16x3 on gfx950's
// v_mfma_f32_16x16x32_bf16 (x = hi + lo, x y ~ hi hi + hi lo + lo hi, FP32
// accumulation).  Operands are read from the complex128 MPO and rounded on
// load; C and T are stored in complex128 for the FP64 SVD stage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "common.h"

namespace bsmpo {

template <typename T>
struct cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cfma(cx<T> acc, cx<T> a, cx<T> b) {
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
  return acc;
}
template <typename T>
__device__ __forceinline__ cx<T> cfmac(cx<T> acc, cx<T> a, cx<T> b) {  // acc + conj(a) b
  acc.x += a.x * b.x + a.y * b.y;
  acc.y += a.x * b.y - a.y * b.x;
  return acc;
}

// offset of W_A in the packed table: sum_{a < A} (a + 1)^2
__host__ __device__ constexpr int wtab_off(int A) { return A * (A + 1) * (2 * A + 1) / 6; }

// ---------------------------------------------------------- MFMA steps --
// v_mfma_f32_16x16x4f32: A[i][k] lane i + 16 k, B[k][j] lane j + 16 k,
// D[4 (l / 16) + r][l % 16] in element r of lane l.
__device__ __forceinline__ void cmfma_f32(flt4 &accr, flt4 &acci, float ar, float ai, float br,
                                          float bi) {
  accr = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, accr, 0, 0, 0);
  accr = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, bi, accr, 0, 0, 0);
  acci = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, acci, 0, 0, 0);
  acci = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, acci, 0, 0, 0);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// round-to-nearest-even float -> bf16, and the residual split x = hi + lo
__device__ __forceinline__ unsigned short bf16_bits(float x) {
  unsigned int b = __float_as_uint(x);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (unsigned short)(b >> 16);
}
__device__ __forceinline__ __bf16 bf16_of(unsigned short h) {
  return __builtin_bit_cast(__bf16, h);
}
__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
  const unsigned short h = bf16_bits(x);
  hi = bf16_of(h);
  lo = bf16_of(bf16_bits(x - __uint_as_float(((unsigned int)h) << 16)));
}
// v_mfma_f32_16x16x32_bf16: lane i + 16 g holds row i of A (column i of B)
// at eight K positions (element e); the accumulator layout is that of the
// 16x16x4 f32 form.
__device__ __forceinline__ flt4 mfma_bf16(bf16x8 a, bf16x8 b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ flt4 mfma3(bf16x8 xh, bf16x8 xl, bf16x8 yh, bf16x8 yl, flt4 acc) {
  acc = mfma_bf16(xl, yh, acc);
  acc = mfma_bf16(xh, yl, acc);
  return mfma_bf16(xh, yh, acc);
}

// ------------------------------------------------------------- phase 1 --
// P = Glc[rows, block] diag(LC) Gcr[block, cols] per centre block, in chunks
// of KC centre rows.  Policy per arithmetic: Frag holds one chunk's operand
// loads (complex128 from the MPO), mma folds a chunk into the accumulators,
// store writes rows [s0, s0 + rp) of the 16 x 16 result to LDS.
template <int PREC, bool LCW>
struct Prod;

template <bool LCW>
struct Frag4 {  // four K = 4 steps: lane (i, kq) holds A[i][4 s + kq], B[4 s + kq][i]
  double2 a[4], b[4];
  double w[LCW ? 4 : 1];
};

template <bool LCW>
__device__ __forceinline__ void load4(Frag4<LCW> &f, gptr<const double2> gl,
                                      gptr<const double2> gc, const double *LC, int64_t chi,
                                      bool rv, bool cv, int k0, int qe, int kq) {
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = k0 + 4 * s + kq;
    const bool kv = k < qe;
    f.a[s] = (kv && rv) ? gload2(gl, k) : make_double2(0.0, 0.0);
    f.b[s] = (kv && cv) ? gload2(gc, (int64_t)k * chi) : make_double2(0.0, 0.0);
    if constexpr (LCW) f.w[s] = kv ? LC[k] : 0.0;
  }
}

template <bool LCW>
struct Prod<BSMPO_C128, LCW> {
  typedef double T;
  typedef Frag4<LCW> Frag;
  struct Acc {
    dbl4 r, i;
  };
  static constexpr int KC = 16;
  static __device__ __forceinline__ void zero(Acc &c) {
    c.r = dbl4{0.0, 0.0, 0.0, 0.0};
    c.i = dbl4{0.0, 0.0, 0.0, 0.0};
  }
  static __device__ __forceinline__ void load(Frag &f, gptr<const double2> gl,
                                              gptr<const double2> gc, const double *LC,
                                              int64_t chi, bool rv, bool cv, int k0, int qe,
                                              int lane) {
    load4<LCW>(f, gl, gc, LC, chi, rv, cv, k0, qe, lane >> 4);
  }
  static __device__ __forceinline__ void mma(Acc &c, const Frag &f) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      double ax = f.a[s].x, ay = f.a[s].y;
      if constexpr (LCW) ax *= f.w[s], ay *= f.w[s];
      cmfma_f64(c.r, c.i, cplx{ax, ay}, cplx{f.b[s].x, f.b[s].y});
    }
  }
  static __device__ __forceinline__ void store(const Acc &c, cx<double> *P, int lane, int s0,
                                               int rp) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = mfma64_row(lane >> 4, r) - s0;
      if (row >= 0 && row < rp) P[row * 16 + (lane & 15)] = cx<double>{c.r[r], c.i[r]};
    }
  }
};

template <bool LCW>
struct Prod<BSMPO_C64, LCW> {
  typedef float T;
  typedef Frag4<LCW> Frag;
  struct Acc {
    flt4 r, i;
  };
  static constexpr int KC = 16;
  static __device__ __forceinline__ void zero(Acc &c) {
    c.r = flt4{0.f, 0.f, 0.f, 0.f};
    c.i = flt4{0.f, 0.f, 0.f, 0.f};
  }
  static __device__ __forceinline__ void load(Frag &f, gptr<const double2> gl,
                                              gptr<const double2> gc, const double *LC,
                                              int64_t chi, bool rv, bool cv, int k0, int qe,
                                              int lane) {
    load4<LCW>(f, gl, gc, LC, chi, rv, cv, k0, qe, lane >> 4);
  }
  static __device__ __forceinline__ void mma(Acc &c, const Frag &f) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      double ax = f.a[s].x, ay = f.a[s].y;
      if constexpr (LCW) ax *= f.w[s], ay *= f.w[s];
      cmfma_f32(c.r, c.i, (float)ax, (float)ay, (float)f.b[s].x, (float)f.b[s].y);
    }
  }
  static __device__ __forceinline__ void store(const Acc &c, cx<float> *P, int lane, int s0,
                                               int rp) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 4 * (lane >> 4) + r - s0;
      if (row >= 0 && row < rp) P[row * 16 + (lane & 15)] = cx<float>{c.r[r], c.i[r]};
    }
  }
};

template <bool LCW>
struct Prod<BSMPO_BF16X3, LCW> {
  typedef float T;
  struct Frag {  // one K = 32 step: lane (i, g) holds A[i][8 g + e], B[8 g + e][i]
    double2 a[8], b[8];
    double w[LCW ? 8 : 1];
  };
  struct Acc {
    flt4 r, i;
  };
  static constexpr int KC = 32;
  static __device__ __forceinline__ void zero(Acc &c) {
    c.r = flt4{0.f, 0.f, 0.f, 0.f};
    c.i = flt4{0.f, 0.f, 0.f, 0.f};
  }
  static __device__ __forceinline__ void load(Frag &f, gptr<const double2> gl,
                                              gptr<const double2> gc, const double *LC,
                                              int64_t chi, bool rv, bool cv, int k0, int qe,
                                              int lane) {
    const int g = lane >> 4;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + 8 * g + e;
      const bool kv = k < qe;
      f.a[e] = (kv && rv) ? gload2(gl, k) : make_double2(0.0, 0.0);
      f.b[e] = (kv && cv) ? gload2(gc, (int64_t)k * chi) : make_double2(0.0, 0.0);
      if constexpr (LCW) f.w[e] = kv ? LC[k] : 0.0;
    }
  }
  static __device__ __forceinline__ void mma(Acc &c, const Frag &f) {
    bf16x8 arh, arl, aih, ail, naih, nail, brh, brl, bih, bil;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      double ax = f.a[e].x, ay = f.a[e].y;
      if constexpr (LCW) ax *= f.w[e], ay *= f.w[e];
      __bf16 h, l;
      split_bf16((float)ax, h, l); arh[e] = h; arl[e] = l;
      split_bf16((float)ay, h, l); aih[e] = h; ail[e] = l;
      split_bf16(-(float)ay, h, l); naih[e] = h; nail[e] = l;
      split_bf16((float)f.b[e].x, h, l); brh[e] = h; brl[e] = l;
      split_bf16((float)f.b[e].y, h, l); bih[e] = h; bil[e] = l;
    }
    // Pr += Ar Br - Ai Bi ; Pi += Ar Bi + Ai Br
    c.r = mfma3(arh, arl, brh, brl, c.r);
    c.r = mfma3(naih, nail, bih, bil, c.r);
    c.i = mfma3(arh, arl, bih, bil, c.i);
    c.i = mfma3(aih, ail, brh, brl, c.i);
  }
  static __device__ __forceinline__ void store(const Acc &c, cx<float> *P, int lane, int s0,
                                               int rp) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 4 * (lane >> 4) + r - s0;
      if (row >= 0 && row < rp) P[row * 16 + (lane & 15)] = cx<float>{c.r[r], c.i[r]};
    }
  }
};

// One wave's share of phase 1: slots s = first, first + step, ...  The
// chunks of all its slots form one sequence whose next chunk is loaded into
// the other fragment while the current one runs on the MFMA (two fragments
// alternate, so no register copy waits on a load in flight).
template <int PREC, bool LCW>
__device__ __forceinline__ void wave_products(const UpdDesc &u, const int *scb, const int *spos,
                                              int nslot,
                                              int first, int step, gptr<const double2> gl,
                                              gptr<const double2> gc, bool rv, bool cv,
                                              int lane, int s0, int rp,
                                              cx<typename Prod<PREC, LCW>::T> *P, int E) {
  typedef Prod<PREC, LCW> Pol;
  if (first >= nslot) return;
  const int64_t chi = u.chi;
  typename Pol::Frag f0, f1;
  typename Pol::Acc acc;
  Pol::zero(acc);
  int s = first;
  Block bk = u.cblk[scb[s]];
  int k0 = bk.start;
  Pol::load(f0, gl, gc, u.LC, chi, rv, cv, k0, bk.end, lane);
  // advance (s, k0, bk) to the next chunk; false at the end of the sequence
  auto next = [&](int &ns, int &nk0, Block &nbk) {
    ns = s;
    nk0 = k0 + Pol::KC;
    nbk = bk;
    if (nk0 >= bk.end) {
      ns = s + step;
      if (ns >= nslot) return false;
      nbk = u.cblk[scb[ns]];
      nk0 = nbk.start;
    }
    return true;
  };
  auto step_with = [&](const typename Pol::Frag &cur, typename Pol::Frag &nxt) {
    int ns, nk0;
    Block nbk;
    const bool more = next(ns, nk0, nbk);
    if (more) Pol::load(nxt, gl, gc, u.LC, chi, rv, cv, nk0, nbk.end, lane);
    Pol::mma(acc, cur);
    if (!more || ns != s) {
      Pol::store(acc, P + (int64_t)spos[s] * E, lane, s0, rp);
      Pol::zero(acc);
    }
    s = ns;
    k0 = nk0;
    bk = nbk;
    return more;
  };
  while (step_with(f0, f1) && step_with(f1, f0)) {
  }
}

// ---------------------------------------------------------------- kernel --
// Dynamic LDS: the packed W table (d_max), then per item the products of
// every union-box position (nbox x E, E = rp x 16) and the output tables
// (sector element offset and row count per position, row offsets of the
// pass's left blocks and column offsets of the tile's right blocks per
// position).
template <int PREC, bool LCW>
__global__ __launch_bounds__(256) void k_contract_tile(const UpdDesc *__restrict__ Ud, WorkSeg ws,
                                                       int wtab_elems) {
  typedef typename Prod<PREC, LCW>::T T;
  extern __shared__ __align__(16) char lds_raw[];
  cx<T> *Wt = reinterpret_cast<cx<T> *>(lds_raw);
  cx<T> *P = Wt + wtab_elems;
  __shared__ int scb[256];     // non-empty union positions: centre block ...
  __shared__ int spos[256];    // ... and union-box position
  __shared__ int einfo[256];   // per element: A0 | A1 << 8 | bq0 << 16 | bq1 << 24, or -1
  __shared__ int eblk[256];    // (a - a_lo) | (b - b_lo) << 16
  __shared__ int eoff[256];    // row in left block | column in right block << 16
  __shared__ double ewt[256];  // LR[j] (LL[i]): the T weight
  __shared__ int wcount[4];
  __shared__ unsigned char pempty[256];  // union-box position without a centre block
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const double *w_u3 = nullptr;  // BS table the W table in LDS was built from
  // XCD-aware order: block b runs on XCD b % 8 (blocks are dispatched round
  // robin over the 8 XCDs and the grid is a multiple of 8), so each XCD walks
  // one contiguous eighth of the list and the tiles that share operand panels
  // share its L2; within the eighth each block takes one contiguous run of
  // items, so that its consecutive items mostly belong to one update and
  // reuse the W table already in LDS (measured: the stage 4% faster than
  // blocks striding through the eighth).
  const int2 sg = *ws.seg;
  const int2 *__restrict__ list = ws.list + sg.x;
  const int nlist = sg.y;
  const int chunk = (nlist + 7) >> 3;
  const int g8 = (int)gridDim.x >> 3, xcd = blockIdx.x & 7, bx = blockIdx.x >> 3;
  const int sub = (chunk + g8 - 1) / g8;
  for (int k = 0; k < sub; ++k) {
    const int loc = bx * sub + k;
    const int idx = xcd * chunk + loc;
    if (loc >= chunk || idx >= nlist) break;  // uniform over the block
    const int2 ref = list[idx];
    const UpdDesc &u = Ud[ref.x];
    const TileItem ti = u.titems[ref.y];
    const int d = u.d, D1 = d + 1;
    const int64_t chi = u.chi;
    // item fields are the same for the whole block: keep them in scalar registers
    const int B0u = __builtin_amdgcn_readfirstlane(ti.hi0 - ti.lo0 + 1);
    const int B1u = __builtin_amdgcn_readfirstlane(ti.hi1 - ti.lo1 + 1);
    const int nbox = B0u * B1u;
    const int E = __builtin_amdgcn_readfirstlane(ti.rp * 16);
    const int na = __builtin_amdgcn_readfirstlane(ti.na), nb = __builtin_amdgcn_readfirstlane(ti.nb);
    int64_t *ts_off = reinterpret_cast<int64_t *>(P + (int64_t)nbox * E);
    int *ts_R = reinterpret_cast<int *>(ts_off + nbox);
    int *roff = ts_R + nbox;
    int *coff = roff + nbox * na;
    __syncthreads();  // the previous item's readers of LDS are done
    if (u.U3 != w_u3) {  // W_A[t][c] = U3[A - t][t][A - c] for every A < d
      // one flat gather: every load is issued before any of them is stored
      w_u3 = u.U3;
      const gptr<const double2> U = gp(reinterpret_cast<const double2 *>(u.U3));
      constexpr int kWq = (wtab_off(kMaxD) + 255) / 256;
      const int nw = wtab_off(d);
      double2 wv[kWq];
#pragma unroll
      for (int q = 0; q < kWq; ++q) {
        const int e = threadIdx.x + 256 * q;
        wv[q] = make_double2(0.0, 0.0);
        if (e < nw) {
          int A = 0;
          while (wtab_off(A + 1) <= e) ++A;
          const int f = e - wtab_off(A), n = A + 1, t = f / n, c = f % n;
          wv[q] = gload2(U, ((A - t) * d + t) * d + (A - c));
        }
      }
#pragma unroll
      for (int q = 0; q < kWq; ++q) {
        const int e = threadIdx.x + 256 * q;
        if (e < nw) Wt[e] = cx<T>{(T)wv[q].x, (T)wv[q].y};
      }
    }
    // non-empty centre blocks of the union box, compacted
    int cb = -1;
    if ((int)threadIdx.x < nbox)
      cb = u.cblk_of[(ti.lo0 + (int)threadIdx.x / B1u) * D1 + ti.lo1 + (int)threadIdx.x % B1u];
    if ((int)threadIdx.x < nbox) pempty[threadIdx.x] = cb < 0;
    const unsigned long long m = __ballot(cb >= 0);
    const int local = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[wave] = __popcll(m);
    // per element: box, block indices, offsets, T weight
    {
      const int e = threadIdx.x;
      const int row = ti.s0 + e / 16, col = e % 16;
      int info = -1;
      if (e < E && row < ti.nr && col < ti.nc) {
        const int i = ti.r0 + row, j = ti.c0 + col;
        const int2 ri = u.row_info[i], ci = u.col_info[j];  // see SlotInfo
        const double lr = u.LR[j], ll = u.LL ? u.LL[i] : 1.0;
        if (ri.x >= 0 && ci.x >= 0) {
          const int cl0 = (ri.x >> 16) & 255, cl1 = ri.x >> 24;
          const int cr0 = (ci.x >> 16) & 255, cr1 = ci.x >> 24;
          const int A0 = cl0 - cr0, A1 = cl1 - cr1;
          if (A0 >= 0 && A1 >= 0) {
            info = A0 | A1 << 8 | (cr0 - ti.lo0) << 16 | (cr1 - ti.lo1) << 24;
            eblk[e] = ((ri.x & 0xffff) - ti.a_lo) | ((ci.x & 0xffff) - ti.b_lo) << 16;
            eoff[e] = ri.y | ci.y << 16;
            ewt[e] = lr * ll;
          }
        }
      }
      einfo[e] = info;
    }
    // output tables per union-box position (all lookups issued together)
    for (int t = threadIdx.x; t < nbox * (1 + na + nb); t += blockDim.x) {
      const int pos = t < nbox ? t : (t < nbox * (1 + na) ? (t - nbox) / na : (t - nbox * (1 + na)) / nb);
      const int sidx = u.tau_sector[(ti.lo0 + pos / B1u) * D1 + ti.lo1 + pos % B1u];
      if (t < nbox) {
        ts_off[pos] = sidx >= 0 ? u.sectors[sidx].c_off : -1;
        ts_R[pos] = sidx >= 0 ? u.sectors[sidx].R : 0;
      } else if (t < nbox * (1 + na)) {
        const int k = (t - nbox) % na;
        roff[pos * na + k] = sidx >= 0 ? u.rowoff[(int64_t)sidx * u.n_lblk + ti.a_lo + k] : -1;
      } else {
        const int k = (t - nbox * (1 + na)) % nb;
        coff[pos * nb + k] = sidx >= 0 ? u.coloff[(int64_t)sidx * u.n_rblk + ti.b_lo + k] : -1;
      }
    }
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += wcount[w];
    const int nslot = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    if (cb >= 0) {
      scb[base + local] = cb;
      spos[base + local] = threadIdx.x;
    }
    if (nslot < nbox) {  // empty positions of the union box hold P = 0 (all threads)
      for (int t = threadIdx.x; t < nbox * E; t += blockDim.x)
        if (pempty[t / E]) P[t] = cx<T>{(T)0, (T)0};
    }
    __syncthreads();

    // phase 1: one MFMA product per non-empty position, waves round robin
    {
      const int i16 = lane & 15;
      const bool rv = i16 < ti.nr, cv = i16 < ti.nc;
      const gptr<const double2> gl =
          gp(reinterpret_cast<const double2 *>(u.Glc)) + (int64_t)(ti.r0 + (rv ? i16 : 0)) * chi;
      const gptr<const double2> gc =
          gp(reinterpret_cast<const double2 *>(u.Gcr)) + (ti.c0 + (cv ? i16 : 0));
      wave_products<PREC, LCW>(u, scb, spos, nslot, wave, 4, gl, gc, rv, cv, lane, ti.s0, ti.rp,
                               P, E);
    }
    __syncthreads();

    // phase 2: Q[q0][t1] = sum_q1 conj(W_A1[t1][q1]) P[q0][q1], in place, per
    // (union row, element).  The loops run over the union box's side, which
    // is the same for the whole block, so they unroll with constant indices
    // (registers) and still stop early; positions outside an element's own
    // box enter with value 0 and clamped (in-table) weight indices.
    for (int t = threadIdx.x; t < B0u * E; t += blockDim.x) {
      const int qu0 = t / E, e = t % E;
      const int info = einfo[e];
      if (info < 0) continue;
      const int A0 = info & 255, A1 = (info >> 8) & 255, bq0 = (info >> 16) & 255,
                bq1 = info >> 24;
      if (qu0 < bq0 || qu0 > bq0 + A0) continue;
      cx<T> *row = P + (int64_t)qu0 * B1u * E + e;
      cx<T> r[kMaxD];
#pragma unroll
      for (int qu1 = 0; qu1 < kMaxD; ++qu1) {
        const int q1 = qu1 - bq1;
        r[qu1] = (qu1 < B1u && q1 >= 0 && q1 <= A1) ? row[(int64_t)qu1 * E] : cx<T>{(T)0, (T)0};
      }
      const cx<T> *W1 = Wt + wtab_off(A1);
#pragma unroll
      for (int tu1 = 0; tu1 < kMaxD; ++tu1) {
        if (tu1 < B1u) {
          const int t1 = tu1 - bq1;
          const int tc = min(max(t1, 0), A1) * (A1 + 1);
          cx<T> acc{(T)0, (T)0};
#pragma unroll
          for (int qu1 = 0; qu1 < kMaxD; ++qu1)
            if (qu1 < B1u) acc = cfmac(acc, W1[tc + min(max(qu1 - bq1, 0), A1)], r[qu1]);
          if (t1 >= 0 && t1 <= A1) row[(int64_t)tu1 * E] = acc;
        }
      }
    }
    __syncthreads();

    // phase 3: C[t0][t1] = sum_q0 W_A0[t0][q0] Q[q0][t1] per (union column,
    // element), written into sector tau = cr + (t0, t1)
    gptr<double2> Cg = gp(reinterpret_cast<double2 *>(u.C));
    gptr<double2> Tg = gp(reinterpret_cast<double2 *>(u.T));
    for (int t = threadIdx.x; t < B1u * E; t += blockDim.x) {
      const int qu1 = t / E, e = t % E;
      const int info = einfo[e];
      if (info < 0) continue;
      const int A0 = info & 255, A1 = (info >> 8) & 255, bq0 = (info >> 16) & 255,
                bq1 = info >> 24;
      if (qu1 < bq1 || qu1 > bq1 + A1) continue;
      const cx<T> *colp = P + (int64_t)qu1 * E + e;
      cx<T> c[kMaxD];
#pragma unroll
      for (int qu0 = 0; qu0 < kMaxD; ++qu0) {
        const int q0 = qu0 - bq0;
        c[qu0] = (qu0 < B0u && q0 >= 0 && q0 <= A0) ? colp[(int64_t)qu0 * B1u * E]
                                                    : cx<T>{(T)0, (T)0};
      }
      const cx<T> *W0 = Wt + wtab_off(A0);
      const int eb = eblk[e], eo = eoff[e];
      const int ka = eb & 0xffff, kb = eb >> 16, li = eo & 0xffff, rj = eo >> 16;
      const double wt = ewt[e];
#pragma unroll
      for (int tu0 = 0; tu0 < kMaxD; ++tu0) {
        const int t0 = tu0 - bq0;
        if (tu0 < B0u && t0 >= 0 && t0 <= A0) {
          cx<T> acc{(T)0, (T)0};
#pragma unroll
          for (int qu0 = 0; qu0 < kMaxD; ++qu0)
            if (qu0 < B0u) acc = cfma(acc, W0[t0 * (A0 + 1) + min(max(qu0 - bq0, 0), A0)], c[qu0]);
          const int pos = tu0 * B1u + qu1;
          const int64_t so = ts_off[pos];
          if (so >= 0) {
            const int64_t o = so + (int64_t)(roff[pos * na + ka] + li) * ts_R[pos] +
                              coff[pos * nb + kb] + rj;
            const double xr = (double)acc.x, xi = (double)acc.y;
            gstore2(Cg, o, xr, xi);
            gstore2(Tg, o, xr * wt, xi * wt);
          }
        }
      }
    }
  }
}

template <int PREC, bool LCW>
int launch_tile(const GroupHeader &g, const char *w, hipStream_t stream) {
  typedef typename Prod<PREC, LCW>::T T;
  const UpdDesc *desc = reinterpret_cast<const UpdDesc *>(w + g.o_desc);
  const int2 *list = reinterpret_cast<const int2 *>(w + g.o_tiles);
  const int2 *segs = reinterpret_cast<const GroupCounts *>(w + g.o_counts)->seg;
  const int wtab_elems = (wtab_off(g.d_max) + 1) / 2 * 2;  // keeps P 16-byte aligned
  const size_t wbytes = sizeof(cx<T>) * (size_t)wtab_elems;
  if (int rc = ensure_lds_limit((const void *)k_contract_tile<PREC, LCW>,
                                (int)(sizeof(cx<T>) * (kTileElems[kTileClasses - 1] +
                                                       (size_t)wtab_off(kMaxD) + 1))))
    return rc;
  // largest class first: the long items start early, the short ones fill in
  for (int c = kTileClasses - 1; c >= 0; --c) {
    const int n = g.tile_off[c + 1] - g.tile_off[c];
    if (n == 0) continue;
    const int grid = std::min((n + 7) / 8 * 8, 8192);
    const size_t lds = sizeof(cx<T>) * (size_t)kTileElems[c] + wbytes;
    hipLaunchKernelGGL((k_contract_tile<PREC, LCW>), dim3(grid), dim3(256), lds, stream, desc,
                       WorkSeg{list, segs + SG_TILE + c}, wtab_elems);
    BSMPO_HIP(hipGetLastError());
  }
  return 0;
}

int launch_contract(const GroupHeader &g, const char *w, int precision, hipStream_t stream) {
  if (g.n_tiles == 0) return 0;
  if (g.d_max > kMaxD) return fail(-1, "contract: d above 16");
  double flops_issued = 0.0;
  if (precision == BSMPO_BF16X3)
    flops_issued = 3.0 * 8.0 * 256.0 * g.mfma_k32;
  else
    flops_issued = 8.0 * 256.0 * g.mfma_k4;
  const int tk = timing_begin(ST_CONTRACT, stream);
  int rc;
  const bool lcw = g.has_lc != 0;
  if (precision == BSMPO_C128)
    rc = lcw ? launch_tile<BSMPO_C128, true>(g, w, stream)
             : launch_tile<BSMPO_C128, false>(g, w, stream);
  else if (precision == BSMPO_C64)
    rc = lcw ? launch_tile<BSMPO_C64, true>(g, w, stream)
             : launch_tile<BSMPO_C64, false>(g, w, stream);
  else if (precision == BSMPO_BF16X3)
    rc = lcw ? launch_tile<BSMPO_BF16X3, true>(g, w, stream)
             : launch_tile<BSMPO_BF16X3, false>(g, w, stream);
  else
    return fail(-1, "unknown precision");
  if (rc) return rc;
  (void)flops_issued;
  timing_end(tk, stream, g.flops_alg,
             (precision == BSMPO_C128 ? 1.0 : 0.5) * g.bytes_alg16);
  return 0;
}

}  // namespace bsmpo
