// Back-transformation of the tridiagonal route's kept eigenvectors on the
// FP64 MFMA (the W = V^H C input of src/beam_splitter_parallel/Rank.py:122-124
// needs the Gram matrix's eigenvectors; the route reduces G = Q Tri Q^H).
//
// The Householder reduction leaves reflectors H_i = I - tau_i v_i v_i^H,
// i = 0 .. n-1-off, with v_i zero above row i + off, one at row i + off and
// the rest in column i of the reduced Gram matrix: off = 1 after the one-stage
// reduction (the reflectors sit below the subdiagonal), off = kBand after the
// band reduction of the two-stage one (below the band; the bulge-chasing
// reflectors are applied to the vectors first).  The kept eigenvectors of G
// are x = H_0 H_1 ... y for the eigenvectors y of the reduced matrix.
// Reflectors are taken 32 at a time (block m holds
// i = 32 m .. 32 m + 31) in the compact WY form
//   H_i0 ... H_i0+31 = I - V T V^H,  T upper triangular,
//   T[j][j] = tau_j,  T[0:j, j] = -tau_j T[0:j, 0:j] (V[:, 0:j]^H v_j)
// (Schreiber and Van Loan, SIAM J. Sci. Stat. Comput. 10 (1989) 53), and the
// blocks are applied last to first: Z <- Z - V (T (V^H Z)).
//   k_wy_tmat    one workgroup per (sector, block): S = V^H V, then T
//   k_wy_apply   one workgroup per (sector, panel of 32 kept vectors): all
//                blocks, each as three 16 x 16-tiled MFMA products over
//                32-row chunks of V and Z staged in LDS with coalesced loads
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"

namespace bsmpo {

constexpr int kWy = 32;       // reflectors per block, vectors per panel
constexpr int kWyRows = 32;   // rows per staged chunk
constexpr int kWyLd = 33;     // LDS row stride (complex): column reads hit distinct banks

struct WyItem {
  int u, sector, idx;  // idx: block (k_wy_tmat) or panel (k_wy_apply)
};
__device__ __forceinline__ WyItem wy_item(int2 r) { return WyItem{r.x, r.y & 0xfffff, r.y >> 20}; }

// V[r][p] of block i0, rows rb .. rb + 31, into Vs[r - rb][p] (coalesced:
// eight threads read 32 consecutive rows of one reflector column).  Reflector
// i has its unit entry at row i + off (off = 1 after the one-stage reduction,
// kBand after the two-stage one) and the rest below it in column i of A.
// A chunk lives in four registers per thread between its fetch and its
// stash, so that the next chunk's loads are in flight while the current one is
// multiplied.
struct WyChunk {
  double2 x[4];
};
__device__ __forceinline__ void wy_fetch_v(WyChunk &c, gptr<const double2> A, int n, int i0, int rb,
                                           int off) {
  const int p = threadIdx.x >> 3, r0 = (threadIdx.x & 7) * 4;
  const int i = i0 + p;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int r = rb + r0 + e;
    c.x[e] = make_double2(0.0, 0.0);
    if (i <= n - 1 - off && r < n && r >= i + off)
      c.x[e] = r == i + off ? make_double2(1.0, 0.0) : gload2(A, (int64_t)i * n + r);
  }
}
__device__ __forceinline__ void wy_stash(double2 *S, const WyChunk &c) {
  const int p = threadIdx.x >> 3, r0 = (threadIdx.x & 7) * 4;
#pragma unroll
  for (int e = 0; e < 4; ++e) S[(r0 + e) * kWyLd + p] = c.x[e];
}

__device__ __forceinline__ int wy_offset(const UpdDesc &u, const Sector &s) {
  return two_stage(s.n, u.ts_min) ? kBand : 1;
}

// Z rows rb .. rb + 31 of vectors j0 .. j0 + kp - 1 (column-major, length n)
__device__ __forceinline__ void wy_fetch_z(WyChunk &c, gptr<const double2> X, int n, int j0, int kp,
                                           int rb) {
  const int j = threadIdx.x >> 3, r0 = (threadIdx.x & 7) * 4;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int r = rb + r0 + e;
    c.x[e] = (j < kp && r < n) ? gload2(X, (int64_t)(j0 + j) * n + r) : make_double2(0.0, 0.0);
  }
}

__device__ __forceinline__ void wy_store_z(const double2 *Zs, gptr<double2> X, int n, int j0,
                                           int kp, int rb) {
  const int j = threadIdx.x >> 3, r0 = (threadIdx.x & 7) * 4;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int r = rb + r0 + e;
    if (j < kp && r < n) {
      const double2 z = Zs[(r0 + e) * kWyLd + j];
      gstore2(X, (int64_t)(j0 + j) * n + r, z.x, z.y);
    }
  }
}

// acc += (Vs^H Zs) tile: rows tp*16 .. of the reflector index, columns tj*16 ..
__device__ __forceinline__ void wy_vhz(dbl4 &ar, dbl4 &ai, const double2 *Vs, const double2 *Zs,
                                       int tp, int tj, int lane) {
  const int i16 = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int s = 0; s < kWyRows / 4; ++s) {
    const int r = 4 * s + kq;
    const double2 v = Vs[r * kWyLd + tp * 16 + i16];
    const double2 z = Zs[r * kWyLd + tj * 16 + i16];
    cmfma_f64(ar, ai, cplx{v.x, -v.y}, cplx{z.x, z.y});
  }
}

// acc = L (32 x 32, rows tl*16 .., LDS row-major) times R (32 x 32, cols tj*16 ..)
__device__ __forceinline__ void wy_mm(dbl4 &ar, dbl4 &ai, const double2 *L, const double2 *R,
                                      int tl, int tj, int lane) {
  const int i16 = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int s = 0; s < kWy / 4; ++s) {
    const int q = 4 * s + kq;
    const double2 a = L[(tl * 16 + i16) * kWyLd + q];
    const double2 b = R[q * kWyLd + tj * 16 + i16];
    cmfma_f64(ar, ai, cplx{a.x, a.y}, cplx{b.x, b.y});
  }
}

__device__ __forceinline__ void wy_put(double2 *M, const dbl4 &ar, const dbl4 &ai, int tr, int tc,
                                       int lane) {
#pragma unroll
  for (int r = 0; r < 4; ++r)
    M[(tr * 16 + mfma64_row(lane >> 4, r)) * kWyLd + tc * 16 + (lane & 15)] =
        make_double2(ar[r], ai[r]);
}

__device__ __forceinline__ gptr<double2> wy_tstore(const UpdDesc &u, const Sector &s) {
  return gp(reinterpret_cast<double2 *>(u.triwork + s.tw_off));
}

// T of block `idx` of a sector, written to its work area (the inverse
// iteration's arrays there are dead by now): blocks after one another,
// 32 x 32 row-major each.
__global__ __launch_bounds__(256) void k_wy_tmat(const UpdDesc *__restrict__ Ud, WorkSeg ws) {
  __shared__ double2 Vs[kWyRows * kWyLd];
  __shared__ double2 Ss[kWy * kWyLd];
  __shared__ double2 Ts[kWy * kWyLd];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int2 sg = *ws.seg;
  const int2 *__restrict__ list = ws.list + sg.x;
  const int nlist = sg.y;
  for (int it = blockIdx.x; it < nlist; it += gridDim.x) {
    const WyItem w = wy_item(list[it]);
    const UpdDesc &u = Ud[w.u];
    const Sector s = u.sectors[w.sector];
    if (u.kcount[w.sector] == 0) continue;  // uniform: nothing kept, nothing to transform
    const int n = s.n, i0 = w.idx * kWy, off = wy_offset(u, s);
    const gptr<const double2> A = gp(reinterpret_cast<const double2 *>(u.G) + s.g_off);
    const gptr<const double2> tau = gp(reinterpret_cast<const double2 *>(u.tau) + s.s_off);
    // S = V^H V over rows i0 + off .. n - 1, one 16 x 16 tile per wave
    const int tp = wave >> 1, tj = wave & 1;
    dbl4 ar = {0.0, 0.0, 0.0, 0.0}, ai = {0.0, 0.0, 0.0, 0.0};
    WyChunk cv;
    wy_fetch_v(cv, A, n, i0, i0 + off, off);
    for (int rb = i0 + off; rb < n; rb += kWyRows) {
      __syncthreads();
      wy_stash(Vs, cv);
      __syncthreads();
      if (rb + kWyRows < n) wy_fetch_v(cv, A, n, i0, rb + kWyRows, off);
      wy_vhz(ar, ai, Vs, Vs, tp, tj, lane);
    }
    wy_put(Ss, ar, ai, tp, tj, lane);
    for (int e = threadIdx.x; e < kWy * kWyLd; e += blockDim.x) Ts[e] = make_double2(0.0, 0.0);
    __syncthreads();
    // forward recurrence, column by column
    for (int j = 0; j < kWy; ++j) {
      const double2 tj_ = i0 + j <= n - 1 - off ? gload2(tau, i0 + j) : make_double2(0.0, 0.0);
      const int t = threadIdx.x;
      if (t < j) {
        double sr = 0.0, si = 0.0;
        for (int l = t; l < j; ++l) {
          const double2 a = Ts[t * kWyLd + l], b = Ss[l * kWyLd + j];
          sr += a.x * b.x - a.y * b.y;
          si += a.x * b.y + a.y * b.x;
        }
        Ts[t * kWyLd + j] = make_double2(-(tj_.x * sr - tj_.y * si), -(tj_.x * si + tj_.y * sr));
      } else if (t == j) {
        Ts[j * kWyLd + j] = tj_;
      }
      __syncthreads();
    }
    gptr<double2> Tw = wy_tstore(u, s) + (int64_t)w.idx * kWy * kWy;
    for (int e = threadIdx.x; e < kWy * kWy; e += blockDim.x) {
      const double2 v = Ts[(e / kWy) * kWyLd + e % kWy];
      gstore2(Tw, e, v.x, v.y);
    }
    __syncthreads();
  }
}

// Z <- B_0 B_1 ... B_last Z for the panel of vectors j0 .. j0 + kp - 1.
__global__ __launch_bounds__(256) void k_wy_apply(const UpdDesc *__restrict__ Ud, WorkSeg ws) {
  __shared__ double2 Vs[kWyRows * kWyLd];
  __shared__ double2 Zs[kWyRows * kWyLd];
  __shared__ double2 Ts[kWy * kWyLd];
  __shared__ double2 Ws[kWy * kWyLd];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ta = wave >> 1, tb = wave & 1;  // this wave's 16 x 16 tile of a 32 x 32 product
  const int2 sg = *ws.seg;
  const int2 *__restrict__ list = ws.list + sg.x;
  const int nlist = sg.y;
  for (int it = blockIdx.x; it < nlist; it += gridDim.x) {
    const WyItem w = wy_item(list[it]);
    const UpdDesc &u = Ud[w.u];
    const Sector s = u.sectors[w.sector];
    const int k = u.kcount[w.sector], j0 = w.idx * kWy;
    if (j0 >= k) continue;  // uniform: a panel past the kept vectors
    const int n = s.n, kp = min(kWy, k - j0);
    gptr<double2> X = gp(s.left_small ? reinterpret_cast<double2 *>(u.V) + s.v_off
                                      : reinterpret_cast<double2 *>(u.W) + s.w_off);
    const gptr<const double2> A = gp(reinterpret_cast<const double2 *>(u.G) + s.g_off);
    const gptr<const double2> Tw = wy_tstore(u, s);
    const int off = wy_offset(u, s);
    const int nblk = (n - off + kWy - 1) / kWy;  // reflectors 0 .. n - 1 - off
    for (int m = nblk - 1; m >= 0; --m) {
      const int i0 = m * kWy;
      __syncthreads();  // the previous block's readers of Ts / Ws are done
      for (int e = threadIdx.x; e < kWy * kWy; e += blockDim.x)
        Ts[(e / kWy) * kWyLd + e % kWy] = gload2(Tw, (int64_t)m * kWy * kWy + e);
      // (1) W = V^H Z over rows i0 + off .. n - 1
      dbl4 ar = {0.0, 0.0, 0.0, 0.0}, ai = {0.0, 0.0, 0.0, 0.0};
      WyChunk cv, cz;
      wy_fetch_v(cv, A, n, i0, i0 + off, off);
      wy_fetch_z(cz, X, n, j0, kp, i0 + off);
      for (int rb = i0 + off; rb < n; rb += kWyRows) {
        __syncthreads();
        wy_stash(Vs, cv);
        wy_stash(Zs, cz);
        __syncthreads();
        if (rb + kWyRows < n) {
          wy_fetch_v(cv, A, n, i0, rb + kWyRows, off);
          wy_fetch_z(cz, X, n, j0, kp, rb + kWyRows);
        }
        wy_vhz(ar, ai, Vs, Zs, ta, tb, lane);
      }
      wy_put(Ws, ar, ai, ta, tb, lane);
      __syncthreads();
      // (2) W <- T W
      ar = dbl4{0.0, 0.0, 0.0, 0.0};
      ai = dbl4{0.0, 0.0, 0.0, 0.0};
      wy_mm(ar, ai, Ts, Ws, ta, tb, lane);
      __syncthreads();
      wy_put(Ws, ar, ai, ta, tb, lane);
      // (3) Z <- Z - V W, chunk by chunk: staged, updated in LDS, written back
      wy_fetch_v(cv, A, n, i0, i0 + off, off);
      wy_fetch_z(cz, X, n, j0, kp, i0 + off);
      for (int rb = i0 + off; rb < n; rb += kWyRows) {
        __syncthreads();
        wy_stash(Vs, cv);
        wy_stash(Zs, cz);
        __syncthreads();
        if (rb + kWyRows < n) {
          wy_fetch_v(cv, A, n, i0, rb + kWyRows, off);
          wy_fetch_z(cz, X, n, j0, kp, rb + kWyRows);
        }
        ar = dbl4{0.0, 0.0, 0.0, 0.0};
        ai = dbl4{0.0, 0.0, 0.0, 0.0};
        wy_mm(ar, ai, Vs, Ws, ta, tb, lane);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double2 &z = Zs[(ta * 16 + mfma64_row(lane >> 4, r)) * kWyLd + tb * 16 + (lane & 15)];
          z.x -= ar[r];
          z.y -= ai[r];
        }
        __syncthreads();
        wy_store_z(Zs, X, n, j0, kp, rb);
      }
    }
  }
}

int launch_wy_backtransform(const GroupHeader &g, const char *w, hipStream_t stream) {
  if (g.n_wyt == 0) return 0;
  const UpdDesc *desc = reinterpret_cast<const UpdDesc *>(w + g.o_desc);
  const int2 *segs = reinterpret_cast<const GroupCounts *>(w + g.o_counts)->seg;
  hipLaunchKernelGGL(k_wy_tmat, dim3(std::min(g.n_wyt, 16384)), dim3(256), 0, stream, desc,
                     WorkSeg{reinterpret_cast<const int2 *>(w + g.o_wyt), segs + SG_WYT});
  BSMPO_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_wy_apply, dim3(std::min(g.n_wya, 16384)), dim3(256), 0, stream, desc,
                     WorkSeg{reinterpret_cast<const int2 *>(w + g.o_wya), segs + SG_WYA});
  BSMPO_HIP(hipGetLastError());
  return 0;
}

}  // namespace bsmpo
