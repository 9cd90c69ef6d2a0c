// Pair formulation of the sector contraction (src/update_MPO_kernel.cu:193-734
// semantics, src/beam_splitter_parallel/Rank.py:114-118 for T = C diag(LR)).
//
// One workgroup per PairItem: a tl x tr tile of (left block a, right block b).
// With the box cr <= q <= cl of the pair,
//   P[q][e]   = sum_{c in centre block q} Glc[l_e][c] (LC[c]) Gcr[c][r_e]
//   Q[q0][t1] = sum_q1 conj(beta[t1][q1]) P[q0][q1]
//   C[t0][t1] = sum_q0 alpha[t0][q0] Q[q0][t1]   -> C_tau, tau = cr + (t0, t1)
// with alpha[t][c] = U3[cl0 - cr0 - t][t][cl0 - cr0 - c] (component 0) and beta
// the same on component 1, so that alpha conj(beta) is the weight
// u(cl, cc, cr; tau) = U3[cl0-t0, t0-cr0, cl0-cc0] conj(U3[cl1-t1, t1-cr1, cl1-cc1]).
// Every product Glc Gcr of the update is formed once for all tau; the tau
// dependence costs B0 B1 (B0 + B1) multiply-adds per element.  The charge
// blocks are a few rows each, so the work runs on the vector ALUs (FP64 for
// complex128, FP32 for complex64) rather than on padded MFMA tiles.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"

namespace bsmpo {

template <typename T>
struct cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cxmul(cx<T> a, cx<T> b) {
  return cx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cxfma(cx<T> acc, cx<T> a, cx<T> b) {
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
  return acc;
}

constexpr int kPairRows = 4;  // centre rows per batch of loads in phase 1

template <typename T>
__global__ __launch_bounds__(256) void k_contract_pair(const UpdDesc *__restrict__ Ud,
                                                       const int2 *__restrict__ list, int nlist) {
  __shared__ cx<T> P[kPairLds];
  __shared__ cx<T> al[kPairMaxB * kPairMaxB];  // alpha[t0][c0]
  __shared__ cx<T> bc_[kPairMaxB * kPairMaxB]; // conj(beta[t1][c1])
  for (int it = blockIdx.x; it < nlist; it += gridDim.x) {
    const int2 ref = list[it];
    const UpdDesc &u = Ud[ref.x];
    const PairItem pi = u.pitems[ref.y];
    const Block la = u.lblk[pi.a], rb = u.rblk[pi.b];
    const int d = u.d, chi = u.chi, D1 = d + 1;
    const int A0 = la.c0 - rb.c0, A1 = la.c1 - rb.c1;  // box sides minus one
    const int B0 = A0 + 1, B1 = A1 + 1;
    const int tr = pi.tr, E = pi.tl * pi.tr;
    const double2 *U = reinterpret_cast<const double2 *>(u.U3);
    for (int e = threadIdx.x; e < B0 * B0; e += blockDim.x) {
      const int t = e / B0, c = e % B0;
      const double2 v = U[((A0 - t) * d + t) * d + (A0 - c)];
      al[t * kPairMaxB + c] = cx<T>{(T)v.x, (T)v.y};
    }
    for (int e = threadIdx.x; e < B1 * B1; e += blockDim.x) {
      const int t = e / B1, c = e % B1;
      const double2 v = U[((A1 - t) * d + t) * d + (A1 - c)];
      bc_[t * kPairMaxB + c] = cx<T>{(T)v.x, (T)-v.y};
    }
    // (1) P[q][e] over the centre blocks of the box; the rows of a block are
    //     taken kPairRows at a time with their loads issued together
    const gptr<const double2> G1 = gp(reinterpret_cast<const double2 *>(u.Glc));
    const gptr<const double2> G2 = gp(reinterpret_cast<const double2 *>(u.Gcr));
    for (int idx = threadIdx.x; idx < B0 * B1 * E; idx += blockDim.x) {
      const int q = idx / E, e = idx % E;
      const int q0 = q / B1, q1 = q % B1;
      const int cb = u.cblk_of[(rb.c0 + q0) * D1 + rb.c1 + q1];
      cx<T> acc{(T)0, (T)0};
      if (cb >= 0) {
        const Block bk = u.cblk[cb];
        const int l = la.start + pi.l0 + e / tr, r = rb.start + pi.r0 + e % tr;
        const gptr<const double2> gl = G1 + (int64_t)l * chi;
        for (int c0 = bk.start; c0 < bk.end; c0 += kPairRows) {
          double2 x[kPairRows], y[kPairRows];
#pragma unroll
          for (int k = 0; k < kPairRows; ++k) {
            const int c = min(c0 + k, bk.end - 1);
            x[k] = gload2(gl, c);
            y[k] = gload2(G2, (int64_t)c * chi + r);
          }
#pragma unroll
          for (int k = 0; k < kPairRows; ++k) {
            if (c0 + k >= bk.end) break;
            if (u.LC) {
              const double w = u.LC[c0 + k];
              x[k].x *= w;
              x[k].y *= w;
            }
            acc = cxfma(acc, cx<T>{(T)x[k].x, (T)x[k].y}, cx<T>{(T)y[k].x, (T)y[k].y});
          }
        }
      }
      P[idx] = acc;
    }
    __syncthreads();
    // (2) Q[q0][t1][e] = sum_q1 conj(beta[t1][q1]) P[q0][q1][e], in place per (q0, e) row
    for (int idx = threadIdx.x; idx < B0 * E; idx += blockDim.x) {
      const int q0 = idx / E, e = idx % E;
      cx<T> row[kPairMaxB];
#pragma unroll
      for (int q1 = 0; q1 < kPairMaxB; ++q1)
        row[q1] = q1 < B1 ? P[(q0 * B1 + q1) * E + e] : cx<T>{(T)0, (T)0};
#pragma unroll
      for (int t1 = 0; t1 < kPairMaxB; ++t1) {
        if (t1 >= B1) break;
        cx<T> acc{(T)0, (T)0};
#pragma unroll
        for (int q1 = 0; q1 < kPairMaxB; ++q1)
          if (q1 < B1) acc = cxfma(acc, bc_[t1 * kPairMaxB + q1], row[q1]);
        P[(q0 * B1 + t1) * E + e] = acc;
      }
    }
    __syncthreads();
    // (3) C[t0][t1][e] = sum_q0 alpha[t0][q0] Q[q0][t1][e] into every sector tau
    for (int idx = threadIdx.x; idx < B1 * E; idx += blockDim.x) {
      const int t1 = idx / E, e = idx % E;
      const int lr = pi.l0 + e / tr, rr = pi.r0 + e % tr;
      const int grow = la.start + lr, gcol = rb.start + rr;
      double wt = u.LR[gcol];
      if (u.LL) wt *= u.LL[grow];
      for (int t0 = 0; t0 < B0; ++t0) {
        const int s = u.tau_sector[(rb.c0 + t0) * D1 + rb.c1 + t1];
        if (s < 0) continue;
        cx<T> acc{(T)0, (T)0};
        for (int q0 = 0; q0 < B0; ++q0)
          acc = cxfma(acc, al[t0 * kPairMaxB + q0], P[(q0 * B1 + t1) * E + e]);
        const Sector &sc = u.sectors[s];
        const int row = u.rowoff[(int64_t)s * u.n_lblk + pi.a] + lr;
        const int col = u.coloff[(int64_t)s * u.n_rblk + pi.b] + rr;
        const int64_t o = sc.c_off + (int64_t)row * sc.R + col;
        const cplx c{(double)acc.x, (double)acc.y};
        cstore(u.C, o, c);
        cstore(u.T, o, cplx{c.x * wt, c.y * wt});
      }
    }
    __syncthreads();
  }
}

int launch_contract_pair(const GroupHeader &g, const char *w, int precision, hipStream_t stream) {
  if (g.n_pitems == 0) return 0;
  const int grid = std::min(g.n_pitems, 16384);
  const UpdDesc *desc = reinterpret_cast<const UpdDesc *>(w + g.o_desc);
  const int2 *list = reinterpret_cast<const int2 *>(w + g.o_pitems);
  if (precision == BSMPO_C128)
    hipLaunchKernelGGL(k_contract_pair<double>, dim3(grid), dim3(256), 0, stream, desc, list,
                       g.n_pitems);
  else
    hipLaunchKernelGGL(k_contract_pair<float>, dim3(grid), dim3(256), 0, stream, desc, list,
                       g.n_pitems);
  BSMPO_HIP(hipGetLastError());
  return 0;
}

}  // namespace bsmpo
