// Observables of the MPO over the ket = bra diagonal (src/summary.py:3-36):
// right / left environment chains, the trace ("norm") and the per-mode
// photon-number marginals P(n_j = n).
#include <hip/hip_runtime.h>

#include "common.h"

namespace bsmpo {

// mask of bond b: diagonal charge (c0 == c1 < d) and, on inner bonds, Lambda > 0
__device__ __forceinline__ bool bond_mask(int b, int x, int M, int d, int chi,
                                          const int32_t *charge, const double *Lambda) {
  const int32_t *c = charge + ((int64_t)b * chi + x) * 2;
  bool m = c[0] == c[1] && c[0] < d;
  if (b > 0 && b < M) m = m && Lambda[(int64_t)(b - 1) * chi + x] > 0.0;
  return m;
}

__global__ void k_env_init(int M, int d, int chi, const int32_t *charge, const double *Lambda,
                           double *R, double *L) {
  for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < chi; x += gridDim.x * blockDim.x) {
    cstore(R, (int64_t)M * chi + x, cplx{x == 0 ? 1.0 : 0.0, 0.0});
    cstore(L, x, cplx{bond_mask(0, x, M, d, chi, charge, Lambda) ? 1.0 : 0.0, 0.0});
  }
}

// R[b][a] = sum_c Gamma[b][a][c] R[b+1][c] mask[b+1][c]   (one wave per row)
__global__ __launch_bounds__(256) void k_renv(int b, int M, int d, int chi, const double *Gamma,
                                              const double *Lambda, const int32_t *charge,
                                              double *R) {
  const int lane = threadIdx.x & 63;
  const double2 *G = reinterpret_cast<const double2 *>(Gamma) + (int64_t)b * chi * chi;
  const double2 *Rn = reinterpret_cast<const double2 *>(R) + (int64_t)(b + 1) * chi;
  for (int a = blockIdx.x * 4 + (threadIdx.x >> 6); a < chi; a += gridDim.x * 4) {
    double sr = 0.0, si = 0.0;
    for (int c = lane; c < chi; c += 64) {
      if (!bond_mask(b + 1, c, M, d, chi, charge, Lambda)) continue;
      const double2 g = G[(int64_t)a * chi + c], r = Rn[c];
      sr += g.x * r.x - g.y * r.y;
      si += g.x * r.y + g.y * r.x;
    }
    for (int o = 32; o >= 1; o >>= 1) {
      sr += __shfl_xor(sr, o);
      si += __shfl_xor(si, o);
    }
    if (lane == 0) cstore(R, (int64_t)b * chi + a, cplx{sr, si});
  }
}

// L[b][c] = (sum_a L[b-1][a] Gamma[b-1][a][c]) mask[b][c]   (one thread per column)
__global__ void k_lenv(int b, int M, int d, int chi, const double *Gamma, const double *Lambda,
                       const int32_t *charge, double *L) {
  const double2 *G = reinterpret_cast<const double2 *>(Gamma) + (int64_t)(b - 1) * chi * chi;
  const double2 *Lp = reinterpret_cast<const double2 *>(L) + (int64_t)(b - 1) * chi;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < chi; c += gridDim.x * blockDim.x) {
    double sr = 0.0, si = 0.0;
    if (bond_mask(b, c, M, d, chi, charge, Lambda)) {
      for (int a = 0; a < chi; ++a) {
        const double2 l = Lp[a];
        if (l.x == 0.0 && l.y == 0.0) continue;
        const double2 g = G[(int64_t)a * chi + c];
        sr += l.x * g.x - l.y * g.y;
        si += l.x * g.y + l.y * g.x;
      }
    }
    cstore(L, (int64_t)b * chi + c, cplx{sr, si});
  }
}

// P[j][n] = Re sum_{a,c: cl0(a) - cr0(c) = n} L[j][a] Gamma[j][a][c] R[j+1][c] mask[j+1][c]
// One block per site; deterministic reduction (fixed wave order).
__global__ __launch_bounds__(256) void k_marginals(int M, int d, int chi, const double *Gamma,
                                                   const double *Lambda, const int32_t *charge,
                                                   const double *R, const double *L,
                                                   double *marg, double *trace) {
  const int j = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __shared__ int start[34];
  __shared__ double bins[4][32];
  const int32_t *cr = charge + (int64_t)(j + 1) * chi * 2;
  const int32_t *cl = charge + (int64_t)j * chi * 2;
  if (threadIdx.x == 0) {
    // first column of each c0 value on the right bond (charges sorted)
    int x = 0;
    for (int v = 0; v <= d; ++v) {
      while (x < chi && cr[2 * x] < v) ++x;
      start[v] = x;
    }
  }
  if (threadIdx.x < 128) bins[threadIdx.x >> 5][threadIdx.x & 31] = 0.0;
  __syncthreads();
  const double2 *G = reinterpret_cast<const double2 *>(Gamma) + (int64_t)j * chi * chi;
  const double2 *Lj = reinterpret_cast<const double2 *>(L) + (int64_t)j * chi;
  const double2 *Rn = reinterpret_cast<const double2 *>(R) + (int64_t)(j + 1) * chi;
  double mybin = 0.0;  // lane n holds bin n
  for (int a = wave; a < chi; a += 4) {
    const double2 l = Lj[a];
    if (l.x == 0.0 && l.y == 0.0) continue;
    const int c0 = cl[2 * a];
    if (c0 >= d) continue;
    for (int v = 0; v <= c0; ++v) {
      double sr = 0.0, si = 0.0;
      for (int c = start[v] + lane; c < start[v + 1]; c += 64) {
        if (!bond_mask(j + 1, c, M, d, chi, charge, Lambda)) continue;
        const double2 g = G[(int64_t)a * chi + c], r = Rn[c];
        sr += g.x * r.x - g.y * r.y;
        si += g.x * r.y + g.y * r.x;
      }
      for (int o = 32; o >= 1; o >>= 1) {
        sr += __shfl_xor(sr, o);
        si += __shfl_xor(si, o);
      }
      if (lane == c0 - v) mybin += l.x * sr - l.y * si;
    }
  }
  if (lane < 32) bins[wave][lane] = mybin;
  __syncthreads();
  if (threadIdx.x < d) {
    const int n = threadIdx.x;
    marg[(int64_t)j * d + n] = ((bins[0][n] + bins[1][n]) + bins[2][n]) + bins[3][n];
  }
  if (j == 0 && threadIdx.x == 0) {
    const double2 *R0 = reinterpret_cast<const double2 *>(R);
    double t = 0.0;
    for (int a = 0; a < chi; ++a)
      if (bond_mask(0, a, M, d, chi, charge, Lambda)) t += R0[a].x;
    *trace = t;
  }
}

// Entanglement entropy of every inner bond (src/summary.py:3-8,28-30):
// S_b = -sum_x p_x log2 p_x with p = Lambda_b^2 / sum_x Lambda_b^2; terms with
// p = 0 add nothing and an all-zero bond gives 0 (the reference's nansum over
// 0 log 0 and 0 / 0 terms).  One block per bond.
__global__ __launch_bounds__(256) void k_entropy(int chi, const double *Lambda, double *S) {
  __shared__ double red[4];
  const double *lam = Lambda + (int64_t)blockIdx.x * chi;
  auto block_sum = [&](double v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
  };
  double t = 0.0;
  for (int x = threadIdx.x; x < chi; x += blockDim.x) t += lam[x] * lam[x];
  const double tot = block_sum(t);
  double h = 0.0;
  if (tot > 0.0)
    for (int x = threadIdx.x; x < chi; x += blockDim.x) {
      const double p = lam[x] * lam[x] / tot;
      if (p > 0.0) h -= p * log2(p);
    }
  const double s = block_sum(h);
  if (threadIdx.x == 0) S[blockIdx.x] = s;
}

int launch_entropy(int M, int chi, const double *Lambda, double *S, hipStream_t stream) {
  if (M < 2 || chi < 1) return fail(-1, "entanglement_entropy: bad sizes");
  hipLaunchKernelGGL(k_entropy, dim3(M - 1), dim3(256), 0, stream, chi, Lambda, S);
  BSMPO_HIP(hipGetLastError());
  return 0;
}

int launch_observables(int M, int d, int chi, const double *Gamma, const double *Lambda,
                       const int32_t *charge, double *R, double *L, double *marg, double *trace,
                       hipStream_t stream) {
  if (M < 2 || d < 1 || d > 32 || chi < 1) return fail(-1, "observables: bad sizes");
  const int gx = (chi + 255) / 256;
  const int tk = timing_begin(ST_OBSERVABLES, stream);
  hipLaunchKernelGGL(k_env_init, dim3(gx), dim3(256), 0, stream, M, d, chi, charge, Lambda, R, L);
  BSMPO_HIP(hipGetLastError());
  for (int b = M - 1; b >= 0; --b) {
    hipLaunchKernelGGL(k_renv, dim3((chi + 3) / 4), dim3(256), 0, stream, b, M, d, chi, Gamma,
                       Lambda, charge, R);
    BSMPO_HIP(hipGetLastError());
  }
  for (int b = 1; b <= M; ++b) {
    hipLaunchKernelGGL(k_lenv, dim3(gx), dim3(256), 0, stream, b, M, d, chi, Gamma, Lambda,
                       charge, L);
    BSMPO_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_marginals, dim3(M), dim3(256), 0, stream, M, d, chi, Gamma, Lambda, charge,
                     (const double *)R, (const double *)L, marg, trace);
  BSMPO_HIP(hipGetLastError());
  timing_end(tk, stream, 0.0, 3.0 * 16.0 * M * (double)chi * chi);
  return 0;
}

}  // namespace bsmpo
