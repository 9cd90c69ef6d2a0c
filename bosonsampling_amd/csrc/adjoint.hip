// Reverse-mode (adjoint) kernels: the site update (sector contraction, the
// Gram-route SVD with truncation, W = V^H C), the BS table, and the output
// observables (-log P(pattern) + sum_j w_j <n_j>).
//
// Gradient convention: for a real loss L and complex z, g = dL/dRe z + i
// dL/dIm z, so dL = Re sum conj(g) dz.  For y = A x: g_x = A^H g_y and
// g_A = g_y x^H.
//
// The per-sector backward runs one workgroup per sector with the dense
// algebra as workgroup loops in FP64 (sector sizes of the differentiated
// workloads are ~100); the sector Gram is re-diagonalised by one-sided
// Jacobi in global memory so that every non-dropped eigenpair is available.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "common.h"
#include "adjoint.h"

namespace bsmpo {

namespace {

constexpr int kVjpThreads = 256;
constexpr int kVjpSweeps = 60;
constexpr double kVjpDegen = 1e-13;   // |lambda_i - lambda_j| <= 1e-13 lambda_max: no coupling

__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return cplx{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return cplx{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return cplx{a.x * s, a.y * s}; }

// C[i, j] = beta C[i, j] + alpha sum_k a(i, k) b(k, j) over one workgroup, with
// a(i, k) = A[i sai + k sak] (conjugated if CA), b(k, j) = B[k sbk + j sbj]
// (conjugated if CB), C[i sci + j scj].  Ends with a barrier.
template <bool CA, bool CB>
__device__ void wg_gemm(int m, int n, int k, const double *A, int64_t sai, int64_t sak,
                        const double *B, int64_t sbk, int64_t sbj, double *Cm, int64_t sci,
                        int64_t scj, cplx alpha, bool accumulate) {
  for (int e = threadIdx.x; e < m * n; e += blockDim.x) {
    const int i = e % m, j = e / m;
    cplx acc{0.0, 0.0};
    for (int kk = 0; kk < k; ++kk) {
      cplx a = cload(A, i * sai + kk * sak);
      cplx b = cload(B, kk * sbk + j * sbj);
      if (CA) a = cconj(a);
      if (CB) b = cconj(b);
      acc = cadd(acc, cmul(a, b));
    }
    acc = cmul(alpha, acc);
    const int64_t o = i * sci + j * scj;
    if (accumulate) acc = cadd(acc, cload(Cm, o));
    cstore(Cm, o, acc);
  }
  __syncthreads();
}

__device__ __forceinline__ void rr_pair_v(int step, int i, int n2, int &p, int &q) {
  if (i == 0) {
    p = n2 - 1;
    q = step;
  } else {
    const int m = n2 - 1;
    p = (step + i) % m;
    q = (step - i + m) % m;
  }
  if (p > q) { const int t = p; p = q; q = t; }
}

// One-sided cyclic Jacobi on the columns of X (n x n, column-major, global
// memory): afterwards X = Y diag(lambda) for the Hermitian PSD matrix it held.
__device__ void wg_jacobi(double *Xd, int n, int *any_rot) {
  double2 *X = reinterpret_cast<double2 *>(Xd);
  const int group = threadIdx.x >> 3, gl = threadIdx.x & 7;
  const int ngroups = blockDim.x >> 3;
  const int n2 = n + (n & 1);
  const double tol = fmax(1e-15, (double)n * 2.220446049250313e-16);
  for (int sweep = 0; sweep < kVjpSweeps && n > 1; ++sweep) {
    if (threadIdx.x == 0) *any_rot = 0;
    __syncthreads();
    for (int step = 0; step < n2 - 1; ++step) {
      for (int pi = group; pi < n2 / 2; pi += ngroups) {
        int p, q;
        rr_pair_v(step, pi, n2, p, q);
        if (q >= n) continue;
        double2 *xp = X + (int64_t)p * n, *xq = X + (int64_t)q * n;
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        for (int i = gl; i < n; i += 8) {
          const double2 a = xp[i], b = xq[i];
          al += a.x * a.x + a.y * a.y;
          be += b.x * b.x + b.y * b.y;
          gr += a.x * b.x + a.y * b.y;
          gi += a.x * b.y - a.y * b.x;
        }
        for (int o = 4; o >= 1; o >>= 1) {
          al += __shfl_xor(al, o);
          be += __shfl_xor(be, o);
          gr += __shfl_xor(gr, o);
          gi += __shfl_xor(gi, o);
        }
        const double ga = sqrt(gr * gr + gi * gi);
        if (ga > tol * sqrt(al * be) && ga > 0.0) {
          const double zeta = (be - al) / (2.0 * ga);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
          const double er = gr / ga, ei = gi / ga;
          for (int i = gl; i < n; i += 8) {
            const double2 a = xp[i], b = xq[i];
            double2 na, nb;
            na.x = c * a.x - sn * (er * b.x + ei * b.y);
            na.y = c * a.y - sn * (er * b.y - ei * b.x);
            nb.x = c * b.x + sn * (er * a.x - ei * a.y);
            nb.y = c * b.y + sn * (er * a.y + ei * a.x);
            xp[i] = na;
            xq[i] = nb;
          }
          if (gl == 0) *any_rot = 1;
        }
      }
      __syncthreads();
    }
    if (!*any_rot) break;
    __syncthreads();
  }
}

}  // namespace

// ============================================================ sector VJP ==
__global__ __launch_bounds__(kVjpThreads) void k_sector_vjp(VjpArgs a) {
  __shared__ int any_rot;
  __shared__ double lam_max;
  const VjpSector s = a.sec[blockIdx.x];
  const int L = s.L, R = s.R, n = s.n, k = s.k, chi = a.chi;
  const double *C = a.C + 2 * s.c_off;
  const double *T = a.T + 2 * s.c_off;
  double *gC = a.gC + 2 * s.c_off;
  const int32_t *rows = a.rows + s.row_off;
  const int32_t *cols = a.cols + s.col_off;
  if (k == 0 || n == 0) {
    for (int e = threadIdx.x; e < L * R; e += blockDim.x) cstore(gC, e, cplx{0.0, 0.0});
    return;
  }
  // scratch (complex, column-major unless noted)
  double *Gm = a.scr + 2 * s.scr;             // n x n: Gram, then Jacobi X
  double *Y = Gm + 2 * (int64_t)n * n;        // n x n: eigenvectors, descending
  double *H = Y + 2 * (int64_t)n * n;         // n x n
  double *A = H + 2 * (int64_t)n * n;         // L x k: forward V_P
  double *gA = A + 2 * (int64_t)L * k;        // L x k
  double *J = gA + 2 * (int64_t)L * k;        // n x k
  double *E = J + 2 * (int64_t)n * k;         // n x k
  double *gYP = E + 2 * (int64_t)n * k;       // n x k
  const int64_t pmax = (int64_t)max(n * R, L * n);
  double *P1 = gYP + 2 * (int64_t)n * k;      // n x R or L x n
  double *P2 = P1 + 2 * pmax;
  double *X2 = P2 + 2 * pmax;                 // k x R or L x k
  const int64_t xmax = (int64_t)max(k * R, L * k);
  double *gT = X2 + 2 * xmax;                 // L x R, row-major like C
  double *gB = gT + 2 * (int64_t)L * R;       // k x R
  double *lam = a.dscr + s.dscr;              // n sorted eigenvalues
  double *sv = lam + n;                       // k forward singular values
  double *gs = sv + k;                        // k
  double *tmp = gs + k;                       // n raw eigenvalues
  int32_t *rank = reinterpret_cast<int32_t *>(tmp + n);  // n
  const cplx one{1.0, 0.0}, zero{0.0, 0.0};

  // 1. Gram of the smaller side, column-major n x n
  if (s.left_small)
    wg_gemm<false, true>(n, n, R, T, R, 1, T, 1, R, Gm, 1, n, one, false);   // T T^H
  else
    wg_gemm<true, false>(n, n, L, T, 1, R, T, R, 1, Gm, 1, n, one, false);   // T^H T
  // 2. all eigenpairs (descending), vectors of dropped ones set to 0
  wg_jacobi(Gm, n, &any_rot);
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    double nr = 0.0;
    for (int i = 0; i < n; ++i) {
      const cplx v = cload(Gm, (int64_t)p * n + i);
      nr += v.x * v.x + v.y * v.y;
    }
    tmp[p] = sqrt(nr);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int p = 0; p < n; ++p) m = fmax(m, tmp[p]);
    lam_max = m;
  }
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const double lp = tmp[p];
    int r = 0;
    for (int q = 0; q < n; ++q) r += (tmp[q] > lp || (tmp[q] == lp && q < p)) ? 1 : 0;
    rank[p] = r;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int p = e / n, i = e % n;
    const double lp = tmp[p];
    const bool keep = lam_max > 0.0 && lp > a.drop_rel2 * lam_max;
    const cplx v = cload(Gm, e);
    cstore(Y, (int64_t)rank[p] * n + i, keep ? cscale(v, 1.0 / lp) : zero);
    if (i == 0) lam[rank[p]] = keep ? lp : 0.0;
  }
  __syncthreads();
  // 3. the forward's kept vectors, values and the incoming gradients
  for (int e = threadIdx.x; e < L * k; e += blockDim.x) {
    const int i = e % L, j = e / L;
    const int64_t g = (int64_t)rows[i] * chi + s.slot0 + j;
    cstore(A, e, cload(a.G0, g));
    cstore(gA, e, cload(a.gG0, g));
  }
  for (int e = threadIdx.x; e < k * R; e += blockDim.x) {
    const int j = e % k, r = e / k;
    cstore(gB, e, cload(a.gG1, (int64_t)(s.slot0 + j) * chi + cols[r]));  // k x R col-major
  }
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    sv[j] = a.lam_new[s.slot0 + j];
    gs[j] = a.gLam[s.slot0 + j];
  }
  __syncthreads();
  for (int j = threadIdx.x; j < k; j += blockDim.x) lam[j] = sv[j] * sv[j];
  // 4. gC = A gB ;  gA += C gB^H
  wg_gemm<false, false>(L, R, k, A, 1, L, gB, 1, k, gC, R, 1, one, false);
  wg_gemm<false, true>(L, k, R, C, R, 1, gB, k, 1, gA, 1, L, one, true);
  if (s.left_small) {
    // Y[:, :k] = A (the forward's gauge); J = Y^H gA; E = (gA - Y J) / lambda
    for (int e = threadIdx.x; e < L * k; e += blockDim.x) cstore(Y, e, cload(A, e));
    __syncthreads();
    wg_gemm<true, false>(n, k, n, Y, n, 1, gA, 1, L, J, 1, n, one, false);
    for (int e = threadIdx.x; e < L * k; e += blockDim.x) cstore(E, e, cload(gA, e));
    __syncthreads();
    wg_gemm<false, false>(n, k, n, Y, 1, n, J, 1, n, E, 1, n, cplx{-1.0, 0.0}, true);
  } else {
    // Y[:, :k] = T^H A / s ; gs -= Re diag(A^H gA) / s ; gA /= s
    wg_gemm<true, false>(n, k, L, T, 1, R, A, 1, L, Y, 1, n, one, false);
    for (int e = threadIdx.x; e < n * k; e += blockDim.x)
      cstore(Y, e, cscale(cload(Y, e), 1.0 / sv[e / n]));
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
      double re = 0.0;
      for (int i = 0; i < L; ++i) {
        const cplx x = cload(A, (int64_t)j * L + i), y = cload(gA, (int64_t)j * L + i);
        re += x.x * y.x + x.y * y.y;
      }
      gs[j] -= re / sv[j];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < L * k; e += blockDim.x)
      cstore(gA, e, cscale(cload(gA, e), 1.0 / sv[e / L]));
    __syncthreads();
    // gT = gA YP^H (row-major L x R) ; gYP = T^H gA
    wg_gemm<false, true>(L, R, k, gA, 1, L, Y, n, 1, gT, R, 1, one, false);
    wg_gemm<true, false>(n, k, L, T, 1, R, gA, 1, L, gYP, 1, n, one, false);
    wg_gemm<true, false>(n, k, n, Y, n, 1, gYP, 1, n, J, 1, n, one, false);
    for (int e = threadIdx.x; e < n * k; e += blockDim.x) cstore(E, e, cload(gYP, e));
    __syncthreads();
    wg_gemm<false, false>(n, k, n, Y, 1, n, J, 1, n, E, 1, n, cplx{-1.0, 0.0}, true);
  }
  // E /= lambda_j ; H_ij = F_ij (J~_ij - conj(J~_ji)) / 2, H_jj = gs_j / (2 s_j)
  for (int e = threadIdx.x; e < n * k; e += blockDim.x)
    cstore(E, e, cscale(cload(E, e), 1.0 / lam[e / n]));
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e % n, j = e / n;
    cplx h{0.0, 0.0};
    if (i == j) {
      if (j < k) h.x = gs[j] / (2.0 * sv[j]);
    } else {
      const double diff = lam[j] - lam[i];
      if (fabs(diff) > kVjpDegen * lam_max && (i < k || j < k)) {
        const cplx jij = j < k ? cload(J, (int64_t)j * n + i) : zero;
        const cplx jji = i < k ? cload(J, (int64_t)i * n + j) : zero;
        h = cscale(csub(jij, cconj(jji)), 0.5 / diff);
      }
    }
    cstore(H, e, h);
  }
  __syncthreads();
  if (s.left_small) {
    // gT = 2 Y H Y^H T + E (A^H T) + A (E^H T)
    wg_gemm<true, false>(n, R, L, Y, n, 1, T, R, 1, P1, 1, n, one, false);     // Y^H T
    wg_gemm<false, false>(n, R, n, H, 1, n, P1, 1, n, P2, 1, n, one, false);   // H P1
    wg_gemm<false, false>(L, R, n, Y, 1, n, P2, 1, n, gT, R, 1, cplx{2.0, 0.0}, false);
    wg_gemm<true, false>(k, R, L, A, L, 1, T, R, 1, X2, 1, k, one, false);     // A^H T
    wg_gemm<false, false>(L, R, k, E, 1, n, X2, 1, k, gT, R, 1, one, true);
    wg_gemm<true, false>(k, R, L, E, n, 1, T, R, 1, X2, 1, k, one, false);     // E^H T
    wg_gemm<false, false>(L, R, k, A, 1, L, X2, 1, k, gT, R, 1, one, true);
  } else {
    // gT += 2 T Y H Y^H + (T E) YP^H + (A s) E^H
    wg_gemm<false, false>(L, n, R, T, R, 1, Y, 1, n, P1, 1, L, one, false);    // T Y
    wg_gemm<false, false>(L, n, n, P1, 1, L, H, 1, n, P2, 1, L, one, false);   // P1 H
    wg_gemm<false, true>(L, R, n, P2, 1, L, Y, n, 1, gT, R, 1, cplx{2.0, 0.0}, true);
    wg_gemm<false, false>(L, k, R, T, R, 1, E, 1, n, X2, 1, L, one, false);    // T E
    wg_gemm<false, true>(L, R, k, X2, 1, L, Y, n, 1, gT, R, 1, one, true);
    for (int e = threadIdx.x; e < L * k; e += blockDim.x)
      cstore(X2, e, cscale(cload(A, e), sv[e / L]));
    __syncthreads();
    wg_gemm<false, true>(L, R, k, X2, 1, L, E, n, 1, gT, R, 1, one, true);
  }
  // gC += gT diag(LR) ; gLR[col] += Re sum_l conj(gT) C
  for (int e = threadIdx.x; e < L * R; e += blockDim.x) {
    const double lr = a.LR[cols[e % R]];
    cstore(gC, e, cadd(cload(gC, e), cscale(cload(gT, e), lr)));
  }
  for (int r = threadIdx.x; r < R; r += blockDim.x) {
    double acc = 0.0;
    for (int l = 0; l < L; ++l) {
      const cplx g = cload(gT, (int64_t)l * R + r), c = cload(C, (int64_t)l * R + r);
      acc += g.x * c.x + g.y * c.y;
    }
    atomicAdd(a.gLR + cols[r], acc);
  }
}

// ======================================================= contraction VJP ==
__device__ __forceinline__ bool ge2(const int32_t *x, const int32_t *y) {
  return x[0] >= y[0] && x[1] >= y[1];
}

__device__ __forceinline__ cplx weight(const VjpArgs &a, const int32_t *cl, const int32_t *cc,
                                       const int32_t *cr, int t0, int t1, cplx *x, cplx *y,
                                       int64_t *i0, int64_t *i1) {
  const int d = a.d;
  *i0 = ((int64_t)(cl[0] - t0) * d + (t0 - cr[0])) * d + (cl[0] - cc[0]);
  *i1 = ((int64_t)(cl[1] - t1) * d + (t1 - cr[1])) * d + (cl[1] - cc[1]);
  *x = cload(a.U3, *i0);
  *y = cload(a.U3, *i1);
  return cmulc(*x, *y);
}

__global__ void k_vjp_pos(VjpArgs a) {
  const int64_t tot = (int64_t)a.n_sectors * a.chi;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    a.rowpos[e] = -1;
    a.colpos[e] = -1;
  }
}

__global__ void k_vjp_pos_set(VjpArgs a) {
  const VjpSector s = a.sec[blockIdx.x];
  for (int i = threadIdx.x; i < s.L; i += blockDim.x)
    a.rowpos[(int64_t)blockIdx.x * a.chi + a.rows[s.row_off + i]] = i;
  for (int i = threadIdx.x; i < s.R; i += blockDim.x)
    a.colpos[(int64_t)blockIdx.x * a.chi + a.cols[s.col_off + i]] = i;
}

// gGlc[l][c] = sum_tau sum_r gC_tau[l, r] conj(u(cl, cc, cr; tau) Gcr[c][r])
__global__ void k_vjp_glc(VjpArgs a) {
  const int chi = a.chi, d = a.d;
  const int64_t tot = (int64_t)chi * chi;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(e / chi), c = (int)(e % chi);
    const int32_t *cl = a.CL + 2 * l, *cc = a.CC + 2 * c;
    cplx acc{0.0, 0.0};
    if (cl[0] < d && cl[1] < d && cc[0] < d && cc[1] < d && ge2(cl, cc)) {
      for (int si = 0; si < a.n_sectors; ++si) {
        const int li = a.rowpos[(int64_t)si * chi + l];
        if (li < 0) continue;
        const VjpSector s = a.sec[si];
        const double *g = a.gC + 2 * (s.c_off + (int64_t)li * s.R);
        for (int ri = 0; ri < s.R; ++ri) {
          const int r = a.cols[s.col_off + ri];
          const int32_t *cr = a.CR + 2 * r;
          if (!ge2(cc, cr)) continue;
          cplx x, y;
          int64_t i0, i1;
          const cplx u = weight(a, cl, cc, cr, s.t0, s.t1, &x, &y, &i0, &i1);
          const cplx w = cmul(u, cload(a.Gcr, (int64_t)c * chi + r));
          acc = cadd(acc, cmulc(cload(g, ri), w));
        }
      }
    }
    cstore(a.gGlc, e, acc);
  }
}

// gGcr[c][r] = sum_tau sum_l conj(Glc[l][c] u(cl, cc, cr; tau)) gC_tau[l, r]
__global__ void k_vjp_gcr(VjpArgs a) {
  const int chi = a.chi, d = a.d;
  const int64_t tot = (int64_t)chi * chi;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / chi), r = (int)(e % chi);
    const int32_t *cc = a.CC + 2 * c, *cr = a.CR + 2 * r;
    cplx acc{0.0, 0.0};
    if (cc[0] < d && cc[1] < d && cr[0] < d && cr[1] < d && ge2(cc, cr)) {
      for (int si = 0; si < a.n_sectors; ++si) {
        const int ri = a.colpos[(int64_t)si * chi + r];
        if (ri < 0) continue;
        const VjpSector s = a.sec[si];
        for (int li = 0; li < s.L; ++li) {
          const int l = a.rows[s.row_off + li];
          const int32_t *cl = a.CL + 2 * l;
          if (!ge2(cl, cc)) continue;
          cplx x, y;
          int64_t i0, i1;
          const cplx u = weight(a, cl, cc, cr, s.t0, s.t1, &x, &y, &i0, &i1);
          const cplx w = cmul(cload(a.Glc, (int64_t)l * chi + c), u);
          acc = cadd(acc, cmulc(cload(a.gC, s.c_off + (int64_t)li * s.R + ri), w));
        }
      }
    }
    cstore(a.gGcr, e, acc);
  }
}

// g_u(a, c, b; tau) = sum_{l in a, r in b} gC_tau[l, r] conj((Glc[a, c] Gcr[c, b])[l, r]),
// one workgroup per (sector, left block, right block) pair; then the product
// rule on u = U3[i0] conj(U3[i1]): gU3[i0] += g_u U3[i1], gU3[i1] += conj(g_u) U3[i0].
__global__ __launch_bounds__(256) void k_vjp_u(VjpArgs a) {
  __shared__ double red[2][256];
  const VjpPair pr = a.pairs[blockIdx.x];
  const VjpSector s = a.sec[pr.s];
  const Block la = a.lblk[pr.a], rb = a.rblk[pr.b];
  const int na = la.end - la.start, nb = rb.end - rb.start;
  const int chi = a.chi;
  const int32_t cl[2] = {la.c0, la.c1}, cr[2] = {rb.c0, rb.c1};
  for (int ci = 0; ci < a.n_cblk; ++ci) {
    const Block cb = a.cblk[ci];
    const int32_t cc[2] = {cb.c0, cb.c1};
    if (!(ge2(cl, cc) && ge2(cc, cr))) continue;
    double sr = 0.0, si = 0.0;
    for (int e = threadIdx.x; e < na * nb; e += blockDim.x) {
      const int li = e / nb, ri = e % nb;
      const int l = la.start + li, r = rb.start + ri;
      cplx p{0.0, 0.0};
      for (int c = cb.start; c < cb.end; ++c)
        p = cadd(p, cmul(cload(a.Glc, (int64_t)l * chi + c), cload(a.Gcr, (int64_t)c * chi + r)));
      const cplx g = cload(a.gC, s.c_off + (int64_t)(pr.ro + li) * s.R + pr.co + ri);
      const cplx v = cmulc(g, p);  // g conj(p)
      sr += v.x;
      si += v.y;
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        red[0][threadIdx.x] += red[0][threadIdx.x + o];
        red[1][threadIdx.x] += red[1][threadIdx.x + o];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const cplx gu{red[0][0], red[1][0]};
      cplx x, y;
      int64_t i0, i1;
      weight(a, cl, cc, cr, s.t0, s.t1, &x, &y, &i0, &i1);
      const cplx g0 = cmul(gu, y), g1 = cmul(cconj(gu), x);
      atomicAdd(a.gU3 + 2 * i0, g0.x);
      atomicAdd(a.gU3 + 2 * i0 + 1, g0.y);
      atomicAdd(a.gU3 + 2 * i1, g1.x);
      atomicAdd(a.gU3 + 2 * i1 + 1, g1.y);
    }
    __syncthreads();
  }
}

__global__ void k_vjp_zero(double *p, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    p[e] = 0.0;
}

int launch_site_vjp(const VjpArgs &a, int n_pairs, hipStream_t st) {
  const int chi = a.chi;
  const int g2 = (int)std::min<int64_t>(((int64_t)chi * chi + 255) / 256, 4096);
  hipLaunchKernelGGL(k_vjp_zero, dim3(64), dim3(256), 0, st, a.gU3,
                     (int64_t)2 * a.d * a.d * a.d);
  BSMPO_HIP(hipGetLastError());
  if (a.n_sectors > 0) {
    hipLaunchKernelGGL(k_sector_vjp, dim3(a.n_sectors), dim3(kVjpThreads), 0, st, a);
    BSMPO_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_vjp_pos, dim3(64), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_vjp_pos_set, dim3(a.n_sectors), dim3(256), 0, st, a);
    BSMPO_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_vjp_glc, dim3(g2), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_vjp_gcr, dim3(g2), dim3(256), 0, st, a);
  BSMPO_HIP(hipGetLastError());
  if (n_pairs > 0) {
    hipLaunchKernelGGL(k_vjp_u, dim3(n_pairs), dim3(256), 0, st, a);
    BSMPO_HIP(hipGetLastError());
  }
  return 0;
}

// ========================================================= BS table VJP ==
// gB[v] = sum_{n1, n2, p} gU3[n1][n2][p] conj(dU3[n1][n2][p] / dB_v), one block
// per beam splitter, v over (B00, B01, B10, B11).
__global__ __launch_bounds__(256) void k_bs_table_vjp(int d, const double *B, const double *gU3,
                                                      double *gB, BsArgs args) {
  __shared__ double red[8][256];
  const int b = blockIdx.x;
  const double *Bb = B + 8 * b;
  const cplx bv[4] = {{Bb[0], Bb[1]}, {Bb[2], Bb[3]}, {Bb[4], Bb[5]}, {Bb[6], Bb[7]}};
  const double *G = gU3 + (int64_t)b * d * d * d * 2;
  cplx acc[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  for (int e = threadIdx.x; e < d * d * d; e += blockDim.x) {
    const int n1 = e / (d * d), n2 = (e / d) % d, p = e % d;
    const int q = n1 + n2 - p;
    if (q < 0 || q >= d) continue;
    const cplx g = cload(G, e);
    const double f = (args.sqf[n1] * args.sqf[n2]) / (args.sqf[p] * args.sqf[q]);
    const int klo = max(0, n1 - q), khi = min(p, n1);
    for (int kk = klo; kk <= khi; ++kk) {
      const double c = f * binom_u64(p, kk) * binom_u64(q, n1 - kk);
      const int ex[4] = {kk, n1 - kk, p - kk, q - n1 + kk};  // powers of B00, B01, B10, B11
      for (int v = 0; v < 4; ++v) {
        if (ex[v] == 0) continue;
        cplx t{c * ex[v], 0.0};
        for (int w = 0; w < 4; ++w) t = cmul(t, cpow_int(bv[w], w == v ? ex[w] - 1 : ex[w]));
        acc[v] = cadd(acc[v], cmulc(g, t));
      }
    }
  }
  for (int v = 0; v < 4; ++v) {
    red[2 * v][threadIdx.x] = acc[v].x;
    red[2 * v + 1][threadIdx.x] = acc[v].y;
  }
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int v = 0; v < 8; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 8) gB[8 * b + threadIdx.x] = red[threadIdx.x][0];
}

int launch_bs_table_vjp(int d, int nb, const double *d_B, const double *d_gU3, double *d_gB,
                        hipStream_t st) {
  if (d < 1 || d > 32 || nb < 0) return fail(-1, "bs_table_vjp: bad d or nb");
  if (nb == 0) return 0;
  hipLaunchKernelGGL(k_bs_table_vjp, dim3(nb), dim3(256), 0, st, d, d_B, d_gU3, d_gB, bs_args());
  BSMPO_HIP(hipGetLastError());
  return 0;
}

// ======================================================= observables VJP ==
__device__ __forceinline__ bool obs_mask(int b, int x, int M, int d, int chi,
                                         const int32_t *charge, const double *Lambda) {
  const int32_t *c = charge + ((int64_t)b * chi + x) * 2;
  bool m = c[0] == c[1] && c[0] < d;
  if (b > 0 && b < M) m = m && Lambda[(int64_t)(b - 1) * chi + x] > 0.0;
  return m;
}

// Environments of the three chains (pattern, plain, weighted occupation), left
// and right, each [(M+1) x chi] complex: env + ((kind * (M+1) + b) * chi + x).
// kind 0 lp, 1 l0, 2 lw, 3 rp, 4 r0, 5 rw.  out = {loss, P, sum w <n>}.
__global__ __launch_bounds__(1024) void k_obs_envs(int M, int d, int chi, const double *Gamma,
                                                   const double *Lambda, const int32_t *charge,
                                                   const int32_t *pattern, const double *w,
                                                   double *env, double *out) {
  auto E = [&](int kind, int b) { return env + 2 * (((int64_t)kind * (M + 1) + b) * chi); };
  for (int x = threadIdx.x; x < chi; x += blockDim.x) {
    const double m0 = obs_mask(0, x, M, d, chi, charge, Lambda) ? 1.0 : 0.0;
    cstore(E(0, 0), x, cplx{m0, 0.0});
    cstore(E(1, 0), x, cplx{m0, 0.0});
    cstore(E(2, 0), x, cplx{0.0, 0.0});
    const double mM = (x == 0 && obs_mask(M, x, M, d, chi, charge, Lambda)) ? 1.0 : 0.0;
    cstore(E(3, M), x, cplx{mM, 0.0});
    cstore(E(4, M), x, cplx{mM, 0.0});
    cstore(E(5, M), x, cplx{0.0, 0.0});
  }
  __syncthreads();
  for (int b = 0; b < M; ++b) {  // left: v[b+1][j] = (sum_i v[b][i] G[i][j]) mask[b+1][j]
    const double2 *G = reinterpret_cast<const double2 *>(Gamma) + (int64_t)b * chi * chi;
    const int32_t *cb = charge + (int64_t)b * chi * 2, *cn = charge + (int64_t)(b + 1) * chi * 2;
    for (int j = threadIdx.x; j < chi; j += blockDim.x) {
      cplx p{0.0, 0.0}, z{0.0, 0.0}, q{0.0, 0.0};
      if (obs_mask(b + 1, j, M, d, chi, charge, Lambda)) {
        for (int i = 0; i < chi; ++i) {
          const cplx l0 = cload(E(1, b), i), lw = cload(E(2, b), i), lp = cload(E(0, b), i);
          if (l0.x == 0.0 && l0.y == 0.0 && lw.x == 0.0 && lw.y == 0.0 && lp.x == 0.0 &&
              lp.y == 0.0)
            continue;
          const double2 gv = G[(int64_t)i * chi + j];
          const cplx g{gv.x, gv.y};
          const int occ = cb[2 * i] - cn[2 * j];
          z = cadd(z, cmul(l0, g));
          q = cadd(q, cadd(cmul(lw, g), cscale(cmul(l0, g), w[b] * occ)));
          if (occ == pattern[b]) p = cadd(p, cmul(lp, g));
        }
      }
      cstore(E(0, b + 1), j, p);
      cstore(E(1, b + 1), j, z);
      cstore(E(2, b + 1), j, q);
    }
    __syncthreads();
  }
  for (int b = M - 1; b >= 0; --b) {  // right: v[b][i] = (sum_j G[i][j] v[b+1][j]) mask[b][i]
    const double2 *G = reinterpret_cast<const double2 *>(Gamma) + (int64_t)b * chi * chi;
    const int32_t *cb = charge + (int64_t)b * chi * 2, *cn = charge + (int64_t)(b + 1) * chi * 2;
    for (int i = threadIdx.x; i < chi; i += blockDim.x) {
      cplx p{0.0, 0.0}, z{0.0, 0.0}, q{0.0, 0.0};
      if (obs_mask(b, i, M, d, chi, charge, Lambda)) {
        for (int j = 0; j < chi; ++j) {
          const cplx r0 = cload(E(4, b + 1), j), rw = cload(E(5, b + 1), j),
                     rp = cload(E(3, b + 1), j);
          if (r0.x == 0.0 && r0.y == 0.0 && rw.x == 0.0 && rw.y == 0.0 && rp.x == 0.0 &&
              rp.y == 0.0)
            continue;
          const double2 gv = G[(int64_t)i * chi + j];
          const cplx g{gv.x, gv.y};
          const int occ = cb[2 * i] - cn[2 * j];
          z = cadd(z, cmul(g, r0));
          q = cadd(q, cadd(cmul(g, rw), cscale(cmul(g, r0), w[b] * occ)));
          if (occ == pattern[b]) p = cadd(p, cmul(g, rp));
        }
      }
      cstore(E(3, b), i, p);
      cstore(E(4, b), i, z);
      cstore(E(5, b), i, q);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double P = 0.0, nw = 0.0;
    for (int x = 0; x < chi; ++x) {
      P += cload(E(3, 0), x).x * cload(E(0, 0), x).x - cload(E(3, 0), x).y * cload(E(0, 0), x).y;
      nw += cload(E(5, 0), x).x * cload(E(1, 0), x).x - cload(E(5, 0), x).y * cload(E(1, 0), x).y;
    }
    out[0] = -log(P) + nw;
    out[1] = P;
    out[2] = nw;
  }
}

// gGamma[b][i][j] = conj(-lp_b[i] keep rp_{b+1}[j] / P + lw_b[i] r0_{b+1}[j]
//                        + l0_b[i] rw_{b+1}[j] + w_b occ l0_b[i] r0_{b+1}[j])
__global__ void k_obs_grad(int M, int d, int chi, const int32_t *charge, const int32_t *pattern,
                           const double *w, const double *env, const double *out, double *gG) {
  const double invP = 1.0 / out[1];
  const int64_t tot = (int64_t)M * chi * chi;
  auto E = [&](int kind, int b) { return env + 2 * (((int64_t)kind * (M + 1) + b) * chi); };
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / ((int64_t)chi * chi));
    const int i = (int)((e / chi) % chi), j = (int)(e % chi);
    const int occ = charge[((int64_t)b * chi + i) * 2] - charge[((int64_t)(b + 1) * chi + j) * 2];
    const cplx lp = cload(E(0, b), i), l0 = cload(E(1, b), i), lw = cload(E(2, b), i);
    const cplx rp = cload(E(3, b + 1), j), r0 = cload(E(4, b + 1), j), rw = cload(E(5, b + 1), j);
    cplx g = cadd(cmul(lw, r0), cmul(l0, rw));
    g = cadd(g, cscale(cmul(l0, r0), w[b] * occ));
    if (occ == pattern[b]) g = csub(g, cscale(cmul(lp, rp), invP));
    cstore(gG, e, cconj(g));
  }
}

int launch_observable_vjp(int M, int d, int chi, const double *Gamma, const double *Lambda,
                          const int32_t *charge, const int32_t *pattern, const double *w,
                          double *env, double *gGamma, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_obs_envs, dim3(1), dim3(1024), 0, st, M, d, chi, Gamma, Lambda, charge,
                     pattern, w, env, out);
  BSMPO_HIP(hipGetLastError());
  const int64_t tot = (int64_t)M * chi * chi;
  const int g = (int)std::min<int64_t>((tot + 255) / 256, 8192);
  hipLaunchKernelGGL(k_obs_grad, dim3(g), dim3(256), 0, st, M, d, chi, charge, pattern, w, env,
                     out, gGamma);
  BSMPO_HIP(hipGetLastError());
  return 0;
}

}  // namespace bsmpo
