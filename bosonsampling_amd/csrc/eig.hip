// Tridiagonal route of the sector eigensolver (Gram matrices with n above
// the Jacobi threshold), FP64:
//   k_tridiag      Householder reduction G = Q T Q^H (reflectors H = I - tau v v^H
//                  chosen so that T is real symmetric), one block per sector,
//                  G column-major in global memory, v / p vectors in LDS
//   k_bisect       every eigenvalue of T by Sturm-count bisection, one thread per
//                  eigenvalue, written sorted descending with the drop rule
//   k_tri_vectors  after the global top-chi: eigenvectors of the KEPT eigenvalues
//                  only, by inverse iteration (partial-pivoting tridiagonal solve,
//                  Gram-Schmidt inside clusters of close eigenvalues), then the
//                  back-transformation x = H_0 H_1 ... H_{n-2} y.
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

#include "common.h"

namespace bsmpo {

constexpr int kTriRows = 8;   // rows per thread in the reduction passes (n <= 2048)
constexpr double kDropRel2Tri = 1e-12;

__device__ double block_sum_d(double v, double *red) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(256) void k_tridiag(PlanHeader h, const char *work, int cls_begin,
                                                 double *__restrict__ Gbuf,
                                                 double *__restrict__ TD,
                                                 double *__restrict__ TE,
                                                 double *__restrict__ TAU) {
  extern __shared__ double2 tsm[];
  __shared__ double red[8];
  const DevPlan P = dev_plan(h, work);
  const Sector s = P.sectors[P.jac_order[cls_begin + blockIdx.x]];
  const int n = s.n;
  double2 *A = reinterpret_cast<double2 *>(Gbuf) + s.g_off;  // A(r, c) = A[c n + r]
  double2 *v = tsm, *p = tsm + n;
  double *d = TD + s.s_off, *e = TE + s.s_off;
  double2 *tau = reinterpret_cast<double2 *>(TAU) + s.s_off;
  const int t = threadIdx.x, nt = blockDim.x;
  for (int i = 0; i + 1 < n; ++i) {
    // Householder of x = A(i+1:n, i): H^H (alpha; x2) = (beta; 0), beta real
    double xn2 = 0.0;
    for (int r = i + 2 + t; r < n; r += nt) {
      const double2 a = A[(int64_t)i * n + r];
      xn2 += a.x * a.x + a.y * a.y;
    }
    xn2 = block_sum_d(xn2, red);
    const double2 alpha = A[(int64_t)i * n + i + 1];
    double2 tu = make_double2(0.0, 0.0), scale = make_double2(0.0, 0.0);
    double beta = alpha.x;
    if (!(xn2 == 0.0 && alpha.y == 0.0)) {
      beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
      tu = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      const double2 den = make_double2(alpha.x - beta, alpha.y);  // 1 / (alpha - beta)
      const double dd = den.x * den.x + den.y * den.y;
      scale = make_double2(den.x / dd, -den.y / dd);
    }
    if (t == 0) {
      d[i] = A[(int64_t)i * n + i].x;
      e[i] = beta;
      tau[i] = tu;
    }
    for (int r = i + 1 + t; r < n; r += nt) {
      double2 vr;
      if (r == i + 1) {
        vr = make_double2(1.0, 0.0);
      } else {
        const double2 a = A[(int64_t)i * n + r];
        vr = make_double2(a.x * scale.x - a.y * scale.y, a.x * scale.y + a.y * scale.x);
        A[(int64_t)i * n + r] = vr;  // reflector kept below the subdiagonal
      }
      v[r] = vr;
    }
    __syncthreads();
    if (tu.x == 0.0 && tu.y == 0.0) continue;  // H = I (uniform)
    // p = tau A22 v over rows r in [i+1, n)
    double2 acc[kTriRows];
#pragma unroll
    for (int k = 0; k < kTriRows; ++k) acc[k] = make_double2(0.0, 0.0);
    for (int c = i + 1; c < n; ++c) {
      const double2 vc = v[c];
#pragma unroll
      for (int k = 0; k < kTriRows; ++k) {
        const int r = i + 1 + t + k * nt;
        if (r < n) {
          const double2 a = A[(int64_t)c * n + r];
          acc[k].x += a.x * vc.x - a.y * vc.y;
          acc[k].y += a.x * vc.y + a.y * vc.x;
        }
      }
    }
    double pr = 0.0, pi = 0.0;  // p^H v, with p = tau A v
#pragma unroll
    for (int k = 0; k < kTriRows; ++k) {
      const int r = i + 1 + t + k * nt;
      if (r < n) {
        const double2 q = make_double2(tu.x * acc[k].x - tu.y * acc[k].y,
                                       tu.x * acc[k].y + tu.y * acc[k].x);
        acc[k] = q;
        const double2 vr = v[r];
        pr += q.x * vr.x + q.y * vr.y;   // conj(q) v
        pi += q.x * vr.y - q.y * vr.x;
      }
    }
    pr = block_sum_d(pr, red);
    pi = block_sum_d(pi, red);
    // alpha2 = -1/2 tau (p^H v);  w = p + alpha2 v  (stored in p)
    const double2 a2 = make_double2(-0.5 * (tu.x * pr - tu.y * pi), -0.5 * (tu.x * pi + tu.y * pr));
#pragma unroll
    for (int k = 0; k < kTriRows; ++k) {
      const int r = i + 1 + t + k * nt;
      if (r < n) {
        const double2 vr = v[r];
        p[r] = make_double2(acc[k].x + a2.x * vr.x - a2.y * vr.y, acc[k].y + a2.x * vr.y + a2.y * vr.x);
      }
    }
    __syncthreads();
    // A22 -= v w^H + w v^H
    double2 vr_[kTriRows], wr_[kTriRows];
#pragma unroll
    for (int k = 0; k < kTriRows; ++k) {
      const int r = i + 1 + t + k * nt;
      vr_[k] = r < n ? v[r] : make_double2(0.0, 0.0);
      wr_[k] = r < n ? p[r] : make_double2(0.0, 0.0);
    }
    for (int c = i + 1; c < n; ++c) {
      const double2 vc = v[c], wc = p[c];
#pragma unroll
      for (int k = 0; k < kTriRows; ++k) {
        const int r = i + 1 + t + k * nt;
        if (r < n) {
          double2 a = A[(int64_t)c * n + r];
          // v_r conj(w_c) + w_r conj(v_c)
          a.x -= vr_[k].x * wc.x + vr_[k].y * wc.y + wr_[k].x * vc.x + wr_[k].y * vc.y;
          a.y -= vr_[k].y * wc.x - vr_[k].x * wc.y + wr_[k].y * vc.x - wr_[k].x * vc.y;
          A[(int64_t)c * n + r] = a;
        }
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    d[n - 1] = A[(int64_t)(n - 1) * n + n - 1].x;
    tau[n - 1] = make_double2(0.0, 0.0);
  }
}

// number of eigenvalues of the symmetric tridiagonal (d, e) below x
__device__ int sturm_below(const double *d, const double *e, int n, double x, double pivmin) {
  int c = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0.0) ++c;
  for (int i = 1; i < n; ++i) {
    q = d[i] - x - e[i - 1] * e[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++c;
  }
  return c;
}

__device__ double bisect_kth(const double *d, const double *e, int n, int j, double lo, double hi,
                             double pivmin) {
  const int target = n - 1 - j;  // j-th largest = (n-1-j)-th smallest
  for (int it = 0; it < 128; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (!(mid > lo && mid < hi)) break;
    if (sturm_below(d, e, n, mid, pivmin) <= target)
      lo = mid;
    else
      hi = mid;
  }
  return 0.5 * (lo + hi);
}

// grid (large sectors, ceil(n_max / 256)); thread -> eigenvalue index j
__global__ __launch_bounds__(256) void k_bisect(PlanHeader h, const char *work, int cls_begin,
                                                const double *__restrict__ TD,
                                                const double *__restrict__ TE,
                                                double *__restrict__ lam,
                                                double *__restrict__ S,
                                                int32_t *__restrict__ flags) {
  const DevPlan P = dev_plan(h, work);
  const Sector s = P.sectors[P.jac_order[cls_begin + blockIdx.x]];
  const int n = s.n;
  const double *d = TD + s.s_off, *e = TE + s.s_off;
  // Gershgorin interval and pivot guard (redundant per thread; O(n))
  double lo = d[0], hi = d[0], emax2 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    if (i + 1 < n) emax2 = fmax(emax2, e[i] * e[i]);
  }
  const double span = fmax(fabs(lo), fabs(hi));
  lo -= 4.0 * DBL_EPSILON * span + DBL_MIN;
  hi += 4.0 * DBL_EPSILON * span + DBL_MIN;
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  const double lmax = bisect_kth(d, e, n, 0, lo, hi, pivmin);
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double l = j == 0 ? lmax : bisect_kth(d, e, n, j, lo, hi, pivmin);
  const bool keep = lmax > 0.0 && l > kDropRel2Tri * lmax;
  lam[s.s_off + j] = l;
  S[s.s_off + j] = keep ? sqrt(l) : 0.0;
  flags[s.s_off + j] = keep ? 1 : 0;
}

// per-sector count of kept singular values (a prefix of the sorted list)
__global__ void k_kcount(PlanHeader h, const char *work, const int32_t *__restrict__ slot,
                         int32_t *__restrict__ kcount) {
  __shared__ int red[8];
  const DevPlan P = dev_plan(h, work);
  const Sector s = P.sectors[blockIdx.x];
  int c = 0;
  for (int p = threadIdx.x; p < s.n; p += blockDim.x) c += slot[s.s_off + p] >= 0 ? 1 : 0;
  for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) tot += red[i];
    kcount[blockIdx.x] = tot;
  }
}

// Solve (T - lam I) x = b in place (b -> x), partial pivoting (as LAPACK dgtsv).
// dd, du, dl: work arrays of length n (dl also holds the fill-in).
__device__ void tri_solve(const double *d, const double *e, int n, double lam, double tiny,
                          double *dd, double *du, double *dl, double *b) {
  for (int i = 0; i < n; ++i) dd[i] = d[i] - lam;
  for (int i = 0; i + 1 < n; ++i) { du[i] = e[i]; dl[i] = e[i]; }
  for (int i = 0; i + 1 < n; ++i) {
    if (fabs(dd[i]) >= fabs(dl[i])) {
      if (dd[i] == 0.0) dd[i] = tiny;
      const double f = dl[i] / dd[i];
      dd[i + 1] -= f * du[i];
      b[i + 1] -= f * b[i];
      dl[i] = 0.0;
    } else {
      const double f = dd[i] / dl[i];
      dd[i] = dl[i];
      const double tmp = dd[i + 1];
      dd[i + 1] = du[i] - f * tmp;
      if (i + 2 < n) {
        dl[i] = du[i + 1];
        du[i + 1] = -f * dl[i];
      } else {
        dl[i] = 0.0;
      }
      du[i] = tmp;
      const double bt = b[i];
      b[i] = b[i + 1];
      b[i + 1] = bt - f * b[i + 1];
    }
  }
  if (dd[n - 1] == 0.0) dd[n - 1] = tiny;
  b[n - 1] /= dd[n - 1];
  if (n >= 2) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / dd[n - 2];
  for (int i = n - 3; i >= 0; --i) b[i] = (b[i] - du[i] * b[i + 1] - dl[i] * b[i + 2]) / dd[i];
}

// One block per large sector: eigenvectors of the kept eigenvalues j < k,
// written as complex columns (n long) into the sector's Y slot, then
// back-transformed by the Householder reflectors stored in G.
__global__ __launch_bounds__(256) void k_tri_vectors(
    PlanHeader h, const char *work, int cls_begin, const double *__restrict__ Gbuf,
    const double *__restrict__ TD, const double *__restrict__ TE, const double *__restrict__ TAU,
    const double *__restrict__ lam, const int32_t *__restrict__ kcount,
    double *__restrict__ Vbuf, double *__restrict__ Wbuf, double *__restrict__ triwork) {
  extern __shared__ double2 vsm[];  // reflector cache (n)
  __shared__ int cstart[2048 + 1];
  __shared__ int ncl;
  const DevPlan P = dev_plan(h, work);
  const int sid = P.jac_order[cls_begin + blockIdx.x];
  const Sector s = P.sectors[sid];
  const int n = s.n, k = kcount[sid];
  if (k == 0) return;
  const double *d = TD + s.s_off, *e = TE + s.s_off, *l = lam + s.s_off;
  double2 *X = s.left_small ? reinterpret_cast<double2 *>(Vbuf) + s.v_off
                            : reinterpret_cast<double2 *>(Wbuf) + s.w_off;  // column j: X[j n + r]
  // clusters of close eigenvalues (gap <= 1e-3 |T|), as LAPACK's stein
  if (threadIdx.x == 0) {
    double tn = 0.0;
    for (int i = 0; i < n; ++i)
      tn = fmax(tn, fabs(d[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0));
    const double ortol = 1e-3 * tn;
    int c = 0;
    for (int j = 0; j < k; ++j)
      if (j == 0 || fabs(l[j - 1] - l[j]) > ortol) cstart[c++] = j;
    cstart[c] = k;
    ncl = c;
  }
  __syncthreads();
  double tn1 = 0.0;
  for (int i = 0; i < n; ++i) tn1 = fmax(tn1, fabs(d[i]) + (i + 1 < n ? 2.0 * fabs(e[i]) : 0.0));
  const double tiny = DBL_EPSILON * fmax(tn1, DBL_MIN);
  double *wk = triwork + s.tw_off + (int64_t)threadIdx.x * 4 * n;
  double *dd = wk, *du = wk + n, *dl = wk + 2 * n, *b = wk + 3 * n;
  for (int c = threadIdx.x; c < ncl; c += blockDim.x) {
    for (int j = cstart[c]; j < cstart[c + 1]; ++j) {
      for (int i = 0; i < n; ++i) b[i] = 1.0 / sqrt((double)n) * (1.0 + 1e-3 * ((i * 7919 + j * 104729) % 1000) / 1000.0);
      for (int it = 0; it < 3; ++it) {
        tri_solve(d, e, n, l[j], tiny, dd, du, dl, b);
        for (int q = cstart[c]; q < j; ++q) {  // Gram-Schmidt against earlier members
          double dot = 0.0;
          for (int i = 0; i < n; ++i) dot += X[(int64_t)q * n + i].x * b[i];
          for (int i = 0; i < n; ++i) b[i] -= dot * X[(int64_t)q * n + i].x;
        }
        double nr = 0.0;
        for (int i = 0; i < n; ++i) nr += b[i] * b[i];
        nr = nr > 0.0 ? 1.0 / sqrt(nr) : 0.0;
        for (int i = 0; i < n; ++i) b[i] *= nr;
      }
      for (int i = 0; i < n; ++i) X[(int64_t)j * n + i] = make_double2(b[i], 0.0);
    }
  }
  __syncthreads();
  // x = H_0 H_1 ... H_{n-2} y : apply reflectors from the last to the first
  const double2 *A = reinterpret_cast<const double2 *>(Gbuf) + s.g_off;
  const double2 *tau = reinterpret_cast<const double2 *>(TAU) + s.s_off;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int i = n - 2; i >= 0; --i) {
    const double2 ti = tau[i];
    if (ti.x == 0.0 && ti.y == 0.0) continue;
    for (int r = i + 1 + threadIdx.x; r < n; r += blockDim.x)
      vsm[r] = r == i + 1 ? make_double2(1.0, 0.0) : A[(int64_t)i * n + r];
    __syncthreads();
    for (int j = wave; j < k; j += nw) {
      double2 *x = X + (int64_t)j * n;
      double sr = 0.0, si = 0.0;  // v^H x
      for (int r = i + 1 + lane; r < n; r += 64) {
        const double2 a = vsm[r], b2 = x[r];
        sr += a.x * b2.x + a.y * b2.y;
        si += a.x * b2.y - a.y * b2.x;
      }
      for (int o = 32; o >= 1; o >>= 1) {
        sr += __shfl_xor(sr, o);
        si += __shfl_xor(si, o);
      }
      const double2 f = make_double2(ti.x * sr - ti.y * si, ti.x * si + ti.y * sr);  // tau (v^H x)
      for (int r = i + 1 + lane; r < n; r += 64) {
        const double2 a = vsm[r];
        double2 b2 = x[r];
        b2.x -= a.x * f.x - a.y * f.y;
        b2.y -= a.x * f.y + a.y * f.x;
        x[r] = b2;
      }
    }
    __syncthreads();
  }
}

int launch_tri_values(const PlanHeader &h, void *d_work, hipStream_t stream) {
  const int c = kTriClass;
  const int nsec = h.jac_off[c + 1] - h.jac_off[c];
  if (nsec == 0) return 0;
  if (h.jac_max_n[c] > 256 * kTriRows) return fail(-1, "sector too large for the tridiagonal route");
  char *w = reinterpret_cast<char *>(d_work);
  double *G = reinterpret_cast<double *>(w + h.w_G);
  double *TD = reinterpret_cast<double *>(w + h.w_td);
  double *TE = reinterpret_cast<double *>(w + h.w_te);
  double *TAU = reinterpret_cast<double *>(w + h.w_tau);
  const size_t lds = (size_t)2 * h.jac_max_n[c] * 16;
  hipLaunchKernelGGL(k_tridiag, dim3(nsec), dim3(256), lds, stream, h, (const char *)d_work,
                     h.jac_off[c], G, TD, TE, TAU);
  BSMPO_HIP(hipGetLastError());
  double *LAM = reinterpret_cast<double *>(w + h.w_lam);
  double *S = reinterpret_cast<double *>(w + h.w_S);
  int32_t *F = reinterpret_cast<int32_t *>(w + h.w_flag);
  hipLaunchKernelGGL(k_bisect, dim3(nsec, (h.jac_max_n[c] + 255) / 256), dim3(256), 0, stream, h,
                     (const char *)d_work, h.jac_off[c], (const double *)TD, (const double *)TE,
                     LAM, S, F);
  BSMPO_HIP(hipGetLastError());
  return 0;
}

int launch_kcount(const PlanHeader &h, void *d_work, hipStream_t stream) {
  if (h.n_sectors == 0) return 0;
  char *w = reinterpret_cast<char *>(d_work);
  hipLaunchKernelGGL(k_kcount, dim3(h.n_sectors), dim3(256), 0, stream, h, (const char *)d_work,
                     (const int32_t *)(w + h.w_slot), (int32_t *)(w + h.w_kcount));
  BSMPO_HIP(hipGetLastError());
  return 0;
}

int launch_tri_vectors(const PlanHeader &h, void *d_work, hipStream_t stream) {
  const int c = kTriClass;
  const int nsec = h.jac_off[c + 1] - h.jac_off[c];
  if (nsec == 0) return 0;
  char *w = reinterpret_cast<char *>(d_work);
  const size_t lds = (size_t)h.jac_max_n[c] * 16;
  hipLaunchKernelGGL(k_tri_vectors, dim3(nsec), dim3(256), lds, stream, h, (const char *)d_work,
                     h.jac_off[c], (const double *)(w + h.w_G), (const double *)(w + h.w_td),
                     (const double *)(w + h.w_te), (const double *)(w + h.w_tau),
                     (const double *)(w + h.w_lam), (const int32_t *)(w + h.w_kcount),
                     (double *)(w + h.w_V), (double *)(w + h.w_W), (double *)(w + h.w_triwork));
  BSMPO_HIP(hipGetLastError());
  return 0;
}

}  // namespace bsmpo
