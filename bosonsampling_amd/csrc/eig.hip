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

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "common.h"

namespace bsmpo {

constexpr double kClusterRel = 1e-6;

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

// Householder generator (LAPACK zlarfg form): H^H (alpha; x) = (beta; 0) with
// beta real, H = I - tau v v^H, v = (1; x / (alpha - beta)).
struct Reflector {
  double2 tau, scale;
  double beta;
};
__device__ __forceinline__ Reflector make_reflector(double2 alpha, double xn2) {
  Reflector R;
  R.tau = make_double2(0.0, 0.0);
  R.scale = make_double2(0.0, 0.0);
  R.beta = alpha.x;
  if (!(xn2 == 0.0 && alpha.y == 0.0)) {
    R.beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
    R.tau = make_double2((R.beta - alpha.x) / R.beta, -alpha.y / R.beta);
    const double2 den = make_double2(alpha.x - R.beta, alpha.y);
    const double dd = den.x * den.x + den.y * den.y;
    R.scale = make_double2(den.x / dd, -den.y / dd);
  }
  return R;
}

__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Householder tridiagonalisation G = Q T Q^H, one block per sector, blocked
// by panels of kTriNB columns (the LAPACK zhetrd / zlatrd organisation,
// restated):
//   column j of a panel: apply the panel's earlier reflectors to column j,
//   form its reflector v_j, then w_j from p = A22 v_j, where A22 is the
//   trailing matrix as of the panel start and the panel's pending rank-2
//   updates enter as corrections -V (W^H v) - W (V^H v);
//   after the panel: one pass A22 -= V W^H + W V^H over the trailing matrix.
// Only the lower triangle of the (Hermitian) trailing matrix is read or
// written: per column it is read once (the product A22 v, each entry below the
// diagonal serving A v and A^H v alike); it is written once per panel.  V and
// W of the panel (n x kTriNB each, column-major) live in LDS when they fit the
// launch's LDS (tri_lds), else in the sector's tridiagonal work area.  The
// product is the only step of a column that streams memory: the steps around
// it are a chain of short dependent ones (five barriers; block reductions
// alternate between two LDS buffers, so each costs one barrier), which is
// why the column's data stay on chip between them.  Two block shapes: 256
// threads x 2 rows (n <= 512) and 512 threads x 4 rows (n <= 2048); each
// launch takes its own segment of the sector list, which is sorted largest
// first (n outside [nmin, nmax] is skipped, a guard of the LDS sizing).
constexpr int kTriNB = kTriPanel;
// Streaming shape of the reduction.  Measured on the M=64, chi=1024 grouped
// workload (one launching stream): strips of 4 columns at 3 waves per SIMD
// (168 VGPRs, one trailing tile per pass) 1.7% faster, 4 waves per SIMD
// (spills) 25% slower, strips of 4 at 2 waves 4% slower; the stage streams
// ~76% of the copy bandwidth either way, so the plain shape stays.
constexpr int kTriSW = 8;    // columns per strip of the streamed product A22 v (8 or 4)
constexpr int kTriOcc = 2;   // waves per SIMD the launch bound asks for
constexpr int kTriTT = 2;    // trailing-update tiles per wave pass

__device__ __forceinline__ double2 csub_mul(double2 a, double2 x, double2 y) {  // a - x conj(y)
  a.x -= x.x * y.x + x.y * y.y;
  a.y -= x.y * y.x - x.x * y.y;
  return a;
}
__device__ __forceinline__ double2 cfma2(double2 acc, double2 a, double2 b) {  // acc + a b
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
  return acc;
}

template <int NW>
__device__ __forceinline__ double2 block_sum2(double a, double b, double2 (*red)[16], int &par) {
  for (int o = 32; o >= 1; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  if ((threadIdx.x & 63) == 0) red[par][threadIdx.x >> 6] = make_double2(a, b);
  __syncthreads();
  double2 t = make_double2(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    t.x += red[par][i].x;
    t.y += red[par][i].y;
  }
  par ^= 1;
  return t;
}

// x[q] (q = 0..7) summed over the 64 lanes of a wave by recursive halving:
// lane l returns the total of x[(l >> 3) & 7].
__device__ __forceinline__ double2 wave_colsum8(const double2 (&x)[8], int lane) {
  const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8;
  double2 y4[4], y2[2];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double2 snd = h5 ? x[q] : x[q + 4], kp = h5 ? x[q + 4] : x[q];
    y4[q] = make_double2(kp.x + __shfl_xor(snd.x, 32), kp.y + __shfl_xor(snd.y, 32));
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const double2 snd = h4 ? y4[q] : y4[q + 2], kp = h4 ? y4[q + 2] : y4[q];
    y2[q] = make_double2(kp.x + __shfl_xor(snd.x, 16), kp.y + __shfl_xor(snd.y, 16));
  }
  const double2 snd = h3 ? y2[0] : y2[1], kp = h3 ? y2[1] : y2[0];
  double2 y = make_double2(kp.x + __shfl_xor(snd.x, 8), kp.y + __shfl_xor(snd.y, 8));
#pragma unroll
  for (int o = 4; o >= 1; o >>= 1) {
    y.x += __shfl_xor(y.x, o);
    y.y += __shfl_xor(y.y, o);
  }
  return y;
}

// x[q] (q = 0..3) summed over the 64 lanes of a wave by recursive halving:
// lane l returns the total of x[(l >> 4) & 3].
__device__ __forceinline__ double2 wave_colsum4(const double2 (&x)[4], int lane) {
  const bool h5 = lane & 32, h4 = lane & 16;
  double2 y2[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const double2 snd = h5 ? x[q] : x[q + 2], kp = h5 ? x[q + 2] : x[q];
    y2[q] = make_double2(kp.x + __shfl_xor(snd.x, 32), kp.y + __shfl_xor(snd.y, 32));
  }
  const double2 snd = h4 ? y2[0] : y2[1], kp = h4 ? y2[1] : y2[0];
  double2 y = make_double2(kp.x + __shfl_xor(snd.x, 16), kp.y + __shfl_xor(snd.y, 16));
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) {
    y.x += __shfl_xor(y.x, o);
    y.y += __shfl_xor(y.y, o);
  }
  return y;
}
// x[q] (q = 0..31) summed over the 64 lanes of a wave by recursive halving:
// at each exchange distance o = 32 .. 2 a lane keeps the half of its values
// selected by bit o of its index and adds the partner's copy of that half;
// lanes l and l ^ 1 end with the total of x[(l >> 1) & 31].
template <int H>
__device__ __forceinline__ void halve32(double (&x)[32], int lane) {
  constexpr int o = 2 * H;
  const bool up = (lane & o) != 0;
#pragma unroll
  for (int q = 0; q < H; ++q) {
    const double keep = up ? x[q + H] : x[q];
    const double send = up ? x[q] : x[q + H];
    x[q] = keep + __shfl_xor(send, o);
  }
}
__device__ __forceinline__ double wave_sum32(double (&x)[32], int lane) {
  halve32<16>(x, lane);
  halve32<8>(x, lane);
  halve32<4>(x, lane);
  halve32<2>(x, lane);
  halve32<1>(x, lane);
  return x[0] + __shfl_xor(x[0], 1);
}

template <int SW>
__device__ __forceinline__ double2 wave_colsum(const double2 (&x)[SW], int lane) {
  if constexpr (SW == 8) return wave_colsum8(x, lane);
  else return wave_colsum4(x, lane);
}

// LDS of a sector's reduction: v and A22 v (2 n complex), and with VWL the
// panel's V and W (2 n kTriNB complex) too.
__host__ __device__ constexpr int64_t tri_lds_need(int n, bool vwl) {
  return 16 * (3 * (int64_t)n + 3 + (vwl ? 2 * (int64_t)n * kTriNB : 0));
}

// One sector's reduction by the whole block.  LDS (tsm, complex entries): v,
// A22 v and the stashed next column (n each), the two pivot slots, the strip
// counter of the product (one entry), then with VWL the panel's V and W
// (n kTriNB each); without VWL, V and W live in the sector's tridiagonal work
// area in global memory.
template <int NT, int RPT, int SW, bool VWL>
__device__ __forceinline__ void tridiag_sector(const UpdDesc &u, const Sector &s, double2 *tsm,
                                               double2 (*red)[16], double (*red4)[4 * kTriNB]) {
  constexpr int NW = NT / 64;
  const int n = s.n;
  gptr<double2> A = gp(reinterpret_cast<double2 *>(u.G) + s.g_off);  // A(r, c) = A[c n + r]
  double2 *vs = tsm;                                                  // v_j, n entries
  double2 *ys = tsm + n;                                              // A22 v_j, n entries
  const lptr<double2> nx = (lptr<double2>)(tsm + 2 * n);  // column j + 1 as (c) of column j read it
  double2 *piv = tsm + 3 * n;  // column j's diagonal and subdiagonal entries
  int *ctr = reinterpret_cast<int *>(tsm + 3 * n + 2);  // next strip of the product (c)
  using VWP = typename std::conditional<VWL, lptr<double2>, gptr<double2>>::type;
  VWP Vp;
  if constexpr (VWL)
    Vp = (VWP)(tsm + 3 * n + 3);
  else
    Vp = gp(reinterpret_cast<double2 *>(u.triwork + s.tw_off));
  const VWP Wp = Vp + (int64_t)n * kTriNB;
  gptr<double> d = gp(u.td + s.s_off), e = gp(u.te + s.s_off);
  gptr<double2> tau = gp(reinterpret_cast<double2 *>(u.tau) + s.s_off);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  int par = 0;
  // The reduction works on G scaled by a power of two that brings its largest
  // diagonal entry (its largest entry: G is PSD) into [1, 2); a sector of tiny
  // weights would otherwise underflow in the squared column norms of the
  // reflectors.  The scaling is exact.  It is applied as the first panel reads
  // the matrix (every entry is rewritten, scaled, by then), and e[n - 1], not
  // an off-diagonal, keeps the factor that maps the eigenvalues back (0 for a
  // zero matrix, whose eigenvalues are all dropped).
  double scl;
  {
    double mx = 0.0;
    for (int i = t; i < n; i += NT) mx = fmax(mx, fabs(A[(int64_t)i * n + i].x));
    for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if (lane == 0) red[0][wave].x = mx;
    __syncthreads();
    mx = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) mx = fmax(mx, red[0][w].x);
    const int ex = (mx > 0.0 && isfinite(mx)) ? ilogb(mx) : 0;
    scl = ldexp(1.0, -ex);
    if (t == 0) e[n - 1] = mx > 0.0 ? ldexp(1.0, ex) : 0.0;
    __syncthreads();
  }
  if (n == 1) {
    if (t == 0) { d[0] = A[0].x * scl; tau[0] = make_double2(0.0, 0.0); }
    return;
  }
  for (int j0 = 0; j0 < n - 1; j0 += kTriNB) {
    if (j0 > 0) scl = 1.0;  // entries read from here on were written scaled
    const int pb = min(kTriNB, n - 1 - j0);
    for (int i = 0; i < pb; ++i) {
      const int j = j0 + i;
      // (a) column j, rows r >= j (thread t: rows j + t + k NT), gets the
      //     panel's columns l < i and stays in registers; inside a panel it is
      //     read from the copy the previous column's product (c) left in LDS
      //     (the trailing matrix is unchanged within a panel), at a panel's
      //     first column from global memory.  The diagonal entry and the
      //     subdiagonal one (alpha) go to two LDS slots for every thread.
      double2 col[RPT];
      double xn2 = 0.0;
      // the load as a parameter, so that the LDS and the global variants stay
      // two loops (one selected address would be a flat load)
      auto column = [&](auto load) {
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
          const int r = j + t + k * NT;
          col[k] = make_double2(0.0, 0.0);
          if (r < n) {
            double2 a = load(r);
            a.x *= scl;
            a.y *= scl;
            for (int l = 0; l < i; ++l) {
              a = csub_mul(a, Vp[(int64_t)l * n + r], Wp[(int64_t)l * n + j]);
              a = csub_mul(a, Wp[(int64_t)l * n + r], Vp[(int64_t)l * n + j]);
            }
            col[k] = a;
            if (r >= j + 2) xn2 += a.x * a.x + a.y * a.y;
            if (r <= j + 1) piv[r - j] = a;
          }
        }
      };
      if (i > 0)  // uniform
        column([&](int r) -> double2 { return nx[r]; });
      else
        column([&](int r) -> double2 { return A[(int64_t)j * n + r]; });
      xn2 = block_sum2<NW>(xn2, 0.0, red, par).x;  // also orders the pivot slots
      // (b) reflector of column j; v below the subdiagonal to global memory
      //     (the back-transformation's copy), v and zeroed A22 v to LDS
      const double2 alpha = piv[1];
      const Reflector R = make_reflector(alpha, xn2);
      if (t == 0) {
        d[j] = piv[0].x;
        e[j] = R.beta;
        tau[j] = R.tau;
        *ctr = NW;  // strips 0 .. NW - 1 are the waves' first
      }
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const int r = j + t + k * NT;
        if (r > j && r < n) {
          double2 vr;
          if (r == j + 1) {
            vr = make_double2(1.0, 0.0);
          } else {
            vr = cmul2(col[k], R.scale);
            A[(int64_t)j * n + r] = vr;
          }
          vs[r] = vr;
          ys[r] = make_double2(0.0, 0.0);
        }
      }
      __syncthreads();
      // (c) p = A22 v over rows r >= j+1, A22 as of the panel start with only
      //     its lower triangle read: strips of 8 columns go round the waves;
      //     lane rows r >= c feed p(r) += A(r, c) v(c) and, below the
      //     diagonal, p(c) += conj(A(r, c)) v(r).  Both sums go to ys by LDS
      //     atomics (row sums per lane, column sums after a wave reduction).
      //     A wave's items (strip s, row block r0 = s, s + 64, ...; then the
      //     next strip) form one sequence whose loads run two items ahead,
      //     across strip boundaries.  Wave w starts on strip w; each further
      //     strip is claimed from an LDS counter (strips longest first), so
      //     that waves slowed by memory take fewer and the barrier after the
      //     product waits less for the slowest.
      {
        auto advance = [&](int &s, int &r0) {  // next item of this wave
          r0 += 64;
          if (r0 >= n) {
            int k = 0;
            if (lane == 0) k = atomicAdd(ctr, 1);
            k = __shfl(k, 0);
            s = j + 1 + SW * k;
            r0 = s;
          }
        };
        auto fetch = [&](double2(&x)[SW], int s, int r0) {
          const int r = r0 + lane;
#pragma unroll
          for (int q = 0; q < SW; ++q)
            x[q] = (r < n && s + q <= r) ? A[(int64_t)(s + q) * n + r] : make_double2(0.0, 0.0);
        };
        double2 cp[SW], a[SW], b[SW], c[SW];
#pragma unroll
        for (int q = 0; q < SW; ++q) cp[q] = make_double2(0.0, 0.0);
        int sa = j + 1 + SW * wave, ra = sa;  // item held in a
        int sb = sa, rb = ra;
        advance(sb, rb);                     // item held in b
        int sc = sb, rc = rb;
        advance(sc, rc);                     // item loaded into c
        fetch(a, sa, ra);
        fetch(b, sb, rb);
        while (sa < n) {  // wave-uniform
          fetch(c, sc, rc);
          const int r = ra + lane;
          if (r < n) {
            if (sa == j + 1) nx[r] = a[0];  // wave-uniform: column j + 1, for (a)
            const double2 vr = vs[r];
            double2 rp = make_double2(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < SW; ++q) {
              rp = cfma2(rp, a[q], vs[min(sa + q, n - 1)]);
              if (sa + q < r) cp[q] = cfma2(cp[q], make_double2(a[q].x, -a[q].y), vr);
            }
            atomicAdd(&ys[r].x, rp.x * scl);
            atomicAdd(&ys[r].y, rp.y * scl);
          }
          if (sb != sa) {  // strip sa done: its column sums
            const double2 cs = wave_colsum<SW>(cp, lane);
            const int cq = sa + (lane >> 3) / (8 / SW);
            if ((lane & (64 / SW - 1)) == 0 && cq < n) {
              atomicAdd(&ys[cq].x, cs.x * scl);
              atomicAdd(&ys[cq].y, cs.y * scl);
            }
#pragma unroll
            for (int q = 0; q < SW; ++q) cp[q] = make_double2(0.0, 0.0);
          }
#pragma unroll
          for (int q = 0; q < SW; ++q) {
            a[q] = b[q];
            b[q] = c[q];
          }
          sa = sb;
          ra = rb;
          sb = sc;
          rb = rc;
          advance(sc, rc);
        }
      }
      // (d) corrections for the panel's pending updates: the per-wave parts of
      //     W_l^H v and V_l^H v, before the barrier that completes A22 v
      //     All 4 i sums of a wave are reduced together by recursive halving
      //     (wave_sum32: 32 exchanges instead of 6 per sum, one chain of 6
      //     steps instead of i).
      if (i > 0) {
        double x[4 * kTriNB];
#pragma unroll
        for (int l = 0; l < kTriNB; ++l) {
          double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
          if (l < i) {  // uniform
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
              const int r = j + 1 + t + k * NT;
              if (r >= n) continue;
              const double2 v = vs[r];
              const double2 w_ = Wp[(int64_t)l * n + r], v_ = Vp[(int64_t)l * n + r];
              p0 += w_.x * v.x + w_.y * v.y;  // conj(w_l) v
              p1 += w_.x * v.y - w_.y * v.x;
              p2 += v_.x * v.x + v_.y * v.y;  // conj(v_l) v
              p3 += v_.x * v.y - v_.y * v.x;
            }
          }
          x[4 * l + 0] = p0;
          x[4 * l + 1] = p1;
          x[4 * l + 2] = p2;
          x[4 * l + 3] = p3;
        }
        const double tot = wave_sum32(x, lane);  // total of x[(lane >> 1) & 31]
        if ((lane & 1) == 0 && (lane >> 3) < i) red4[wave][lane >> 1] = tot;
      }
      __syncthreads();
      double2 acc[RPT];
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const int r = j + 1 + t + k * NT;
        acc[k] = r < n ? ys[r] : make_double2(0.0, 0.0);
      }
      if (i > 0) {
        for (int l = 0; l < i; ++l) {
          double2 y = make_double2(0.0, 0.0), z = make_double2(0.0, 0.0);  // W_l^H v, V_l^H v
#pragma unroll
          for (int wv = 0; wv < NW; ++wv) {
            y.x += red4[wv][4 * l + 0];
            y.y += red4[wv][4 * l + 1];
            z.x += red4[wv][4 * l + 2];
            z.y += red4[wv][4 * l + 3];
          }
#pragma unroll
          for (int k = 0; k < RPT; ++k) {
            const int r = j + 1 + t + k * NT;
            if (r >= n) continue;
            const double2 v_ = Vp[(int64_t)l * n + r], w_ = Wp[(int64_t)l * n + r];
            acc[k].x -= v_.x * y.x - v_.y * y.y + w_.x * z.x - w_.y * z.y;
            acc[k].y -= v_.x * y.y + v_.y * y.x + w_.x * z.y + w_.y * z.x;
          }
        }
      }
      // (e) p = tau A v ; w = p - 1/2 tau (p^H v) v ; store the panel's column i
      double pr = 0.0, pi = 0.0;
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const int r = j + 1 + t + k * NT;
        if (r < n) {
          acc[k] = cmul2(R.tau, acc[k]);
          const double2 v = vs[r];
          pr += acc[k].x * v.x + acc[k].y * v.y;
          pi += acc[k].x * v.y - acc[k].y * v.x;
        }
      }
      const double2 ph = block_sum2<NW>(pr, pi, red, par);
      const double2 a2 = make_double2(-0.5 * (R.tau.x * ph.x - R.tau.y * ph.y),
                                      -0.5 * (R.tau.x * ph.y + R.tau.y * ph.x));
#pragma unroll
      for (int k = 0; k < RPT; ++k) {
        const int r = j + 1 + t + k * NT;
        if (r < n) {
          const double2 v = vs[r];
          const double2 q = cmul2(a2, v);
          Wp[(int64_t)i * n + r] = make_double2(acc[k].x + q.x, acc[k].y + q.y);
          Vp[(int64_t)i * n + r] = v;
        }
      }
      __syncthreads();
    }
    // (f) trailing update A(r, c) -= sum_l V(r, l) conj(W(c, l)) + W(r, l) conj(V(c, l))
    //     on the lower triangle n > r >= c >= jn: a product with K = 2 pb on the
    //     FP64 MFMA, 16 x 16 tiles round the waves.  Each tile is formed
    //     transposed, U^T = [conj W  conj V](c, :) [V  W](r, :)^T, so that a
    //     lane's output elements run down a column of A (contiguous rows).
    const int jn = j0 + pb;
    {
      const int nts = (n - jn + 15) / 16;  // tiles per side
      const int ntiles = nts * (nts + 1) / 2;
      const int i16 = lane & 15, kq = lane >> 4;
      // two tiles per pass (tix and tix + NW), every load of both issued
      // before the first product, so that a pass waits on memory once
      constexpr int kTT = kTriTT;
      for (int tix0 = wave; tix0 < ntiles; tix0 += kTT * NW) {
        int R0[kTT], C0[kTT];
        bool live[kTT];
        double2 av[kTT][4];
        cplx a[kTT][4], b[kTT][4];
#pragma unroll
        for (int u = 0; u < kTT; ++u) {
          const int tix = tix0 + u * NW;
          live[u] = tix < ntiles;  // wave-uniform
          const int tq = live[u] ? tix : 0;
          int ti = (int)((sqrt(8.0 * tq + 1.0) - 1.0) * 0.5);  // tile row: ti (ti + 1) / 2 <= tix
          while ((ti + 1) * (ti + 2) / 2 <= tq) ++ti;
          while (ti * (ti + 1) / 2 > tq) --ti;
          const int tj = tq - ti * (ti + 1) / 2;
          R0[u] = jn + 16 * ti;
          C0[u] = jn + 16 * tj;
          const int ca = C0[u] + i16, rb = R0[u] + i16;  // operand rows: a side (columns of A), b side
#pragma unroll
          for (int q = 0; q < 4; ++q) {  // element (r, c) = (R0 + i16, C0 + kq + 4 q)
            const int r = R0[u] + i16, c = C0[u] + mfma64_row(kq, q);
            av[u][q] = (live[u] && r < n && c <= r) ? A[(int64_t)c * n + r] : make_double2(0.0, 0.0);
          }
#pragma unroll
          for (int k4 = 0; k4 < 4; ++k4) {
            const int k = 4 * k4 + kq, l = k % kTriNB;
            const bool vw = k < kTriNB;  // k < 8: V(r) conj(W(c)); k >= 8: W(r) conj(V(c))
            a[u][k4] = cplx{0.0, 0.0};
            b[u][k4] = cplx{0.0, 0.0};
            if (live[u] && l < pb) {
              if (ca < n) {
                const double2 x = vw ? Wp[(int64_t)l * n + ca] : Vp[(int64_t)l * n + ca];
                a[u][k4] = cplx{x.x, -x.y};
              }
              if (rb < n) {
                const double2 y = vw ? Vp[(int64_t)l * n + rb] : Wp[(int64_t)l * n + rb];
                b[u][k4] = cplx{y.x, y.y};
              }
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kTT; ++u) {
          if (!live[u]) break;
          dbl4 ur = {0.0, 0.0, 0.0, 0.0}, ui = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int k4 = 0; k4 < 4; ++k4) cmfma_f64(ur, ui, a[u][k4], b[u][k4]);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int r = R0[u] + i16, c = C0[u] + mfma64_row(kq, q);
            if (r < n && c <= r) {
              double2 x = av[u][q];
              x.x = x.x * scl - ur[q];
              x.y = x.y * scl - ui[q];
              A[(int64_t)c * n + r] = x;
            }
          }
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

// Blocks walk the list (grid-stride, largest sectors first); sectors with n
// outside [nmin, nmax] belong to the other block shape's launch.
// Column steps (a) and (b) hold RPT rows per thread: the block shapes cover
// n <= NT RPT (512 x 4 = 2048 = kTriMaxN; 256 x 2 = 512 >= the block-shape split).
static_assert(kTriMaxN <= 512 * 4, "512-thread Householder blocks hold 4 rows per thread");
template <int NT, int RPT, int SW, int OCC>
__global__ __launch_bounds__(NT, OCC) void k_tridiag(const UpdDesc *__restrict__ Ud, WorkSeg ws,
                                                int nmin, int nmax, int lds_bytes) {
  extern __shared__ double2 tsm[];
  __shared__ double2 red[2][16];
  __shared__ double red4[NT / 64][4 * kTriNB];
  const int2 sg = *ws.seg;
  const int2 *__restrict__ list = ws.list + sg.x;
  const int nlist = sg.y;
  for (int it = blockIdx.x; it < nlist; it += gridDim.x) {
    const int2 ref = list[it];
    const UpdDesc &u = Ud[ref.x];
    const Sector s = u.sectors[ref.y];
    if (s.n < nmin || s.n > nmax) continue;  // uniform
    __syncthreads();  // the previous sector's readers of LDS are done
    if (tri_lds_need(s.n, true) <= lds_bytes)  // uniform: V and W fit in the LDS
      tridiag_sector<NT, RPT, SW, true>(u, s, tsm, red, red4);
    else
      tridiag_sector<NT, RPT, SW, false>(u, s, tsm, red, red4);
  }
}

// 1 / q from the hardware reciprocal and one Newton step (relative error a few
// ulp): the Sturm recurrence below is a chain of these, and IEEE division is
// about twice as long a dependency chain.  |q| >= pivmin keeps 1 / q finite.
__device__ __forceinline__ double recip_nr(double q) {
  const double r = __builtin_amdgcn_rcp(q);
  return fma(fma(-q, r, 1.0), r, r);
}

// number of eigenvalues of the symmetric tridiagonal (d, e) below x; e2 = e^2.
// The pivot guard (|q| >= pivmin, pivmin >= DBL_MIN max(1, max e^2)) keeps every
// e^2 / q finite; the table loads go four rows at a time, off the chain.
__device__ int sturm_below(const double *d, const double *e2, int n, double x, double pivmin) {
  int c = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  c += q < 0.0;
  int i = 1;
  for (; i + 3 < n; i += 4) {
    double dd[4], ff[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      dd[k] = d[i + k] - x;
      ff[k] = e2[i + k - 1];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      q = dd[k] - ff[k] * recip_nr(q);
      if (fabs(q) < pivmin) q = -pivmin;
      c += q < 0.0;
    }
  }
  for (; i < n; ++i) {
    q = d[i] - x - e2[i - 1] * recip_nr(q);
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0;
  }
  return c;
}

// j-th largest eigenvalue in [lo, hi] by bisection on the Sturm count.  Stops
// early once hi < below: such an eigenvalue is dropped, its value not used.
__device__ double bisect_kth(const double *d, const double *e2, int n, int j, double lo, double hi,
                             double pivmin, double below = -1.0) {
  const int target = n - 1 - j;  // j-th largest = (n-1-j)-th smallest
  for (int it = 0; it < 128; ++it) {
    if (hi < below) return hi;
    const double mid = 0.5 * (lo + hi);
    if (!(mid > lo && mid < hi)) break;
    if (sturm_below(d, e2, n, mid, pivmin) <= target)
      lo = mid;
    else
      hi = mid;
  }
  return 0.5 * (lo + hi);
}

// grid (large sectors, ceil(n_max / 256)); thread -> eigenvalue index j.  The
// tridiagonal (d, e^2) is staged in LDS.  Each block first takes Sturm counts
// at 256 points spaced a quarter octave apart below the Gershgorin bound (one
// per thread), so that each eigenvalue starts from a bracket of relative width
// 2^(1/4) - 1 instead of the whole spectrum; an eigenvalue whose bracket falls
// below the drop threshold (u.drop_rel2 times a lower bound of the largest
// eigenvalue) stops there.  The largest eigenvalue itself is bisected once per
// block, by the thread that owns index 0 in the first block and by thread 0 in
// the others.
//   pass 0 (grid (sectors, 1)): the sample counts only, added to the update's
//          histogram of singular values: the count in [x_m, x_(m-1)) goes to
//          the bin of sqrt(x_m unscale) when x_m is above the drop threshold's
//          bound (those values are surely kept by the drop rule)
//   pass 1: the eigenvalues; those whose bracket lies below half the update's
//          cutoff (no value there can be among the chi kept, see k_cutoff) are
//          not refined: flagged dropped, value the bracket midpoint
constexpr int kBisSamples = 256;
constexpr double kBisOctaves = 0.25;

__device__ __forceinline__ int hist_bin(double sv) {  // -1: below the histogram
  const int q = (int)floor(4.0 * log2(sv));
  return q < -kHistZero ? -1 : min(q + kHistZero, kHistBins - 1);
}
__device__ __forceinline__ double hist_cut(const UpdDesc &u) {
  return *reinterpret_cast<const double *>(u.hist + kHistBins);
}

__global__ __launch_bounds__(256) void k_bisect(const UpdDesc *__restrict__ Ud, WorkSeg ws,
                                                int pass) {
  extern __shared__ double bsm[];
  __shared__ double sh_lmax;
  __shared__ double red[3][4];
  __shared__ double xs[kBisSamples + 1];
  __shared__ int cs[kBisSamples + 1];
  const int2 sg = *ws.seg;
  if ((int)blockIdx.x >= sg.y) return;  // a launch sized for the group's capacity
  const int2 ref = ws.list[sg.x + blockIdx.x];
  const UpdDesc &u = Ud[ref.x];
  const Sector s = u.sectors[ref.y];
  const int n = s.n;
  if ((int)(blockIdx.y * blockDim.x) >= n) return;  // no eigenvalue of this sector here
  const double *d = u.td + s.s_off, *e = u.te + s.s_off;
  double *lam = u.lam, *S = u.S;
  int32_t *flags = u.flag;
  double *ds = bsm, *e2s = bsm + n;
  // Gershgorin interval and pivot guard, by a block reduction
  double glo = INFINITY, ghi = -INFINITY, emax2 = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double di = d[i], ei = i + 1 < n ? e[i] : 0.0;
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + fabs(ei);
    ds[i] = di;
    e2s[i] = ei * ei;
    glo = fmin(glo, di - r);
    ghi = fmax(ghi, di + r);
    emax2 = fmax(emax2, ei * ei);
  }
  for (int o = 32; o >= 1; o >>= 1) {
    glo = fmin(glo, __shfl_xor(glo, o));
    ghi = fmax(ghi, __shfl_xor(ghi, o));
    emax2 = fmax(emax2, __shfl_xor(emax2, o));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = glo;
    red[1][threadIdx.x >> 6] = ghi;
    red[2][threadIdx.x >> 6] = emax2;
  }
  __syncthreads();
  glo = red[0][0];
  ghi = red[1][0];
  emax2 = red[2][0];
#pragma unroll
  for (int w = 1; w < 4; ++w) {
    glo = fmin(glo, red[0][w]);
    ghi = fmax(ghi, red[1][w]);
    emax2 = fmax(emax2, red[2][w]);
  }
  const double span = fmax(fabs(glo), fabs(ghi));
  const double lo = glo - (4.0 * DBL_EPSILON * span + DBL_MIN);
  const double hi = ghi + (4.0 * DBL_EPSILON * span + DBL_MIN);
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  const double unscale = e[n - 1];  // 0: a zero matrix
  {  // sample counts: xs[m] = hi 2^(-m / 4), m < 256, decreasing; xs[256] = lo
    const int m = threadIdx.x;
    const double x = hi * exp2(-kBisOctaves * m);
    xs[m] = x;
    cs[m] = (hi > 0.0 && x > lo) ? sturm_below(ds, e2s, n, x, pivmin) : 0;
    if (m == 0) {
      xs[kBisSamples] = lo;
      cs[kBisSamples] = 0;
    }
  }
  __syncthreads();
  if (pass == 0) {
    const int m = threadIdx.x + 1;  // eigenvalues in [xs[m], xs[m - 1])
    const int cnt = cs[m - 1] - cs[m];
    const double x = xs[m];
    if (cnt > 0 && unscale > 0.0 && x > 0.0 && x > u.drop_rel2 * hi) {
      const int b = hist_bin(sqrt(x * unscale));
      if (b >= 0) atomicAdd(&u.hist[b], cnt);
    }
    return;
  }
  // a lower bound of the largest eigenvalue: the largest sample with an
  // eigenvalue at or above it
  double lmax_lb = 0.0;
  for (int m = 0; m < kBisSamples; ++m)
    if (cs[m] < n) {
      lmax_lb = xs[m];
      break;
    }
  // bracket of the j-th largest: [xs[m + 1], xs[m]] with cs[m] > n-1-j >= cs[m + 1]
  // (cs[0] = n, cs[256] = 0); the whole [lo, hi] when no eigenvalue is positive
  auto bracket = [&](int jj, double &blo, double &bhi) {
    if (!(cs[0] == n)) {
      blo = lo;
      bhi = hi;
      return;
    }
    const int target = n - 1 - jj;
    int m = 0;
    while (cs[m + 1] > target) ++m;  // stops at m = 255 at the latest
    blo = fmax(xs[m + 1], lo);
    bhi = xs[m];
  };
  const double tcut = 0.5 * hist_cut(u);  // in singular-value units
  const int j = blockIdx.y * blockDim.x + threadIdx.x;
  double l = 0.0;
  bool refined = false;
  if (j < n) {
    double blo, bhi;
    bracket(j, blo, bhi);
    if (tcut > 0.0 && !(sqrt(fmax(bhi, 0.0) * unscale) >= tcut)) {
      l = 0.5 * (blo + bhi);  // below every kept value: not refined
    } else {
      l = bisect_kth(ds, e2s, n, j, blo, bhi, pivmin, u.drop_rel2 * lmax_lb);
      refined = true;
    }
  }
  if (threadIdx.x == 0) {
    if (blockIdx.y == 0 && refined) {
      sh_lmax = l;
    } else {
      double blo, bhi;
      bracket(0, blo, bhi);
      sh_lmax = bisect_kth(ds, e2s, n, 0, blo, bhi, pivmin);
    }
  }
  __syncthreads();
  if (j >= n) return;
  const double lmax = sh_lmax;
  const bool keep = refined && unscale > 0.0 && lmax > 0.0 && l > u.drop_rel2 * lmax;
  lam[s.s_off + j] = l;  // of the scaled matrix, as the inverse iteration uses it
  S[s.s_off + j] = keep ? sqrt(l * unscale) : 0.0;
  flags[s.s_off + j] = keep ? 1 : 0;
}

// the histogram of every update of the group, and its cutoff, cleared
__global__ void k_hist_zero(const UpdDesc *__restrict__ Ud, const int2 *__restrict__ useg) {
  if ((int)blockIdx.x >= useg->y) return;  // a skipped group
  int32_t *h = Ud[blockIdx.x].hist;
  for (int i = threadIdx.x; i < kHistBins + 2; i += blockDim.x) h[i] = 0;
}

// the kept (flagged) singular values of the Jacobi-route sectors, exact,
// added to their update's histogram: one block per sector
__global__ void k_hist_jacobi(const UpdDesc *__restrict__ Ud, WorkSeg ws) {
  const int2 sg = *ws.seg;
  if ((int)blockIdx.x >= sg.y) return;
  const int2 ref = ws.list[sg.x + blockIdx.x];
  const UpdDesc &u = Ud[ref.x];
  const Sector s = u.sectors[ref.y];
  for (int r = threadIdx.x; r < s.n; r += blockDim.x) {
    if (!u.flag[s.s_off + r]) continue;
    const double sv = u.S[s.s_off + r];
    if (!(sv > 0.0)) continue;
    const int b = hist_bin(sv);
    if (b >= 0) atomicAdd(&u.hist[b], 1);
  }
}

// per update: the cutoff t = 2^((b - kHistZero) / 4) for the highest bin b
// with at least chi surely-kept values at or above t (each value counted in a
// bin is at least that bin's bound), 0 if there is none.  The chi-th largest
// kept singular value is then >= t, so nothing below t (1 - tie band) is
// kept; pass 1 of k_bisect refines every eigenvalue above t / 2.
__global__ __launch_bounds__(256) void k_cutoff(const UpdDesc *__restrict__ Ud,
                                                const int2 *__restrict__ useg) {
  __shared__ int part[256];
  if ((int)blockIdx.x >= useg->y) return;  // a skipped group
  const UpdDesc &u = Ud[blockIdx.x];
  constexpr int kPer = kHistBins / 256;
  int a = 0;
  for (int i = 0; i < kPer; ++i) a += u.hist[threadIdx.x * kPer + i];
  part[threadIdx.x] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int chi = u.chi;
    int cum = 0, b = -1;
    for (int c = 255; c >= 0 && b < 0; --c) {
      if (cum + part[c] >= chi) {
        for (int i = kPer - 1; i >= 0; --i) {
          cum += u.hist[c * kPer + i];
          if (cum >= chi) {
            b = c * kPer + i;
            break;
          }
        }
      } else {
        cum += part[c];
      }
    }
    *reinterpret_cast<double *>(u.hist + kHistBins) = b >= 0 ? exp2(0.25 * (b - kHistZero)) : 0.0;
  }
}

// per-sector count of kept singular values (a prefix of the sorted list)
// Algorithmic flops of the stages that work on kept pairs only, accumulated
// while stage timing is on (complex MAC = 8 flops): W = V^H C, 8 k L R;
// V = T Y / sigma (sectors with L > R), 8 k L R; back-transformation of the
// tridiagonal route, 8 k n^2 (each reflector of length m costs 2 m MACs per
// vector).  k = kept pairs of the sector.
__device__ double g_kept_flops[3];

int kept_flops_read(double *out, bool reset) {
  BSMPO_HIP(hipMemcpyFromSymbol(out, HIP_SYMBOL(g_kept_flops), sizeof(double) * 3));
  if (reset) {
    const double z[3] = {0.0, 0.0, 0.0};
    BSMPO_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_kept_flops), z, sizeof(z)));
  }
  return 0;
}

__global__ void k_kcount(const UpdDesc *__restrict__ Ud, WorkSeg ws, int count_flops) {
  __shared__ int red[8];
  const int2 sg = *ws.seg;
  if ((int)blockIdx.x >= sg.y) return;
  const int2 ref = ws.list[sg.x + blockIdx.x];
  const UpdDesc &u = Ud[ref.x];
  const Sector s = u.sectors[ref.y];
  int c = 0;
  for (int p = threadIdx.x; p < s.n; p += blockDim.x) c += u.slot[s.s_off + p] >= 0 ? 1 : 0;
  for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) tot += red[i];
    u.kcount[ref.y] = tot;
    if (count_flops && tot > 0) {
      const double k = tot, L = s.L, R = s.R, n = s.n;
      atomicAdd(&g_kept_flops[0], 8.0 * k * L * R);
      if (!s.left_small) atomicAdd(&g_kept_flops[1], 8.0 * k * L * R);
      // back-transformation: 8 n^2 per vector (the bulge-chasing reflectors
      // of a two-stage sector as much again)
      if (s.n > u.eig_thr)
        atomicAdd(&g_kept_flops[2], (two_stage(s.n, u.ts_min) ? 16.0 : 8.0) * k * n * n);
    }
  }
}

// Inverse iteration for the kept eigenvalues j < k of a large sector (the
// LAPACK dlagtf / dlagts / dstein organisation, restated).  Per eigenvalue,
// P (T - lam I) = L U with partial pivoting is formed once: step i keeps
// 1/U(i,i), U(i,i+1), U(i,i+2), the multiplier and the row swap.  Three solves
// U x = L^-1 P b follow, each from the previous iterate; its norm is divided
// out while the next forward pass reads it (and on output).  Clusters of
// close eigenvalues run in one lane, Gram-Schmidt against earlier members.
// The five factor arrays and the iterate (6 n doubles per lane) live in LDS
// for sectors with n <= nsplit and few kept vectors per lane; the other
// sectors keep them in their tridiagonal work area, interleaved over 128
// lanes (global address space, so that LDS waits do not wait on them).
constexpr int kVecArrays = 6;
constexpr int kVecLds = 40960;        // dynamic LDS of the LDS-resident variant
constexpr int kVecLdsThreads = 64;
constexpr int kVecSerial = 4;         // at most this many vectors per LDS lane
constexpr int kVecStepLds = 8;        // solve steps whose loads are issued together (LDS)
constexpr int kVecStepGlobal = 16;    // the same for arrays in global memory
constexpr int kVecGs = 8;             // cluster vectors projected out per pass

__host__ __device__ inline int vec_lds_head(int n, int k) {  // d, e, lambda, cluster starts (bytes)
  return (16 * n + 8 * k + 4 * (k + 1) + 15) / 16 * 16;
}
__host__ __device__ inline int vec_lds_lanes(int n, int k) {
  const int room = kVecLds - vec_lds_head(n, k);
  return room <= 0 ? 0 : min(kVecLdsThreads, room / (kVecArrays * 8 * n));
}
__host__ __device__ inline bool vec_in_lds(int n, int k, int nsplit) {
  const int p = vec_lds_lanes(n, k);
  return n <= nsplit && p > 0 && k <= kVecSerial * p;
}

__device__ __forceinline__ double vec_start(int i, int j, int n) {
  return 1.0 / sqrt((double)n) * (1.0 + 1e-3 * ((i * 7919 + j * 104729) % 1000) / 1000.0);
}

// One lane's clusters: factor arrays at base + a * sa + i * si (a = 0..5).
template <int kVecStep, typename PT>
__device__ __forceinline__ void vec_lanes(PT base, int64_t sa, int si, int P, int n, int ncl,
                                          const int *cstart, const double *d, const double *e,
                                          const double *l, double tiny, gptr<double2> X) {
  PT rd = base, du = base + sa, dl = base + 2 * sa, fm = base + 3 * sa, sw = base + 4 * sa,
     b = base + 5 * sa;
  for (int c = threadIdx.x; c < ncl; c += P) {
    for (int j = cstart[c]; j < cstart[c + 1]; ++j) {
      const double lam = l[j];
      // P (T - lam I) = L U, the running row [cd, cu] carried in registers
      double cd = d[0] - lam, cu = n > 1 ? e[0] : 0.0;
      for (int i = 0; i + 1 < n; ++i) {
        const double sub = e[i], nd = d[i + 1] - lam, nu = i + 2 < n ? e[i + 1] : 0.0;
        const bool swp = fabs(cd) < fabs(sub);
        if (!swp && cd == 0.0) cd = tiny;
        const double r = 1.0 / (swp ? sub : cd);
        const double f = (swp ? cd : sub) * r;
        rd[i * si] = r;
        du[i * si] = swp ? nd : cu;
        dl[i * si] = swp ? nu : 0.0;
        fm[i * si] = f;
        sw[i * si] = swp ? 1.0 : 0.0;
        const double ncd = swp ? cu - f * nd : nd - f * cu;
        cu = swp ? -f * nu : nu;
        cd = ncd;
      }
      if (cd == 0.0) cd = tiny;
      rd[(n - 1) * si] = 1.0 / cd;
      for (int i = 0; i < n; ++i) b[i * si] = vec_start(i, j, n);
      double scale = 1.0;
      for (int it = 0; it < 3; ++it) {
        // y = L^-1 P (scale b), in place; steps in blocks of kVecStep whose
        // loads are all issued before the block's dependent chain
        double cb = b[0] * scale;
        for (int i0 = 0; i0 + 1 < n; i0 += kVecStep) {
          double nbv[kVecStep], fv[kVecStep], sv[kVecStep];
#pragma unroll
          for (int q = 0; q < kVecStep; ++q) {
            const int i = min(i0 + q, n - 2);
            nbv[q] = b[(i + 1) * si];
            fv[q] = fm[i * si];
            sv[q] = sw[i * si];
          }
#pragma unroll
          for (int q = 0; q < kVecStep; ++q) {
            const int i = i0 + q;
            if (i + 1 < n) {
              const double nb = nbv[q] * scale;
              const bool swp = sv[q] != 0.0;
              b[i * si] = swp ? nb : cb;
              cb = swp ? cb - fv[q] * nb : nb - fv[q] * cb;
            }
          }
        }
        // x = U^-1 y, in place, with |x|^2
        double x1 = cb * rd[(n - 1) * si], x2 = 0.0, nr = x1 * x1;
        b[(n - 1) * si] = x1;
        for (int i0 = n - 2; i0 >= 0; i0 -= kVecStep) {
          double yv[kVecStep], uv[kVecStep], lv[kVecStep], rv[kVecStep];
#pragma unroll
          for (int q = 0; q < kVecStep; ++q) {
            const int i = max(i0 - q, 0);
            yv[q] = b[i * si];
            uv[q] = du[i * si];
            lv[q] = dl[i * si];
            rv[q] = rd[i * si];
          }
#pragma unroll
          for (int q = 0; q < kVecStep; ++q) {
            const int i = i0 - q;
            if (i >= 0) {
              const double x = (yv[q] - uv[q] * x1 - lv[q] * x2) * rv[q];
              b[i * si] = x;
              nr += x * x;
              x2 = x1;
              x1 = x;
            }
          }
        }
        if (j > cstart[c]) {
          // Gram-Schmidt against the cluster's earlier (unit) vectors, kVecGs
          // at a time: all their projections in one pass over the rows (the
          // loads of a row are independent), then one combined update.  Block
          // classical Gram-Schmidt loses orthogonality like kappa^2 where the
          // modified form loses kappa (and the iterate of a tight cluster is
          // nearly inside the earlier vectors' span), so it runs twice
          // ("twice is enough")
          for (int pass = 0; pass < 2; ++pass)
          for (int qb = cstart[c]; qb < j; qb += kVecGs) {
            const int nq = min(kVecGs, j - qb);
            double dot[kVecGs];
#pragma unroll
            for (int t = 0; t < kVecGs; ++t) dot[t] = 0.0;
            for (int i = 0; i < n; ++i) {
              const double bi = b[i * si];
#pragma unroll
              for (int t = 0; t < kVecGs; ++t)
                if (t < nq) dot[t] += X[(int64_t)(qb + t) * n + i].x * bi;
            }
            for (int i = 0; i < n; ++i) {
              double bi = b[i * si];
#pragma unroll
              for (int t = 0; t < kVecGs; ++t)
                if (t < nq) bi -= dot[t] * X[(int64_t)(qb + t) * n + i].x;
              b[i * si] = bi;
            }
          }
          nr = 0.0;
          for (int i = 0; i < n; ++i) nr += b[i * si] * b[i * si];
        }
        scale = nr > 0.0 ? 1.0 / sqrt(nr) : 0.0;
      }
      for (int i = 0; i < n; ++i) X[(int64_t)j * n + i] = make_double2(b[i * si] * scale, 0.0);
    }
  }}

template <bool kInLds>
__global__ __launch_bounds__(kInLds ? kVecLdsThreads : kTriVecThreads) void k_tri_vectors(
    const UpdDesc *__restrict__ Ud, WorkSeg ws, int nsplit) {
  extern __shared__ double tvs[];  // d, e (2 n), lambda (k), cluster starts (k + 1), lane arrays
  __shared__ int ncl;
  __shared__ double sh_tn[2][kTriVecThreads / 64];
  const int2 sg = *ws.seg;
  if ((int)blockIdx.x >= sg.y) return;  // a launch sized for the group's capacity
  const int2 ref = ws.list[sg.x + blockIdx.x];
  const UpdDesc &u = Ud[ref.x];
  const Sector s = u.sectors[ref.y];
  const int n = s.n, k = u.kcount[ref.y];
  if (k == 0 || vec_in_lds(n, k, nsplit) != kInLds) return;  // nothing kept, or the other variant's
  double *d = tvs, *e = tvs + n, *l = tvs + 2 * n;
  int *cstart = reinterpret_cast<int *>(tvs + 2 * n + k);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    d[i] = u.td[s.s_off + i];
    e[i] = u.te[s.s_off + i];
  }
  for (int j = threadIdx.x; j < k; j += blockDim.x) l[j] = u.lam[s.s_off + j];
  __syncthreads();
  double2 *X = s.left_small ? reinterpret_cast<double2 *>(u.V) + s.v_off
                            : reinterpret_cast<double2 *>(u.W) + s.w_off;  // column j: X[j n + r]
  // |T| bounds (row sums) by a block max
  double tn = 0.0, tn1 = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double ei = i + 1 < n ? fabs(e[i]) : 0.0;
    tn = fmax(tn, fabs(d[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0) + ei);
    tn1 = fmax(tn1, fabs(d[i]) + 2.0 * ei);
  }
  for (int o = 32; o >= 1; o >>= 1) {
    tn = fmax(tn, __shfl_xor(tn, o));
    tn1 = fmax(tn1, __shfl_xor(tn1, o));
  }
  if ((threadIdx.x & 63) == 0) {
    sh_tn[0][threadIdx.x >> 6] = tn;
    sh_tn[1][threadIdx.x >> 6] = tn1;
  }
  __syncthreads();
  tn = tn1 = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    tn = fmax(tn, sh_tn[0][w]);
    tn1 = fmax(tn1, sh_tn[1][w]);
  }
  // clusters of close eigenvalues are orthogonalised explicitly.  Outside
  // them inverse iteration leaves an overlap of about eps |T| / gap between
  // vectors, which enters the truncated state weighted by the singular value
  // sigma = sqrt(lambda).  The cluster test gap <= kClusterRel sqrt(|T| lambda)
  // (lambda the larger of the pair) keeps that weighted overlap below
  // 2e-10 sigma_max for every pair: kClusterRel |T| at the top of the
  // spectrum, tighter below it, so that the many small kept eigenvalues, a
  // small absolute gap apart, do not chain into one long serial cluster
  // (LAPACK's stein uses 1e-3 |T| throughout)
  for (int j = threadIdx.x; j < k; j += blockDim.x)
    cstart[j] = (j == 0 || fabs(l[j - 1] - l[j]) > kClusterRel * sqrt(tn * fmax(l[j - 1], 0.0)))
                    ? 1 : 0;  // cluster head flags
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0;
    for (int j = 0; j < k; ++j)
      if (cstart[j]) cstart[c++] = j;  // compaction in place (c <= j)
    cstart[c] = k;
    ncl = c;
  }
  __syncthreads();
  const double tiny = DBL_EPSILON * fmax(tn1, DBL_MIN);

  int P, si;
  int64_t sa;
  double *base;
  if (kInLds) {
    P = vec_lds_lanes(n, k);
    base = reinterpret_cast<double *>(reinterpret_cast<char *>(tvs) + vec_lds_head(n, k)) +
           (int64_t)threadIdx.x * kVecArrays * n;
    sa = n;
    si = 1;
  } else {
    P = min(n, (int)blockDim.x);
    base = u.triwork + s.tw_off + threadIdx.x;
    sa = (int64_t)n * P;
    si = P;
  }
  if ((int)threadIdx.x >= P) return;
  if constexpr (kInLds)
    vec_lanes<kVecStepLds>(base, sa, si, P, n, ncl, cstart, d, e, l, tiny, gp(X));
  else
    vec_lanes<kVecStepGlobal>(gp(base), sa, si, P, n, ncl, cstart, d, e, l, tiny, gp(X));
}

// largest n whose LDS variant holds at least one lane (k = n at most)
static int vec_lds_nsplit() {
  int n = 1;
  while (n < kTriMaxN && vec_lds_lanes(n + 1, n + 1) > 0) ++n;
  return n;
}


static inline const UpdDesc *gdesc(const GroupHeader &g, const char *w) {
  return reinterpret_cast<const UpdDesc *>(w + g.o_desc);
}
static inline const int2 *gseg(const GroupHeader &g, const char *w, int seg) {
  return reinterpret_cast<const GroupCounts *>(w + g.o_counts)->seg + seg;
}
static inline WorkSeg wseg(const GroupHeader &g, const char *w, int64_t o_list, int seg) {
  return WorkSeg{reinterpret_cast<const int2 *>(w + o_list), gseg(g, w, seg)};
}

// Side stream (default priority, non-blocking) with a fork and a join event,
// one per launching stream (so that groups run on different streams do not
// share one), created on first use and kept for the life of the process.
struct SideStream {
  hipStream_t parent, stream;
  hipEvent_t fork, join;
};
static int side_stream(hipStream_t parent, SideStream **out) {
  constexpr int kMaxSides = 64;
  static std::mutex mu;
  static SideStream sides[kMaxSides];
  static int dev_of[kMaxSides];
  static int n_sides = 0;
  int dev = 0;
  BSMPO_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  for (int i = 0; i < n_sides; ++i)
    if (sides[i].parent == parent && dev_of[i] == dev) {
      *out = &sides[i];
      return 0;
    }
  if (n_sides == kMaxSides) return fail(-1, "too many launching streams for side streams");
  SideStream &x = sides[n_sides];
  x.parent = parent;
  BSMPO_HIP(hipStreamCreateWithFlags(&x.stream, hipStreamNonBlocking));
  BSMPO_HIP(hipEventCreateWithFlags(&x.fork, hipEventDisableTiming));
  BSMPO_HIP(hipEventCreateWithFlags(&x.join, hipEventDisableTiming));
  dev_of[n_sides++] = dev;
  *out = &x;
  return 0;
}

// Dynamic LDS of a Householder launch of NT-thread blocks whose largest
// sector has order nmax: v and A22 v, plus the panel's V and W up to the
// block's share of the CU's LDS at its occupancy (one 512-thread block or two
// 256-thread blocks per CU; static arrays 2.5 / 1.5 KiB); sectors whose V and
// W do not fit keep them in global memory.
template <int NT>
static int tri_lds(int nmax) {
  constexpr int share = (160 * 1024 / (NT >= 512 ? 1 : 2) - (NT >= 512 ? 2560 : 1536)) & ~255;
  const int64_t want = tri_lds_need(nmax, true);
  return (int)std::max<int64_t>(tri_lds_need(nmax, false), std::min<int64_t>(want, share));
}

static int tri_limits(const GroupHeader &g) {
  if (g.tri_max_n > kTriMaxN) return fail(-1, "sector too large for the tridiagonal route (n > 2048)");
  if (int rc = ensure_lds_limit((const void *)k_tridiag<256, 2, kTriSW, kTriOcc>, tri_lds<256>(512)))
    return rc;
  return ensure_lds_limit((const void *)k_tridiag<512, 4, kTriSW, kTriOcc>, tri_lds<512>(kTriMaxN));
}

// One of the two Householder launches of a device-planned group (its
// capacities for counts): part 0 the sectors above the block-shape split on
// 512-thread blocks, part 1 the rest on 256-thread blocks, each with a grid
// for the whole list; the kernels read the split (SG_TRIBIG, SG_TRISMALL)
// from the device.  A layer program makes them two independent graph nodes.
int launch_tridiag_part(const GroupHeader &g, const char *w, hipStream_t stream, int part) {
  if (g.n_tri == 0) return 0;
  if (!g.dev_planned || g.n_ts != 0) return fail(-1, "tridiag part: a device-planned one-stage group only");
  if (int rc = tri_limits(g)) return rc;
  if (part == 0) {
    const int lds = tri_lds<512>(g.tri_max_n);
    hipLaunchKernelGGL((k_tridiag<512, 4, kTriSW, kTriOcc>), dim3(g.n_tri), dim3(512), (size_t)lds,
                       stream, gdesc(g, w), wseg(g, w, g.o_tri, SG_TRIBIG), 1, kTriMaxN, lds);
  } else {
    const int lds = tri_lds<256>(std::min(g.tri_split, g.tri_max_n));
    hipLaunchKernelGGL((k_tridiag<256, 2, kTriSW, kTriOcc>), dim3(g.n_tri), dim3(256), (size_t)lds,
                       stream, gdesc(g, w), wseg(g, w, g.o_tri, SG_TRISMALL), 1, g.tri_split, lds);
  }
  BSMPO_HIP(hipGetLastError());
  return 0;
}

// After the reductions: the update's cutoff from sample counts (tridiagonal
// route) and exact values (Jacobi route), then the eigenvalues above half of it.
int launch_tri_sturm(const GroupHeader &g, const char *w, hipStream_t stream) {
  if (g.n_tri == 0) return 0;
  const int tk = timing_begin(ST_BISECT, stream);
  hipLaunchKernelGGL(k_hist_zero, dim3(g.n_upd), dim3(256), 0, stream, gdesc(g, w),
                     gseg(g, w, SG_UPD));
  BSMPO_HIP(hipGetLastError());
  const int njac = g.jac_off[kTriClass];
  if (njac > 0) {
    hipLaunchKernelGGL(k_hist_jacobi, dim3(njac), dim3(64), 0, stream, gdesc(g, w),
                       wseg(g, w, g.o_jac, SG_JACALL));
    BSMPO_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_bisect, dim3(g.n_tri, 1), dim3(256), (size_t)g.tri_max_n * 16, stream,
                     gdesc(g, w), wseg(g, w, g.o_tri, SG_TRI), 0);
  BSMPO_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_cutoff, dim3(g.n_upd), dim3(256), 0, stream, gdesc(g, w),
                     gseg(g, w, SG_UPD));
  BSMPO_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_bisect, dim3(g.n_tri, (g.tri_max_n + 255) / 256), dim3(256),
                     (size_t)g.tri_max_n * 16, stream, gdesc(g, w), wseg(g, w, g.o_tri, SG_TRI), 1);
  BSMPO_HIP(hipGetLastError());
  timing_end(tk, stream, 0.0, 0.0);
  return 0;
}

int launch_tri_values(const GroupHeader &g, const char *w, hipStream_t stream) {
  if (g.n_tri == 0) return 0;
  if (int rc = tri_limits(g)) return rc;
  int tk = timing_begin(ST_TRIDIAG, stream);
  // The list holds the two-stage sectors first, then the one-stage ones, each
  // part largest first.  The two-stage sectors (band reduction, then bulge
  // chasing) run on the launching stream.  One-stage sectors: with few of them
  // (a single update), the larger block shape for every sector, so that each
  // column chain gets more threads; with many (a group across instances), the
  // sectors above the split (bsmpo_set_tri_split, default 256) take the
  // 512-thread shape and the rest 256-thread blocks, more per CU.  Work beside
  // the launching stream's goes to a side stream forked before the first
  // launch (fork / join by events), so that it fills the slots the long
  // chains leave.  A device-planned group launches both shapes with grids for
  // its capacity (see launch_tridiag_part).
  const int n1 = g.n_tri - g.n_ts;
  const bool few = !g.dev_planned && n1 < kTriFewSectors;
  const int nbig = few || g.dev_planned ? n1 : g.tri_nbig;  // grid of the 512-thread launch
  const int nsmall = g.dev_planned ? n1 : n1 - nbig;         // grid of the 256-thread launch
  const int split = few ? kTriMaxN : g.tri_split;
  const bool big_on_side = few && g.n_ts > 0;
  const bool main_work = g.n_ts > 0 || (nbig > 0 && !big_on_side);
  const bool side_work = nsmall > 0 || (big_on_side && nbig > 0);
  const bool both = main_work && side_work;
  SideStream *ss = nullptr;
  if (both) {
    int rc = side_stream(stream, &ss);
    if (rc) return rc;
    BSMPO_HIP(hipEventRecord(ss->fork, stream));
    BSMPO_HIP(hipStreamWaitEvent(ss->stream, ss->fork, 0));
  }
  const hipStream_t sst = both ? ss->stream : stream;
  if (int rc = launch_band(g, w, stream)) return rc;
  if (nbig > 0) {
    const int lds = tri_lds<512>(g.tri_max_n);
    hipLaunchKernelGGL((k_tridiag<512, 4, kTriSW, kTriOcc>), dim3(nbig), dim3(512), (size_t)lds,
                       big_on_side ? sst : stream, gdesc(g, w), wseg(g, w, g.o_tri, SG_TRIBIG), 1,
                       kTriMaxN, lds);
    BSMPO_HIP(hipGetLastError());
  }
  if (nsmall > 0) {
    const int lds = tri_lds<256>(std::min(split, g.tri_max_n));
    hipLaunchKernelGGL((k_tridiag<256, 2, kTriSW, kTriOcc>), dim3(nsmall), dim3(256), (size_t)lds,
                       sst, gdesc(g, w), wseg(g, w, g.o_tri, SG_TRISMALL), 1, split, lds);
    BSMPO_HIP(hipGetLastError());
  }
  if (both) {
    BSMPO_HIP(hipEventRecord(ss->join, ss->stream));
    BSMPO_HIP(hipStreamWaitEvent(stream, ss->join, 0));
  }
  timing_end(tk, stream, g.flops_tridiag, g.bytes_tridiag);
  return launch_tri_sturm(g, w, stream);
}

int launch_kcount(const GroupHeader &g, const char *w, hipStream_t stream) {
  if (g.n_sec == 0) return 0;
  hipLaunchKernelGGL(k_kcount, dim3(g.n_sec), dim3(256), 0, stream, gdesc(g, w),
                     wseg(g, w, g.o_sec, SG_SEC), timing_on() ? 1 : 0);
  BSMPO_HIP(hipGetLastError());
  return 0;
}

int launch_tri_vectors(const GroupHeader &g, const char *w, hipStream_t stream) {
  if (g.n_tri == 0) return 0;
  int tk = timing_begin(ST_TRIVEC, stream);
  static const int nsplit = vec_lds_nsplit();
  hipLaunchKernelGGL((k_tri_vectors<true>), dim3(g.n_tri), dim3(kVecLdsThreads), (size_t)kVecLds,
                     stream, gdesc(g, w), wseg(g, w, g.o_tri, SG_TRI), nsplit);
  BSMPO_HIP(hipGetLastError());
  // sectors past nsplit, or with more kept vectors than the LDS lanes take
  hipLaunchKernelGGL((k_tri_vectors<false>), dim3(g.n_tri), dim3(kTriVecThreads),
                     (size_t)vec_lds_head(g.tri_max_n, g.tri_max_n), stream, gdesc(g, w),
                     wseg(g, w, g.o_tri, SG_TRI), nsplit);
  BSMPO_HIP(hipGetLastError());
  timing_end(tk, stream, 0.0, 0.0);
  tk = timing_begin(ST_BACKTRANS, stream);
  if (int rc = launch_bulge_apply(g, w, stream)) return rc;  // Q2 first (two-stage sectors)
  if (int rc = launch_wy_backtransform(g, w, stream)) return rc;
  timing_end(tk, stream, 0.0, 0.0);
  return 0;
}

}  // namespace bsmpo
