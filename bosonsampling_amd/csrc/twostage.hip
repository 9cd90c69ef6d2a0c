// Two-stage tridiagonalisation of the large Gram matrices of the tridiagonal
// route (sectors with n >= the two-stage threshold), FP64.  The one-stage
// reduction reads the trailing matrix once per column (the product A22 v);
// here the matrix is first reduced to a band of width kBand by panels of
// kBand columns, each read with MFMA products twice (the product A22 V and the
// two-sided update), and the band is then reduced to real tridiagonal form by
// bulge chasing.  The eigenvectors of G are x = Q1 Q2 y for the tridiagonal
// eigenvectors y.
//
//   k_band         stage 1, one workgroup per sector: per panel, Householder
//                  QR of the m x kBand block below the band (rows in
//                  registers; one block reduction per reflector), the compact
//                  WY factor T, then Y = A22 V, Z = V^H Y, W = Y T - V (T^H Z T) / 2
//                  and A22 -= V W^H + W V^H on the FP64 MFMA over the full
//                  (both triangles) trailing matrix; finally the band is copied
//                  to the sector's work area.  Reflector i (column i) has its
//                  unit entry at row i + kBand and the rest below it in A.
//   k_bulge        stage 2, one workgroup per sector: sweep j annihilates
//                  column j below the subdiagonal with a reflector on rows
//                  j + 1 .. j + kBand, then chases the bulge down kBand rows at a
//                  time (right-apply the previous reflector to the block below,
//                  annihilate its first column, left-apply, two-sided update of
//                  the diagonal block).  A wave runs one sweep at a time; sweep
//                  j may take its step k once sweep j - 1 has done k + 2 steps,
//                  which keeps the blocks of the steps in flight disjoint, so the
//                  waves work on several sweeps in lockstep (one barrier per
//                  step).  Band storage: offsets 0 .. 2 kBand - 1 below the
//                  diagonal per column (band plus the bulge).  Sweep j's
//                  reflectors fill rows 0 .. n - 2 - j of column n - 1 - j of A
//                  (the upper triangle, free after stage 1): the first slot of
//                  each holds tau, the others v (the unit entry is implicit).
//   k_bulge_apply  Z <- Q2 Z for the kept vectors: sweeps last to first; the
//                  reflectors of one sweep act on disjoint rows and are applied
//                  together.  Vectors staged in LDS in chunks, the next sweep's
//                  reflectors loaded while the current one is applied.
// The WY back-transformation then applies Q1 (reflector offset kBand).
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "common.h"

namespace bsmpo {

constexpr int kBandLd = 2 * kBand;  // band storage per column: offsets 0 .. 2 kBand - 1
constexpr int kBandWaves = kBandThreads / 64;
constexpr int kBulgeWaves = 8;
constexpr int kBcThreads = 256;

// Householder generator (LAPACK zlarfg form, as in the one-stage reduction):
// H^H (alpha; x) = (beta; 0), beta real, H = I - tau v v^H, v = (1; x scale).
struct Refl {
  double2 tau, scale;
  double beta;
};
__device__ __forceinline__ Refl refl_make(double2 alpha, double xn2) {
  Refl R;
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

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zmulc(double2 a, double2 b) {  // conj(a) b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 zsub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 zshfl(double2 a, int src) {
  return make_double2(__shfl(a.x, src), __shfl(a.y, src));
}
__device__ __forceinline__ double2 zxor(double2 a, int m) {
  return make_double2(a.x + __shfl_xor(a.x, m), a.y + __shfl_xor(a.y, m));
}

// Work area of a two-stage sector (complex elements): the band (n x kBandLd),
// then V, Y, W of the current panel (m x kBand each, row-major, m < n).
struct BandWork {
  gptr<double2> band, Vr, Yr, Wr;
};
__device__ __forceinline__ BandWork band_work(const UpdDesc &u, const Sector &s) {
  double2 *base = reinterpret_cast<double2 *>(u.triwork + s.tw_off);
  const int64_t n = s.n;
  BandWork b;
  b.band = gp(base);
  b.Vr = gp(base + kBandLd * n);
  b.Yr = gp(base + (kBandLd + kBand) * n);
  b.Wr = gp(base + (kBandLd + 2 * kBand) * n);
  return b;
}

// x[0..31] summed over the 64 lanes of a wave by recursive halving: lanes l
// and l ^ 1 return the total of x[(l >> 1) & 31].
__device__ __forceinline__ double wave_sum32(double (&x)[32], int lane) {
  double y16[16], y8[8], y4[4], y2[2];
  const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4, b1 = lane & 2;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const double snd = b5 ? x[q] : x[q + 16], kp = b5 ? x[q + 16] : x[q];
    y16[q] = kp + __shfl_xor(snd, 32);
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double snd = b4 ? y16[q] : y16[q + 8], kp = b4 ? y16[q + 8] : y16[q];
    y8[q] = kp + __shfl_xor(snd, 16);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double snd = b3 ? y8[q] : y8[q + 4], kp = b3 ? y8[q + 4] : y8[q];
    y4[q] = kp + __shfl_xor(snd, 8);
  }
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const double snd = b2 ? y4[q] : y4[q + 2], kp = b2 ? y4[q + 2] : y4[q];
    y2[q] = kp + __shfl_xor(snd, 4);
  }
  const double snd = b1 ? y2[0] : y2[1], kp = b1 ? y2[1] : y2[0];
  double y = kp + __shfl_xor(snd, 2);
  return y + __shfl_xor(y, 1);
}

// ------------------------------------------------------------- stage 1 --
template <int RPT>
__global__ __launch_bounds__(kBandThreads) void k_band(const UpdDesc *__restrict__ Ud, WorkSeg ws) {
  constexpr int NT = kBandThreads, NW = kBandWaves;
  __shared__ double red[NW][32];
  __shared__ double fin[32];
  __shared__ double2 rowk[kBand];
  __shared__ double2 Ts[kBand][kBand + 1];
  __shared__ double2 Zs[kBand][kBand + 1];
  __shared__ double2 Ms[kBand][kBand + 1];
  __shared__ double2 Bw[2 * kBand][kBand + 1];  // [T; -(T^H Z T) / 2]
  __shared__ double mxs[NW];
  const int2 sg = *ws.seg;
  if ((int)blockIdx.x >= sg.y) return;
  const int2 ref = ws.list[sg.x + blockIdx.x];
  const UpdDesc &u = Ud[ref.x];
  const Sector s = u.sectors[ref.y];
  const int n = s.n;
  gptr<double2> A = gp(reinterpret_cast<double2 *>(u.G) + s.g_off);  // A(r, c) = A[c n + r]
  gptr<double> e = gp(u.te + s.s_off);
  gptr<double2> tau1 = gp(reinterpret_cast<double2 *>(u.tau) + s.s_off);
  const BandWork bw = band_work(u, s);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, i16 = lane & 15, kq = lane >> 4;

  // (0) G scaled by the power of two that brings its largest diagonal entry
  // into [1, 2) (exact; e[n - 1] keeps the factor that maps the eigenvalues
  // back, 0 for a zero matrix), and mirrored into the upper triangle
  {
    double mx = 0.0;
    for (int i = t; i < n; i += NT) mx = fmax(mx, fabs(A[(int64_t)i * n + i].x));
    for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if (lane == 0) mxs[wave] = mx;
    __syncthreads();
    mx = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) mx = fmax(mx, mxs[w]);
    const int ex = (mx > 0.0 && isfinite(mx)) ? ilogb(mx) : 0;
    const double scl = ldexp(1.0, -ex);
    if (t == 0) e[n - 1] = mx > 0.0 ? ldexp(1.0, ex) : 0.0;
    for (int i = t; i < n; i += NT) tau1[i] = make_double2(0.0, 0.0);
    const int nts = (n + 15) / 16, ntl = nts * (nts + 1) / 2;
    for (int tix = wave; tix < ntl; tix += NW) {
      int ti = (int)((sqrt(8.0 * tix + 1.0) - 1.0) * 0.5);
      while ((ti + 1) * (ti + 2) / 2 <= tix) ++ti;
      while (ti * (ti + 1) / 2 > tix) --ti;
      const int tj = tix - ti * (ti + 1) / 2;
      const int r = 16 * ti + i16;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = 16 * tj + kq + 4 * q;
        if (r < n && c <= r) {
          double2 a = A[(int64_t)c * n + r];
          a.x *= scl;
          a.y *= scl;
          A[(int64_t)c * n + r] = a;
          if (r > c) A[(int64_t)r * n + c] = zconj(a);
        }
      }
    }
    __syncthreads();
  }

  for (int j0 = 0; n - j0 - kBand >= 2; j0 += kBand) {
    const int m = n - j0 - kBand, rg = j0 + kBand;  // panel rows rg .. n - 1 (local 0 .. m - 1)
    const int nb = min(kBand, m - 1);               // reflectors of this panel
    for (int x = t; x < kBand * (kBand + 1); x += NT) {
      (&Ts[0][0])[x] = make_double2(0.0, 0.0);
      (&Zs[0][0])[x] = make_double2(0.0, 0.0);
    }
    // (a) the panel in registers: thread t holds local rows t + NT i
    double2 P[RPT][kBand];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = t + NT * i;
#pragma unroll
      for (int c = 0; c < kBand; ++c)
        P[i][c] = r < m ? A[(int64_t)(j0 + c) * n + rg + r] : make_double2(0.0, 0.0);
    }
    // (b) Householder QR, one block reduction per reflector.  With x the
    //     column below the diagonal (alpha = x_k) and v = scale (x - beta e_k):
    //     v^H P(:, c) = conj(scale) (x^H P(:, c) - beta P(k, c))   (c > k)
    //     V_l^H v    = scale (V_l^H x - beta conj(P(k, l)))        (l < k)
    //     so the sums x^H P(:, c), V_l^H x and |x below k|^2 (slot k) are
    //     reduced together before the reflector is known.
    for (int k = 0; k < nb; ++k) {
      if (t == k) {
#pragma unroll
        for (int c = 0; c < kBand; ++c) rowk[c] = P[0][c];
      }
      // column k of the panel rows (k is a run-time index: select by compare,
      // so that P stays in registers)
      double2 xk[RPT];
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        xk[i] = P[i][0];
#pragma unroll
        for (int c = 1; c < kBand; ++c)
          if (c == k) xk[i] = P[i][c];
      }
      double q[32];
#pragma unroll
      for (int c = 0; c < 32; ++c) q[c] = 0.0;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int r = t + NT * i;
        if (r >= k && r < m) {
          const double2 x = xk[i];
#pragma unroll
          for (int c = 0; c < kBand; ++c) {
            if (c == k) {
              if (r > k) q[2 * c] += x.x * x.x + x.y * x.y;
            } else if (c > k) {
              const double2 p = zmulc(x, P[i][c]);
              q[2 * c] += p.x;
              q[2 * c + 1] += p.y;
            } else {
              const double2 p = zmulc(P[i][c], x);
              q[2 * c] += p.x;
              q[2 * c + 1] += p.y;
            }
          }
        }
      }
      const double qs = wave_sum32(q, lane);
      if ((lane & 1) == 0) red[wave][lane >> 1] = qs;
      __syncthreads();
      if (t < 32) {
        double a = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) a += red[w][t];
        fin[t] = a;
      }
      __syncthreads();
      const double2 alpha = rowk[k];
      const Refl R = refl_make(alpha, fin[2 * k]);
      const double2 ctau = zconj(R.tau), cscale = zconj(R.scale);
      double2 v[RPT], nk[RPT];  // v, and the new column k (beta at row k, v below)
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int r = t + NT * i;
        v[i] = make_double2(0.0, 0.0);
        nk[i] = xk[i];
        if (r >= k && r < m) {
          if (r == k) {
            v[i] = make_double2(1.0, 0.0);
            nk[i] = make_double2(R.beta, 0.0);
          } else {
            v[i] = zmul(xk[i], R.scale);
            nk[i] = v[i];
          }
        }
      }
      // P(:, c) -= conj(tau) v w_c,  w_c = v^H P(:, c)  (c > k); column k replaced
#pragma unroll
      for (int c = 0; c < kBand; ++c) {
        if (c == k) {
#pragma unroll
          for (int i = 0; i < RPT; ++i) P[i][c] = nk[i];
        } else if (c > k) {
          const double2 qc = make_double2(fin[2 * c], fin[2 * c + 1]);
          const double2 rk = rowk[c];
          const double2 wc =
              zmul(ctau, zmul(cscale, make_double2(qc.x - R.beta * rk.x, qc.y - R.beta * rk.y)));
#pragma unroll
          for (int i = 0; i < RPT; ++i) P[i][c] = zsub(P[i][c], zmul(v[i], wc));
        }
      }
      // column k of T: T(k, k) = tau, T(l, k) = -tau sum_{p = l}^{k - 1} T(l, p) (V_p^H v)
      if (t < k) {
        double2 acc = make_double2(0.0, 0.0);
        for (int p = t; p < k; ++p) {
          const double2 qp = make_double2(fin[2 * p], fin[2 * p + 1]);
          const double2 rk = rowk[p];
          const double2 sp =
              zmul(R.scale, make_double2(qp.x - R.beta * rk.x, qp.y + R.beta * rk.y));
          acc = zadd(acc, zmul(Ts[t][p], sp));
        }
        Ts[t][k] = make_double2(-(R.tau.x * acc.x - R.tau.y * acc.y),
                                -(R.tau.x * acc.y + R.tau.y * acc.x));
      } else if (t == k) {
        Ts[k][k] = R.tau;
        tau1[j0 + k] = R.tau;
      }
      __syncthreads();  // rowk and fin are rewritten by the next reflector
    }
    // (c) the panel back to A (R on and above row k of column k, reflectors
    //     below), V row-major with its unit diagonal and zeros above
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = t + NT * i;
      if (r < m) {
#pragma unroll
        for (int c = 0; c < kBand; ++c) {
          A[(int64_t)(j0 + c) * n + rg + r] = P[i][c];
          const double2 v = c < nb ? (r > c ? P[i][c]
                                            : (r == c ? make_double2(1.0, 0.0) : make_double2(0.0, 0.0)))
                                   : make_double2(0.0, 0.0);
          bw.Vr[(int64_t)r * kBand + c] = v;
        }
      }
    }
    __syncthreads();
    // (d) Y = A22 V (row tiles in pairs round the waves, FP64 MFMA over
    //     K = m), stored row-major, and this wave's part of Z = V^H Y
    dbl4 zr = {0.0, 0.0, 0.0, 0.0}, zi = {0.0, 0.0, 0.0, 0.0};
    for (int rb = 32 * wave; rb < m; rb += 32 * NW) {
      dbl4 y0r = {0.0, 0.0, 0.0, 0.0}, y0i = {0.0, 0.0, 0.0, 0.0};
      dbl4 y1r = {0.0, 0.0, 0.0, 0.0}, y1i = {0.0, 0.0, 0.0, 0.0};
      const int ra = rb + i16, rb2 = rb + 16 + i16;
      for (int k0 = 0; k0 < m; k0 += 16) {
        cplx a0[4], a1[4], bv[4];
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const int kk = k0 + 4 * st + kq;
          const bool ok = kk < m;
          const int64_t col = (int64_t)(rg + kk) * n + rg;
          const double2 x0 = (ok && ra < m) ? A[col + ra] : make_double2(0.0, 0.0);
          const double2 x1 = (ok && rb2 < m) ? A[col + rb2] : make_double2(0.0, 0.0);
          const double2 vv = ok ? bw.Vr[(int64_t)kk * kBand + i16] : make_double2(0.0, 0.0);
          a0[st] = cplx{x0.x, x0.y};
          a1[st] = cplx{x1.x, x1.y};
          bv[st] = cplx{vv.x, vv.y};
        }
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          cmfma_f64(y0r, y0i, a0[st], bv[st]);
          cmfma_f64(y1r, y1i, a1[st], bv[st]);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row0 = rb + mfma64_row(kq, r), row1 = row0 + 16;
        if (row0 < m) bw.Yr[(int64_t)row0 * kBand + i16] = make_double2(y0r[r], y0i[r]);
        if (row1 < m) bw.Yr[(int64_t)row1 * kBand + i16] = make_double2(y1r[r], y1i[r]);
      }
      // Z += V_tile^H Y_tile: K runs over the tile's rows, Y's accumulator
      // element s is its row kq + 4 s, the B operand of step s as it stands
#pragma unroll
      for (int sgl = 0; sgl < 2; ++sgl) {
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const int rv = rb + 16 * sgl + 4 * st + kq;
          const double2 vv = rv < m ? bw.Vr[(int64_t)rv * kBand + i16] : make_double2(0.0, 0.0);
          const cplx b = sgl == 0 ? cplx{y0r[st], y0i[st]} : cplx{y1r[st], y1i[st]};
          cmfma_f64(zr, zi, cplx{vv.x, -vv.y}, b);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      atomicAdd(&Zs[mfma64_row(kq, r)][i16].x, zr[r]);
      atomicAdd(&Zs[mfma64_row(kq, r)][i16].y, zi[r]);
    }
    __syncthreads();
    // (e) M = Z T, then B_W = [T; -(T^H M) / 2]
    if (t < kBand * kBand) {
      const int a = t / kBand, b = t % kBand;
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int l = 0; l < kBand; ++l) acc = zadd(acc, zmul(Zs[a][l], Ts[l][b]));
      Ms[a][b] = acc;
    }
    __syncthreads();
    if (t < kBand * kBand) {
      const int a = t / kBand, b = t % kBand;
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int l = 0; l < kBand; ++l) acc = zadd(acc, zmulc(Ts[l][a], Ms[l][b]));
      Bw[a][b] = Ts[a][b];
      Bw[kBand + a][b] = make_double2(-0.5 * acc.x, -0.5 * acc.y);
    }
    __syncthreads();
    // (f) W = [Y V] B_W per 16-row tile (K = 32)
    for (int rb = 16 * wave; rb < m; rb += 16 * NW) {
      dbl4 wr = {0.0, 0.0, 0.0, 0.0}, wi = {0.0, 0.0, 0.0, 0.0};
      const int ra = rb + i16;
      cplx a[8], b[8];
#pragma unroll
      for (int st = 0; st < 8; ++st) {
        const int kk = 4 * (st & 3) + kq;
        const double2 x = ra < m ? (st < 4 ? bw.Yr[(int64_t)ra * kBand + kk]
                                           : bw.Vr[(int64_t)ra * kBand + kk])
                                 : make_double2(0.0, 0.0);
        a[st] = cplx{x.x, x.y};
        const double2 y = Bw[4 * st + kq][i16];
        b[st] = cplx{y.x, y.y};
      }
#pragma unroll
      for (int st = 0; st < 8; ++st) cmfma_f64(wr, wi, a[st], b[st]);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb + mfma64_row(kq, r);
        if (row < m) bw.Wr[(int64_t)row * kBand + i16] = make_double2(wr[r], wi[r]);
      }
    }
    __syncthreads();
    // (g) A22 -= V W^H + W V^H over every 16 x 16 tile, formed transposed,
    //     U^T = [conj W  conj V](c, :) [V ; W](r, :), so that a lane's
    //     elements run down a column of A; two tiles per pass, all loads first
    {
      const int nts = (m + 15) / 16, ntiles = nts * nts;
      constexpr int kTT = 1;  // two tiles per pass spill (the panel's registers)
      for (int tix0 = wave; tix0 < ntiles; tix0 += kTT * NW) {
        int R0[kTT], C0[kTT];
        bool live[kTT];
        double2 av[kTT][4];
        cplx a[kTT][8], b[kTT][8];
#pragma unroll
        for (int uu = 0; uu < kTT; ++uu) {
          const int tix = tix0 + uu * NW;
          live[uu] = tix < ntiles;
          const int tq = live[uu] ? tix : 0;
          R0[uu] = 16 * (tq / nts);
          C0[uu] = 16 * (tq % nts);
          const int r = R0[uu] + i16;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int c = C0[uu] + mfma64_row(kq, q);
            av[uu][q] = (live[uu] && r < m && c < m) ? A[(int64_t)(rg + c) * n + rg + r]
                                                    : make_double2(0.0, 0.0);
          }
          const int ca = C0[uu] + i16;
#pragma unroll
          for (int st = 0; st < 8; ++st) {
            const int kk = 4 * (st & 3) + kq;
            double2 x = make_double2(0.0, 0.0), y = make_double2(0.0, 0.0);
            if (live[uu]) {
              if (ca < m) x = st < 4 ? bw.Wr[(int64_t)ca * kBand + kk] : bw.Vr[(int64_t)ca * kBand + kk];
              if (r < m) y = st < 4 ? bw.Vr[(int64_t)r * kBand + kk] : bw.Wr[(int64_t)r * kBand + kk];
            }
            a[uu][st] = cplx{x.x, -x.y};
            b[uu][st] = cplx{y.x, y.y};
          }
        }
#pragma unroll
        for (int uu = 0; uu < kTT; ++uu) {
          if (!live[uu]) break;
          dbl4 ur = {0.0, 0.0, 0.0, 0.0}, ui = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int st = 0; st < 8; ++st) cmfma_f64(ur, ui, a[uu][st], b[uu][st]);
          const int r = R0[uu] + i16;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int c = C0[uu] + mfma64_row(kq, q);
            if (r < m && c < m)
              A[(int64_t)(rg + c) * n + rg + r] = make_double2(av[uu][q].x - ur[q], av[uu][q].y - ui[q]);
          }
        }
      }
    }
    __syncthreads();
  }
  // (h) the band (offsets 0 .. kBand of the lower triangle) to the work area
  for (int x = t; x < n * kBandLd; x += NT) {
    const int c = x / kBandLd, o = x % kBandLd;
    double2 v = make_double2(0.0, 0.0);
    if (o <= kBand && c + o < n) v = A[(int64_t)c * n + c + o];
    bw.band[x] = v;
  }
}

// ------------------------------------------------------------- stage 2 --
// Element mapping of a 16 x 16 block in a wave: lane l holds row l & 15,
// columns (l >> 4) + 4 i, i = 0..3.  Row sums: over i, then lanes xor 16, 32;
// column sums: lanes xor 1, 2, 4, 8.
__global__ __launch_bounds__(64 * kBulgeWaves) void k_bulge(const UpdDesc *__restrict__ Ud, WorkSeg ws) {
  constexpr int NWv = kBulgeWaves;
  __shared__ int prog[2][NWv][2];  // per wave: current sweep, steps done in it
  const int2 sg = *ws.seg;
  if ((int)blockIdx.x >= sg.y) return;
  const int2 ref = ws.list[sg.x + blockIdx.x];
  const UpdDesc &u = Ud[ref.x];
  const Sector s = u.sectors[ref.y];
  const int n = s.n;
  gptr<double2> A = gp(reinterpret_cast<double2 *>(u.G) + s.g_off);
  gptr<double> d = gp(u.td + s.s_off), e = gp(u.te + s.s_off);
  const gptr<double2> band = band_work(u, s).band;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r16 = lane & 15, g = lane >> 4;
  const int nsw = n - 1;  // sweeps 0 .. n - 2
  int j = wave, ks = 0;   // this wave's sweep and its next step
  double2 vp = make_double2(0.0, 0.0), taup = make_double2(0.0, 0.0);  // previous reflector
  int stp = 0, edp = 0;
  if (lane == 0) {
    prog[0][wave][0] = j;
    prog[0][wave][1] = 0;
  }
  __syncthreads();
  int par = 0;
  // every step that can run does (sweep 0 always can), so the sweeps finish
  // within their task count (about n^2 / (2 kBand) + n steps); the cap only
  // bounds the loop
  const int cap = 4 * n * n + 1024;
  for (int iter = 0; iter < cap; ++iter) {
    bool done = true;
#pragma unroll
    for (int w = 0; w < NWv; ++w) done = done && prog[par][w][0] >= nsw;
    if (done) break;  // uniform: every wave reads the same snapshot
    bool run = false;
    if (j < nsw) {
      if (j == 0) {
        run = true;
      } else {
        const int w2 = (j - 1) % NWv;
        const int j2 = prog[par][w2][0], k2 = prog[par][w2][1];
        run = j2 > j - 1 || (j2 == j - 1 && k2 >= ks + 2);
      }
    }
    if (run) {  // uniform over the wave
      const int st = ks == 0 ? j + 1 : edp + 1;
      const int ed = min(st + kBand - 1, n - 1), len = ed - st + 1;
      double2 x;
      double2 O[4];
      const int lenp = edp - stp + 1;
      if (ks == 0) {
        x = r16 < len ? band[(int64_t)j * kBandLd + 1 + r16] : make_double2(0.0, 0.0);
      } else {
        // block rows st .. ed, columns stp .. edp: element (r, c) at offset
        // st - stp + r - c of column stp + c
        double2 vpc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int c = g + 4 * i;
          O[i] = (r16 < len && c < lenp)
                     ? band[(int64_t)(stp + c) * kBandLd + (st - stp) + r16 - c]
                     : make_double2(0.0, 0.0);
          vpc[i] = zshfl(vp, c);
        }
        // O <- O H_p = O - tau_p (O v_p) v_p^H
        double2 sr = make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < 4; ++i) sr = zadd(sr, zmul(O[i], vpc[i]));
        sr = zxor(sr, 16);
        sr = zxor(sr, 32);
        const double2 ts = zmul(taup, sr);
#pragma unroll
        for (int i = 0; i < 4; ++i) O[i] = zsub(O[i], zmul(ts, zconj(vpc[i])));
        x = zshfl(O[0], r16);  // column 0 of the block: lanes 0..15, element 0
      }
      // reflector annihilating x below its first entry
      double xn2 = (r16 >= 1 && r16 < len) ? x.x * x.x + x.y * x.y : 0.0;
#pragma unroll
      for (int o = 1; o <= 8; o <<= 1) xn2 += __shfl_xor(xn2, o);
      const double2 alpha = zshfl(x, 0);
      const Refl R = refl_make(alpha, xn2);
      const double2 v = r16 == 0 ? make_double2(1.0, 0.0)
                                 : (r16 < len ? zmul(x, R.scale) : make_double2(0.0, 0.0));
      if (g == 0 && r16 < len) {
        const bool first = r16 == 0;
        A[(int64_t)(n - 1 - j) * n + ks * kBand + r16] =
            make_double2(first ? R.tau.x : v.x, first ? R.tau.y : v.y);
      }
      if (ks == 0) {
        if (g == 0 && r16 < len)
          band[(int64_t)j * kBandLd + 1 + r16] =
              r16 == 0 ? make_double2(R.beta, 0.0) : make_double2(0.0, 0.0);
      } else {
        // columns 1.. : O <- H^H O = O - conj(tau) v (v^H O); column 0 = (beta, 0, ..)
        const double2 ct = zconj(R.tau);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int c = g + 4 * i;
          double2 tc = zmulc(v, O[i]);
#pragma unroll
          for (int o = 1; o <= 8; o <<= 1) tc = zxor(tc, o);
          if (c == 0)
            O[i] = r16 == 0 ? make_double2(R.beta, 0.0) : make_double2(0.0, 0.0);
          else
            O[i] = zsub(O[i], zmul(ct, zmul(v, tc)));
          if (r16 < len && c < lenp)
            band[(int64_t)(stp + c) * kBandLd + (st - stp) + r16 - c] = O[i];
        }
      }
      // diagonal block rows / columns st .. ed: D <- H^H D H = D - v w^H - w v^H,
      // w = tau y - |tau|^2 (v^H y) v / 2, y = D v
      double2 D[4], vc[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = g + 4 * i;
        D[i] = make_double2(0.0, 0.0);
        if (r16 < len && c < len) {
          // lower entry (r, c), or the conjugate of (c, r)
          const bool lo = r16 >= c;
          const double2 b = band[(int64_t)(st + (lo ? c : r16)) * kBandLd + (lo ? r16 - c : c - r16)];
          D[i] = make_double2(b.x, lo ? b.y : -b.y);
        }
        vc[i] = zshfl(v, c);
      }
      double2 y = make_double2(0.0, 0.0);
#pragma unroll
      for (int i = 0; i < 4; ++i) y = zadd(y, zmul(D[i], vc[i]));
      y = zxor(y, 16);
      y = zxor(y, 32);
      double vy = v.x * y.x + v.y * y.y;  // Re(conj(v) y); v^H D v is real
#pragma unroll
      for (int o = 1; o <= 8; o <<= 1) vy += __shfl_xor(vy, o);
      const double t2 = 0.5 * (R.tau.x * R.tau.x + R.tau.y * R.tau.y) * vy;
      const double2 w = make_double2(R.tau.x * y.x - R.tau.y * y.y - t2 * v.x,
                                     R.tau.x * y.y + R.tau.y * y.x - t2 * v.y);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = g + 4 * i;
        const double2 wc = zshfl(w, c);
        D[i] = zsub(D[i], zadd(zmul(v, zconj(wc)), zmul(w, zconj(vc[i]))));
        if (r16 < len && c <= r16) band[(int64_t)(st + c) * kBandLd + r16 - c] = D[i];
      }
      vp = v;
      taup = R.tau;
      stp = st;
      edp = ed;
      ++ks;
      if (ed == n - 1) {  // sweep done
        j += NWv;
        ks = 0;
      }
    }
    if (lane == 0) {
      prog[par ^ 1][wave][0] = j;
      prog[par ^ 1][wave][1] = ks;
    }
    __syncthreads();  // this step's band writes and progress visible to every wave
    par ^= 1;
  }
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    d[c] = band[(int64_t)c * kBandLd].x;
    if (c < n - 1) e[c] = band[(int64_t)c * kBandLd + 1].x;  // e[n - 1] keeps the scale
  }
}

// ------------------------------------------------- Q2 back-transformation --
__global__ __launch_bounds__(kBcThreads) void k_bulge_apply(const UpdDesc *__restrict__ Ud, WorkSeg ws) {
  extern __shared__ double2 bsm[];
  const int2 sg = *ws.seg;
  const int2 *__restrict__ list = ws.list + sg.x;
  const int nlist = sg.y;
  for (int it = blockIdx.x; it < nlist; it += gridDim.x) {
    const int2 ref = list[it];
    const int sector = ref.y & 0xfffff, chunk = ref.y >> 20;
    const UpdDesc &u = Ud[ref.x];
    const Sector s = u.sectors[sector];
    const int n = s.n, NV = bc_vectors(n);
    const int k = u.kcount[sector], v0 = chunk * NV;
    if (v0 >= k) continue;  // uniform
    const int nv = min(NV, k - v0);
    double2 *Z = bsm;                    // Z[r * NV + vec]
    double2 *Rf = bsm + (int64_t)n * NV;  // two buffers of n reflector slots
    gptr<double2> X = gp(s.left_small ? reinterpret_cast<double2 *>(u.V) + s.v_off
                                      : reinterpret_cast<double2 *>(u.W) + s.w_off);
    const gptr<const double2> A = gp(reinterpret_cast<const double2 *>(u.G) + s.g_off);
    __syncthreads();  // the previous item's readers of LDS are done
    for (int x = threadIdx.x; x < n * NV; x += blockDim.x) {
      const int r = x % n, vec = x / n;
      Z[r * NV + vec] = vec < nv ? gload2(X, (int64_t)(v0 + vec) * n + r) : make_double2(0.0, 0.0);
    }
    // sweep jj's reflectors: column n - 1 - jj of A, slots 0 .. n - 2 - jj;
    // the last sweep (jj = n - 2) first
    if (threadIdx.x == 0) Rf[0] = gload2(A, (int64_t)1 * n);
    __syncthreads();
    int p = 0;
    for (int jj = n - 2; jj >= 0; --jj) {
      // prefetch the next (earlier) sweep's reflectors into registers
      constexpr int kPf = 8;  // n <= kPf * kBcThreads
      double2 pf[kPf];
      const int nxt = jj - 1, nlen = nxt >= 0 ? n - 1 - nxt : 0;
#pragma unroll
      for (int q = 0; q < kPf; ++q) {
        const int x = threadIdx.x + q * kBcThreads;
        pf[q] = x < nlen ? gload2(A, (int64_t)(n - 1 - nxt) * n + x) : make_double2(0.0, 0.0);
      }
      const double2 *rf = Rf + p * n;
      const int len = n - 1 - jj, ntask = (len + kBand - 1) / kBand;
      for (int unit = threadIdx.x; unit < ntask * NV; unit += blockDim.x) {
        const int task = unit / NV, vec = unit % NV;
        const int st = jj + 1 + task * kBand, L = min(kBand, n - st);
        const double2 *vt = rf + task * kBand;
        double2 *zc = Z + (int64_t)st * NV + vec;
        double2 w = zc[0];
        for (int i = 1; i < L; ++i) w = zadd(w, zmulc(vt[i], zc[i * NV]));
        const double2 tw = zmul(vt[0], w);
        zc[0] = zsub(zc[0], tw);
        for (int i = 1; i < L; ++i) zc[i * NV] = zsub(zc[i * NV], zmul(vt[i], tw));
      }
#pragma unroll
      for (int q = 0; q < kPf; ++q) {
        const int x = threadIdx.x + q * kBcThreads;
        if (x < nlen) Rf[(p ^ 1) * n + x] = pf[q];
      }
      __syncthreads();
      p ^= 1;
    }
    for (int x = threadIdx.x; x < n * nv; x += blockDim.x) {
      const int r = x % n, vec = x / n;
      const double2 z = Z[r * NV + vec];
      gstore2(X, (int64_t)(v0 + vec) * n + r, z.x, z.y);
    }
  }
}

static inline const UpdDesc *ts_desc(const GroupHeader &g, const char *w) {
  return reinterpret_cast<const UpdDesc *>(w + g.o_desc);
}
static inline const int2 *ts_segs(const GroupHeader &g, const char *w) {
  return reinterpret_cast<const GroupCounts *>(w + g.o_counts)->seg;
}

// Stage 1 and 2 over the two-stage head of the tridiagonal list (largest
// first: the sectors with two panel rows per thread lead).
int launch_band(const GroupHeader &g, const char *w, hipStream_t stream) {
  if (g.n_ts == 0) return 0;
  if (g.ts_max_n > kBandMaxN) return fail(-1, "two-stage sector too large");
  const WorkSeg tl{reinterpret_cast<const int2 *>(w + g.o_tri), ts_segs(g, w) + SG_TS};
  if (g.ts_nbig > 0) return fail(-1, "two-stage sector with more panel rows than threads");
  hipLaunchKernelGGL(k_band<1>, dim3(g.n_ts), dim3(kBandThreads), 0, stream, ts_desc(g, w), tl);
  BSMPO_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_bulge, dim3(g.n_ts), dim3(64 * kBulgeWaves), 0, stream, ts_desc(g, w), tl);
  BSMPO_HIP(hipGetLastError());
  return 0;
}

int launch_bulge_apply(const GroupHeader &g, const char *w, hipStream_t stream) {
  if (g.n_bc == 0) return 0;
  const int lds = (int)sizeof(double2) * (2048 + 2 * g.ts_max_n);
  if (int rc = ensure_lds_limit((const void *)k_bulge_apply, (int)sizeof(double2) * (2048 + 2 * kBandMaxN)))
    return rc;
  hipLaunchKernelGGL(k_bulge_apply, dim3(std::min(g.n_bc, 16384)), dim3(kBcThreads), (size_t)lds,
                     stream, ts_desc(g, w),
                     WorkSeg{reinterpret_cast<const int2 *>(w + g.o_bc), ts_segs(g, w) + SG_BC});
  BSMPO_HIP(hipGetLastError());
  return 0;
}

}  // namespace bsmpo
