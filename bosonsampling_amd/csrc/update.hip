// gfx950 kernels of the beam-splitter site update (complex128 path).
//
//   k_bs_table     Fock BS elements, FP64            (src/u_kernel.cu:37-70)
//   k_gram         G = T T^H or T^H T per sector on FP64 MFMA
//   k_jacobi       one-sided Jacobi on G (LDS-resident when it fits)
//   k_vectors      V = T Y diag(1/sigma) for sectors with L > R
//   k_wmat         W = V^H C                         (Rank.py:123-124)
//   k_topk         global top-chi + slot assignment  (Rank.py:139-145,182-196)
//   k_zero/k_scatter  write Gamma0Out / Gamma1Out    (Rank.py:147-179, Full.py:130-135)
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <vector>

#include "common.h"

namespace bsmpo {

// ============================================================ BS table ==
// One block per beam splitter; U3[n1][n2][p] = <n1, n2|U(B)|p, q>, q = n1+n2-p:
//   sqrt(n1! n2! / (p! q!)) sum_k C(p,k) C(q,n1-k) B00^k B10^(p-k) B01^(n1-k) B11^(q-n1+k)
__device__ void bs_table_one(int d, const double *Bb, double *U, const BsArgs &args) {
  const cplx b00{Bb[0], Bb[1]}, b01{Bb[2], Bb[3]}, b10{Bb[4], Bb[5]}, b11{Bb[6], Bb[7]};
  for (int e = threadIdx.x; e < d * d * d; e += blockDim.x) {
    const int n1 = e / (d * d), n2 = (e / d) % d, p = e % d;
    const int q = n1 + n2 - p;
    cplx acc{0.0, 0.0};
    if (q >= 0 && q < d) {
      const int klo = max(0, n1 - q), khi = min(p, n1);
      for (int k = klo; k <= khi; ++k) {
        const double c = binom_u64(p, k) * binom_u64(q, n1 - k);
        cplx t = cmul(cmul(cpow_int(b00, k), cpow_int(b10, p - k)),
                      cmul(cpow_int(b01, n1 - k), cpow_int(b11, q - n1 + k)));
        acc.x += c * t.x;
        acc.y += c * t.y;
      }
      const double f = (args.sqf[n1] * args.sqf[n2]) / (args.sqf[p] * args.sqf[q]);
      acc.x *= f;
      acc.y *= f;
    }
    cstore(U, e, acc);
  }
}

__global__ void k_bs_table(int d, const double *B, double *U3, BsArgs args) {
  const int b = blockIdx.x;
  bs_table_one(d, B + 8 * b, U3 + (int64_t)b * d * d * d * 2, args);
}

// The tables of a layer program's replay r = state[1]: its matrices are
// B[r][update]; each update's table goes to its own U3 buffer.  A replay past
// the n_layers provided records DP_NO_TABLE in state[3] (the gate then skips
// the replay).
__global__ void k_bs_table_layer(DpParams P, const DpUpdate *__restrict__ U, BsArgs args) {
  const int r = P.state[1];
  if (r < 0 || r >= P.n_layers) {
    if (blockIdx.x == 0 && threadIdx.x == 0) P.state[3] = DP_NO_TABLE;
    return;
  }
  const int set = r % P.n_sets, b = blockIdx.x;
  if (b >= P.n_upd[set]) return;
  bs_table_one(P.d, P.B + 8 * ((int64_t)r * P.max_upd + b), U[P.off[set] + b].U3, args);
}

int launch_bs_table_layer(const DpParams &p, const DpUpdate *d_upd, hipStream_t stream) {
  if (p.d < 1 || p.d > 32 || p.max_upd < 1) return fail(-1, "bs_table_layer: bad d or n_upd");
  hipLaunchKernelGGL(k_bs_table_layer, dim3(p.max_upd), dim3(256), 0, stream, p, d_upd, bs_args());
  BSMPO_HIP(hipGetLastError());
  return 0;
}

int launch_bs_table(int d, int nb, const double *d_B, double *d_U3, hipStream_t stream) {
  if (d < 1 || d > 32 || nb < 0) return fail(-1, "bs_table: bad d or nb");
  if (nb == 0) return 0;
  const BsArgs a = bs_args();
  const int tk = timing_begin(ST_BSTABLE, stream);
  hipLaunchKernelGGL(k_bs_table, dim3(nb), dim3(256), 0, stream, d, d_B, d_U3, a);
  BSMPO_HIP(hipGetLastError());
  timing_end(tk, stream, 0.0, 16.0 * nb * d * d * d);
  return 0;
}

// ============================================ 32 x 32 block products ==
// D (32 x 32 block) = sum_k A(i, k) B(k, j) on the FP64 MFMA, one workgroup
// per block (its four waves one 16 x 16 tile each), the two 32-row operand
// panels P(i, k) = A(i, k) and Q(j, k) = B(k, j) staged through LDS in chunks
// of kGramK along k with coalesced loads one chunk ahead: each panel is read
// once per block instead of once per 16 x 16 tile.  A panel is "k-major" when
// its memory runs along i (resp. j) for fixed k, else along k; the staging
// order follows memory.  fetch(x, k) returns the panel element (masked).
constexpr int kGramK = 16;
constexpr int kGramLd = kGramK + 1;  // LDS row stride (complex)

struct BlockLds {
  double2 P[2][kGramBlock * kGramLd], Q[2][kGramBlock * kGramLd];
};

template <bool PK, bool QK, bool PC, bool QC, class FP, class FQ>
__device__ __forceinline__ void block_product(BlockLds &sm, int K, bool skip, FP fp, FQ fq,
                                              dbl4 &accr, dbl4 &acci) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int tr = wave >> 1, tc = wave & 1;
  auto xk = [](bool kmajor, int e, int &x, int &k) {
    x = kmajor ? e % kGramBlock : e / kGramK;
    k = kmajor ? e / kGramBlock : e % kGramK;
  };
  double2 pv[2], qv[2];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int e = threadIdx.x + 256 * h;
      int x, k;
      xk(PK, e, x, k);
      pv[h] = k0 + k < K ? fp(x, k0 + k) : make_double2(0.0, 0.0);
      xk(QK, e, x, k);
      qv[h] = k0 + k < K ? fq(x, k0 + k) : make_double2(0.0, 0.0);
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int e = threadIdx.x + 256 * h;
      int x, k;
      xk(PK, e, x, k);
      sm.P[buf][x * kGramLd + k] = pv[h];
      xk(QK, e, x, k);
      sm.Q[buf][x * kGramLd + k] = qv[h];
    }
  };
  fetch(0);
  __syncthreads();  // the previous block's readers of LDS are done
  stash(0);
  __syncthreads();
  accr = dbl4{0.0, 0.0, 0.0, 0.0};
  acci = dbl4{0.0, 0.0, 0.0, 0.0};
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += kGramK) {
    const bool more = k0 + kGramK < K;  // uniform
    if (more) fetch(k0 + kGramK);
    if (!skip) {
      const double2 *Pr = sm.P[buf] + (16 * tr + i16) * kGramLd;
      const double2 *Qr = sm.Q[buf] + (16 * tc + i16) * kGramLd;
#pragma unroll
      for (int st = 0; st < kGramK / 4; ++st) {
        const double2 a = Pr[4 * st + kq], b = Qr[4 * st + kq];
        cmfma_f64(accr, acci, cplx{a.x, PC ? -a.y : a.y}, cplx{b.x, QC ? -b.y : b.y});
      }
    }
    if (more) {
      stash(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
}

// ================================================================ Gram ==
// G (n x n, column-major) = T T^H if L <= R else T^H T, T row-major L x R.
// Tridiagonal-route sectors (n above the eigensolver threshold) need only the
// lower triangle: the strictly upper tile of a diagonal block is skipped.
__global__ __launch_bounds__(256) void k_gram(const UpdDesc *__restrict__ Ud, WorkSeg ws) {
  const int2 sg = *ws.seg;
  const int2 *__restrict__ list = ws.list + sg.x;
  const int nlist = sg.y;
  __shared__ BlockLds sm;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int tr = wave >> 1, tc = wave & 1;
  for (int pos = blockIdx.x; pos < xcd_positions(nlist); pos += gridDim.x) {
    const int it = xcd_item(pos, nlist);
    if (it >= nlist) continue;  // uniform over the block
    const int2 ref = list[it];
    const UpdDesc &u = Ud[ref.x];
    const Item item = u.gram_items[ref.y];
    const Sector s = u.sectors[item.sector];
    const gptr<const double2> Tg = gp(reinterpret_cast<const double2 *>(u.T) + s.c_off);
    const bool skip = s.n > u.eig_thr && item.r0 == item.c0 && tr < tc;  // strictly upper
    const int R = s.R;
    dbl4 accr, acci;
    if (s.left_small) {  // P = T(r0 + i, k), Q = T(c0 + j, k) conjugated; rows of T contiguous
      block_product<false, false, false, true>(
          sm, R, skip,
          [&](int x, int k) { return x < item.nr ? gload2(Tg, (int64_t)(item.r0 + x) * R + k) : make_double2(0.0, 0.0); },
          [&](int x, int k) { return x < item.nc ? gload2(Tg, (int64_t)(item.c0 + x) * R + k) : make_double2(0.0, 0.0); },
          accr, acci);
    } else {  // P = T(k, r0 + i) conjugated, Q = T(k, c0 + j); columns of T along a row
      block_product<true, true, true, false>(
          sm, s.L, skip,
          [&](int x, int k) { return x < item.nr ? gload2(Tg, (int64_t)k * R + item.r0 + x) : make_double2(0.0, 0.0); },
          [&](int x, int k) { return x < item.nc ? gload2(Tg, (int64_t)k * R + item.c0 + x) : make_double2(0.0, 0.0); },
          accr, acci);
    }
    if (!skip) {
      const int qc = 16 * tc + i16;  // column of G in the block
      if (qc < item.nc) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * tr + mfma64_row(kq, r);
          if (row < item.nr)
            cstore(u.G, s.g_off + (int64_t)(item.c0 + qc) * s.n + item.r0 + row,
                   cplx{accr[r], acci[r]});
        }
      }
    }
  }
}

// ============================================================== Jacobi ==
// One-sided (Hestenes) Jacobi on the Hermitian PSD Gram matrix G = X, held
// in LDS: rotate column pairs until all are orthogonal.  Then
// X = Y diag(lambda) with Y the eigenvectors of G and lambda = sigma^2.
// Column pairs follow the round-robin tournament; 8 lanes per pair.
constexpr int kJacSweeps = 40;
constexpr double kJacTol = 1e-15;
// singular values with sigma <= drop_rel sigma_max(sector) (default 1e-6: the
// Gram route's noise floor; bsmpo_set_drop_rel) are dropped (sigma := 0,
// vectors 0).  UpdDesc::drop_rel2 holds drop_rel^2, on lambda = sigma^2.

__device__ __forceinline__ void rr_pair(int step, int i, int n2, int &p, int &q) {
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

__global__ __launch_bounds__(256) void k_jacobi(const UpdDesc *__restrict__ Ud, WorkSeg ws) {
  extern __shared__ double2 smem[];
  const int2 sg = *ws.seg;
  if ((int)blockIdx.x >= sg.y) return;  // a launch sized for the group's capacity
  const int2 ref = ws.list[sg.x + blockIdx.x];
  const UpdDesc &u = Ud[ref.x];
  const Sector s = u.sectors[ref.y];
  const int n = s.n;
  double2 *X = smem;
  __shared__ int any_rot;
  __shared__ double lam_max;
  if (n == 0) return;
  double *lam = u.lam;
  int32_t *rank = u.slot;  // scratch until the truncation stage
  {
    const double2 *Gs = reinterpret_cast<const double2 *>(u.G) + s.g_off;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) X[e] = Gs[e];
    __syncthreads();
  }
  const int group = threadIdx.x >> 3, gl = threadIdx.x & 7;
  const int ngroups = blockDim.x >> 3;
  const int n2 = n + (n & 1);
  const double tol = fmax(kJacTol, (double)n * 2.220446049250313e-16);
  for (int sweep = 0; sweep < kJacSweeps && n > 1; ++sweep) {
    if (threadIdx.x == 0) any_rot = 0;
    __syncthreads();
    for (int step = 0; step < n2 - 1; ++step) {
      for (int pi = group; pi < n2 / 2; pi += ngroups) {
        int p, q;
        rr_pair(step, pi, n2, p, q);
        if (q >= n) continue;
        double2 *xp = X + (int64_t)p * n, *xq = X + (int64_t)q * n;
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        for (int i = gl; i < n; i += 8) {
          const double2 a = xp[i], b = xq[i];
          al += a.x * a.x + a.y * a.y;
          be += b.x * b.x + b.y * b.y;
          gr += a.x * b.x + a.y * b.y;  // conj(a) b
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
          const double er = gr / ga, ei = gi / ga;  // e^{i phi}
          for (int i = gl; i < n; i += 8) {
            const double2 a = xp[i], b = xq[i];
            // a' = c a - s conj(e) b ; b' = s e a + c b
            double2 na, nb;
            na.x = c * a.x - sn * (er * b.x + ei * b.y);
            na.y = c * a.y - sn * (er * b.y - ei * b.x);
            nb.x = c * b.x + sn * (er * a.x - ei * a.y);
            nb.y = c * b.y + sn * (er * a.y + ei * a.x);
            xp[i] = na;
            xq[i] = nb;
          }
          if (gl == 0) any_rot = 1;
        }
      }
      __syncthreads();
    }
    if (!any_rot) break;
    __syncthreads();
  }
  // eigenvalues = column norms (X = Y diag(lambda)); order them descending,
  // ties by index, and write sigma = sqrt(lambda), flags and unit vectors to
  // the sorted positions.
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    double nr = 0.0;
    for (int i = 0; i < n; ++i) {
      const double2 a = X[(int64_t)p * n + i];
      nr += a.x * a.x + a.y * a.y;
    }
    lam[s.s_off + p] = sqrt(nr);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int p = 0; p < n; ++p) m = fmax(m, lam[s.s_off + p]);
    lam_max = m;
  }
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const double lp = lam[s.s_off + p];
    int r = 0;
    for (int q = 0; q < n; ++q) {
      const double lq = lam[s.s_off + q];
      r += (lq > lp || (lq == lp && q < p)) ? 1 : 0;
    }
    rank[s.s_off + p] = r;
  }
  __syncthreads();
  double2 *Yout = s.left_small ? reinterpret_cast<double2 *>(u.V) + s.v_off
                               : reinterpret_cast<double2 *>(u.W) + s.w_off;
  const double thr = u.drop_rel2 * lam_max;
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const double lp = lam[s.s_off + p];
    const bool keep = lam_max > 0.0 && lp > thr;
    const int r = rank[s.s_off + p];
    u.S[s.s_off + r] = keep ? sqrt(lp) : 0.0;
    u.flag[s.s_off + r] = keep ? 1 : 0;
  }
  for (int p = 0; p < n; ++p) {
    const double lp = lam[s.s_off + p];
    const bool keep = lam_max > 0.0 && lp > thr;
    const double inv = keep ? 1.0 / lp : 0.0;
    const int r = rank[s.s_off + p];
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const double2 a = X[(int64_t)p * n + i];
      Yout[(int64_t)r * n + i] = make_double2(a.x * inv, a.y * inv);
    }
  }
}

// V (L x n column-major) = T Y diag(1 / sigma) for sectors with L > R (Y
// the Gram eigenvectors, n x n column-major in W's storage), kept columns only.
__global__ __launch_bounds__(256) void k_vectors(const UpdDesc *__restrict__ Ud, WorkSeg ws) {
  const int2 sg = *ws.seg;
  const int2 *__restrict__ list = ws.list + sg.x;
  const int nlist = sg.y;
  __shared__ BlockLds sm;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int tr = wave >> 1, tc = wave & 1;
  for (int pos = blockIdx.x; pos < xcd_positions(nlist); pos += gridDim.x) {
    const int it = xcd_item(pos, nlist);
    if (it >= nlist) continue;  // uniform over the block
    const int2 ref = list[it];
    const UpdDesc &u = Ud[ref.x];
    const Item item = u.v_items[ref.y];
    if (item.c0 >= u.kcount[item.sector]) continue;  // uniform: only kept vectors are needed
    const Sector s = u.sectors[item.sector];
    const gptr<const double2> Tg = gp(reinterpret_cast<const double2 *>(u.T) + s.c_off);
    const gptr<const double2> Yg = gp(reinterpret_cast<const double2 *>(u.W) + s.w_off);
    const int R = s.R, n = s.n;
    dbl4 accr, acci;
    // P = T(r0 + i, k) (rows of T contiguous), Q = Y(k, c0 + p) (columns of Y contiguous)
    block_product<false, false, false, false>(
        sm, R, false,
        [&](int x, int k) { return x < item.nr ? gload2(Tg, (int64_t)(item.r0 + x) * R + k) : make_double2(0.0, 0.0); },
        [&](int x, int k) { return x < item.nc ? gload2(Yg, (int64_t)(item.c0 + x) * n + k) : make_double2(0.0, 0.0); },
        accr, acci);
    const int pc = 16 * tc + i16;
    if (pc < item.nc) {
      const int p = item.c0 + pc;
      const double sg = u.S[s.s_off + p];
      const double inv = sg > 0.0 ? 1.0 / sg : 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * tr + mfma64_row(kq, r);
        if (row < item.nr)
          cstore(u.V, s.v_off + (int64_t)p * s.L + item.r0 + row, cplx{accr[r] * inv, acci[r] * inv});
      }
    }
  }
}

// W (n x R row-major) = V^H C, V L x n column-major, C L x R row-major.
__global__ __launch_bounds__(256) void k_wmat(const UpdDesc *__restrict__ Ud, WorkSeg ws) {
  const int2 sg = *ws.seg;
  const int2 *__restrict__ list = ws.list + sg.x;
  const int nlist = sg.y;
  __shared__ BlockLds sm;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int tr = wave >> 1, tc = wave & 1;
  for (int pos = blockIdx.x; pos < xcd_positions(nlist); pos += gridDim.x) {
    const int it = xcd_item(pos, nlist);
    if (it >= nlist) continue;  // uniform over the block
    const int2 ref = list[it];
    const UpdDesc &u = Ud[ref.x];
    const Item item = u.w_items[ref.y];
    if (item.r0 >= u.kcount[item.sector]) continue;  // uniform: only kept rows of W are needed
    const Sector s = u.sectors[item.sector];
    const gptr<const double2> Vg = gp(reinterpret_cast<const double2 *>(u.V) + s.v_off);
    const gptr<const double2> Cg = gp(reinterpret_cast<const double2 *>(u.C) + s.c_off);
    const int L = s.L, R = s.R;
    dbl4 accr, acci;
    // P = V(k, r0 + p) conjugated (columns of V contiguous), Q = C(k, c0 + j) (rows of C)
    block_product<false, true, true, false>(
        sm, L, false,
        [&](int x, int k) { return x < item.nr ? gload2(Vg, (int64_t)(item.r0 + x) * L + k) : make_double2(0.0, 0.0); },
        [&](int x, int k) { return x < item.nc ? gload2(Cg, (int64_t)k * R + item.c0 + x) : make_double2(0.0, 0.0); },
        accr, acci);
    const int jc = 16 * tc + i16;
    if (jc < item.nc) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * tr + mfma64_row(kq, r);
        if (row < item.nr)
          cstore(u.W, s.w_off + (int64_t)(item.r0 + row) * R + item.c0 + jc, cplx{accr[r], acci[r]});
      }
    }
  }
}

// ========================================================= truncation ==
__device__ int block_sum_int(int v, int *scratch) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = v;
  __syncthreads();
  int t = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += scratch[i];
  __syncthreads();
  return t;
}

// exclusive prefix sum over a 256-thread block (Hillis-Steele in LDS)
__device__ int block_scan_excl(int v, int *sc, int &total) {
  __syncthreads();
  sc[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int add = threadIdx.x >= (unsigned)o ? sc[threadIdx.x - o] : 0;
    __syncthreads();
    sc[threadIdx.x] += add;
    __syncthreads();
  }
  total = sc[255];
  const int incl = sc[threadIdx.x];
  __syncthreads();
  return incl - v;
}

// One block per update: global top-chi over all sectors' singular values,
// then slots in concatenated order, which is the (tau0, tau1, descending
// sigma) order of the stable charge sort.  Values within tie_rel (relative)
// of the chi-th largest t are ties, broken by the lower concatenated index
// (degenerate twins of Hermitian-conjugate sectors).
__global__ __launch_bounds__(256) void k_topk(const UpdDesc *__restrict__ Ud,
                                              const int2 *__restrict__ useg) {
  __shared__ int scratch[16];
  __shared__ int sc[256];
  if ((int)blockIdx.x >= useg->y) return;  // a skipped group
  const UpdDesc &u = Ud[blockIdx.x];
  const int64_t nt = u.sv_total;
  const double *S = u.S;
  const int32_t *flags = u.flag;
  const unsigned long long *key = reinterpret_cast<const unsigned long long *>(S);
  int valid = 0;
  for (int64_t i = threadIdx.x; i < nt; i += blockDim.x) valid += flags[i] ? 1 : 0;
  valid = block_sum_int(valid, scratch);
  const int K = min(valid, u.chi);
  unsigned long long t = 0ull;  // K-th largest key, by bitwise binary search
  if (K > 0) {
    for (int bit = 62; bit >= 0; --bit) {
      const unsigned long long cand = t | (1ull << bit);
      int cnt = 0;
      for (int64_t i = threadIdx.x; i < nt; i += blockDim.x) cnt += (flags[i] && key[i] >= cand) ? 1 : 0;
      cnt = block_sum_int(cnt, scratch);
      if (cnt >= K) t = cand;
    }
  }
  const double tv = __longlong_as_double((long long)t);
  const double hi = tv * (1.0 + u.tie_rel), lo = tv * (1.0 - u.tie_rel);
  int gt = 0;
  for (int64_t i = threadIdx.x; i < nt; i += blockDim.x) gt += (flags[i] && S[i] > hi) ? 1 : 0;
  gt = block_sum_int(gt, scratch);
  const int need_eq = K - gt;
  int carry_eq = 0, carry_kept = 0;
  for (int64_t base = 0; base < nt; base += blockDim.x) {
    const int64_t i = base + threadIdx.x;
    const bool v = i < nt && flags[i] && K > 0;
    const bool gtf = v && S[i] > hi;
    const bool eqf = v && !gtf && S[i] >= lo;
    int tot_eq;
    const int eq_rank = carry_eq + block_scan_excl(eqf ? 1 : 0, sc, tot_eq);
    const bool kept = gtf || (eqf && eq_rank < need_eq);
    int tot_kept;
    const int pos = carry_kept + block_scan_excl(kept ? 1 : 0, sc, tot_kept);
    carry_eq += tot_eq;
    carry_kept += tot_kept;
    if (i < nt) {
      u.slot[i] = kept ? pos : -1;
      if (kept) {
        const Sector s = u.sectors[u.sv_sector[i]];
        u.charge_c[2 * pos] = s.t0;
        u.charge_c[2 * pos + 1] = s.t1;
        u.lambda_c[pos] = S[i];
      }
    }
  }
  for (int j = K + threadIdx.x; j < u.chi; j += blockDim.x) {
    u.charge_c[2 * j] = u.d;
    u.charge_c[2 * j + 1] = u.d;
    u.lambda_c[j] = 0.0;
  }
  if (threadIdx.x == 0) *u.num_lambda = K;
}

// stored part of the outputs (Full.py:130-135 edge storage)
__device__ __forceinline__ void out_region(const UpdDesc &u, int &w0, int &w1, int &c1) {
  const int wd = min(u.chi, u.d * u.d);
  w0 = u.edge ? wd : u.chi;
  w1 = u.edge ? wd : u.chi;
  c1 = u.edge ? 1 : u.chi;
}

// grid (x, n_updates): zero the stored part of Gamma0Out / Gamma1Out
__global__ void k_zero(const UpdDesc *__restrict__ Ud, const int2 *__restrict__ useg) {
  if ((int)blockIdx.y >= useg->y) return;  // a skipped group
  const UpdDesc &u = Ud[blockIdx.y];
  int w0, w1, c1;
  out_region(u, w0, w1, c1);
  const int chi = u.chi;
  const int64_t tot0 = (int64_t)chi * w0, tot1 = (int64_t)w1 * c1;
  double2 *a = reinterpret_cast<double2 *>(u.Glc);
  double2 *b = reinterpret_cast<double2 *>(u.Gcr);
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot0 + tot1;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (e < tot0) {
      const int64_t r = e / w0, c = e % w0;
      a[r * chi + c] = make_double2(0.0, 0.0);
    } else {
      const int64_t f = e - tot0;
      const int64_t r = f / c1, c = f % c1;
      b[r * chi + c] = make_double2(0.0, 0.0);
    }
  }
}

// grid (x, n_updates): one block per kept singular-value slot: V column, W row
__global__ __launch_bounds__(256) void k_scatter(const UpdDesc *__restrict__ Ud,
                                                 const int2 *__restrict__ useg) {
  if ((int)blockIdx.y >= useg->y) return;  // a skipped group
  const UpdDesc &u = Ud[blockIdx.y];
  int w0, w1, c1;
  out_region(u, w0, w1, c1);
  const int chi = u.chi;
  for (int64_t i = blockIdx.x; i < u.sv_total; i += gridDim.x) {
    const int j = u.slot[i];
    if (j < 0) continue;
    const Sector s = u.sectors[u.sv_sector[i]];
    const int p = (int)(i - s.s_off);
    const double2 *V = reinterpret_cast<const double2 *>(u.V) + s.v_off + (int64_t)p * s.L;
    const double2 *W = reinterpret_cast<const double2 *>(u.W) + s.w_off + (int64_t)p * s.R;
    double2 *a = reinterpret_cast<double2 *>(u.Glc);
    double2 *b = reinterpret_cast<double2 *>(u.Gcr);
    if (j < w0)
      for (int r = threadIdx.x; r < s.L; r += blockDim.x)
        a[(int64_t)u.rows[s.row_off + r] * chi + j] = V[r];
    if (j < w1)
      for (int c = threadIdx.x; c < s.R; c += blockDim.x) {
        const int gc = u.cols[s.col_off + c];
        if (gc < c1) b[(int64_t)j * chi + gc] = W[c];
      }
  }
}

// ============================================================ launchers ==
int launch_tri_values(const GroupHeader &g, const char *d_work, hipStream_t stream);
int launch_svd_head(const GroupHeader &g, const char *d_work, hipStream_t stream);
int launch_kcount(const GroupHeader &g, const char *d_work, hipStream_t stream);
int launch_tri_vectors(const GroupHeader &g, const char *d_work, hipStream_t stream);

static inline const UpdDesc *gdesc(const GroupHeader &g, const char *w) {
  return reinterpret_cast<const UpdDesc *>(w + g.o_desc);
}
static inline const int2 *gseg(const GroupHeader &g, const char *w, int seg) {
  return reinterpret_cast<const GroupCounts *>(w + g.o_counts)->seg + seg;
}
static inline WorkSeg wseg(const GroupHeader &g, const char *w, int64_t o_list, int seg) {
  return WorkSeg{reinterpret_cast<const int2 *>(w + o_list), gseg(g, w, seg)};
}

// The header's counts size the grids (exact for a host-built group, the
// capacity for one planned on the device); the kernels read the counts of
// their work from the group's GroupCounts.
int launch_svd(const GroupHeader &g, const char *w, hipStream_t stream) {
  if (g.n_sec == 0) return 0;
  if (int rc = launch_svd_head(g, w, stream)) return rc;
  return launch_tri_values(g, w, stream);
}

// Gram matrices and the one-sided Jacobi route (the SVD stage up to the
// tridiagonal route's reductions)
int launch_svd_head(const GroupHeader &g, const char *w, hipStream_t stream) {
  if (g.n_sec == 0) return 0;
  int tk = timing_begin(ST_GRAM, stream);
  hipLaunchKernelGGL(k_gram, dim3(xcd_grid(g.n_gram)), dim3(256), 0, stream, gdesc(g, w),
                     wseg(g, w, g.o_gram, SG_GRAM));
  BSMPO_HIP(hipGetLastError());
  timing_end(tk, stream, g.flops_gram, 16.0 * (g.c_elems + g.g_elems));
  if (int rc = ensure_lds_limit((const void *)k_jacobi, 96 * 96 * 16)) return rc;
  tk = timing_begin(ST_JACOBI, stream);
  for (int c = 0; c < kTriClass; ++c) {
    const int nsec = g.jac_off[c + 1] - g.jac_off[c];
    if (nsec == 0) continue;
    const size_t lds = (size_t)g.jac_max_n[c] * g.jac_max_n[c] * 16;
    hipLaunchKernelGGL(k_jacobi, dim3(nsec), dim3(256), lds, stream, gdesc(g, w),
                       wseg(g, w, g.o_jac, SG_JAC + c));
    BSMPO_HIP(hipGetLastError());
  }
  timing_end(tk, stream, g.flops_jacobi, 0.0);
  return 0;
}

int launch_truncate(const GroupHeader &g, const char *w, hipStream_t stream) {
  int tk = timing_begin(ST_TOPK, stream);
  hipLaunchKernelGGL(k_topk, dim3(g.n_upd), dim3(256), 0, stream, gdesc(g, w), gseg(g, w, SG_UPD));
  BSMPO_HIP(hipGetLastError());
  timing_end(tk, stream, 0.0, 12.0 * g.sv_total);
  int rc = launch_kcount(g, w, stream);
  if (rc) return rc;
  if ((rc = launch_tri_vectors(g, w, stream))) return rc;
  if (g.n_v) {
    tk = timing_begin(ST_VECTORS, stream);
    hipLaunchKernelGGL(k_vectors, dim3(xcd_grid(g.n_v)), dim3(256), 0, stream, gdesc(g, w),
                       wseg(g, w, g.o_v, SG_V));
    BSMPO_HIP(hipGetLastError());
    timing_end(tk, stream, 0.0, 0.0);  // flops: kept pairs, counted by k_kcount
  }
  if (g.n_w) {
    tk = timing_begin(ST_WMAT, stream);
    hipLaunchKernelGGL(k_wmat, dim3(xcd_grid(g.n_w)), dim3(256), 0, stream, gdesc(g, w),
                       wseg(g, w, g.o_w, SG_W));
    BSMPO_HIP(hipGetLastError());
    timing_end(tk, stream, 0.0, 16.0 * (g.v_elems + g.c_elems + g.w_elems));  // flops: k_kcount
  }
  tk = timing_begin(ST_SCATTER, stream);
  hipLaunchKernelGGL(k_zero, dim3(256, g.n_upd), dim3(256), 0, stream, gdesc(g, w), gseg(g, w, SG_UPD));
  BSMPO_HIP(hipGetLastError());
  if (g.max_sv > 0) {
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)std::min<int64_t>(g.max_sv, 4096), g.n_upd),
                       dim3(256), 0, stream, gdesc(g, w), gseg(g, w, SG_UPD));
    BSMPO_HIP(hipGetLastError());
  }
  timing_end(tk, stream, 0.0, 16.0 * 2.0 * g.n_upd * (double)g.chi * g.chi);
  return 0;
}

}  // namespace bsmpo
