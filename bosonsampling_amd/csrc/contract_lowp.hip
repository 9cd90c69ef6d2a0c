// Reduced-precision sector contraction (gfx950), same semantics as
// k_contract_c128 (src/update_MPO_kernel.cu:193-734, Rank.py:114-118):
//   k_contract_f32     complex FP32 on v_mfma_f32_16x16x4f32, FP32 accumulation
//   k_contract_bf16x3  split-BF16x3 on v_mfma_f32_16x16x16bf16_1k: each real
//                      operand x = hi + lo (bf16 each), x y ~ hi hi + hi lo + lo hi,
//                      FP32 accumulation
// Operands are read from the complex128 MPO and rounded on load; C and
// T = C diag(LR) are written back in complex128 for the FP64 SVD stage.
#include <hip/hip_runtime.h>

#include "common.h"

namespace bsmpo {

typedef short bf16x4 __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x4f32: A[i][k] lane i + 16 k, B[k][j] lane j + 16 k,
// D[4 (l / 16) + r][l % 16] in element r of lane l.
__device__ __forceinline__ int mfma32_row(int kq, int r) { return 4 * kq + r; }

__device__ __forceinline__ void cmfma_f32(flt4 &accr, flt4 &acci, float ar, float ai, float br,
                                          float bi) {
  accr = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, accr, 0, 0, 0);
  accr = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, bi, accr, 0, 0, 0);
  acci = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bi, acci, 0, 0, 0);
  acci = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, br, acci, 0, 0, 0);
}

__global__ __launch_bounds__(256) void k_contract_f32(const UpdDesc *__restrict__ Ud,
                                                      const int2 *__restrict__ list, int nlist) {
  const int lane = threadIdx.x & 63;
  const int i16 = lane & 15, kq = lane >> 4;
  const int nwaves = gridDim.x * 4;
  for (int it = blockIdx.x * 4 + (threadIdx.x >> 6); it < nlist; it += nwaves) {
    const int2 ref = list[it];
    const UpdDesc &u = Ud[ref.x];
    const int d = u.d, chi = u.chi;
    const int32_t *CR = u.charges + 4 * chi;
    const Item item = u.items[ref.y];
    const Sector s = u.sectors[item.sector];
    const Block bl = u.lblk[item.blk];
    const bool rv = i16 < item.nr, cv = i16 < item.nc;
    const int grow = rv ? u.rows[s.row_off + item.r0 + i16] : 0;
    const int gcol = cv ? u.cols[s.col_off + item.c0 + i16] : 0;
    const int cr0 = cv ? CR[2 * gcol] : (1 << 20);
    const int cr1 = cv ? CR[2 * gcol + 1] : (1 << 20);
    const int mincr0 = __shfl(cr0, 0);
    int mincr1 = cr1;
    for (int o = 8; o >= 1; o >>= 1) mincr1 = min(mincr1, __shfl_xor(mincr1, o));
    const double2 *G1 = reinterpret_cast<const double2 *>(u.Glc);
    const double2 *G2 = reinterpret_cast<const double2 *>(u.Gcr);
    const double2 *U = reinterpret_cast<const double2 *>(u.U3);
    flt4 accr = {0.f, 0.f, 0.f, 0.f}, acci = {0.f, 0.f, 0.f, 0.f};
    for (int cb = 0; cb < u.n_cblk; ++cb) {
      const Block bc = u.cblk[cb];
      if (bc.c0 > bl.c0 || bc.c1 > bl.c1) continue;
      if (bc.c0 < mincr0 || bc.c1 < mincr1) continue;
      cplx w{0.0, 0.0};
      if (cv && bc.c0 >= cr0 && bc.c1 >= cr1) {
        const double2 a = U[((bl.c0 - s.t0) * d + (s.t0 - cr0)) * d + (bl.c0 - bc.c0)];
        const double2 b = U[((bl.c1 - s.t1) * d + (s.t1 - cr1)) * d + (bl.c1 - bc.c1)];
        w = cmulc(cplx{a.x, a.y}, cplx{b.x, b.y});
      }
      const float wr = (float)w.x, wi = (float)w.y;
      for (int k0 = bc.start; k0 < bc.end; k0 += 4) {
        const int k = k0 + kq;
        const bool kv = k < bc.end;
        float ar = 0.f, ai = 0.f, gr = 0.f, gi = 0.f;
        if (kv && rv) {
          const double2 v = G1[(int64_t)grow * chi + k];
          ar = (float)v.x;
          ai = (float)v.y;
        }
        if (kv && cv) {
          const double2 v = G2[(int64_t)k * chi + gcol];
          gr = (float)v.x;
          gi = (float)v.y;
        }
        cmfma_f32(accr, acci, ar, ai, gr * wr - gi * wi, gr * wi + gi * wr);
      }
    }
    if (cv) {
      const double lr = u.LR[gcol];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = mfma32_row(kq, r);
        if (row < item.nr) {
          const int64_t idx = s.c_off + (int64_t)(item.r0 + row) * s.R + item.c0 + i16;
          cstore(u.C, idx, cplx{(double)accr[r], (double)acci[r]});
          cstore(u.T, idx, cplx{(double)accr[r] * lr, (double)acci[r] * lr});
        }
      }
    }
  }
}

// round-to-nearest-even float -> bf16 bits, and the residual split
__device__ __forceinline__ unsigned short bf16_bits(float x) {
  unsigned int b = __float_as_uint(x);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (unsigned short)(b >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) {
  return __uint_as_float(((unsigned int)h) << 16);
}
__device__ __forceinline__ void split_bf16(float x, short &hi, short &lo) {
  const unsigned short h = bf16_bits(x);
  hi = (short)h;
  lo = (short)bf16_bits(x - bf16_val(h));
}

// v_mfma_f32_16x16x16bf16_1k: A[i][k] lane i + 16 (k / 4), element k % 4;
// B[k][j] lane j + 16 (k / 4), element k % 4; D as the f32 16x16x4 form.
__device__ __forceinline__ flt4 mfma_bf16(bf16x4 a, bf16x4 b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
// acc += x y with x = xh + xl, y = yh + yl (the xl yl term is dropped)
__device__ __forceinline__ flt4 mfma3(bf16x4 xh, bf16x4 xl, bf16x4 yh, bf16x4 yl, flt4 acc) {
  acc = mfma_bf16(xl, yh, acc);
  acc = mfma_bf16(xh, yl, acc);
  return mfma_bf16(xh, yh, acc);
}

// Centre rows of every qualifying block are compacted into LDS per wave, so
// the K = 16 steps are not padded at charge-block boundaries.
__global__ __launch_bounds__(256) void k_contract_bf16x3(const UpdDesc *__restrict__ Ud,
                                                         const int2 *__restrict__ list,
                                                         int nlist) {
  extern __shared__ int cmp[];  // per wave: rows[chi] then block index[chi]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int nwaves = gridDim.x * 4;
  for (int it = blockIdx.x * 4 + wave; it < nlist; it += nwaves) {
    const int2 ref = list[it];
    const UpdDesc &u = Ud[ref.x];
    const int d = u.d, chi = u.chi;
    int *crow = cmp + (size_t)wave * 2 * chi, *cblk = crow + chi;
    const int32_t *CR = u.charges + 4 * chi;
    const Item item = u.items[ref.y];
    const Sector s = u.sectors[item.sector];
    const Block bl = u.lblk[item.blk];
    const bool rv = i16 < item.nr, cv = i16 < item.nc;
    const int grow = rv ? u.rows[s.row_off + item.r0 + i16] : 0;
    const int gcol = cv ? u.cols[s.col_off + item.c0 + i16] : 0;
    const int cr0 = cv ? CR[2 * gcol] : (1 << 20);
    const int cr1 = cv ? CR[2 * gcol + 1] : (1 << 20);
    const int mincr0 = __shfl(cr0, 0);
    int mincr1 = cr1;
    for (int o = 8; o >= 1; o >>= 1) mincr1 = min(mincr1, __shfl_xor(mincr1, o));
    int Q = 0;
    for (int cb = 0; cb < u.n_cblk; ++cb) {
      const Block bc = u.cblk[cb];
      if (bc.c0 > bl.c0 || bc.c1 > bl.c1) continue;
      if (bc.c0 < mincr0 || bc.c1 < mincr1) continue;
      for (int c = bc.start + lane; c < bc.end; c += 64) {
        crow[Q + c - bc.start] = c;
        cblk[Q + c - bc.start] = cb;
      }
      Q += bc.end - bc.start;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const double2 *G1 = reinterpret_cast<const double2 *>(u.Glc);
    const double2 *G2 = reinterpret_cast<const double2 *>(u.Gcr);
    const double2 *U = reinterpret_cast<const double2 *>(u.U3);
    flt4 accr = {0.f, 0.f, 0.f, 0.f}, acci = {0.f, 0.f, 0.f, 0.f};
    for (int q0 = 0; q0 < Q; q0 += 16) {
      bf16x4 arh, arl, aih, ail, naih, nail, brh, brl, bih, bil;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int q = q0 + 4 * kq + m;
        float ar = 0.f, ai = 0.f, br = 0.f, bi = 0.f;
        if (q < Q) {
          const int c = crow[q];
          if (rv) {
            const double2 v = G1[(int64_t)grow * chi + c];
            ar = (float)v.x;
            ai = (float)v.y;
          }
          if (cv) {
            const Block bc = u.cblk[cblk[q]];
            if (bc.c0 >= cr0 && bc.c1 >= cr1) {
              const double2 a = U[((bl.c0 - s.t0) * d + (s.t0 - cr0)) * d + (bl.c0 - bc.c0)];
              const double2 b = U[((bl.c1 - s.t1) * d + (s.t1 - cr1)) * d + (bl.c1 - bc.c1)];
              const cplx w = cmulc(cplx{a.x, a.y}, cplx{b.x, b.y});
              const double2 g = G2[(int64_t)c * chi + gcol];
              br = (float)(g.x * w.x - g.y * w.y);
              bi = (float)(g.x * w.y + g.y * w.x);
            }
          }
        }
        short h, l;
        split_bf16(ar, h, l); arh[m] = h; arl[m] = l;
        split_bf16(ai, h, l); aih[m] = h; ail[m] = l;
        split_bf16(-ai, h, l); naih[m] = h; nail[m] = l;
        split_bf16(br, h, l); brh[m] = h; brl[m] = l;
        split_bf16(bi, h, l); bih[m] = h; bil[m] = l;
      }
      // Cr += Ar Br - Ai Bi ; Ci += Ar Bi + Ai Br
      accr = mfma3(arh, arl, brh, brl, accr);
      accr = mfma3(naih, nail, bih, bil, accr);
      acci = mfma3(arh, arl, bih, bil, acci);
      acci = mfma3(aih, ail, brh, brl, acci);
    }
    if (cv) {
      const double lr = u.LR[gcol];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = mfma32_row(kq, r);
        if (row < item.nr) {
          const int64_t idx = s.c_off + (int64_t)(item.r0 + row) * s.R + item.c0 + i16;
          cstore(u.C, idx, cplx{(double)accr[r], (double)acci[r]});
          cstore(u.T, idx, cplx{(double)accr[r] * lr, (double)acci[r] * lr});
        }
      }
    }
    __builtin_amdgcn_wave_barrier();  // crow / cblk are reused by the next item
  }
}

int launch_contract_lowp(const GroupHeader &g, const char *w, int precision,
                         hipStream_t stream) {
  const UpdDesc *desc = reinterpret_cast<const UpdDesc *>(w + g.o_desc);
  const int2 *list = reinterpret_cast<const int2 *>(w + g.o_items);
  if (precision == BSMPO_C64) {
    hipLaunchKernelGGL(k_contract_f32, dim3(grid_for(g.n_items)), dim3(256), 0, stream, desc, list,
                       g.n_items);
  } else if (precision == BSMPO_BF16X3) {
    const size_t lds = (size_t)4 * 2 * g.chi * sizeof(int);
    static bool attr = false;
    if (!attr) {
      BSMPO_HIP(hipFuncSetAttribute((const void *)k_contract_bf16x3,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 4 * 2 * 4096 * 4));
      attr = true;
    }
    if (g.chi > 4096) return fail(-1, "bf16x3 contraction supports chi <= 4096");
    hipLaunchKernelGGL(k_contract_bf16x3, dim3(grid_for(g.n_items)), dim3(256), lds, stream, desc,
                       list, g.n_items);
  } else {
    return fail(-1, "unknown precision");
  }
  BSMPO_HIP(hipGetLastError());
  return 0;
}

}  // namespace bsmpo
