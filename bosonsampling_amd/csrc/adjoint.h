// Shared declarations of the reverse-mode (adjoint) path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace bsmpo {

// One sector of an update's backward pass (built on the host from the plan
// and the forward's new centre charges).
struct VjpSector {
  int32_t t0, t1, L, R;
  int32_t n, k;          // min(L, R); kept count (slots slot0 .. slot0+k-1)
  int32_t slot0, left_small;
  int32_t row_off, col_off;  // into the rows / cols pools
  int64_t c_off;         // complex element offset of C / T / gC (L x R row-major)
  int64_t scr;           // complex element offset of the sector's scratch
  int64_t dscr;          // double offset of the sector's real scratch
};

// (sector, left block, right block) with the blocks' offsets inside the sector
struct VjpPair {
  int32_t s, a, b, ro, co, pad;
};

struct VjpArgs {
  int32_t d, chi, n_sectors, n_cblk;
  const VjpSector *sec;
  const VjpPair *pairs;
  const int32_t *rows, *cols;
  const int32_t *CL, *CC, *CR;      // chi x 2 each
  const Block *lblk, *cblk, *rblk;
  int32_t *rowpos, *colpos;         // n_sectors x chi
  const double *C, *T;              // packed sector contraction and T = C diag(LR)
  double *gC;                       // packed gradient of C
  const double *Glc, *Gcr, *LR, *U3;  // the update's inputs
  const double *G0, *G1, *lam_new;  // the update's outputs (new Gamma[mode], [mode+1], Lambda)
  const double *gG0, *gG1, *gLam;   // gradients of the outputs
  double *gGlc, *gGcr, *gLR, *gU3;  // gradients of the inputs (written)
  double *scr, *dscr;               // scratch
  double drop_rel2;                 // the forward's drop rule, on lambda
};

int launch_site_vjp(const VjpArgs &a, int n_pairs, hipStream_t st);
int launch_bs_table_vjp(int d, int nb, const double *d_B, const double *d_gU3, double *d_gB,
                        hipStream_t st);
int launch_observable_vjp(int M, int d, int chi, const double *Gamma, const double *Lambda,
                          const int32_t *charge, const int32_t *pattern, const double *w,
                          double *env, double *gGamma, double *out, hipStream_t st);

}  // namespace bsmpo
