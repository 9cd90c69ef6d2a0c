/*
 * bsmpo.h -- C ABI of the MI355X (gfx950) beam-splitter site update for the
 * lossy boson-sampling MPO.  Plain pointers and sizes only.
 *
 * Conventions
 *  - complex numbers are complex128 (interleaved re, im doubles) unless a
 *    function says complex64 (interleaved floats);
 *  - matrices are row-major; Gamma tensors are chi x chi per site;
 *  - charges are int32 pairs (ket, bra) per bond slot, lexicographically
 *    sorted, unused slots padded with d;
 *  - every pointer named d_* is device memory owned by the caller; the
 *    library never allocates or frees caller memory.  Functions that need
 *    scratch take a caller-provided workspace (d_work, work_bytes) whose
 *    size is obtained from the matching *_query call;
 *  - all device work is enqueued on `stream` (a hipStream_t passed as
 *    void*; NULL = the default stream) and nothing synchronises the host
 *    unless the function says so;
 *  - return value: 0 ok, > 0 a hipError_t, < 0 invalid argument; the text
 *    of the last error of the calling thread is bsmpo_last_error().
 */
#ifndef BSMPO_H
#define BSMPO_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- misc -- */

/* Text of the last error raised on this thread ("" if none). */
const char *bsmpo_last_error(void);

/* ABI version (increments on any signature change). */
int bsmpo_abi_version(void);

/*
 * Sector-SVD route selection (process-wide, affects plans built afterwards):
 * Gram matrices with n <= threshold are diagonalised by one-sided Jacobi in
 * LDS, larger ones by the tridiagonal route (Householder reduction, Sturm
 * bisection for all eigenvalues, inverse iteration + back-transformation for
 * the kept pairs only).  Default 96, allowed [0, 96].  No counterpart in the
 * reference (which calls cuSOLVER gesvd per sector, Rank.py:123).
 */
int bsmpo_set_eig_threshold(int n);

/*
 * Householder block-shape split (process-wide, affects groups built
 * afterwards): in a grouped launch with many tridiagonal-route sectors, those
 * with n > split run on 512-thread blocks, launched first, the rest on
 * 256-thread blocks on a side stream.  Default 256, allowed [1, 512]; every
 * value gives the same results to rounding.  No counterpart in the reference
 * (a launch-geometry knob of this build's eigensolver, like
 * bsmpo_set_eig_threshold).
 */
int bsmpo_set_tri_split(int n);
/* The block-shape split and the route threshold in effect. */
int bsmpo_tri_split(void);
int bsmpo_eig_threshold(void);

/*
 * Contraction tile order (process-wide, affects plans built afterwards): the
 * contraction's 16 x 16 tiles of (left row, right column) elements are listed
 * in super-tiles of order x order tiles, row by row, so that neighbouring work
 * items (which run on one XCD and share its L2) read the same operand panels.
 * Default 8, allowed 2, 4, 8, 16; every value gives the same results.  No
 * counterpart in the reference (its kernel tiles each sector on its own,
 * src/update_MPO_kernel.cu:193-734); engine.autotune_layer picks it.
 */
int bsmpo_set_tile_order(int n);
int bsmpo_tile_order(void);

/*
 * Two-stage reduction threshold (process-wide, affects plans built
 * afterwards): tridiagonal-route Gram matrices with threshold <= n <= 528 are
 * reduced to a band of width 16 (panels of 16 columns, FP64 MFMA), then to
 * tridiagonal form by bulge chasing, instead of the one-stage Householder
 * reduction (which reads the trailing matrix once per column).  0 turns it
 * off; other values are raised to at least 64.  Allowed [0, 528].  Same
 * results to rounding.  No counterpart in the reference (cuSOLVER gesvd per
 * sector, Rank.py:123).
 */
int bsmpo_set_two_stage(int n);
/* The two-stage threshold in effect (0: off). */
int bsmpo_two_stage_threshold(void);

/*
 * Truncation tie band (process-wide, affects plans built afterwards): the
 * kept set is every singular value above t (1 + rel), t the chi-th largest,
 * then the values within [t (1 - rel), t (1 + rel)] in concatenated order
 * (sector order, descending sigma).  Exactly degenerate twins are the rule
 * (sectors (a, b) and (b, a) of the Hermitian density matrix), and rounding
 * noise must not decide between them.  Default 1e-6; replaces the
 * unspecified argpartition tie order of src/beam_splitter_parallel/Rank.py:142.
 */
int bsmpo_set_tie_rel(double rel);

/*
 * Drop rule (process-wide, affects plans built afterwards): a singular value
 * sigma <= rel * sigma_max of its sector is dropped with its vectors before
 * the truncation; the site update's adjoint uses the same rule.  Default
 * 1e-6, the noise floor of the Gram route in FP64 (sigma_min / sigma_max ~
 * sqrt(eps)); a contraction of lower precision has a higher floor (split
 * BF16x3: about 1e-5), below which the values are rounding noise that the
 * inverse iteration would otherwise treat as one long cluster.  Accepted
 * range [1e-8, 1e-3].  The reference keeps every value
 * (src/beam_splitter_parallel/Rank.py:139-145).
 */
int bsmpo_set_drop_rel(double rel);
/* The drop rule in effect. */
double bsmpo_drop_rel(void);

/* Contraction arithmetic of the sector GEMMs (the 16 x 16 block products
 * P_q = Glc Gcr of the sector contraction; the BS-weight transforms that
 * follow run on the vector ALUs in the same precision). */
enum bsmpo_precision {
  BSMPO_C128 = 0,    /* complex128 on v_mfma_f64_16x16x4f64                   */
  BSMPO_C64 = 1,     /* complex64 on v_mfma_f32_16x16x4f32, FP32 accumulation */
  BSMPO_BF16X3 = 2   /* split-BF16x3 (hi hi + hi lo + lo hi) on gfx950's
                        v_mfma_f32_16x16x32_bf16, FP32 accumulation          */
};

/* ------------------------------------------------------------- timing -- */

/*
 * Per-stage device timing (HIP events on the launch stream), for the live
 * roofline of bench.py.  Stages: 0 bs_table, 1 contract, 2 gram, 3 jacobi,
 * 4 vectors, 5 wmat, 6 topk, 7 scatter, 8 observables, 9 tridiag,
 * 10 bisect, 11 trivec, 12 backtransform (BSMPO_N_STAGES).
 * bsmpo_timing_read synchronises on the recorded events and returns the
 * accumulated milliseconds, algorithmic flops and bytes and launch counts
 * per stage (arrays of BSMPO_N_STAGES, any may be NULL); reset != 0 zeroes
 * the accumulators.  No counterpart in the reference (which times with
 * time.time(), src/beam_splitter_parallel/Rank.py:93-127).
 */
#define BSMPO_N_STAGES 13
int bsmpo_timing_enable(int on);
int bsmpo_timing_read(double *ms, double *flops, double *bytes, int64_t *count, int reset);

/* ------------------------------------------------- beam-splitter table -- */

/*
 * Fock-space beam-splitter elements for nb beam splitters.
 *   d_B   : device, nb x 2 x 2 complex128, single-particle matrix, new = B old
 *   d_U3  : device, nb x d x d x d complex128,
 *           U3[n1][n2][p] = <n1, n2 | U(B) | p, n1 + n2 - p>  (0 outside)
 * Replaces src/cuda_kernels.py:18-50 (Rand_U) + src/u_kernel.cu:37-70
 * (`unitary`); the reference's B = [[t, r], [-r*, t*]] is one such matrix.
 * Evaluated in FP64 (the reference stores float).
 */
int bsmpo_bs_table(int d, int nb, const double *d_B, double *d_U3, void *stream);

/* ------------------------------------------------------ update planning -- */

/*
 * Host-side plan of one site update from the three bonds' charges.
 *   h_CL, h_CC, h_CR : host int32 [chi][2] (left, centre, right bond)
 * The plan is written into an opaque host buffer of plan_bytes bytes
 * (size from bsmpo_plan_query with the same arguments) and contains the
 * sector list (tau, rows cl >= tau, cols cr <= tau), the charge-block tables
 * and the contraction work list.  tau0 < 0 plans every sector of the update;
 * tau0, tau1 >= 0 plans that one sector only (the update_MPO drop-in).
 * Replaces the Aligner of src/mpo_sort.py:39-418 and the tau loop bounds of
 * src/beam_splitter_parallel/Rank.py:60-105.  d <= 16 (the local Fock
 * cutoff; the reference's contraction kernel has the same bound: 256-entry
 * centre-charge tables, src/update_MPO_kernel.cu:236-238).
 */
int bsmpo_plan_query(int d, int chi, const int32_t *h_CL, const int32_t *h_CC,
                     const int32_t *h_CR, int tau0, int tau1, size_t *plan_bytes,
                     size_t *work_bytes);
int bsmpo_plan_build(int d, int chi, const int32_t *h_CL, const int32_t *h_CC,
                     const int32_t *h_CR, int tau0, int tau1, void *h_plan, size_t plan_bytes);

/* Summary numbers of a built plan (host, no device work). */
typedef struct {
  int n_sectors;        /* non-empty tau sectors                                */
  int n_items;          /* contraction work items                               */
  int64_t c_elems;      /* sum_tau L_tau R_tau                                  */
  int64_t g_elems;      /* sum_tau n_tau^2, n = min(L, R)                      */
  int64_t sv_total;     /* sum_tau n_tau                                        */
  double flops_alg;     /* 8 sum_{cl>=ch>=cr} nL nC nR (complex MAC = 8 flops) */
  double flops_tau;     /* 8 sum_tau sum_{cl>=tau>=cr} sum_{cl>=ch>=cr} nL nC nR */
  double bytes_alg16;   /* B_alg with s = 16 bytes                               */
} bsmpo_plan_info;
int bsmpo_plan_info_get(const void *h_plan, bsmpo_plan_info *info);

/* ------------------------------------------------------- the site update -- */

/*
 * One beam-splitter site update on the pair (mode, mode+1), RankProcess
 * semantics (src/beam_splitter_parallel/Rank.py:32-206) with the master's
 * merge (src/beam_splitter_parallel/Full.py:113-137):
 *   C_tau = sum_c Glc[l, c] Gcr[c, r] u(cl, cc, cr; tau),  T = C diag(LR),
 *   T = V S W' (FP64), W = V^H C, global top-chi of S over all sectors,
 *   stable sort by charge, scatter.
 * In place on the caller's MPO buffers:
 *   d_Glc, d_Gcr : device chi x chi complex128, Gamma[mode], Gamma[mode+1]
 *                  (read, then overwritten with Gamma0Out / Gamma1Out)
 *   d_LR         : device chi float64 (Lambda[mode+1], or ones at the edge)
 *   d_U3         : device d^3 complex128 (bsmpo_bs_table)
 *   d_charge_c   : device int32 [chi][2], new centre charges (padded d)
 *   d_lambda_c   : device float64 [chi], new centre Lambda (padded 0)
 *   edge         : nonzero for the last pair (mode == M-2): only columns /
 *                  rows < min(chi, d*d) of the outputs are stored
 *                  (Full.py:130-132)
 *   precision    : enum bsmpo_precision for the sector contraction; the
 *                  SVD, truncation, BS elements stay FP64
 *   d_work       : device workspace of work_bytes (bsmpo_plan_query)
 *   d_num_lambda : device int32[1], number of kept singular values
 * The host plan must come from bsmpo_plan_build on the same charges.
 */
int bsmpo_site_update(int d, int chi, const void *h_plan, double *d_Glc, double *d_Gcr,
                      const double *d_LR, const double *d_U3, int32_t *d_charge_c,
                      double *d_lambda_c, int32_t *d_num_lambda, int edge, int precision,
                      void *d_work, size_t work_bytes, void *stream);

/*
 * Stage entry point of the same update (used by the update_MPO / update_MPS
 * drop-ins and by tests): the sector contraction alone.  Writes the packed
 * C_tau and T_tau = diag(LL[rows]) C_tau diag(LR[cols]) (row-major
 * L_tau x R_tau each, sectors in plan order) into d_C and d_T, complex128.
 * Optional d_LL (chi float64, row weights of T) and d_LC (chi float64,
 * centre weights inside the sum) may be NULL; they give the pure-state
 * variant of src/update_kernel.cu:594-595,706-715.
 * Replaces src/cuda_kernels.py:101-127 (update_MPO) + :72-90 (update_MPS),
 * src/update_MPO_kernel.cu:193-734 and src/update_kernel.cu:193-769
 * (`kernel`), and the LR weight and compaction of
 * src/beam_splitter_parallel/Rank.py:115-118.
 */
int bsmpo_contract(int d, int chi, const void *h_plan, const double *d_Glc,
                   const double *d_Gcr, const double *d_LR, const double *d_U3,
                   double *d_C, double *d_T, const double *d_LL, const double *d_LC,
                   int precision, void *d_work, size_t work_bytes, void *stream);

/*
 * bsmpo_contract without a workspace: the same contraction and outputs, with
 * the staging it needs (update descriptor, work lists and the plan's device
 * part) kept in an 8 MiB array the library holds on each device it runs on
 * (until the process ends).  It calls no allocating routine (no hipMalloc,
 * hipFree, pinned or pageable heap allocation on the success path), which
 * suits callers that cannot allocate: tight serving loops and harnesses that
 * forbid allocation.  Plans whose staging exceeds 8 MiB are rejected (< 0,
 * with the size in bsmpo_last_error); chi <= 2048 at d = 13 fits.  The call
 * synchronises `stream` once, after uploading the plan, because the host
 * image it uploads from is shared by all calls.  Calls that overlap on one
 * device race on its array: issue them one at a time (the kernels of one
 * call must finish before the next call's upload overwrites the staging,
 * which the synchronisation gives on a single stream).  Stage timing must be
 * off.  No counterpart in the reference.
 */
int bsmpo_contract_noalloc(int d, int chi, const void *h_plan, const double *d_Glc,
                           const double *d_Gcr, const double *d_LR, const double *d_U3,
                           double *d_C, double *d_T, const double *d_LL, const double *d_LC,
                           int precision, void *stream);

/* Copies the device offsets of the packed per-sector buffers into h_out:
 * for sector s: h_out[4 s + 0..3] = tau0, tau1, L, R, and the packed C/T
 * element offset is returned in h_coff[s].  Host only. */
int bsmpo_plan_sectors(const void *h_plan, int32_t *h_out, int64_t *h_coff,
                       int32_t *h_rows, int32_t *h_cols);

/* ------------------------------------------------------ grouped updates -- */

/*
 * Any number of independent updates (all beam splitters of a brick-wall
 * layer, across any number of interferometer instances) run by ONE launch per
 * stage.  Each update has its own plan (bsmpo_plan_build) and its own MPO
 * buffers; the updates must not share Gamma / charge / Lambda rows.
 *   bsmpo_group_query : sizes of the host image and of the device workspace
 *   bsmpo_group_build : writes the host image (h_group, preferably pinned)
 *                       for this d_work; the image holds device addresses,
 *                       so it is only valid with the d_work it was built for
 *   bsmpo_group_run   : one host-to-device copy of the image, then the
 *                       contraction, SVD and truncation stages on `stream`
 * The per-update semantics are those of bsmpo_site_update.  No counterpart in
 * the reference, which dispatches one update per MPI rank
 * (src/beam_splitter_parallel/Full.py:149-166).
 */
typedef struct {
  double *Glc;            /* Gamma[mode]   (chi x chi complex128, in/out)   */
  double *Gcr;            /* Gamma[mode+1] (chi x chi complex128, in/out)   */
  const double *LR;       /* Lambda[mode+1] or ones at the edge (chi f64)   */
  const double *U3;       /* d^3 complex128 BS table                         */
  int32_t *charge_c;      /* new centre charges [chi][2] (out)               */
  double *lambda_c;       /* new centre Lambda [chi] (out)                   */
  int32_t *num_lambda;    /* kept count [1] (out)                            */
  int32_t edge;           /* last pair of the chain (Full.py:130-132)        */
  int32_t reserved;
} bsmpo_update_io;
int bsmpo_group_query(int n_upd, const void *const *h_plans, size_t *host_bytes,
                      size_t *work_bytes);
int bsmpo_group_build(int n_upd, const void *const *h_plans, const bsmpo_update_io *io,
                      void *d_work, void *h_group, size_t host_bytes);
int bsmpo_group_run(const void *h_group, int precision, void *d_work, size_t work_bytes,
                    void *stream);

/*
 * A group run captured as a HIP graph: the image upload and the fixed launch
 * sequence of bsmpo_group_run, recorded on `stream` (a non-default stream,
 * timing disabled) and instantiated.  bsmpo_graph_launch replays it on any
 * stream; the image upload reads h_group at replay time, so rebuilding the
 * same group (same plans, same d_work) into h_group and replaying repeats
 * the run on the buffers the image names.  No reference counterpart (the
 * reference launches per sector, src/beam_splitter_parallel/Rank.py:90-127).
 */
int bsmpo_group_graph_create(const void *h_group, int precision, void *d_work, size_t work_bytes,
                             void *stream, void **graph_exec);
int bsmpo_graph_launch(void *graph_exec, void *stream);
int bsmpo_graph_destroy(void *graph_exec);

/* ----------------------------------------------------- layer programs -- */

/*
 * A layer program: the site updates of brick-wall layers (all beam splitters
 * of a layer, across any number of interferometer instances), planned ON THE
 * DEVICE from the bond charges in HBM and run by the grouped stages, captured
 * once as HIP graphs that are replayed for every layer.  Replay r applies
 * layer r: update set r % n_sets (a brick-wall mesh has two, the even and
 * the odd pairs), BS tables from d_B[r].  No host round trip per layer: no
 * host planning, no charge read-back, no image upload.  A replay is a fixed
 * launch sequence of captured graphs (the planning; per group its head, its
 * two Householder block shapes and its tail) on the caller's stream and
 * three streams of the program, ordered by events (HIP's graph executor runs
 * one graph's nodes one after another, so the layer's concurrency comes from
 * the streams).  Replaces, for the
 * brick-wall sweep of src/beam_splitter_parallel/Full.py:149-166 (one
 * RankProcess per beam splitter with host-side Aligner bookkeeping,
 * src/mpo_sort.py:215-384 and src/beam_splitter_parallel/Rank.py:78-105),
 * the per-layer host work; the per-update semantics are bsmpo_site_update's.
 *
 * Each update set is split into n_groups contiguous groups that run as
 * n_streams chains (1 to 4; group g on chain g % n_streams, each chain with
 * its own scratch arena); every group has the capacities given (bytes of its
 * variable plan part and of its scratch, entries of each work list).
 * A replay whose plan does not fit (or has bad charges, or no BS matrices)
 * changes nothing: the program records the status and the replay index, and
 * every later replay does nothing too until bsmpo_layer_seek.
 * bsmpo_layer_state_read returns that record and what each group needed in
 * the last planned replay (a synchronising read).
 */
typedef struct {
  const int32_t *CL, *CC, *CR; /* charges of bonds mode, mode+1, mode+2 [chi][2]  */
  double *Glc, *Gcr;           /* Gamma[mode], Gamma[mode+1] (in/out)             */
  const double *LR;            /* Lambda[mode+1], or ones at the edge             */
  int32_t *charge_c;           /* = CC: the new centre charges are written there */
  double *lambda_c;            /* Lambda[mode] (out)                              */
  int32_t *num_lambda;         /* kept count [1] (out)                            */
  int32_t edge;
  int32_t first_replay;        /* replays r < first_replay skip the update (the
                                  light cone, Full.py:152-153); 0: every replay   */
} bsmpo_layer_update;
typedef struct {
  int64_t var_bytes, work_bytes;          /* plan pools and items; scratch          */
  int32_t tiles, gram, jac, tri, sec, v, w, wyt, wya, status; /* list entries     */
} bsmpo_layer_caps;  /* capacities per group; also what a group needed (+ status) */
typedef struct {
  int d, chi, precision;
  int n_sets;                          /* 1 or 2 update sets: replay r runs set r % n_sets */
  int n_upd[2];                        /* updates of each set                        */
  const bsmpo_layer_update *upd[2];    /* host arrays of n_upd[s] records            */
  int n_groups, n_streams;             /* groups per set; chains / arenas (1..4)    */
  int n_layers;                        /* replays with BS matrices in d_B            */
  const double *d_B;                   /* device [n_layers][max n_upd][2][2] c128:
                                          replay r, update u of its set              */
  bsmpo_layer_caps caps;               /* per group                                  */
} bsmpo_layer_spec;
typedef struct {
  int32_t status;      /* 0 ok; 1 over capacity; 2 bad charges; 3 tile too big;
                          4 sector too big; 5 no BS matrices; 6 skipped          */
  int32_t replays;     /* replay counter (the next replay's layer)               */
  int32_t failed_at;   /* replay of the first failure (when status != 0)         */
  int32_t n_groups;
} bsmpo_layer_state;
/* Bytes of the caller workspace a program needs. */
int bsmpo_layer_query(const bsmpo_layer_spec *spec, size_t *work_bytes);
/* Builds and captures the program (on `stream`, a non-default stream, stage
 * timing disabled; the call synchronises that stream).  d_work stays owned
 * by the caller and must outlive the program. */
int bsmpo_layer_create(const bsmpo_layer_spec *spec, void *d_work, size_t work_bytes,
                       void *stream, void **prog);
/* Sets the replay counter to `replay` and clears the recorded failure. */
int bsmpo_layer_seek(void *prog, int replay, void *stream);
/* Replays the layer on `stream` (no host synchronisation; the program's own
 * streams join `stream` before it returns control to later work on it). */
int bsmpo_layer_launch(void *prog, void *stream);
/* Synchronising read of the program's state; group_needs (may be NULL) gets
 * n_groups records: what each group needed in the last planned replay. */
int bsmpo_layer_state_read(void *prog, bsmpo_layer_state *state, bsmpo_layer_caps *group_needs);
int bsmpo_layer_destroy(void *prog);
/* What a group of host plans needs as a device-planned group (to size
 * capacities), in the layout bsmpo_layer_create uses. */
int bsmpo_layer_needs(int n_upd, const void *const *h_plans, bsmpo_layer_caps *needs);

/* --------------------------------------------------------- observables -- */

/*
 * Right environments over the ket = bra diagonal of an M-site MPO
 * (src/summary.py:11-26 chain):  R[M] = e_0,
 * R[b] = Gamma[b] (R[b+1] * mask[b+1]), mask = (c0 == c1 < d) and, on inner
 * bonds, Lambda > 0.  Left environments the same from the left.
 *   d_Gamma  : device M x chi x chi complex128
 *   d_Lambda : device (M-1) x chi float64
 *   d_charge : device (M+1) x chi x 2 int32
 *   d_R, d_L : device (M+1) x chi complex128
 *   d_marg   : device M x d float64, P(n photons in mode j) (unnormalised)
 *   d_trace  : device float64[1], sum over bond 0 of R[0] (the "norm")
 */
int bsmpo_observables(int M, int d, int chi, const double *d_Gamma, const double *d_Lambda,
                      const int32_t *d_charge, double *d_R, double *d_L, double *d_marg,
                      double *d_trace, void *stream);

/*
 * Entanglement entropy of every inner bond of an M-site MPO:
 *   d_S[b] = -sum_x p_x log2 p_x,  p = Lambda[b]^2 / sum_x Lambda[b]^2,
 * b = 0..M-2 (d_Lambda device (M-1) x chi float64, d_S device M-1 float64);
 * an all-zero bond gives 0.  Replaces src/summary.py:3-8,28-30
 * (EntanglementEntropy, EntropyFromColumn, ColumnSumToOne).
 */
int bsmpo_entanglement_entropy(int M, int chi, const double *d_Lambda, double *d_S, void *stream);

/* ------------------------------------------------------ reverse mode -- */
/*
 * The reference has no derivative code.  These entry points give the
 * adjoint of the site update above (src/beam_splitter_parallel/Rank.py:90-196
 * semantics with its truncation held fixed), of the BS table
 * (src/u_kernel.cu:37-70 values) and of the observables
 * (src/summary.py:11-26 chain), so that a caller can run reverse mode through
 * a whole mesh.  Gradients of a real loss L are complex arrays
 * g = dL/dRe z + i dL/dIm z (PyTorch's convention for complex tensors).
 *
 * bsmpo_site_update_vjp: one update backwards.
 *   h_plan          : plan of the update's INPUT charges (bsmpo_plan_build)
 *   h_charge_new    : host int32 [chi][2], the forward's new centre charges
 *   d_Glc, d_Gcr,   : the forward's inputs (Gamma[mode], Gamma[mode+1],
 *   d_LR, d_U3        LR, BS table), unchanged copies
 *   d_G0new, d_G1new, d_lambda_new : the forward's outputs (separate buffers)
 *   d_gG0, d_gG1, d_gLam           : gradients of those outputs
 *   d_gGlc, d_gGcr (chi x chi complex), d_gU3 (d^3 complex): gradients of
 *                     the inputs, written (gGlc may alias gG0, gGcr gG1)
 *   d_gLR (chi)     : gradient of LR, ADDED to what the buffer holds
 *   precision       : the forward's contraction precision (recomputed)
 * The workspace size depends on the kept slots: query it with the same
 * plan and new charges.  The call synchronises the stream once (it uploads
 * per-call tables from pageable host memory).
 */
int bsmpo_site_update_vjp_query(int d, int chi, const void *h_plan, const int32_t *h_charge_new,
                                size_t *work_bytes);
int bsmpo_site_update_vjp(int d, int chi, const void *h_plan, const int32_t *h_charge_new,
                          const double *d_Glc, const double *d_Gcr, const double *d_LR,
                          const double *d_U3, const double *d_G0new, const double *d_G1new,
                          const double *d_lambda_new, const double *d_gG0, const double *d_gG1,
                          const double *d_gLam, double *d_gGlc, double *d_gGcr, double *d_gLR,
                          double *d_gU3, int precision, void *d_work, size_t work_bytes,
                          void *stream);

/* d_gB[nb][2][2] complex128 = sum gU3 conj(dU3/dB) for the tables of
 * bsmpo_bs_table(d, nb, d_B, ...); d_gU3 is [nb][d][d][d] complex128. */
int bsmpo_bs_table_vjp(int d, int nb, const double *d_B, const double *d_gU3, double *d_gB,
                       void *stream);

/*
 * L = -log P(pattern) + sum_j w_j <n_j> of an M-site MPO and dL/dGamma:
 *   P(pattern) : the diagonal chain with site j's local occupation (ket charge
 *                difference across the site) restricted to pattern[j]
 *   <n_j>      : mean photon number of mode j (the marginals' first moment)
 *   d_pattern  : device int32 [M];  d_weights : device float64 [M]
 *   d_env      : device scratch, 6 (M+1) chi complex128
 *   d_gGamma   : device M x chi x chi complex128 (written)
 *   d_out      : device float64 [3] = {L, P, sum_j w_j <n_j>}
 */
int bsmpo_observable_vjp(int M, int d, int chi, const double *d_Gamma, const double *d_Lambda,
                         const int32_t *d_charge, const int32_t *d_pattern,
                         const double *d_weights, double *d_env, double *d_gGamma, double *d_out,
                         void *stream);

#ifdef __cplusplus
}
#endif
#endif /* BSMPO_H */
