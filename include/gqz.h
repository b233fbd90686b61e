/*
 * gqz.h -- C ABI of the MI355X multishift QZ (libgqz.so).
 *
 * Drop-in boundary for the PDHGEQZ / PDGGHRD path of NLAFET/GEVP-PDHGEQZ on a 1x1 process grid.
 * All matrices are column-major FP64.  "_dev" entry points take device (HBM) pointers and run
 * on the library's HIP stream; the host-pointer entry points copy in, run, and copy out.
 * Indices in this header are 1-based where they mirror a Fortran argument (ILO, IHI) and
 * 0-based nowhere else.  No torch types cross this boundary.
 * Re-entrant across host threads (the reference keeps its tolerances in COMMON blocks and is not):
 * each host thread has its own context (streams, workspace arena, pinned staging, settings), created
 * on device 0 by its first call unless it called gqz_init.  Calls from one thread are sequential.
 */
#ifndef GQZ_H
#define GQZ_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Tuning of one call.  A zero field selects the PIPARMQ default for the active block size
 * (src/pdhgeqz/piparmq.f:445-517: NS table, NW = 3NS/2 above 500, NIBBLE). */
typedef struct {
    int nmin;      /* blocks <= nmin use the LDS-resident base QZ (<= 64)                     */
    int ns;        /* shifts per sweep (ISPEC 53)                                               */
    int nw;        /* AED window (ISPEC 51)                                                     */
    int nibble;    /* sweep-skip threshold in percent (ISPEC 52)                                */
    int chain;     /* shifts per bulge chain (SHIFTPERWND, PDHGEQZ0.f:1145-1148); <= 40; 0: NS/6 in [10, 20] */
    int maxit;     /* iteration cap factor (MAXIT = maxit * NH, 30 in the reference)           */
    int use_policy;/* 1: consult the installed parameter policy at every top-level iteration  */
} gqz_params;

typedef struct {
    long iters, aed_calls, sweeps, shifts_used, small_calls, swaps, swap_fail, inf_defl;
    long chase_windows, slab_launches;
    double t_total, t_aed, t_sweep, t_small, t_inf;   /* seconds (host-timed, stream-synced) */
} gqz_stats;

/* Per top-level iteration record (the learned-parameter trainer's input). */
typedef struct {
    int nh, ktop, kbot, ns, nw, nibble, nd, ndinf, swept;
    float feat[12];
    double t_iter;
} gqz_trace_rec;

/* Parameter policy: called at each top-level iteration with the 12 deflation-history features and
 * the order nh of the active block; ns, nw, nibble hold the default values (PIPARMQ of the block
 * passed to the QZ call) on entry and may be overwritten (<= 0 keeps the default). */
typedef void (*gqz_policy_fn)(void *user, const float feat[12], int nh, int *ns, int *nw, int *nibble);

/* ---- library ---------------------------------------------------------------------------- */
int gqz_init(int device);                 /* select the device, create stream + workspace    */
int gqz_finalize(void);                   /* releases the device state; -1 after a HIP error  */
const char *gqz_version(void);
void *gqz_stream(void);                   /* hipStream_t used by the _dev entry points        */
int gqz_sync(void);
int gqz_set_policy(gqz_policy_fn fn, void *user);
int gqz_trace_enable(int cap);            /* record up to cap iterations (0 disables)         */
int gqz_trace_get(gqz_trace_rec *out, int cap);   /* returns the number of records          */
/* Progress of long QZ calls (no counterpart in the reference): with seconds > 0, a QZ call writes
 * one line to stderr (top-level iteration, active block, sweeps and AED calls so far) at most every
 * `seconds`; 0 turns it off (the default). */
int gqz_progress(double seconds);
/* Sweep update overlap (no counterpart in the reference): with on = 1 (the default) the updates
 * outside the chase windows of one round run on a second stream of the library (same priority)
 * while the next round is chased; 0 runs them on the library stream after each chase (same
 * products, in a different order).  on < 0 only queries.  Returns the previous setting; per host
 * thread, like the rest of the context. */
int gqz_overlap(int on);
/* Order of the smallest active block whose sweeps use the overlap (default 2048; smaller blocks, such as
 * the nested QZs of AED windows, keep their updates on the library stream).  n < 0 only queries.
 * Returns the previous value; per host thread.  No counterpart in the reference. */
int gqz_overlap_min(int n);
/* Replay of the sweep's chain scheduler on the host, without a device (a test hook; no counterpart in the
 * reference): nch chains of nb[c] bulges (1 <= nb[c], 3 nb[c] + 2 <= 64) chased over rows [ktop, kbot]
 * of an order-n pencil, with the overlapped updates (ovl = 1) or without.  Writes one record of 8 ints
 * per window launched: round, chain, p, w, steps, last, nc (near columns its U leaves pending), near
 * (the pending columns of its previous window that this launch applies first); a record with chain = -1
 * marks a point where the scheduler let the second stream finish.  Returns the number of records
 * (only the first cap are written) or -1 (bad arguments, or the scheduler stalled). */
int gqz_sched_replay(int n, int ktop, int kbot, int nch, const int *nb, int ovl, int *out, int cap);
const char *gqz_last_error(void);

/* ---- QZ ----------------------------------------------------------------------------------
 * Replaces PDHGEQZ (src/pdhgeqz/PDHGEQZ.f:15-31) for a 1x1 grid.
 * job: 'E' or 'S'; compq, compz: 'N', 'I', 'V' (PDHGEQZ.f:132-161).  ilo, ihi are 1-based.
 * On exit H = Q^T H_in Z (Schur form), T likewise, alphar/alphai/beta per PDHGEQZ.f:51-66
 * (beta >= 0, beta = 0 exactly for infinite eigenvalues).  Returns INFO (PDHGEQZ.f:496-502):
 * < 0 argument error, > 0 no convergence (index of the unconverged bottom), 0 success.
 * The device entry point takes one leading dimension for H and T: ldt must equal ldh (INFO = -8
 * otherwise; PDHGEQZ's DESCH and DESCT are separate).  The host entry points (gqz_hgeqz,
 * gqz_pdhgeqz_host) accept any ldh, ldt, ldq, ldz: they copy into device arrays of their own. */
int gqz_hgeqz_dev(char job, char compq, char compz, int n, int ilo, int ihi,
                  double *H, int ldh, double *T, int ldt,
                  double *alphar, double *alphai, double *beta,
                  double *Q, int ldq, double *Z, int ldz,
                  const gqz_params *prm, gqz_stats *stats);
int gqz_hgeqz(char job, char compq, char compz, int n, int ilo, int ihi,
              double *H, int ldh, double *T, int ldt,
              double *alphar, double *alphai, double *beta,
              double *Q, int ldq, double *Z, int ldz,
              const gqz_params *prm, gqz_stats *stats);

/* ---- HT reduction --------------------------------------------------------------------------
 * Replaces PDGGHRD (src/ab2ht/PDGGHRD.f:13-14).  B must be upper triangular on entry.
 * Q/Z convention here is PDHGEQZ's (A_out = Q^T A_in Z); the Fortran shim transposes Q to
 * PDGGHRD's own convention (PDGGHRD.f:25-27). */
int gqz_gghrd_dev(char compq, char compz, int n, int ilo, int ihi, double *A, int lda,
                  double *B, int ldb, double *Q, int ldq, double *Z, int ldz);
int gqz_gghrd(char compq, char compz, int n, int ilo, int ihi, double *A, int lda,
              double *B, int ldb, double *Q, int ldq, double *Z, int ldz);
/* QR of B applied to A (and accumulated into Q): the pre-step of the drivers
 * (testing/EXRAND2.f:296-329; PDGEQRF + PDORMQR). compq: 'N', 'I', 'V'. */
int gqz_geqrb_dev(char compq, int n, double *A, int lda, double *B, int ldb, double *Q, int ldq);

/* ---- ScaLAPACK-descriptor surface for the Fortran shim (1x1 grid) ----------------------------
 * desc[9] = DTYPE, CTXT, M, N, MB, NB, RSRC, CSRC, LLD (PDHGEQZ.f:88-127).  Host pointers.
 * gqz_pdhgeqz_host: workspace query when ldwork == -1 or liwork == -1, after the argument checks
 * (PDHGEQZ.f:384-430).  gqz_pdgghrd_host: lwork == -1 returns MB in work[0] before any check
 * (PDGGHRD.f:256-266); INFO -1..-7 as PDGGHRD.f:318-336. */
int gqz_pdhgeqz_host(char job, char compq, char compz, int n, int ilo, int ihi,
                     double *H, const int *desch, double *T, const int *desct,
                     double *alphar, double *alphai, double *beta,
                     double *Q, const int *descq, double *Z, const int *descz,
                     double *dwork, int ldwork, int *iwork, int liwork);
int gqz_pdgghrd_host(char compq, char compz, int n, int ilo, int ihi,
                     double *A, const int *desca, double *B, const int *descb,
                     double *Q, const int *descq, double *Z, const int *descz,
                     double *work, int lwork);

/* PDGGHRD with the routine's implemented Q convention (rotations applied to Q from the left;
 * the drivers transpose Q afterwards, testing/EXRAND2.f:346-356): COMPQ='I' returns Q_ht^T,
 * COMPQ='V' returns Q_ht^T * Q1 where A_out = Q_ht^T A_in Z.  Host pointers; runs on the device. */
int gqz_gghrd_rowq(char compq, char compz, int n, int ilo, int ihi, double *A, int lda,
                   double *B, int ldb, double *Q, int ldq, double *Z, int ldz);
/* ---- eigenvalue reordering (PDTGORD, src/reorder/PDTGORD.f:1-4) --------------------------------
 * Reorders a generalized real Schur form (S, T) so that the selected eigenvalues come first; Q, Z
 * are postmultiplied by the transformation when wantq / wantz.  select (host, n ints, in/out): a 2x2
 * block is selected when either of its rows is; on exit it shows the final positions.  para (host,
 * 6 ints, PILAENVX ISPEC 80-85 values; may be null): para[2] caps the window (<= 64 here), para[1]
 * the eigenvalues carried per pass.  *m = number of selected eigenvalues.  alphar/alphai/beta
 * (device) receive the eigenvalues of the reordered pencil (beta >= 0).  Returns 0, or the 1-based
 * row of the pair whose swap was rejected (partially reordered pencil, still an orthogonal
 * equivalence), or < 0 for an argument error / -100 for a HIP error.  Device pointers, lds == ldt. */
int gqz_tgord_dev(int wantq, int wantz, int *select, const int *para, int n, double *S, int lds,
                  double *T, int ldt, double *Q, int ldq, double *Z, int ldz,
                  double *alphar, double *alphai, double *beta, int *m);
/* The descriptor surface of PDTGORD (host pointers, 1x1 grid): argument order of PDTGORD.f:1-4,
 * workspace query with LDWORK = -1 or LIWORK = -1, INFO = -31 / -33 for too small workspaces. */
int gqz_pdtgord_host(int wantq, int wantz, int *select, const int *para, int n,
                     double *S, const int *descs, double *T, const int *desct,
                     double *Q, const int *descq, double *Z, const int *descz,
                     double *alphar, double *alphai, double *beta, int *m,
                     double *dwork, int ldwork, int *iwork, int liwork);

/* ---- Matrix Market input (tools/mmio.f mminfo2 / mmread2, tools/PDRDMM.f:6-90), host only ----------
 * gqz_mm_info: header of a .mtx file (rep: "coordinate" | "array", field: "real" | "integer" |
 * "pattern", symm: "general" | "symmetric" | "skew-symmetric"; buffers of at least 11, 8 and 20
 * chars, may be null).  gqz_mm_read_dense: the whole matrix into a dense column-major host array
 * (cleared first; a symmetric file sets (i, j) and (j, i), as PDRDMM does).  Returns 0, or -1 file not
 * readable, -2 not a Matrix Market matrix, -3 unsupported type (complex, Hermitian), -4 bad size line,
 * -5 an entry is missing or malformed, -6 an index is out of range, -7 size differs from
 * nrows x ncols, -8 A null or lda too small.  A Fortran caller passes the file name: Fortran unit
 * numbers do not cross the C ABI. */
int gqz_mm_info(const char *path, char *rep, char *field, char *symm, int *nrows, int *ncols, long *nnz);
int gqz_mm_read_dense(const char *path, int nrows, int ncols, double *A, int lda);

/* PIPARMQ tuning values (src/pdhgeqz/piparmq.f:300-531) for ISPEC 50..55 and active size nh;
 * the Fortran shim's PILAENVX answers from it. */
int gqz_piparmq(int ispec, int nh);

/* ---- residual check (PCOMPRES, tools/PQZHLPS.f:19-319) on device data ------------------------
 * out[0] = ||S - Q^T A Z||_F/||A||_F, out[1] likewise for B, out[2] = ||I - Q^T Q||_F/(n eps),
 * out[3] likewise for Z. */
int gqz_compres_dev(int n, const double *A, const double *B, const double *S, const double *T,
                    const double *Q, const double *Z, int ld, double out[4]);

/* ---- kernel-level entry points (parity tests of single kernels) --------------------------------
 * Slab update (the PDHGEQZ6 role): side 'R': C[m x k] <- C F ; side 'L': C[k x m] <- F^T C. */
int gqz_slab_update_dev(char side, int m, int k, double *C, int ldc, const double *F, int ldf);
/* One window of the chain chase (PDHGEQZB + DHGEQZ5 on a window): the w x w window at (p, p) of
 * H, T (ld) holds nb bulges with tail columns t0 + 3b (intro = 1: introduce them first);
 * every bulge advances `steps` columns (last = 1: push off at the window bottom).  U, V (w x w,
 * ld w) receive the window's accumulated transforms; shifts are 2*nb (sr, si, sb) on the device.
 * Only the window is modified. */
int gqz_chase_window_dev(double *H, double *T, int ld, int p, int w, int nb, int t0, int steps,
                         int intro, int last, double *U, double *V, const double *sr,
                         const double *si, const double *sb);
/* One swap of adjacent diagonal blocks of orders n1, n2 in {1, 2} (DTGEX2, src/reorder/dtgex2.f:221-860)
 * on the (n1+n2) x (n1+n2) corner (S, T) at ld (device pointers), computed by the lane that does
 * it inside the AED reorder.  out (65 device doubles): out[0] = 1 when the swap is rejected, else 0
 * and out[1..64] = Qa | Za | S' | T' (m x m, ld m) with S' = Qa^T S Za, T' = Qa^T T Za. */
int gqz_tgex2_dev(int n1, int n2, const double *S, const double *T, int ld, double *out);
/* LDS-resident base QZ of an m x m block (m <= 64) at (H, T) with U, V (m x m, ld m) and the
 * block's eigenvalues; returns the kernel's INFO. */
int gqz_small_qz_dev(int m, double *H, double *T, int ld, double *U, double *V,
                     double *alphar, double *alphai, double *beta, double ttol);

/* ---- learned parameter policy (replaces PILAENVX ISPEC 51/52/53, pilaenvx.f:1; piparmq.f:440-517) ----
 * A 12 -> 64 -> 64 -> 3 perceptron, GELU, FP32, parameters as one block of gqz_mlp_nparams_c()
 * floats: W1[64][12] b1[64] W2[64][64] b2[64] W3[3][64] b3[3] (row-major).  Inputs: the 12
 * deflation-history features of a top-level iteration (gqz_trace_rec.feat); outputs: log2
 * multipliers of the default NS, NW and NIBBLE (clamped to [-2, 1], [-5, 1], [-2, 4]).
 * All pointers of the _dev calls are device pointers. */
int gqz_mlp_nparams_c(void);
int gqz_mlp_forward_dev(const float *params, const float *x, int batch, float *y);
/* adds dL/dparams of L = mean over the batch and the 3 outputs of (y - target)^2 to grad; loss
 * (device float, may be null) receives L */
int gqz_mlp_grad_dev(const float *params, const float *x, const float *target, int batch, float *grad,
                     float *loss);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected at step >= 1); grad must hold
 * zeros on entry and is left zeroed; asynchronous on the library stream */
int gqz_mlp_train_step_dev(float *params, float *grad, float *m, float *v, int step, float lr, const float *x,
                           const float *target, int batch, float *loss);
/* install (params_host != null) and enable/disable the learned policy; it is used by QZ calls
 * whose gqz_params.use_policy is non-zero (before any gqz_set_policy callback) */
int gqz_mlp_policy(const float *params_host, int enable);

/* ---- kernel timing (no counterpart in the reference) -----------------------------------------
 * HIP events recorded on the library stream around the launches of the enabled kernel classes
 * (bit c of mask = class c; every launch, or one in k with gqz_ktime_sample); per class: launches,
 * event time and the algorithmic work of those launches (flops; bytes = each operand read once and
 * each result written once). */
enum { GQZ_KC_SLAB = 0, GQZ_KC_CHASE, GQZ_KC_SMALL, GQZ_KC_REORDER, GQZ_KC_SPIKE, GQZ_KC_HT, GQZ_KC_INF,
       GQZ_KC_COUNT };
typedef struct {
    const char *name;
    long launches;
    double ms, flops, bytes;
} gqz_ktime;
int gqz_ktime_enable(unsigned mask);          /* resets the totals */
/* time one launch in `every` (pseudo-random choice; default 1: all); ms is then the mean event time
 * of the timed launches times the number of launches, flops and bytes still cover every launch */
int gqz_ktime_sample(int every);
/* synchronises; returns the number of entries (2 * GQZ_KC_COUNT: each class at the top level, then
 * inside AED windows, named "<class>@aed") */
int gqz_ktime_get(gqz_ktime *out, int cap);

/* ---- batched pencils (no counterpart in the reference) ---------------------------------------
 * count pencils of orders n[i] stored back to back (column-major, ld = n[i]) at offsets off[i]
 * (in doubles) of the device arrays A and B; full pipeline QR(B) + HT + QZ per pencil in one
 * persistent launch.  Q, Z (same layout) may be null.  Eigenvalues at eoff[i]. */
int gqz_batched_dev(int count, const int *n, const long *off, const long *eoff,
                    double *A, double *B, double *Q, double *Z,
                    double *alphar, double *alphai, double *beta, int *info, int wg_threads);
/* n, off, eoff are host arrays; A, B, Q, Z, alphar, alphai, beta, info are device arrays.
 * wg_threads: 0 = tuned shape (or the size heuristic before tuning), else 128/256/512/1024.
 * Repeated calls with the same arguments replay a captured hipGraph. */

/* Allocation-free variant of gqz_batched_dev (no counterpart in the reference): for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, same kernels and results; the work counter and the per-pencil tables
 * live in a fixed per-GPU device array of the library (about 192 KiB, held until the process ends)
 * instead of buffers allocated per call, and the launch is direct (no graph instantiation).
 * Limits: count <= 8192 (larger counts return -1 with a message); gqz_init must have run on the
 * device (it sets the scratch up; otherwise -100).  Calls that overlap on one GPU share that array and
 * race: serialise them.  Every call resets the counter it uses. */
int gqz_batched_noalloc_dev(int count, const int *n, const long *off, const long *eoff,
                            double *A, double *B, double *Q, double *Z,
                            double *alphar, double *alphai, double *beta, int *info, int wg_threads);

/* Autotuner for the batched launch: times every workgroup size (128, 256, 512) on a
 * synthetic sample drawn at the quantiles of the size list n (host array) and keeps the fastest
 * for later gqz_batched_dev calls with wg_threads = 0 (sample: up to 64 pencils, orders capped at
 * 512; the grid then holds 4 (<= 256 threads) or 2 workgroups per CU). */
int gqz_batched_tune(int count, const int *n, int *best_threads, int *best_wgs, double *best_ms);

#ifdef __cplusplus
}
#endif
#endif
