/*
 * qz_oracle.h -- CPU restatement of the PDHGEQZ / PDGGHRD path on a 1x1 process grid.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP implementation in
 * gevp-pdhgeqz_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product never links or calls it.
 *
 * Conventions: column-major FP64, 0-based indices, leading dimensions >= n.
 * Reference citations are relative to the NLAFET/GEVP-PDHGEQZ source tree.
 *
 * What pins each function (tests/test_oracle.py unless named; the reference ships no golden
 * vectors, so the pins are closed forms, invariants, LAPACK via scipy and brute force):
 *   orc_fill_*, orc_splitmix64    committed generator fixture (test_generator_fixture)
 *   orc_lartg, orc_larfg          closed forms and annihilation (test_lartg_larfg)
 *   orc_shift_col                 closed form of the shift polynomial's first column
 *   orc_invhse3                   annihilation of T's first column (test_invhse_annihilates)
 *   orc_lagv2                     standardised 2x2 forms, real and complex (test_lagv2_2x2_forms)
 *   orc_eig22                     roots of the 2x2 characteristic quadratic
 *   orc_tgex2, orc_tgexc          swapped spectra, structure and residuals
 *   orc_geqrf_apply, orc_gghrd    HT structure and PCOMPRES residuals
 *   orc_qz_ds                     Schur form, residuals and LAPACK eigenvalues
 *   orc_hgeqz                     LAPACK eigenvalues (n = 1..260), the golden LAPACK fixture,
 *                                 brute-force roots at n <= 6, exact triangular and B = I cases,
 *                                 exactly known infinite counts
 *   orc_chase_window              invariants of a chased chain (test_oracle_chase_window_equivalence)
 *   orc_compres                   agreement with an independent numpy PCOMPRES
 *   orc_piparmq_*                 the published PIPARMQ table (test_piparmq_table)
 *   orc_mlp_forward, orc_mlp_grad closed form and central finite differences
 * Parity unpinned (would need a run of the reference): the iteration trajectory of orc_hgeqz
 * (per-iteration KTOP, KBOT, ND), the order of its eigenvalues and the particular S, T, Q, Z.
 */
#ifndef QZ_ORACLE_H
#define QZ_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-call tuning (PIPARMQ ISPEC 51/52/53 and the small-block cut-over).  A zero field
 * means "use the PIPARMQ default for the active block size" (src/pdhgeqz/piparmq.f:445-517). */
typedef struct {
    int nmin;      /* blocks <= nmin go to the double-shift base QZ (KKQZ/QZLAP role)          */
    int ns;        /* shifts per sweep                                                          */
    int nw;        /* AED window                                                                */
    int nibble;    /* skip the sweep when 100*ND > NIBBLE*JW (PDHGEQZ0.f:989-990)              */
    int chain;     /* shifts per bulge chain (SHIFTPERWND, PDHGEQZ0.f:1145-1148)               */
    int maxit;     /* iteration cap per active-block element (30 in the reference)             */
} orc_params;

typedef struct {
    long iters;          /* outer QZ iterations (all levels)     */
    long aed_calls;
    long sweeps;
    long shifts_used;
    long ds_iters;       /* double-shift base iterations          */
    long swaps;          /* accepted block swaps during AED        */
    long swap_fail;      /* rejected swaps                          */
    long inf_defl;       /* infinite eigenvalues deflated           */
} orc_stats;

/* ---- generators ---- */
uint64_t orc_splitmix64(uint64_t *state);
void orc_fill_normal(uint64_t seed, size_t count, double *out);
void orc_fill_uniform(uint64_t seed, size_t count, double *out);

/* ---- elementary kernels (exported for kernel-level parity tests) ---- */
void orc_lartg(double f, double g, double *c, double *s, double *r);
void orc_larfg(int n, double *alpha, double *x, int incx, double *tau);
/* first column of the double-shift polynomial at the top-left of (H,T) */
void orc_shift_col(const double *H, int ldh, const double *T, int ldt,
                   double sr1, double si1, double sb1, double sr2, double si2, double sb2,
                   double v[3]);
/* INVHSE: opposite Householder v,tau with T(0:3,0:3)*(I - tau v v^T) zero in rows 1:3 of col 0 */
void orc_invhse3(const double *T, int ldt, double v[3], double *tau);
/* DLAGV2: standardise a 2x2 block of (A,B) in place: A <- QL^T A ZR, B <- QL^T B ZR. */
void orc_lagv2(double *A, int lda, double *B, int ldb, double alphar[2], double alphai[2],
               double beta[2], double QL[4], double ZR[4]);
/* eigenvalues (alpha_r, alpha_i, beta) of a 2x2 pencil with B upper triangular */
void orc_eig22(const double *A, int lda, const double *B, int ldb, double ar[2], double ai[2],
               double b[2]);

/* PDTGORD (src/reorder/PDTGORD.f:1-4) restated serially: selected eigenvalues to the top by DTGEXC
 * moves, 2x2 blocks selected when either row is; returns 0 or the 1-based row of a rejected swap. */
int orc_tgord(int n, double *S, int lds, double *T, int ldt, double *Q, int ldq, double *Z, int ldz,
              int *sel, int *m);

/* ---- HT reduction (PDGGHRD path) ---- */
/* QR of B (Householder), A <- Q^T A; Q <- Q*Q_B when compq != 0 (EXRAND2.f:296-329). */
int orc_geqrf_apply(int n, double *A, int lda, double *B, int ldb, double *Q, int ldq, int compq);
/* Givens HT reduction of (A,B) with B upper triangular (PDGGHRD.f:365-642 semantics,
 * unblocked).  Q and Z follow PDHGEQZ's convention: A_out = Q^T A_in Z. */
int orc_gghrd(int n, int ilo, int ihi, double *A, int lda, double *B, int ldb,
              double *Q, int ldq, double *Z, int ldz, int compq, int compz);

/* ---- QZ ---- */
/* job: 0 = eigenvalues only ('E'), 1 = Schur form ('S');
 * compq/compz: 0 = 'N', 1 = 'I', 2 = 'V'.  Returns INFO as PDHGEQZ.f:496-502. */
int orc_hgeqz(int job, int compq, int compz, int n, int ilo, int ihi,
              double *H, int ldh, double *T, int ldt,
              double *alphar, double *alphai, double *beta,
              double *Q, int ldq, double *Z, int ldz,
              const orc_params *prm, orc_stats *st);

/* double-shift base QZ on the active block [ilo,ihi] of an n x n pencil; transforms applied
 * to full rows/columns and to Q,Z (n rows) when non-NULL. */
int orc_qz_ds(int n, int ilo, int ihi, double *H, int ldh, double *T, int ldt,
              double *alphar, double *alphai, double *beta,
              double *Q, int ldq, double *Z, int ldz, double ttol, orc_stats *st);

/* DTGEX2-style swap of adjacent blocks at j1 (sizes n1, n2 in {1,2}).  Returns 0 on success,
 * 1 when the swap is rejected by the stability tests (src/reorder/dtgex2.f). */
int orc_tgex2(int n, double *A, int lda, double *B, int ldb, double *Q, int ldq,
              double *Z, int ldz, int j1, int n1, int n2);
/* DTGEXC: move the block at ifst to ilst < ifst by adjacent swaps (src/reorder/dtgexc.f). */
int orc_tgexc(int n, double *A, int lda, double *B, int ldb, double *Q, int ldq,
              double *Z, int ldz, int *ifst, int *ilst);

/* Chase one tightly packed chain of nb bulges through a window (PDHGEQZB.f:281-491 with
 * DHGEQZ5.f:128-275): the window wsize x wsize pencil (Aw,Bw) holds bulges at rows
 * 1+3b .. 3+3b (b=0..nb-1); each is moved movedist steps, bottom bulge first.  U,V
 * (wsize x wsize) accumulate the left/right transforms.  last != 0 pushes the bulges off the
 * bottom of the window (NCOLS=2 handling).  shifts: 2*nb (sr,si,sb) triples for re-introduction. */
void orc_chase_window(int wsize, double *Aw, int lda, double *Bw, int ldb, double *U, double *V,
                      int ldu, int nb, int movedist, int last, const double *sr,
                      const double *si, const double *sb);

/* Introduce nb bulges at the top of a wsize x wsize window (PDHGEQZA.f:245-384): bulge k from
 * shift pair (2k, 2k+1); before each new bulge the ones already introduced (oldest first) move 3
 * steps.  U, V (wsize x wsize, ld ldu) accumulate the left / right transforms. */
void orc_intro_window(int wsize, double *Aw, int lda, double *Bw, int ldb, double *U, double *V,
                      int ldu, int nb, const double *sr, const double *si, const double *sb);

/* PCOMPRES (tools/PQZHLPS.f:19-319): resa = ||S - Q^T A Z||_F/||A||_F, resb likewise,
 * resq = ||I - Q^T Q||_F/(n eps), resz likewise.  out[0..3]. */
void orc_compres(int n, const double *A, const double *B, const double *S, const double *T,
                 const double *Q, const double *Z, int ld, double out[4]);

/* PIPARMQ defaults (src/pdhgeqz/piparmq.f:445-517) */
int orc_piparmq_ns(int nh);
int orc_piparmq_nw(int nh);
int orc_piparmq_nibble(int nh);

/* perceptron of the learned parameter policy (mlp_oracle.c) */
int orc_mlp_nparams(void);
void orc_mlp_forward(const double *P, const double *X, int B, double *Y);
double orc_mlp_grad(const double *P, const double *X, const double *T, int B, double *G);

#ifdef __cplusplus
}
#endif
#endif
