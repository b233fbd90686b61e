// gqz_kernels.h -- kernel entry points and host-side launch helpers shared by the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct GqzSmallJob {
    double *H, *T;          // block top-left, leading dimension ld
    int ld, m;
    double *U, *V;          // m x m outputs (may be null), leading dimension ldu
    int ldu, pad;
    double *ar, *ai, *be;   // eigenvalue outputs (m entries each)
    int *info;
    double ttol, smlnum;
};

// dynamic LDS bytes of k_small_qz for a block of order m (H, T, U, V with ld = m + 1 + control)
size_t gqz_small_lds_bytes(int m);

#ifdef __HIPCC__
extern "C" __global__ void k_small_qz(const GqzSmallJob *jobs);
extern "C" __global__ void k_small_qz1(GqzSmallJob job);
extern "C" __global__ void k_gghrd_wg(double *A, double *B, int ld, int n, int ilo, int ihi,
                                      double *Q, double *Z, int ldq, int nq);
#endif

// in-place slab updates sharing one launch: C_q <- C_q F_q (right) or C_q <- F_q^T C_q (left);
// m[q] = rows (right) or columns (left) of slab q, k[q] = order of its factor F_q.  Slabs of one
// launch must not overlap.
#define GQZ_SLAB_MAXJOBS 64
struct GqzSlabJobs {
    int njobs;
    double *C[GQZ_SLAB_MAXJOBS];
    const double *F[GQZ_SLAB_MAXJOBS];
    int ldc[GQZ_SLAB_MAXJOBS];
    int m[GQZ_SLAB_MAXJOBS];
    int ldf[GQZ_SLAB_MAXJOBS];
    int k[GQZ_SLAB_MAXJOBS];
    int blk0[GQZ_SLAB_MAXJOBS + 1];   // first workgroup of each job (filled by the launcher)
};
// add a slab to a job list (ignored when empty); returns false when the list is full
bool gqz_slab_add(GqzSlabJobs &J, double *C, int ldc, int m, const double *F, int ldf, int k);
void gqz_slab_right_multi(hipStream_t st, GqzSlabJobs &J);
void gqz_slab_left_multi(hipStream_t st, GqzSlabJobs &J);

// ---- host launchers (defined in the .hip files; all asynchronous on `st`) -----------------
// C[m x k] <- C * V[k x k]   (in place, k <= 1024)
void gqz_slab_right(hipStream_t st, double *C, int ldc, int m, int k, const double *V, int ldv,
                    double *ws);
// C[k x n] <- U^T[k x k] * C  (in place)
void gqz_slab_left(hipStream_t st, double *C, int ldc, int k, int n, const double *U, int ldu,
                   double *ws);
// general out-of-place C = op(A) * op(B) (m x n, inner k); ta/tb transpose flags
void gqz_gemm(hipStream_t st, int ta, int tb, int m, int n, int k, const double *A, int lda,
              const double *B, int ldb, double *C, int ldc);
// C = alpha op(A) op(B) + beta C (beta = 0: C is only written)
void gqz_gemm_ab(hipStream_t st, int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc);

// chain chase inside an LDS window (see k_chase.hip)
struct GqzChainArgs {
    double *H, *T;     // full matrices
    int ld;
    int p;             // window top row/col
    int w;             // window order
    int nb;            // bulges in the chain
    int t0;            // tail column (window-relative) of the top bulge at entry, -1 = intro
    int steps;         // chase steps of every bulge (capped by push-off)
    int intro;         // 1: introduce all bulges at the top of the window first
    int last;          // 1: window bottom is the bottom of the active block (push-off)
    double *U, *V;     // w x w outputs (ld = w)
    const double *sr, *si, *sb;   // 2*nb shifts
    double smlnum;
};
void gqz_chase_window(hipStream_t st, const GqzChainArgs &a);

// several chains chased at once, one workgroup each, in disjoint diagonal windows
#define GQZ_MAXCH 16
struct GqzChainSet {
    double *H, *T;
    int ld, nch;
    double *U, *V;            // chain c: U + uslot[c] * ustride (w_c x w_c, ld w_c); same for V
    int ustride, pad;
    const double *sr, *si, *sb;
    double smlnum;
    int p[GQZ_MAXCH], w[GQZ_MAXCH], nb[GQZ_MAXCH], t0[GQZ_MAXCH], steps[GQZ_MAXCH];
    int intro[GQZ_MAXCH], last[GQZ_MAXCH], soff[GQZ_MAXCH], uslot[GQZ_MAXCH];
    // deferred left update of the chain's previous window (near[c] > 0): before its window is
    // loaded, chain c applies U + pslot[c] * ustride (pw[c] x pw[c]) to rows [pp, pp + pw) of H and
    // T in the near[c] columns right of that window, the columns its window reads
    int near[GQZ_MAXCH], pslot[GQZ_MAXCH], pp[GQZ_MAXCH], pw[GQZ_MAXCH];
};
void gqz_chase_set(hipStream_t st, const GqzChainSet &s);
int gqz_chase_max_window();
