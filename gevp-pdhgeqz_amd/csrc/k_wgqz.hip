// k_wgqz.hip -- single-workgroup QZ kernels for gfx950.
//
//  * k_small_qz   : double-shift QZ of a small active block held entirely in LDS (the
//                   KKQZ / QZLAP role for blocks <= 64, PDHGEQZ4.f:14-405 semantics); emits the
//                   block's Schur form, the accumulated U, V and its eigenvalues.
//  * k_gghrd_wg   : Givens Hessenberg-triangular reduction of a pencil by one workgroup
//                   (PDGGHRD semantics, used to restore HT form inside AED, PDHGEQZ3.f:964).
//
// One lane computes each small transform; the workgroup applies it (see gqz_wg.h).
#include "gqz_ds.h"
#include "gqz_kernels.h"
// --------------------------------------------------------------------------------------------
// k_small_qz: one job per workgroup; the block is copied into LDS (ld = m + 1)
// --------------------------------------------------------------------------------------------
__device__ void small_qz_body(const GqzSmallJob &J) {
    extern __shared__ double lds[];
    const int m = J.m, ld = m + 1;
    double *sH = lds, *sT = sH + (size_t)ld * m, *sU = sT + (size_t)ld * m, *sV = sU + (size_t)ld * m;
    WgCtl *C = (WgCtl *)(sV + (size_t)ld * m);
    for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
        int i = idx % m, j = idx / m;
        sH[i + ld * j] = (i <= j + 1) ? J.H[GQZ_IX(i, j, J.ld)] : 0.0;
        sT[i + ld * j] = (i <= j) ? J.T[GQZ_IX(i, j, J.ld)] : 0.0;
        sU[i + ld * j] = (i == j) ? 1.0 : 0.0;
        sV[i + ld * j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    WgPencil P;
    P.H = sH; P.T = sT; P.U = sU; P.V = sV; P.ld = ld; P.ldu = ld; P.n = m; P.nq = m;
    int info = wg_qz_ds(P, C, 0, m - 1, J.ar, J.ai, J.be, J.ttol, J.smlnum);
    for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
        int i = idx % m, j = idx / m;
        J.H[GQZ_IX(i, j, J.ld)] = sH[i + ld * j];
        J.T[GQZ_IX(i, j, J.ld)] = sT[i + ld * j];
        if (J.U) J.U[GQZ_IX(i, j, J.ldu)] = sU[i + ld * j];
        if (J.V) J.V[GQZ_IX(i, j, J.ldu)] = sV[i + ld * j];
    }
    if (threadIdx.x == 0 && J.info) *J.info = info;
}

extern "C" __global__ void __launch_bounds__(256) k_small_qz(const GqzSmallJob *jobs) {
    small_qz_body(jobs[blockIdx.x]);
}
extern "C" __global__ void __launch_bounds__(256) k_small_qz1(GqzSmallJob J) { small_qz_body(J); }

// --------------------------------------------------------------------------------------------
// k_gghrd_wg: Givens HT reduction of A (active rows/cols [ilo, ihi]) with B upper triangular,
// rows of A, B updated over columns up to n, columns over rows [0, ihi]; Q, Z (nq rows).
// Each rotation pair costs two barriers: lane 0 computes the row rotation and, from the
// updated 2x2 diagonal block of B, the column rotation, and applies both to the 2x2 corners;
// the workgroup applies them everywhere else (PDGGHRD.f:365-642 / KRNL* semantics).
// --------------------------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(1024) k_gghrd_wg(double *A, double *B, int ld, int n,
                                                               int ilo, int ihi, double *Q,
                                                               double *Z, int ldq, int nq) {
    __shared__ double sh[4];
    wg_gghrd(A, B, ld, n, ilo, ihi, Q, Z, ldq, nq, sh);
}
