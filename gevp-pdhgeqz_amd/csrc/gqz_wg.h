// gqz_wg.h -- workgroup-cooperative building blocks of the QZ engine.
//
// One lane (threadIdx.x == 0) owns the control flow and computes the small transforms; all
// lanes of the workgroup apply them.  Pointers are generic (flat), so the same code runs on
// LDS-resident windows (small blocks) and on HBM/L2-resident windows (AED post-processing).
#pragma once
#include "gqz_dev.h"

struct WgPencil {
    double *H, *T;     // n x n, leading dimension ld
    double *U, *V;     // nq x n accumulators (may be null), leading dimension ldu
    int ld, ldu;
    int n, nq;
};

// control block shared by the workgroup (lives in LDS)
struct WgCtl {
    double v[4], w[4];
    double tau, tau2;
    double c, s;
    double G[8];
    double S[16], Tm[16], Qa[16], Za[16];
    int op, i0, i1, i2, i3, i4;
    int lo, ilast, ifirst, iiter, its, maxit, info;
    int sp;
    int scan_f, scan_t;      // results of the parallel deflation scans of wg_qz_ds
    int stk_lo[96], stk_hi[96];
    double sr[2], si[2], sb[2];
    double red[64];
};

// ---- parallel applications (no barrier inside; callers synchronise) -----------------------
__device__ __forceinline__ void wg_rows_rot(double *M, int ld, int i1, int i2, int c0, int c1,
                                            double c, double s) {
    for (int j = c0 + (int)threadIdx.x; j < c1; j += blockDim.x) {
        double x = M[GQZ_IX(i1, j, ld)], y = M[GQZ_IX(i2, j, ld)];
        M[GQZ_IX(i1, j, ld)] = c * x + s * y;
        M[GQZ_IX(i2, j, ld)] = c * y - s * x;
    }
}
__device__ __forceinline__ void wg_cols_rot(double *M, int ld, int j1, int j2, int r0, int r1,
                                            double c, double s) {
    double *x = M + GQZ_IX(0, j1, ld), *y = M + GQZ_IX(0, j2, ld);
    for (int i = r0 + (int)threadIdx.x; i < r1; i += blockDim.x) {
        double a = x[i], b = y[i];
        x[i] = c * a + s * b;
        y[i] = c * b - s * a;
    }
}

// rows k..k+2 of two matrices (columns [c0a,c1) of A and [c0b,c1) of B) and columns k..k+2 of
// the accumulator U (rows [0,nq)) in one pass: the left half of a bulge step.
__device__ inline void wg_left3(const WgPencil &P, int k, int c0h, int c0t, int c1,
                                const double *v, double tau) {
    if (tau == 0.0) return;
    int nh = c1 - c0h, nt_ = c1 - c0t, nu = P.U ? P.nq : 0;
    int tot = nh + nt_ + nu;
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        if (idx < nh + nt_) {
            double *M; int j;
            if (idx < nh) { M = P.H; j = c0h + idx; } else { M = P.T; j = c0t + idx - nh; }
            double *p = M + GQZ_IX(k, j, P.ld);
            double t = tau * (p[0] + v[1] * p[1] + v[2] * p[2]);
            p[0] -= t; p[1] -= t * v[1]; p[2] -= t * v[2];
        } else {
            int i = idx - nh - nt_;
            double *x = P.U + GQZ_IX(i, k, P.ldu);
            double a = x[0], b = x[P.ldu], cc = x[2 * (size_t)P.ldu];
            double t = tau * (a + v[1] * b + v[2] * cc);
            x[0] = a - t; x[P.ldu] = b - t * v[1]; x[2 * (size_t)P.ldu] = cc - t * v[2];
        }
    }
}
// columns k..k+2: rows [0,r1h) of H, [0,r1t) of T, [0,nq) of V
__device__ inline void wg_right3(const WgPencil &P, int k, int r0, int r1h, int r1t,
                                 const double *v, double tau) {
    if (tau == 0.0) return;
    int nh = r1h - r0, nt_ = r1t - r0, nv = P.V ? P.nq : 0;
    int tot = nh + nt_ + nv;
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        double *M; int i; int ld;
        if (idx < nh) { M = P.H; i = r0 + idx; ld = P.ld; }
        else if (idx < nh + nt_) { M = P.T; i = r0 + idx - nh; ld = P.ld; }
        else { M = P.V; i = idx - nh - nt_; ld = P.ldu; }
        double *x = M + GQZ_IX(i, k, ld);
        double a = x[0], b = x[ld], cc = x[2 * (size_t)ld];
        double t = tau * (a + v[1] * b + v[2] * cc);
        x[0] = a - t; x[ld] = b - t * v[1]; x[2 * (size_t)ld] = cc - t * v[2];
    }
}
// plane rotation on rows (i, i+1): H cols [c0h,n), T cols [c0t,n), U cols (i,i+1)
__device__ inline void wg_rowrot_all(const WgPencil &P, int i, int c0h, int c0t, double c, double s) {
    int nh = P.n - c0h, nt_ = P.n - c0t, nu = P.U ? P.nq : 0;
    for (int idx = threadIdx.x; idx < nh + nt_ + nu; idx += blockDim.x) {
        if (idx < nh + nt_) {
            double *M; int j;
            if (idx < nh) { M = P.H; j = c0h + idx; } else { M = P.T; j = c0t + idx - nh; }
            double x = M[GQZ_IX(i, j, P.ld)], y = M[GQZ_IX(i + 1, j, P.ld)];
            M[GQZ_IX(i, j, P.ld)] = c * x + s * y;
            M[GQZ_IX(i + 1, j, P.ld)] = c * y - s * x;
        } else {
            int r = idx - nh - nt_;
            double x = P.U[GQZ_IX(r, i, P.ldu)], y = P.U[GQZ_IX(r, i + 1, P.ldu)];
            P.U[GQZ_IX(r, i, P.ldu)] = c * x + s * y;
            P.U[GQZ_IX(r, i + 1, P.ldu)] = c * y - s * x;
        }
    }
}
// plane rotation on columns (jx, jy) (x' = c x + s y, y' = c y - s x): H rows [0,r1h),
// T rows [0,r1t), V all rows
__device__ inline void wg_colrot_all(const WgPencil &P, int jx, int jy, int r1h, int r1t,
                                     double c, double s) {
    int nv = P.V ? P.nq : 0;
    for (int idx = threadIdx.x; idx < r1h + r1t + nv; idx += blockDim.x) {
        double *M; int i, ld;
        if (idx < r1h) { M = P.H; i = idx; ld = P.ld; }
        else if (idx < r1h + r1t) { M = P.T; i = idx - r1h; ld = P.ld; }
        else { M = P.V; i = idx - r1h - r1t; ld = P.ldu; }
        double a = M[GQZ_IX(i, jx, ld)], b = M[GQZ_IX(i, jy, ld)];
        M[GQZ_IX(i, jx, ld)] = c * a + s * b;
        M[GQZ_IX(i, jy, ld)] = c * b - s * a;
    }
}
// general 2x2 orthogonal: rows (j,j+1) by QL^T over cols [c0,n) of H and T, columns by ZR
// over rows [0,r1) of H and T; U <- U QL, V <- V ZR
__device__ inline void wg_g2_all(const WgPencil &P, int j, int c0, int r1, const double *QL,
                                 const double *ZR) {
    int nr = P.n - c0;
    int nu = P.U ? P.nq : 0, nv = P.V ? P.nq : 0;
    int tot = 2 * nr + 2 * r1 + nu + nv;
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        if (idx < 2 * nr) {
            double *M = idx < nr ? P.H : P.T;
            int col = c0 + (idx < nr ? idx : idx - nr);
            double x = M[GQZ_IX(j, col, P.ld)], y = M[GQZ_IX(j + 1, col, P.ld)];
            M[GQZ_IX(j, col, P.ld)] = QL[0] * x + QL[1] * y;
            M[GQZ_IX(j + 1, col, P.ld)] = QL[2] * x + QL[3] * y;
        } else if (idx < 2 * nr + 2 * r1) {
            int t = idx - 2 * nr;
            double *M = t < r1 ? P.H : P.T;
            int row = t < r1 ? t : t - r1;
            double a = M[GQZ_IX(row, j, P.ld)], b = M[GQZ_IX(row, j + 1, P.ld)];
            M[GQZ_IX(row, j, P.ld)] = a * ZR[0] + b * ZR[1];
            M[GQZ_IX(row, j + 1, P.ld)] = a * ZR[2] + b * ZR[3];
        } else {
            int t = idx - 2 * nr - 2 * r1;
            double *M; const double *G;
            if (t < nu) { M = P.U; G = QL; } else { M = P.V; G = ZR; t -= nu; }
            double a = M[GQZ_IX(t, j, P.ldu)], b = M[GQZ_IX(t, j + 1, P.ldu)];
            M[GQZ_IX(t, j, P.ldu)] = a * G[0] + b * G[1];
            M[GQZ_IX(t, j + 1, P.ldu)] = a * G[2] + b * G[3];
        }
    }
}
// general m x m (m <= 4) orthogonal on the block at j1: rows by Qa^T over cols [j1+m, n),
// columns by Za over rows [0, j1); U <- U Qa, V <- V Za (block interior written by caller)
__device__ inline void wg_gm_all(const WgPencil &P, int j1, int m, const double *Qa, const double *Za) {
    int nr = P.n - (j1 + m);
    int nc = j1;
    int nu = P.U ? P.nq : 0, nv = P.V ? P.nq : 0;
    int tot = 2 * nr + 2 * nc + nu + nv;
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        double x[4], y[4];
        if (idx < 2 * nr) {
            double *M = idx < nr ? P.H : P.T;
            int col = j1 + m + (idx < nr ? idx : idx - nr);
            for (int i = 0; i < m; ++i) x[i] = M[GQZ_IX(j1 + i, col, P.ld)];
            for (int i = 0; i < m; ++i) { double s = 0; for (int k = 0; k < m; ++k) s += Qa[k + m * i] * x[k]; y[i] = s; }
            for (int i = 0; i < m; ++i) M[GQZ_IX(j1 + i, col, P.ld)] = y[i];
        } else {
            double *M; const double *G; int row, ld;
            int t = idx - 2 * nr;
            if (t < 2 * nc) { M = t < nc ? P.H : P.T; row = t < nc ? t : t - nc; G = Za; ld = P.ld; }
            else {
                t -= 2 * nc;
                if (t < nu) { M = P.U; G = Qa; } else { M = P.V; G = Za; t -= nu; }
                row = t; ld = P.ldu;
            }
            for (int j = 0; j < m; ++j) x[j] = M[GQZ_IX(row, j1 + j, ld)];
            for (int j = 0; j < m; ++j) { double s = 0; for (int k = 0; k < m; ++k) s += x[k] * G[k + m * j]; y[j] = s; }
            for (int j = 0; j < m; ++j) M[GQZ_IX(row, j1 + j, ld)] = y[j];
        }
    }
}
