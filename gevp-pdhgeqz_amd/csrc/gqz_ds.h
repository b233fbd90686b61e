// gqz_ds.h -- workgroup-cooperative double-shift QZ of a pencil block (KKQZ / QZLAP role,
// PDHGEQZ4.f:14-405 semantics): deflation decisions by one lane, transforms applied by the
// whole workgroup.  Used on LDS-resident blocks (k_small_qz) and on HBM-resident pencils
// (k_batched).
#pragma once
#include "gqz_wg.h"
enum { OP_NOP = 0, OP_SEGDONE, OP_COLROT, OP_ROWROT, OP_G2, OP_NEG, OP_ZCHASE, OP_SWEEP, OP_FAIL };
enum { POST_NONE = 0, POST_G2 };

// --------------------------------------------------------------------------------------------
// double-shift sweep on [f, l] (intro at f, 3x3 chase, push-off at l): DHGEQZ5.f:128-251,
// PDHGEQZA.f:290-330, PDHGEQZB.f:368-406.
// --------------------------------------------------------------------------------------------
static __device__ void wg_ds_sweep(const WgPencil &P, WgCtl *C, int f, int l, double smlnum) {
    const int tid = threadIdx.x;
    for (int k = f; k <= l - 2; ++k) {
        if (tid == 0) {
            double *H = P.H, *T = P.T;
            int ld = P.ld;
            // all reads first (the writes below touch column k-1 only), then the arithmetic
            double t3[9], h3[9];
            for (int jj = 0; jj < 3; ++jj)
                for (int ii = 0; ii < 3; ++ii) {
                    t3[ii + 3 * jj] = T[GQZ_IX(k + ii, k + jj, ld)];
                    h3[ii + 3 * jj] = H[GQZ_IX(k + ii, k + jj, ld)];
                }
            double v0, v1, v2, tau;
            if (k == f) {
                double v[3];
                d_shift_col(H + GQZ_IX(k, k, ld), ld, T + GQZ_IX(k, k, ld), ld, C->sr[0], C->si[0],
                            C->sb[0], C->sr[1], C->sb[1], v);
                v0 = v[0]; v1 = v[1]; v2 = v[2];
                d_larfg3f(v0, v1, v2, tau);
            } else {
                const double hp = H[GQZ_IX(k - 1, k - 1, ld)];
                const double c0 = H[GQZ_IX(k, k - 1, ld)], c1 = H[GQZ_IX(k + 1, k - 1, ld)];
                const double c2 = H[GQZ_IX(k + 2, k - 1, ld)];
                T[GQZ_IX(k, k - 1, ld)] = 0.0; T[GQZ_IX(k + 1, k - 1, ld)] = 0.0;
                v1 = c1; v2 = c2;
                double beta = c0;
                d_larfg3f(beta, v1, v2, tau);
                double hk = beta;
                if (c2 == 0.0 && h3[2] == 0.0 && h3[5] != 0.0) {
                    // collapsed bulge: try to re-introduce it (DHGEQZ5.f:141-189)
                    double w[3];
                    d_shift_col(H + GQZ_IX(k, k, ld), ld, T + GQZ_IX(k, k, ld), ld, C->sr[0], C->si[0],
                                C->sb[0], C->sr[1], C->sb[1], w);
                    double a1 = w[0], x1 = w[1], x2 = w[2], tau1;
                    d_larfg3f(a1, x1, x2, tau1);
                    double refsum = tau1 * (c0 + x1 * c1);
                    if (!(fabs(c1 - refsum * x1) + fabs(refsum * x2) >
                          GQZ_ULP * (fabs(hp) + fabs(h3[0]) + fabs(h3[4])))) {
                        hk = c0 - refsum;
                        tau = tau1; v1 = x1; v2 = x2;
                    }
                }
                H[GQZ_IX(k, k - 1, ld)] = hk;
                H[GQZ_IX(k + 1, k - 1, ld)] = 0.0; H[GQZ_IX(k + 2, k - 1, ld)] = 0.0;
            }
            // the right (opposite) reflector is computed from the 3x3 block of T after the left
            // reflector, which lane 0 forms privately; lane 0 also updates the 3x3 corners of H and
            // T (left then right), so that all other entries can be updated in a single phase
            if (tau != 0.0) {
                for (int jj = 0; jj < 3; ++jj) {
                    double *tc = t3 + 3 * jj, *hc = h3 + 3 * jj;
                    double a = tau * (tc[0] + v1 * tc[1] + v2 * tc[2]);
                    tc[0] -= a; tc[1] -= a * v1; tc[2] -= a * v2;
                    double b = tau * (hc[0] + v1 * hc[1] + v2 * hc[2]);
                    hc[0] -= b; hc[1] -= b * v1; hc[2] -= b * v2;
                }
            }
            double w[3], tau2;
            d_nullrefl3(t3[1], t3[4], t3[7], t3[2], t3[5], t3[8], w, tau2);
            if (tau2 != 0.0) {
                for (int ii = 0; ii < 3; ++ii) {
                    double a = tau2 * (t3[ii] + w[1] * t3[ii + 3] + w[2] * t3[ii + 6]);
                    t3[ii] -= a; t3[ii + 3] -= a * w[1]; t3[ii + 6] -= a * w[2];
                    double b = tau2 * (h3[ii] + w[1] * h3[ii + 3] + w[2] * h3[ii + 6]);
                    h3[ii] -= b; h3[ii + 3] -= b * w[1]; h3[ii + 6] -= b * w[2];
                }
            }
            for (int jj = 0; jj < 3; ++jj)
                for (int ii = 0; ii < 3; ++ii) {
                    T[GQZ_IX(k + ii, k + jj, ld)] = t3[ii + 3 * jj];
                    H[GQZ_IX(k + ii, k + jj, ld)] = h3[ii + 3 * jj];
                }
            C->v[0] = 1.0; C->v[1] = v1; C->v[2] = v2; C->tau = tau;
            C->w[0] = 1.0; C->w[1] = w[1]; C->w[2] = w[2]; C->tau2 = tau2;
        }
        __syncthreads();
        {
            // one phase: left reflector on rows k..k+2 (columns right of the corner, and U), right
            // reflector on columns k..k+2 (rows above the corner, H's fill row k+3, and V); the
            // two sets are disjoint outside the corner
            const double tau = C->tau, v1 = C->v[1], v2 = C->v[2];
            const double tau2 = C->tau2, w1 = C->w[1], w2 = C->w[2];
            const int ld = P.ld, ldu = P.ldu;
            const int nlh = P.n - (k + 3), nlt = nlh, nu = P.U ? P.nq : 0;
            const int hbot = min(k + 4, l + 1);           // rows [0, hbot) of H see the right reflector
            const int nrh = k + (hbot > k + 3 ? 1 : 0), nrt = k, nv = P.V ? P.nq : 0;
            const int tot = nlh + nlt + nu + nrh + nrt + nv;
            for (int idx = tid; idx < tot; idx += blockDim.x) {
                int t = idx;
                if (t < nlh + nlt) {
                    if (tau == 0.0) continue;
                    double *M = t < nlh ? P.H : P.T;
                    const int j = k + 3 + (t < nlh ? t : t - nlh);
                    double *p = M + GQZ_IX(k, j, ld);
                    double a = tau * (p[0] + v1 * p[1] + v2 * p[2]);
                    p[0] -= a; p[1] -= a * v1; p[2] -= a * v2;
                    continue;
                }
                t -= nlh + nlt;
                if (t < nu) {
                    if (tau == 0.0) continue;
                    double *x = P.U + GQZ_IX(t, k, ldu);
                    double a0 = x[0], b0 = x[ldu], c0 = x[2 * (size_t)ldu];
                    double a = tau * (a0 + v1 * b0 + v2 * c0);
                    x[0] = a0 - a; x[ldu] = b0 - a * v1; x[2 * (size_t)ldu] = c0 - a * v2;
                    continue;
                }
                t -= nu;
                if (tau2 == 0.0) continue;
                double *x; int ldx;
                if (t < nrh) { x = P.H + GQZ_IX(t < k ? t : k + 3, k, ld); ldx = ld; }
                else if (t < nrh + nrt) { x = P.T + GQZ_IX(t - nrh, k, ld); ldx = ld; }
                else { x = P.V + GQZ_IX(t - nrh - nrt, k, ldu); ldx = ldu; }
                double a0 = x[0], b0 = x[ldx], c0 = x[2 * (size_t)ldx];
                double a = tau2 * (a0 + w1 * b0 + w2 * c0);
                x[0] = a0 - a; x[ldx] = b0 - a * w1; x[2 * (size_t)ldx] = c0 - a * w2;
            }
        }
        __syncthreads();
    }
    // push-off at the bottom
    if (tid == 0) {
        double *H = P.H, *T = P.T;
        int ld = P.ld;
        T[GQZ_IX(l - 1, l - 2, ld)] = 0.0; T[GQZ_IX(l, l - 2, ld)] = 0.0;
        double c, s, r;
        d_lartgf(H[GQZ_IX(l - 1, l - 2, ld)], H[GQZ_IX(l, l - 2, ld)], c, s, r);
        H[GQZ_IX(l - 1, l - 2, ld)] = r; H[GQZ_IX(l, l - 2, ld)] = 0.0;
        C->c = c; C->s = s;
    }
    __syncthreads();
    wg_rowrot_all(P, l - 1, l - 1, l - 1, C->c, C->s);
    __syncthreads();
    if (tid == 0) {
        double *T = P.T;
        int ld = P.ld;
        double c, s, r;
        d_lartgf(T[GQZ_IX(l, l, ld)], T[GQZ_IX(l, l - 1, ld)], c, s, r);
        T[GQZ_IX(l, l, ld)] = r; T[GQZ_IX(l, l - 1, ld)] = 0.0;
        C->c = c; C->s = s;
    }
    __syncthreads();
    wg_colrot_all(P, l, l - 1, l + 1, l, C->c, C->s);
    __syncthreads();
    {
        // vigilant deflation (DHGEQZ5.f:254-275); the test of H(j, j-1) reads no entry that
        // another test writes, so the workgroup runs them all at once
        double *H = P.H;
        int ld = P.ld;
        for (int j = f + 1 + tid; j <= l; j += blockDim.x) {
            double hj = H[GQZ_IX(j, j - 1, ld)];
            if (hj == 0.0) continue;
            double a = H[GQZ_IX(j - 1, j - 1, ld)], b = H[GQZ_IX(j, j, ld)];
            if (fabs(hj) <= fmax(smlnum, GQZ_ULP * (fabs(a) + fabs(b)))) {
                double up = H[GQZ_IX(j - 1, j, ld)];
                double h12 = fmax(fabs(hj), fabs(up)), h21 = fmin(fabs(hj), fabs(up));
                double h11 = fmax(fabs(b), fabs(a - b)), h22 = fmin(fabs(b), fabs(a - b));
                double scl = h11 + h12, tst2 = h22 * (h11 / scl);
                if (tst2 == 0.0 || h21 * (h12 / scl) <= fmax(smlnum, GQZ_ULP * tst2))
                    H[GQZ_IX(j, j - 1, ld)] = 0.0;
            }
        }
    }
    __syncthreads();
}

// lane-0 decision step of the base QZ; fills C->op and its parameters.
// parallel part of the deflation decision: the largest j in (lo, ilast] with a negligible H(j, j-1)
// (PDLASMSUB.f:368; lo when there is none) and, in [that j, ilast], the largest i with a
// negligible T(i, i) (-1 when none); negligible T(i, i) are set to zero
static __device__ void ds_scan(const WgPencil &P, WgCtl *C, double ttol, double smlnum) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const double *H = P.H;
    double *T = P.T;
    const int ld = P.ld;
    const int lo = C->lo, il = C->ilast;
    if (tid == 0) { C->scan_f = lo; C->scan_t = -1; }
    __syncthreads();
    for (int j = lo + 1 + tid; j <= il; j += nt) {
        double tst = fabs(H[GQZ_IX(j - 1, j - 1, ld)]) + fabs(H[GQZ_IX(j, j, ld)]);
        if (fabs(H[GQZ_IX(j, j - 1, ld)]) <= fmax(smlnum, GQZ_ULP * tst)) atomicMax(&C->scan_f, j);
    }
    __syncthreads();
    const double tol = fmax(ttol, smlnum);
    for (int i = C->scan_f + tid; i <= il; i += nt)
        if (fabs(T[GQZ_IX(i, i, ld)]) <= tol) { T[GQZ_IX(i, i, ld)] = 0.0; atomicMax(&C->scan_t, i); }
    __syncthreads();
}

static __device__ void ds_decide(const WgPencil &P, WgCtl *C, double *ar, double *ai, double *be,
                          double ttol, double smlnum) {
    double *H = P.H, *T = P.T;
    int ld = P.ld;
    C->op = OP_NOP;
    if (C->ilast < C->lo) { C->op = OP_SEGDONE; return; }
    if (++C->its > C->maxit) { C->op = OP_FAIL; C->info = C->ilast + 1; return; }
    int lo = C->lo, il = C->ilast;
    // the bottom-up search for a negligible subdiagonal entry and the scan of diag(T)
    // (PDHGEQZ7.f:380) were done by the whole workgroup (ds_scan)
    const int f = C->scan_f;
    if (f > lo) H[GQZ_IX(f, f - 1, ld)] = 0.0;
    C->ifirst = f;
    const int found = C->scan_t;
    if (found >= 0) {
        C->iiter = 0;
        if (il == f) {
            ar[il] = H[GQZ_IX(il, il, ld)]; ai[il] = 0.0; be[il] = 0.0;
            C->ilast = il - 1; return;
        }
        double c, s, r;
        if (T[GQZ_IX(il, il, ld)] == 0.0) {
            d_lartgf(H[GQZ_IX(il, il, ld)], H[GQZ_IX(il, il - 1, ld)], c, s, r);
            H[GQZ_IX(il, il, ld)] = r; H[GQZ_IX(il, il - 1, ld)] = 0.0;
            ar[il] = r; ai[il] = 0.0; be[il] = 0.0;
            C->op = OP_COLROT; C->i0 = il; C->i1 = il - 1; C->i2 = il; C->i3 = il; C->c = c; C->s = s;
            C->ilast = il - 1; return;
        }
        // top deflation splits the segment; with the split stack full the zero is chased down instead
        if (T[GQZ_IX(f, f, ld)] == 0.0 && (f == lo || C->sp < 96)) {
            d_lartgf(H[GQZ_IX(f, f, ld)], H[GQZ_IX(f + 1, f, ld)], c, s, r);
            H[GQZ_IX(f, f, ld)] = r; H[GQZ_IX(f + 1, f, ld)] = 0.0;
            ar[f] = r; ai[f] = 0.0; be[f] = 0.0;
            C->op = OP_ROWROT; C->i0 = f; C->i1 = f + 1; C->i2 = f + 1; C->c = c; C->s = s;
            if (f > lo) { C->stk_lo[C->sp] = lo; C->stk_hi[C->sp] = f - 1; C->sp++; }
            C->lo = f + 1; return;
        }
        C->op = OP_ZCHASE; C->i0 = found; C->i1 = f; C->i2 = il; return;
    }
    if (f == il) {
        if (T[GQZ_IX(il, il, ld)] < 0.0) { C->op = OP_NEG; C->i0 = il; }
        ar[il] = H[GQZ_IX(il, il, ld)]; ai[il] = 0.0; be[il] = T[GQZ_IX(il, il, ld)];
        if (C->op == OP_NEG) { ar[il] = -ar[il]; be[il] = -be[il]; }
        C->ilast = il - 1; C->iiter = 0; return;
    }
    if (f == il - 1) {
        double A[4] = { H[GQZ_IX(f, f, ld)], H[GQZ_IX(f + 1, f, ld)], H[GQZ_IX(f, f + 1, ld)], H[GQZ_IX(f + 1, f + 1, ld)] };
        double B[4] = { T[GQZ_IX(f, f, ld)], 0.0, T[GQZ_IX(f, f + 1, ld)], T[GQZ_IX(f + 1, f + 1, ld)] };
        int cplx = d_lagv2(A, B, C->G, C->G + 4);
        for (int q = 0; q < 4; ++q) { C->S[q] = A[q]; C->Tm[q] = B[q]; }
        C->op = OP_G2; C->i0 = f; C->i1 = cplx;
        if (cplx) {
            double a2[2], i2[2], b2[2];
            d_pair_eigs(A, B, a2, i2, b2);
            ar[f] = a2[0]; ai[f] = i2[0]; be[f] = b2[0];
            ar[f + 1] = a2[1]; ai[f + 1] = i2[1]; be[f + 1] = b2[1];
            C->ilast = il - 2;
        }
        C->iiter = 0;
        return;
    }
    // double-shift sweep on [f, il]
    C->iiter++;
    double er[2], ei[2];
    d_eig22(H[GQZ_IX(il - 1, il - 1, ld)], H[GQZ_IX(il, il - 1, ld)], H[GQZ_IX(il - 1, il, ld)],
            H[GQZ_IX(il, il, ld)], T[GQZ_IX(il - 1, il - 1, ld)], T[GQZ_IX(il - 1, il, ld)],
            T[GQZ_IX(il, il, ld)], er, ei);
    if (C->iiter % 10 == 0) {
        double e = fabs(H[GQZ_IX(il, il - 1, ld)] / T[GQZ_IX(il - 1, il - 1, ld)]) +
                   fabs(H[GQZ_IX(il - 1, il - 2, ld)] / T[GQZ_IX(il - 2, il - 2, ld)]);
        double lam = H[GQZ_IX(il, il, ld)] / T[GQZ_IX(il, il, ld)] + 0.75 * e;
        er[0] = er[1] = lam; ei[0] = ei[1] = 0.0;
        if (C->iiter % 20 == 0) er[1] = -0.4375 * e + H[GQZ_IX(il, il, ld)] / T[GQZ_IX(il, il, ld)];
    }
    if (!isfinite(er[0]) || !isfinite(er[1]) || !isfinite(ei[0])) { er[0] = er[1] = 1.0; ei[0] = ei[1] = 0.0; }
    C->sr[0] = er[0]; C->sr[1] = er[1]; C->si[0] = ei[0]; C->si[1] = ei[1]; C->sb[0] = C->sb[1] = 1.0;
    C->op = OP_SWEEP; C->i0 = f; C->i1 = il;
}

// interior zero of diag(T) chased to the bottom (DHGEQZ7 / PDHGEQZ8 direction)
static __device__ void wg_zchase(const WgPencil &P, WgCtl *C, int j, int kt, int kb) {
    const int tid = threadIdx.x;
    for (int jch = j; jch < kb; ++jch) {
        if (tid == 0) {
            double c, s, r;
            d_lartgf(P.T[GQZ_IX(jch, jch + 1, P.ld)], P.T[GQZ_IX(jch + 1, jch + 1, P.ld)], c, s, r);
            P.T[GQZ_IX(jch, jch + 1, P.ld)] = r; P.T[GQZ_IX(jch + 1, jch + 1, P.ld)] = 0.0;
            C->c = c; C->s = s;
        }
        __syncthreads();
        wg_rowrot_all(P, jch, jch > kt ? jch - 1 : jch, jch + 2, C->c, C->s);
        __syncthreads();
        if (jch > kt) {
            if (tid == 0) {
                double c, s, r;
                d_lartgf(P.H[GQZ_IX(jch + 1, jch, P.ld)], P.H[GQZ_IX(jch + 1, jch - 1, P.ld)], c, s, r);
                P.H[GQZ_IX(jch + 1, jch, P.ld)] = r; P.H[GQZ_IX(jch + 1, jch - 1, P.ld)] = 0.0;
                C->c = c; C->s = s;
            }
            __syncthreads();
            wg_colrot_all(P, jch, jch - 1, jch + 1, jch, C->c, C->s);
            __syncthreads();
        }
    }
}

// Base QZ of the pencil's block [ilo, ihi] with full-row/column updates inside the pencil and
// U, V accumulation.  Returns INFO (0 = converged).
static __device__ int wg_qz_ds(const WgPencil &P, WgCtl *C, int ilo, int ihi, double *ar, double *ai,
                        double *be, double ttol, double smlnum) {
    const int tid = threadIdx.x;
    if (tid == 0) {
        C->sp = 0; C->info = 0; C->its = 0; C->maxit = 40 * (ihi - ilo + 1) + 20;
        C->stk_lo[0] = ilo; C->stk_hi[0] = ihi; C->sp = 1;
        C->op = OP_NOP;
    }
    __syncthreads();
    for (;;) {
        if (tid == 0) {
            if (C->sp == 0 || C->info) C->op = OP_SEGDONE;
            else { C->sp--; C->lo = C->stk_lo[C->sp]; C->ilast = C->stk_hi[C->sp]; C->iiter = 0; C->op = OP_NOP; }
        }
        __syncthreads();
        if (C->op == OP_SEGDONE) break;
        for (;;) {
            ds_scan(P, C, ttol, smlnum);
            if (tid == 0) ds_decide(P, C, ar, ai, be, ttol, smlnum);
            __syncthreads();
            int op = C->op;
            if (op == OP_SEGDONE || op == OP_FAIL) break;
            switch (op) {
            case OP_COLROT: wg_colrot_all(P, C->i0, C->i1, C->i2, C->i3, C->c, C->s); break;
            case OP_ROWROT: wg_rowrot_all(P, C->i0, C->i1, C->i2, C->c, C->s); break;
            case OP_NEG: {
                int j = C->i0;
                for (int idx = tid; idx < 2 * (j + 1) + (P.V ? P.nq : 0); idx += blockDim.x) {
                    if (idx <= j) P.H[GQZ_IX(idx, j, P.ld)] = -P.H[GQZ_IX(idx, j, P.ld)];
                    else if (idx < 2 * (j + 1)) P.T[GQZ_IX(idx - j - 1, j, P.ld)] = -P.T[GQZ_IX(idx - j - 1, j, P.ld)];
                    else P.V[GQZ_IX(idx - 2 * (j + 1), j, P.ldu)] = -P.V[GQZ_IX(idx - 2 * (j + 1), j, P.ldu)];
                }
                break;
            }
            case OP_G2: {
                wg_g2_all(P, C->i0, C->i0 + 2, C->i0, C->G, C->G + 4);
                __syncthreads();
                if (tid == 0) {
                    int f = C->i0;
                    P.H[GQZ_IX(f, f, P.ld)] = C->S[0]; P.H[GQZ_IX(f + 1, f, P.ld)] = C->i1 ? C->S[1] : 0.0;
                    P.H[GQZ_IX(f, f + 1, P.ld)] = C->S[2]; P.H[GQZ_IX(f + 1, f + 1, P.ld)] = C->S[3];
                    P.T[GQZ_IX(f, f, P.ld)] = C->Tm[0]; P.T[GQZ_IX(f + 1, f, P.ld)] = 0.0;
                    P.T[GQZ_IX(f, f + 1, P.ld)] = C->Tm[2]; P.T[GQZ_IX(f + 1, f + 1, P.ld)] = C->Tm[3];
                }
                break;
            }
            case OP_ZCHASE: wg_zchase(P, C, C->i0, C->i1, C->i2); break;
            case OP_SWEEP: wg_ds_sweep(P, C, C->i0, C->i1, smlnum); break;
            default: break;
            }
            __syncthreads();
        }
        if (C->op == OP_FAIL) break;
    }
    __syncthreads();
    return C->info;
}


// --------------------------------------------------------------------------------------------
// wg_gghrd: Givens HT reduction of A (active rows/cols [ilo, ihi]) with B upper triangular,
// rows of A, B updated over columns up to n, columns over rows [0, ihi]; Q, Z (nq rows).
// Each rotation pair costs two barriers: lane 0 computes the row rotation and, from the
// updated 2x2 diagonal block of B, the column rotation, and applies both to the 2x2 corners;
// the workgroup applies them everywhere else (PDGGHRD.f:365-642 / KRNL* semantics).
// --------------------------------------------------------------------------------------------
static __device__ void wg_gghrd(double *A, double *B, int ld, int n, int ilo, int ihi, double *Q, double *Z,
                                int ldq, int nq, double *sh) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int j = ilo; j <= ihi - 2; ++j) {
        for (int i = ihi; i >= j + 2; --i) {
            if (tid == 0) {
                // all reads first (none of them is written before it is read), then the arithmetic
                const double f = A[GQZ_IX(i - 1, j, ld)], g = A[GQZ_IX(i, j, ld)];
                const double a00 = A[GQZ_IX(i - 1, i - 1, ld)], a01 = A[GQZ_IX(i - 1, i, ld)];
                const double a10 = A[GQZ_IX(i, i - 1, ld)], a11 = A[GQZ_IX(i, i, ld)];
                const double b00 = B[GQZ_IX(i - 1, i - 1, ld)], b01 = B[GQZ_IX(i - 1, i, ld)];
                const double b11 = B[GQZ_IX(i, i, ld)];
                double c, s, r, cc, ss;
                d_lartgf(f, g, c, s, r);
                A[GQZ_IX(i - 1, j, ld)] = r; A[GQZ_IX(i, j, ld)] = 0.0;
                // 2x2 corners (rows i-1,i x cols i-1,i) of A and B: row rotation first
                double na00 = c * a00 + s * a10, na10 = c * a10 - s * a00;
                double na01 = c * a01 + s * a11, na11 = c * a11 - s * a01;
                double nb00 = c * b00, nb10 = -s * b00;
                double nb01 = c * b01 + s * b11, nb11 = c * b11 - s * b01;
                // column rotation (i, i-1) zeroing B(i, i-1)
                d_lartgf(nb11, nb10, cc, ss, r);
                // x = col i, y = col i-1: x' = cc x + ss y ; y' = cc y - ss x
                double xa0 = na01, ya0 = na00, xa1 = na11, ya1 = na10;
                A[GQZ_IX(i - 1, i, ld)] = cc * xa0 + ss * ya0; A[GQZ_IX(i - 1, i - 1, ld)] = cc * ya0 - ss * xa0;
                A[GQZ_IX(i, i, ld)] = cc * xa1 + ss * ya1; A[GQZ_IX(i, i - 1, ld)] = cc * ya1 - ss * xa1;
                B[GQZ_IX(i - 1, i, ld)] = cc * nb01 + ss * nb00; B[GQZ_IX(i - 1, i - 1, ld)] = cc * nb00 - ss * nb01;
                B[GQZ_IX(i, i, ld)] = r; B[GQZ_IX(i, i - 1, ld)] = 0.0;
                sh[0] = c; sh[1] = s; sh[2] = cc; sh[3] = ss;
            }
            __syncthreads();
            const double c = sh[0], s = sh[1], cc = sh[2], ss = sh[3];
            // row rotation: A cols [j+1, n) \ {i-1, i}; B cols [i+1, n); Q cols (i-1, i)
            // column rotation: A rows [0, ihi] \ {i-1, i}; B rows [0, i-1); Z cols (i, i-1)
            // Every item is a pair of entries (x, y) -> (c x + s y, c y - s x); a thread gathers
            // four items and issues all their loads before its first store.
            const int na = n - (j + 1), nb = n - (i + 1), nc = ihi + 1, nd = i - 1;
            const int nqq = Q ? nq : 0, nzz = Z ? nq : 0;
            const int tot = na + nb + nqq + nc + nd + nzz;
            for (int base = tid; base < tot; base += 4 * nt) {
                double *px[4], *py[4], cq[4], sq[4];
                bool ok[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    int t = base + q * nt;
                    ok[q] = t < tot; px[q] = py[q] = A; cq[q] = c; sq[q] = s;
                    if (!ok[q]) continue;
                    if (t < na) {
                        const int col = j + 1 + t;
                        ok[q] = (col != i - 1 && col != i);
                        px[q] = A + GQZ_IX(i - 1, col, ld); py[q] = A + GQZ_IX(i, col, ld);
                        continue;
                    }
                    t -= na;
                    if (t < nb) {
                        const int col = i + 1 + t;
                        px[q] = B + GQZ_IX(i - 1, col, ld); py[q] = B + GQZ_IX(i, col, ld);
                        continue;
                    }
                    t -= nb;
                    if (t < nqq) { px[q] = Q + GQZ_IX(t, i - 1, ldq); py[q] = Q + GQZ_IX(t, i, ldq); continue; }
                    t -= nqq;
                    cq[q] = cc; sq[q] = ss;
                    if (t < nc) {
                        ok[q] = (t != i - 1 && t != i);
                        px[q] = A + GQZ_IX(t, i, ld); py[q] = A + GQZ_IX(t, i - 1, ld);
                        continue;
                    }
                    t -= nc;
                    if (t < nd) { px[q] = B + GQZ_IX(t, i, ld); py[q] = B + GQZ_IX(t, i - 1, ld); continue; }
                    t -= nd;
                    px[q] = Z + GQZ_IX(t, i, ldq); py[q] = Z + GQZ_IX(t, i - 1, ldq);
                }
                double x[4], y[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    x[q] = ok[q] ? *px[q] : 0.0;
                    y[q] = ok[q] ? *py[q] : 0.0;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (ok[q]) {
                        *px[q] = cq[q] * x[q] + sq[q] * y[q];
                        *py[q] = cq[q] * y[q] - sq[q] * x[q];
                    }
            }
            __syncthreads();
        }
    }
}
