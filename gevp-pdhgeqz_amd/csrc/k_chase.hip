// k_chase.hip -- multishift sweep: a chain of tightly packed 3x3 bulges chased inside an
// LDS-resident window by one workgroup (PDHGEQZ5/A/B + DHGEQZ5 semantics).
//
// Window: rows/cols [p, p+w) of H and T, w <= 64, held in LDS together with its accumulated
// orthogonal factors U, V (w x w).  Bulge b (b = 0 top) sits at "tail column" c_b (window
// relative): its next left reflector acts on rows c+1..c+3 and is computed from H(c+1:c+3, c)
// (c = -1 means introduction from the shift polynomial).  Bulges are 3 columns apart.
//
// Every chase step advances ALL active bulges by one column:
//   one lane per bulge computes its left reflector and, from the 3x3 block of T that this
//   reflector produces (formed privately), its opposite (right) reflector;
//   the workgroup applies all left reflectors (disjoint row sets), then all right reflectors
//   (disjoint column sets), one wave per (bulge, matrix) segment.
// Left and right multiplications commute, and no reflector reads an entry that a neighbouring
// bulge's same-step reflector writes before it in the reference order (bottom bulge first,
// PDHGEQZB.f:357-407; introduction order of PDHGEQZA.f:290-330), so the result equals the
// reference's sequential chase up to rounding.
//
// The parts of the pencil outside the window are updated afterwards by the MFMA slab kernels
// with U (rows to the right) and V (columns above, Q, Z).  When the host defers the left update of
// the columns that a chain's next window reads (near[c] > 0), the chain's next launch applies it
// while it loads its window (FP64 MFMA from the previous U staged in LDS), so that no separate
// launch sits between two chases of a chain.
#include "gqz_dev.h"
#include "gqz_kernels.h"
#include "gqz_misc.h"

#define CH_MAXW 64
#define CH_MAXB 21
#define CH_NT 1024         // 16 waves: the LDS round trips of the update phases overlap
#define CH_LDX 66          // column stride of the staged near block (2-way bank conflicts at most)
#define CH_NEAR_LDS (sizeof(double) * (4096 + 2 * 64 * CH_LDX))

typedef double dbl4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(CH_NT) k_chase(GqzChainSet S) {
    extern __shared__ double lds[];
    const int cidx = blockIdx.x;
    GqzChainArgs a;
    a.H = S.H; a.T = S.T; a.ld = S.ld; a.p = S.p[cidx]; a.w = S.w[cidx]; a.nb = S.nb[cidx];
    a.t0 = S.t0[cidx]; a.steps = S.steps[cidx]; a.intro = S.intro[cidx]; a.last = S.last[cidx];
    a.U = S.U + (size_t)S.uslot[cidx] * S.ustride; a.V = S.V + (size_t)S.uslot[cidx] * S.ustride;
    a.sr = S.sr + S.soff[cidx]; a.si = S.si + S.soff[cidx]; a.sb = S.sb + S.soff[cidx];
    a.smlnum = S.smlnum;
    const int w = a.w, ld = w + 1, nb = a.nb;
    double *sH = lds, *sT = sH + (size_t)ld * w, *sU = sT + (size_t)ld * w, *sV = sU + (size_t)ld * w;
    double *bv = sV + (size_t)ld * w;          // per bulge: v1 v2 tau w1 w2 tau2 c s  (8)
    int *bc = (int *)(bv + 8 * CH_MAXB);       // per bulge: current tail column
    int *bk = bc + CH_MAXB;                    // per bulge: kind of this step (0 idle, 1 refl, 2 pushoff, 3 intro)
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const double *gH = a.H + GQZ_IX(a.p, a.p, a.ld);
    const double *gT = a.T + GQZ_IX(a.p, a.p, a.ld);
    // Deferred left update of the previous window [pp, pp + pw) (near columns, nc > 0): X <- U^T X for
    // X = rows [pp, pp + pw) x columns [pp + pw, pp + pw + nc) of H and T, with this chain's previous U
    // (pw x pw, ld pw).  The rows from p on (the first pw - d rows of this window, d = p - pp) are this
    // window's block [0, pw - d) x [pw - d, w) and go to LDS; the rows above go back to HBM.  U, X and
    // the window are loaded together (one memory latency), X by whole columns; U is staged in LDS (sF,
    // 64 x 64) in the operand order of FP64 MFMA 16x16x4: entry (r, c) at
    // ((c / 16) * 16 + r / 4) * 64 + (r % 4) * 16 + c % 16, zero beyond pw; X in LDS by columns (sX,
    // leading dimension CH_LDX, H then T).  Work unit (matrix, 16-column tile ct, 16-row output block
    // it), at most two per wave: lane (li, lk) = (l % 16, l / 16) takes X(4 s + lk, 16 ct + li),
    // s = 0..15, and receives output rows 16 it + li of columns 16 ct + lk + 4 r.  Every wave has read
    // its operands before the barrier after which the window overwrites sF and sX.
    const int nc = S.near[cidx], pp = S.pp[cidx], pw = S.pw[cidx], d = a.p - pp, nb0 = pw - d;
    const int nct = (nc + 15) >> 4, nunit = 8 * nct;      // nc <= 64: at most 32 units
    double *XH = a.H + GQZ_IX(pp, pp + pw, a.ld), *XT = a.T + GQZ_IX(pp, pp + pw, a.ld);
    const int li = lane & 15, lk = lane >> 4;
    double uv[4096 / CH_NT], xv[2 * 64 * 64 / CH_NT];
    dbl4 acc[2];
    if (nc > 0) {
        const double *Up = S.U + (size_t)S.pslot[cidx] * S.ustride;
#pragma unroll
        for (int q = 0; q < 4096 / CH_NT; ++q) {
            const int idx = tid + CH_NT * q, r = idx & 63, c = idx >> 6;
            uv[q] = (r < pw && c < pw) ? Up[r + pw * c] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 2 * 64 * 64 / CH_NT; ++q) {
            const int idx = tid + CH_NT * q, r = idx & 63, col = (idx >> 6) & 63, m = idx >> 12;
            xv[q] = (r < pw && col < nc) ? (m == 0 ? XH : XT)[r + (long)col * a.ld] : 0.0;
        }
    }
    // window: every load of a thread is issued before its first LDS write
    double hv[CH_MAXW * CH_MAXW / CH_NT], tv[CH_MAXW * CH_MAXW / CH_NT];
#pragma unroll
    for (int q = 0; q < CH_MAXW * CH_MAXW / CH_NT; ++q) {
        const int idx = tid + CH_NT * q, i = idx % w, j = idx / w;
        const bool ok = idx < w * w;
        hv[q] = ok ? gH[GQZ_IX(i, j, a.ld)] : 0.0;
        tv[q] = ok ? gT[GQZ_IX(i, j, a.ld)] : 0.0;
    }
    if (nc > 0) {
        double *sF = lds, *sX = lds + 4096;
#pragma unroll
        for (int q = 0; q < 4096 / CH_NT; ++q) {
            const int idx = tid + CH_NT * q, r = idx & 63, c = idx >> 6;
            sF[(((c >> 4) * 16 + (r >> 2)) << 6) + (r & 3) * 16 + (c & 15)] = uv[q];
        }
#pragma unroll
        for (int q = 0; q < 2 * 64 * 64 / CH_NT; ++q) {
            const int idx = tid + CH_NT * q, r = idx & 63, cm = idx >> 6;
            sX[r + CH_LDX * cm] = xv[q];
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int u = wave + (CH_NT / 64) * h;
            acc[h] = dbl4{ 0.0, 0.0, 0.0, 0.0 };
            if (u >= nunit) continue;
            const int m = u / (4 * nct), rem = u - 4 * nct * m, ct = rem >> 2, it = rem & 3;
            const double *xc = sX + CH_LDX * (64 * m + 16 * ct + li) + lk;
#pragma unroll
            for (int s2 = 0; s2 < 16; ++s2)
                acc[h] = __builtin_amdgcn_mfma_f64_16x16x4f64(xc[4 * s2], sF[((it * 16 + s2) << 6) + lane], acc[h], 0, 0, 0);
        }
        __syncthreads();               // sF and sX are read: the window may overwrite them
    }
#pragma unroll
    for (int q = 0; q < CH_MAXW * CH_MAXW / CH_NT; ++q) {
        const int idx = tid + CH_NT * q, i = idx % w, j = idx / w;
        if (idx < w * w) {
            if (!(nc > 0 && i < nb0 && j >= nb0)) {      // the near block comes from the update
                sH[i + ld * j] = hv[q];
                sT[i + ld * j] = tv[q];
            }
            sU[i + ld * j] = (i == j) ? 1.0 : 0.0;
            sV[i + ld * j] = (i == j) ? 1.0 : 0.0;
        }
    }
    if (nc > 0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int u = wave + (CH_NT / 64) * h;
            if (u >= nunit) continue;
            const int m = u / (4 * nct), rem = u - 4 * nct * m, ct = rem >> 2, it = rem & 3;
            const int i = 16 * it + li;
            if (i >= pw) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int cj = 16 * ct + lk + 4 * r;
                if (cj >= nc) continue;
                if (i >= d) (m == 0 ? sH : sT)[(i - d) + ld * (nb0 + cj)] = acc[h][r];
                else (m == 0 ? XH : XT)[i + (long)cj * a.ld] = acc[h][r];
            }
        }
    }
    __syncthreads();
    const int total = a.intro ? 3 * nb - 2 + a.steps : a.steps;
    for (int s = 0; s < total; ++s) {
        // ---------------- phase L: left reflectors ----------------
        if (tid < nb) {
            const int b = tid;
            const int start = a.intro ? 3 * (nb - 1 - b) : 0;
            const int nst = a.intro ? 3 * b + 1 + a.steps : a.steps;
            const int cst = a.intro ? -1 : a.t0 + 3 * b;
            const int q = s - start;
            int kind = 0, c = cst + q;
            if (q >= 0 && q < nst) {
                if (c + 3 <= w - 1) kind = (c < 0) ? 3 : 1;
                else if (a.last && c + 2 == w - 1) kind = 2;
                else kind = 0;
            }
            // shifts of this bulge (pair index nb-1-b)
            const int sp = 2 * (nb - 1 - b);
            double *B = bv + 8 * b;
            if (kind == 1 || kind == 3) {
                const int k = c + 1;
                // all reads first (the writes below touch column c of H and column k-1 of T only),
                // then the chain: left reflector, the 3x3 block of T after it (the other bulges'
                // left reflectors act on other rows), right reflector
                double t3[9];
#pragma unroll
                for (int jj = 0; jj < 3; ++jj)
#pragma unroll
                    for (int ii = 0; ii < 3; ++ii) t3[ii + 3 * jj] = sT[(k + ii) + ld * (k + jj)];
                double h0 = 0.0, h1 = 0.0, h2 = 0.0, hk2 = 0.0, hk21 = 0.0;
                if (kind == 1) {
                    h0 = sH[k + ld * c]; h1 = sH[(k + 1) + ld * c]; h2 = sH[(k + 2) + ld * c];
                    hk2 = sH[(k + 2) + ld * k]; hk21 = sH[(k + 2) + ld * (k + 1)];
                }
                double x1, x2, tau;
                if (kind == 3) {
                    double v[3];
                    d_shift_col(sH, ld, sT, ld, a.sr[sp], a.si[sp], a.sb[sp], a.sr[sp + 1], a.sb[sp + 1], v);
                    double al = v[0];
                    x1 = v[1]; x2 = v[2];
                    d_larfg3f(al, x1, x2, tau);
                } else {
                    // zero the fill left in T by this bulge's previous right reflector
                    if (q > 0) { sT[k + ld * (k - 1)] = 0.0; sT[(k + 1) + ld * (k - 1)] = 0.0; }
                    double beta = h0;
                    x1 = h1; x2 = h2;
                    d_larfg3f(beta, x1, x2, tau);
                    double hnew = beta;
                    if (h2 == 0.0 && hk2 == 0.0 && hk21 != 0.0) {
                        // collapsed bulge: try to re-introduce it (DHGEQZ5.f:141-189)
                        double wv[3];
                        d_shift_col(sH + k + ld * k, ld, sT + k + ld * k, ld, a.sr[sp], a.si[sp], a.sb[sp],
                                    a.sr[sp + 1], a.sb[sp + 1], wv);
                        double a1 = wv[0], y1 = wv[1], y2 = wv[2], tau1;
                        d_larfg3f(a1, y1, y2, tau1);
                        double refsum = tau1 * (h0 + y1 * h1);
                        double tst = fabs(sH[c + ld * c]) + fabs(sH[k + ld * k]) + fabs(sH[(k + 1) + ld * (k + 1)]);
                        if (!(fabs(h1 - refsum * y1) + fabs(refsum * y2) > GQZ_ULP * tst)) {
                            hnew = h0 - refsum;
                            tau = tau1; x1 = y1; x2 = y2;
                        }
                    }
                    sH[k + ld * c] = hnew; sH[(k + 1) + ld * c] = 0.0; sH[(k + 2) + ld * c] = 0.0;
                }
                double r1[3], r2[3];
#pragma unroll
                for (int jj = 0; jj < 3; ++jj) {
                    const double t0 = t3[3 * jj], t1 = t3[1 + 3 * jj], t2 = t3[2 + 3 * jj];
                    const double sj = tau * (t0 + x1 * t1 + x2 * t2);
                    r1[jj] = t1 - sj * x1; r2[jj] = t2 - sj * x2;
                }
                double wv[3], tau2;
                d_nullrefl3(r1[0], r1[1], r1[2], r2[0], r2[1], r2[2], wv, tau2);
                B[0] = x1; B[1] = x2; B[2] = tau; B[3] = wv[1]; B[4] = wv[2]; B[5] = tau2;
            } else if (kind == 2) {
                const int l = w - 1;
                const double x0 = sT[(l - 1) + ld * (l - 1)], y0 = sT[l + ld * (l - 1)];
                const double xx1 = sT[(l - 1) + ld * l], yy1 = sT[l + ld * l];
                if (q > 0) { sT[(l - 1) + ld * (l - 2)] = 0.0; sT[l + ld * (l - 2)] = 0.0; }
                double cs, sn, r;
                d_lartgf(sH[(l - 1) + ld * (l - 2)], sH[l + ld * (l - 2)], cs, sn, r);
                sH[(l - 1) + ld * (l - 2)] = r; sH[l + ld * (l - 2)] = 0.0;
                // the closing column rotation, from T(l-1:l, l-1:l) after the row rotation
                double c2, s2, r2;
                d_lartgf(cs * yy1 - sn * xx1, cs * y0 - sn * x0, c2, s2, r2);
                B[6] = cs; B[7] = sn; B[3] = c2; B[4] = s2; B[5] = r2;
            }
            bc[b] = c; bk[b] = kind;
        }
        __syncthreads();
        // apply all left transforms: one wave per (bulge, matrix) segment; the lane is the column
        // of H or T (rows k..k+2) or the row of U (columns k..k+2)
        for (int seg = wave; seg < 3 * nb; seg += CH_NT / 64) {
            const int b = seg / 3, typ = seg - 3 * b;
            const double *B = bv + 8 * b;
            const int kind = bk[b], k = bc[b] + 1;
            const double v1 = B[0], v2 = B[1], tau = B[2], cs = B[6], sn = B[7];
            if (kind == 0 || lane >= w) continue;
            if (kind == 2) {
                const int i = w - 2;
                if (typ < 2) {
                    if (lane < w - 2) continue;
                    double *M = typ == 0 ? sH : sT;
                    double x = M[i + ld * lane], y = M[(i + 1) + ld * lane];
                    M[i + ld * lane] = cs * x + sn * y; M[(i + 1) + ld * lane] = cs * y - sn * x;
                } else {
                    double x = sU[lane + ld * i], y = sU[lane + ld * (i + 1)];
                    sU[lane + ld * i] = cs * x + sn * y; sU[lane + ld * (i + 1)] = cs * y - sn * x;
                }
                continue;
            }
            if (tau == 0.0) continue;
            if (typ < 2) {
                if (lane < k) continue;                 // columns k.. (column c set explicitly)
                double *p = (typ == 0 ? sH : sT) + k + ld * lane;
                double t = tau * (p[0] + v1 * p[1] + v2 * p[2]);
                p[0] -= t; p[1] -= t * v1; p[2] -= t * v2;
            } else {
                double *x = sU + lane + ld * k;
                double x0 = x[0], x1 = x[ld], x2 = x[2 * ld];
                double t = tau * (x0 + v1 * x1 + v2 * x2);
                x[0] = x0 - t; x[ld] = x1 - t * v1; x[2 * ld] = x2 - t * v2;
            }
        }
        __syncthreads();
        // apply all right transforms: one wave per (bulge, matrix) segment; the lane is the row
        for (int seg = wave; seg < 3 * nb; seg += CH_NT / 64) {
            const int b = seg / 3, typ = seg - 3 * b;
            const double *B = bv + 8 * b;
            const int kind = bk[b], k = bc[b] + 1;
            const double w1 = B[3], w2 = B[4], tau2 = B[5];    // kind 2: c, s, r of the rotation
            if (kind == 0 || lane >= w) continue;
            double *M = typ == 0 ? sH : (typ == 1 ? sT : sV);
            if (kind == 2) {
                const int l = w - 1;
                if (typ == 1 && lane >= l) {
                    if (lane == l) { sT[l + ld * l] = tau2; sT[l + ld * (l - 1)] = 0.0; }
                    continue;
                }
                double x = M[lane + ld * l], y = M[lane + ld * (l - 1)];
                M[lane + ld * l] = w1 * x + w2 * y; M[lane + ld * (l - 1)] = w1 * y - w2 * x;
                continue;
            }
            if (tau2 == 0.0) continue;
            if (typ == 0 && lane >= min(k + 4, w)) continue;
            if (typ == 1 && lane >= k + 3) continue;
            double *x = M + lane + ld * k;
            double x0 = x[0], x1 = x[ld], x2 = x[2 * ld];
            double t = tau2 * (x0 + w1 * x1 + w2 * x2);
            x[0] = x0 - t; x[ld] = x1 - t * w1; x[2 * ld] = x2 - t * w2;
        }
        __syncthreads();
    }
    // clean the exact zeros left by the last right reflector of every bulge, then the
    // vigilant deflation check of DHGEQZ5.f:254-275 over the window
    if (tid < nb) {
        const int b = tid;
        const int start = a.intro ? 3 * (nb - 1 - b) : 0;
        const int nst = a.intro ? 3 * b + 1 + a.steps : a.steps;
        const int cst = a.intro ? -1 : a.t0 + 3 * b;
        int last_c = cst + min(nst, total - start) - 1;
        int k = last_c + 1;
        if (k + 2 <= w - 1 && k >= 0) { sT[(k + 1) + ld * k] = 0.0; sT[(k + 2) + ld * k] = 0.0; }
    }
    __syncthreads();
    {
        // (the test of H(j, j-1) reads no entry that another test writes: all at once)
        for (int j = 1 + tid; j < w; j += blockDim.x) {
            double hj = sH[j + ld * (j - 1)];
            if (hj == 0.0) continue;
            double x = sH[(j - 1) + ld * (j - 1)], y = sH[j + ld * j];
            if (fabs(hj) <= fmax(a.smlnum, GQZ_ULP * (fabs(x) + fabs(y)))) {
                double up = sH[(j - 1) + ld * j];
                double h12 = fmax(fabs(hj), fabs(up)), h21 = fmin(fabs(hj), fabs(up));
                double h11 = fmax(fabs(y), fabs(x - y)), h22 = fmin(fabs(y), fabs(x - y));
                double scl = h11 + h12, tst2 = h22 * (h11 / scl);
                if (tst2 == 0.0 || h21 * (h12 / scl) <= fmax(a.smlnum, GQZ_ULP * tst2)) sH[j + ld * (j - 1)] = 0.0;
            }
        }
    }
    __syncthreads();
    double *oH = a.H + GQZ_IX(a.p, a.p, a.ld);
    double *oT = a.T + GQZ_IX(a.p, a.p, a.ld);
    for (int idx = tid; idx < w * w; idx += blockDim.x) {
        int i = idx % w, j = idx / w;
        oH[GQZ_IX(i, j, a.ld)] = sH[i + ld * j];
        oT[GQZ_IX(i, j, a.ld)] = sT[i + ld * j];
        a.U[i + (size_t)w * j] = sU[i + ld * j];
        a.V[i + (size_t)w * j] = sV[i + ld * j];
    }
}

int gqz_chase_max_window() { return CH_MAXW; }

static size_t chase_lds(int w) {
    return sizeof(double) * (4 * (size_t)(w + 1) * w + 8 * CH_MAXB) + sizeof(int) * 2 * CH_MAXB;
}

void gqz_chase_set(hipStream_t st, const GqzChainSet &s) {
    if (s.nch <= 0) return;
    int wm = 0;
    bool nr = false;
    for (int c = 0; c < s.nch; ++c) {
        wm = max(wm, s.w[c]);
        nr = nr || s.near[c] > 0;
        // the near block must be exactly the block [0, pw - d) x [pw - d, w) of the window, d = p - pp
        if (s.near[c] > 0 && (s.pw[c] > CH_MAXW || s.pw[c] < 1 || s.p[c] <= s.pp[c] || s.p[c] >= s.pp[c] + s.pw[c] ||
                              s.pp[c] + s.pw[c] + s.near[c] != s.p[c] + s.w[c]))
            throw std::runtime_error("chase: deferred update does not match the window");
    }
    size_t lds = chase_lds(wm);
    if (nr) lds = max(lds, CH_NEAR_LDS);      // the staged U and near block of the near update
    static size_t lds_set[64] = {};
    gqz_need_lds((const void *)k_chase, lds, lds_set);
    hipLaunchKernelGGL(k_chase, dim3(s.nch), dim3(CH_NT), lds, st, s);
}

void gqz_chase_window(hipStream_t st, const GqzChainArgs &a) {
    GqzChainSet s = {};
    s.H = a.H; s.T = a.T; s.ld = a.ld; s.nch = 1;
    s.U = a.U; s.V = a.V; s.ustride = 0;
    s.sr = a.sr; s.si = a.si; s.sb = a.sb; s.smlnum = a.smlnum;
    s.p[0] = a.p; s.w[0] = a.w; s.nb[0] = a.nb; s.t0[0] = a.t0; s.steps[0] = a.steps;
    s.intro[0] = a.intro; s.last[0] = a.last; s.soff[0] = 0;
    gqz_chase_set(st, s);
}
