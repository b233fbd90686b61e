// k_aed.hip -- aggressive early deflation post-processing on a window in HBM (QZEARLY,
// src/kkqz/qzearly.f:250-444; PDHGEQZ3.f:630-935 on a 1x1 grid).
//
// k_aed_reorder: deflation test of the spike from the bottom of the window; undeflatable
//   blocks are moved to the top.  QZEARLY moves each undeflatable block to the top before the
//   next test (O(jw^2) dependent swaps).  Here every undeflatable block found so far moves up by
//   one block per round ("train"): in a round all pairs (unexamined block above, moving block
//   below) swap simultaneously - their index sets are disjoint, rows are updated before
//   columns - and the block that a moving block has just passed is tested as soon as it sits
//   at the bottom.  Each moving block passes the same blocks in the same order as in QZEARLY,
//   and every test sees the same spike entries, so the result equals QZEARLY's up to rounding
//   while the number of dependent steps drops from O(jw^2) to O(jw).
// k_aed_spike: eigenvalues of the window blocks, the spike reflector (qzearly.f:369-386) and
//   the RQ factorisation of T (qzearly.f:387-392).
#include "gqz_wg.h"
#include "gqz_swap.h"
#include "gqz_misc.h"

#define AED_MAXW 1024
enum { ST_U = 0, ST_X = 1, ST_A = 2, ST_D = 3 };

// per-pair scratch record in HBM: Qa[16] Za[16] S[16] T[16] flag
#define PREC 65

template <int N1, int N2>
__device__ __forceinline__ void pair_swap(const double *Hw, const double *Tw, int jw, int j1, double *rec) {
    constexpr int M = N1 + N2;
    double S[M * M], T[M * M], Qa[M * M], Za[M * M];
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
        for (int i = 0; i < M; ++i) {
            S[i + M * j] = Hw[GQZ_IX(j1 + i, j1 + j, jw)];
            T[i + M * j] = Tw[GQZ_IX(j1 + i, j1 + j, jw)];
        }
    int f = r_tgex2<N1, N2>(S, T, Qa, Za);
    rec[64] = f ? 1.0 : 0.0;
    if (f) return;
#pragma unroll
    for (int q = 0; q < M * M; ++q) { rec[q] = Qa[q]; rec[16 + q] = Za[q]; rec[32 + q] = S[q]; rec[48 + q] = T[q]; }
}

extern "C" __global__ void __launch_bounds__(256) k_aed_reorder(double *Hw, double *Tw, double *U, double *V,
                                                                int jw, const double *sptr, double smlnum,
                                                                double *pscr, int *out) {
    __shared__ unsigned char st[AED_MAXW];
    __shared__ short tgt[AED_MAXW];
    __shared__ short pj1[AED_MAXW / 2];
    __shared__ unsigned char pn1[AED_MAXW / 2], pn2[AED_MAXW / 2];
    __shared__ int s_np, s_kb, s_ilst, s_knt, s_fail, s_swaps;
    const int tid = threadIdx.x;
    const double s = sptr ? *sptr : 0.0;
    for (int i = tid; i < jw; i += blockDim.x) { st[i] = ST_U; tgt[i] = 0; }
    if (tid == 0) { s_kb = jw - 1; s_ilst = 0; s_knt = 0; s_fail = 0; s_swaps = 0; }
    __syncthreads();
    for (int round = 0; round < 4 * jw + 8; ++round) {
        if (tid == 0) {
            // ---- test phase: examine blocks at the bottom while they are unexamined ----
            while (s_kb >= 0 && s_knt < jw) {
                int kb = s_kb;
                int b0 = (kb > 0 && Hw[GQZ_IX(kb, kb - 1, jw)] != 0.0) ? kb - 1 : kb;
                if (st[b0] != ST_U) break;
                int bulge = (b0 != kb);
                bool dfl;
                if (!bulge) {
                    s_knt += 1;
                    double tmp = fabs(s) + fabs(Hw[GQZ_IX(kb, kb, jw)]);
                    dfl = fabs(s * U[GQZ_IX(0, kb, jw)]) <= fmax(GQZ_ULP * tmp, smlnum);
                } else {
                    s_knt += 2;
                    double h00 = Hw[GQZ_IX(kb - 1, kb - 1, jw)], h10 = Hw[GQZ_IX(kb, kb - 1, jw)];
                    double h01 = Hw[GQZ_IX(kb - 1, kb, jw)], h11 = Hw[GQZ_IX(kb, kb, jw)];
                    double rtdet;
                    if (h00 == 0.0 && h10 == 0.0) rtdet = 0.0;
                    else if (fabs(h00) >= fabs(h10)) rtdet = sqrt(fabs(h00)) * sqrt(fabs(h11 - (h10 / h00) * h01));
                    else rtdet = sqrt(fabs(h10)) * sqrt(fabs(h01 - (h00 / h10) * h11));
                    double tmp = fabs(s) + rtdet;
                    dfl = fmax(fabs(s * U[GQZ_IX(0, kb, jw)]), fabs(s * U[GQZ_IX(0, kb - 1, jw)])) <=
                          fmax(GQZ_ULP * tmp, smlnum);
                }
                if (dfl) { st[b0] = ST_D; s_kb = b0 - 1; }
                else {
                    st[b0] = (b0 == s_ilst) ? ST_A : ST_X;
                    tgt[b0] = (short)s_ilst;
                    s_ilst += bulge ? 2 : 1;
                    break;
                }
            }
            // ---- pair phase: (unexamined block above, moving block below) ----
            int np = 0;
            int pst = -1, psz = 0, pstate = ST_A;
            for (int p = 0; p <= s_kb;) {
                int sz = (p + 1 < jw && Hw[GQZ_IX(p + 1, p, jw)] != 0.0) ? 2 : 1;
                int stt = st[p];
                if (stt == ST_X && p > tgt[p] && pstate == ST_U && pst >= 0) {
                    pj1[np] = (short)pst; pn1[np] = (unsigned char)psz; pn2[np] = (unsigned char)sz; ++np;
                    // the pair consumes both blocks; the next block has no usable predecessor
                    pst = -1; pstate = ST_A; p += sz;
                    continue;
                }
                pst = p; psz = sz; pstate = stt;
                p += sz;
            }
            s_np = np;
        }
        __syncthreads();
        const int np = s_np;
        if (np == 0) break;
        // ---- compute the swaps (one lane per pair, register resident) ----
        for (int q = tid; q < np; q += blockDim.x) {
            const int j1 = pj1[q], n1 = pn1[q], n2 = pn2[q];
            double *rec = pscr + (size_t)PREC * q;
            if (n1 == 1 && n2 == 1) pair_swap<1, 1>(Hw, Tw, jw, j1, rec);
            else if (n1 == 1 && n2 == 2) pair_swap<1, 2>(Hw, Tw, jw, j1, rec);
            else if (n1 == 2 && n2 == 1) pair_swap<2, 1>(Hw, Tw, jw, j1, rec);
            else pair_swap<2, 2>(Hw, Tw, jw, j1, rec);
        }
        __syncthreads();
        // ---- rows of every pair (cols right of its block) and U columns ----
        for (size_t idx = tid; idx < (size_t)np * 3 * jw; idx += blockDim.x) {
            const int q = (int)(idx / (3 * jw)), r = (int)(idx % (3 * jw));
            const double *rec = pscr + (size_t)PREC * q;
            if (rec[64] != 0.0) continue;
            const int j1 = pj1[q], m = pn1[q] + pn2[q];
            const double *Qa = rec;
            double x[4], y[4];
            if (r < 2 * jw) {
                double *M = r < jw ? Hw : Tw;
                const int col = r < jw ? r : r - jw;
                if (col < j1 + m) continue;
                for (int i = 0; i < m; ++i) x[i] = M[GQZ_IX(j1 + i, col, jw)];
                for (int i = 0; i < m; ++i) { double a = 0; for (int k = 0; k < m; ++k) a += Qa[k + m * i] * x[k]; y[i] = a; }
                for (int i = 0; i < m; ++i) M[GQZ_IX(j1 + i, col, jw)] = y[i];
            } else {
                const int row = r - 2 * jw;
                for (int j = 0; j < m; ++j) x[j] = U[GQZ_IX(row, j1 + j, jw)];
                for (int j = 0; j < m; ++j) { double a = 0; for (int k = 0; k < m; ++k) a += x[k] * Qa[k + m * j]; y[j] = a; }
                for (int j = 0; j < m; ++j) U[GQZ_IX(row, j1 + j, jw)] = y[j];
            }
        }
        __syncthreads();
        // ---- columns of every pair (rows above its block), V, and the new diagonal blocks ----
        for (size_t idx = tid; idx < (size_t)np * (3 * jw + 16); idx += blockDim.x) {
            const int q = (int)(idx / (3 * jw + 16)), r = (int)(idx % (3 * jw + 16));
            const double *rec = pscr + (size_t)PREC * q;
            if (rec[64] != 0.0) continue;
            const int j1 = pj1[q], m = pn1[q] + pn2[q];
            const double *Za = rec + 16;
            double x[4], y[4];
            if (r >= 3 * jw) {
                const int e = r - 3 * jw;
                if (e < m * m) {
                    const int i = e % m, j = e / m;
                    Hw[GQZ_IX(j1 + i, j1 + j, jw)] = rec[32 + e];
                    Tw[GQZ_IX(j1 + i, j1 + j, jw)] = rec[48 + e];
                }
                continue;
            }
            double *M; int row;
            if (r < jw) { M = Hw; row = r; if (row >= j1) continue; }
            else if (r < 2 * jw) { M = Tw; row = r - jw; if (row >= j1) continue; }
            else { M = V; row = r - 2 * jw; }
            for (int j = 0; j < m; ++j) x[j] = M[GQZ_IX(row, j1 + j, jw)];
            for (int j = 0; j < m; ++j) { double a = 0; for (int k = 0; k < m; ++k) a += x[k] * Za[k + m * j]; y[j] = a; }
            for (int j = 0; j < m; ++j) M[GQZ_IX(row, j1 + j, jw)] = y[j];
        }
        __syncthreads();
        // ---- state update ----
        if (tid == 0) {
            for (int q = 0; q < np; ++q) {
                const double *rec = pscr + (size_t)PREC * q;
                if (rec[64] != 0.0) { s_fail = 1; continue; }
                const int j1 = pj1[q], n1 = pn1[q], n2 = pn2[q];
                const int t = tgt[j1 + n1];
                s_swaps++;
                // the moving block now starts at j1 (order n2), the passed block at j1 + n2 (n1)
                bool xsplit = (n2 == 2) && Hw[GQZ_IX(j1 + 1, j1, jw)] == 0.0;
                st[j1] = (j1 == t) ? ST_A : ST_X; tgt[j1] = (short)t;
                if (xsplit) { st[j1 + 1] = (j1 + 1 == t + 1) ? ST_A : ST_X; tgt[j1 + 1] = (short)(t + 1); }
                st[j1 + n2] = ST_U;
                if (n1 == 2) st[j1 + n2 + 1] = ST_U;
            }
        }
        __syncthreads();
        if (s_fail) break;
    }
    if (tid == 0) {
        int kbc = s_kb + 1;
        if (!s_fail && kbc < s_ilst) kbc = s_ilst;
        out[0] = kbc; out[1] = s_swaps; out[2] = s_fail; out[3] = (s != 0.0);
    }
}

// eigenvalues of the window's diagonal blocks (the shifts and the deflated eigenvalues)
extern "C" __global__ void __launch_bounds__(256) k_aed_eigs(const double *Hw, const double *Tw, int jw,
                                                             double *war, double *wai, double *wb) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < jw; j += gridDim.x * blockDim.x) {
        const bool lower2 = (j > 0) && Hw[GQZ_IX(j, j - 1, jw)] != 0.0;
        const bool upper2 = (j + 1 < jw) && Hw[GQZ_IX(j + 1, j, jw)] != 0.0;
        if (lower2) continue;
        if (upper2) {
            double A[4] = { Hw[GQZ_IX(j, j, jw)], Hw[GQZ_IX(j + 1, j, jw)], Hw[GQZ_IX(j, j + 1, jw)], Hw[GQZ_IX(j + 1, j + 1, jw)] };
            double B[4] = { Tw[GQZ_IX(j, j, jw)], 0.0, Tw[GQZ_IX(j, j + 1, jw)], Tw[GQZ_IX(j + 1, j + 1, jw)] };
            double er[2], ei[2];
            d_eig22(A[0], A[1], A[2], A[3], B[0], B[2], B[3], er, ei);
            double b0 = B[0], b1 = B[3];
            if (ei[0] != 0.0) {
                war[j] = er[0] * b0; wai[j] = fabs(ei[0]) * b0; wb[j] = b0;
                war[j + 1] = er[0] * b1; wai[j + 1] = -fabs(ei[0]) * b1; wb[j + 1] = b1;
            } else {
                war[j] = er[0] * b0; wai[j] = 0.0; wb[j] = b0;
                war[j + 1] = er[1] * b1; wai[j + 1] = 0.0; wb[j + 1] = b1;
            }
        } else {
            war[j] = Hw[GQZ_IX(j, j, jw)]; wai[j] = 0.0; wb[j] = Tw[GQZ_IX(j, j, jw)];
        }
    }
}

// Spike reflector and RQ update (qzearly.f:372-436 with PARA_AGGRQUP = .TRUE., kkqzconf.f:138):
//   P = I - tau v v^T maps the spike s U(0, 0:kbc) onto a multiple of e_0 and is applied to the
//   rows 0..kbc-1 of Hw (all columns) and T's columns right of the block, and to U's columns.
//   P T(0:kbc, 0:kbc) = T + v w^T with w = -tau T^T v is brought back to triangular form by
//   2 kbc - 2 column rotations: the first sweep (i = 0 .. kbc-2) folds w into its last entry
//   (T becomes Hessenberg, the rank-one term lives in the last column), the second sweep
//   (i = kbc-2 .. 0) removes the subdiagonal.  Hw's block columns and V see the same rotations.
// One workgroup; kbc <= AED_MAXW.  Rotations are kept in scr (4 kbc doubles).
extern "C" __global__ void __launch_bounds__(1024) k_aed_spike(double *Hw, double *Tw, double *U, double *V,
                                                               int jw, const double *sptr, double *scr,
                                                               const int *out) {
    __shared__ double sv[AED_MAXW], sw[AED_MAXW];
    __shared__ double red[1024];
    __shared__ double s_tau, s_cs, s_sn;
    const int tid = threadIdx.x, nt = blockDim.x;
    const double s = sptr ? *sptr : 0.0;
    const int kbc = out[0];
    if (kbc <= 0 || s == 0.0 || kbc == jw) return;
    // ---- reflector (DLARFG on s U(0, 0:kbc)) ----
    for (int i = tid; i < kbc; i += nt) sv[i] = s * U[GQZ_IX(0, i, jw)];
    __syncthreads();
    double ss = 0.0;
    for (int i = 1 + tid; i < kbc; i += nt) ss += sv[i] * sv[i];
    red[tid] = ss;
    __syncthreads();
    for (int h = nt / 2; h > 0; h >>= 1) { if (tid < h) red[tid] += red[tid + h]; __syncthreads(); }
    if (tid == 0) {
        const double xnorm = sqrt(red[0]), alpha = sv[0];
        if (xnorm == 0.0) { s_tau = 0.0; s_cs = 1.0; }
        else {
            const double beta = -copysign(hypot(alpha, xnorm), alpha);
            s_tau = (beta - alpha) / beta;
            s_cs = 1.0 / (alpha - beta);
        }
    }
    __syncthreads();
    const double tau = s_tau;
    if (tau == 0.0) return;
    for (int i = 1 + tid; i < kbc; i += nt) sv[i] *= s_cs;
    if (tid == 0) sv[0] = 1.0;
    __syncthreads();
    // ---- apply P: H rows (all columns), T rows (columns right of the block), U columns; and
    //      w = -tau T(0:kbc,0:kbc)^T v (T upper triangular) ----
    for (int idx = tid; idx < jw + (jw - kbc) + jw + kbc; idx += nt) {
        if (idx < jw + (jw - kbc)) {
            double *p = (idx < jw) ? Hw + GQZ_IX(0, idx, jw) : Tw + GQZ_IX(0, kbc + idx - jw, jw);
            double t = 0.0;
            for (int i = 0; i < kbc; ++i) t += sv[i] * p[i];
            t *= tau;
            for (int i = 0; i < kbc; ++i) p[i] -= t * sv[i];
        } else if (idx < 3 * jw - kbc) {
            const int row = idx - (2 * jw - kbc);
            double t = 0.0;
            for (int j = 0; j < kbc; ++j) t += U[GQZ_IX(row, j, jw)] * sv[j];
            t *= tau;
            for (int j = 0; j < kbc; ++j) U[GQZ_IX(row, j, jw)] -= t * sv[j];
        } else {
            const int j = idx - (3 * jw - kbc);
            double t = 0.0;
            for (int i = 0; i <= j; ++i) t += Tw[GQZ_IX(i, j, jw)] * sv[i];
            sw[j] = -tau * t;
        }
    }
    __syncthreads();
    // ---- first sweep: rotations (i, i+1) fold w into w(kbc-1) (qzearly.f:406-412) ----
    double *rc = scr, *rs = scr + kbc, *qc = scr + 2 * kbc, *qs = scr + 3 * kbc;
    if (tid == 0) {
        for (int i = 0; i < kbc - 1; ++i) {
            double c, sn, r;
            d_lartg(sw[i + 1], sw[i], c, sn, r);
            sw[i + 1] = r;
            rc[i] = c; rs[i] = -sn;
        }
    }
    __syncthreads();
    // T columns (i, i+1), rows 0..i+1, in order i = 0..kbc-2: each row applies the rotations that
    // reach it (rotation i touches rows <= i + 1) in order
    for (int row = tid; row < kbc; row += nt) {
        const int i0 = max(0, row - 1);
        if (i0 >= kbc - 1) continue;
        double x = Tw[GQZ_IX(row, i0, jw)];
        for (int i = i0; i < kbc - 1; ++i) {
            double y = Tw[GQZ_IX(row, i + 1, jw)];
            Tw[GQZ_IX(row, i, jw)] = rc[i] * x + rs[i] * y;
            x = rc[i] * y - rs[i] * x;
        }
        Tw[GQZ_IX(row, kbc - 1, jw)] = x;
    }
    __syncthreads();
    // rank-one term now lives in the last column: T(:, kbc-1) += w(kbc-1) v
    for (int i = tid; i < kbc; i += nt) Tw[GQZ_IX(i, kbc - 1, jw)] += sw[kbc - 1] * sv[i];
    // H block columns (rows 0..kbc-1; H below row kbc is zero in these columns) and V (all rows)
    for (int row = tid; row < kbc + jw; row += nt) {
        double *M = row < kbc ? Hw : V;
        const int r = row < kbc ? row : row - kbc;
        double x = M[GQZ_IX(r, 0, jw)];
        for (int i = 0; i < kbc - 1; ++i) {
            double y = M[GQZ_IX(r, i + 1, jw)];
            M[GQZ_IX(r, i, jw)] = rc[i] * x + rs[i] * y;
            x = rc[i] * y - rs[i] * x;
        }
        M[GQZ_IX(r, kbc - 1, jw)] = x;
    }
    __syncthreads();
    // ---- second sweep: i = kbc-2 .. 0 removes T(i+1, i) (qzearly.f:421-429) ----
    for (int i = kbc - 2; i >= 0; --i) {
        if (tid == 0) {
            double c, sn, r;
            d_lartg(Tw[GQZ_IX(i + 1, i + 1, jw)], Tw[GQZ_IX(i + 1, i, jw)], c, sn, r);
            Tw[GQZ_IX(i + 1, i + 1, jw)] = r;
            Tw[GQZ_IX(i + 1, i, jw)] = 0.0;
            qc[i] = c; qs[i] = -sn;
            s_cs = c; s_sn = -sn;
        }
        __syncthreads();
        const double c = s_cs, sn = s_sn;
        for (int row = tid; row <= i; row += nt) {
            double x = Tw[GQZ_IX(row, i, jw)], y = Tw[GQZ_IX(row, i + 1, jw)];
            Tw[GQZ_IX(row, i, jw)] = c * x + sn * y;
            Tw[GQZ_IX(row, i + 1, jw)] = c * y - sn * x;
        }
        __syncthreads();
    }
    for (int row = tid; row < kbc + jw; row += nt) {
        double *M = row < kbc ? Hw : V;
        const int r = row < kbc ? row : row - kbc;
        double y = M[GQZ_IX(r, kbc - 1, jw)];
        for (int i = kbc - 2; i >= 0; --i) {
            double x = M[GQZ_IX(r, i, jw)];
            M[GQZ_IX(r, i + 1, jw)] = qc[i] * y - qs[i] * x;
            y = qc[i] * x + qs[i] * y;
        }
        M[GQZ_IX(r, 0, jw)] = y;
    }
}

void gqz_launch_aed_reorder(hipStream_t st, double *Hw, double *Tw, double *U, double *V, int jw,
                            const double *sptr, double smlnum, double *pscr, int *out) {
    hipLaunchKernelGGL(k_aed_reorder, dim3(1), dim3(256), 0, st, Hw, Tw, U, V, jw, sptr, smlnum, pscr, out);
}
void gqz_launch_aed_eigs(hipStream_t st, const double *Hw, const double *Tw, int jw, double *war, double *wai,
                         double *wb) {
    hipLaunchKernelGGL(k_aed_eigs, dim3((jw + 255) / 256), dim3(256), 0, st, Hw, Tw, jw, war, wai, wb);
}
void gqz_launch_aed_spike(hipStream_t st, double *Hw, double *Tw, double *U, double *V, int jw,
                          const double *sptr, double *scr, const int *out) {
    hipLaunchKernelGGL(k_aed_spike, dim3(1), dim3(1024), 0, st, Hw, Tw, U, V, jw, sptr, scr, out);
}
