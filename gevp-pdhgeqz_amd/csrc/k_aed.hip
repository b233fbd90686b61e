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

// eigenvalues of the window blocks + spike reflector + RQ of T(0:kbc, 0:kbc)
extern "C" __global__ void __launch_bounds__(1024) k_aed_spike(double *Hw, double *Tw, double *U, double *V,
                                                               int jw, const double *sptr, double *war,
                                                               double *wai, double *wb, double *scr,
                                                               const int *out) {
    __shared__ double s_tau;
    const int tid = threadIdx.x;
    const double s = sptr ? *sptr : 0.0;
    const int kbc = out[0];
    for (int j = tid; j < jw; j += blockDim.x) {
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
    __syncthreads();
    if (kbc <= 0 || s == 0.0) return;
    double *vsp = scr;
    if (tid == 0) {
        for (int i = 0; i < kbc; ++i) vsp[i] = s * U[GQZ_IX(0, i, jw)];
        double alpha = vsp[0], tau;
        d_larfg(kbc, alpha, vsp + 1, 1, tau);
        vsp[0] = 1.0;
        s_tau = tau;
    }
    __syncthreads();
    {
        const double tau = s_tau;
        for (int idx = tid; idx < 3 * jw; idx += blockDim.x) {
            if (idx < 2 * jw) {
                double *M = idx < jw ? Hw : Tw;
                const int col = idx < jw ? idx : idx - jw;
                double *p = M + GQZ_IX(0, col, jw);
                double t = 0.0;
                for (int i = 0; i < kbc; ++i) t += vsp[i] * p[i];
                t *= tau;
                for (int i = 0; i < kbc; ++i) p[i] -= t * vsp[i];
            } else {
                const int row = idx - 2 * jw;
                double t = 0.0;
                for (int j = 0; j < kbc; ++j) t += U[GQZ_IX(row, j, jw)] * vsp[j];
                t *= tau;
                for (int j = 0; j < kbc; ++j) U[GQZ_IX(row, j, jw)] -= t * vsp[j];
            }
        }
    }
    __syncthreads();
    for (int r = kbc - 1; r >= 1; --r) {
        if (tid == 0) {
            for (int j = 0; j < r; ++j) vsp[j] = Tw[GQZ_IX(r, j, jw)];
            double alpha = Tw[GQZ_IX(r, r, jw)], tau;
            d_larfg(r + 1, alpha, vsp, 1, tau);
            vsp[r] = 1.0;
            Tw[GQZ_IX(r, r, jw)] = alpha;
            for (int j = 0; j < r; ++j) Tw[GQZ_IX(r, j, jw)] = 0.0;
            s_tau = tau;
        }
        __syncthreads();
        const double tau = s_tau;
        for (int idx = tid; idx < r + 2 * jw; idx += blockDim.x) {
            double *M; int row;
            if (idx < r) { M = Tw; row = idx; }
            else if (idx < r + jw) { M = Hw; row = idx - r; }
            else { M = V; row = idx - r - jw; }
            double t = 0.0;
            for (int j = 0; j <= r; ++j) t += M[GQZ_IX(row, j, jw)] * vsp[j];
            t *= tau;
            for (int j = 0; j <= r; ++j) M[GQZ_IX(row, j, jw)] -= t * vsp[j];
        }
        __syncthreads();
    }
}

void gqz_launch_aed_reorder(hipStream_t st, double *Hw, double *Tw, double *U, double *V, int jw,
                            const double *sptr, double smlnum, double *pscr, int *out) {
    hipLaunchKernelGGL(k_aed_reorder, dim3(1), dim3(256), 0, st, Hw, Tw, U, V, jw, sptr, smlnum, pscr, out);
}
void gqz_launch_aed_spike(hipStream_t st, double *Hw, double *Tw, double *U, double *V, int jw,
                          const double *sptr, double *war, double *wai, double *wb, double *scr, const int *out) {
    hipLaunchKernelGGL(k_aed_spike, dim3(1), dim3(1024), 0, st, Hw, Tw, U, V, jw, sptr, war, wai, wb, scr, out);
}
