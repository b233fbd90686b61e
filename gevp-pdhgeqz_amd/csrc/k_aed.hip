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
#include <hip/hip_cooperative_groups.h>
#include <algorithm>

#include "gqz_ds.h"
#include "gqz_swap.h"
#include "gqz_misc.h"

extern "C" void *gqz_stream(void);

#define AED_MAXW 1024
enum { ST_U = 0, ST_X = 1, ST_A = 2, ST_D = 3 };

// per-pair scratch record in HBM: Qa[16] Za[16] S[16] T[16] flag
#define PREC 65

template <int N1, int N2>
__device__ __forceinline__ void pair_swap(const double *Hw, const double *Tw, int ld, int j1, double *rec) {
    constexpr int M = N1 + N2;
    double S[M * M], T[M * M], Qa[M * M], Za[M * M];
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
        for (int i = 0; i < M; ++i) {
            S[i + M * j] = Hw[GQZ_IX(j1 + i, j1 + j, ld)];
            T[i + M * j] = Tw[GQZ_IX(j1 + i, j1 + j, ld)];
        }
    int f = r_tgex2<N1, N2>(S, T, Qa, Za);
    rec[64] = f ? 1.0 : 0.0;
    if (f) return;
#pragma unroll
    for (int q = 0; q < M * M; ++q) { rec[q] = Qa[q]; rec[16 + q] = Za[q]; rec[32 + q] = S[q]; rec[48 + q] = T[q]; }
}

// Block structure lives in LDS: bsz[p] = order of the block starting at row p (0 inside a 2x2
// block), st[p] / tgt[p] = state and target row of that block.  out[2] = 0 (done), 1 (the window is
// not quasi-triangular: nothing was swapped) or 2 (a swap was rejected: the round's other swaps are
// in place, the rejected pair wrote nothing, and every row down to the bottom of the unexamined
// region counts as undeflatable, kbc = s_kb + 1, as QZEARLY goes on to the HT restore after a
// failed DTGEXC, qzearly.f:318-326).  One round:
//   test (lane 0)  -> pair detection (all lanes, LDS) -> swap factors (one lane per pair)
//   -> rows of every pair -> columns of every pair -> block-structure update (one lane per pair).
template <int MAXW>
__device__ void aed_reorder_body(double *Hw, double *Tw, double *U, double *V, int ld, int jw, double s,
                                 double smlnum, double *pscr, int *out) {
    __shared__ unsigned char st[MAXW], bsz[MAXW];
    __shared__ short tgt[MAXW];
    __shared__ short pj1[MAXW / 2];
    __shared__ unsigned char pn1[MAXW / 2], pn2[MAXW / 2];
    __shared__ int s_np, s_kb, s_ilst, s_knt, s_fail, s_swaps;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < jw; i += nt) {
        st[i] = ST_U; tgt[i] = 0;
        const bool lower = (i > 0) && Hw[GQZ_IX(i, i - 1, ld)] != 0.0;
        const bool upper = (i + 1 < jw) && Hw[GQZ_IX(i + 1, i, ld)] != 0.0;
        bsz[i] = lower ? 0 : (upper ? 2 : 1);
    }
    if (tid == 0) { s_kb = jw - 1; s_ilst = 0; s_knt = 0; s_fail = 0; s_swaps = 0; s_np = 0; }
    __syncthreads();
    // a window that is not quasi-triangular (two consecutive non-zero subdiagonal entries: an
    // unreduced 3x3 or larger block) has no valid block structure: no reordering, reported as
    // a rejected swap (the caller then discards the AED)
    for (int i = 1 + tid; i + 1 < jw; i += nt)
        if (bsz[i] == 0 && bsz[i + 1] == 0) s_fail = 1;
    __syncthreads();
    for (int round = 0; round < (s_fail ? 0 : 4 * jw + 8); ++round) {
        if (tid == 0) {
            // ---- test phase: examine blocks at the bottom while they are unexamined ----
            while (s_kb >= 0 && s_knt < jw) {
                const int kb = s_kb;
                const int b0 = (bsz[kb] == 0) ? kb - 1 : kb;
                if (st[b0] != ST_U) break;
                const int bulge = (b0 != kb);
                bool dfl;
                if (!bulge) {
                    s_knt += 1;
                    double tmp = fabs(s) + fabs(Hw[GQZ_IX(kb, kb, ld)]);
                    dfl = fabs(s * U[GQZ_IX(0, kb, ld)]) <= fmax(GQZ_ULP * tmp, smlnum);
                } else {
                    s_knt += 2;
                    double h00 = Hw[GQZ_IX(kb - 1, kb - 1, ld)], h10 = Hw[GQZ_IX(kb, kb - 1, ld)];
                    double h01 = Hw[GQZ_IX(kb - 1, kb, ld)], h11 = Hw[GQZ_IX(kb, kb, ld)];
                    double rtdet;
                    if (h00 == 0.0 && h10 == 0.0) rtdet = 0.0;
                    else if (fabs(h00) >= fabs(h10)) rtdet = sqrt(fabs(h00)) * sqrt(fabs(h11 - (h10 / h00) * h01));
                    else rtdet = sqrt(fabs(h10)) * sqrt(fabs(h01 - (h00 / h10) * h11));
                    double tmp = fabs(s) + rtdet;
                    dfl = fmax(fabs(s * U[GQZ_IX(0, kb, ld)]), fabs(s * U[GQZ_IX(0, kb - 1, ld)])) <=
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
            s_np = 0;
        }
        __syncthreads();
        // ---- pair detection: a moving block whose upper neighbour is unexamined ----
        const int kbnow = s_kb;
        for (int p = tid; p <= kbnow; p += nt) {
            if (bsz[p] == 0 || st[p] != ST_X || p <= tgt[p] || p == 0) continue;
            const int q = (bsz[p - 1] != 0) ? p - 1 : p - 2;
            if (q < 0 || st[q] != ST_U) continue;
            const int k = atomicAdd(&s_np, 1);
            pj1[k] = (short)q; pn1[k] = bsz[q]; pn2[k] = bsz[p];
        }
        __syncthreads();
        const int np = s_np;
        if (np == 0) break;
        // ---- compute the swaps (one lane per pair, register resident) ----
        for (int q = tid; q < np; q += nt) {
            const int j1 = pj1[q], n1 = pn1[q], n2 = pn2[q];
            double *rec = pscr + (size_t)PREC * q;
            if (n1 == 1 && n2 == 1) pair_swap<1, 1>(Hw, Tw, ld, j1, rec);
            else if (n1 == 1 && n2 == 2) pair_swap<1, 2>(Hw, Tw, ld, j1, rec);
            else if (n1 == 2 && n2 == 1) pair_swap<2, 1>(Hw, Tw, ld, j1, rec);
            else pair_swap<2, 2>(Hw, Tw, ld, j1, rec);
        }
        __syncthreads();
        // ---- rows of every pair (cols right of its block) and U columns; the (pair, item) index
        //      space is flattened so that all lanes work on independent items ----
        const int d1 = 3 * jw;
        const double inv1 = 1.0 / d1;
        for (int idx = tid; idx < np * d1; idx += nt) {
            const int q = (int)((idx + 0.5) * inv1), r = idx - q * d1;
            const double *rec = pscr + (size_t)PREC * q;
            if (rec[64] != 0.0) continue;
            const int j1 = pj1[q], m = pn1[q] + pn2[q];
            const double *Qa = rec;
            double x[4], y[4];
            if (r < 2 * jw) {
                double *M = r < jw ? Hw : Tw;
                const int col = r < jw ? r : r - jw;
                if (col < j1 + m) continue;
                for (int i = 0; i < m; ++i) x[i] = M[GQZ_IX(j1 + i, col, ld)];
                for (int i = 0; i < m; ++i) { double a = 0; for (int k = 0; k < m; ++k) a += Qa[k + m * i] * x[k]; y[i] = a; }
                for (int i = 0; i < m; ++i) M[GQZ_IX(j1 + i, col, ld)] = y[i];
            } else {
                const int row = r - 2 * jw;
                for (int j = 0; j < m; ++j) x[j] = U[GQZ_IX(row, j1 + j, ld)];
                for (int j = 0; j < m; ++j) { double a = 0; for (int k = 0; k < m; ++k) a += x[k] * Qa[k + m * j]; y[j] = a; }
                for (int j = 0; j < m; ++j) U[GQZ_IX(row, j1 + j, ld)] = y[j];
            }
        }
        __syncthreads();
        // ---- columns of every pair (rows above its block), V, and the new diagonal blocks ----
        const int d2 = 3 * jw + 16;
        const double inv2 = 1.0 / d2;
        for (int idx = tid; idx < np * d2; idx += nt) {
            const int q = (int)((idx + 0.5) * inv2), r = idx - q * d2;
            const double *rec = pscr + (size_t)PREC * q;
            if (rec[64] != 0.0) continue;
            const int j1 = pj1[q], m = pn1[q] + pn2[q];
            const double *Za = rec + 16;
            double x[4], y[4];
            if (r >= 3 * jw) {
                const int e = r - 3 * jw;
                if (e < m * m) {
                    const int i = e % m, j = e / m;
                    Hw[GQZ_IX(j1 + i, j1 + j, ld)] = rec[32 + e];
                    Tw[GQZ_IX(j1 + i, j1 + j, ld)] = rec[48 + e];
                }
                continue;
            }
            double *M; int row;
            if (r < jw) { M = Hw; row = r; if (row >= j1) continue; }
            else if (r < 2 * jw) { M = Tw; row = r - jw; if (row >= j1) continue; }
            else { M = V; row = r - 2 * jw; }
            for (int j = 0; j < m; ++j) x[j] = M[GQZ_IX(row, j1 + j, ld)];
            for (int j = 0; j < m; ++j) { double a = 0; for (int k = 0; k < m; ++k) a += x[k] * Za[k + m * j]; y[j] = a; }
            for (int j = 0; j < m; ++j) M[GQZ_IX(row, j1 + j, ld)] = y[j];
        }
        __syncthreads();
        // ---- block-structure update (one lane per pair; pairs are disjoint) ----
        for (int q = tid; q < np; q += nt) {
            const double *rec = pscr + (size_t)PREC * q;
            if (rec[64] != 0.0) { s_fail = 2; continue; }
            const int j1 = pj1[q], n1 = pn1[q], n2 = pn2[q];
            const int t = tgt[j1 + n1];
            atomicAdd(&s_swaps, 1);
            // the moving block now starts at j1 (order n2), the passed block at j1 + n2 (order n1);
            // a 2x2 block whose eigenvalues turned real after the swap splits into two 1x1 blocks
            const bool msplit = (n2 == 2) && Hw[GQZ_IX(j1 + 1, j1, ld)] == 0.0;
            const bool psplit = (n1 == 2) && Hw[GQZ_IX(j1 + n2 + 1, j1 + n2, ld)] == 0.0;
            if (msplit) {
                bsz[j1] = 1; bsz[j1 + 1] = 1;
                st[j1] = (j1 == t) ? ST_A : ST_X; tgt[j1] = (short)t;
                st[j1 + 1] = (j1 + 1 == t + 1) ? ST_A : ST_X; tgt[j1 + 1] = (short)(t + 1);
            } else {
                bsz[j1] = (unsigned char)n2;
                if (n2 == 2) bsz[j1 + 1] = 0;
                st[j1] = (j1 == t) ? ST_A : ST_X; tgt[j1] = (short)t;
            }
            const int pb = j1 + n2;
            if (n1 == 2 && !psplit) { bsz[pb] = 2; bsz[pb + 1] = 0; st[pb] = ST_U; }
            else if (n1 == 2) { bsz[pb] = 1; bsz[pb + 1] = 1; st[pb] = ST_U; st[pb + 1] = ST_U; }
            else { bsz[pb] = 1; st[pb] = ST_U; }
        }
        __syncthreads();
        if (s_fail) break;
    }
    if (tid == 0) {
        int kbc = s_kb + 1;
        if (s_fail != 1 && kbc < s_ilst) kbc = s_ilst;
        out[0] = kbc; out[1] = s_swaps; out[2] = s_fail; out[3] = (s != 0.0);
    }
    __syncthreads();
}

extern "C" __global__ void __launch_bounds__(512) k_aed_reorder(double *Hw, double *Tw, double *U, double *V,
                                                                 int jw, const double *sptr, double smlnum,
                                                                 double *pscr, int *out) {
    aed_reorder_body<AED_MAXW>(Hw, Tw, U, V, jw, jw, sptr ? *sptr : 0.0, smlnum, pscr, out);
}

// ---------------------------------------------------------------------------------------------
// k_aed_reorder_coop: the same train on a large window with the row/column updates spread over
// a cooperative grid.  Every workgroup keeps its own copy of the block structure and of every
// pair's swap factors in LDS and recomputes them identically (same inputs after each grid
// barrier), so a round needs only two grid barriers: one after the row updates and one after the
// column updates.  Pairs are numbered by position (a scan, not atomics) so that all workgroups
// agree on the numbering.  The new diagonal blocks travel through pscr, written by the pair's
// owner workgroup.
// ---------------------------------------------------------------------------------------------
#define COOP_MAXW 1024
#define COOP_NT 512

extern "C" __global__ void __launch_bounds__(COOP_NT) k_aed_reorder_coop(double *Hw, double *Tw, double *U,
                                                                         double *V, int jw, const double *sptr,
                                                                         double smlnum, double *pscr, int *out) {
    namespace cg = cooperative_groups;
    __shared__ unsigned char st[COOP_MAXW], bsz[COOP_MAXW];
    __shared__ short tgt[COOP_MAXW];
    __shared__ short pj1[COOP_MAXW / 2];
    __shared__ unsigned char pn1[COOP_MAXW / 2], pn2[COOP_MAXW / 2], pfl[COOP_MAXW / 2];
    __shared__ double pq[COOP_MAXW / 2][32];          // Qa (m x m, ld m) | Za
    __shared__ int scan[COOP_NT];
    __shared__ int s_np, s_kb, s_ilst, s_knt, s_fail, s_swaps;
    cg::grid_group grid = cg::this_grid();
    const int tid = threadIdx.x, nt = blockDim.x;
    const int gtid = blockIdx.x * nt + tid, gnt = gridDim.x * nt;
    const double s = sptr ? *sptr : 0.0;
    const int ld = jw;
    for (int i = tid; i < jw; i += nt) {
        st[i] = ST_U; tgt[i] = 0;
        const bool lower = (i > 0) && Hw[GQZ_IX(i, i - 1, ld)] != 0.0;
        const bool upper = (i + 1 < jw) && Hw[GQZ_IX(i + 1, i, ld)] != 0.0;
        bsz[i] = lower ? 0 : (upper ? 2 : 1);
    }
    if (tid == 0) { s_kb = jw - 1; s_ilst = 0; s_knt = 0; s_fail = 0; s_swaps = 0; s_np = 0; }
    __syncthreads();
    // not quasi-triangular: no reordering (every workgroup sees the same window)
    for (int i = 1 + tid; i + 1 < jw; i += nt)
        if (bsz[i] == 0 && bsz[i + 1] == 0) s_fail = 1;
    __syncthreads();
    for (int round = 0; round < (s_fail ? 0 : 4 * jw + 8); ++round) {
        if (tid == 0) {
            while (s_kb >= 0 && s_knt < jw) {
                const int kb = s_kb;
                const int b0 = (bsz[kb] == 0) ? kb - 1 : kb;
                if (st[b0] != ST_U) break;
                const int bulge = (b0 != kb);
                bool dfl;
                if (!bulge) {
                    s_knt += 1;
                    double tmp = fabs(s) + fabs(Hw[GQZ_IX(kb, kb, ld)]);
                    dfl = fabs(s * U[GQZ_IX(0, kb, ld)]) <= fmax(GQZ_ULP * tmp, smlnum);
                } else {
                    s_knt += 2;
                    double h00 = Hw[GQZ_IX(kb - 1, kb - 1, ld)], h10 = Hw[GQZ_IX(kb, kb - 1, ld)];
                    double h01 = Hw[GQZ_IX(kb - 1, kb, ld)], h11 = Hw[GQZ_IX(kb, kb, ld)];
                    double rtdet;
                    if (h00 == 0.0 && h10 == 0.0) rtdet = 0.0;
                    else if (fabs(h00) >= fabs(h10)) rtdet = sqrt(fabs(h00)) * sqrt(fabs(h11 - (h10 / h00) * h01));
                    else rtdet = sqrt(fabs(h10)) * sqrt(fabs(h01 - (h00 / h10) * h11));
                    double tmp = fabs(s) + rtdet;
                    dfl = fmax(fabs(s * U[GQZ_IX(0, kb, ld)]), fabs(s * U[GQZ_IX(0, kb - 1, ld)])) <=
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
        }
        __syncthreads();
        // ---- pair detection, numbered by position: per-lane chunk counts + exclusive scan ----
        const int kbnow = s_kb;
        const int npos = kbnow + 1;
        const int chunk = (npos + nt - 1) / nt;
        const int p0 = tid * chunk, p1 = min(npos, p0 + chunk);
        int cnt = 0;
        for (int p = p0; p < p1; ++p) {
            if (bsz[p] == 0 || st[p] != ST_X || p <= tgt[p] || p == 0) continue;
            const int q = (bsz[p - 1] != 0) ? p - 1 : p - 2;
            if (q >= 0 && st[q] == ST_U) ++cnt;
        }
        scan[tid] = cnt;
        __syncthreads();
        for (int off = 1; off < nt; off <<= 1) {
            const int v = (tid >= off) ? scan[tid - off] : 0;
            __syncthreads();
            scan[tid] += v;
            __syncthreads();
        }
        int k = scan[tid] - cnt;
        for (int p = p0; p < p1; ++p) {
            if (bsz[p] == 0 || st[p] != ST_X || p <= tgt[p] || p == 0) continue;
            const int q = (bsz[p - 1] != 0) ? p - 1 : p - 2;
            if (q < 0 || st[q] != ST_U) continue;
            pj1[k] = (short)q; pn1[k] = bsz[q]; pn2[k] = bsz[p]; ++k;
        }
        if (tid == nt - 1) s_np = scan[nt - 1];
        __syncthreads();
        const int np = s_np;
        if (np == 0) break;
        // ---- swap factors of every pair (every workgroup, identically) ----
        for (int q = tid; q < np; q += nt) {
            const int j1 = pj1[q], n1 = pn1[q], n2 = pn2[q];
            double rec[PREC];
            if (n1 == 1 && n2 == 1) pair_swap<1, 1>(Hw, Tw, ld, j1, rec);
            else if (n1 == 1 && n2 == 2) pair_swap<1, 2>(Hw, Tw, ld, j1, rec);
            else if (n1 == 2 && n2 == 1) pair_swap<2, 1>(Hw, Tw, ld, j1, rec);
            else pair_swap<2, 2>(Hw, Tw, ld, j1, rec);
            const int m = n1 + n2;
            pfl[q] = rec[64] != 0.0;
            if (!pfl[q]) {
                for (int e = 0; e < m * m; ++e) { pq[q][e] = rec[e]; pq[q][16 + e] = rec[16 + e]; }
            }
            if (q % gridDim.x == blockIdx.x) {
                double *g = pscr + (size_t)PREC * q;
                g[64] = rec[64];
                if (!pfl[q]) for (int e = 0; e < m * m; ++e) { g[32 + e] = rec[32 + e]; g[48 + e] = rec[48 + e]; }
            }
        }
        __syncthreads();
        grid.sync();
        // ---- rows of every pair (cols right of its block) and U columns ----
        const int d1 = 3 * jw;
        const double inv1 = 1.0 / d1;
        for (int idx = gtid; idx < np * d1; idx += gnt) {
            const int q = (int)((idx + 0.5) * inv1), r = idx - q * d1;
            if (pfl[q]) continue;
            const int j1 = pj1[q], m = pn1[q] + pn2[q];
            const double *Qa = pq[q];
            double x[4], y[4];
            if (r < 2 * jw) {
                double *M = r < jw ? Hw : Tw;
                const int col = r < jw ? r : r - jw;
                if (col < j1 + m) continue;
                for (int i = 0; i < m; ++i) x[i] = M[GQZ_IX(j1 + i, col, ld)];
                for (int i = 0; i < m; ++i) { double a = 0; for (int kk = 0; kk < m; ++kk) a += Qa[kk + m * i] * x[kk]; y[i] = a; }
                for (int i = 0; i < m; ++i) M[GQZ_IX(j1 + i, col, ld)] = y[i];
            } else {
                const int row = r - 2 * jw;
                for (int j = 0; j < m; ++j) x[j] = U[GQZ_IX(row, j1 + j, ld)];
                for (int j = 0; j < m; ++j) { double a = 0; for (int kk = 0; kk < m; ++kk) a += x[kk] * Qa[kk + m * j]; y[j] = a; }
                for (int j = 0; j < m; ++j) U[GQZ_IX(row, j1 + j, ld)] = y[j];
            }
        }
        grid.sync();
        // ---- columns of every pair (rows above its block), V, and the new diagonal blocks ----
        const int d2 = 3 * jw + 16;
        const double inv2 = 1.0 / d2;
        for (int idx = gtid; idx < np * d2; idx += gnt) {
            const int q = (int)((idx + 0.5) * inv2), r = idx - q * d2;
            if (pfl[q]) continue;
            const int j1 = pj1[q], m = pn1[q] + pn2[q];
            const double *Za = pq[q] + 16;
            double x[4], y[4];
            if (r >= 3 * jw) {
                const int e = r - 3 * jw;
                if (e < m * m) {
                    const double *g = pscr + (size_t)PREC * q;
                    const int i = e % m, j = e / m;
                    Hw[GQZ_IX(j1 + i, j1 + j, ld)] = g[32 + e];
                    Tw[GQZ_IX(j1 + i, j1 + j, ld)] = g[48 + e];
                }
                continue;
            }
            double *M; int row;
            if (r < jw) { M = Hw; row = r; if (row >= j1) continue; }
            else if (r < 2 * jw) { M = Tw; row = r - jw; if (row >= j1) continue; }
            else { M = V; row = r - 2 * jw; }
            for (int j = 0; j < m; ++j) x[j] = M[GQZ_IX(row, j1 + j, ld)];
            for (int j = 0; j < m; ++j) { double a = 0; for (int kk = 0; kk < m; ++kk) a += x[kk] * Za[kk + m * j]; y[j] = a; }
            for (int j = 0; j < m; ++j) M[GQZ_IX(row, j1 + j, ld)] = y[j];
        }
        grid.sync();
        // ---- block-structure update (every workgroup, one lane per pair) ----
        for (int q = tid; q < np; q += nt) {
            if (pfl[q]) { s_fail = 2; continue; }
            const int j1 = pj1[q], n1 = pn1[q], n2 = pn2[q];
            const int t = tgt[j1 + n1];
            atomicAdd(&s_swaps, 1);
            const bool msplit = (n2 == 2) && Hw[GQZ_IX(j1 + 1, j1, ld)] == 0.0;
            const bool psplit = (n1 == 2) && Hw[GQZ_IX(j1 + n2 + 1, j1 + n2, ld)] == 0.0;
            if (msplit) {
                bsz[j1] = 1; bsz[j1 + 1] = 1;
                st[j1] = (j1 == t) ? ST_A : ST_X; tgt[j1] = (short)t;
                st[j1 + 1] = (j1 + 1 == t + 1) ? ST_A : ST_X; tgt[j1 + 1] = (short)(t + 1);
            } else {
                bsz[j1] = (unsigned char)n2;
                if (n2 == 2) bsz[j1 + 1] = 0;
                st[j1] = (j1 == t) ? ST_A : ST_X; tgt[j1] = (short)t;
            }
            const int pb = j1 + n2;
            if (n1 == 2 && !psplit) { bsz[pb] = 2; bsz[pb + 1] = 0; st[pb] = ST_U; }
            else if (n1 == 2) { bsz[pb] = 1; bsz[pb + 1] = 1; st[pb] = ST_U; st[pb + 1] = ST_U; }
            else { bsz[pb] = 1; st[pb] = ST_U; }
        }
        __syncthreads();
        if (s_fail) break;
    }
    if (blockIdx.x == 0 && tid == 0) {
        int kbc = s_kb + 1;
        if (s_fail != 1 && kbc < s_ilst) kbc = s_ilst;
        out[0] = kbc; out[1] = s_swaps; out[2] = s_fail; out[3] = (s != 0.0);
    }
}

// grid of the cooperative reorder for a window of order jw (0: use the single-workgroup kernel)
int gqz_aed_reorder_coop_grid(int jw) {
    if (jw < 160 || jw > COOP_MAXW) return 0;
    return std::min(128, std::max(16, jw / 6));
}

// eigenvalues of the window's diagonal blocks (the shifts and the deflated eigenvalues)
__device__ void aed_eigs_body(const double *Hw, const double *Tw, int ld, int jw, double *war, double *wai,
                              double *wb, int j0, int stride) {
    for (int j = j0; j < jw; j += stride) {
        const bool lower2 = (j > 0) && Hw[GQZ_IX(j, j - 1, ld)] != 0.0;
        const bool upper2 = (j + 1 < jw) && Hw[GQZ_IX(j + 1, j, ld)] != 0.0;
        if (lower2) continue;
        if (upper2) {
            double A[4] = { Hw[GQZ_IX(j, j, ld)], Hw[GQZ_IX(j + 1, j, ld)], Hw[GQZ_IX(j, j + 1, ld)], Hw[GQZ_IX(j + 1, j + 1, ld)] };
            double B[4] = { Tw[GQZ_IX(j, j, ld)], 0.0, Tw[GQZ_IX(j, j + 1, ld)], Tw[GQZ_IX(j + 1, j + 1, ld)] };
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
            war[j] = Hw[GQZ_IX(j, j, ld)]; wai[j] = 0.0; wb[j] = Tw[GQZ_IX(j, j, ld)];
        }
    }
}

extern "C" __global__ void __launch_bounds__(256) k_aed_eigs(const double *Hw, const double *Tw, int jw,
                                                             double *war, double *wai, double *wb) {
    aed_eigs_body(Hw, Tw, jw, jw, war, wai, wb, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Spike reflector and RQ update (qzearly.f:372-436 with PARA_AGGRQUP = .TRUE., kkqzconf.f:138):
//   P = I - tau v v^T maps the spike s U(0, 0:kbc) onto a multiple of e_0 and is applied to the
//   rows 0..kbc-1 of Hw (all columns) and T's columns right of the block, and to U's columns.
//   P T(0:kbc, 0:kbc) = T + v w^T with w = -tau T^T v is brought back to triangular form by
//   2 kbc - 2 column rotations: the first sweep (i = 0 .. kbc-2) folds w into its last entry
//   (T becomes Hessenberg, the rank-one term lives in the last column), the second sweep
//   (i = kbc-2 .. 0) removes the subdiagonal.  Hw's block columns and V see the same rotations.
// One workgroup; kbc <= AED_MAXW.  Rotations are kept in scr (4 kbc doubles).
// sv, sw: kbc doubles each, red: blockDim.x doubles (LDS); scr: 4 kbc doubles
__device__ void aed_spike_body(double *Hw, double *Tw, double *U, double *V, int ld, int jw, double s, int kbc,
                               double *scr, double *sv, double *sw, double *red) {
    __shared__ double s_tau, s_cs, s_sn;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (kbc <= 0 || s == 0.0 || kbc == jw) return;
    // ---- reflector (DLARFG on s U(0, 0:kbc)) ----
    for (int i = tid; i < kbc; i += nt) sv[i] = s * U[GQZ_IX(0, i, ld)];
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
            double *p = (idx < jw) ? Hw + GQZ_IX(0, idx, ld) : Tw + GQZ_IX(0, kbc + idx - jw, ld);
            double t = 0.0;
            for (int i = 0; i < kbc; ++i) t += sv[i] * p[i];
            t *= tau;
            for (int i = 0; i < kbc; ++i) p[i] -= t * sv[i];
        } else if (idx < 3 * jw - kbc) {
            const int row = idx - (2 * jw - kbc);
            double t = 0.0;
            for (int j = 0; j < kbc; ++j) t += U[GQZ_IX(row, j, ld)] * sv[j];
            t *= tau;
            for (int j = 0; j < kbc; ++j) U[GQZ_IX(row, j, ld)] -= t * sv[j];
        } else {
            const int j = idx - (3 * jw - kbc);
            double t = 0.0;
            for (int i = 0; i <= j; ++i) t += Tw[GQZ_IX(i, j, ld)] * sv[i];
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
        double x = Tw[GQZ_IX(row, i0, ld)];
        for (int i = i0; i < kbc - 1; ++i) {
            double y = Tw[GQZ_IX(row, i + 1, ld)];
            Tw[GQZ_IX(row, i, ld)] = rc[i] * x + rs[i] * y;
            x = rc[i] * y - rs[i] * x;
        }
        Tw[GQZ_IX(row, kbc - 1, ld)] = x;
    }
    __syncthreads();
    // rank-one term now lives in the last column: T(:, kbc-1) += w(kbc-1) v
    for (int i = tid; i < kbc; i += nt) Tw[GQZ_IX(i, kbc - 1, ld)] += sw[kbc - 1] * sv[i];
    // H block columns (rows 0..kbc-1; H below row kbc is zero in these columns) and V (all rows)
    for (int row = tid; row < kbc + jw; row += nt) {
        double *M = row < kbc ? Hw : V;
        const int r = row < kbc ? row : row - kbc;
        double x = M[GQZ_IX(r, 0, ld)];
        for (int i = 0; i < kbc - 1; ++i) {
            double y = M[GQZ_IX(r, i + 1, ld)];
            M[GQZ_IX(r, i, ld)] = rc[i] * x + rs[i] * y;
            x = rc[i] * y - rs[i] * x;
        }
        M[GQZ_IX(r, kbc - 1, ld)] = x;
    }
    __syncthreads();
    // ---- second sweep: i = kbc-2 .. 0 removes T(i+1, i) (qzearly.f:421-429) ----
    for (int i = kbc - 2; i >= 0; --i) {
        if (tid == 0) {
            double c, sn, r;
            d_lartg(Tw[GQZ_IX(i + 1, i + 1, ld)], Tw[GQZ_IX(i + 1, i, ld)], c, sn, r);
            Tw[GQZ_IX(i + 1, i + 1, ld)] = r;
            Tw[GQZ_IX(i + 1, i, ld)] = 0.0;
            qc[i] = c; qs[i] = -sn;
            s_cs = c; s_sn = -sn;
        }
        __syncthreads();
        const double c = s_cs, sn = s_sn;
        for (int row = tid; row <= i; row += nt) {
            double x = Tw[GQZ_IX(row, i, ld)], y = Tw[GQZ_IX(row, i + 1, ld)];
            Tw[GQZ_IX(row, i, ld)] = c * x + sn * y;
            Tw[GQZ_IX(row, i + 1, ld)] = c * y - sn * x;
        }
        __syncthreads();
    }
    for (int row = tid; row < kbc + jw; row += nt) {
        double *M = row < kbc ? Hw : V;
        const int r = row < kbc ? row : row - kbc;
        double y = M[GQZ_IX(r, kbc - 1, ld)];
        for (int i = kbc - 2; i >= 0; --i) {
            double x = M[GQZ_IX(r, i, ld)];
            M[GQZ_IX(r, i + 1, ld)] = qc[i] * y - qs[i] * x;
            y = qc[i] * x + qs[i] * y;
        }
        M[GQZ_IX(r, 0, ld)] = y;
    }
    __syncthreads();
}

extern "C" __global__ void __launch_bounds__(1024) k_aed_spike(double *Hw, double *Tw, double *U, double *V,
                                                               int jw, const double *sptr, double *scr,
                                                               const int *out) {
    __shared__ double sv[AED_MAXW], sw[AED_MAXW];
    __shared__ double red[1024];
    aed_spike_body(Hw, Tw, U, V, jw, jw, sptr ? *sptr : 0.0, out[0], scr, sv, sw, red);
}

// ---------------------------------------------------------------------------------------------
// k_aed_small: the whole AED of a window of order jw <= AEDS_MAXW in one workgroup, with the
// window and its U, V resident in LDS (QZEARLY, qzearly.f:1-455): window QZ (double-shift base
// QZ), deflation test + reordering (the train above), eigenvalues, and - when something deflated
// or there is no spike (PDHGEQZ3.f:840) - the spike reflector with the RQ update, the HT restore
// of the undeflatable block and the write-back of the window and of the new spike entry.
//   out[0] = kbc, out[1] = swaps, out[2] = reorder status (0 done, 1 window not quasi-triangular:
//   nothing written back, 2 a swap rejected: deflations so far kept), out[3] = spike present,
//   out[4] = INFO of the window QZ (non-zero: nothing was written back).
// U, V (jw x jw, ld jw) receive the window's orthogonal factors for the off-window updates.
// ---------------------------------------------------------------------------------------------
#define AEDS_MAXW 64
#define AEDS_NT 256

size_t gqz_aed_small_lds(int jw) {
    const int ld = jw + 1;
    return sizeof(double) * (4 * (size_t)ld * jw + 2 * (size_t)jw + AEDS_NT) + sizeof(WgCtl) + 16;
}
int gqz_aed_small_max() { return AEDS_MAXW; }

extern "C" __global__ void __launch_bounds__(AEDS_NT) k_aed_small(double *H, double *T, int ldh, int kwtop, int jw,
                                                                  int has_spike, double ttol, double smlnum,
                                                                  double *U, double *V, double *war, double *wai,
                                                                  double *wb, double *pscr, double *scr, int *out) {
    extern __shared__ double lds[];
    __shared__ double sh[8];
    __shared__ int res[4];
    const int ld = jw + 1;
    double *sH = lds, *sT = sH + (size_t)ld * jw, *sU = sT + (size_t)ld * jw, *sV = sU + (size_t)ld * jw;
    double *sv = sV + (size_t)ld * jw, *sw = sv + jw, *red = sw + jw;
    WgCtl *C = (WgCtl *)(red + AEDS_NT);
    const int tid = threadIdx.x, nt = blockDim.x;
    const double s = has_spike ? H[GQZ_IX(kwtop, kwtop - 1, ldh)] : 0.0;
    for (int idx = tid; idx < jw * jw; idx += nt) {
        const int i = idx % jw, j = idx / jw;
        sH[i + ld * j] = (i <= j + 1) ? H[GQZ_IX(kwtop + i, kwtop + j, ldh)] : 0.0;
        sT[i + ld * j] = (i <= j) ? T[GQZ_IX(kwtop + i, kwtop + j, ldh)] : 0.0;
        sU[i + ld * j] = (i == j) ? 1.0 : 0.0;
        sV[i + ld * j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    WgPencil P;
    P.H = sH; P.T = sT; P.U = sU; P.V = sV; P.ld = ld; P.ldu = ld; P.n = jw; P.nq = jw;
    const int info = wg_qz_ds(P, C, 0, jw - 1, war, wai, wb, ttol, smlnum);
    if (info != 0) {
        if (tid == 0) { out[0] = jw; out[1] = 0; out[2] = 0; out[3] = (s != 0.0); out[4] = info; }
        return;
    }
    aed_reorder_body<AEDS_MAXW>(sH, sT, sU, sV, ld, jw, s, smlnum, pscr, res);
    aed_eigs_body(sH, sT, ld, jw, war, wai, wb, tid, nt);
    // a window that is not quasi-triangular (res[2] == 1) is discarded (nothing written back, no
    // deflation); after a rejected swap (res[2] == 2) the deflations found so far are kept
    const int kbc = res[2] == 1 ? jw : res[0];
    if (res[2] != 1 && (kbc < jw || s == 0.0)) {
        if (kbc > 0 && s != 0.0) aed_spike_body(sH, sT, sU, sV, ld, jw, s, kbc, scr, sv, sw, red);
        if (kbc > 2 && s != 0.0) wg_gghrd(sH, sT, ld, jw, 0, kbc - 1, sU, sV, ld, jw, sh);
        __syncthreads();
        for (int idx = tid; idx < jw * jw; idx += nt) {
            const int i = idx % jw, j = idx / jw;
            H[GQZ_IX(kwtop + i, kwtop + j, ldh)] = sH[i + ld * j];
            T[GQZ_IX(kwtop + i, kwtop + j, ldh)] = sT[i + ld * j];
            U[i + (size_t)jw * j] = sU[i + ld * j];
            V[i + (size_t)jw * j] = sV[i + ld * j];
        }
        // the spike after its reflector: s U(0,0) on the first row; dropped when the whole window
        // deflated (kbc = 0; PDHGEQZ3.f:831, :1346): a tiny non-zero entry here would join two
        // converged blocks across the deflation boundary for good
        if (tid == 0 && s != 0.0) H[GQZ_IX(kwtop, kwtop - 1, ldh)] = kbc > 0 ? s * sU[0] : 0.0;
    }
    if (tid == 0) { out[0] = kbc; out[1] = res[1]; out[2] = res[2]; out[3] = res[2] == 1 ? 1 : res[3]; out[4] = 0; }
}

void gqz_launch_aed_small(hipStream_t st, double *H, double *T, int ldh, int kwtop, int jw, int has_spike,
                          double ttol, double smlnum, double *U, double *V, double *war, double *wai, double *wb,
                          double *pscr, double *scr, int *out) {
    const size_t lds = gqz_aed_small_lds(jw);
    static size_t lds_set[64] = {};
    gqz_need_lds((const void *)k_aed_small, lds, lds_set);
    hipLaunchKernelGGL(k_aed_small, dim3(1), dim3(AEDS_NT), lds, st, H, T, ldh, kwtop, jw, has_spike, ttol, smlnum,
                       U, V, war, wai, wb, pscr, scr, out);
}

void gqz_launch_aed_reorder(hipStream_t st, double *Hw, double *Tw, double *U, double *V, int jw,
                            const double *sptr, double smlnum, double *pscr, int *out) {
    const int g = gqz_aed_reorder_coop_grid(jw);
    if (g > 0) {
        void *args[] = { &Hw, &Tw, &U, &V, &jw, (void *)&sptr, &smlnum, &pscr, &out };
        GQZ_HIP(hipLaunchCooperativeKernel((const void *)k_aed_reorder_coop, dim3(g), dim3(COOP_NT), args, 0, st));
        return;
    }
    hipLaunchKernelGGL(k_aed_reorder, dim3(1), dim3(512), 0, st, Hw, Tw, U, V, jw, sptr, smlnum, pscr, out);
}
void gqz_launch_aed_eigs(hipStream_t st, const double *Hw, const double *Tw, int jw, double *war, double *wai,
                         double *wb) {
    hipLaunchKernelGGL(k_aed_eigs, dim3((jw + 255) / 256), dim3(256), 0, st, Hw, Tw, jw, war, wai, wb);
}
void gqz_launch_aed_spike(hipStream_t st, double *Hw, double *Tw, double *U, double *V, int jw,
                          const double *sptr, double *scr, const int *out) {
    hipLaunchKernelGGL(k_aed_spike, dim3(1), dim3(1024), 0, st, Hw, Tw, U, V, jw, sptr, scr, out);
}

// ---------------------------------------------------------------------------------------------
// k_tgord_window: eigenvalue reordering inside one LDS-resident diagonal window (the window step of
// PDTGORD, src/reorder/PDTGORD.f:600-1600, with its local DTGEXC/DTGEX2 swaps).  The window
// [p, p + w) of (S, T), w <= 64, is loaded with U = V = I; selected blocks (bit r of `sel` for window
// row r; a 2x2 block is selected when its first row is) move to the top of the window: every round
// swaps all pairs (unselected block directly above a selected block) at once - their index sets are
// disjoint, rows are updated before columns - until no such pair is left.  The window and U, V
// (w x w, ld w) are written back; the parts of the pencil outside the window are updated by the
// caller with U (rows right of the window, Q) and V (columns above it, Z).
//   out[0] = 0, or 1 + the window row of the pair whose swap was rejected (the window is then
//            partially reordered and consistent: the rejected pair wrote nothing);
//   out[1], out[2] = selection mask after the reordering (low, high 32 bits);
//   out[3], out[4] = mask of the rows r with S(r + 1, r) != 0 (2x2 blocks) after the reordering;
//   out[5] = number of swaps.
// ---------------------------------------------------------------------------------------------
#define TGO_MAXW 64
#define TGO_NT 256

extern "C" __global__ void __launch_bounds__(TGO_NT) k_tgord_window(double *S, double *T, int ld, int p, int w,
                                                                    unsigned long long sel, double *U, double *V,
                                                                    int *out) {
    constexpr int LDW = TGO_MAXW + 1;
    __shared__ double sS[TGO_MAXW * LDW], sT[TGO_MAXW * LDW], sU[TGO_MAXW * LDW], sV[TGO_MAXW * LDW];
    __shared__ double rec[TGO_MAXW / 2][PREC];
    __shared__ unsigned char bsz[TGO_MAXW], sl[TGO_MAXW];
    __shared__ short pj1[TGO_MAXW / 2];
    __shared__ unsigned char pn1[TGO_MAXW / 2], pn2[TGO_MAXW / 2];
    __shared__ int s_np, s_fail, s_swaps;
    const int tid = threadIdx.x, nt = blockDim.x;
    double *gS = S + GQZ_IX(p, p, ld), *gT = T + GQZ_IX(p, p, ld);
    for (int idx = tid; idx < w * w; idx += nt) {
        const int i = idx % w, j = idx / w;
        sS[i + LDW * j] = gS[GQZ_IX(i, j, ld)];
        sT[i + LDW * j] = gT[GQZ_IX(i, j, ld)];
        sU[i + LDW * j] = (i == j) ? 1.0 : 0.0;
        sV[i + LDW * j] = (i == j) ? 1.0 : 0.0;
    }
    if (tid == 0) { s_fail = 0; s_swaps = 0; }
    __syncthreads();
    for (int i = tid; i < w; i += nt) {
        const bool lower = (i > 0) && sS[i + LDW * (i - 1)] != 0.0;
        const bool upper = (i + 1 < w) && sS[(i + 1) + LDW * i] != 0.0;
        bsz[i] = lower ? 0 : (upper ? 2 : 1);
        const int b0 = lower ? i - 1 : i;
        // a 2x2 block is selected when either of its rows is (PDTGORD.f:512-527)
        const bool s0 = (sel >> b0) & 1ull, s1 = upper || lower ? ((sel >> (b0 + 1)) & 1ull) : false;
        sl[i] = (s0 || s1) ? 1 : 0;
    }
    __syncthreads();
    for (int round = 0; round < 2 * TGO_MAXW + 2; ++round) {
        if (tid == 0) s_np = 0;
        __syncthreads();
        // pairs: a selected block whose upper neighbour block is unselected
        for (int q = tid; q < w; q += nt) {
            if (bsz[q] == 0 || sl[q] || q + bsz[q] >= w) continue;
            const int r = q + bsz[q];
            if (!sl[r]) continue;
            const int k = atomicAdd(&s_np, 1);
            pj1[k] = (short)q; pn1[k] = bsz[q]; pn2[k] = bsz[r];
        }
        __syncthreads();
        const int np = s_np;
        if (np == 0) break;
        for (int q = tid; q < np; q += nt) {
            const int j1 = pj1[q], n1 = pn1[q], n2 = pn2[q];
            if (n1 == 1 && n2 == 1) pair_swap<1, 1>(sS, sT, LDW, j1, rec[q]);
            else if (n1 == 1 && n2 == 2) pair_swap<1, 2>(sS, sT, LDW, j1, rec[q]);
            else if (n1 == 2 && n2 == 1) pair_swap<2, 1>(sS, sT, LDW, j1, rec[q]);
            else pair_swap<2, 2>(sS, sT, LDW, j1, rec[q]);
        }
        __syncthreads();
        // rows of every pair (columns right of its corner) and U columns
        for (int idx = tid; idx < np * 2 * w + np * w; idx += nt) {
            const int q = idx < np * 2 * w ? idx / (2 * w) : (idx - np * 2 * w) / w;
            if (rec[q][64] != 0.0) continue;
            const int j1 = pj1[q], m = pn1[q] + pn2[q];
            const double *Qa = rec[q];
            double x[4], y[4];
            if (idx < np * 2 * w) {
                const int r = idx - q * 2 * w;
                double *M = r < w ? sS : sT;
                const int col = r < w ? r : r - w;
                if (col < j1 + m) continue;
                for (int i = 0; i < m; ++i) x[i] = M[(j1 + i) + LDW * col];
                for (int i = 0; i < m; ++i) { double a = 0; for (int k = 0; k < m; ++k) a += Qa[k + m * i] * x[k]; y[i] = a; }
                for (int i = 0; i < m; ++i) M[(j1 + i) + LDW * col] = y[i];
            } else {
                const int row = idx - np * 2 * w - q * w;
                for (int j = 0; j < m; ++j) x[j] = sU[row + LDW * (j1 + j)];
                for (int j = 0; j < m; ++j) { double a = 0; for (int k = 0; k < m; ++k) a += x[k] * Qa[k + m * j]; y[j] = a; }
                for (int j = 0; j < m; ++j) sU[row + LDW * (j1 + j)] = y[j];
            }
        }
        __syncthreads();
        // columns of every pair (rows above its corner), V, and the new corners
        for (int idx = tid; idx < np * (3 * w + 16); idx += nt) {
            const int q = idx / (3 * w + 16), r = idx - q * (3 * w + 16);
            if (rec[q][64] != 0.0) continue;
            const int j1 = pj1[q], m = pn1[q] + pn2[q];
            const double *Za = rec[q] + 16;
            double x[4], y[4];
            if (r >= 3 * w) {
                const int e = r - 3 * w;
                if (e < m * m) {
                    const int i = e % m, j = e / m;
                    sS[(j1 + i) + LDW * (j1 + j)] = rec[q][32 + e];
                    sT[(j1 + i) + LDW * (j1 + j)] = rec[q][48 + e];
                }
                continue;
            }
            double *M; int row;
            if (r < w) { M = sS; row = r; if (row >= j1) continue; }
            else if (r < 2 * w) { M = sT; row = r - w; if (row >= j1) continue; }
            else { M = sV; row = r - 2 * w; }
            for (int j = 0; j < m; ++j) x[j] = M[row + LDW * (j1 + j)];
            for (int j = 0; j < m; ++j) { double a = 0; for (int k = 0; k < m; ++k) a += x[k] * Za[k + m * j]; y[j] = a; }
            for (int j = 0; j < m; ++j) M[row + LDW * (j1 + j)] = y[j];
        }
        __syncthreads();
        // block structure and selection flags travel with the blocks
        for (int q = tid; q < np; q += nt) {
            if (rec[q][64] != 0.0) { atomicCAS(&s_fail, 0, 1 + pj1[q]); continue; }
            atomicAdd(&s_swaps, 1);
            const int j1 = pj1[q], n1 = pn1[q], n2 = pn2[q];
            // the selected block now starts at j1 (order n2), the passed one at j1 + n2 (order n1); a
            // 2x2 block whose eigenvalues turned real splits into two 1x1 blocks
            const bool msplit = (n2 == 2) && sS[(j1 + 1) + LDW * j1] == 0.0;
            const bool psplit = (n1 == 2) && sS[(j1 + n2 + 1) + LDW * (j1 + n2)] == 0.0;
            if (n2 == 2 && !msplit) { bsz[j1] = 2; bsz[j1 + 1] = 0; }
            else if (n2 == 2) { bsz[j1] = 1; bsz[j1 + 1] = 1; }
            else bsz[j1] = 1;
            for (int i = 0; i < n2; ++i) sl[j1 + i] = 1;
            const int pb = j1 + n2;
            if (n1 == 2 && !psplit) { bsz[pb] = 2; bsz[pb + 1] = 0; }
            else if (n1 == 2) { bsz[pb] = 1; bsz[pb + 1] = 1; }
            else bsz[pb] = 1;
            for (int i = 0; i < n1; ++i) sl[pb + i] = 0;
        }
        __syncthreads();
        if (s_fail) break;
    }
    for (int idx = tid; idx < w * w; idx += nt) {
        const int i = idx % w, j = idx / w;
        gS[GQZ_IX(i, j, ld)] = sS[i + LDW * j];
        gT[GQZ_IX(i, j, ld)] = sT[i + LDW * j];
        U[i + (size_t)w * j] = sU[i + LDW * j];
        V[i + (size_t)w * j] = sV[i + LDW * j];
    }
    if (tid == 0) {
        unsigned long long ms = 0ull, mb = 0ull;
        for (int i = 0; i < w; ++i) {
            if (sl[i]) ms |= 1ull << i;
            if (i + 1 < w && sS[(i + 1) + LDW * i] != 0.0) mb |= 1ull << i;
        }
        out[0] = s_fail;
        out[1] = (int)(unsigned)(ms & 0xffffffffull); out[2] = (int)(unsigned)(ms >> 32);
        out[3] = (int)(unsigned)(mb & 0xffffffffull); out[4] = (int)(unsigned)(mb >> 32);
        out[5] = s_swaps;
    }
}

int gqz_tgord_max_window() { return TGO_MAXW; }

void gqz_launch_tgord_window(hipStream_t st, double *S, double *T, int ld, int p, int w, unsigned long long sel,
                             double *U, double *V, int *out) {
    hipLaunchKernelGGL(k_tgord_window, dim3(1), dim3(TGO_NT), 0, st, S, T, ld, p, w, sel, U, V, out);
}

// ---- kernel-level entry point: one block swap (the pair_swap lane) for parity tests ---------------
// out[0] = 1 when the swap is rejected; otherwise out[1..64] = Qa | Za | S | T (m x m each, ld m,
// m = n1 + n2) with the new corner S = Qa^T S_in Za, T likewise.
__global__ void k_tgex2_one(int n1, int n2, const double *S, const double *T, int ld, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double rec[PREC];
    for (int i = 0; i < PREC; ++i) rec[i] = 0.0;
    if (n1 == 1 && n2 == 1) pair_swap<1, 1>(S, T, ld, 0, rec);
    else if (n1 == 1 && n2 == 2) pair_swap<1, 2>(S, T, ld, 0, rec);
    else if (n1 == 2 && n2 == 1) pair_swap<2, 1>(S, T, ld, 0, rec);
    else pair_swap<2, 2>(S, T, ld, 0, rec);
    out[0] = rec[64];
    for (int i = 0; i < 64; ++i) out[1 + i] = rec[i];
}

extern "C" int gqz_tgex2_dev(int n1, int n2, const double *S, const double *T, int ld, double *out) {
    if (n1 < 1 || n1 > 2 || n2 < 1 || n2 > 2 || ld < n1 + n2 || !S || !T || !out) return -1;
    try {
        if (gqz_init_current()) return -100;
        hipStream_t st = (hipStream_t)gqz_stream();
        hipLaunchKernelGGL(k_tgex2_one, dim3(1), dim3(64), 0, st, n1, n2, S, T, ld, out);
        GQZ_HIP(hipGetLastError());
        GQZ_HIP(hipStreamSynchronize(st));
        return 0;
    } catch (const std::exception &e) {
        gqz_set_error(e.what());
        return -100;
    }
}

