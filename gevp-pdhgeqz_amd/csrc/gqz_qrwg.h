// gqz_qrwg.h -- blocked Householder QR of B inside one workgroup (the QR pre-step of the drivers,
// testing/EXRAND2.f:296-329, PDGEQRF + PDORMQR on a 1x1 grid), for pencils held in HBM.
//
// Panels of QRW_NB columns are factored column by column (one wave per panel column for the
// reflector applications, lanes over rows); the panel's reflectors H_j = I - tau_j v_j v_j^T are
// then combined into the compact WY form H_j0 ... H_j0+nb-1 = I - V T V^T (T upper triangular,
// from the Gram matrix V^T V by the forward recurrence of LAPACK's DLARFT) and applied with FP64
// MFMA (v_mfma_f64_16x16x4f64) to
//     the trailing columns of B and all columns of A:  C <- (I - V T V^T)^T C = C - V (T^T (V^T C)),
//     Q (accumulated, Q = I on entry):                 Q <- Q (I - V T V^T)   = Q - ((Q V) T) V^T,
// so every trailing entry is read twice and written once per panel instead of once per column.
// V lives in B below the diagonal (unit diagonal implicit) and is staged through LDS in chunks of
// QRW_RC rows in MFMA operand order.
//
// MFMA 16x16x4 FP64 operand layout (gfx950): lane l = 16 lk + li holds A[li][lk] and B[lk][li]
// of D += A (16x4) B (4x16); element r of the accumulator is D[lk + 4 r][li].
#pragma once
#include "gqz_dev.h"

#define QRW_NB 32          // reflectors per panel
#define QRW_CB 64          // columns of C per block of the left update (4 tiles of 16)
#define QRW_RC 64          // rows of a staged V chunk
#define QRW_LV 33          // leading dimension of the staged chunk (row-major, padded)
#define QRW_LW 65          // leading dimension of the W buffers (padded)

typedef double qrw_d4 __attribute__((ext_vector_type(4)));

struct QrwLds {
    double sV[QRW_RC * QRW_LV];       // V chunk: sV[r * QRW_LV + b] = V(r0 + r, b)
    double sW[QRW_NB * QRW_LW];       // W (left: V^T C of a column block; right: (Q V)^T of a row block)
    double sW2[QRW_NB * QRW_LW];      // -T^T W, or -(Q V T)^T; the Gram matrix before the updates
    double sT[QRW_NB * QRW_LV];       // T (nb x nb, upper triangular)
    double tau[QRW_NB];
    double red[32];
    double shv[2];
};

static __device__ __forceinline__ qrw_d4 qrw_mfma(double a, double b, qrw_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

static __device__ __forceinline__ double qrw_wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// sum over the workgroup (every thread receives the total); red holds one slot per wave
static __device__ __forceinline__ double qrw_block_sum(double x, double *red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    x = qrw_wave_sum(x);
    if (lane == 0) red[wave] = x;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += red[w];
    __syncthreads();
    return t;
}

// V(i, b) of the panel at column j0 (rows relative to j0): zero above the diagonal, one on it
static __device__ __forceinline__ double qrw_v(const double *B, int ld, int j0, int m, int nb, int i, int b) {
    if (i >= m || b >= nb || i < b) return 0.0;
    if (i == b) return 1.0;
    return B[GQZ_IX(j0 + i, j0 + b, ld)];
}

// stage rows [r0, r0 + QRW_RC) of V into sV (zero outside the panel); caller synchronises
static __device__ __forceinline__ void qrw_stage(QrwLds &S, const double *B, int ld, int j0, int m, int nb, int r0) {
    for (int idx = threadIdx.x; idx < QRW_RC * QRW_NB; idx += blockDim.x) {
        const int r = idx % QRW_RC, b = idx / QRW_RC;
        S.sV[r * QRW_LV + b] = qrw_v(B, ld, j0, m, nb, r0 + r, b);
    }
}

// C <- C - V T^T V^T C for the columns [c0, c0 + ncb) (ncb <= QRW_CB) of the m x . matrix C (rows
// j0.. of the pencil matrix M, leading dimension ld)
template <int NT>
static __device__ __forceinline__ void qrw_left_block(QrwLds &S, const double *B, double *M, int ld, int j0, int m, int nb, int c0,
                                      int ncb) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    double *C = M + GQZ_IX(j0, c0, ld);
    // step 1: W^T = C^T V, items (column tile t, half h of V's columns)
    constexpr int NWV = NT / 64, NQ = (8 + NWV - 1) / NWV;     // waves, items per wave
    qrw_d4 acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = qrw_d4{ 0.0, 0.0, 0.0, 0.0 };
    for (int r0 = 0; r0 < m; r0 += QRW_RC) {
        qrw_stage(S, B, ld, j0, m, nb, r0);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int it = wave + q * NWV;
            const int t = (it >> 1) & 3, h = it & 1;
            const int col = 16 * t + li;
            if (it < 8 && 16 * t < ncb) {
#pragma unroll 4
                for (int s = 0; s < QRW_RC / 4; ++s) {
                    const int r = r0 + 4 * s + lk;
                    const double a = (r < m && col < ncb) ? C[GQZ_IX(r, col, ld)] : 0.0;
                    const double b = S.sV[(4 * s + lk) * QRW_LV + 16 * h + li];
                    acc[q] = qrw_mfma(a, b, acc[q]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int it = wave + q * NWV;
        const int t = (it >> 1) & 3, h = it & 1;
        if (it < 8) {
#pragma unroll
            for (int r = 0; r < 4; ++r) S.sW[(16 * h + li) * QRW_LW + 16 * t + lk + 4 * r] = acc[q][r];
        }
    }
    __syncthreads();
    // step 2: W2 = -T^T W  (T^T lower triangular)
    for (int idx = tid; idx < QRW_NB * QRW_CB; idx += blockDim.x) {
        const int a = idx / QRW_CB, c = idx % QRW_CB;
        double s = 0.0;
        for (int b = 0; b <= a && b < nb; ++b) s += S.sT[b * QRW_LV + a] * S.sW[b * QRW_LW + c];
        S.sW2[a * QRW_LW + c] = (a < nb) ? -s : 0.0;
    }
    __syncthreads();
    // step 3: C^T tile (16 columns x 16 rows) += W2^T V^T
    const int nrb = (m + 15) / 16, nct = (ncb + 15) / 16;
    for (int it = wave; it < nrb * nct; it += nw) {
        const int rb = it / nct, t = it % nct;
        const int row = 16 * rb + li;
        qrw_d4 d;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int cc = 16 * t + lk + 4 * r;
            d[r] = (row < m && cc < ncb) ? C[GQZ_IX(row, cc, ld)] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < QRW_NB / 4; ++u) {
            const double a = S.sW2[(4 * u + lk) * QRW_LW + 16 * t + li];
            const double b = qrw_v(B, ld, j0, m, nb, row, 4 * u + lk);
            d = qrw_mfma(a, b, d);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int cc = 16 * t + lk + 4 * r;
            if (row < m && cc < ncb) C[GQZ_IX(row, cc, ld)] = d[r];
        }
    }
    __syncthreads();
}

// Q(q0 : q0 + nr, j0 : n) <- Q - ((Q V) T) V^T for one block of nr <= 64 rows of Q
template <int NT>
static __device__ __forceinline__ void qrw_right_block(QrwLds &S, const double *B, double *Q, int ld, int j0, int m, int nb, int q0,
                                       int nr) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    double *C = Q + GQZ_IX(q0, j0, ld);
    // step 1: Y = C V (nr x nb), items (row tile t, half h)
    constexpr int NWV = NT / 64, NQ = (8 + NWV - 1) / NWV;     // waves, items per wave
    qrw_d4 acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = qrw_d4{ 0.0, 0.0, 0.0, 0.0 };
    for (int r0 = 0; r0 < m; r0 += QRW_RC) {
        qrw_stage(S, B, ld, j0, m, nb, r0);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int it = wave + q * NWV;
            const int t = (it >> 1) & 3, h = it & 1;
            const int row = 16 * t + li;
            if (it < 8 && 16 * t < nr) {
#pragma unroll 4
                for (int s = 0; s < QRW_RC / 4; ++s) {
                    const int k = r0 + 4 * s + lk;
                    const double a = (row < nr && k < m) ? C[GQZ_IX(row, k, ld)] : 0.0;
                    const double b = S.sV[(4 * s + lk) * QRW_LV + 16 * h + li];
                    acc[q] = qrw_mfma(a, b, acc[q]);
                }
            }
        }
        __syncthreads();
    }
    // D[lk + 4r][li] = Y[16 t + lk + 4 r][16 h + li]  ->  sW[b][row]
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int it = wave + q * NWV;
        const int t = (it >> 1) & 3, h = it & 1;
        if (it < 8) {
#pragma unroll
            for (int r = 0; r < 4; ++r) S.sW[(16 * h + li) * QRW_LW + 16 * t + lk + 4 * r] = acc[q][r];
        }
    }
    __syncthreads();
    // step 2: Y2 = -Y T  ->  sW2[a][row] = -sum_{b <= a} Y[row][b] T[b][a]
    for (int idx = tid; idx < QRW_NB * 64; idx += blockDim.x) {
        const int a = idx / 64, row = idx % 64;
        double s = 0.0;
        for (int b = 0; b <= a && b < nb; ++b) s += S.sW[b * QRW_LW + row] * S.sT[b * QRW_LV + a];
        S.sW2[a * QRW_LW + row] = (a < nb) ? -s : 0.0;
    }
    __syncthreads();
    // step 3: C^T tile (16 columns of C x 16 rows) += V Y2^T
    const int nct = (m + 15) / 16, nrt = (nr + 15) / 16;
    for (int it = wave; it < nct * nrt; it += nw) {
        const int ct = it / nrt, t = it % nrt;
        const int row = 16 * t + li;
        qrw_d4 d;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int cc = 16 * ct + lk + 4 * r;
            d[r] = (row < nr && cc < m) ? C[GQZ_IX(row, cc, ld)] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < QRW_NB / 4; ++u) {
            const double a = qrw_v(B, ld, j0, m, nb, 16 * ct + li, 4 * u + lk);
            const double b = S.sW2[(4 * u + lk) * QRW_LW + 16 * t + li];
            d = qrw_mfma(a, b, d);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int cc = 16 * ct + lk + 4 * r;
            if (row < nr && cc < m) C[GQZ_IX(row, cc, ld)] = d[r];
        }
    }
    __syncthreads();
}

// Householder QR of B (n x n, ld) applied to A (B <- Q^T B = R, A <- Q^T A) and accumulated into Q
// (Q = I on entry; may be null).  Reflectors as LAPACK's DGEQR2/DLARFG: beta = -sign(alpha) ||x||,
// tau = (beta - alpha) / beta, v = x / (alpha - beta) below the diagonal.
template <int NT>
static __device__ __forceinline__ void wg_qr_b_blk(double *A, double *B, double *Q, int n, int ld, QrwLds &S) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    for (int j0 = 0; j0 < n - 1; j0 += QRW_NB) {
        const int nb = min(QRW_NB, n - 1 - j0);      // reflectors of this panel
        const int m = n - j0;                         // rows of the panel
        // ---- panel factorisation
        for (int jj = 0; jj < nb; ++jj) {
            const int j = j0 + jj;
            double *col = B + GQZ_IX(0, j, ld);
            double ss = 0.0;
            for (int i = j + 1 + tid; i < n; i += blockDim.x) ss += col[i] * col[i];
            ss = qrw_block_sum(ss, S.red);
            if (tid == 0) {
                const double xnorm = sqrt(ss), alpha = col[j];
                if (xnorm == 0.0) { S.shv[0] = 0.0; S.shv[1] = 0.0; }
                else {
                    const double beta = -copysign(hypot(alpha, xnorm), alpha);
                    S.shv[0] = (beta - alpha) / beta;
                    S.shv[1] = 1.0 / (alpha - beta);
                    col[j] = beta;
                }
                S.tau[jj] = S.shv[0];
            }
            __syncthreads();
            const double tau = S.shv[0], sc = S.shv[1];
            if (tau != 0.0)
                for (int i = j + 1 + tid; i < n; i += blockDim.x) col[i] *= sc;
            __syncthreads();
            if (tau != 0.0) {
                // the remaining panel columns: one wave per column, lanes over rows
                for (int c = j + 1 + wave; c < j0 + nb; c += nw) {
                    double *mc = B + GQZ_IX(0, c, ld);
                    double s = (lane == 0) ? mc[j] : 0.0;
                    for (int i = j + 1 + lane; i < n; i += 64) s += col[i] * mc[i];
                    s = qrw_wave_sum(s) * tau;
                    if (lane == 0) mc[j] -= s;
                    for (int i = j + 1 + lane; i < n; i += 64) mc[i] -= s * col[i];
                }
            }
            __syncthreads();
        }
        // ---- T: Gram matrix G = V^T V (upper blocks (0,0), (0,1), (1,1) of 16 x 16 by MFMA, one wave
        // each; plain dot products with fewer than three waves), then the recurrence
        if (nw >= 3) {
            qrw_d4 g = qrw_d4{ 0.0, 0.0, 0.0, 0.0 };
            const int a_h = (wave == 2) ? 1 : 0, b_h = (wave == 0) ? 0 : 1;
            for (int r0 = 0; r0 < m; r0 += QRW_RC) {
                qrw_stage(S, B, ld, j0, m, nb, r0);
                __syncthreads();
                if (wave < 3) {
#pragma unroll 4
                    for (int s = 0; s < QRW_RC / 4; ++s) {
                        const double a = S.sV[(4 * s + lk) * QRW_LV + 16 * a_h + li];
                        const double b = S.sV[(4 * s + lk) * QRW_LV + 16 * b_h + li];
                        g = qrw_mfma(a, b, g);
                    }
                }
                __syncthreads();
            }
            if (wave < 3) {
#pragma unroll
                for (int r = 0; r < 4; ++r) S.sW2[(16 * a_h + lk + 4 * r) * QRW_LW + 16 * b_h + li] = g[r];
            }
        } else {
            for (int idx = tid; idx < nb * nb; idx += blockDim.x) {
                const int a = idx % nb, b = idx / nb;
                if (a >= b) continue;
                double s = 0.0;
                for (int i = b; i < m; ++i) s += qrw_v(B, ld, j0, m, nb, i, a) * qrw_v(B, ld, j0, m, nb, i, b);
                S.sW2[a * QRW_LW + b] = s;
            }
        }
        __syncthreads();
        if (wave == 0) {
            // lane a owns row a of T: T(a, i) = -tau_i sum_{a <= b < i} T(a, b) G(b, i), T(i, i) = tau_i
            const int a = lane;
            for (int i = 0; i < nb; ++i) {
                if (a < i) {
                    double s = 0.0;
                    for (int b = a; b < i; ++b) s += S.sT[a * QRW_LV + b] * S.sW2[b * QRW_LW + i];
                    S.sT[a * QRW_LV + i] = -S.tau[i] * s;
                } else if (a == i) {
                    S.sT[a * QRW_LV + a] = S.tau[a];
                } else if (a < QRW_NB) {
                    S.sT[a * QRW_LV + i] = 0.0;
                }
            }
        }
        __syncthreads();
        // ---- trailing columns of B, all columns of A
        const int cB = j0 + nb, ncB = n - cB;
        for (int c0 = 0; c0 < ncB; c0 += QRW_CB) qrw_left_block<NT>(S, B, B, ld, j0, m, nb, cB + c0, min(QRW_CB, ncB - c0));
        for (int c0 = 0; c0 < n; c0 += QRW_CB) qrw_left_block<NT>(S, B, A, ld, j0, m, nb, c0, min(QRW_CB, n - c0));
        if (Q)
            for (int q0 = 0; q0 < n; q0 += 64) qrw_right_block<NT>(S, B, Q, ld, j0, m, nb, q0, min(64, n - q0));
    }
    for (int idx = tid; idx < n * n; idx += blockDim.x) {
        const int i = idx % n, j = idx / n;
        if (i > j) B[GQZ_IX(i, j, ld)] = 0.0;
    }
    __syncthreads();
}
