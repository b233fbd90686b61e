// gqz_swap.h -- register-resident swap of adjacent diagonal blocks (DTGEX2 semantics,
// src/reorder/dtgex2.f; direct swapping method of Kagstrom and Poromaa).
//
// For blocks of orders N1 (upper) and N2 (lower) of the (N1+N2) x (N1+N2) corner (S, T):
//   1. solve the generalized Sylvester equation  S11 R - L S22 = g S12,  T11 R - L T22 = g T12
//      by Gaussian elimination with complete pivoting on its Kronecker form (DTGSY2/DGETC2),
//   2. Za = orthogonal factor of QR([-R; g I]), Qa = orthogonal factor of QR([-L; g I]),
//   3. S <- Qa^T S Za, T <- Qa^T T Za, with the weak (lower-left block) and strong
//      (backward error of the reconstruction) stability tests of DTGEX2,
//   4. re-standardise 2x2 blocks (DLAGV2).
// Everything is templated on the block sizes and fully unrolled so that the arrays live in
// registers (no scratch memory).
#pragma once
#include "gqz_dev.h"

// Gaussian elimination with complete pivoting (DGETC2 + DGESC2 semantics); returns the scale g.
template <int M>
__device__ __forceinline__ double r_solve_cp(double (&K)[M * M], double (&rhs)[M]) {
    int cperm[M];
#pragma unroll
    for (int i = 0; i < M; ++i) cperm[i] = i;
    double xmax = 0.0;
#pragma unroll
    for (int i = 0; i < M * M; ++i) xmax = fmax(xmax, fabs(K[i]));
    const double smlnum = GQZ_SAFMIN / GQZ_ULP;
    const double smin = fmax(GQZ_ULP * xmax, smlnum);
#pragma unroll
    for (int k = 0; k < M; ++k) {
        // pivot search over the trailing block
        double pv = -1.0;
        int pi = k, pj = k;
#pragma unroll
        for (int j = k; j < M; ++j)
#pragma unroll
            for (int i = k; i < M; ++i) {
                double a = fabs(K[i + M * j]);
                bool t = a > pv;
                pv = t ? a : pv; pi = t ? i : pi; pj = t ? j : pj;
            }
        // swap rows k <-> pi (all columns) and rhs
#pragma unroll
        for (int j = 0; j < M; ++j) {
            double rk = K[k + M * j], rp = rk;
#pragma unroll
            for (int i = k + 1; i < M; ++i) rp = (i == pi) ? K[i + M * j] : rp;
#pragma unroll
            for (int i = k + 1; i < M; ++i) K[i + M * j] = (i == pi) ? rk : K[i + M * j];
            K[k + M * j] = rp;
        }
        {
            double rk = rhs[k], rp = rk;
#pragma unroll
            for (int i = k + 1; i < M; ++i) rp = (i == pi) ? rhs[i] : rp;
#pragma unroll
            for (int i = k + 1; i < M; ++i) rhs[i] = (i == pi) ? rk : rhs[i];
            rhs[k] = rp;
        }
        // swap columns k <-> pj (all rows) and record the permutation
#pragma unroll
        for (int i = 0; i < M; ++i) {
            double ck = K[i + M * k], cp = ck;
#pragma unroll
            for (int j = k + 1; j < M; ++j) cp = (j == pj) ? K[i + M * j] : cp;
#pragma unroll
            for (int j = k + 1; j < M; ++j) K[i + M * j] = (j == pj) ? ck : K[i + M * j];
            K[i + M * k] = cp;
        }
        {
            int ck = cperm[k], cp = ck;
#pragma unroll
            for (int j = k + 1; j < M; ++j) cp = (j == pj) ? cperm[j] : cp;
#pragma unroll
            for (int j = k + 1; j < M; ++j) cperm[j] = (j == pj) ? ck : cperm[j];
            cperm[k] = cp;
        }
        if (fabs(K[k + M * k]) < smin) K[k + M * k] = smin;
        const double inv = 1.0 / K[k + M * k];
#pragma unroll
        for (int i = k + 1; i < M; ++i) {
            const double f = K[i + M * k] * inv;
            K[i + M * k] = f;
#pragma unroll
            for (int j = k + 1; j < M; ++j) K[i + M * j] -= f * K[k + M * j];
            rhs[i] -= f * rhs[k];
        }
    }
    double scale = 1.0, bmax = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) bmax = fmax(bmax, fabs(rhs[i]));
    if (2.0 * smlnum * bmax > fabs(K[(M - 1) + M * (M - 1)])) {
        const double t = 0.5 / bmax;
#pragma unroll
        for (int i = 0; i < M; ++i) rhs[i] *= t;
        scale = t;
    }
    double x[M];
#pragma unroll
    for (int i = M - 1; i >= 0; --i) {
        double s = rhs[i];
#pragma unroll
        for (int j = i + 1; j < M; ++j) s -= K[i + M * j] * x[j];
        x[i] = s / K[i + M * i];
    }
    // undo the column permutation: unknown cperm[i] = x[i]
#pragma unroll
    for (int u = 0; u < M; ++u) {
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < M; ++i) v = (cperm[i] == u) ? x[i] : v;
        rhs[u] = v;
    }
    return scale;
}

// orthogonal factor Qo (M x M) of the Householder QR of an M x KC matrix X (destroyed)
template <int M, int KC>
__device__ __forceinline__ void r_qr_full(double (&X)[M * KC], double (&Qo)[M * M]) {
    double vs[KC][M], taus[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) {
        // Householder on X(j:M, j)
        double alpha = X[j + M * j];
        double ss = 0.0;
#pragma unroll
        for (int i = j + 1; i < M; ++i) ss += X[i + M * j] * X[i + M * j];
        double tau = 0.0, beta = alpha, sc = 0.0;
        if (ss != 0.0) {
            beta = -copysign(sqrt(alpha * alpha + ss), alpha);
            sc = 1.0 / (alpha - beta);
        }
        // tau = 2 / (v^T v) of the stored v, so that the reflector is orthogonal whatever the
        // rounding of v: (beta - alpha) / beta equals it only in exact arithmetic, and with the
        // squares in the subnormal range the two differ by many percent (a collapsed column of the
        // swap factor, accepted by the stability tests when the block is nearly singular)
        double vv = 1.0;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            vs[j][i] = (i < j) ? 0.0 : (i == j ? 1.0 : X[i + M * j] * sc);
            if (i > j) vv += vs[j][i] * vs[j][i];
        }
        if (ss != 0.0) tau = 2.0 / vv;
        taus[j] = tau;
        X[j + M * j] = beta;
#pragma unroll
        for (int jj = j + 1; jj < KC; ++jj) {
            double s = 0.0;
#pragma unroll
            for (int i = j; i < M; ++i) s += vs[j][i] * X[i + M * jj];
            s *= tau;
#pragma unroll
            for (int i = j; i < M; ++i) X[i + M * jj] -= s * vs[j][i];
        }
    }
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
        for (int i = 0; i < M; ++i) Qo[i + M * j] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int j = KC - 1; j >= 0; --j)
#pragma unroll
        for (int jj = 0; jj < M; ++jj) {
            double s = 0.0;
#pragma unroll
            for (int i = j; i < M; ++i) s += vs[j][i] * Qo[i + M * jj];
            s *= taus[j];
#pragma unroll
            for (int i = j; i < M; ++i) Qo[i + M * jj] -= s * vs[j][i];
        }
}

template <int M>
__device__ __forceinline__ void r_mm(const double (&A)[M * M], bool ta, const double (&B)[M * M], bool tb,
                                     double (&C)[M * M]) {
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
        for (int i = 0; i < M; ++i) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < M; ++k) s += (ta ? A[k + M * i] : A[i + M * k]) * (tb ? B[j + M * k] : B[k + M * j]);
            C[i + M * j] = s;
        }
}

// 2x2 standardisation of the diagonal block at offset O of an M x M register pencil,
// accumulating the rotations into Qa (left) and Za (right)
template <int M, int O>
__device__ __forceinline__ void r_std2(double (&S)[M * M], double (&T)[M * M], double (&Qa)[M * M],
                                       double (&Za)[M * M]) {
    // make T's block upper triangular first
    double c, s, r;
    d_lartg(T[(O + 1) + M * (O + 1)], T[(O + 1) + M * O], c, s, r);
#pragma unroll
    for (int i = 0; i < M; ++i) {
        double x, y;
        x = S[i + M * (O + 1)]; y = S[i + M * O]; S[i + M * (O + 1)] = c * x + s * y; S[i + M * O] = c * y - s * x;
        x = T[i + M * (O + 1)]; y = T[i + M * O]; T[i + M * (O + 1)] = c * x + s * y; T[i + M * O] = c * y - s * x;
        x = Za[i + M * (O + 1)]; y = Za[i + M * O]; Za[i + M * (O + 1)] = c * x + s * y; Za[i + M * O] = c * y - s * x;
    }
    T[(O + 1) + M * O] = 0.0;
    double Ab[4] = { S[O + M * O], S[(O + 1) + M * O], S[O + M * (O + 1)], S[(O + 1) + M * (O + 1)] };
    double Bb[4] = { T[O + M * O], 0.0, T[O + M * (O + 1)], T[(O + 1) + M * (O + 1)] };
    double QL[4], ZR[4];
    d_lagv2(Ab, Bb, QL, ZR);
#pragma unroll
    for (int j = 0; j < M; ++j) {   // rows O, O+1 by QL^T
        double x = S[O + M * j], y = S[(O + 1) + M * j];
        S[O + M * j] = QL[0] * x + QL[1] * y; S[(O + 1) + M * j] = QL[2] * x + QL[3] * y;
        x = T[O + M * j]; y = T[(O + 1) + M * j];
        T[O + M * j] = QL[0] * x + QL[1] * y; T[(O + 1) + M * j] = QL[2] * x + QL[3] * y;
    }
#pragma unroll
    for (int i = 0; i < M; ++i) {   // columns O, O+1 by ZR; Qa by QL, Za by ZR
        double a = S[i + M * O], b = S[i + M * (O + 1)];
        S[i + M * O] = a * ZR[0] + b * ZR[1]; S[i + M * (O + 1)] = a * ZR[2] + b * ZR[3];
        a = T[i + M * O]; b = T[i + M * (O + 1)];
        T[i + M * O] = a * ZR[0] + b * ZR[1]; T[i + M * (O + 1)] = a * ZR[2] + b * ZR[3];
        a = Qa[i + M * O]; b = Qa[i + M * (O + 1)];
        Qa[i + M * O] = a * QL[0] + b * QL[1]; Qa[i + M * (O + 1)] = a * QL[2] + b * QL[3];
        a = Za[i + M * O]; b = Za[i + M * (O + 1)];
        Za[i + M * O] = a * ZR[0] + b * ZR[1]; Za[i + M * (O + 1)] = a * ZR[2] + b * ZR[3];
    }
    S[O + M * O] = Ab[0]; S[(O + 1) + M * O] = Ab[1]; S[O + M * (O + 1)] = Ab[2]; S[(O + 1) + M * (O + 1)] = Ab[3];
    T[O + M * O] = Bb[0]; T[(O + 1) + M * O] = 0.0; T[O + M * (O + 1)] = Bb[2]; T[(O + 1) + M * (O + 1)] = Bb[3];
}

// The swap itself.  S, T: the M x M corner (in/out).  Returns 0 on success (Qa, Za set).
template <int N1, int N2>
__device__ __forceinline__ int r_tgex2(double (&S)[(N1 + N2) * (N1 + N2)], double (&T)[(N1 + N2) * (N1 + N2)],
                                       double (&Qa)[(N1 + N2) * (N1 + N2)], double (&Za)[(N1 + N2) * (N1 + N2)]) {
    constexpr int M = N1 + N2, NN = N1 * N2, M2 = 2 * NN;
    double S0[M * M], T0[M * M];
    double sa = 0.0, sb = 0.0;
#pragma unroll
    for (int i = 0; i < M * M; ++i) { S0[i] = S[i]; T0[i] = T[i]; sa += S[i] * S[i]; sb += T[i] * T[i]; }
    const double smlnum = GQZ_SAFMIN / GQZ_ULP;
    const double tha = fmax(20.0 * GQZ_ULP * sqrt(sa), smlnum), thb = fmax(20.0 * GQZ_ULP * sqrt(sb), smlnum);
    double K[M2 * M2], rhs[M2];
#pragma unroll
    for (int i = 0; i < M2 * M2; ++i) K[i] = 0.0;
#pragma unroll
    for (int q = 0; q < N2; ++q)
#pragma unroll
        for (int p = 0; p < N1; ++p) {
            const int row = p + N1 * q;
#pragma unroll
            for (int k = 0; k < N1; ++k) {
                K[row + M2 * (k + N1 * q)] += S0[p + M * k];
                K[row + NN + M2 * (k + N1 * q)] += T0[p + M * k];
            }
#pragma unroll
            for (int k = 0; k < N2; ++k) {
                K[row + M2 * (NN + p + N1 * k)] -= S0[(N1 + k) + M * (N1 + q)];
                K[row + NN + M2 * (NN + p + N1 * k)] -= T0[(N1 + k) + M * (N1 + q)];
            }
            rhs[row] = S0[p + M * (N1 + q)];
            rhs[row + NN] = T0[p + M * (N1 + q)];
        }
    const double g = r_solve_cp<M2>(K, rhs);
    double XR[M * N2], XL[M * N2];
#pragma unroll
    for (int j = 0; j < N2; ++j)
#pragma unroll
        for (int i = 0; i < M; ++i) {
            XR[i + M * j] = (i < N1) ? -rhs[i + N1 * j] : ((i - N1) == j ? g : 0.0);
            XL[i + M * j] = (i < N1) ? -rhs[NN + i + N1 * j] : ((i - N1) == j ? g : 0.0);
        }
    r_qr_full<M, N2>(XR, Za);
    r_qr_full<M, N2>(XL, Qa);
    double W[M * M], S2[M * M], T2[M * M];
    r_mm<M>(Qa, true, S0, false, W); r_mm<M>(W, false, Za, false, S2);
    r_mm<M>(Qa, true, T0, false, W); r_mm<M>(W, false, Za, false, T2);
    double ea = 0.0, eb = 0.0;
#pragma unroll
    for (int j = 0; j < N2; ++j)
#pragma unroll
        for (int i = N2; i < M; ++i) { ea += S2[i + M * j] * S2[i + M * j]; eb += T2[i + M * j] * T2[i + M * j]; }
    if (!(sqrt(ea) <= tha && sqrt(eb) <= thb)) return 1;
#pragma unroll
    for (int j = 0; j < N2; ++j)
#pragma unroll
        for (int i = N2; i < M; ++i) { S2[i + M * j] = 0.0; T2[i + M * j] = 0.0; }
    if constexpr (N2 == 2) r_std2<M, 0>(S2, T2, Qa, Za);
    if constexpr (N1 == 2) r_std2<M, N2>(S2, T2, Qa, Za);
    double S3[M * M], T3[M * M];
    r_mm<M>(Qa, false, S2, false, W); r_mm<M>(W, false, Za, true, S3);
    r_mm<M>(Qa, false, T2, false, W); r_mm<M>(W, false, Za, true, T3);
    double da = 0.0, db = 0.0;
#pragma unroll
    for (int i = 0; i < M * M; ++i) { da += (S3[i] - S0[i]) * (S3[i] - S0[i]); db += (T3[i] - T0[i]) * (T3[i] - T0[i]); }
    if (!(sqrt(da) <= tha && sqrt(db) <= thb)) return 1;
#pragma unroll
    for (int j = 0; j < M; ++j)
#pragma unroll
        for (int i = j + 1; i < M; ++i) T2[i + M * j] = 0.0;
#pragma unroll
    for (int i = 0; i < M * M; ++i) { S[i] = S2[i]; T[i] = T2[i]; }
    return 0;
}
