// k_ht.hip -- Householder QR of B applied to A and Q (the pre-step of the PDGGHRD drivers,
// testing/EXRAND2.f:296-329: PDGEQRF + PDORMQR), plus norms for PCOMPRES-style checks.
//
// Blocked (gqz_launch_house_qr_blk): per panel of QR_NB columns the reflectors are formed one column
// at a time (k_house_make, one workgroup) and applied inside the panel only (k_house_left); the panel's
// product H_k ... H_{k+nb-1} = I - V T V^T (compact WY, T by DLARFT's forward recurrence in
// k_qr_tmat) then goes to the trailing columns of B, to A and to Q as FP64-MFMA GEMMs (k_gemm64):
// B <- B - V (T^T (V^T B)), A likewise, Q <- Q - ((Q V) T) V^T.  The unblocked form (one reflector per
// column applied everywhere, memory bound) remains for small orders.
#include <algorithm>
#include "gqz_dev.h"
#include "gqz_misc.h"

// reflector for column j of B (rows j..n-1): B(j,j) <- beta, B(j+1:n, j) <- v(2:), tau[j]
extern "C" __global__ void __launch_bounds__(1024) k_house_make(double *B, int ld, int n, int j,
                                                                double *tau) {
    __shared__ double red[1024];
    __shared__ double sh[2];
    const int tid = threadIdx.x;
    double *col = B + GQZ_IX(0, j, ld);
    // scaled sum of squares of B(j+1:n, j)
    double mx = 0.0;
    for (int i = j + 1 + tid; i < n; i += blockDim.x) mx = fmax(mx, fabs(col[i]));
    red[tid] = mx;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) { if (tid < s) red[tid] = fmax(red[tid], red[tid + s]); __syncthreads(); }
    const double scale = red[0];
    __syncthreads();
    double ss = 0.0;
    if (scale > 0.0)
        for (int i = j + 1 + tid; i < n; i += blockDim.x) { double t = col[i] / scale; ss += t * t; }
    red[tid] = ss;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) { if (tid < s) red[tid] += red[tid + s]; __syncthreads(); }
    if (tid == 0) {
        double xnorm = scale * sqrt(red[0]);
        double alpha = col[j];
        if (xnorm == 0.0) { sh[0] = 0.0; sh[1] = 0.0; tau[j] = 0.0; }
        else {
            double beta = -copysign(hypot(alpha, xnorm), alpha);
            tau[j] = (beta - alpha) / beta;
            sh[0] = 1.0 / (alpha - beta);
            sh[1] = 1.0;
            col[j] = beta;
        }
    }
    __syncthreads();
    if (sh[1] != 0.0) {
        const double sc = sh[0];
        for (int i = j + 1 + tid; i < n; i += blockDim.x) col[i] *= sc;
    }
}

// M(j:n, c) <- (I - tau v v^T) M(j:n, c) for columns c in [c0, c1); v = [1; B(j+1:n, j)]
// one wavefront per column
extern "C" __global__ void __launch_bounds__(256) k_house_left(const double *B, int ldb, int n, int j,
                                                               const double *tau, double *M, int ldm,
                                                               int c0, int c1) {
    const double t = tau[j];
    if (t == 0.0) return;
    const int lane = threadIdx.x & 63;
    const int wv = (blockIdx.x * (blockDim.x >> 6)) + (threadIdx.x >> 6);
    const int c = c0 + wv;
    if (c >= c1) return;
    const double *v = B + GQZ_IX(0, j, ldb);
    double *m = M + GQZ_IX(0, c, ldm);
    double s = 0.0;
    for (int i = j + lane; i < n; i += 64) s += (i == j ? 1.0 : v[i]) * m[i];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    s *= t;
    for (int i = j + lane; i < n; i += 64) m[i] -= s * (i == j ? 1.0 : v[i]);
}

// Q(r, j:n) <- Q(r, j:n) (I - tau v v^T) for rows r in [0, nq); one thread per row
extern "C" __global__ void __launch_bounds__(256) k_house_right(const double *B, int ldb, int n, int j,
                                                                const double *tau, double *Q, int ldq,
                                                                int nq) {
    const double t = tau[j];
    if (t == 0.0) return;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nq) return;
    const double *v = B + GQZ_IX(0, j, ldb);
    double s = Q[GQZ_IX(r, j, ldq)];
    for (int i = j + 1; i < n; ++i) s += Q[GQZ_IX(r, i, ldq)] * v[i];
    s *= t;
    Q[GQZ_IX(r, j, ldq)] -= s;
    for (int i = j + 1; i < n; ++i) Q[GQZ_IX(r, i, ldq)] -= s * v[i];
}

extern "C" __global__ void k_zero_lower(double *B, int ld, int n) {
    const int j = blockIdx.x;
    for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) B[GQZ_IX(i, j, ld)] = 0.0;
}

// sum of squares of an m x n matrix (optionally only the upper triangle) into out (atomic)
extern "C" __global__ void k_sumsq(const double *A, int ld, int m, int n, int upper, double *out) {
    __shared__ double red[256];
    double s = 0.0;
    const size_t total = (size_t)m * n;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(idx % m), j = (int)(idx / m);
        if (upper && i > j) continue;
        double v = A[GQZ_IX(i, j, ld)];
        s += v * v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) { if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k]; __syncthreads(); }
    if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

// R = R - S (elementwise), or R = R - I when S == null
extern "C" __global__ void k_sub(double *R, int ldr, const double *S, int lds, int n) {
    const size_t total = (size_t)n * n;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(idx % n), j = (int)(idx / n);
        R[GQZ_IX(i, j, ldr)] -= S ? S[GQZ_IX(i, j, lds)] : (i == j ? 1.0 : 0.0);
    }
}

// V (m x nbk, ld m): the panel's reflectors as a unit lower-trapezoidal matrix from B(k.., k..)
extern "C" __global__ void k_qr_vcopy(const double *B, int ldb, int k, int m, int nbk, double *V) {
    const long total = (long)m * nbk;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        const int i = (int)(idx % m), c = (int)(idx / m);
        V[idx] = (i == c) ? 1.0 : (i > c ? B[GQZ_IX(k + i, k + c, ldb)] : 0.0);
    }
}

// T (nbk x nbk upper triangular, ld nbk) with H_0 ... H_{nbk-1} = I - V T V^T, from G = V^T V and tau
// (DLARFT, DIRECT = 'F', STOREV = 'C'): T(i,i) = tau_i, T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i)
#define QR_NB 32
extern "C" __global__ void __launch_bounds__(64) k_qr_tmat(const double *G, int nbk, const double *tau, double *T) {
    __shared__ double sT[QR_NB * QR_NB], sG[QR_NB * QR_NB];
    const int r = threadIdx.x;
    for (int idx = r; idx < nbk * nbk; idx += blockDim.x) { sG[idx] = G[idx]; sT[idx] = 0.0; }
    __syncthreads();
    for (int i = 0; i < nbk; ++i) {
        const double ti = tau[i];
        double t = 0.0;
        if (r < i) {
            for (int c = r; c < i; ++c) t = fma(sT[r + nbk * c], sG[c + nbk * i], t);
            t *= -ti;
        }
        __syncthreads();
        if (r < i) sT[r + nbk * i] = t;
        if (r == i) sT[i + nbk * i] = ti;
        __syncthreads();
    }
    for (int idx = r; idx < nbk * nbk; idx += blockDim.x) T[idx] = sT[idx];
}

size_t gqz_house_qr_ws_doubles(int n) { return 5 * (size_t)n * QR_NB + 2 * (size_t)QR_NB * QR_NB + 256; }

void gqz_launch_house_qr_blk(hipStream_t st, int n, double *A, int lda, double *B, int ldb, double *Q, int ldq,
                             double *tau, double *ws) {
    double *V = ws, *W = V + (size_t)n * QR_NB, *W2 = W + (size_t)n * QR_NB, *Y = W2 + (size_t)n * QR_NB;
    double *Y2 = Y + (size_t)n * QR_NB, *G = Y2 + (size_t)n * QR_NB, *T = G + QR_NB * QR_NB;
    for (int k = 0; k < n - 1; k += QR_NB) {
        const int nbk = std::min(QR_NB, n - 1 - k), m = n - k;
        for (int j = k; j < k + nbk; ++j) {
            hipLaunchKernelGGL(k_house_make, dim3(1), dim3(1024), 0, st, B, ldb, n, j, tau);
            const int nc = k + nbk - (j + 1);
            if (nc > 0)
                hipLaunchKernelGGL(k_house_left, dim3((nc + 3) / 4), dim3(256), 0, st, B, ldb, n, j, tau, B, ldb,
                                   j + 1, k + nbk);
        }
        {
            const long tot = (long)m * nbk;
            int g = (int)((tot + 255) / 256); if (g > 4096) g = 4096;
            hipLaunchKernelGGL(k_qr_vcopy, dim3(g), dim3(256), 0, st, B, ldb, k, m, nbk, V);
        }
        gqz_gemm(st, 1, 0, nbk, nbk, m, V, m, V, m, G, nbk);
        hipLaunchKernelGGL(k_qr_tmat, dim3(1), dim3(64), 0, st, G, nbk, tau + k, T);
        // trailing columns of B, then all columns of A: C <- C - V (T^T (V^T C)), rows k..n-1
        const int nct = n - (k + nbk);
        double *Cs[2] = { B + GQZ_IX(k, k + nbk, ldb), A + GQZ_IX(k, 0, lda) };
        const int ncs[2] = { nct, n }, lds[2] = { ldb, lda };
        for (int q = 0; q < 2; ++q) {
            if (ncs[q] <= 0) continue;
            gqz_gemm(st, 1, 0, nbk, ncs[q], m, V, m, Cs[q], lds[q], W, nbk);
            gqz_gemm(st, 1, 0, nbk, ncs[q], nbk, T, nbk, W, nbk, W2, nbk);
            gqz_gemm_ab(st, 0, 0, m, ncs[q], nbk, -1.0, V, m, W2, nbk, 1.0, Cs[q], lds[q]);
        }
        // Q(:, k..n-1) <- Q(:, k..n-1) - ((Q V) T) V^T
        if (Q) {
            double *C = Q + GQZ_IX(0, k, ldq);
            gqz_gemm(st, 0, 0, n, nbk, m, C, ldq, V, m, Y, n);
            gqz_gemm(st, 0, 0, n, nbk, nbk, Y, n, T, nbk, Y2, n);
            gqz_gemm_ab(st, 0, 1, n, m, nbk, -1.0, Y2, n, V, m, 1.0, C, ldq);
        }
    }
    hipLaunchKernelGGL(k_zero_lower, dim3(n), dim3(256), 0, st, B, ldb, n);
}

void gqz_launch_house_qr(hipStream_t st, int n, double *A, int lda, double *B, int ldb, double *Q, int ldq,
                         double *tau) {
    for (int j = 0; j < n - 1; ++j) {
        hipLaunchKernelGGL(k_house_make, dim3(1), dim3(1024), 0, st, B, ldb, n, j, tau);
        int ncb = n - (j + 1);
        if (ncb > 0) hipLaunchKernelGGL(k_house_left, dim3((ncb + 3) / 4), dim3(256), 0, st, B, ldb, n, j, tau, B, ldb, j + 1, n);
        hipLaunchKernelGGL(k_house_left, dim3((n + 3) / 4), dim3(256), 0, st, B, ldb, n, j, tau, A, lda, 0, n);
        if (Q) hipLaunchKernelGGL(k_house_right, dim3((n + 255) / 256), dim3(256), 0, st, B, ldb, n, j, tau, Q, ldq, n);
    }
    hipLaunchKernelGGL(k_zero_lower, dim3(n), dim3(256), 0, st, B, ldb, n);
}
void gqz_launch_sumsq(hipStream_t st, const double *A, int ld, int m, int n, int upper, double *out) {
    size_t total = (size_t)m * n;
    int nb = (int)((total + 255) / 256); if (nb > 2048) nb = 2048; if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, st, A, ld, m, n, upper, out);
}
void gqz_launch_sub(hipStream_t st, double *R, int ldr, const double *S, int lds, int n) {
    size_t total = (size_t)n * n;
    int nb = (int)((total + 255) / 256); if (nb > 4096) nb = 4096; if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_sub, dim3(nb), dim3(256), 0, st, R, ldr, S, lds, n);
}
