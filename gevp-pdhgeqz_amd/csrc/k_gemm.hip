// k_gemm.hip -- FP64 MFMA slab updates (the PDHGEQZ6 role, PDHGEQZ6.f:276-380 on a 1x1 grid).
//
// After a window of order k has been processed, its accumulated orthogonal factors are
// applied to the rest of the pencil:
//     rows of the window, columns to the right :  C <- U^T C     (k_slab_left)
//     columns of the window, rows above, Q, Z  :  C <- C V       (k_slab_right)
// Both are done IN PLACE: a workgroup owns a panel of rows (right update) or columns (left
// update), stages the panel in LDS once, and writes the product straight back, so every
// element of the slab is read once and written once (algorithmic traffic 16 bytes/element,
// 2k flops/element -> arithmetic intensity k/8 flop/byte).
//
// MFMA: v_mfma_f64_16x16x4_f64 (D[16x16] += A[16x4] B[4x16]).  Operand layout (gfx90a/942/950):
//   A: lane l holds A[l % 16][l / 16];  B: lane l holds B[l / 16][l % 16];
//   D: lane l, element r holds D[l / 16 + 4 * r][l % 16]  (rows interleaved by 4; the FP32
//      16x16x4 form uses D[4 * (l / 16) + r] instead).
#include <hip/hip_runtime.h>
#include "gqz_kernels.h"

typedef double dbl4 __attribute__((ext_vector_type(4)));

static __device__ __forceinline__ dbl4 mfma16(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// workgroup -> (job, panel): jobs are laid out back to back in blockIdx.x
static __device__ __forceinline__ int slab_job(const GqzSlabJobs &J, int &bx) {
    int lo = 0, hi = J.njobs - 1;
    while (lo < hi) {                         // last job with blk0 <= bx
        int mid = (lo + hi + 1) >> 1;
        if (J.blk0[mid] <= bx) lo = mid; else hi = mid - 1;
    }
    bx -= J.blk0[lo];
    return lo;
}

// C[r0 .. r0+RB) x [0,k) <- C * F ;  panel in LDS with leading dimension RB.
template <int RB>
__global__ void __launch_bounds__(256) k_slab_right(GqzSlabJobs J) {
    extern __shared__ double sP[];
    const int tid = threadIdx.x;
    int bx = blockIdx.x;
    const int jb = slab_job(J, bx);
    double *__restrict__ C = J.C[jb];
    const int ldc = J.ldc[jb], m = J.m[jb], k = J.k[jb];
    const double *__restrict__ V = J.F[jb];
    const int ldv = J.ldf[jb];
    const long r0 = (long)bx * RB;
    if (r0 >= m) return;
    const int rows = (int)min((long)RB, (long)m - r0);
    for (int idx = tid; idx < RB * k; idx += 256) {
        int i = idx % RB, j = idx / RB;
        sP[i + RB * j] = (i < rows) ? C[r0 + i + (long)j * ldc] : 0.0;
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    constexpr int RS = RB / 16;
    const int nct = (k + 15) >> 4;
    for (int item = wave; item < RS * nct; item += 4) {
        const int rs = item % RS, ct = item / RS;
        const int jcol = ct * 16 + li;
        dbl4 acc = { 0.0, 0.0, 0.0, 0.0 };
        const double *vcol = V + (long)(jcol < k ? jcol : 0) * ldv;
        const double *prow = sP + rs * 16 + li;
        int kk = 0;
        for (; kk + 4 <= k; kk += 4) {
            const int kc = kk + lk;
            double a = prow[RB * kc];
            double b = (jcol < k) ? vcol[kc] : 0.0;
            acc = mfma16(a, b, acc);
        }
        if (kk < k) {
            const int kc = kk + lk;
            double a = (kc < k) ? prow[RB * kc] : 0.0;
            double b = (kc < k && jcol < k) ? vcol[kc] : 0.0;
            acc = mfma16(a, b, acc);
        }
        if (jcol < k) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = rs * 16 + lk + 4 * r;
                if (row < rows) C[r0 + row + (long)jcol * ldc] = acc[r];
            }
        }
    }
}

// C[0,k) x [c0 .. c0+CB) <- F^T * C ; panel in LDS with leading dimension kmax + 1
template <int CB>
__global__ void __launch_bounds__(256) k_slab_left(GqzSlabJobs J, int kmax) {
    extern __shared__ double sP[];
    const int tid = threadIdx.x;
    int bx = blockIdx.x;
    const int jb = slab_job(J, bx);
    double *__restrict__ C = J.C[jb];
    const int ldc = J.ldc[jb], n = J.m[jb], k = J.k[jb];
    const double *__restrict__ U = J.F[jb];
    const int ldu = J.ldf[jb];
    const long c0 = (long)bx * CB;
    if (c0 >= n) return;
    const int cols = (int)min((long)CB, (long)n - c0);
    const int kp = kmax + 1;
    for (int idx = tid; idx < k * CB; idx += 256) {
        int i = idx % k, j = idx / k;
        sP[i + kp * j] = (j < cols) ? C[i + (c0 + j) * (long)ldc] : 0.0;
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    constexpr int CS = CB / 16;
    const int nrt = (k + 15) >> 4;
    for (int item = wave; item < CS * nrt; item += 4) {
        const int cs = item % CS, rt = item / CS;
        const int irow = rt * 16 + li;
        dbl4 acc = { 0.0, 0.0, 0.0, 0.0 };
        const double *ucol = U + (long)(irow < k ? irow : 0) * ldu;
        const double *pcol = sP + kp * (cs * 16 + li);
        int kk = 0;
        for (; kk + 4 <= k; kk += 4) {
            const int kc = kk + lk;
            double a = (irow < k) ? ucol[kc] : 0.0;
            double b = pcol[kc];
            acc = mfma16(a, b, acc);
        }
        if (kk < k) {
            const int kc = kk + lk;
            double a = (kc < k && irow < k) ? ucol[kc] : 0.0;
            double b = (kc < k) ? pcol[kc] : 0.0;
            acc = mfma16(a, b, acc);
        }
        const int col = cs * 16 + li;
        if (col < cols) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = rt * 16 + lk + 4 * r;
                if (row < k) C[row + (c0 + col) * (long)ldc] = acc[r];
            }
        }
    }
}

bool gqz_slab_add(GqzSlabJobs &J, double *C, int ldc, int m, const double *F, int ldf, int k) {
    if (m <= 0 || k <= 0) return true;
    if (J.njobs >= GQZ_SLAB_MAXJOBS) return false;
    const int q = J.njobs++;
    J.C[q] = C; J.ldc[q] = ldc; J.m[q] = m; J.F[q] = F; J.ldf[q] = ldf; J.k[q] = k;
    return true;
}

static int slab_kmax(const GqzSlabJobs &J) {
    int km = 0;
    for (int q = 0; q < J.njobs; ++q) km = max(km, J.k[q]);
    return km;
}

template <int RB>
static void launch_right(hipStream_t st, GqzSlabJobs &J, int kmax) {
    int tot = 0;
    for (int q = 0; q < J.njobs; ++q) { J.blk0[q] = tot; tot += (J.m[q] + RB - 1) / RB; }
    J.blk0[J.njobs] = tot;
    if (tot == 0) return;
    size_t lds = sizeof(double) * RB * (size_t)kmax;
    static size_t lds_set = 0;
    if (lds > lds_set) {
        hipFuncSetAttribute((const void *)k_slab_right<RB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        lds_set = lds;
    }
    hipLaunchKernelGGL(k_slab_right<RB>, dim3(tot), dim3(256), lds, st, J);
}
template <int CB>
static void launch_left(hipStream_t st, GqzSlabJobs &J, int kmax) {
    int tot = 0;
    for (int q = 0; q < J.njobs; ++q) { J.blk0[q] = tot; tot += (J.m[q] + CB - 1) / CB; }
    J.blk0[J.njobs] = tot;
    if (tot == 0) return;
    size_t lds = sizeof(double) * (size_t)(kmax + 1) * CB;
    static size_t lds_set = 0;
    if (lds > lds_set) {
        hipFuncSetAttribute((const void *)k_slab_left<CB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        lds_set = lds;
    }
    hipLaunchKernelGGL(k_slab_left<CB>, dim3(tot), dim3(256), lds, st, J, kmax);
}

void gqz_slab_right_multi(hipStream_t st, GqzSlabJobs &J) {
    if (J.njobs <= 0) return;
    const int k = slab_kmax(J);
    if (k <= 128) launch_right<64>(st, J, k);
    else if (k <= 256) launch_right<32>(st, J, k);
    else launch_right<16>(st, J, k);
}
void gqz_slab_left_multi(hipStream_t st, GqzSlabJobs &J) {
    if (J.njobs <= 0) return;
    const int k = slab_kmax(J);
    if (k <= 128) launch_left<64>(st, J, k);
    else if (k <= 256) launch_left<32>(st, J, k);
    else launch_left<16>(st, J, k);
}
void gqz_slab_right(hipStream_t st, double *C, int ldc, int m, int k, const double *V, int ldv,
                    double *ws) {
    (void)ws;
    GqzSlabJobs J;
    J.njobs = 0;
    gqz_slab_add(J, C, ldc, m, V, ldv, k);
    gqz_slab_right_multi(st, J);
}
void gqz_slab_left(hipStream_t st, double *C, int ldc, int k, int n, const double *U, int ldu,
                   double *ws) {
    (void)ws;
    GqzSlabJobs J;
    J.njobs = 0;
    gqz_slab_add(J, C, ldc, n, U, ldu, k);
    gqz_slab_left_multi(st, J);
}

// ---- general out-of-place GEMM: C = op(A) op(B), 64x64 tiles, k-steps of 16 via LDS ---------
__global__ void __launch_bounds__(256) k_gemm64(int ta, int tb, int m, int n, int k,
                                                const double *__restrict__ A, int lda,
                                                const double *__restrict__ B, int ldb,
                                                double *__restrict__ C, int ldc) {
    __shared__ double sA[64 * 17];   // 64 rows x 16 k  (ld 64 + padding by k-major 17?)
    __shared__ double sB[16 * 65];   // 16 k x 64 cols
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const long i0 = (long)blockIdx.y * 64, j0 = (long)blockIdx.x * 64;
    const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;   // each wave: 32x32 = 2x2 tiles
    dbl4 acc[2][2];
    for (int a = 0; a < 2; ++a) for (int b = 0; b < 2; ++b) acc[a][b] = dbl4{ 0.0, 0.0, 0.0, 0.0 };
    for (int kk = 0; kk < k; kk += 16) {
        for (int idx = tid; idx < 64 * 16; idx += 256) {
            int i = idx % 64, q = idx / 64;
            long gi = i0 + i; int gk = kk + q;
            double va = 0.0;
            if (gi < m && gk < k) va = ta ? A[gk + gi * (long)lda] : A[gi + (long)gk * lda];
            sA[q * 64 + i] = va;
            int jj = idx % 64, q2 = idx / 64;
            long gj = j0 + jj; int gk2 = kk + q2;
            double vb = 0.0;
            if (gj < n && gk2 < k) vb = tb ? B[gj + (long)gk2 * ldb] : B[gk2 + gj * (long)ldb];
            sB[q2 * 65 + jj] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q += 4) {
            double a0 = sA[(q + lk) * 64 + wr + li], a1 = sA[(q + lk) * 64 + wr + 16 + li];
            double b0 = sB[(q + lk) * 65 + wc + li], b1 = sB[(q + lk) * 65 + wc + 16 + li];
            acc[0][0] = mfma16(a0, b0, acc[0][0]);
            acc[0][1] = mfma16(a0, b1, acc[0][1]);
            acc[1][0] = mfma16(a1, b0, acc[1][0]);
            acc[1][1] = mfma16(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            long col = j0 + wc + 16 * b + li;
            if (col >= n) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                long row = i0 + wr + 16 * a + lk + 4 * r;
                if (row < m) C[row + col * ldc] = acc[a][b][r];
            }
        }
}

void gqz_gemm(hipStream_t st, int ta, int tb, int m, int n, int k, const double *A, int lda,
              const double *B, int ldb, double *C, int ldc) {
    if (m <= 0 || n <= 0) return;
    dim3 grid((n + 63) / 64, (m + 63) / 64);
    hipLaunchKernelGGL(k_gemm64, grid, dim3(256), 0, st, ta, tb, m, n, k, A, lda, B, ldb, C, ldc);
}
