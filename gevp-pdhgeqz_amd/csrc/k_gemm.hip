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
#include "gqz_misc.h"

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

// ---- factors of order k <= 64: streaming tiles through MFMA registers --------------------------
// A workgroup stages the factor in LDS once, in MFMA operand order (entry [t][s][lane] is what
// lane `lane` feeds to the k-step s of output tile t: conflict-free 512-byte reads); each of its
// four waves then streams 16-row (right update) or 16-column (left update) tiles of the slab:
// the tile is read once into registers, multiplied (at most 64 MFMAs) and written back, with no
// further barrier.  The products are formed transposed (D^T = F^T C^T for the right update,
// D^T = C^T U for the left one), so that the D layout stores 16 consecutive rows of a column:
// full 128-byte lines.
#define SLAB_RK 64

// (all 16 loads of a lane are issued before the first LDS write: one memory latency, not 16)
static __device__ __forceinline__ void stage_factor(double *sF, const double *__restrict__ F, int ldf, int k) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int idx = threadIdx.x + 256 * q, r = idx & 63, c = idx >> 6;   // F[r][c], coalesced along r
        v[q] = (r < k && c < k) ? F[r + (long)c * ldf] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int idx = threadIdx.x + 256 * q, r = idx & 63, c = idx >> 6;
        const int t = c >> 4, li = c & 15, s = r >> 2, lk = r & 3;
        sF[((t * 16 + s) << 6) + lk * 16 + li] = v[q];
    }
    __syncthreads();
}

// C[m x k] <- C F : tile = 16 rows of C; D^T[j][i] = sum_kk F[kk][j] C[i][kk].
// NJ = ceil(kmax / 16) output tiles of 16 columns, 4 NJ k-steps; the operands are zero beyond k,
// so the MFMA sequence is straight-line code (no guard between the MFMAs of a tile).
template <int NJ>
__global__ void __launch_bounds__(256) k_slab_right_reg(GqzSlabJobs J, int tpw) {
    __shared__ double sF[64 * 64];
    int bx = blockIdx.x;
    const int jb = slab_job(J, bx);
    double *__restrict__ C = J.C[jb];
    const int ldc = J.ldc[jb], m = J.m[jb], k = J.k[jb];
    stage_factor(sF, J.F[jb], J.ldf[jb], k);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const long tb = ((long)bx * 4 + wave) * tpw;        // first tile of this wave
    for (int t = 0; t < tpw; ++t) {
        const long r0 = (tb + t) * 16;
        if (r0 >= m) break;
        const long row = r0 + li;
        const bool rok = row < m;
        double b[4 * NJ];
#pragma unroll
        for (int s = 0; s < 4 * NJ; ++s) {
            const int c = 4 * s + lk;
            b[s] = (rok && c < k) ? C[row + (long)c * ldc] : 0.0;
        }
        dbl4 acc[NJ];
#pragma unroll
        for (int jt = 0; jt < NJ; ++jt) acc[jt] = dbl4{ 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
        for (int s = 0; s < 4 * NJ; ++s)
#pragma unroll
            for (int jt = 0; jt < NJ; ++jt) acc[jt] = mfma16(sF[((jt * 16 + s) << 6) + lane], b[s], acc[jt]);
        if (rok) {
#pragma unroll
            for (int jt = 0; jt < NJ; ++jt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int c = 16 * jt + lk + 4 * r;
                    if (c < k) C[row + (long)c * ldc] = acc[jt][r];
                }
        }
    }
}

// C[k x n] <- U^T C : tile = 16 columns of C; D^T[j][i] = sum_kk C[kk][j] U[kk][i]
template <int NJ>
__global__ void __launch_bounds__(256) k_slab_left_reg(GqzSlabJobs J, int tpw) {
    __shared__ double sU[64 * 64];
    int bx = blockIdx.x;
    const int jb = slab_job(J, bx);
    double *__restrict__ C = J.C[jb];
    const int ldc = J.ldc[jb], n = J.m[jb], k = J.k[jb];
    stage_factor(sU, J.F[jb], J.ldf[jb], k);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const long tb = ((long)bx * 4 + wave) * tpw;
    for (int t = 0; t < tpw; ++t) {
        const long c0 = (tb + t) * 16;
        if (c0 >= n) break;
        const long col = c0 + li;                      // column of this lane's A operand
        const bool cok = col < n;
        double a[4 * NJ];
#pragma unroll
        for (int s = 0; s < 4 * NJ; ++s) {
            const int r = 4 * s + lk;
            a[s] = (cok && r < k) ? C[r + col * (long)ldc] : 0.0;
        }
        dbl4 acc[NJ];
#pragma unroll
        for (int it = 0; it < NJ; ++it) acc[it] = dbl4{ 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
        for (int s = 0; s < 4 * NJ; ++s)
#pragma unroll
            for (int it = 0; it < NJ; ++it) acc[it] = mfma16(a[s], sU[((it * 16 + s) << 6) + lane], acc[it]);
        // D^T lane l, element r: column lk + 4r of the tile, row 16 it + li
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const long cj = c0 + lk + 4 * r;
            if (cj >= n) continue;
#pragma unroll
            for (int it = 0; it < NJ; ++it) {
                const int i = 16 * it + li;
                if (i < k) C[i + cj * (long)ldc] = acc[it][r];
            }
        }
    }
}

static int slab_kmax(const GqzSlabJobs &J);
static void launch_reg(hipStream_t st, GqzSlabJobs &J, bool right) {
    long tiles = 0;
    for (int q = 0; q < J.njobs; ++q) tiles += (J.m[q] + 15) / 16;
    // tiles per wave: about 2 workgroups per CU before runs get longer, at most 8 tiles per wave
    const int tpw = (int)min(8L, max(1L, tiles / 2048));
    int tot = 0;
    for (int q = 0; q < J.njobs; ++q) {
        J.blk0[q] = tot;
        const int t = (J.m[q] + 15) / 16;
        tot += (t + 4 * tpw - 1) / (4 * tpw);
    }
    J.blk0[J.njobs] = tot;
    if (tot == 0) return;
    const int nj = (slab_kmax(J) + 15) / 16;
#define GQZ_SLAB_REG(NJ)                                                                             \
    do {                                                                                             \
        if (right) hipLaunchKernelGGL(k_slab_right_reg<NJ>, dim3(tot), dim3(256), 0, st, J, tpw);    \
        else hipLaunchKernelGGL(k_slab_left_reg<NJ>, dim3(tot), dim3(256), 0, st, J, tpw);           \
    } while (0)
    switch (nj) {
    case 1: GQZ_SLAB_REG(1); break;
    case 2: GQZ_SLAB_REG(2); break;
    case 3: GQZ_SLAB_REG(3); break;
    default: GQZ_SLAB_REG(4); break;
    }
#undef GQZ_SLAB_REG
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
    static size_t lds_set[64] = {};
    gqz_need_lds((const void *)k_slab_right<RB>, lds, lds_set);
    hipLaunchKernelGGL(k_slab_right<RB>, dim3(tot), dim3(256), lds, st, J);
}
template <int CB>
static void launch_left(hipStream_t st, GqzSlabJobs &J, int kmax) {
    int tot = 0;
    for (int q = 0; q < J.njobs; ++q) { J.blk0[q] = tot; tot += (J.m[q] + CB - 1) / CB; }
    J.blk0[J.njobs] = tot;
    if (tot == 0) return;
    size_t lds = sizeof(double) * (size_t)(kmax + 1) * CB;
    static size_t lds_set[64] = {};
    gqz_need_lds((const void *)k_slab_left<CB>, lds, lds_set);
    hipLaunchKernelGGL(k_slab_left<CB>, dim3(tot), dim3(256), lds, st, J, kmax);
}

void gqz_slab_right_multi(hipStream_t st, GqzSlabJobs &J) {
    if (J.njobs <= 0) return;
    const int k = slab_kmax(J);
    if (k <= SLAB_RK) launch_reg(st, J, true);
    else if (k <= 128) launch_right<64>(st, J, k);
    else if (k <= 256) launch_right<32>(st, J, k);
    else launch_right<16>(st, J, k);
}
void gqz_slab_left_multi(hipStream_t st, GqzSlabJobs &J) {
    if (J.njobs <= 0) return;
    const int k = slab_kmax(J);
    if (k <= SLAB_RK) launch_reg(st, J, false);
    else if (k <= 128) launch_left<64>(st, J, k);
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

// ---- general GEMM: C = alpha op(A) op(B) + beta C (beta = 0: C not read), 64x64 tiles, k-steps
// of 16 via LDS ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gemm64(int ta, int tb, int m, int n, int k,
                                                const double *__restrict__ A, int lda,
                                                const double *__restrict__ B, int ldb,
                                                double *__restrict__ C, int ldc, double alpha, double beta) {
    __shared__ double sA[64 * 17];   // k-major: 16 k-rows of 64 rows (entries 1024.. unused)
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
                if (row < m) {
                    double *c = C + row + col * ldc;
                    *c = (beta == 0.0) ? alpha * acc[a][b][r] : fma(alpha, acc[a][b][r], beta * *c);
                }
            }
        }
}

void gqz_gemm(hipStream_t st, int ta, int tb, int m, int n, int k, const double *A, int lda,
              const double *B, int ldb, double *C, int ldc) {
    gqz_gemm_ab(st, ta, tb, m, n, k, 1.0, A, lda, B, ldb, 0.0, C, ldc);
}
void gqz_gemm_ab(hipStream_t st, int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc) {
    if (m <= 0 || n <= 0) return;
    dim3 grid((n + 63) / 64, (m + 63) / 64);
    hipLaunchKernelGGL(k_gemm64, grid, dim3(256), 0, st, ta, tb, m, n, k, A, lda, B, ldb, C, ldc, alpha, beta);
}
