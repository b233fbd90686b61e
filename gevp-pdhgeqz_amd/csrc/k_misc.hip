// k_misc.hip -- deflation scan, infinite-eigenvalue deflation, window copies, finalisation.
#include "gqz_wg.h"
#include "gqz_misc.h"

// ----------------------------------------------------------------------------------------------
// k_inf_defl: one workgroup.  (1) bottom-up search for a negligible subdiagonal
// (PDLASMSUB.f:368) giving ktop; (2) zeros of diag(T) inside [ktop, kbot] (|T_ii| <=
// max(TTOL, SMLNUM), PDHGEQZ7.f:380) are deflated: bottom zero by a column rotation, top zero
// by a row rotation (beta = 0 exactly), interior zeros chased to the bottom first (DHGEQZ7 /
// PDHGEQZ8 direction).  Rotations are applied to the full pencil (Schur form) and to Q, Z.
// kt0 >= 0: the search for ktop is skipped and the active block starts at kt0 (a repeated call after
// an interior zero was chased by the caller).  chase = 0: an interior zero is not chased here; the
// kernel stops and reports its row in out[4] (the caller chases it in LDS windows, k_zchase_window).
//   out[0] = ktop, out[1] = kt, out[2] = kb (remaining active block), out[3] = #deflated,
//   out[4] = row of the interior zero left to the caller (-1: none)
// ----------------------------------------------------------------------------------------------
extern "C" __global__ void __launch_bounds__(1024) k_inf_defl(double *H, double *T, double *Q,
                                                              double *Z, int ld, int ldq, int n,
                                                              int nq, int ilo, int kbot, double ttol,
                                                              double smlnum, double *ar, double *ai,
                                                              double *be, int *out, int kt0, int chase) {
    __shared__ int s_i[8];
    __shared__ int s_max;
    __shared__ double s_d[4];
    const int tid = threadIdx.x, nt = blockDim.x;
    WgPencil P;
    P.H = H; P.T = T; P.U = Q; P.V = Z; P.ld = ld; P.ldu = ldq; P.n = n; P.nq = nq;
    // lowest negligible subdiagonal entry (the bottom-up search of PDLASMSUB finds the largest k),
    // found by all lanes at once
    if (tid == 0) s_max = ilo;
    __syncthreads();
    if (kt0 < 0) {
        for (int k = ilo + 1 + tid; k <= kbot; k += nt) {
            double tst = fabs(H[GQZ_IX(k - 1, k - 1, ld)]) + fabs(H[GQZ_IX(k, k, ld)]);
            if (fabs(H[GQZ_IX(k, k - 1, ld)]) <= fmax(smlnum, GQZ_ULP * tst)) atomicMax(&s_max, k);
        }
    }
    __syncthreads();
    if (tid == 0) {
        const int ktop = kt0 >= 0 ? kt0 : s_max;
        if (kt0 < 0 && ktop > ilo) H[GQZ_IX(ktop, ktop - 1, ld)] = 0.0;
        s_i[0] = ktop; s_i[1] = ktop; s_i[2] = kbot; s_i[3] = 0; s_i[6] = -1;
    }
    __syncthreads();
    const double tol = fmax(ttol, smlnum);
    for (;;) {
        // negligible diagonal entries of T in [kt, kb] are set to zero; s_max = the lowest one
        if (tid == 0) s_max = -1;
        __syncthreads();
        {
            const int kt = s_i[1], kb = s_i[2];
            for (int i = kt + tid; i <= kb; i += nt)
                if (fabs(T[GQZ_IX(i, i, ld)]) <= tol) { T[GQZ_IX(i, i, ld)] = 0.0; atomicMax(&s_max, i); }
        }
        __syncthreads();
        if (tid == 0) {
            int kt = s_i[1], kb = s_i[2];
            s_i[4] = 0;   // op: 0 done, 1 colrot bottom, 2 rowrot top, 3 chase
            if (kb >= kt) {
                const int found = s_max;
                if (found >= 0) {
                    double c, s, r;
                    if (kb == kt) {
                        ar[kb] = H[GQZ_IX(kb, kb, ld)]; ai[kb] = 0.0; be[kb] = 0.0;
                        s_i[2] = kb - 1; s_i[3]++; s_i[4] = 4;
                    } else if (T[GQZ_IX(kb, kb, ld)] == 0.0) {
                        d_lartg(H[GQZ_IX(kb, kb, ld)], H[GQZ_IX(kb, kb - 1, ld)], c, s, r);
                        H[GQZ_IX(kb, kb, ld)] = r; H[GQZ_IX(kb, kb - 1, ld)] = 0.0;
                        ar[kb] = r; ai[kb] = 0.0; be[kb] = 0.0;
                        s_d[0] = c; s_d[1] = s; s_i[5] = kb;
                        s_i[2] = kb - 1; s_i[3]++; s_i[4] = 1;
                    } else if (T[GQZ_IX(kt, kt, ld)] == 0.0) {
                        d_lartg(H[GQZ_IX(kt, kt, ld)], H[GQZ_IX(kt + 1, kt, ld)], c, s, r);
                        H[GQZ_IX(kt, kt, ld)] = r; H[GQZ_IX(kt + 1, kt, ld)] = 0.0;
                        ar[kt] = r; ai[kt] = 0.0; be[kt] = 0.0;
                        s_d[0] = c; s_d[1] = s; s_i[5] = kt;
                        s_i[1] = kt + 1; s_i[3]++; s_i[4] = 2;
                    } else if (chase) {
                        s_i[5] = found; s_i[4] = 3;
                    } else {
                        s_i[6] = found;     // left to the caller; op stays 0 (done)
                    }
                }
            }
        }
        __syncthreads();
        const int op = s_i[4];
        if (op == 0) break;
        if (op == 1) {
            const int kb = s_i[5];
            wg_colrot_all(P, kb, kb - 1, kb, kb, s_d[0], s_d[1]);
        } else if (op == 2) {
            const int kt = s_i[5];
            wg_rowrot_all(P, kt, kt + 1, kt + 1, s_d[0], s_d[1]);
        } else if (op == 3) {
            const int j = s_i[5], kt = s_i[1], kb = s_i[2];
            for (int jch = j; jch < kb; ++jch) {
                if (tid == 0) {
                    double c, s, r;
                    d_lartg(T[GQZ_IX(jch, jch + 1, ld)], T[GQZ_IX(jch + 1, jch + 1, ld)], c, s, r);
                    T[GQZ_IX(jch, jch + 1, ld)] = r; T[GQZ_IX(jch + 1, jch + 1, ld)] = 0.0;
                    s_d[0] = c; s_d[1] = s;
                }
                __syncthreads();
                wg_rowrot_all(P, jch, jch > kt ? jch - 1 : jch, jch + 2, s_d[0], s_d[1]);
                __syncthreads();
                if (jch > kt) {
                    if (tid == 0) {
                        double c, s, r;
                        d_lartg(H[GQZ_IX(jch + 1, jch, ld)], H[GQZ_IX(jch + 1, jch - 1, ld)], c, s, r);
                        H[GQZ_IX(jch + 1, jch, ld)] = r; H[GQZ_IX(jch + 1, jch - 1, ld)] = 0.0;
                        s_d[0] = c; s_d[1] = s;
                    }
                    __syncthreads();
                    wg_colrot_all(P, jch, jch - 1, jch + 1, jch, s_d[0], s_d[1]);
                    __syncthreads();
                }
            }
        }
        __syncthreads();
    }
    if (tid == 0) { out[0] = s_i[0]; out[1] = s_i[1]; out[2] = s_i[2]; out[3] = s_i[3]; out[4] = s_i[6]; }
}

// ----------------------------------------------------------------------------------------------
// k_zchase_window: chase a zero diagonal entry of T down through the LDS-resident window
// [p, p + w) of (H, T), w <= 64 (the interior-zero chase of k_inf_defl, DHGEQZ7 / PDHGEQZ8 on an NB
// window, applied by the caller through PDHGEQZ6's updates).  T(p + 1, p + 1) = 0 on entry; each
// step j = 1 .. w - 2 (window rows) moves it from j to j + 1 by the row rotation of rows (j, j + 1)
// that zeroes T(j + 1, j + 1), then removes the fill H(j + 1, j - 1) with the column rotation of
// columns (j, j - 1); on exit the zero sits at T(p + w - 1, p + w - 1).  The rotations are those of
// k_inf_defl's chase in the same order; lane 0 forms both and applies them to the 2x2 corner of H
// they share, the workgroup applies them to the rest of the window and accumulates U (row
// rotations, as columns) and V (column rotations); the caller applies U to the rows of the window
// right of it and to Q, V to the columns above it and to Z.  U, V: w x w, ld w.
// ----------------------------------------------------------------------------------------------
#define ZC_MAXW 64
#define ZC_NT 256
extern "C" __global__ void __launch_bounds__(ZC_NT) k_zchase_window(double *H, double *T, int ld, int p, int w,
                                                                     double *U, double *V) {
    extern __shared__ double lds[];
    __shared__ double rot[4];
    const int tid = threadIdx.x, nt = blockDim.x, lw = w + 1;
    double *sH = lds, *sT = sH + (size_t)lw * w, *sU = sT + (size_t)lw * w, *sV = sU + (size_t)lw * w;
    const double *gH = H + GQZ_IX(p, p, ld), *gT = T + GQZ_IX(p, p, ld);
    for (int idx = tid; idx < w * w; idx += nt) {
        const int i = idx % w, j = idx / w;
        sH[i + lw * j] = gH[GQZ_IX(i, j, ld)];
        sT[i + lw * j] = gT[GQZ_IX(i, j, ld)];
        sU[i + lw * j] = (i == j) ? 1.0 : 0.0;
        sV[i + lw * j] = (i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int j = 1; j <= w - 2; ++j) {
        if (tid == 0) {
            double c, s, r, c2, s2, r2;
            d_lartg(sT[j + lw * (j + 1)], sT[(j + 1) + lw * (j + 1)], c, s, r);
            sT[j + lw * (j + 1)] = r; sT[(j + 1) + lw * (j + 1)] = 0.0;
            // corner rows (j, j+1) x columns (j-1, j) of H: row rotation, then the column rotation
            // formed from it (x' = c x + s y, y' = c y - s x on rows; on columns x = column j)
            const double ha = sH[j + lw * (j - 1)], hb = sH[(j + 1) + lw * (j - 1)];
            const double hc = sH[j + lw * j], hd = sH[(j + 1) + lw * j];
            const double a1 = c * ha + s * hb, b1 = c * hb - s * ha;
            const double c1 = c * hc + s * hd, d1 = c * hd - s * hc;
            d_lartg(d1, b1, c2, s2, r2);
            sH[(j + 1) + lw * j] = r2; sH[(j + 1) + lw * (j - 1)] = 0.0;
            sH[j + lw * j] = c2 * c1 + s2 * a1; sH[j + lw * (j - 1)] = c2 * a1 - s2 * c1;
            rot[0] = c; rot[1] = s; rot[2] = c2; rot[3] = s2;
        }
        __syncthreads();
        const double c = rot[0], s = rot[1], c2 = rot[2], s2 = rot[3];
        // rows (j, j+1): H columns j+1.., T columns j+2.., U columns (j, j+1); columns (j, j-1): H
        // and T rows 0..j-1, V columns (j, j-1)
        const int nlh = w - (j + 1), nlt = w - (j + 2), nr = j;
        const int tot = nlh + nlt + w + 2 * nr + w;
        for (int idx = tid; idx < tot; idx += nt) {
            int t = idx;
            if (t < nlh + nlt) {
                double *M = t < nlh ? sH : sT;
                const int col = t < nlh ? j + 1 + t : j + 2 + (t - nlh);
                const double x = M[j + lw * col], y = M[(j + 1) + lw * col];
                M[j + lw * col] = c * x + s * y; M[(j + 1) + lw * col] = c * y - s * x;
                continue;
            }
            t -= nlh + nlt;
            if (t < w) {
                const double x = sU[t + lw * j], y = sU[t + lw * (j + 1)];
                sU[t + lw * j] = c * x + s * y; sU[t + lw * (j + 1)] = c * y - s * x;
                continue;
            }
            t -= w;
            double *M = (t < nr) ? sH : (t < 2 * nr ? sT : sV);
            const int row = (t < nr) ? t : (t < 2 * nr ? t - nr : t - 2 * nr);
            const double x = M[row + lw * j], y = M[row + lw * (j - 1)];
            M[row + lw * j] = c2 * x + s2 * y; M[row + lw * (j - 1)] = c2 * y - s2 * x;
        }
        __syncthreads();
    }
    double *oH = H + GQZ_IX(p, p, ld), *oT = T + GQZ_IX(p, p, ld);
    for (int idx = tid; idx < w * w; idx += nt) {
        const int i = idx % w, j = idx / w;
        oH[GQZ_IX(i, j, ld)] = sH[i + lw * j];
        oT[GQZ_IX(i, j, ld)] = sT[i + lw * j];
        U[i + (size_t)w * j] = sU[i + lw * j];
        V[i + (size_t)w * j] = sV[i + lw * j];
    }
}

// copy the jw x jw window at (k0, k0) into contiguous Hw/Tw (zero below the Hessenberg /
// triangular structure) and set U = V = I
extern "C" __global__ void k_window_in(const double *H, const double *T, int ld, int k0, int jw,
                                       double *Hw, double *Tw, double *U, double *V) {
    size_t total = (size_t)jw * jw;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(idx % jw), j = (int)(idx / jw);
        Hw[idx] = (i <= j + 1) ? H[GQZ_IX(k0 + i, k0 + j, ld)] : 0.0;
        Tw[idx] = (i <= j) ? T[GQZ_IX(k0 + i, k0 + j, ld)] : 0.0;
        U[idx] = (i == j) ? 1.0 : 0.0;
        V[idx] = (i == j) ? 1.0 : 0.0;
    }
}

// write the processed window back and set the new spike H(k0, k0-1) <- H(k0, k0-1) * U(0,0)
extern "C" __global__ void k_window_out(double *H, double *T, int ld, int k0, int jw,
                                        const double *Hw, const double *Tw, const double *U,
                                        int has_spike) {
    size_t total = (size_t)jw * jw;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(idx % jw), j = (int)(idx / jw);
        H[GQZ_IX(k0 + i, k0 + j, ld)] = Hw[idx];
        T[GQZ_IX(k0 + i, k0 + j, ld)] = Tw[idx];
        // the spike value is read and rewritten by the same thread (i = 0); has_spike = 2: the
        // whole window deflated, the spike is dropped (PDHGEQZ3.f:831, :1346: SCAL = 0 when NS = 0)
        if (has_spike && j == 0) {
            if (i == 0) H[GQZ_IX(k0, k0 - 1, ld)] = has_spike == 1 ? H[GQZ_IX(k0, k0 - 1, ld)] * U[0] : 0.0;
            else H[GQZ_IX(k0 + i, k0 - 1, ld)] = 0.0;
        }
    }
}

extern "C" __global__ void k_identity(double *A, int ld, int m, int n) {
    size_t total = (size_t)m * n;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(idx % m), j = (int)(idx / m);
        A[GQZ_IX(i, j, ld)] = (i == j) ? 1.0 : 0.0;
    }
}

extern "C" __global__ void k_fill(double *x, int n, double v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

// Final pass over the Schur form (PDHGEQZ.f:51-57 output convention): one thread per column.
// A 1x1 block with T(j,j) < 0 is negated (column j of S and T, rows 0..j, and of Z); the
// eigenvalues of every block are read off the diagonal blocks.
extern "C" __global__ void k_finalize(double *H, double *T, double *Z, int ld, int ldz, int nz,
                                      int ilo, int ihi, double *ar, double *ai, double *be) {
    const int j = ilo + blockIdx.x;
    if (j > ihi) return;
    const bool lower2 = (j > ilo) && H[GQZ_IX(j, j - 1, ld)] != 0.0;   // second column of a 2x2
    const bool upper2 = (j < ihi) && H[GQZ_IX(j + 1, j, ld)] != 0.0;   // first column of a 2x2
    if (lower2) return;
    if (upper2) {
        if (threadIdx.x == 0) {
            double A[4] = { H[GQZ_IX(j, j, ld)], H[GQZ_IX(j + 1, j, ld)], H[GQZ_IX(j, j + 1, ld)], H[GQZ_IX(j + 1, j + 1, ld)] };
            double B[4] = { T[GQZ_IX(j, j, ld)], 0.0, T[GQZ_IX(j, j + 1, ld)], T[GQZ_IX(j + 1, j + 1, ld)] };
            double a2[2], i2[2], b2[2];
            d_pair_eigs(A, B, a2, i2, b2);
            ar[j] = a2[0]; ai[j] = i2[0]; be[j] = b2[0];
            ar[j + 1] = a2[1]; ai[j + 1] = i2[1]; be[j + 1] = b2[1];
        }
        return;
    }
    const double t = T[GQZ_IX(j, j, ld)];
    if (t < 0.0) {
        for (int i = threadIdx.x; i <= j; i += blockDim.x) {
            H[GQZ_IX(i, j, ld)] = -H[GQZ_IX(i, j, ld)];
            T[GQZ_IX(i, j, ld)] = -T[GQZ_IX(i, j, ld)];
        }
        if (Z) for (int i = threadIdx.x; i < nz; i += blockDim.x) Z[GQZ_IX(i, j, ldz)] = -Z[GQZ_IX(i, j, ldz)];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ar[j] = H[GQZ_IX(j, j, ld)];     // already negated above when t < 0
        ai[j] = 0.0;
        be[j] = fabs(t);
    }
}

// ---- host launchers ---------------------------------------------------------------------------
static int nblk(size_t total) { size_t b = (total + 255) / 256; return (int)(b > 4096 ? 4096 : (b < 1 ? 1 : b)); }

void gqz_launch_inf_defl(hipStream_t st, double *H, double *T, double *Q, double *Z, int ld, int ldq,
                         int n, int nq, int ilo, int kbot, double ttol, double smlnum, double *ar,
                         double *ai, double *be, int *out, int kt0, int chase) {
    hipLaunchKernelGGL(k_inf_defl, dim3(1), dim3(1024), 0, st, H, T, Q, Z, ld, ldq, n, nq, ilo, kbot,
                       ttol, smlnum, ar, ai, be, out, kt0, chase);
}
int gqz_zchase_max_window() { return ZC_MAXW; }
void gqz_launch_zchase_window(hipStream_t st, double *H, double *T, int ld, int p, int w, double *U, double *V) {
    const size_t lds = sizeof(double) * 4 * (size_t)(w + 1) * w;
    static size_t lds_set[64] = {};
    gqz_need_lds((const void *)k_zchase_window, lds, lds_set);
    hipLaunchKernelGGL(k_zchase_window, dim3(1), dim3(ZC_NT), lds, st, H, T, ld, p, w, U, V);
}
void gqz_launch_window_in(hipStream_t st, const double *H, const double *T, int ld, int k0, int jw,
                          double *Hw, double *Tw, double *U, double *V) {
    hipLaunchKernelGGL(k_window_in, dim3(nblk((size_t)jw * jw)), dim3(256), 0, st, H, T, ld, k0, jw, Hw, Tw, U, V);
}
void gqz_launch_window_out(hipStream_t st, double *H, double *T, int ld, int k0, int jw, const double *Hw,
                           const double *Tw, const double *U, int has_spike) {
    hipLaunchKernelGGL(k_window_out, dim3(nblk((size_t)jw * jw)), dim3(256), 0, st, H, T, ld, k0, jw, Hw, Tw, U,
                       has_spike);
}
void gqz_launch_identity(hipStream_t st, double *A, int ld, int m, int n) {
    hipLaunchKernelGGL(k_identity, dim3(nblk((size_t)m * n)), dim3(256), 0, st, A, ld, m, n);
}
void gqz_launch_fill(hipStream_t st, double *x, int n, double v) {
    hipLaunchKernelGGL(k_fill, dim3(nblk((size_t)n)), dim3(256), 0, st, x, n, v);
}
void gqz_launch_finalize(hipStream_t st, double *H, double *T, double *Z, int ld, int ldz, int nz, int ilo,
                         int ihi, double *ar, double *ai, double *be) {
    if (ihi < ilo) return;
    hipLaunchKernelGGL(k_finalize, dim3(ihi - ilo + 1), dim3(256), 0, st, H, T, Z, ld, ldz, nz, ilo, ihi, ar, ai, be);
}
void gqz_launch_small(hipStream_t st, const GqzSmallJob *djobs, int njobs, int mmax) {
    size_t lds = gqz_small_lds_bytes(mmax);
    hipFuncSetAttribute((const void *)k_small_qz, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_small_qz, dim3(njobs), dim3(256), lds, st, djobs);
}
size_t gqz_small_lds_bytes(int m) { return sizeof(double) * 4 * (size_t)(m + 1) * m + sizeof(WgCtl) + 64; }
void gqz_launch_small1(hipStream_t st, const GqzSmallJob &job) {
    size_t lds = gqz_small_lds_bytes(job.m);
    hipFuncSetAttribute((const void *)k_small_qz1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_small_qz1, dim3(1), dim3(256), lds, st, job);
}
void gqz_launch_gghrd_wg(hipStream_t st, double *A, double *B, int ld, int n, int ilo, int ihi, double *Q,
                         double *Z, int ldq, int nq) {
    hipLaunchKernelGGL(k_gghrd_wg, dim3(1), dim3(1024), 0, st, A, B, ld, n, ilo, ihi, Q, Z, ldq, nq);
}
