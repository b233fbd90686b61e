// qz_desc.cpp -- the ScaLAPACK-descriptor surface used by the Fortran drop-in (1x1 grid).
//
// gqz_pdhgeqz_host mirrors PDHGEQZ's argument checks and INFO codes (PDHGEQZ.f:384-444):
//   -1 JOB, -2 COMPQ, -3 COMPZ, -4 N, -5 ILO, -6 IHI, -7 NB < 5, -8 MB != NB,
//   -9 LDWORK too small, -10 LIWORK too small; LDWORK = -1 or LIWORK = -1 is a workspace query
//   returning the optimum in DWORK(1) and IWORK(1) (PDHGEQZ.f:409-430).  The reported optimum is
//   that of PDHGEQZ1 (N*NB doubles, 2N integers, PDHGEQZ1.f:369-372): the device workspace is
//   internal, so the caller's DWORK/IWORK are only checked for size.
// gqz_pdgghrd_host mirrors PDGGHRD (PDGGHRD.f:13-14, 256-336): LWORK = -1 returns DESCA(MB_) in
//   WORK(1) before the checks; then -1 COMPQ, -2 COMPZ, -3 N, -4 ILO < 1, -5 IHI outside
//   [ILO-1, N], -6 MB != NB, -7 LWORK < MB.  An empty range (IHI < ILO + 2) reduces nothing.  Its Q follows the routine's
//   implemented convention (rotations applied to Q from the left, so the drivers transpose Q
//   afterwards: testing/EXRAND2.f:346-356): COMPQ='I' returns Q_ht^T and COMPQ='V' returns
//   Q_ht^T * Q1, where A_out = Q_ht^T A_in Z.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>

#include "../../include/gqz.h"
#include "gqz_kernels.h"
#include "gqz_misc.h"

namespace {
enum { DTYPE_ = 0, CTXT_, M_, N_, MB_, NB_, RSRC_, CSRC_, LLD_ };
int decode(char c, const char *opts) {
    char u = (char)toupper((unsigned char)c);
    for (int i = 0; opts[i]; ++i) if (opts[i] == u) return i;
    return -1;
}
bool desc_1x1_ok(const int *d, int n) {
    return d && d[RSRC_] == 0 && d[CSRC_] == 0 && d[LLD_] >= std::max(1, n);
}
}  // namespace

extern "C" {

int gqz_gghrd_rowq(char compq, char compz, int n, int ilo, int ihi, double *A, int lda, double *B, int ldb,
                   double *Q, int ldq, double *Z, int ldz);

int gqz_pdhgeqz_host(char job, char compq, char compz, int n, int ilo, int ihi, double *H, const int *desch,
                     double *T, const int *desct, double *alphar, double *alphai, double *beta, double *Q,
                     const int *descq, double *Z, const int *descz, double *dwork, int ldwork, int *iwork,
                     int liwork) {
    int ijob = decode(job, "ES"), icq = decode(compq, "NVI"), icz = decode(compz, "NVI");
    if (ijob < 0) return -1;
    if (icq < 0) return -2;
    if (icz < 0) return -3;
    if (n < 0) return -4;
    if (ilo < 1) return -5;
    if (ihi > n || ihi < ilo - 1) return -6;
    if (!desch || desch[NB_] < 5) return -7;
    if (desch[NB_] != desch[MB_]) return -8;
    const int nb = desch[NB_];
    const int ldwkopt = std::max(1, n * nb), liwkopt = std::max(1, 2 * n);
    if (ldwork == -1 || liwork == -1) {
        if (dwork) dwork[0] = (double)ldwkopt;
        if (iwork) iwork[0] = liwkopt;
        return 0;
    }
    if (ldwork < ldwkopt) return -9;
    if (liwork < liwkopt) return -10;
    if (!desc_1x1_ok(desch, n) || !desc_1x1_ok(desct, n)) return -8;
    if (icq > 0 && !desc_1x1_ok(descq, n)) return -8;
    if (icz > 0 && !desc_1x1_ok(descz, n)) return -8;
    if (n == 0) return 0;
    const char cq = "NVI"[icq], cz = "NVI"[icz];
    return gqz_hgeqz(job, cq, cz, n, ilo, ihi, H, desch[LLD_], T, desct[LLD_], alphar, alphai, beta,
                     icq ? Q : nullptr, icq ? descq[LLD_] : n, icz ? Z : nullptr, icz ? descz[LLD_] : n,
                     nullptr, nullptr);
}

int gqz_pdgghrd_host(char compq, char compz, int n, int ilo, int ihi, double *A, const int *desca, double *B,
                     const int *descb, double *Q, const int *descq, double *Z, const int *descz, double *work,
                     int lwork) {
    // the workspace query is answered before any argument check (PDGGHRD.f:256-266)
    if (lwork == -1) { if (work && desca) work[0] = desca[MB_]; return 0; }
    int icq = decode(compq, "NVI"), icz = decode(compz, "NVI");
    if (icq < 0) return -1;
    if (icz < 0) return -2;
    if (n < 0) return -3;
    if (ilo < 1) return -4;
    if (ihi > n || ihi < ilo - 1) return -5;
    if (!desca || desca[NB_] != desca[MB_]) return -6;
    if (lwork < desca[MB_]) return -7;
    if (!desc_1x1_ok(desca, n) || !desc_1x1_ok(descb, n)) return -6;
    if (icq > 0 && !desc_1x1_ok(descq, n)) return -6;
    if (icz > 0 && !desc_1x1_ok(descz, n)) return -6;
    if (n == 0) return 0;
    const char cq = "NVI"[icq], cz = "NVI"[icz];
    return gqz_gghrd_rowq(cq, cz, n, ilo, ihi, A, desca[LLD_], B, descb[LLD_], icq ? Q : nullptr,
                          icq ? descq[LLD_] : n, icz ? Z : nullptr, icz ? descz[LLD_] : n);
}

// PDTGORD (src/reorder/PDTGORD.f:1-4) on a 1x1 grid.  WANTQ, WANTZ, SELECT (INTEGER, in/out), PARA(6),
// N, S, DESCS, T, DESCT, Q, DESCQ, Z, DESCZ, ALPHAR, ALPHAI, BETA, M, DWORK, LDWORK, IWORK, LIWORK.
// SELECT is normalised for 2x2 blocks and M counted first (PDTGORD.f:504-527); the workspace sizes
// are the routine's own formulas for a 1x1 grid (PDTGORD.f:548-556: LWMIN = 2N + 14 NB^2 + 8 N PARA(3)
// with LOCr = LOCc = N, LIWMIN = 5 PARA(1) + PARA(2) PARA(3) - PARA(2)(PARA(2)+1)/2);
// LDWORK = -1 or LIWORK = -1 is a query; too small gives INFO = -31 / -33 (the routine's codes,
// PDTGORD.f:567-572).  INFO > 0: the 1-based row where a swap was rejected, the pencil partially
// reordered (and SELECT showing the partial order).  The device workspace is internal.
int gqz_tgord_dev(int wantq, int wantz, int *sel, const int *para, int n, double *S, int lds, double *T, int ldt,
                  double *Q, int ldq, double *Z, int ldz, double *alphar, double *alphai, double *beta, int *m);

int gqz_pdtgord_host(int wantq, int wantz, int *select, const int *para, int n, double *S, const int *descs,
                     double *T, const int *desct, double *Q, const int *descq, double *Z, const int *descz,
                     double *alphar, double *alphai, double *beta, int *m, double *dwork, int ldwork, int *iwork,
                     int liwork) {
    if (!descs || !para) return -7;
    const int nb = descs[MB_];
    const long wrk1 = 2L * n + 14L * nb * nb + 4L * n * para[2] + 2L * n * para[2] + 2L * n * para[2];
    const long iwrk1 = 5L * para[0] + (long)para[1] * para[2] - (long)para[1] * (para[1] + 1) / 2;
    const long lwmin = std::max(1L, wrk1), liwmin = std::max(1L, iwrk1);
    if (ldwork == -1 || liwork == -1) {
        if (dwork) dwork[0] = (double)lwmin;
        if (iwork) iwork[0] = (int)liwmin;
        return 0;
    }
    if (n < 0) return -5;
    if (ldwork < lwmin) return -31;
    if (liwork < liwmin) return -33;
    if (!desc_1x1_ok(descs, n)) return -7;
    if (!desc_1x1_ok(desct, n)) return -9;
    if (wantq && !desc_1x1_ok(descq, n)) return -11;
    if (wantz && !desc_1x1_ok(descz, n)) return -13;
    if (m) *m = 0;
    if (n == 0) return 0;
    if (gqz_init_current()) return -100;
    // device copies (ld = n), run, copy back
    try {
        const size_t nn = (size_t)n * n;
        double *d = nullptr;
        if (hipMalloc((void **)&d, sizeof(double) * (4 * nn + 3 * (size_t)n)) != hipSuccess) {
            gqz_set_error("gqz_pdtgord_host: device allocation failed");
            return -100;
        }
        double *dS = d, *dT = d + nn, *dQ = d + 2 * nn, *dZ = d + 3 * nn, *de = d + 4 * nn;
        hipStream_t st = (hipStream_t)gqz_stream();
        auto in = [&](double *dst, const double *src, int ld) {
            return hipMemcpy2DAsync(dst, n * sizeof(double), src, ld * sizeof(double), n * sizeof(double), n,
                                    hipMemcpyHostToDevice, st);
        };
        auto out = [&](double *dst, int ld, const double *src) {
            return hipMemcpy2DAsync(dst, ld * sizeof(double), src, n * sizeof(double), n * sizeof(double), n,
                                    hipMemcpyDeviceToHost, st);
        };
        hipError_t e = in(dS, S, descs[LLD_]);
        if (e == hipSuccess) e = in(dT, T, desct[LLD_]);
        if (e == hipSuccess && wantq) e = in(dQ, Q, descq[LLD_]);
        if (e == hipSuccess && wantz) e = in(dZ, Z, descz[LLD_]);
        int info = -100;
        if (e == hipSuccess)
            info = gqz_tgord_dev(wantq, wantz, select, para, n, dS, n, dT, n, wantq ? dQ : nullptr, n,
                                 wantz ? dZ : nullptr, n, de, de + n, de + 2 * n, m);
        if (info >= 0) {
            e = out(S, descs[LLD_], dS);
            if (e == hipSuccess) e = out(T, desct[LLD_], dT);
            if (e == hipSuccess && wantq) e = out(Q, descq[LLD_], dQ);
            if (e == hipSuccess && wantz) e = out(Z, descz[LLD_], dZ);
            if (e == hipSuccess) e = hipMemcpyAsync(alphar, de, n * sizeof(double), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipMemcpyAsync(alphai, de + n, n * sizeof(double), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipMemcpyAsync(beta, de + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost, st);
        }
        const hipError_t e2 = hipStreamSynchronize(st);
        (void)hipFree(d);
        if (e != hipSuccess || e2 != hipSuccess) {
            gqz_set_error("gqz_pdtgord_host: HIP copy failed");
            return -100;
        }
        return info;
    } catch (...) {
        return -100;
    }
}

}  // extern "C"
