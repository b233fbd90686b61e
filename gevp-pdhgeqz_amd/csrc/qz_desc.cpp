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

}  // extern "C"
