/*
 * qz_oracle.c -- CPU restatement of PDHGEQZ's multishift QZ (1x1 process grid) and of the
 * PDGGHRD Hessenberg-triangular reduction.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation, and the "port" CPU
 * baseline.  Nothing in the product path links or calls this file.
 *
 * Every routine names the reference routine (file:line, relative to the NLAFET/GEVP-PDHGEQZ
 * tree) whose semantics it restates.  Column-major, 0-based.
 */
#include "qz_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define IX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))
#define MAXI(a, b) ((a) > (b) ? (a) : (b))
#define MINI(a, b) ((a) < (b) ? (a) : (b))

static const double ULP = DBL_EPSILON;          /* PDLAMCH('P') = eps*base = 2^-52 */
static const double SAFMIN = DBL_MIN;           /* PDLAMCH('S') */

/* ===================================================================================== */
/* generators (specified identically in the product's Python/C++ test helpers)           */
/* ===================================================================================== */
uint64_t orc_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double u53(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

void orc_fill_uniform(uint64_t seed, size_t count, double *out) {
    uint64_t st = seed;
    for (size_t i = 0; i < count; ++i) out[i] = u53(orc_splitmix64(&st));
}

/* Box-Muller on pairs of 53-bit uniforms; u1 is mapped to (0,1]. */
void orc_fill_normal(uint64_t seed, size_t count, double *out) {
    uint64_t st = seed;
    for (size_t i = 0; i < count; i += 2) {
        double u1 = 1.0 - u53(orc_splitmix64(&st));
        double u2 = u53(orc_splitmix64(&st));
        double r = sqrt(-2.0 * log(u1));
        double t = 6.283185307179586476925286766559 * u2;
        out[i] = r * cos(t);
        if (i + 1 < count) out[i + 1] = r * sin(t);
    }
}

/* ===================================================================================== */
/* elementary transforms                                                                 */
/* ===================================================================================== */

/* Plane rotation with [c s; -s c][f; g] = [r; 0] (LAPACK DLARTG semantics). */
void orc_lartg(double f, double g, double *c, double *s, double *r) {
    if (g == 0.0) { *c = 1.0; *s = 0.0; *r = f; return; }
    if (f == 0.0) { *c = 0.0; *s = (g > 0) ? 1.0 : -1.0; *r = fabs(g); return; }
    double d = hypot(f, g);
    double rr = copysign(d, f);
    *c = fabs(f) / d;
    *s = g / rr;
    *r = rr;
}

/* Householder: (I - tau [1;v][1 v^T]) [alpha; x] = [beta; 0] (LAPACK DLARFG semantics). */
void orc_larfg(int n, double *alpha, double *x, int incx, double *tau) {
    if (n <= 1) { *tau = 0.0; return; }
    double xnorm = 0.0, scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n - 1; ++i) {
        double v = fabs(x[(size_t)i * incx]);
        if (v != 0.0) {
            if (scale < v) { ssq = 1.0 + ssq * (scale / v) * (scale / v); scale = v; }
            else ssq += (v / scale) * (v / scale);
        }
    }
    xnorm = scale * sqrt(ssq);
    if (xnorm == 0.0) { *tau = 0.0; return; }
    double beta = -copysign(hypot(*alpha, xnorm), *alpha);
    *tau = (beta - *alpha) / beta;
    double sc = 1.0 / (*alpha - beta);
    for (int i = 0; i < n - 1; ++i) x[(size_t)i * incx] *= sc;
    *alpha = beta;
}

/* rows i1,i2 of M, columns [c0,c1):  x' = c x + s y, y' = c y - s x  (DROT on rows) */
static void rot_rows(double *M, int ld, int i1, int i2, int c0, int c1, double c, double s) {
    for (int j = c0; j < c1; ++j) {
        double x = M[IX(i1, j, ld)], y = M[IX(i2, j, ld)];
        M[IX(i1, j, ld)] = c * x + s * y;
        M[IX(i2, j, ld)] = c * y - s * x;
    }
}
/* columns j1,j2 of M, rows [r0,r1) */
static void rot_cols(double *M, int ld, int j1, int j2, int r0, int r1, double c, double s) {
    double *x = M + IX(0, j1, ld), *y = M + IX(0, j2, ld);
    for (int i = r0; i < r1; ++i) {
        double a = x[i], b = y[i];
        x[i] = c * a + s * b;
        y[i] = c * b - s * a;
    }
}
/* (I - tau v v^T) applied from the left to rows k..k+2, columns [c0,c1); v[0] = 1 */
static void refl3_left(double *M, int ld, int k, int c0, int c1, const double v[3], double tau) {
    if (tau == 0.0) return;
    for (int j = c0; j < c1; ++j) {
        double *p = M + IX(k, j, ld);
        double t = tau * (p[0] + v[1] * p[1] + v[2] * p[2]);
        p[0] -= t; p[1] -= t * v[1]; p[2] -= t * v[2];
    }
}
/* from the right to columns k..k+2, rows [r0,r1) */
static void refl3_right(double *M, int ld, int k, int r0, int r1, const double v[3], double tau) {
    if (tau == 0.0) return;
    double *x = M + IX(0, k, ld), *y = M + IX(0, k + 1, ld), *z = M + IX(0, k + 2, ld);
    for (int i = r0; i < r1; ++i) {
        double t = tau * (x[i] + v[1] * y[i] + v[2] * z[i]);
        x[i] -= t; y[i] -= t * v[1]; z[i] -= t * v[2];
    }
}
/* general Householder of length m from the left on rows k..k+m-1 / right on cols */
static void reflm_left(double *M, int ld, int k, int m, int c0, int c1, const double *v, double tau) {
    if (tau == 0.0) return;
    for (int j = c0; j < c1; ++j) {
        double *p = M + IX(k, j, ld);
        double t = 0.0;
        for (int i = 0; i < m; ++i) t += v[i] * p[i];
        t *= tau;
        for (int i = 0; i < m; ++i) p[i] -= t * v[i];
    }
}
static void reflm_right(double *M, int ld, int k, int m, int r0, int r1, const double *v, double tau) {
    if (tau == 0.0) return;
    for (int i = r0; i < r1; ++i) {
        double t = 0.0;
        for (int j = 0; j < m; ++j) t += M[IX(i, k + j, ld)] * v[j];
        t *= tau;
        for (int j = 0; j < m; ++j) M[IX(i, k + j, ld)] -= t * v[j];
    }
}

/* ===================================================================================== */
/* 2x2 pencils: eigenvalues (DLAG2 role) and standardisation (DLAGV2 role)               */
/* ===================================================================================== */

/* Eigenvalues of (A,B) 2x2, B upper triangular with non-zero diagonal, as (ar, ai, b) with
 * lambda = (ar + i ai)/b.  Computed from M = B^{-1} A with the standard 2x2 formula. */
void orc_eig22(const double *A, int lda, const double *B, int ldb, double ar[2], double ai[2],
               double b[2]) {
    double a00 = A[IX(0, 0, lda)], a01 = A[IX(0, 1, lda)], a10 = A[IX(1, 0, lda)], a11 = A[IX(1, 1, lda)];
    double b00 = B[IX(0, 0, ldb)], b01 = B[IX(0, 1, ldb)], b11 = B[IX(1, 1, ldb)];
    double m10 = a10 / b11, m11 = a11 / b11;
    double m00 = (a00 - b01 * m10) / b00, m01 = (a01 - b01 * m11) / b00;
    double p = 0.5 * (m00 - m11);
    double bcmax = fmax(fabs(m01), fabs(m10));
    double bcmis = fmin(fabs(m01), fabs(m10)) * copysign(1.0, m01) * copysign(1.0, m10);
    double scale = fmax(fabs(p), bcmax);
    double z = (scale == 0.0) ? 0.0 : (p / scale) * p + (bcmax / scale) * bcmis;
    b[0] = b[1] = 1.0;
    if (z >= 0.0) {          /* real */
        double zz = p + copysign(sqrt(scale) * sqrt(z), p);
        double l1 = m11 + zz;
        double l2 = (zz != 0.0) ? m11 - (bcmax / zz) * bcmis : m11;
        ar[0] = l1; ar[1] = l2; ai[0] = ai[1] = 0.0;
    } else {                 /* complex pair */
        double re = 0.5 * (m00 + m11);
        double disc = -((m00 - m11) * (m00 - m11) * 0.25 + m01 * m10);
        double im = sqrt(fmax(disc, 0.0));
        ar[0] = ar[1] = re; ai[0] = im; ai[1] = -im;
    }
}

/* [x; y] <- G^T [x; y] on rows i,i+1 (columns [c0,c1)) and [x y] <- [x y] G on columns */
static void g2_rows(double *M, int ld, int i, int c0, int c1, const double G[4]) {
    for (int j = c0; j < c1; ++j) {
        double x = M[IX(i, j, ld)], y = M[IX(i + 1, j, ld)];
        M[IX(i, j, ld)] = G[0] * x + G[1] * y;
        M[IX(i + 1, j, ld)] = G[2] * x + G[3] * y;
    }
}
static void g2_cols(double *M, int ld, int j, int r0, int r1, const double G[4]) {
    double *x = M + IX(0, j, ld), *y = M + IX(0, j + 1, ld);
    for (int i = r0; i < r1; ++i) {
        double a = x[i], b = y[i];
        x[i] = a * G[0] + b * G[1];
        y[i] = a * G[2] + b * G[3];
    }
}

/* Standardise the 2x2 pencil (DLAGV2 role, used at PDHGEQZ0.f:622-674 and
 * PDHGEQZ1.f:686-725).  On exit A <- QL^T A ZR, B <- QL^T B ZR with QL, ZR orthogonal 2x2
 * (column-major).  Real eigenvalues: A and B upper triangular.  Complex pair: B diagonal
 * with positive entries (the right factor of B's SVD, obtained by one Jacobi rotation of
 * B^T B, and the left factor from the normalised columns of B ZR). */
#define orc_lartg_v(f, g, c, s, r) orc_lartg((f), (g), &(c), &(s), &(r))
void orc_lagv2(double *A, int lda, double *B, int ldb, double alphar[2], double alphai[2],
               double beta[2], double QL[4], double ZR[4]) {
    double ar[2], ai[2], bb[2];
    orc_eig22(A, lda, B, ldb, ar, ai, bb);
    double cl, sl, cr, sr, r;
    double anorm = fabs(A[IX(0, 0, lda)]) + fabs(A[IX(1, 0, lda)]) + fabs(A[IX(0, 1, lda)]) + fabs(A[IX(1, 1, lda)]);
    double bnorm = fabs(B[IX(0, 0, ldb)]) + fabs(B[IX(0, 1, ldb)]) + fabs(B[IX(1, 1, ldb)]);
    if (ai[0] == 0.0) {
        double lam = ar[0];
        double h00 = A[IX(0, 0, lda)] - lam * B[IX(0, 0, ldb)];
        double h01 = A[IX(0, 1, lda)] - lam * B[IX(0, 1, ldb)];
        double h10 = A[IX(1, 0, lda)];
        double h11 = A[IX(1, 1, lda)] - lam * B[IX(1, 1, ldb)];
        double rr = fabs(h00) + fabs(h01), qq = fabs(h10) + fabs(h11);
        /* ZR e0 spans the null vector of the dominant row of A - lambda B */
        if (rr >= qq) orc_lartg(h01, -h00, &cr, &sr, &r);
        else orc_lartg(h11, -h10, &cr, &sr, &r);
        ZR[0] = cr; ZR[1] = sr; ZR[2] = -sr; ZR[3] = cr;
        g2_cols(A, lda, 0, 0, 2, ZR);
        g2_cols(B, ldb, 0, 0, 2, ZR);
        /* left rotation zeroing the (1,0) entries, from the relatively larger column */
        double na = fabs(A[IX(0, 0, lda)]) + fabs(A[IX(1, 0, lda)]);
        double nb = fabs(B[IX(0, 0, ldb)]) + fabs(B[IX(1, 0, ldb)]);
        if (na * bnorm >= nb * anorm)
            orc_lartg(A[IX(0, 0, lda)], A[IX(1, 0, lda)], &cl, &sl, &r);
        else
            orc_lartg(B[IX(0, 0, ldb)], B[IX(1, 0, ldb)], &cl, &sl, &r);
        QL[0] = cl; QL[1] = sl; QL[2] = -sl; QL[3] = cl;
        g2_rows(A, lda, 0, 0, 2, QL);
        g2_rows(B, ldb, 0, 0, 2, QL);
        A[IX(1, 0, lda)] = 0.0;
        B[IX(1, 0, ldb)] = 0.0;
        alphar[0] = A[IX(0, 0, lda)]; alphar[1] = A[IX(1, 1, lda)];
        alphai[0] = alphai[1] = 0.0;
        beta[0] = B[IX(0, 0, ldb)]; beta[1] = B[IX(1, 1, ldb)];
        return;
    }
    double f = B[IX(0, 0, ldb)], g = B[IX(0, 1, ldb)], h = B[IX(1, 1, ldb)];
    /* Jacobi rotation diagonalising B^T B = [f^2, fg; fg, g^2+h^2] */
    double p = f * f, q = f * g, w = g * g + h * h;
    double c2 = 1.0, s2 = 0.0;
    if (q != 0.0) {
        double theta = 0.5 * (w - p) / q;
        double t = copysign(1.0, theta) / (fabs(theta) + sqrt(1.0 + theta * theta));
        { double rr; orc_lartg(1.0, t, &c2, &s2, &rr); }
    }
    ZR[0] = c2; ZR[1] = -s2; ZR[2] = s2; ZR[3] = c2;
    g2_cols(A, lda, 0, 0, 2, ZR);
    g2_cols(B, ldb, 0, 0, 2, ZR);
    double b0x = B[IX(0, 0, ldb)], b0y = B[IX(1, 0, ldb)], b1x = B[IX(0, 1, ldb)], b1y = B[IX(1, 1, ldb)];
    double s0 = hypot(b0x, b0y), s1 = hypot(b1x, b1y);
    /* B ZR has orthogonal columns b0, b1.  QL^T is the plane rotation (or, when needed to keep both
     * diagonal entries positive, reflection) that maps the longer of them onto a coordinate axis.
     * (Normalising both columns separately, Gram-Schmidt style, leaves an orthogonality defect of
     * one sign that accumulates linearly over many standardisations.) */
    double c, s;
    double m00, m01, m10, m11;
    if (s0 >= s1) {
        orc_lartg_v(b0x, b0y, c, s, r);
        if (r < 0.0) { c = -c; s = -s; }
        m00 = c; m01 = s;
        if (-s * b1x + c * b1y >= 0.0) { m10 = -s; m11 = c; }   /* rotation */
        else { m10 = s; m11 = -c; }                             /* reflection */
    } else {
        orc_lartg_v(b1y, -b1x, c, s, r);
        if (r < 0.0) { c = -c; s = -s; }
        m10 = -s; m11 = c;
        if (c * b0x + s * b0y >= 0.0) { m00 = c; m01 = s; }
        else { m00 = -c; m01 = -s; }
    }
    QL[0] = m00; QL[1] = m01; QL[2] = m10; QL[3] = m11;
    g2_rows(A, lda, 0, 0, 2, QL);
    g2_rows(B, ldb, 0, 0, 2, QL);
    B[IX(1, 0, ldb)] = 0.0; B[IX(0, 1, ldb)] = 0.0;
    double e_ar[2], e_ai[2], e_b[2];
    orc_eig22(A, lda, B, ldb, e_ar, e_ai, e_b);
    double b0 = B[IX(0, 0, ldb)], b1 = B[IX(1, 1, ldb)];
    if (e_ai[0] != 0.0) {
        alphar[0] = e_ar[0] * b0; alphai[0] = fabs(e_ai[0]) * b0; beta[0] = b0;
        alphar[1] = e_ar[0] * b1; alphai[1] = -fabs(e_ai[0]) * b1; beta[1] = b1;
    } else {
        alphar[0] = e_ar[0] * b0; alphai[0] = 0.0; beta[0] = b0;
        alphar[1] = e_ar[1] * b1; alphai[1] = 0.0; beta[1] = b1;
    }
}

/* ===================================================================================== */
/* shift polynomial first column (QZFCOL role, src/kkqz/qzfcol.f:1-116)                  */
/* ===================================================================================== */
/* v = [b1 b2 M^2 - c1 M + c0] e1 with M = H T^{-1}, homogeneous in (alpha, beta) so that
 * shifts at infinity stay finite; real pair: c1 = ar1 b2 + ar2 b1, c0 = ar1 ar2;
 * complex pair: b1^2 (M - l)(M - conj l) with l = (ar1 + i ai1)/b1. */
void orc_shift_col(const double *H, int ldh, const double *T, int ldt,
                   double sr1, double si1, double sb1, double sr2, double si2, double sb2,
                   double v[3]) {
    double h00 = H[IX(0, 0, ldh)], h01 = H[IX(0, 1, ldh)], h10 = H[IX(1, 0, ldh)];
    double h11 = H[IX(1, 1, ldh)], h21 = H[IX(2, 1, ldh)];
    double t00 = T[IX(0, 0, ldt)], t01 = T[IX(0, 1, ldt)], t11 = T[IX(1, 1, ldt)];
    double c2, c1, c0;
    if (si1 != 0.0) {
        c2 = sb1 * sb1; c1 = 2.0 * sb1 * sr1; c0 = sr1 * sr1 + si1 * si1;
    } else {
        c2 = sb1 * sb2; c1 = sr1 * sb2 + sr2 * sb1; c0 = sr1 * sr2;
    }
    (void)si2;
    /* x = M e1 */
    double x0 = h00 / t00, x1 = h10 / t00;
    /* y = M x : solve T z = x (2x2 upper), y = H z */
    double z1 = x1 / t11;
    double z0 = (x0 - t01 * z1) / t00;
    double y0 = h00 * z0 + h01 * z1, y1 = h10 * z0 + h11 * z1, y2 = h21 * z1;
    v[0] = c2 * y0 - c1 * x0 + c0;
    v[1] = c2 * y1 - c1 * x1;
    v[2] = c2 * y2;
    double m = fmax(fabs(v[0]), fmax(fabs(v[1]), fabs(v[2])));
    if (m > 0.0 && isfinite(m)) { v[0] /= m; v[1] /= m; v[2] /= m; }
}

/* INVHSE with ITOPP=2 (src/kkqz/invhse.f:61-71): reflector whose first column spans the
 * first row of the orthogonal factor of an RQ decomposition of the 3x3 T, i.e. the vector
 * orthogonal to rows 1,2 of T. */
void orc_invhse3(const double *T, int ldt, double vout[3], double *tau) {
    double t[9];
    for (int j = 0; j < 3; ++j) for (int i = 0; i < 3; ++i) t[i + 3 * j] = T[IX(i, j, ldt)];
    /* H1 maps row 2 onto e2: reflector over (col2, col0, col1) ordering */
    double a = t[2 + 3 * 2];
    double x[2] = { t[2 + 3 * 0], t[2 + 3 * 1] };
    double tau1;
    orc_larfg(3, &a, x, 1, &tau1);
    /* u1 = (x0, x1, 1) in positions (0,1,2) */
    double u1[3] = { x[0], x[1], 1.0 };
    /* t <- t * (I - tau1 u1 u1^T) */
    for (int i = 0; i < 3; ++i) {
        double s = tau1 * (t[i] * u1[0] + t[i + 3] * u1[1] + t[i + 6] * u1[2]);
        t[i] -= s * u1[0]; t[i + 3] -= s * u1[1]; t[i + 6] -= s * u1[2];
    }
    /* H2 maps row 1 (cols 0,1) onto e1 */
    double a2 = t[1 + 3 * 1];
    double x2[1] = { t[1 + 3 * 0] };
    double tau2;
    orc_larfg(2, &a2, x2, 1, &tau2);
    double u2[3] = { x2[0], 1.0, 0.0 };
    /* first column of Z = H1 H2 e0 */
    double w[3] = { 1.0, 0.0, 0.0 };
    double s2 = tau2 * (w[0] * u2[0] + w[1] * u2[1]);
    w[0] -= s2 * u2[0]; w[1] -= s2 * u2[1];
    double s1 = tau1 * (w[0] * u1[0] + w[1] * u1[1] + w[2] * u1[2]);
    w[0] -= s1 * u1[0]; w[1] -= s1 * u1[1]; w[2] -= s1 * u1[2];
    /* w is a unit vector: reflect it onto -sign(w0) e0 with v = w + sign(w0) e0 scaled to v0 = 1 and
     * tau = 2 / (v^T v) taken from v itself.  (DLARFG on a unit vector rounds its norm to exactly 1 in a
     * biased way; the accumulated Z then loses orthogonality coherently, about 3x faster than Q.) */
    double v0 = w[0] + (w[0] >= 0.0 ? 1.0 : -1.0);
    double v1 = w[1] / v0, v2 = w[2] / v0;
    *tau = 2.0 / (1.0 + v1 * v1 + v2 * v2);
    vout[0] = 1.0; vout[1] = v1; vout[2] = v2;
}

/* ===================================================================================== */
/* context                                                                               */
/* ===================================================================================== */
typedef struct {
    int n;              /* full order (rows of Q and Z, columns of H and T) */
    double *H, *T, *Q, *Z;
    int ldh, ldt, ldq, ldz;
    double smlnum, ttol;
    double *alphar, *alphai, *beta;
    orc_params prm;
    orc_stats *st;
    int level;
} qzctx;

#define Hc(i, j) c->H[IX(i, j, c->ldh)]
#define Tc(i, j) c->T[IX(i, j, c->ldt)]

static void set_eig(qzctx *c, int j, double ar, double ai, double b) {
    c->alphar[j] = ar; c->alphai[j] = ai; c->beta[j] = b;
}

/* 1x1 block at j: normalise beta >= 0 by negating column j of H, T (rows 0..j) and Z. */
static void defl_1x1(qzctx *c, int j) {
    if (Tc(j, j) < 0.0) {
        for (int i = 0; i <= j; ++i) { Hc(i, j) = -Hc(i, j); Tc(i, j) = -Tc(i, j); }
        if (c->Z) for (int i = 0; i < c->n; ++i) c->Z[IX(i, j, c->ldz)] = -c->Z[IX(i, j, c->ldz)];
    }
    set_eig(c, j, Hc(j, j), 0.0, Tc(j, j));
}

/* standardise the 2x2 block at j (DLAGV2 + PDROT, PDHGEQZ1.f:686-725); returns 1 when the
 * block is a complex pair, 0 when it split into two 1x1 blocks. */
static int defl_2x2(qzctx *c, int j) {
    double A[4] = { Hc(j, j), Hc(j + 1, j), Hc(j, j + 1), Hc(j + 1, j + 1) };
    double B[4] = { Tc(j, j), 0.0, Tc(j, j + 1), Tc(j + 1, j + 1) };
    double ar[2], ai[2], bt[2], QL[4], ZR[4];
    orc_lagv2(A, 2, B, 2, ar, ai, bt, QL, ZR);
    g2_rows(c->H, c->ldh, j, j + 2, c->n, QL);
    g2_rows(c->T, c->ldt, j, j + 2, c->n, QL);
    g2_cols(c->H, c->ldh, j, 0, j, ZR);
    g2_cols(c->T, c->ldt, j, 0, j, ZR);
    if (c->Q) g2_cols(c->Q, c->ldq, j, 0, c->n, QL);
    if (c->Z) g2_cols(c->Z, c->ldz, j, 0, c->n, ZR);
    Hc(j, j) = A[0]; Hc(j + 1, j) = A[1]; Hc(j, j + 1) = A[2]; Hc(j + 1, j + 1) = A[3];
    Tc(j, j) = B[0]; Tc(j + 1, j) = 0.0; Tc(j, j + 1) = B[2]; Tc(j + 1, j + 1) = B[3];
    if (ai[0] != 0.0) {
        set_eig(c, j, ar[0], ai[0], bt[0]);
        set_eig(c, j + 1, ar[1], ai[1], bt[1]);
        return 1;
    }
    Hc(j + 1, j) = 0.0;
    defl_1x1(c, j);
    defl_1x1(c, j + 1);
    return 0;
}

/* ===================================================================================== */
/* infinite eigenvalues (PDHGEQZ7.f:50-330, PDHGEQZ8/9, DHGEQZ7)                        */
/* ===================================================================================== */
/* Zeros on diag(T) within [*ktop,*kbot] (|T_ii| <= max(TTOL, SMLNUM), PDHGEQZ7.f:380) are
 * set to zero; a zero at the bottom is deflated by a column rotation, at the top by a row
 * rotation (beta = 0 exactly); interior zeros are chased down to the bottom first
 * (PDHGEQZ8 direction; the upward PDHGEQZ9 chase is not used).  Returns the count. */
static int inf_deflate(qzctx *c, int *ktop, int *kbot) {
    int nd = 0;
    double tol = fmax(c->ttol, c->smlnum);
    for (;;) {
        int kt = *ktop, kb = *kbot;
        if (kb < kt) break;
        int found = -1;
        for (int i = kb; i >= kt; --i)
            if (fabs(Tc(i, i)) <= tol) { Tc(i, i) = 0.0; if (found < 0) found = i; }
        if (found < 0) break;
        if (kb == kt) {     /* 1x1 with zero T */
            set_eig(c, kb, Hc(kb, kb), 0.0, 0.0);
            (*kbot)--; nd++; continue;
        }
        if (Tc(kb, kb) == 0.0) {   /* bottom: column rotation zeroing H(kb,kb-1) */
            double cs, sn, r;
            orc_lartg(Hc(kb, kb), Hc(kb, kb - 1), &cs, &sn, &r);
            Hc(kb, kb) = r; Hc(kb, kb - 1) = 0.0;
            rot_cols(c->H, c->ldh, kb, kb - 1, 0, kb, cs, sn);
            rot_cols(c->T, c->ldt, kb, kb - 1, 0, kb, cs, sn);
            if (c->Z) rot_cols(c->Z, c->ldz, kb, kb - 1, 0, c->n, cs, sn);
            set_eig(c, kb, Hc(kb, kb), 0.0, 0.0);
            (*kbot)--; nd++;
            if (c->st) c->st->inf_defl++;
            continue;
        }
        if (Tc(kt, kt) == 0.0) {   /* top: row rotation zeroing H(kt+1,kt) */
            double cs, sn, r;
            orc_lartg(Hc(kt, kt), Hc(kt + 1, kt), &cs, &sn, &r);
            Hc(kt, kt) = r; Hc(kt + 1, kt) = 0.0;
            rot_rows(c->H, c->ldh, kt, kt + 1, kt + 1, c->n, cs, sn);
            rot_rows(c->T, c->ldt, kt, kt + 1, kt + 1, c->n, cs, sn);
            if (c->Q) rot_cols(c->Q, c->ldq, kt, kt + 1, 0, c->n, cs, sn);
            set_eig(c, kt, Hc(kt, kt), 0.0, 0.0);
            (*ktop)++; nd++;
            if (c->st) c->st->inf_defl++;
            continue;
        }
        /* interior zero at j = found (the lowest one): chase it down to kb (DHGEQZ7) */
        int j = found;
        for (int jch = j; jch < kb; ++jch) {
            double cs, sn, r;
            orc_lartg(Tc(jch, jch + 1), Tc(jch + 1, jch + 1), &cs, &sn, &r);
            Tc(jch, jch + 1) = r; Tc(jch + 1, jch + 1) = 0.0;
            rot_rows(c->T, c->ldt, jch, jch + 1, jch + 2, c->n, cs, sn);
            rot_rows(c->H, c->ldh, jch, jch + 1, jch > kt ? jch - 1 : jch, c->n, cs, sn);
            if (c->Q) rot_cols(c->Q, c->ldq, jch, jch + 1, 0, c->n, cs, sn);
            if (jch > kt) {
                orc_lartg(Hc(jch + 1, jch), Hc(jch + 1, jch - 1), &cs, &sn, &r);
                Hc(jch + 1, jch) = r; Hc(jch + 1, jch - 1) = 0.0;
                rot_cols(c->H, c->ldh, jch, jch - 1, 0, jch + 1, cs, sn);
                rot_cols(c->T, c->ldt, jch, jch - 1, 0, jch, cs, sn);
                if (c->Z) rot_cols(c->Z, c->ldz, jch, jch - 1, 0, c->n, cs, sn);
            }
        }
        /* now T(kb,kb) = 0: the bottom case handles it on the next pass */
    }
    return nd;
}

/* ===================================================================================== */
/* bulge chasing (DHGEQZ5.f:128-275, PDHGEQZA.f:245-484, PDHGEQZB.f:281-491)             */
/* ===================================================================================== */
typedef struct {     /* column/row ranges for the far updates */
    int r0;          /* first row for right transforms (0 for Schur form) */
    int c1;          /* one past the last column for left transforms (n) */
    int ilast;       /* bottom of the active block (rows of H below are zero) */
} chase_rng;

/* DLAQR5-style collapse handling used at DHGEQZ5.f:141-189 */
static void bulge_step(qzctx *c, int k, const chase_rng *g, const double *shift6) {
    double v[3], tau, beta;
    v[0] = Hc(k, k - 1); v[1] = Hc(k + 1, k - 1); v[2] = Hc(k + 2, k - 1);
    double x[2] = { v[1], v[2] };
    beta = v[0];
    orc_larfg(3, &beta, x, 1, &tau);
    v[0] = 1.0; v[1] = x[0]; v[2] = x[1];
    if (Hc(k + 2, k - 1) != 0.0 || Hc(k + 2, k) != 0.0 || Hc(k + 2, k + 1) == 0.0) {
        Hc(k, k - 1) = beta; Hc(k + 1, k - 1) = 0.0; Hc(k + 2, k - 1) = 0.0;
    } else {
        double v1[3], tau1;
        orc_shift_col(&Hc(k, k), c->ldh, &Tc(k, k), c->ldt, shift6[0], shift6[1], shift6[2],
                      shift6[3], shift6[4], shift6[5], v1);
        double a1 = v1[0], x1[2] = { v1[1], v1[2] };
        orc_larfg(3, &a1, x1, 1, &tau1);
        v1[0] = 1.0; v1[1] = x1[0]; v1[2] = x1[1];
        double refsum = tau1 * (Hc(k, k - 1) + v1[1] * Hc(k + 1, k - 1));
        if (fabs(Hc(k + 1, k - 1) - refsum * v1[1]) + fabs(refsum * v1[2]) >
            ULP * (fabs(Hc(k - 1, k - 1)) + fabs(Hc(k, k)) + fabs(Hc(k + 1, k + 1)))) {
            Hc(k, k - 1) = beta; Hc(k + 1, k - 1) = 0.0; Hc(k + 2, k - 1) = 0.0;
        } else {
            Hc(k, k - 1) -= refsum; Hc(k + 1, k - 1) = 0.0; Hc(k + 2, k - 1) = 0.0;
            tau = tau1; v[1] = v1[1]; v[2] = v1[2];
        }
    }
    refl3_left(c->H, c->ldh, k, k, g->c1, v, tau);
    refl3_left(c->T, c->ldt, k, k, g->c1, v, tau);
    if (c->Q) refl3_right(c->Q, c->ldq, k, 0, c->n, v, tau);
    orc_invhse3(&Tc(k, k), c->ldt, v, &tau);
    refl3_right(c->H, c->ldh, k, g->r0, MINI(k + 4, g->ilast + 1), v, tau);
    refl3_right(c->T, c->ldt, k, g->r0, k + 3, v, tau);
    Tc(k + 1, k) = 0.0; Tc(k + 2, k) = 0.0;
    if (c->Z) refl3_right(c->Z, c->ldz, k, 0, c->n, v, tau);
}

/* bulge introduction at k (PDHGEQZA.f:290-330) */
static void bulge_intro(qzctx *c, int k, const chase_rng *g, const double *shift6) {
    double v[3], tau;
    orc_shift_col(&Hc(k, k), c->ldh, &Tc(k, k), c->ldt, shift6[0], shift6[1], shift6[2],
                  shift6[3], shift6[4], shift6[5], v);
    double a = v[0], x[2] = { v[1], v[2] };
    orc_larfg(3, &a, x, 1, &tau);
    v[0] = 1.0; v[1] = x[0]; v[2] = x[1];
    refl3_left(c->H, c->ldh, k, k, g->c1, v, tau);
    refl3_left(c->T, c->ldt, k, k, g->c1, v, tau);
    if (c->Q) refl3_right(c->Q, c->ldq, k, 0, c->n, v, tau);
    orc_invhse3(&Tc(k, k), c->ldt, v, &tau);
    refl3_right(c->H, c->ldh, k, g->r0, MINI(k + 4, g->ilast + 1), v, tau);
    refl3_right(c->T, c->ldt, k, g->r0, k + 3, v, tau);
    Tc(k + 1, k) = 0.0; Tc(k + 2, k) = 0.0;
    if (c->Z) refl3_right(c->Z, c->ldz, k, 0, c->n, v, tau);
}

/* push-off at the bottom (PDHGEQZB.f:368-406): rows ilast-1, ilast */
static void bulge_pushoff(qzctx *c, const chase_rng *g) {
    int l = g->ilast;
    double cs, sn, r;
    orc_lartg(Hc(l - 1, l - 2), Hc(l, l - 2), &cs, &sn, &r);
    Hc(l - 1, l - 2) = r; Hc(l, l - 2) = 0.0;
    rot_rows(c->H, c->ldh, l - 1, l, l - 1, g->c1, cs, sn);
    rot_rows(c->T, c->ldt, l - 1, l, l - 1, g->c1, cs, sn);
    if (c->Q) rot_cols(c->Q, c->ldq, l - 1, l, 0, c->n, cs, sn);
    orc_lartg(Tc(l, l), Tc(l, l - 1), &cs, &sn, &r);
    Tc(l, l) = r; Tc(l, l - 1) = 0.0;
    rot_cols(c->H, c->ldh, l, l - 1, g->r0, l + 1, cs, sn);
    rot_cols(c->T, c->ldt, l, l - 1, g->r0, l, cs, sn);
    if (c->Z) rot_cols(c->Z, c->ldz, l, l - 1, 0, c->n, cs, sn);
}

/* vigilant deflation check over (k0, k1] (DHGEQZ5.f:254-275; Ahues-Tisseur test) */
static void vigilant(qzctx *c, int k0, int k1) {
    for (int j = k0 + 1; j <= k1; ++j) {
        double hj = Hc(j, j - 1);
        if (hj == 0.0) continue;
        double tst1 = fabs(Hc(j - 1, j - 1)) + fabs(Hc(j, j));
        if (fabs(hj) <= fmax(c->smlnum, ULP * tst1)) {
            double h12 = fmax(fabs(hj), fabs(Hc(j - 1, j)));
            double h21 = fmin(fabs(hj), fabs(Hc(j - 1, j)));
            double h11 = fmax(fabs(Hc(j, j)), fabs(Hc(j - 1, j - 1) - Hc(j, j)));
            double h22 = fmin(fabs(Hc(j, j)), fabs(Hc(j - 1, j - 1) - Hc(j, j)));
            double scl = h11 + h12;
            double tst2 = h22 * (h11 / scl);
            if (tst2 == 0.0 || h21 * (h12 / scl) <= fmax(c->smlnum, ULP * tst2)) Hc(j, j - 1) = 0.0;
        }
    }
}

/* Introduction of a chain of nb bulges at ktop (PDHGEQZA.f:245-384): bulge k is made from shift
 * pair (2k, 2k+1) (QZFCOL's first column, the left reflector and INVHSE's right one) and, before the
 * next one enters, every bulge already in the chain (oldest = bottom first) moves 3 steps; a bulge
 * that reaches kbot is pushed off.  tail[k] (bulge k's tail column relative to ktop) and done[k]
 * are outputs. */
static void chain_intro(qzctx *c, int ktop, int kbot, int nb, const double *sr, const double *si,
                        const double *sb, const chase_rng *g, int *tail, int *done) {
    double sh[6];
    for (int k = 0; k < nb; ++k) {
        sh[0] = sr[2 * k]; sh[1] = si[2 * k]; sh[2] = sb[2 * k];
        sh[3] = sr[2 * k + 1]; sh[4] = si[2 * k + 1]; sh[5] = sb[2 * k + 1];
        bulge_intro(c, ktop, g, sh);
        tail[k] = 0;
        done[k] = 0;
        if (k + 1 < nb) {
            /* move existing bulges (bottom = oldest first) 3 steps */
            for (int b = 0; b <= k; ++b) {
                for (int s = 0; s < 3; ++s) {
                    int col = ktop + tail[b];           /* reflector rows col+1..col+3 */
                    if (done[b]) break;
                    if (col + 3 > kbot) {
                        if (col + 2 == kbot) bulge_pushoff(c, g);
                        done[b] = 1; break;
                    }
                    double s6[6] = { sr[2 * b], si[2 * b], sb[2 * b], sr[2 * b + 1], si[2 * b + 1], sb[2 * b + 1] };
                    bulge_step(c, col + 1, g, s6);
                    tail[b]++;
                }
            }
        }
    }
}

/* Kernel-level entry for parity tests: introduce nb bulges at the top of a wsize x wsize window
 * (chain_intro with ktop = 0, kbot = wsize - 1), accumulating the left / right transforms in U, V. */
void orc_intro_window(int wsize, double *Aw, int lda, double *Bw, int ldb, double *U, double *V,
                      int ldu, int nb, const double *sr, const double *si, const double *sb) {
    qzctx cc; memset(&cc, 0, sizeof cc);
    cc.n = wsize; cc.H = Aw; cc.T = Bw; cc.ldh = lda; cc.ldt = ldb;
    cc.Q = U; cc.Z = V; cc.ldq = ldu; cc.ldz = ldu;
    cc.smlnum = SAFMIN * ((double)wsize / ULP);
    chase_rng g = { 0, wsize, wsize - 1 };
    int *tail = (int *)malloc(sizeof(int) * (size_t)(nb > 0 ? nb : 1));
    int *done = (int *)calloc((size_t)(nb > 0 ? nb : 1), sizeof(int));
    chain_intro(&cc, 0, wsize - 1, nb, sr, si, sb, &g, tail, done);
    free(tail); free(done);
}

/* One chain of nb = nsh/2 bulges through [ktop, kbot] (PDHGEQZ5 with NUMWIN = 1): the
 * bulges are introduced at ktop (PDHGEQZA: each new bulge after the previous ones moved 3
 * steps, bottom bulge first), then chased window by window with the bottom bulge moved
 * first (PDHGEQZB), and pushed off at kbot.  shifts: nsh triples (sr, si, sb). */
static void chain_sweep(qzctx *c, int ktop, int kbot, int nsh, const double *sr, const double *si,
                        const double *sb, int movedist) {
    int nb = nsh / 2;
    chase_rng g = { 0, c->n, kbot };
    int m = kbot - ktop + 1;
    if (m < 3) return;
    /* tail column of each bulge relative to ktop; -1 = not yet introduced */
    int *tail = (int *)malloc(sizeof(int) * (size_t)nb);
    int *done = (int *)calloc((size_t)nb, sizeof(int));
    chain_intro(c, ktop, kbot, nb, sr, si, sb, &g, tail, done);
    /* chase: bottom bulge (b = 0, oldest) moved movedist steps first, then the next */
    for (;;) {
        int any = 0;
        for (int b = 0; b < nb; ++b) {
            for (int s = 0; s < movedist; ++s) {
                if (done[b]) break;
                int col = ktop + tail[b];
                if (col + 3 > kbot) {
                    if (col + 2 == kbot) bulge_pushoff(c, &g);
                    done[b] = 1; break;
                }
                double s6[6] = { sr[2 * b], si[2 * b], sb[2 * b], sr[2 * b + 1], si[2 * b + 1], sb[2 * b + 1] };
                bulge_step(c, col + 1, &g, s6);
                tail[b]++;
                any = 1;
            }
        }
        if (!any) break;
        int all = 1;
        for (int b = 0; b < nb; ++b) all &= done[b];
        if (all) break;
    }
    vigilant(c, ktop, kbot);
    free(tail); free(done);
}

/* Kernel-level entry for parity tests: one chain in a window (see header). */
void orc_chase_window(int wsize, double *Aw, int lda, double *Bw, int ldb, double *U, double *V,
                      int ldu, int nb, int movedist, int last, const double *sr,
                      const double *si, const double *sb) {
    qzctx cc; memset(&cc, 0, sizeof cc);
    cc.n = wsize; cc.H = Aw; cc.T = Bw; cc.ldh = lda; cc.ldt = ldb;
    cc.Q = U; cc.Z = V; cc.ldq = ldu; cc.ldz = ldu;
    cc.smlnum = SAFMIN * ((double)wsize / ULP);
    qzctx *c = &cc;
    chase_rng g = { 0, wsize, wsize - 1 };
    /* bulge b (b = 0 top) has tail column 3b; bottom bulge first (PDHGEQZB.f:357-407) */
    for (int b = nb - 1; b >= 0; --b) {
        int tail = 3 * b;
        for (int s = 0; s < movedist; ++s) {
            int col = tail + s;
            if (col + 3 > wsize - 1) {
                if (last && col + 2 == wsize - 1) bulge_pushoff(c, &g);
                break;
            }
            double s6[6] = { sr[2 * b], si[2 * b], sb[2 * b], sr[2 * b + 1], si[2 * b + 1], sb[2 * b + 1] };
            bulge_step(c, col + 1, &g, s6);
        }
    }
}

/* ===================================================================================== */
/* double-shift base QZ (KKQZ / QZLAP role for small blocks, src/kkqz/kkqz.f:320-519)     */
/* ===================================================================================== */
static int qz_ds_ctx(qzctx *c, int ilo, int ihi) {
    int ilast = ihi;
    int maxit = 30 * (ihi - ilo + 1);
    int iiter = 0, jiter = 0;
    while (ilast >= ilo) {
        if (++jiter > maxit + 10) return ilast + 1;
        /* split point */
        int ifirst = ilo;
        for (int j = ilast; j > ilo; --j) {
            double tst = fabs(Hc(j - 1, j - 1)) + fabs(Hc(j, j));
            if (fabs(Hc(j, j - 1)) <= fmax(c->smlnum, ULP * tst)) { Hc(j, j - 1) = 0.0; ifirst = j; break; }
        }
        int kt = ifirst, kb = ilast;
        int ninf = inf_deflate(c, &kt, &kb);
        if (ninf > 0) {
            /* eigenvalues at the top of the block were deflated: the rows [ifirst, kt) are
             * finished; any block above ifirst is still pending */
            if (kb < kt) { ilast = ifirst - 1; }
            else if (kt > ifirst) {
                /* top deflations: the remaining block is [kt, kb]; solve it now by
                 * recursion on the range, then continue above ifirst */
                int info = qz_ds_ctx(c, kt, kb);
                if (info) return info;
                ilast = ifirst - 1;
            } else ilast = kb;
            iiter = 0;
            continue;
        }
        if (ifirst == ilast) { defl_1x1(c, ilast); ilast--; iiter = 0; continue; }
        if (ifirst == ilast - 1) {
            defl_2x2(c, ilast - 1);
            ilast -= 2; iiter = 0; continue;
        }
        /* double-shift sweep on [ifirst, ilast] */
        iiter++;
        if (c->st) c->st->ds_iters++;
        double ar[2], ai[2], bb[2];
        double Ab[4] = { Hc(ilast - 1, ilast - 1), Hc(ilast, ilast - 1), Hc(ilast - 1, ilast), Hc(ilast, ilast) };
        double Bb[4] = { Tc(ilast - 1, ilast - 1), 0.0, Tc(ilast - 1, ilast), Tc(ilast, ilast) };
        orc_eig22(Ab, 2, Bb, 2, ar, ai, bb);
        if (iiter % 10 == 0) {
            /* exceptional shifts (ad hoc, as DHGEQZ's every-10-iterations rule) */
            double e = fabs(Hc(ilast, ilast - 1) / Tc(ilast - 1, ilast - 1)) +
                       fabs(Hc(ilast - 1, ilast - 2) / Tc(ilast - 2, ilast - 2));
            double l = Hc(ilast, ilast) / Tc(ilast, ilast) + 0.75 * e;
            ar[0] = ar[1] = l; ai[0] = ai[1] = 0.0; bb[0] = bb[1] = 1.0;
            if (iiter % 20 == 0) { ar[0] = l; ar[1] = -0.4375 * e + Hc(ilast, ilast) / Tc(ilast, ilast); }
        }
        if (!isfinite(ar[0]) || !isfinite(ar[1]) || !isfinite(ai[0])) {
            ar[0] = ar[1] = 1.0; ai[0] = ai[1] = 0.0; bb[0] = bb[1] = 1.0;
        }
        double sr[2] = { ar[0], ar[1] }, si[2] = { ai[0], ai[1] }, sb[2] = { bb[0], bb[1] };
        chain_sweep(c, ifirst, ilast, 2, sr, si, sb, ilast - ifirst + 1);
    }
    return 0;
}

int orc_qz_ds(int n, int ilo, int ihi, double *H, int ldh, double *T, int ldt,
              double *alphar, double *alphai, double *beta,
              double *Q, int ldq, double *Z, int ldz, double ttol, orc_stats *st) {
    qzctx c; memset(&c, 0, sizeof c);
    c.n = n; c.H = H; c.T = T; c.Q = Q; c.Z = Z; c.ldh = ldh; c.ldt = ldt; c.ldq = ldq; c.ldz = ldz;
    c.smlnum = SAFMIN * ((double)(ihi - ilo + 1) / ULP);
    c.ttol = ttol; c.alphar = alphar; c.alphai = alphai; c.beta = beta; c.st = st;
    return qz_ds_ctx(&c, ilo, ihi);
}

/* ===================================================================================== */
/* block swapping (DTGEX2, src/reorder/dtgex2.f; Kagstrom-Poromaa direct swap)            */
/* ===================================================================================== */

/* Gaussian elimination with complete pivoting on an m x m system (DGETC2/DGESC2 roles):
 * tiny pivots are replaced by smin; returns the scale applied to the right-hand side. */
static double solve_cp(int m, double *K, double *rhs) {
    int jpiv[8];
    double eps = ULP, smlnum = SAFMIN / eps;
    double xmax = 0.0;
    for (int i = 0; i < m * m; ++i) xmax = fmax(xmax, fabs(K[i]));
    double smin = fmax(eps * xmax, smlnum);
    for (int i = 0; i < m; ++i) jpiv[i] = i;
    for (int k = 0; k < m; ++k) {
        int pi = k, pj = k; double pv = -1.0;
        for (int j = k; j < m; ++j) for (int i = k; i < m; ++i)
            if (fabs(K[i + m * j]) > pv) { pv = fabs(K[i + m * j]); pi = i; pj = j; }
        if (pi != k) { for (int j = 0; j < m; ++j) { double t = K[k + m * j]; K[k + m * j] = K[pi + m * j]; K[pi + m * j] = t; }
                       double t = rhs[k]; rhs[k] = rhs[pi]; rhs[pi] = t; }
        if (pj != k) { for (int i = 0; i < m; ++i) { double t = K[i + m * k]; K[i + m * k] = K[i + m * pj]; K[i + m * pj] = t; }
                       int t = jpiv[k]; jpiv[k] = jpiv[pj]; jpiv[pj] = t; }
        if (fabs(K[k + m * k]) < smin) K[k + m * k] = smin;
        for (int i = k + 1; i < m; ++i) {
            double f = K[i + m * k] / K[k + m * k];
            K[i + m * k] = f;
            for (int j = k + 1; j < m; ++j) K[i + m * j] -= f * K[k + m * j];
            rhs[i] -= f * rhs[k];
        }
    }
    /* back substitution with overflow guard */
    double scale = 1.0;
    double bmax = 0.0;
    for (int i = 0; i < m; ++i) bmax = fmax(bmax, fabs(rhs[i]));
    if (2.0 * smlnum * bmax > fabs(K[(m - 1) + m * (m - 1)])) {
        double t = 0.5 / bmax;
        for (int i = 0; i < m; ++i) rhs[i] *= t;
        scale = t;
    }
    double x[8];
    for (int i = m - 1; i >= 0; --i) {
        double s = rhs[i];
        for (int j = i + 1; j < m; ++j) s -= K[i + m * j] * x[j];
        x[i] = s / K[i + m * i];
    }
    for (int i = 0; i < m; ++i) rhs[jpiv[i]] = x[i];
    return scale;
}

/* full m x m orthogonal from the Householder QR of an m x k matrix X (m <= 4) */
static void qr_full(int m, int k, double *X, int ldx, double *Qo /* m x m */) {
    double vs[4][4], taus[4];
    for (int j = 0; j < k; ++j) {
        double a = X[j + ldx * j];
        double x[4];
        for (int i = j + 1; i < m; ++i) x[i - j - 1] = X[i + ldx * j];
        orc_larfg(m - j, &a, x, 1, &taus[j]);
        X[j + ldx * j] = a;
        vs[j][j] = 1.0;
        for (int i = j + 1; i < m; ++i) { vs[j][i] = x[i - j - 1]; X[i + ldx * j] = 0.0; }
        for (int jj = j + 1; jj < k; ++jj) {
            double s = 0.0;
            for (int i = j; i < m; ++i) s += vs[j][i] * X[i + ldx * jj];
            s *= taus[j];
            for (int i = j; i < m; ++i) X[i + ldx * jj] -= s * vs[j][i];
        }
    }
    for (int j = 0; j < m; ++j) for (int i = 0; i < m; ++i) Qo[i + m * j] = (i == j) ? 1.0 : 0.0;
    for (int j = k - 1; j >= 0; --j) {
        for (int jj = 0; jj < m; ++jj) {
            double s = 0.0;
            for (int i = j; i < m; ++i) s += vs[j][i] * Qo[i + m * jj];
            s *= taus[j];
            for (int i = j; i < m; ++i) Qo[i + m * jj] -= s * vs[j][i];
        }
    }
}

static void mm(int m, const double *A, int ta, const double *B, int tb, double *C) {
    for (int j = 0; j < m; ++j) for (int i = 0; i < m; ++i) {
        double s = 0.0;
        for (int k = 0; k < m; ++k) {
            double a = ta ? A[k + m * i] : A[i + m * k];
            double b = tb ? B[j + m * k] : B[k + m * j];
            s += a * b;
        }
        C[i + m * j] = s;
    }
}

/* RQ factorisation of a k x m matrix X (k <= m, ld ldx) by Householder reflectors from the right,
 * bottom row first (DGERQ2 role): X Qo^T = [0 R] with R k x k upper triangular; Qo (m x m, ld m) is
 * the full orthogonal factor (DORGR2 role with M = N = m, K = k). */
static void rq_full(int k, int m, double *X, int ldx, double *Qo) {
    for (int j = 0; j < m; ++j) for (int i = 0; i < m; ++i) Qo[i + m * j] = (i == j) ? 1.0 : 0.0;
    for (int r = k - 1; r >= 0; --r) {
        const int len = m - (k - 1 - r);      /* columns 0..len-1, pivot at column len-1 */
        double al = X[r + ldx * (len - 1)], u[4], tau;
        for (int j = 0; j < len - 1; ++j) u[j] = X[r + ldx * j];
        orc_larfg(len, &al, u, 1, &tau);
        double v[4];
        for (int j = 0; j < len - 1; ++j) v[j] = u[j];
        v[len - 1] = 1.0;
        X[r + ldx * (len - 1)] = al;
        for (int j = 0; j < len - 1; ++j) X[r + ldx * j] = 0.0;
        /* rows above: X(0:r-1, 0:len-1) <- X H */
        for (int i = 0; i < r; ++i) {
            double t = 0.0;
            for (int j = 0; j < len; ++j) t += X[i + ldx * j] * v[j];
            t *= tau;
            for (int j = 0; j < len; ++j) X[i + ldx * j] -= t * v[j];
        }
        /* Qo <- H Qo (X = R H_k ... H_1 style accumulation: Qo = H_r ... applied last-first) */
        for (int jj = 0; jj < m; ++jj) {
            double t = 0.0;
            for (int j = 0; j < len; ++j) t += v[j] * Qo[j + m * jj];
            t *= tau;
            for (int j = 0; j < len; ++j) Qo[j + m * jj] -= t * v[j];
        }
    }
}

/* Frobenius norm of the lower-left (m - n2) x n2 block of an m x m matrix */
static double ll_norm(int m, int n2, const double *S) {
    double s = 0.0;
    for (int j = 0; j < n2; ++j) for (int i = n2; i < m; ++i) s += S[i + m * j] * S[i + m * j];
    return sqrt(s);
}

/* DTGEX2 (src/reorder/dtgex2.f:221-860, the tree's modified LAPACK 3.4.2 version): swap the adjacent
 * diagonal blocks of orders n1 (upper) and n2 (lower) at j1 of the pencil (A, B), updating Q and Z.
 *   - the m x m corner (m = n1 + n2) is scaled by its largest entries MA, MB (dtgex2.f:317-345) and one
 *     threshold THRESH = max(20 eps ||(S, T)||_F, SMLNUM) judges both matrices (:347-363);
 *   - 1x1 / 1x1: two Givens rotations (:365-395), weak test |S21| + |T21| <= THRESH (:401-407),
 *     strong test on the reconstruction of both blocks (:409-445), then the rotations applied to the
 *     unscaled pencil (:451-474);
 *   - otherwise: generalized Sylvester equation (DTGSY2), QR of [-L; scale I] (:523-538) and RQ of
 *     [scale I, R] (:546-563), the tentative swap (:567-574), T re-triangularised both by RQ and by QR
 *     and the variant with the smaller S21 kept when it passes THRESH (:582-647), the strong test on
 *     the reconstruction (:655-686), the copy back with the scale factors (:697-701) and DLAGV2
 *     standardisation of the new 2x2 blocks with the off-diagonal block, Q and Z updated (:706-800).
 * Returns 0 when the swap was done, 1 when it was rejected (the pencil is then unchanged). */
int orc_tgex2(int n, double *A, int lda, double *B, int ldb, double *Q, int ldq,
              double *Z, int ldz, int j1, int n1, int n2) {
    const int m = n1 + n2;
    if (n <= 1 || n1 <= 0 || n2 <= 0) return 0;
    if (n1 > n || j1 + n1 >= n) return 0;
    const double eps = ULP, smlnum = SAFMIN / eps, bignum = 1.0 / smlnum;
    double S[16], T[16], LI[16], IR[16];
    memset(LI, 0, sizeof LI); memset(IR, 0, sizeof IR);
    for (int j = 0; j < m; ++j) for (int i = 0; i < m; ++i) {
        S[i + m * j] = A[IX(j1 + i, j1 + j, lda)];
        T[i + m * j] = B[IX(j1 + i, j1 + j, ldb)];
    }
    /* scale the local copy (dtgex2.f:317-345) */
    double ma = 0.0, mb = 0.0;
    for (int i = 0; i < m * m; ++i) { ma = fmax(ma, fabs(S[i])); mb = fmax(mb, fabs(T[i])); }
    ma = fmin(ma, bignum); mb = fmin(mb, bignum);
    if (ma > 0.0) { ma = fmax(ma, smlnum); for (int i = 0; i < m * m; ++i) S[i] /= ma; }
    if (mb > 0.0) { mb = fmax(mb, smlnum); for (int j = 0; j < m; ++j) for (int i = 0; i <= j; ++i) T[i + m * j] /= mb; }
    double dn = 0.0;
    for (int i = 0; i < m * m; ++i) dn += S[i] * S[i] + T[i] * T[i];
    const double thresh = fmax(20.0 * eps * sqrt(dn), smlnum);
    /* the scaled copies of the original corner for the strong tests */
    double S0[16], T0[16];
    for (int j = 0; j < m; ++j) for (int i = 0; i < m; ++i) {
        S0[i + m * j] = A[IX(j1 + i, j1 + j, lda)] * (ma > 0.0 ? 1.0 / ma : 1.0);
        T0[i + m * j] = B[IX(j1 + i, j1 + j, ldb)] * (mb > 0.0 ? 1.0 / mb : 1.0);
    }
    double W[16], R3[16];
    if (m == 2) {
        /* CASE 1: 1x1 and 1x1 by Givens rotations (dtgex2.f:365-395) */
        const double f = S[3] * T[0] - T[3] * S[0];
        const double g = S[3] * T[2] - T[3] * S[2];
        const double sb = fabs(T[3]), sa = fabs(S[3]);
        double cs, sn, r;
        orc_lartg(f, g, &cs, &sn, &r);
        IR[0 + 2 * 1] = cs; IR[0] = sn; IR[1] = -cs; IR[1 + 2 * 1] = sn;
        const double rc = IR[0], rs = IR[1];            /* DROT(c = IR(1,1), s = IR(2,1)) */
        rot_cols(S, 2, 0, 1, 0, 2, rc, rs);
        rot_cols(T, 2, 0, 1, 0, 2, rc, rs);
        double lc, ls;
        if (sa >= sb) orc_lartg(S[0], S[1], &lc, &ls, &r);
        else orc_lartg(T[0], T[1], &lc, &ls, &r);
        rot_rows(S, 2, 0, 1, 0, 2, lc, ls);
        rot_rows(T, 2, 0, 1, 0, 2, lc, ls);
        LI[0] = lc; LI[1] = ls; LI[3] = lc; LI[2] = -ls;
        /* weak stability test (dtgex2.f:401-407) */
        if (!(fabs(S[1]) + fabs(T[1]) <= thresh)) return 1;
        /* strong stability test: ||(A - LI S IR^T, B - LI T IR^T)||_F / scales (dtgex2.f:409-445) */
        double ss = 0.0;
        mm(2, LI, 0, S, 0, W); mm(2, W, 0, IR, 1, R3);
        for (int i = 0; i < 4; ++i) ss += (S0[i] - R3[i]) * (S0[i] - R3[i]);
        mm(2, LI, 0, T, 0, W); mm(2, W, 0, IR, 1, R3);
        for (int i = 0; i < 4; ++i) ss += (T0[i] - R3[i]) * (T0[i] - R3[i]);
        if (!(sqrt(ss) <= thresh)) return 1;
        /* apply to the unscaled pencil (dtgex2.f:451-474) */
        rot_cols(A, lda, j1, j1 + 1, 0, j1 + 2, rc, rs);
        rot_cols(B, ldb, j1, j1 + 1, 0, j1 + 2, rc, rs);
        rot_rows(A, lda, j1, j1 + 1, j1, n, lc, ls);
        rot_rows(B, ldb, j1, j1 + 1, j1, n, lc, ls);
        A[IX(j1 + 1, j1, lda)] = 0.0;
        B[IX(j1 + 1, j1, ldb)] = 0.0;
        if (Z) rot_cols(Z, ldz, j1, j1 + 1, 0, n, rc, rs);
        if (Q) rot_cols(Q, ldq, j1, j1 + 1, 0, n, lc, ls);
        return 0;
    }
    /* CASE 2: generalized Sylvester equation S11 R - L S22 = scale S12, T11 R - L T22 = scale T12
     * (DTGSY2 with one block on each side: unknowns vec(R), vec(L); DGETC2 + DGESC2) */
    const int nn = n1 * n2, mm2 = 2 * nn;
    double K[64], rhs[8];
    memset(K, 0, sizeof K);
    for (int q = 0; q < n2; ++q) for (int p = 0; p < n1; ++p) {
        const int row = p + n1 * q;
        for (int k = 0; k < n1; ++k) {
            K[row + mm2 * (k + n1 * q)] += S[p + m * k];
            K[row + nn + mm2 * (k + n1 * q)] += T[p + m * k];
        }
        for (int k = 0; k < n2; ++k) {
            K[row + mm2 * (nn + p + n1 * k)] -= S[(n1 + k) + m * (n1 + q)];
            K[row + nn + mm2 * (nn + p + n1 * k)] -= T[(n1 + k) + m * (n1 + q)];
        }
        rhs[row] = S[p + m * (n1 + q)];
        rhs[row + nn] = T[p + m * (n1 + q)];
    }
    const double scale = solve_cp(mm2, K, rhs);
    /* LI: full orthogonal factor of the QR of [-L; scale I_n2] (m x n2) (dtgex2.f:523-538) */
    double X[16];
    for (int j = 0; j < n2; ++j) for (int i = 0; i < m; ++i)
        X[i + m * j] = (i < n1) ? -rhs[nn + i + n1 * j] : ((i - n1) == j ? scale : 0.0);
    qr_full(m, n2, X, m, LI);
    /* IR: full orthogonal factor of the RQ of [scale I_n1, R] (n1 x m) (dtgex2.f:546-563) */
    double Y[16];
    for (int j = 0; j < m; ++j) for (int i = 0; i < n1; ++i)
        Y[i + n1 * j] = (j < n1) ? (i == j ? scale : 0.0) : rhs[i + n1 * (j - n1)];
    rq_full(n1, m, Y, n1, IR);
    /* tentative swap S <- LI^T S IR^T, T likewise (dtgex2.f:567-574) */
    mm(m, LI, 1, S, 0, W); mm(m, W, 0, IR, 1, S);
    mm(m, LI, 1, T, 0, W); mm(m, W, 0, IR, 1, T);
    double Sq[16], Tq[16], LIq[16], IRq[16];
    memcpy(Sq, S, sizeof S); memcpy(Tq, T, sizeof T); memcpy(LIq, LI, sizeof LI); memcpy(IRq, IR, sizeof IR);
    /* variant RQ: T = R Qt, S <- S Qt^T, IR <- Qt IR (dtgex2.f:582-605) */
    for (int r = m - 1; r >= 1; --r) {
        double al = T[r + m * r], u[4], tau, v[4];
        for (int j = 0; j < r; ++j) u[j] = T[r + m * j];
        orc_larfg(r + 1, &al, u, 1, &tau);
        for (int j = 0; j < r; ++j) v[j] = u[j];
        v[r] = 1.0;
        T[r + m * r] = al;
        for (int j = 0; j < r; ++j) T[r + m * j] = 0.0;
        reflm_right(T, m, 0, r + 1, 0, r, v, tau);
        reflm_right(S, m, 0, r + 1, 0, m, v, tau);
        reflm_left(IR, m, 0, r + 1, 0, m, v, tau);
    }
    const double brqa21 = ll_norm(m, n2, S);
    /* variant QR: T = Qt R, S <- Qt^T S, LI <- LI Qt (dtgex2.f:613-636) */
    for (int c = 0; c < m - 1; ++c) {
        double al = Tq[c + m * c], u[4], tau, v[4];
        for (int i = c + 1; i < m; ++i) u[i - c - 1] = Tq[i + m * c];
        orc_larfg(m - c, &al, u, 1, &tau);
        v[0] = 1.0;
        for (int i = c + 1; i < m; ++i) v[i - c] = u[i - c - 1];
        Tq[c + m * c] = al;
        for (int i = c + 1; i < m; ++i) Tq[i + m * c] = 0.0;
        reflm_left(Tq, m, c, m - c, c + 1, m, v, tau);
        reflm_left(Sq, m, c, m - c, 0, m, v, tau);
        reflm_right(LIq, m, c, m - c, 0, m, v, tau);
    }
    const double bqra21 = ll_norm(m, n2, Sq);
    /* choice and weak stability test (dtgex2.f:643-650) */
    if (bqra21 <= brqa21 && bqra21 <= thresh) {
        memcpy(S, Sq, sizeof S); memcpy(T, Tq, sizeof T); memcpy(IR, IRq, sizeof IR); memcpy(LI, LIq, sizeof LI);
    } else if (brqa21 >= thresh) {
        return 1;
    }
    for (int j = 0; j < m; ++j) for (int i = j + 1; i < m; ++i) T[i + m * j] = 0.0;
    /* strong stability test: ||(A - LI S IR, B - LI T IR)||_F / scales (dtgex2.f:655-686) */
    double ss = 0.0;
    mm(m, LI, 0, S, 0, W); mm(m, W, 0, IR, 0, R3);
    for (int i = 0; i < m * m; ++i) ss += (S0[i] - R3[i]) * (S0[i] - R3[i]);
    mm(m, LI, 0, T, 0, W); mm(m, W, 0, IR, 0, R3);
    for (int i = 0; i < m * m; ++i) ss += (T0[i] - R3[i]) * (T0[i] - R3[i]);
    if (!(sqrt(ss) <= thresh)) return 1;
    /* accepted: zero S21, copy back with the scales (dtgex2.f:692-701) */
    for (int j = 0; j < n2; ++j) for (int i = n2; i < m; ++i) S[i + m * j] = 0.0;
    for (int j = 0; j < m; ++j) for (int i = 0; i < m; ++i) {
        A[IX(j1 + i, j1 + j, lda)] = S[i + m * j] * ma;
        B[IX(j1 + i, j1 + j, ldb)] = T[i + m * j] * mb;
    }
    /* standardise the new 2x2 blocks (dtgex2.f:706-727): WL (left) and WR (right) m x m */
    double WL[16], WR[16];
    for (int j = 0; j < m; ++j) for (int i = 0; i < m; ++i) WL[i + m * j] = WR[i + m * j] = (i == j) ? 1.0 : 0.0;
    double ar2[2], ai2[2], be2[2], QL[4], ZR[4];
    if (n2 > 1) {
        orc_lagv2(&A[IX(j1, j1, lda)], lda, &B[IX(j1, j1, ldb)], ldb, ar2, ai2, be2, QL, ZR);
        WL[0] = QL[0]; WL[1] = QL[1]; WL[m] = QL[2]; WL[m + 1] = QL[3];
        WR[0] = ZR[0]; WR[1] = ZR[1]; WR[m] = ZR[2]; WR[m + 1] = ZR[3];
    }
    if (n1 > 1) {
        orc_lagv2(&A[IX(j1 + n2, j1 + n2, lda)], lda, &B[IX(j1 + n2, j1 + n2, ldb)], ldb, ar2, ai2, be2, QL, ZR);
        const int o = n2;
        WL[o + m * o] = QL[0]; WL[(o + 1) + m * o] = QL[1]; WL[o + m * (o + 1)] = QL[2]; WL[(o + 1) + m * (o + 1)] = QL[3];
        WR[o + m * o] = ZR[0]; WR[(o + 1) + m * o] = ZR[1]; WR[o + m * (o + 1)] = ZR[2]; WR[(o + 1) + m * (o + 1)] = ZR[3];
    }
    /* off-diagonal block A(j1:j1+n2-1, j1+n2:j1+m-1) <- WL11^T (.) WR22 (dtgex2.f:728-745) */
    for (int t = 0; t < 2; ++t) {
        double *M = t ? B : A;
        const int ldm = t ? ldb : lda;
        double O[4], O2[4];
        for (int j = 0; j < n1; ++j) for (int i = 0; i < n2; ++i) O[i + n2 * j] = M[IX(j1 + i, j1 + n2 + j, ldm)];
        for (int j = 0; j < n1; ++j) for (int i = 0; i < n2; ++i) {
            double s = 0.0;
            for (int k = 0; k < n2; ++k) s += WL[k + m * i] * O[k + n2 * j];
            O2[i + n2 * j] = s;
        }
        for (int j = 0; j < n1; ++j) for (int i = 0; i < n2; ++i) {
            double s = 0.0;
            for (int k = 0; k < n1; ++k) s += O2[i + n2 * k] * WR[(n2 + k) + m * (n2 + j)];
            M[IX(j1 + i, j1 + n2 + j, ldm)] = s;
        }
    }
    /* LI <- LI WL, IR <- IR^T WR (dtgex2.f:740, :750-752): the final left and right factors */
    mm(m, LI, 0, WL, 0, W); memcpy(LI, W, sizeof(double) * m * m);
    mm(m, IR, 1, WR, 0, W); memcpy(IR, W, sizeof(double) * m * m);
    /* Q <- Q LI, Z <- Z IR, rows right of the corner <- LI^T (.), columns above <- (.) IR
     * (dtgex2.f:756-800) */
    double tmp[4];
    for (int j = j1 + m; j < n; ++j) {
        for (int t = 0; t < 2; ++t) {
            double *M = t ? B : A;
            const int ldm = t ? ldb : lda;
            for (int i = 0; i < m; ++i) { double s = 0; for (int k = 0; k < m; ++k) s += LI[k + m * i] * M[IX(j1 + k, j, ldm)]; tmp[i] = s; }
            for (int i = 0; i < m; ++i) M[IX(j1 + i, j, ldm)] = tmp[i];
        }
    }
    for (int i = 0; i < j1; ++i) {
        for (int t = 0; t < 2; ++t) {
            double *M = t ? B : A;
            const int ldm = t ? ldb : lda;
            for (int j = 0; j < m; ++j) { double s = 0; for (int k = 0; k < m; ++k) s += M[IX(i, j1 + k, ldm)] * IR[k + m * j]; tmp[j] = s; }
            for (int j = 0; j < m; ++j) M[IX(i, j1 + j, ldm)] = tmp[j];
        }
    }
    if (Q) for (int i = 0; i < n; ++i) {
        for (int j = 0; j < m; ++j) { double s = 0; for (int k = 0; k < m; ++k) s += Q[IX(i, j1 + k, ldq)] * LI[k + m * j]; tmp[j] = s; }
        for (int j = 0; j < m; ++j) Q[IX(i, j1 + j, ldq)] = tmp[j];
    }
    if (Z) for (int i = 0; i < n; ++i) {
        for (int j = 0; j < m; ++j) { double s = 0; for (int k = 0; k < m; ++k) s += Z[IX(i, j1 + k, ldz)] * IR[k + m * j]; tmp[j] = s; }
        for (int j = 0; j < m; ++j) Z[IX(i, j1 + j, ldz)] = tmp[j];
    }
    return 0;
}

static int blksize_at(const double *A, int lda, int n, int j) {
    return (j + 1 < n && A[IX(j + 1, j, lda)] != 0.0) ? 2 : 1;
}

/* move the 1x1 block at 'from' up to 'to' by adjacent swaps */
static int move1_up(int n, double *A, int lda, double *B, int ldb, double *Q, int ldq,
                    double *Z, int ldz, int from, int to) {
    int here = from;
    while (here > to) {
        int nbn = (here >= 2 && A[IX(here - 1, here - 2, lda)] != 0.0) ? 2 : 1;
        if (orc_tgex2(n, A, lda, B, ldb, Q, ldq, Z, ldz, here - nbn, nbn, 1)) return 1;
        here -= nbn;
    }
    return 0;
}

/* DTGEXC for ilst < ifst (src/reorder/dtgexc.f): move the block starting at *ifst up to
 * *ilst (a block boundary); a 2x2 block that splits on the way continues as two 1x1 blocks
 * (dtgexc.f "current block consists of two 1-by-1 blocks" branch). */
int orc_tgexc(int n, double *A, int lda, double *B, int ldb, double *Q, int ldq,
              double *Z, int ldz, int *ifst, int *ilst) {
    int here = *ifst;
    if (here > 0 && A[IX(here, here - 1, lda)] != 0.0) here--;
    int nbf = blksize_at(A, lda, n, here);
    int target = *ilst;
    while (here > target) {
        int nbn = (here >= 2 && A[IX(here - 1, here - 2, lda)] != 0.0) ? 2 : 1;
        if (orc_tgex2(n, A, lda, B, ldb, Q, ldq, Z, ldz, here - nbn, nbn, nbf)) { *ilst = here; return 1; }
        here -= nbn;
        if (nbf == 2 && A[IX(here + 1, here, lda)] == 0.0) {
            if (move1_up(n, A, lda, B, ldb, Q, ldq, Z, ldz, here, target)) { *ilst = here; return 1; }
            if (move1_up(n, A, lda, B, ldb, Q, ldq, Z, ldz, here + 1, target + 1)) { *ilst = here; return 1; }
            *ilst = target;
            return 0;
        }
    }
    *ilst = here;
    return 0;
}

/* Reordering of a generalized Schur form so that the selected eigenvalues lead (PDTGORD semantics,
 * src/reorder/PDTGORD.f:1-4, restated serially as the reorder loop of a DTGSEN: every selected block,
 * top to bottom, is moved by DTGEXC to the next free leading position).  A 2x2 block is selected when
 * either of its rows is (PDTGORD.f:512-527).  sel (n ints) is normalised and, on exit, shows the final
 * positions; *m = number of selected eigenvalues.  Returns 0 or the 1-based row at which a swap was
 * rejected (the pencil is then partially reordered). */
int orc_tgord(int n, double *S, int lds, double *T, int ldt, double *Q, int ldq, double *Z, int ldz,
              int *sel, int *m) {
    for (int k = 0; k + 1 < n; ++k)
        if (S[IX(k + 1, k, lds)] != 0.0 && (sel[k] || sel[k + 1])) sel[k] = sel[k + 1] = 1;
    int cnt = 0;
    for (int k = 0; k < n; ++k) { sel[k] = sel[k] != 0; cnt += sel[k]; }
    if (m) *m = cnt;
    int ks = 0, info = 0;
    for (int k = 0; k < n; ) {
        const int bs = (k + 1 < n && S[IX(k + 1, k, lds)] != 0.0) ? 2 : 1;
        if (sel[k]) {
            if (k != ks) {
                int ifst = k, ilst = ks;
                if (orc_tgexc(n, S, lds, T, ldt, Q, ldq, Z, ldz, &ifst, &ilst)) { info = ilst + 1; break; }
            }
            /* the block may have split on its way up: count what now sits at ks */
            ks += bs;
        }
        k += bs;
    }
    /* final positions of the selected eigenvalues */
    for (int k = 0; k < n; ++k) sel[k] = (k < ks) ? 1 : 0;
    if (info) {
        /* partial: positions below ks that still hold selected blocks are not tracked here */
    }
    return info;
}

/* ===================================================================================== */
/* HT reduction (EXRAND2.f:296-337 QR of B, then PDGGHRD.f:365-642 semantics)            */
/* ===================================================================================== */
int orc_geqrf_apply(int n, double *A, int lda, double *B, int ldb, double *Q, int ldq, int compq) {
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    if (compq == 1) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) Q[IX(i, j, ldq)] = (i == j);
    for (int j = 0; j < n - 1; ++j) {
        int m = n - j;
        double alpha = B[IX(j, j, ldb)], tau;
        for (int i = 1; i < m; ++i) v[i] = B[IX(j + i, j, ldb)];
        orc_larfg(m, &alpha, v + 1, 1, &tau);
        v[0] = 1.0;
        B[IX(j, j, ldb)] = alpha;
        for (int i = 1; i < m; ++i) B[IX(j + i, j, ldb)] = 0.0;
        reflm_left(B, ldb, j, m, j + 1, n, v, tau);
        reflm_left(A, lda, j, m, 0, n, v, tau);
        if (compq) reflm_right(Q, ldq, j, m, 0, n, v, tau);
    }
    free(v);
    return 0;
}

int orc_gghrd(int n, int ilo, int ihi, double *A, int lda, double *B, int ldb,
              double *Q, int ldq, double *Z, int ldz, int compq, int compz) {
    if (compq == 1) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) Q[IX(i, j, ldq)] = (i == j);
    if (compz == 1) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) Z[IX(i, j, ldz)] = (i == j);
    for (int j = ilo; j <= ihi - 2; ++j) {
        for (int i = ihi; i >= j + 2; --i) {
            double cs, sn, r;
            /* row rotation (i-1, i) zeroing A(i, j)  (KRNLUPDATEANDREDUCECOLUMN.f:47-67) */
            orc_lartg(A[IX(i - 1, j, lda)], A[IX(i, j, lda)], &cs, &sn, &r);
            A[IX(i - 1, j, lda)] = r; A[IX(i, j, lda)] = 0.0;
            rot_rows(A, lda, i - 1, i, j + 1, n, cs, sn);
            rot_rows(B, ldb, i - 1, i, i - 1, n, cs, sn);
            if (compq) rot_cols(Q, ldq, i - 1, i, 0, n, cs, sn);
            /* column rotation (i, i-1) zeroing B(i, i-1)  (KRNLCOLUMNANNIHILATE.f:78-90) */
            orc_lartg(B[IX(i, i, ldb)], B[IX(i, i - 1, ldb)], &cs, &sn, &r);
            B[IX(i, i, ldb)] = r; B[IX(i, i - 1, ldb)] = 0.0;
            rot_cols(A, lda, i, i - 1, 0, ihi + 1, cs, sn);
            rot_cols(B, ldb, i, i - 1, 0, i, cs, sn);
            if (compz) rot_cols(Z, ldz, i, i - 1, 0, n, cs, sn);
        }
    }
    return 0;
}

/* ===================================================================================== */
/* PIPARMQ defaults                                                                      */
/* ===================================================================================== */
int orc_piparmq_ns(int nh) {                /* piparmq.f:445-476 */
    int ns = 2;
    if (nh >= 30) ns = 4;
    if (nh >= 60) ns = 10;
    if (nh >= 150) { int lg = (int)lround(log((double)nh) / log(2.0)); ns = MAXI(10, nh / lg); }
    if (nh >= 590) ns = 64;
    if (nh >= 3000) ns = 128;
    if (nh >= 6000) ns = 256;
    if (nh >= 12000) ns = 512;
    if (nh >= 24000) ns = 1024;
    if (nh >= 48000) ns = 2048;
    ns = MAXI(2, ns - ns % 2);
    return ns;
}
int orc_piparmq_nw(int nh) {                /* piparmq.f:505-510, PDHGEQZ0.f:457-464 */
    int ns = orc_piparmq_ns(nh);
    return nh <= 500 ? ns : 3 * ns / 2;
}
int orc_piparmq_nibble(int nh) {            /* piparmq.f:488-497, 1x1 grid */
    if (nh < 6000) return 14;
    int v = (int)ceil(335.0 * pow((double)nh, -0.44));
    return MINI(100, v);
}

/* ===================================================================================== */
/* AED (QZEARLY, src/kkqz/qzearly.f:1-455, and PDHGEQZ2/3 1x1 paths)                     */
/* ===================================================================================== */
static int qz_run(qzctx *c, int ilo, int ihi);

static void gemm_nn(int m, int n, int k, const double *A, int lda, const double *B, int ldb,
                    double *C, int ldc) {
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < m; ++i) C[IX(i, j, ldc)] = 0.0;
        for (int p = 0; p < k; ++p) {
            double b = B[IX(p, j, ldb)];
            if (b == 0.0) continue;
            const double *a = A + IX(0, p, lda);
            double *cc = C + IX(0, j, ldc);
            for (int i = 0; i < m; ++i) cc[i] += a[i] * b;
        }
    }
}
/* C (k x n) := U^T (k x k) * C   via workspace */
static void apply_left_t(double *C, int ldc, int k, int n, const double *U, int ldu, double *w) {
    if (n <= 0 || k <= 0) return;
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < k; ++i) {
            double s = 0.0;
            const double *u = U + IX(0, i, ldu);
            const double *cc = C + IX(0, j, ldc);
            for (int p = 0; p < k; ++p) s += u[p] * cc[p];
            w[i] = s;
        }
        for (int i = 0; i < k; ++i) C[IX(i, j, ldc)] = w[i];
    }
}
/* C (m x k) := C * V (k x k) */
static void apply_right(double *C, int ldc, int m, int k, const double *V, int ldv, double *w) {
    if (m <= 0 || k <= 0) return;
    gemm_nn(m, k, k, C, ldc, V, ldv, w, m);
    for (int j = 0; j < k; ++j) memcpy(C + IX(0, j, ldc), w + IX(0, j, m), sizeof(double) * (size_t)m);
}

/* returns ND; shifts (undeflated window eigenvalues, ns_out of them) into sr/si/sb */
static int aed(qzctx *c, int ktop, int kbot, int jw, double *sr, double *si, double *sb, int *ns_out) {
    int kwtop = kbot - jw + 1;
    double s = (kwtop == ktop) ? 0.0 : Hc(kwtop, kwtop - 1);
    size_t jj = (size_t)jw * jw;
    double *Hw = (double *)malloc(sizeof(double) * jj * 4);
    double *Tw = Hw + jj, *U = Tw + jj, *V = U + jj;
    double *war = (double *)malloc(sizeof(double) * (size_t)jw * 3);
    double *wai = war + jw, *wb = wai + jw;
    for (int j = 0; j < jw; ++j) for (int i = 0; i < jw; ++i) {
        Hw[i + (size_t)jw * j] = (i <= j + 1) ? Hc(kwtop + i, kwtop + j) : 0.0;
        Tw[i + (size_t)jw * j] = (i <= j) ? Tc(kwtop + i, kwtop + j) : 0.0;
        U[i + (size_t)jw * j] = V[i + (size_t)jw * j] = (i == j);
    }
    /* window QZ (PDHGEQZ3.f:588 -> PDHGEQZ1; QZEARLY -> QZDACK/QZLAP) */
    qzctx w; memset(&w, 0, sizeof w);
    w.n = jw; w.H = Hw; w.T = Tw; w.Q = U; w.Z = V; w.ldh = w.ldt = w.ldq = w.ldz = jw;
    w.smlnum = SAFMIN * ((double)jw / ULP); w.ttol = c->ttol;
    w.alphar = war; w.alphai = wai; w.beta = wb; w.st = c->st; w.level = c->level + 1;
    w.prm = c->prm; w.prm.ns = 0; w.prm.nw = 0; w.prm.nibble = 0;
    int info = qz_run(&w, 0, jw - 1);
    if (info != 0) {            /* soft recovery: no deflation (PDHGEQZ0.f:944-949) */
        free(Hw); free(war); *ns_out = 0; return 0;
    }
    /* deflation test from the bottom; undeflatable blocks moved to the top (qzearly.f:250-367) */
    int ilst = 0, kb = jw - 1, knt = 0;
    int stop = 0;
    while (knt < jw && !stop) {
        int bulge = (kb > 0) && Hw[kb + (size_t)jw * (kb - 1)] != 0.0;
        int dfl;
        if (!bulge) {
            knt += 1;
            double tmp = fabs(s) + fabs(Hw[kb + (size_t)jw * kb]);
            dfl = fabs(s * U[(size_t)jw * kb]) <= fmax(ULP * tmp, c->smlnum);
        } else {
            knt += 2;
            double h00 = Hw[(kb - 1) + (size_t)jw * (kb - 1)], h10 = Hw[kb + (size_t)jw * (kb - 1)];
            double h01 = Hw[(kb - 1) + (size_t)jw * kb], h11 = Hw[kb + (size_t)jw * kb];
            double rtdet;
            if (h00 == 0.0 && h10 == 0.0) rtdet = 0.0;
            else if (fabs(h00) >= fabs(h10)) rtdet = sqrt(fabs(h00)) * sqrt(fabs(h11 - (h10 / h00) * h01));
            else rtdet = sqrt(fabs(h10)) * sqrt(fabs(h01 - (h00 / h10) * h11));
            double tmp = fabs(s) + rtdet;
            dfl = fmax(fabs(s * U[(size_t)jw * kb]), fabs(s * U[(size_t)jw * (kb - 1)])) <= fmax(ULP * tmp, c->smlnum);
        }
        if (kb < ilst) break;
        if (dfl) { kb -= bulge ? 2 : 1; }
        else {
            int ifst = kb, il = ilst;
            if (orc_tgexc(jw, Hw, jw, Tw, jw, U, jw, V, jw, &ifst, &il)) {
                if (c->st) c->st->swap_fail++;
                stop = 1; break;
            }
            if (c->st) c->st->swaps++;
            ilst += bulge ? 2 : 1;
        }
    }
    int kbc = kb + 1;          /* rows 0..kbc-1: undeflatable (+ unexamined on failure) */
    if (kbc < ilst) kbc = ilst;
    int nd = jw - kbc;
    /* eigenvalues of the whole window Schur form (blocks), top part = shifts */
    for (int j = 0; j < jw; ) {
        if (j + 1 < jw && Hw[(j + 1) + (size_t)jw * j] != 0.0) {
            double A2[4] = { Hw[j + (size_t)jw * j], Hw[(j + 1) + (size_t)jw * j], Hw[j + (size_t)jw * (j + 1)], Hw[(j + 1) + (size_t)jw * (j + 1)] };
            double B2[4] = { Tw[j + (size_t)jw * j], 0.0, Tw[j + (size_t)jw * (j + 1)], Tw[(j + 1) + (size_t)jw * (j + 1)] };
            double ar[2], ai[2], bt[2];
            orc_eig22(A2, 2, B2, 2, ar, ai, bt);
            double b0 = B2[0], b1 = B2[3];
            if (ai[0] != 0.0) {
                war[j] = ar[0] * b0; wai[j] = fabs(ai[0]) * b0; wb[j] = b0;
                war[j + 1] = ar[0] * b1; wai[j + 1] = -fabs(ai[0]) * b1; wb[j + 1] = b1;
            } else {
                war[j] = ar[0] * b0; wai[j] = 0.0; wb[j] = b0;
                war[j + 1] = ar[1] * b1; wai[j + 1] = 0.0; wb[j + 1] = b1;
            }
            j += 2;
        } else { war[j] = Hw[j + (size_t)jw * j]; wai[j] = 0.0; wb[j] = Tw[j + (size_t)jw * j]; j += 1; }
    }
    for (int i = 0; i < kbc; ++i) { sr[i] = war[i]; si[i] = wai[i]; sb[i] = wb[i]; }
    *ns_out = kbc;
    /* spike reflector, RQ of T, HT restore (qzearly.f:369-451) */
    double *wk = (double *)malloc(sizeof(double) * (size_t)jw * (size_t)(jw > c->n ? jw : c->n) + 16);
    if (kbc > 0 && s != 0.0) {
        double *v = (double *)malloc(sizeof(double) * (size_t)kbc);
        for (int i = 0; i < kbc; ++i) v[i] = s * U[(size_t)jw * i];
        double alpha = v[0], tau;
        orc_larfg(kbc, &alpha, v + 1, 1, &tau);
        v[0] = 1.0;
        reflm_left(Hw, jw, 0, kbc, 0, jw, v, tau);
        reflm_right(U, jw, 0, kbc, 0, jw, v, tau);
        reflm_left(Tw, jw, 0, kbc, 0, jw, v, tau);
        free(v);
        /* RQ of Tw(0:kbc,0:kbc) by Householders from the right, bottom row first */
        double *u = (double *)malloc(sizeof(double) * (size_t)kbc);
        for (int r = kbc - 1; r >= 1; --r) {
            /* zero Tw(r, 0:r-1) using Tw(r, r) */
            double al = Tw[r + (size_t)jw * r], tu;
            for (int j = 0; j < r; ++j) u[j] = Tw[r + (size_t)jw * j];
            orc_larfg(r + 1, &al, u, 1, &tu);
            /* reflector vector in column order (0..r-1, r) = (u[0..r-1], 1) */
            double *vv = (double *)malloc(sizeof(double) * (size_t)(r + 1));
            for (int j = 0; j < r; ++j) vv[j] = u[j];
            vv[r] = 1.0;
            reflm_right(Tw, jw, 0, r + 1, 0, r + 1, vv, tu);
            reflm_right(Hw, jw, 0, r + 1, 0, jw, vv, tu);
            reflm_right(V, jw, 0, r + 1, 0, jw, vv, tu);
            for (int j = 0; j < r; ++j) Tw[r + (size_t)jw * j] = 0.0;
            free(vv);
        }
        free(u);
        /* Hessenberg restore of the leading kbc block (DGGHRD, qzearly.f:449) */
        for (int j = 0; j <= kbc - 3; ++j) {
            for (int i = kbc - 1; i >= j + 2; --i) {
                double cs, sn, r;
                orc_lartg(Hw[(i - 1) + (size_t)jw * j], Hw[i + (size_t)jw * j], &cs, &sn, &r);
                Hw[(i - 1) + (size_t)jw * j] = r; Hw[i + (size_t)jw * j] = 0.0;
                rot_rows(Hw, jw, i - 1, i, j + 1, jw, cs, sn);
                rot_rows(Tw, jw, i - 1, i, i - 1, jw, cs, sn);
                rot_cols(U, jw, i - 1, i, 0, jw, cs, sn);
                orc_lartg(Tw[i + (size_t)jw * i], Tw[i + (size_t)jw * (i - 1)], &cs, &sn, &r);
                Tw[i + (size_t)jw * i] = r; Tw[i + (size_t)jw * (i - 1)] = 0.0;
                rot_cols(Hw, jw, i, i - 1, 0, kbc, cs, sn);
                rot_cols(Tw, jw, i, i - 1, 0, i, cs, sn);
                rot_cols(V, jw, i, i - 1, 0, jw, cs, sn);
            }
        }
    }
    /* deflated eigenvalues are final */
    for (int i = kbc; i < jw; ++i) set_eig(c, kwtop + i, war[i], wai[i], wb[i]);
    /* write the window back and update outside it (PDHGEQZ3.f:1571-1725, PDHGEQZ6 role) */
    for (int j = 0; j < jw; ++j) for (int i = 0; i < jw; ++i) {
        Hc(kwtop + i, kwtop + j) = Hw[i + (size_t)jw * j];
        Tc(kwtop + i, kwtop + j) = Tw[i + (size_t)jw * j];
    }
    /* the spike's first entry; zero when the whole window deflated (PDHGEQZ3.f:831, :1346) */
    if (kwtop > ktop) Hc(kwtop, kwtop - 1) = kbc > 0 ? s * U[0] : 0.0;
    if (kwtop > ktop) for (int i = 1; i < jw; ++i) Hc(kwtop + i, kwtop - 1) = 0.0;
    int ncol = c->n - (kbot + 1);
    apply_left_t(&Hc(kwtop, kbot + 1), c->ldh, jw, ncol, U, jw, wk);
    apply_left_t(&Tc(kwtop, kbot + 1), c->ldt, jw, ncol, U, jw, wk);
    apply_right(&Hc(0, kwtop), c->ldh, kwtop, jw, V, jw, wk);
    apply_right(&Tc(0, kwtop), c->ldt, kwtop, jw, V, jw, wk);
    if (c->Q) apply_right(c->Q + IX(0, kwtop, c->ldq), c->ldq, c->n, jw, U, jw, wk);
    if (c->Z) apply_right(c->Z + IX(0, kwtop, c->ldz), c->ldz, c->n, jw, V, jw, wk);
    free(wk); free(Hw); free(war);
    if (c->st) c->st->aed_calls++;
    return nd;
}

/* ===================================================================================== */
/* driver (PDHGEQZ0.f:505-1202 / PDHGEQZ1.f:466-763 semantics)                           */
/* ===================================================================================== */
static int qz_run(qzctx *c, int ilo, int ihi) {
    int nmin = c->prm.nmin > 0 ? c->prm.nmin : 75;
    int nh0 = ihi - ilo + 1;
    if (nh0 <= nmin) return qz_ds_ctx(c, ilo, ihi);
    int ns_cap = c->prm.ns > 0 ? c->prm.ns : orc_piparmq_ns(nh0);
    int nw = c->prm.nw > 0 ? c->prm.nw : orc_piparmq_nw(nh0);
    int nibble = c->prm.nibble > 0 ? c->prm.nibble : orc_piparmq_nibble(nh0);
    int chain = c->prm.chain > 0 ? c->prm.chain : 20;
    int maxit = (c->prm.maxit > 0 ? c->prm.maxit : 30) * nh0;
    double *sr = (double *)malloc(sizeof(double) * (size_t)(nw + 4) * 3);
    double *si = sr + nw + 4, *sb = si + nw + 4;
    int kbot = ihi;
    int info = 0;
    for (int it = 0; ; ++it) {
        if (kbot < ilo) break;
        if (it >= maxit) { info = kbot + 1; break; }
        if (c->st) c->st->iters++;
        /* deflation scan (PDLASMSUB.f:368) */
        int ktop = ilo;
        for (int k = kbot; k > ilo; --k) {
            double tst = fabs(Hc(k - 1, k - 1)) + fabs(Hc(k, k));
            if (fabs(Hc(k, k - 1)) <= fmax(c->smlnum, ULP * tst)) { Hc(k, k - 1) = 0.0; ktop = k; break; }
        }
        /* infinite eigenvalues (PDHGEQZ0.f:540 -> PDHGEQZ7) */
        int kt = ktop, kb = kbot;
        int ninf = inf_deflate(c, &kt, &kb);
        if (ninf > 0) {
            if (kt > ktop && kb >= kt) {
                /* eigenvalues deflated at the top of [ktop,kbot]: finish [kt,kb] separately */
                int r = qz_run(c, kt, kb);
                if (r) { info = r; break; }
                kbot = ktop - 1;
            } else kbot = (kb < kt) ? ktop - 1 : kb;
            continue;
        }
        int nh = kbot - ktop + 1;
        if (nh <= nmin) {      /* small block (PDHGEQZ0.f:597-875) */
            int r = qz_ds_ctx(c, ktop, kbot);
            if (r) { info = r; break; }
            kbot = ktop - 1;
            continue;
        }
        int jw = MINI(nw, nh);
        int ls = 0;
        int nd = aed(c, ktop, kbot, jw, sr, si, sb, &ls);
        kbot -= nd;
        if (kbot < ktop) continue;
        /* sweep unless AED deflated enough (PDHGEQZ0.f:989-990) */
        if (nd == 0 || (100 * nd <= jw * nibble && kbot - ktop + 1 > nmin)) {
            int ns = MINI(ns_cap, MAXI(2, kbot - ktop - 1));
            ns -= ns % 2;
            int avail = ls;
            if (avail > jw - nd) avail = jw - nd;
            if (avail < 2) {
                /* too few shifts: use the bottom 2x2 eigenvalues (exceptional fall-back) */
                double Ab[4] = { Hc(kbot - 1, kbot - 1), Hc(kbot, kbot - 1), Hc(kbot - 1, kbot), Hc(kbot, kbot) };
                double Bb[4] = { Tc(kbot - 1, kbot - 1), 0.0, Tc(kbot - 1, kbot), Tc(kbot, kbot) };
                double ar[2], ai[2], bb[2];
                orc_eig22(Ab, 2, Bb, 2, ar, ai, bb);
                sr[0] = ar[0]; sr[1] = ar[1]; si[0] = ai[0]; si[1] = ai[1]; sb[0] = sb[1] = 1.0;
                avail = 2;
            }
            /* sort by decreasing |lambda| (PDHGEQZ0.f:1068-1102) */
            if (avail > ns) {
                for (int k = avail - 1; k > 0; --k) {
                    int sorted = 1;
                    for (int i = 0; i < k; ++i) {
                        double d1 = fabs((fabs(sr[i]) + fabs(si[i])) / sb[i]);
                        double d2 = fabs((fabs(sr[i + 1]) + fabs(si[i + 1])) / sb[i + 1]);
                        if (d1 < d2) {
                            double t;
                            t = sr[i]; sr[i] = sr[i + 1]; sr[i + 1] = t;
                            t = si[i]; si[i] = si[i + 1]; si[i + 1] = t;
                            t = sb[i]; sb[i] = sb[i + 1]; sb[i + 1] = t;
                            sorted = 0;
                        }
                    }
                    if (sorted) break;
                }
            }
            /* pair shuffle (PDHGEQZ0.f:1111-1126): keep conjugate pairs together */
            for (int i = avail - 1; i >= 2; i -= 2) {
                if (si[i] == 0.0 && si[i - 1] != 0.0) {
                    double t;
                    t = sr[i]; sr[i] = sr[i - 1]; sr[i - 1] = sr[i - 2]; sr[i - 2] = t;
                    t = si[i]; si[i] = si[i - 1]; si[i - 1] = si[i - 2]; si[i - 2] = t;
                    t = sb[i]; sb[i] = sb[i - 1]; sb[i - 1] = sb[i - 2]; sb[i - 2] = t;
                }
            }
            ns = MINI(ns, avail);
            ns -= ns % 2;
            if (ns < 2) continue;
            int ks = avail - ns;            /* use the last ns (smallest) shifts */
            /* replace non-finite shifts by the bottom eigenvalue */
            for (int i = ks; i < avail; ++i) {
                if (!isfinite(sr[i] / sb[i]) || !isfinite(si[i] / sb[i])) { sr[i] = Hc(kbot, kbot); si[i] = 0.0; sb[i] = Tc(kbot, kbot) != 0.0 ? Tc(kbot, kbot) : 1.0; }
            }
            if (c->st) { c->st->sweeps++; c->st->shifts_used += ns; }
            int movedist = 64;
            for (int off = 0; off < ns; off += chain) {
                int cnt = MINI(chain, ns - off);
                cnt -= cnt % 2;
                if (cnt < 2) break;
                chain_sweep(c, ktop, kbot, cnt, sr + ks + off, si + ks + off, sb + ks + off, movedist);
            }
        }
    }
    free(sr);
    return info;
}

int orc_hgeqz(int job, int compq, int compz, int n, int ilo, int ihi,
              double *H, int ldh, double *T, int ldt,
              double *alphar, double *alphai, double *beta,
              double *Q, int ldq, double *Z, int ldz,
              const orc_params *prm, orc_stats *st) {
    (void)job;
    if (n < 0) return -4;
    if (ilo < 0 || ilo > MAXI(0, n - 1)) return -5;
    if (ihi < MINI(ilo, n - 1) || ihi >= n) return -6;
    if (n == 0) return 0;
    qzctx c; memset(&c, 0, sizeof c);
    c.n = n; c.H = H; c.T = T; c.ldh = ldh; c.ldt = ldt;
    c.Q = compq ? Q : NULL; c.Z = compz ? Z : NULL; c.ldq = ldq; c.ldz = ldz;
    if (compq == 1) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) Q[IX(i, j, ldq)] = (i == j);
    if (compz == 1) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) Z[IX(i, j, ldz)] = (i == j);
    /* COMMON /PREC/ (PDHGEQZ.f:449-458): TFNORM is PDLANGE('F', NH, NH, T, 1, 1, ...), the leading
     * NH x NH block of T (not the active block T(ILO:IHI, ILO:IHI) when ILO > 1) */
    int nh = ihi - ilo + 1;
    double tf = 0.0;
    for (int j = 0; j < nh; ++j) for (int i = 0; i <= j; ++i) tf += T[IX(i, j, ldt)] * T[IX(i, j, ldt)];
    c.smlnum = SAFMIN * ((double)nh / ULP);
    c.ttol = fmax(SAFMIN, ULP * sqrt(tf));
    c.alphar = alphar; c.alphai = alphai; c.beta = beta; c.st = st;
    if (prm) c.prm = *prm;
    for (int j = ilo; j <= ihi; ++j) alphar[j] = alphai[j] = beta[j] = -1.0;
    /* eigenvalues outside [ilo,ihi] are read off the diagonal (LAPACK DHGEQZ convention) */
    for (int j = 0; j < ilo; ++j) { alphar[j] = H[IX(j, j, ldh)]; alphai[j] = 0.0; beta[j] = T[IX(j, j, ldt)]; }
    for (int j = ihi + 1; j < n; ++j) { alphar[j] = H[IX(j, j, ldh)]; alphai[j] = 0.0; beta[j] = T[IX(j, j, ldt)]; }
    int info = qz_run(&c, ilo, ihi);
    /* final pass: beta >= 0 for 1x1 blocks, standardised 2x2 blocks, eigenvalues read off
     * the Schur form (PDHGEQZ.f:51-57 output convention) */
    for (int j = ilo; j <= ihi; ) {
        if (j < ihi && H[IX(j + 1, j, ldh)] != 0.0) {
            if (T[IX(j, j + 1, ldt)] != 0.0 || T[IX(j, j, ldt)] < 0.0 || T[IX(j + 1, j + 1, ldt)] < 0.0) {
                defl_2x2(&c, j);
            } else {
                double A2[4] = { H[IX(j, j, ldh)], H[IX(j + 1, j, ldh)], H[IX(j, j + 1, ldh)], H[IX(j + 1, j + 1, ldh)] };
                double B2[4] = { T[IX(j, j, ldt)], 0.0, 0.0, T[IX(j + 1, j + 1, ldt)] };
                double ar[2], ai[2], bt[2];
                orc_eig22(A2, 2, B2, 2, ar, ai, bt);
                double b0 = B2[0], b1 = B2[3];
                if (ai[0] != 0.0) {
                    set_eig(&c, j, ar[0] * b0, fabs(ai[0]) * b0, b0);
                    set_eig(&c, j + 1, ar[0] * b1, -fabs(ai[0]) * b1, b1);
                } else defl_2x2(&c, j);
            }
            j += 2;
        } else {
            if (T[IX(j, j, ldt)] == 0.0) set_eig(&c, j, H[IX(j, j, ldh)], 0.0, 0.0);
            else defl_1x1(&c, j);
            j += 1;
        }
    }
    return info;
}

/* ===================================================================================== */
/* PCOMPRES (tools/PQZHLPS.f:19-319)                                                     */
/* ===================================================================================== */
static double fro(int m, int n, const double *A, int ld) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) for (int i = 0; i < m; ++i) s += A[IX(i, j, ld)] * A[IX(i, j, ld)];
    return sqrt(s);
}
void orc_compres(int n, const double *A, const double *B, const double *S, const double *T,
                 const double *Q, const double *Z, int ld, double out[4]) {
    size_t nn = (size_t)n * n;
    double *W = (double *)malloc(sizeof(double) * nn * 2);
    double *R = W + nn;
    /* R = Q^T A Z - S */
    const double *X[2] = { A, B }, *Y[2] = { S, T };
    for (int t = 0; t < 2; ++t) {
        gemm_nn(n, n, n, X[t], ld, Z, ld, W, n);
        for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int p = 0; p < n; ++p) s += Q[IX(p, i, ld)] * W[IX(p, j, n)];
            R[IX(i, j, n)] = s - Y[t][IX(i, j, ld)];
        }
        out[t] = fro(n, n, R, n) / fro(n, n, X[t], ld);
    }
    const double *O[2] = { Q, Z };
    for (int t = 0; t < 2; ++t) {
        double s2 = 0.0;
        for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int p = 0; p < n; ++p) s += O[t][IX(p, i, ld)] * O[t][IX(p, j, ld)];
            s -= (i == j);
            s2 += s * s;
        }
        out[2 + t] = sqrt(s2) / ((double)n * ULP);
    }
    free(W);
}
