// gqz_dev.h -- device-side scalar primitives of the multishift QZ (gfx950).
//
// Every helper here is computed by ONE lane (the lane that owns the control flow of a
// workgroup); the O(n) applications of the resulting transforms are spread over the whole
// workgroup by the callers.  Column-major storage, 0-based indices.
//
// Reference semantics (NLAFET/GEVP-PDHGEQZ tree):
//   d_lartg  ~ DLARTG (used e.g. at PDHGEQZB.f:373-406)
//   d_larfg  ~ DLARFG (DHGEQZ5.f:135)
//   d_shift_col ~ QZFCOL (src/kkqz/qzfcol.f:1-116), homogeneous in (alpha, beta)
//   d_invhse3f  ~ INVHSE with ITOPP=2 (src/kkqz/invhse.f:61-71); d_nullrefl3: the same reflector
//                 from the cross product of the two rows when they are well apart
//   d_lagv2     ~ DLAGV2 (PDHGEQZ0.f:622-674)
//   r_tgex2 (gqz_swap.h) ~ DTGEX2 (src/reorder/dtgex2.f), direct swap via the generalized Sylvester eq.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GQZ_IX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

#define GQZ_ULP 2.220446049250313080847e-16
#define GQZ_SAFMIN 2.2250738585072013831e-308

__device__ __forceinline__ void d_lartg(double f, double g, double &c, double &s, double &r) {
    if (g == 0.0) { c = 1.0; s = 0.0; r = f; return; }
    if (f == 0.0) { c = 0.0; s = (g > 0.0) ? 1.0 : -1.0; r = fabs(g); return; }
    double d = hypot(f, g);
    double rr = copysign(d, f);
    c = fabs(f) / d;
    s = g / rr;
    r = rr;
}

// Householder of length n: [alpha; x] -> [beta; 0], v = [1; x'] ; x overwritten with x'.
__device__ inline void d_larfg(int n, double &alpha, double *x, int incx, double &tau) {
    if (n <= 1) { tau = 0.0; return; }
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n - 1; ++i) {
        double v = fabs(x[(size_t)i * incx]);
        if (v != 0.0) {
            if (scale < v) { double r = scale / v; ssq = 1.0 + ssq * r * r; scale = v; }
            else { double r = v / scale; ssq += r * r; }
        }
    }
    double xnorm = scale * sqrt(ssq);
    if (xnorm == 0.0) { tau = 0.0; return; }
    double beta = -copysign(hypot(alpha, xnorm), alpha);
    tau = (beta - alpha) / beta;
    double sc = 1.0 / (alpha - beta);
    for (int i = 0; i < n - 1; ++i) x[(size_t)i * incx] *= sc;
    alpha = beta;
}

// sqrt(a^2 + b^2): the plain formula where neither square can overflow or underflow, hypot elsewhere
__device__ __forceinline__ double d_hyp2(double a, double b) {
    const double m = fmax(fabs(a), fabs(b));
    if (m > 1e-140 && m < 1e140) return sqrt(a * a + b * b);
    return hypot(a, b);
}

// DLARFG for a 3-vector [alpha; x1; x2] (same reflector as d_larfg(3, ...), fewer divisions)
__device__ __forceinline__ void d_larfg3(double &alpha, double &x1, double &x2, double &tau) {
    const double xn = d_hyp2(x1, x2);
    if (xn == 0.0) { tau = 0.0; return; }
    const double beta = -copysign(d_hyp2(alpha, xn), alpha);
    tau = (beta - alpha) / beta;
    const double sc = 1.0 / (alpha - beta);
    x1 *= sc; x2 *= sc;
    alpha = beta;
}

// DLARFG for a 2-vector [alpha; x1]
__device__ __forceinline__ void d_larfg2(double &alpha, double &x1, double &tau) {
    if (x1 == 0.0) { tau = 0.0; return; }
    const double beta = -copysign(d_hyp2(alpha, x1), alpha);
    tau = (beta - alpha) / beta;
    x1 *= 1.0 / (alpha - beta);
    alpha = beta;
}

// ---- short-latency variants for the one-lane chains of the chase and the window QZ ----------
// A division compiles to a dozen dependent instructions, a square root to about as many (with
// scaling for the extremes) and hypot to a library sequence; the chains below use one hardware
// reciprocal refined by two Newton steps (the correctly rounded reciprocal in all but rare cases,
// error below one ulp otherwise) and a square root from the hardware reciprocal square root with
// one coupled Newton step and one residual correction (Markstein; error within one ulp, correctly
// rounded in all but rare cases), on sums of squares in the range where neither can under- or
// overflow (the exact routines elsewhere).  Every reflector takes tau = 2 / (v^T v) of its own
// rounded v, so that it is orthogonal to working precision whatever the rounding of v; the
// rounding of tau itself stays unbiased (round to nearest in all but rare cases), since a biased
// tau makes long products of reflectors lose orthogonality coherently.

// 1/x for 1e-290 < |x| < 1e290, without the range test: the callers below have bounded their
// operands already (|alpha - beta| and d within a small factor of an operand checked to lie in
// (1e-140, 1e140), |p0| at least sqrt(nw) of a checked nw, every tau denominator in [1, 3])
__device__ __forceinline__ double d_rcpn(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// sqrt(x) for 1e-290 < x < 1e290 (no scaling needed): g ~ sqrt(x), h ~ 1 / (2 sqrt(x)) from the
// hardware reciprocal square root; r = 1/2 - g h; g, h <- g + g r, h + h r; g <- g + (x - g^2) h
__device__ __forceinline__ double d_sqrtf(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    return fma(fma(-g, g, x), h, g);
}

// DLARFG for [alpha; x1; x2] (the d_larfg3 reflector: v = [1; x / (alpha - beta)])
__device__ __forceinline__ void d_larfg3f(double &alpha, double &x1, double &x2, double &tau) {
    const double m = fmax(fabs(alpha), fmax(fabs(x1), fabs(x2)));
    if (!(m > 1e-140 && m < 1e140)) { d_larfg3(alpha, x1, x2, tau); return; }
    const double s = fma(x1, x1, x2 * x2);
    if (s == 0.0) { tau = 0.0; return; }
    const double beta = -copysign(d_sqrtf(fma(alpha, alpha, s)), alpha);
    const double rd = d_rcpn(alpha - beta);
    x1 *= rd; x2 *= rd;
    tau = 2.0 * d_rcpn(fma(x1, x1, fma(x2, x2, 1.0)));
    alpha = beta;
}

// DLARFG for [alpha; x1]; tau from the fast reciprocal (for intermediate reflectors whose
// orthogonality only enters through a vector that is reflected again)
__device__ __forceinline__ void d_larfg2f(double &alpha, double &x1, double &tau) {
    if (x1 == 0.0) { tau = 0.0; return; }
    const double m = fmax(fabs(alpha), fabs(x1));
    if (!(m > 1e-140 && m < 1e140)) { d_larfg2(alpha, x1, tau); return; }
    const double beta = -copysign(d_sqrtf(fma(alpha, alpha, x1 * x1)), alpha);
    x1 *= d_rcpn(alpha - beta);
    tau = 2.0 * d_rcpn(fma(x1, x1, 1.0));
    alpha = beta;
}

// DLARTG: c = |f| / d, s = sign(f) g / d, r = sign(f) d with d = sqrt(f^2 + g^2)
__device__ __forceinline__ void d_lartgf(double f, double g, double &c, double &s, double &r) {
    if (g == 0.0) { c = 1.0; s = 0.0; r = f; return; }
    if (f == 0.0) { c = 0.0; s = (g > 0.0) ? 1.0 : -1.0; r = fabs(g); return; }
    const double m = fmax(fabs(f), fabs(g));
    if (!(m > 1e-140 && m < 1e140)) { d_lartg(f, g, c, s, r); return; }
    const double d = d_sqrtf(fma(f, f, g * g));
    const double rd = d_rcpn(d);
    c = fabs(f) * rd;
    s = (f > 0.0 ? g : -g) * rd;
    r = copysign(d, f);
}

// Eigenvalues of a 2x2 pencil (A, B), B upper triangular with non-zero diagonal:
// lambda_k = (ar_k + i ai_k), b = 1.  Standard 2x2 formula on M = B^{-1} A.
__device__ inline void d_eig22(double a00, double a10, double a01, double a11,
                               double b00, double b01, double b11,
                               double ar[2], double ai[2]) {
    double m10 = a10 / b11, m11 = a11 / b11;
    double m00 = (a00 - b01 * m10) / b00, m01 = (a01 - b01 * m11) / b00;
    double p = 0.5 * (m00 - m11);
    double bcmax = fmax(fabs(m01), fabs(m10));
    double bcmis = fmin(fabs(m01), fabs(m10)) * copysign(1.0, m01) * copysign(1.0, m10);
    double scale = fmax(fabs(p), bcmax);
    double z = (scale == 0.0) ? 0.0 : (p / scale) * p + (bcmax / scale) * bcmis;
    if (z >= 0.0) {
        double zz = p + copysign(sqrt(scale) * sqrt(z), p);
        ar[0] = m11 + zz;
        ar[1] = (zz != 0.0) ? m11 - (bcmax / zz) * bcmis : m11;
        ai[0] = ai[1] = 0.0;
    } else {
        double re = 0.5 * (m00 + m11);
        double disc = -((m00 - m11) * (m00 - m11) * 0.25 + m01 * m10);
        double im = sqrt(fmax(disc, 0.0));
        ar[0] = ar[1] = re; ai[0] = im; ai[1] = -im;
    }
}

// First column of the double-shift polynomial of (H, T) at the top-left corner.
// real pair: b1 b2 M^2 - (a1 b2 + a2 b1) M + a1 a2 ; complex: b1^2 M^2 - 2 b1 ar M + |a|^2.
__device__ inline void d_shift_col(const double *H, int ldh, const double *T, int ldt,
                                   double sr1, double si1, double sb1, double sr2, double sb2,
                                   double v[3]) {
    double h00 = H[GQZ_IX(0, 0, ldh)], h01 = H[GQZ_IX(0, 1, ldh)], h10 = H[GQZ_IX(1, 0, ldh)];
    double h11 = H[GQZ_IX(1, 1, ldh)], h21 = H[GQZ_IX(2, 1, ldh)];
    double t00 = T[GQZ_IX(0, 0, ldt)], t01 = T[GQZ_IX(0, 1, ldt)], t11 = T[GQZ_IX(1, 1, ldt)];
    double c2, c1, c0;
    if (si1 != 0.0) { c2 = sb1 * sb1; c1 = 2.0 * sb1 * sr1; c0 = sr1 * sr1 + si1 * si1; }
    else { c2 = sb1 * sb2; c1 = sr1 * sb2 + sr2 * sb1; c0 = sr1 * sr2; }
    double x0 = h00 / t00, x1 = h10 / t00;
    double z1 = x1 / t11;
    double z0 = (x0 - t01 * z1) / t00;
    double y0 = h00 * z0 + h01 * z1, y1 = h10 * z0 + h11 * z1, y2 = h21 * z1;
    v[0] = c2 * y0 - c1 * x0 + c0;
    v[1] = c2 * y1 - c1 * x1;
    v[2] = c2 * y2;
    double m = fmax(fabs(v[0]), fmax(fabs(v[1]), fabs(v[2])));
    if (m > 0.0 && isfinite(m)) { v[0] /= m; v[1] /= m; v[2] /= m; }
}

// Opposite Householder of a 3x3 block of T (INVHSE): its first column spans the vector orthogonal
// to rows 1 and 2 (first row of the orthogonal factor of an RQ decomposition), built from rows 1
// and 2 only with the short-latency primitives.  The reflector maps the unit vector w onto
// -sign(w0) e0 with v = w + sign(w0) e0 scaled to v0 = 1 and tau = 2 / (v^T v) from v itself (a
// norm-based DLARFG on a unit vector rounds the norm to 1 with a bias that makes the accumulated Z
// lose orthogonality coherently).
__device__ __forceinline__ void d_invhse3f(double r10, double r11, double r12, double r20, double r21, double r22,
                                           double v[3], double &tau) {
    double a = r22, x0 = r20, x1 = r21, tau1;
    d_larfg3f(a, x0, x1, tau1);                         // u1 = (x0, x1, 1) on columns 0, 1, 2
    const double s = tau1 * (r10 * x0 + r11 * x1 + r12);
    double a2 = r11 - s * x1, y = r10 - s * x0, tau2;  // row 1 after the first reflector
    d_larfg2f(a2, y, tau2);                             // u2 = (y, 1) on columns 0, 1
    const double s2 = tau2 * y;
    double w0 = 1.0 - s2 * y, w1 = -s2;
    const double s1 = tau1 * (w0 * x0 + w1 * x1);
    w0 -= s1 * x0; w1 -= s1 * x1;
    const double w2 = -s1;
    const double p0 = w0 + (w0 >= 0.0 ? 1.0 : -1.0);
    const double rp = d_rcpn(p0);
    const double p1 = w1 * rp, p2 = w2 * rp;
    tau = 2.0 * d_rcpn(fma(p1, p1, fma(p2, p2, 1.0)));
    v[0] = 1.0; v[1] = p1; v[2] = p2;
}

// The same right reflector from the cross product w = r1 x r2 of rows 1 and 2 (w spans their
// common null space), reflected onto -sign(w0) |w| e0.  The computed w carries an error of order
// ulp |r1| |r2|, i.e. a residual r_i . w / |w| of order ulp |r_i| / sin(r1, r2); the cross product
// is used where the rows are at least 1/16 (in sine) from parallel, d_invhse3f elsewhere.
__device__ __forceinline__ void d_nullrefl3(double r10, double r11, double r12, double r20, double r21, double r22,
                                            double v[3], double &tau) {
    const double w0 = fma(r11, r22, -r12 * r21);
    const double w1 = fma(r12, r20, -r10 * r22);
    const double w2 = fma(r10, r21, -r11 * r20);
    const double n1 = fma(r10, r10, fma(r11, r11, r12 * r12));
    const double n2 = fma(r20, r20, fma(r21, r21, r22 * r22));
    const double nw = fma(w0, w0, fma(w1, w1, w2 * w2));
    if (!(fmin(n1, n2) > 1e-130 && fmax(n1, n2) < 1e130 && nw >= 0.00390625 * (n1 * n2))) {
        d_invhse3f(r10, r11, r12, r20, r21, r22, v, tau);
        return;
    }
    const double p0 = w0 + copysign(d_sqrtf(nw), w0);
    const double rp = d_rcpn(p0);
    const double p1 = w1 * rp, p2 = w2 * rp;
    tau = 2.0 * d_rcpn(fma(p1, p1, fma(p2, p2, 1.0)));
    v[0] = 1.0; v[1] = p1; v[2] = p2;
}

// 2x2 transforms of two rows / two columns of small column-major arrays
__device__ inline void d_g2_rows(double *M, int ld, int i, int c0, int c1, const double G[4]) {
    for (int j = c0; j < c1; ++j) {
        double x = M[GQZ_IX(i, j, ld)], y = M[GQZ_IX(i + 1, j, ld)];
        M[GQZ_IX(i, j, ld)] = G[0] * x + G[1] * y;
        M[GQZ_IX(i + 1, j, ld)] = G[2] * x + G[3] * y;
    }
}
__device__ inline void d_g2_cols(double *M, int ld, int j, int r0, int r1, const double G[4]) {
    for (int i = r0; i < r1; ++i) {
        double a = M[GQZ_IX(i, j, ld)], b = M[GQZ_IX(i, j + 1, ld)];
        M[GQZ_IX(i, j, ld)] = a * G[0] + b * G[1];
        M[GQZ_IX(i, j + 1, ld)] = a * G[2] + b * G[3];
    }
}

// Standardise a 2x2 pencil held in A (2x2, col-major, ld 2) and B (upper triangular):
// A <- QL^T A ZR, B <- QL^T B ZR; real eigenvalues -> triangular, complex -> B diagonal > 0.
// Returns 1 for a complex pair.
__device__ inline int d_lagv2(double *A, double *B, double QL[4], double ZR[4]) {
    double ar[2], ai[2];
    d_eig22(A[0], A[1], A[2], A[3], B[0], B[2], B[3], ar, ai);
    double anorm = fabs(A[0]) + fabs(A[1]) + fabs(A[2]) + fabs(A[3]);
    double bnorm = fabs(B[0]) + fabs(B[2]) + fabs(B[3]);
    double c, s, r;
    if (ai[0] == 0.0) {
        double lam = ar[0];
        double h00 = A[0] - lam * B[0], h01 = A[2] - lam * B[2], h10 = A[1], h11 = A[3] - lam * B[3];
        if (fabs(h00) + fabs(h01) >= fabs(h10) + fabs(h11)) d_lartg(h01, -h00, c, s, r);
        else d_lartg(h11, -h10, c, s, r);
        ZR[0] = c; ZR[1] = s; ZR[2] = -s; ZR[3] = c;
        d_g2_cols(A, 2, 0, 0, 2, ZR);
        d_g2_cols(B, 2, 0, 0, 2, ZR);
        double na = fabs(A[0]) + fabs(A[1]), nb = fabs(B[0]) + fabs(B[1]);
        if (na * bnorm >= nb * anorm) d_lartg(A[0], A[1], c, s, r);
        else d_lartg(B[0], B[1], c, s, r);
        QL[0] = c; QL[1] = s; QL[2] = -s; QL[3] = c;
        d_g2_rows(A, 2, 0, 0, 2, QL);
        d_g2_rows(B, 2, 0, 0, 2, QL);
        A[1] = 0.0; B[1] = 0.0;
        return 0;
    }
    double f = B[0], g = B[2], h = B[3];
    double p = f * f, q = f * g, w = g * g + h * h;
    double c2 = 1.0, s2 = 0.0;
    if (q != 0.0) {
        double theta = 0.5 * (w - p) / q;
        double t = copysign(1.0, theta) / (fabs(theta) + sqrt(1.0 + theta * theta));
        { double rr; d_lartg(1.0, t, c2, s2, rr); }
    }
    ZR[0] = c2; ZR[1] = -s2; ZR[2] = s2; ZR[3] = c2;
    d_g2_cols(A, 2, 0, 0, 2, ZR);
    d_g2_cols(B, 2, 0, 0, 2, ZR);
    double b0x = B[0], b0y = B[1], b1x = B[2], b1y = B[3];
    double s0 = hypot(b0x, b0y), s1 = hypot(b1x, b1y);
    // B ZR has orthogonal columns b0, b1.  QL^T is the plane rotation (or, when needed to keep both
    // diagonal entries positive, reflection) that maps the longer of them onto a coordinate axis.
    // (Normalising both columns separately, Gram-Schmidt style, leaves an orthogonality defect of
    // one sign that accumulates linearly over many standardisations.)
    double m00, m01, m10, m11;
    if (s0 >= s1) {
        d_lartg(b0x, b0y, c, s, r);
        if (r < 0.0) { c = -c; s = -s; }
        m00 = c; m01 = s;
        if (-s * b1x + c * b1y >= 0.0) { m10 = -s; m11 = c; }   // rotation
        else { m10 = s; m11 = -c; }                             // reflection
    } else {
        d_lartg(b1y, -b1x, c, s, r);
        if (r < 0.0) { c = -c; s = -s; }
        m10 = -s; m11 = c;
        if (c * b0x + s * b0y >= 0.0) { m00 = c; m01 = s; }
        else { m00 = -c; m01 = -s; }
    }
    QL[0] = m00; QL[1] = m01; QL[2] = m10; QL[3] = m11;
    d_g2_rows(A, 2, 0, 0, 2, QL);
    d_g2_rows(B, 2, 0, 0, 2, QL);
    B[1] = 0.0; B[2] = 0.0;
    return 1;
}

// (alpha_r, alpha_i, beta) of a standardised complex 2x2 block (B diagonal)
__device__ inline void d_pair_eigs(const double *A, const double *B, double ar[2], double ai[2],
                                   double be[2]) {
    double er[2], ei[2];
    d_eig22(A[0], A[1], A[2], A[3], B[0], 0.0, B[3], er, ei);
    double b0 = B[0], b1 = B[3];
    if (ei[0] != 0.0) {
        ar[0] = er[0] * b0; ai[0] = fabs(ei[0]) * b0; be[0] = b0;
        ar[1] = er[0] * b1; ai[1] = -fabs(ei[0]) * b1; be[1] = b1;
    } else {
        ar[0] = er[0] * b0; ai[0] = 0.0; be[0] = b0;
        ar[1] = er[1] * b1; ai[1] = 0.0; be[1] = b1;
    }
}

