/* CPU oracle (test infrastructure only) of the learned parameter policy's perceptron:
 * 12 -> 64 -> 64 -> 3, GELU (exact, erf form), parameters W1[64][12] b1[64] W2[64][64] b2[64]
 * W3[3][64] b3[3] (row-major, out x in), loss L = mean over samples and outputs of (y - t)^2.
 * Double precision, straightforward loops: the checker for the FP32 HIP kernels (k_mlp.hip).
 * It has no counterpart in the reference; it is pinned by central finite differences of L in the
 * tests. */
#include <math.h>
#include <string.h>

#include "qz_oracle.h"

#define NI 12
#define NH 64
#define NO 3
enum { OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NO * NH,
       NPAR = OB3 + NO };

static double gelu(double x) { return 0.5 * x * (1.0 + erf(x / sqrt(2.0))); }
static double gelu_d(double x) { return 0.5 * (1.0 + erf(x / sqrt(2.0))) + x * exp(-0.5 * x * x) / 2.5066282746310002; }

int orc_mlp_nparams(void) { return NPAR; }

void orc_mlp_forward(const double *P, const double *X, int B, double *Y) {
    for (int s = 0; s < B; ++s) {
        const double *x = X + (size_t)s * NI;
        double a1[NH], a2[NH];
        for (int j = 0; j < NH; ++j) {
            double z = P[OB1 + j];
            for (int i = 0; i < NI; ++i) z += P[OW1 + j * NI + i] * x[i];
            a1[j] = gelu(z);
        }
        for (int k = 0; k < NH; ++k) {
            double z = P[OB2 + k];
            for (int j = 0; j < NH; ++j) z += P[OW2 + k * NH + j] * a1[j];
            a2[k] = gelu(z);
        }
        for (int o = 0; o < NO; ++o) {
            double y = P[OB3 + o];
            for (int k = 0; k < NH; ++k) y += P[OW3 + o * NH + k] * a2[k];
            Y[(size_t)s * NO + o] = y;
        }
    }
}

/* G (NPAR) is overwritten with dL/dP; returns L */
double orc_mlp_grad(const double *P, const double *X, const double *T, int B, double *G) {
    memset(G, 0, sizeof(double) * NPAR);
    double L = 0.0;
    const double sc = 2.0 / (3.0 * B);
    for (int s = 0; s < B; ++s) {
        const double *x = X + (size_t)s * NI;
        double z1[NH], a1[NH], z2[NH], a2[NH], dy[NO], d2[NH], d1[NH];
        for (int j = 0; j < NH; ++j) {
            double z = P[OB1 + j];
            for (int i = 0; i < NI; ++i) z += P[OW1 + j * NI + i] * x[i];
            z1[j] = z; a1[j] = gelu(z);
        }
        for (int k = 0; k < NH; ++k) {
            double z = P[OB2 + k];
            for (int j = 0; j < NH; ++j) z += P[OW2 + k * NH + j] * a1[j];
            z2[k] = z; a2[k] = gelu(z);
        }
        for (int o = 0; o < NO; ++o) {
            double y = P[OB3 + o];
            for (int k = 0; k < NH; ++k) y += P[OW3 + o * NH + k] * a2[k];
            double e = y - T[(size_t)s * NO + o];
            L += e * e / (3.0 * B);
            dy[o] = sc * e;
        }
        for (int o = 0; o < NO; ++o) {
            G[OB3 + o] += dy[o];
            for (int k = 0; k < NH; ++k) G[OW3 + o * NH + k] += dy[o] * a2[k];
        }
        for (int k = 0; k < NH; ++k) {
            double d = 0.0;
            for (int o = 0; o < NO; ++o) d += P[OW3 + o * NH + k] * dy[o];
            d2[k] = d * gelu_d(z2[k]);
            G[OB2 + k] += d2[k];
            for (int j = 0; j < NH; ++j) G[OW2 + k * NH + j] += d2[k] * a1[j];
        }
        for (int j = 0; j < NH; ++j) {
            double d = 0.0;
            for (int k = 0; k < NH; ++k) d += P[OW2 + k * NH + j] * d2[k];
            d1[j] = d * gelu_d(z1[j]);
            G[OB1 + j] += d1[j];
            for (int i = 0; i < NI; ++i) G[OW1 + j * NI + i] += d1[j] * x[i];
        }
    }
    return L;
}
