// k_mlp.hip -- the learned parameter policy: a 12 -> 64 -> 64 -> 3 perceptron with GELU
// activations in FP32, with hand-written inference, gradient and Adam kernels.
//
// It replaces the PIPARMQ/PILAENVX choice of NS, NW and NIBBLE (piparmq.f:440-517) per top-level
// iteration: the three outputs o are log2 multipliers of the default values,
//     NS = NS_def * 2^clamp(o0), NW = NW_def * 2^clamp(o1), NIBBLE = NIBBLE_def * 2^clamp(o2),
// and the inputs are the deflation-history features of the iteration (see features() in
// qz_host.cpp).  The orthogonal arithmetic of the QZ is unchanged; only these counts move.
//
// Parameter block (5187 floats): W1[64][12] b1[64] W2[64][64] b2[64] W3[3][64] b3[3] (row-major,
// out x in).
#include <hip/hip_runtime.h>
#include "gqz_misc.h"

#define MLP_IN 12
#define MLP_H 64
#define MLP_OUT 3
#define MLP_NPAR (MLP_H * MLP_IN + MLP_H + MLP_H * MLP_H + MLP_H + MLP_OUT * MLP_H + MLP_OUT)

namespace {
constexpr int OW1 = 0, OB1 = OW1 + MLP_H * MLP_IN, OW2 = OB1 + MLP_H, OB2 = OW2 + MLP_H * MLP_H,
              OW3 = OB2 + MLP_H, OB3 = OW3 + MLP_OUT * MLP_H;
static_assert(OB3 + MLP_OUT == MLP_NPAR, "parameter layout");
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// inference: one thread per sample (x: B x 12, y: B x 3, row-major)
__global__ void __launch_bounds__(64) k_mlp_fwd(const float *__restrict__ P, const float *__restrict__ X, int B,
                                                float *__restrict__ Y) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    float x[MLP_IN], a1[MLP_H];
    for (int i = 0; i < MLP_IN; ++i) x[i] = X[(size_t)s * MLP_IN + i];
    for (int j = 0; j < MLP_H; ++j) {
        float z = P[OB1 + j];
        for (int i = 0; i < MLP_IN; ++i) z += P[OW1 + j * MLP_IN + i] * x[i];
        a1[j] = gelu(z);
    }
    float y[MLP_OUT];
    for (int o = 0; o < MLP_OUT; ++o) y[o] = P[OB3 + o];
    for (int k = 0; k < MLP_H; ++k) {
        float z = P[OB2 + k];
        for (int j = 0; j < MLP_H; ++j) z += P[OW2 + k * MLP_H + j] * a1[j];
        const float a2 = gelu(z);
        for (int o = 0; o < MLP_OUT; ++o) y[o] += P[OW3 + o * MLP_H + k] * a2;
    }
    for (int o = 0; o < MLP_OUT; ++o) Y[(size_t)s * MLP_OUT + o] = y[o];
}

// Gradient of L = (1 / (3 B)) sum_s |y_s - t_s|^2 over a batch: 64 samples per workgroup, their
// activations in LDS; the weight gradients are formed as small products over the workgroup's
// samples (no per-sample atomics) and added to G (which must hold zeros or a running sum).
// loss receives the workgroup's share of L.
__global__ void __launch_bounds__(64) k_mlp_grad(const float *__restrict__ P, const float *__restrict__ X,
                                                 const float *__restrict__ Tg, int B, float *__restrict__ G,
                                                 float *__restrict__ loss) {
    __shared__ float sx[64][MLP_IN];
    __shared__ float z1[64][MLP_H + 1], a1[64][MLP_H + 1], z2[64][MLP_H + 1], a2[64][MLP_H + 1];
    __shared__ float dy[64][MLP_OUT];
    __shared__ float red[64];
    const int t = threadIdx.x;
    const int s = blockIdx.x * 64 + t;
    const bool act = s < B;
    const float scale = 2.0f / (3.0f * (float)B);
    // ---- forward (thread = sample) ----
    for (int i = 0; i < MLP_IN; ++i) sx[t][i] = act ? X[(size_t)s * MLP_IN + i] : 0.0f;
    for (int j = 0; j < MLP_H; ++j) {
        float z = P[OB1 + j];
        for (int i = 0; i < MLP_IN; ++i) z += P[OW1 + j * MLP_IN + i] * sx[t][i];
        z1[t][j] = z; a1[t][j] = gelu(z);
    }
    for (int k = 0; k < MLP_H; ++k) {
        float z = P[OB2 + k];
        for (int j = 0; j < MLP_H; ++j) z += P[OW2 + k * MLP_H + j] * a1[t][j];
        z2[t][k] = z; a2[t][k] = gelu(z);
    }
    float l = 0.0f;
    for (int o = 0; o < MLP_OUT; ++o) {
        float y = P[OB3 + o];
        for (int k = 0; k < MLP_H; ++k) y += P[OW3 + o * MLP_H + k] * a2[t][k];
        const float e = act ? y - Tg[(size_t)s * MLP_OUT + o] : 0.0f;
        l += e * e;
        dy[t][o] = e * scale;
    }
    red[t] = l;
    // ---- backward (thread = sample): dz2, then dz1 (kept in z2 / z1) ----
    for (int k = 0; k < MLP_H; ++k) {
        float d = 0.0f;
        for (int o = 0; o < MLP_OUT; ++o) d += P[OW3 + o * MLP_H + k] * dy[t][o];
        z2[t][k] = d * gelu_d(z2[t][k]);
    }
    for (int j = 0; j < MLP_H; ++j) {
        float d = 0.0f;
        for (int k = 0; k < MLP_H; ++k) d += P[OW2 + k * MLP_H + j] * z2[t][k];
        z1[t][j] = d * gelu_d(z1[t][j]);
    }
    __syncthreads();
    // ---- parameter gradients (thread = parameter row), summed over the 64 samples ----
    const int nsmp = 64;
    // W3, b3: 3 x 64 + 3
    for (int e = t; e < MLP_OUT * MLP_H; e += 64) {
        const int o = e / MLP_H, k = e % MLP_H;
        float g = 0.0f;
        for (int q = 0; q < nsmp; ++q) g += dy[q][o] * a2[q][k];
        atomicAdd(G + OW3 + e, g);
    }
    if (t < MLP_OUT) {
        float g = 0.0f;
        for (int q = 0; q < nsmp; ++q) g += dy[q][t];
        atomicAdd(G + OB3 + t, g);
    }
    // W2 (k = row), b2
    {
        const int k = t;
        float gb = 0.0f;
        for (int q = 0; q < nsmp; ++q) gb += z2[q][k];
        atomicAdd(G + OB2 + k, gb);
        for (int j = 0; j < MLP_H; ++j) {
            float g = 0.0f;
            for (int q = 0; q < nsmp; ++q) g += z2[q][k] * a1[q][j];
            atomicAdd(G + OW2 + k * MLP_H + j, g);
        }
    }
    // W1 (j = row), b1
    {
        const int j = t;
        float gb = 0.0f;
        for (int q = 0; q < nsmp; ++q) gb += z1[q][j];
        atomicAdd(G + OB1 + j, gb);
        for (int i = 0; i < MLP_IN; ++i) {
            float g = 0.0f;
            for (int q = 0; q < nsmp; ++q) g += z1[q][j] * sx[q][i];
            atomicAdd(G + OW1 + j * MLP_IN + i, g);
        }
    }
    // ---- loss ----
    for (int h = 32; h > 0; h >>= 1) {
        __syncthreads();
        if (t < h) red[t] += red[t + h];
    }
    if (t == 0 && loss) atomicAdd(loss, red[0] / (3.0f * (float)B));
}

// Adam (bias-corrected) on all parameters; G is cleared for the next accumulation
__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ G, float *__restrict__ M,
                           float *__restrict__ V, int step, float lr, float b1, float b2, float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_NPAR) return;
    const float g = G[i];
    const float m = b1 * M[i] + (1.0f - b1) * g;
    const float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m; V[i] = v;
    const float mh = m / (1.0f - powf(b1, (float)step)), vh = v / (1.0f - powf(b2, (float)step));
    P[i] -= lr * mh / (sqrtf(vh) + eps);
    G[i] = 0.0f;
}

__global__ void k_fill_f(float *x, int n, float v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

int gqz_mlp_nparams() { return MLP_NPAR; }

void gqz_launch_mlp_fwd(hipStream_t st, const float *P, const float *X, int B, float *Y) {
    if (B <= 0) return;
    hipLaunchKernelGGL(k_mlp_fwd, dim3((B + 63) / 64), dim3(64), 0, st, P, X, B, Y);
}
void gqz_launch_mlp_grad(hipStream_t st, const float *P, const float *X, const float *Tg, int B, float *G,
                         float *loss) {
    if (B <= 0) return;
    if (loss) hipLaunchKernelGGL(k_fill_f, dim3(1), dim3(64), 0, st, loss, 1, 0.0f);
    hipLaunchKernelGGL(k_mlp_grad, dim3((B + 63) / 64), dim3(64), 0, st, P, X, Tg, B, G, loss);
}
void gqz_launch_mlp_adam(hipStream_t st, float *P, float *G, float *M, float *V, int step, float lr) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPAR + 255) / 256), dim3(256), 0, st, P, G, M, V, step, lr, 0.9f,
                       0.999f, 1e-8f);
}
void gqz_launch_fill_f(hipStream_t st, float *x, int n, float v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_f, dim3((n + 255) / 256), dim3(256), 0, st, x, n, v);
}
