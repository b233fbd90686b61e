// k_batched.hip -- thousands of independent small pencils in one persistent launch (no
// counterpart in the reference, which solves one pencil per call).
//
// Each workgroup repeatedly takes the next pencil from an atomic work counter (pencils are
// ordered by decreasing order so the longest jobs start first) and runs the whole pipeline on
// it: blocked Householder QR of B applied to A (the EXRAND2.f:296-329 pre-step, compact WY panels with
// FP64-MFMA updates inside the workgroup, gqz_qrwg.h), Givens HT reduction
// (PDGGHRD semantics) and the double-shift base QZ with infinite-eigenvalue deflation
// (KKQZ/QZLAP role), accumulating Q and Z when requested.  Matrices stay in HBM (L2-resident
// for the sizes of the batched configuration); every step is workgroup-cooperative.
#include <algorithm>
#include <numeric>
#include <vector>

#include "gqz_ds.h"
#include "gqz_misc.h"
#include "gqz_qrwg.h"

extern "C" void *gqz_stream(void);

struct BatchArgs {
    int count;
    const int *order;         // processing order (indices into n/off/eoff)
    const int *n;
    const long *off, *eoff;
    double *A, *B, *Q, *Z;
    double *ar, *ai, *be;
    int *info;
    int *counter;
};

template <int NT>
__global__ void __launch_bounds__(NT) k_batched(BatchArgs a) {
    __shared__ WgCtl Cs;
    __shared__ QrwLds qs;
    __shared__ double red[NT];
    __shared__ double sh[8];
    __shared__ int s_job;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) {
            int k = atomicAdd(a.counter, 1);
            s_job = (k < a.count) ? a.order[k] : -1;
        }
        __syncthreads();
        const int job = s_job;
        if (job < 0) break;
        const int n = a.n[job];
        const long off = a.off[job], eo = a.eoff[job];
        double *A = a.A + off, *B = a.B + off;
        double *Q = a.Q ? a.Q + off : nullptr, *Z = a.Z ? a.Z + off : nullptr;
        double *ar = a.ar + eo, *ai = a.ai + eo, *be = a.be + eo;
        // Q = I, Z = I
        for (int idx = tid; idx < n * n; idx += NT) {
            int i = idx % n, j = idx / n;
            if (Q) Q[GQZ_IX(i, j, n)] = (i == j) ? 1.0 : 0.0;
            if (Z) Z[GQZ_IX(i, j, n)] = (i == j) ? 1.0 : 0.0;
        }
        __syncthreads();
        wg_qr_b_blk<NT>(A, B, Q, n, n, qs);
        wg_gghrd(A, B, n, n, 0, n - 1, Q, Z, n, n, sh);
        // TTOL = ulp * ||T||_F (PDHGEQZ.f:456-458)
        double ss = 0.0;
        for (int idx = tid; idx < n * n; idx += NT) { double v = B[idx]; ss += v * v; }
        red[tid] = ss;
        __syncthreads();
        for (int s = NT / 2; s > 0; s >>= 1) { if (tid < s) red[tid] += red[tid + s]; __syncthreads(); }
        const double ttol = fmax(GQZ_SAFMIN, GQZ_ULP * sqrt(red[0]));
        const double smlnum = GQZ_SAFMIN * ((double)n / GQZ_ULP);
        __syncthreads();
        WgPencil P;
        P.H = A; P.T = B; P.U = Q; P.V = Z; P.ld = n; P.ldu = n; P.n = n; P.nq = n;
        int info = wg_qz_ds(P, &Cs, 0, n - 1, ar, ai, be, ttol, smlnum);
        // ds_decide already returns beta >= 0 (columns negated as blocks deflate)
        if (tid == 0 && a.info) a.info[job] = info;
        __syncthreads();
    }
}

// counter reset as a kernel (a graph replays it; no memset node)
__global__ void k_zero_int(int *p) { if (threadIdx.x == 0) *p = 0; }

// N(0,1) fill for the tuner's sample pencils: SplitMix64 hash of (seed, index) + Box-Muller
__global__ void k_randn(double *x, long n, unsigned long long seed) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(2 * i + 1);
    auto mix = [](unsigned long long v) {
        v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
        v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
        return v ^ (v >> 31);
    };
    double u1 = ((mix(z) >> 11) + 1) * (1.0 / 9007199254740993.0);
    double u2 = (mix(z + 0x9E3779B97F4A7C15ull) >> 11) * (1.0 / 9007199254740992.0);
    x[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// ---- scratch of the allocation-free entry point ------------------------------------------------
// One file-scope device array per GPU holds the work counter and the per-pencil tables (processing
// order, orders, matrix and eigenvalue offsets) of gqz_batched_noalloc_dev.  In host code the name
// `scratch` is only a stand-in: its device address is taken once per device with
// hipGetSymbolAddress (gqz_batched_scratch_setup, called by gqz_init) and only that address is used.
#define GQZ_NOALLOC_MAXC 8192
#define NOALLOC_ORDER 256
#define NOALLOC_N (NOALLOC_ORDER + 4 * GQZ_NOALLOC_MAXC)
#define NOALLOC_OFF (NOALLOC_N + 4 * GQZ_NOALLOC_MAXC)
#define NOALLOC_EOFF (NOALLOC_OFF + 8 * GQZ_NOALLOC_MAXC)
#define NOALLOC_BYTES (NOALLOC_EOFF + 8 * GQZ_NOALLOC_MAXC)
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

#define NOALLOC_MAXDEV 64
unsigned char *g_scratch_dev[NOALLOC_MAXDEV] = {};   // device address of `scratch` on each GPU
unsigned char *g_stage = nullptr;                      // pinned host staging of the tables

struct Batch {
    int *d_meta = nullptr;        // order[count] | n[count] | counter
    long *d_off = nullptr;        // off[count] | eoff[count]
    int cap = 0;
    // graph cache
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<long> key;
    // tuned launch shape (0 = untuned)
    int tuned_nt = 0, tuned_nwg = 0;
};
thread_local Batch B;             // per host thread (its buffers and graph belong to its stream)

void launch(hipStream_t st, const BatchArgs &a, int nt, int nwg) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, a.counter);
    switch (nt) {
    case 128: hipLaunchKernelGGL(k_batched<128>, dim3(nwg), dim3(128), 0, st, a); break;
    case 512: hipLaunchKernelGGL(k_batched<512>, dim3(nwg), dim3(512), 0, st, a); break;
    case 1024: hipLaunchKernelGGL(k_batched<1024>, dim3(nwg), dim3(1024), 0, st, a); break;
    default: hipLaunchKernelGGL(k_batched<256>, dim3(nwg), dim3(256), 0, st, a); break;
    }
}

void ensure_cap(int count) {
    if (B.cap >= count) return;
    if (B.d_meta) { GQZ_HIP(hipFree(B.d_meta)); GQZ_HIP(hipFree(B.d_off)); }
    GQZ_HIP(hipMalloc((void **)&B.d_meta, (2 * (size_t)count + 1) * sizeof(int)));
    GQZ_HIP(hipMalloc((void **)&B.d_off, 2 * (size_t)count * sizeof(long)));
    B.cap = count;
}

// heuristic shape when the tuner has not run: wider groups for larger pencils
void default_shape(int maxn, int count, int &nt, int &nwg) {
    nt = maxn <= 96 ? 128 : (maxn <= 384 ? 256 : 512);
    nwg = std::min(count, 256 * (nt <= 256 ? 4 : 2));
}

}  // namespace

extern "C" {

int gqz_batched_dev(int count, const int *n, const long *off, const long *eoff, double *A, double *Bm, double *Q,
                    double *Z, double *alphar, double *alphai, double *beta, int *info, int wg_threads) {
    if (count < 0) return -1;
    if (count == 0) return 0;
    if (!n || !off || !eoff) return -2;
    if (!A || !Bm || !alphar || !alphai || !beta) return -5;
    if (wg_threads != 0 && wg_threads != 128 && wg_threads != 256 && wg_threads != 512 && wg_threads != 1024)
        return -13;
    int maxn = 0;
    for (int i = 0; i < count; ++i) {
        if (n[i] < 1) return -2;
        maxn = std::max(maxn, n[i]);
    }
    try {
        if (gqz_init_current()) return -100;
        hipStream_t st = (hipStream_t)gqz_stream();
        ensure_cap(count);
        // largest pencils first so that the tail of the persistent loop is short
        std::vector<int> order(count);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
        std::vector<int> meta(2 * (size_t)count);
        std::copy(order.begin(), order.end(), meta.begin());
        std::copy(n, n + count, meta.begin() + count);
        std::vector<long> offs(2 * (size_t)count);
        std::copy(off, off + count, offs.begin());
        std::copy(eoff, eoff + count, offs.begin() + count);
        GQZ_HIP(hipMemcpyAsync(B.d_meta, meta.data(), meta.size() * sizeof(int), hipMemcpyHostToDevice, st));
        GQZ_HIP(hipMemcpyAsync(B.d_off, offs.data(), offs.size() * sizeof(long), hipMemcpyHostToDevice, st));
        int nt, nwg;
        default_shape(maxn, count, nt, nwg);
        if (wg_threads) nt = wg_threads;
        else if (B.tuned_nt) { nt = B.tuned_nt; nwg = std::min(count, B.tuned_nwg); }
        BatchArgs a;
        a.count = count; a.order = B.d_meta; a.n = B.d_meta + count; a.off = B.d_off; a.eoff = B.d_off + count;
        a.A = A; a.B = Bm; a.Q = Q; a.Z = Z; a.ar = alphar; a.ai = alphai; a.be = beta; a.info = info;
        a.counter = B.d_meta + 2 * count;
        // the launch pair (counter reset + persistent kernel) is captured once per argument set and
        // replayed as a hipGraph afterwards
        std::vector<long> key = { (long)count, (long)nt, (long)nwg, (long)(size_t)A, (long)(size_t)Bm,
                                  (long)(size_t)Q, (long)(size_t)Z, (long)(size_t)alphar, (long)(size_t)alphai,
                                  (long)(size_t)beta, (long)(size_t)info, (long)(size_t)B.d_meta };
        if (key != B.key || !B.gexec) {
            if (B.gexec) { GQZ_HIP(hipGraphExecDestroy(B.gexec)); B.gexec = nullptr; }
            if (B.graph) { GQZ_HIP(hipGraphDestroy(B.graph)); B.graph = nullptr; }
            GQZ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            launch(st, a, nt, nwg);
            GQZ_HIP(hipStreamEndCapture(st, &B.graph));
            GQZ_HIP(hipGraphInstantiate(&B.gexec, B.graph, nullptr, nullptr, 0));
            B.key = key;
        }
        GQZ_HIP(hipGraphLaunch(B.gexec, st));
        GQZ_HIP(hipStreamSynchronize(st));
        return 0;
    } catch (const std::exception &e) {
        gqz_set_error(e.what());
        return -100;
    }
}

// Allocation-free variant of gqz_batched_dev: the same kernels, launched directly (no graph, whose
// instantiation allocates), with the counter and the tables in the per-device `scratch` array and the
// host-side tables in the pinned staging buffer that gqz_init set up.  No hipMalloc / hipFree.
int gqz_batched_noalloc_dev(int count, const int *n, const long *off, const long *eoff, double *A, double *Bm,
                            double *Q, double *Z, double *alphar, double *alphai, double *beta, int *info,
                            int wg_threads) {
    if (count < 0) return -1;
    if (count == 0) return 0;
    if (count > GQZ_NOALLOC_MAXC) {
        gqz_set_error("gqz_batched_noalloc_dev: more pencils than the allocation-free limit of 8192");
        return -1;
    }
    if (!n || !off || !eoff) return -2;
    if (!A || !Bm || !alphar || !alphai || !beta) return -5;
    if (wg_threads != 0 && wg_threads != 128 && wg_threads != 256 && wg_threads != 512 && wg_threads != 1024)
        return -13;
    int maxn = 0;
    for (int i = 0; i < count; ++i) {
        if (n[i] < 1) return -2;
        maxn = std::max(maxn, n[i]);
    }
    try {
        const int dev = gqz_ready_device();
        if (dev < 0 || dev >= NOALLOC_MAXDEV || !g_scratch_dev[dev] || !g_stage) {
            gqz_set_error("gqz_batched_noalloc_dev: call gqz_init on this device first (it sets up the scratch)");
            return -100;
        }
        unsigned char *scr = g_scratch_dev[dev];
        hipStream_t st = (hipStream_t)gqz_stream();
        int *h_order = (int *)(g_stage + NOALLOC_ORDER), *h_n = (int *)(g_stage + NOALLOC_N);
        long *h_off = (long *)(g_stage + NOALLOC_OFF), *h_eoff = (long *)(g_stage + NOALLOC_EOFF);
        for (int i = 0; i < count; ++i) { h_order[i] = i; h_n[i] = n[i]; h_off[i] = off[i]; h_eoff[i] = eoff[i]; }
        // largest pencils first, ties by index (std::sort works in place)
        std::sort(h_order, h_order + count, [&](int x, int y) { return n[x] != n[y] ? n[x] > n[y] : x < y; });
        GQZ_HIP(hipMemcpyAsync(scr + NOALLOC_ORDER, h_order, 4 * (size_t)count, hipMemcpyHostToDevice, st));
        GQZ_HIP(hipMemcpyAsync(scr + NOALLOC_N, h_n, 4 * (size_t)count, hipMemcpyHostToDevice, st));
        GQZ_HIP(hipMemcpyAsync(scr + NOALLOC_OFF, h_off, 8 * (size_t)count, hipMemcpyHostToDevice, st));
        GQZ_HIP(hipMemcpyAsync(scr + NOALLOC_EOFF, h_eoff, 8 * (size_t)count, hipMemcpyHostToDevice, st));
        int nt, nwg;
        default_shape(maxn, count, nt, nwg);
        if (wg_threads) nt = wg_threads;
        else if (B.tuned_nt) { nt = B.tuned_nt; nwg = std::min(count, B.tuned_nwg); }
        BatchArgs a;
        a.count = count; a.order = (const int *)(scr + NOALLOC_ORDER); a.n = (const int *)(scr + NOALLOC_N);
        a.off = (const long *)(scr + NOALLOC_OFF); a.eoff = (const long *)(scr + NOALLOC_EOFF);
        a.A = A; a.B = Bm; a.Q = Q; a.Z = Z; a.ar = alphar; a.ai = alphai; a.be = beta; a.info = info;
        a.counter = (int *)scr;            // cleared by the k_zero_int launch in launch()
        launch(st, a, nt, nwg);
        GQZ_HIP(hipGetLastError());
        GQZ_HIP(hipStreamSynchronize(st));
        return 0;
    } catch (const std::exception &e) {
        gqz_set_error(e.what());
        return -100;
    }
}

// Called by gqz_init with `device` current: the device address of `scratch` on this GPU, and (once per
// process) the pinned staging of the host-side tables.  0 on success.
int gqz_batched_scratch_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAXDEV) return -1;
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            gqz_set_error((std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e)).c_str());
            return -1;
        }
        g_scratch_dev[device] = (unsigned char *)p;
    }
    if (!g_stage) {
        const hipError_t e = hipHostMalloc((void **)&g_stage, NOALLOC_BYTES, hipHostMallocDefault);
        if (e != hipSuccess) {
            g_stage = nullptr;
            gqz_set_error((std::string("hipHostMalloc (staging) failed: ") + hipGetErrorString(e)).c_str());
            return -1;
        }
    }
    return 0;
}

int gqz_batched_tune(int count, const int *n, int *best_threads, int *best_wgs, double *best_ms) {
    if (count < 1 || !n) return -1;
    try {
        if (gqz_init_current()) return -100;
        hipStream_t st = (hipStream_t)gqz_stream();
        // sample: up to 64 pencils at the quantiles of the requested size distribution, orders capped
        // at 512 so that one candidate stays short
        std::vector<int> s(n, n + count);
        std::sort(s.begin(), s.end());
        const int ns = std::min(count, 64);
        std::vector<int> sn(ns);
        std::vector<long> so(ns), se(ns);
        long tot = 0, et = 0;
        for (int i = 0; i < ns; ++i) {
            sn[i] = std::min(512, s[(size_t)i * count / ns]);
            so[i] = tot; se[i] = et;
            tot += (long)sn[i] * sn[i]; et += sn[i];
        }
        double *buf = nullptr;
        GQZ_HIP(hipMalloc((void **)&buf, (6 * (size_t)tot + 3 * (size_t)et) * sizeof(double)));
        double *A0 = buf, *B0 = A0 + tot, *A = B0 + tot, *Bm = A + tot, *Q = Bm + tot, *Z = Q + tot;
        double *ar = Z + tot, *ai = ar + et, *be = ai + et;
        hipLaunchKernelGGL(k_randn, dim3((unsigned)((2 * tot + 255) / 256)), dim3(256), 0, st, A0, 2 * tot,
                           0x5EEDull);
        hipEvent_t e0, e1;
        GQZ_HIP(hipEventCreate(&e0)); GQZ_HIP(hipEventCreate(&e1));
        const int cand_nt[3] = { 128, 256, 512 };
        float best = 1e30f;
        int bnt = 256, bnwg = std::min(ns, 512);
        ensure_cap(ns);
        std::vector<int> meta(2 * (size_t)ns);
        std::vector<int> order(ns);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return sn[x] > sn[y]; });
        std::copy(order.begin(), order.end(), meta.begin());
        std::copy(sn.begin(), sn.end(), meta.begin() + ns);
        std::vector<long> offs(2 * (size_t)ns);
        std::copy(so.begin(), so.end(), offs.begin());
        std::copy(se.begin(), se.end(), offs.begin() + ns);
        GQZ_HIP(hipMemcpyAsync(B.d_meta, meta.data(), meta.size() * sizeof(int), hipMemcpyHostToDevice, st));
        GQZ_HIP(hipMemcpyAsync(B.d_off, offs.data(), offs.size() * sizeof(long), hipMemcpyHostToDevice, st));
        BatchArgs a;
        a.count = ns; a.order = B.d_meta; a.n = B.d_meta + ns; a.off = B.d_off; a.eoff = B.d_off + ns;
        a.A = A; a.B = Bm; a.Q = Q; a.Z = Z; a.ar = ar; a.ai = ai; a.be = be; a.info = nullptr;
        a.counter = B.d_meta + 2 * ns;
        // every sample pencil gets its own workgroup, as the leading (largest) pencils of a real batch
        // do; the workgroup size is what the sample decides, the grid follows from it
        for (int ci = 0; ci < 3; ++ci) {
            const int nt = cand_nt[ci];
            GQZ_HIP(hipMemcpyAsync(A, A0, 2 * (size_t)tot * sizeof(double), hipMemcpyDeviceToDevice, st));
            GQZ_HIP(hipEventRecord(e0, st));
            launch(st, a, nt, ns);
            GQZ_HIP(hipEventRecord(e1, st));
            GQZ_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            GQZ_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; bnt = nt; }
        }
        bnwg = 256 * (bnt <= 256 ? 4 : 2);
        GQZ_HIP(hipEventDestroy(e0)); GQZ_HIP(hipEventDestroy(e1));
        GQZ_HIP(hipFree(buf));
        B.tuned_nt = bnt; B.tuned_nwg = bnwg;
        B.key.clear();
        if (best_threads) *best_threads = bnt;
        if (best_wgs) *best_wgs = bnwg;
        if (best_ms) *best_ms = best;
        return 0;
    } catch (const std::exception &e) {
        gqz_set_error(e.what());
        return -100;
    }
}

}  // extern "C"
