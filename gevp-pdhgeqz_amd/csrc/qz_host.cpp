// qz_host.cpp -- host side of the MI355X multishift QZ: the PDHGEQZ0/PDHGEQZ1 iteration
// (deflation scan, infinite eigenvalues, aggressive early deflation, multishift sweeps) driving
// the HIP kernels on one stream, and the C ABI declared in include/gqz.h.
//
// Reference structure (NLAFET/GEVP-PDHGEQZ, 1x1 grid):
//   qz_run        ~ PDHGEQZ0.f:505-1202 / PDHGEQZ1.f:466-763 (recursive: the AED window is
//                   solved by the same routine, PDHGEQZ3.f:588)
//   small_block   ~ PDHGEQZ4 + KKQZ for blocks <= NMIN (PDHGEQZ4.f:14-405)
//   aed           ~ PDHGEQZ3/PDHGEQZ2 + QZEARLY (PDHGEQZ3.f:330-997, qzearly.f:1-455)
//   sweep         ~ PDHGEQZ5 with chains of SHIFTPERWND shifts (PDHGEQZ0.f:1068-1187)
//   apply_window  ~ PDHGEQZ6 (PDHGEQZ6.f:276-380)
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gqz.h"
#include "gqz_kernels.h"
#include "gqz_misc.h"

namespace {

const double kUlp = 2.220446049250313080847e-16;
// HT reductions of order above this use the kernels that delay and bulk-apply each sweep's rotations
const int kHtBlockedMin = 96;
const double kSafmin = 2.2250738585072013831e-308;

struct GqzError : std::runtime_error {
    explicit GqzError(const std::string &s) : std::runtime_error(s) {}
};

#define HIPCHK(x)                                                                               \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess)                                                                   \
            throw GqzError(std::string(#x) + " failed: " + hipGetErrorString(e_));              \
    } while (0)

// Device allocation owned by a scope: released when the scope ends, also when a HIP error unwinds
// it through HIPCHK (stream-ordered when it was allocated on a stream).
struct DevBuf {
    double *p = nullptr;
    hipStream_t st = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    double *alloc(size_t bytes) { HIPCHK(hipMalloc((void **)&p, bytes)); return p; }
    double *alloc_async(size_t bytes, hipStream_t s) {
        HIPCHK(hipMallocAsync((void **)&p, bytes, s));
        st = s;
        return p;
    }
    ~DevBuf() {
        if (!p) return;
        if (st) (void)hipFreeAsync(p, st);
        else (void)hipFree(p);
    }
};

struct Arena {
    char *base = nullptr;
    size_t cap = 0, top = 0;
    double *alloc(size_t n) {
        size_t b = (n * sizeof(double) + 255) & ~(size_t)255;
        if (top + b > cap) throw GqzError("workspace exhausted");
        double *p = (double *)(base + top);
        top += b;
        return p;
    }
    size_t mark() const { return top; }
    void release(size_t m) { top = m; }
};

struct Lib {
    int dev = -1;
    hipStream_t st = nullptr;
    // second stream (default priority) for the slab updates of a sweep, overlapped with the chase
    hipStream_t st2 = nullptr;
    bool overlap = true;
    int overlap_min = 2048;     // smallest active block whose sweeps overlap
    std::vector<hipEvent_t> sev;     // ordering events between the two streams (no timing)
    Arena ws;
    int *d_out = nullptr, *h_out = nullptr;
    double *h_buf = nullptr;   // pinned staging (readbacks and shift uploads)
    size_t h_buf_n = 0;
    double *d_norm = nullptr;
    gqz_policy_fn pol = nullptr;
    void *pol_user = nullptr;
    // learned policy: parameters, one feature row and one output row on the device
    float *d_mlp = nullptr, *d_mlp_x = nullptr, *d_mlp_y = nullptr;
    float *h_mlp_io = nullptr;       // pinned: 12 features + 3 outputs
    bool mlp_on = false;
    std::vector<gqz_trace_rec> trace;
    int trace_cap = 0;
    double progress_s = 0.0;         // > 0: top-level progress lines on stderr at this interval
    std::string err;
    std::mutex mu;
};
// One context per host thread: its own stream, workspace arena, pinned staging and settings, so
// calls from different host threads run concurrently (each thread's first call initialises its
// context on device 0 unless it called gqz_init; gqz_finalize releases the calling thread's).
thread_local Lib L;

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int piparmq_ns(int nh) {            // piparmq.f:445-476
    int ns = 2;
    if (nh >= 30) ns = 4;
    if (nh >= 60) ns = 10;
    if (nh >= 150) {
        int lg = (int)std::lround(std::log((double)nh) / std::log(2.0));
        ns = std::max(10, nh / lg);
    }
    if (nh >= 590) ns = 64;
    if (nh >= 3000) ns = 128;
    if (nh >= 6000) ns = 256;
    if (nh >= 12000) ns = 512;
    if (nh >= 24000) ns = 1024;
    if (nh >= 48000) ns = 2048;
    return std::max(2, ns - ns % 2);
}
int piparmq_nw(int nh) { int ns = piparmq_ns(nh); return nh <= 500 ? ns : 3 * ns / 2; }
int piparmq_nibble(int nh) {
    if (nh < 6000) return 14;
    return std::min(100, (int)std::ceil(335.0 * std::pow((double)nh, -0.44)));
}

struct View {
    double *H, *T, *Q, *Z;
    int ld, ldq, n, nq;
    double *ar, *ai, *be;
    double smlnum;
};

struct Run {
    double ttol;
    gqz_params prm;
    gqz_stats st;
    int n_top;
    // deflation history for the parameter policy (top level)
    std::vector<double> hist_nd, hist_frac;
    int last_ns = 0, last_nw = 0, last_nibble = 0, last_swept = 0, last_ndinf = 0;
    double t_start = 0.0, t_progress = 0.0;
};

// ---- per-class kernel timing: HIP events on the library stream around each launch --------------
struct KTimer {
    unsigned mask = 0;
    struct Pend { int cls; hipEvent_t e0, e1; double flops, bytes; hipStream_t st; };
    std::vector<Pend> pend;
    std::vector<hipEvent_t> pool;
    // classes 0..GQZ_KC_COUNT-1: top level; +GQZ_KC_COUNT: inside AED windows (nested levels)
    long launches[2 * GQZ_KC_COUNT] = {};
    double ms[2 * GQZ_KC_COUNT] = {}, flops[2 * GQZ_KC_COUNT] = {}, bytes[2 * GQZ_KC_COUNT] = {};
    // one launch in `every` is timed, chosen by a linear congruential generator (a fixed stride
    // could lock onto the period of the launch sequence, e.g. left / right slab updates); events
    // add queue packets and host calls.  `timed` counts the timed launches, ms sums their times
    int every = 1;
    unsigned long long rng = 1;
    long timed[2 * GQZ_KC_COUNT] = {};
    int level = 0;
};
thread_local KTimer KT;           // per host thread, like the stream it times

struct KtLevel {            // scoped recursion level for the timing classes
    int saved;
    explicit KtLevel(int l) : saved(KT.level) { KT.level = l; }
    ~KtLevel() { KT.level = saved; }
};

hipEvent_t kt_event() {
    if (!KT.pool.empty()) { hipEvent_t e = KT.pool.back(); KT.pool.pop_back(); return e; }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
}

// returns a slot (or -1 when the class is not timed); the matching kt_end records the stop event.
// Slots <= -2 stand for an untimed launch of class -slot - 2: only its work is counted.
int kt_begin(int cls, hipStream_t st = nullptr) {
    if (!(KT.mask & (1u << cls))) return -1;
    const int c = cls + (KT.level > 0 ? GQZ_KC_COUNT : 0);
    if (KT.every > 1) {
        KT.rng = KT.rng * 6364136223846793005ULL + 1442695040888963407ULL;
        if ((KT.rng >> 33) % (unsigned long long)KT.every != 0) return -2 - c;
    }
    KTimer::Pend p = { c, kt_event(), kt_event(), 0.0, 0.0, st ? st : L.st };
    HIPCHK(hipEventRecord(p.e0, p.st));
    KT.pend.push_back(p);
    return (int)KT.pend.size() - 1;
}

void kt_end(int slot, double flops, double bytes) {
    if (slot == -1) return;
    if (slot < -1) {
        const int c = -slot - 2;
        KT.launches[c]++; KT.flops[c] += flops; KT.bytes[c] += bytes;
        return;
    }
    KTimer::Pend &p = KT.pend[slot];
    p.flops = flops; p.bytes = bytes;
    HIPCHK(hipEventRecord(p.e1, p.st));
}

// fold completed launches into the totals (the stream has been synchronised)
void kt_drain() {
    for (auto &p : KT.pend) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, p.e0, p.e1));
        KT.launches[p.cls]++; KT.ms[p.cls] += t; KT.flops[p.cls] += p.flops; KT.bytes[p.cls] += p.bytes;
        KT.timed[p.cls]++;
        KT.pool.push_back(p.e0); KT.pool.push_back(p.e1);
    }
    KT.pend.clear();
}

void sync_read(void *dst, const void *src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, L.st));
    HIPCHK(hipStreamSynchronize(L.st));
    if (!KT.pend.empty()) kt_drain();
}

// algorithmic work of one multi-slab launch: 2 m k^2 flops per slab; each slab read and written
// once (16 m k bytes) plus its k x k factor (8 k^2 bytes)
void slab_work(const GqzSlabJobs &J, double &flops, double &bytes) {
    flops = 0.0; bytes = 0.0;
    for (int q = 0; q < J.njobs; ++q) {
        const double k = J.k[q];
        flops += 2.0 * J.m[q] * k * k;
        bytes += 16.0 * J.m[q] * k + 8.0 * k * k;
    }
}

// ---- PDHGEQZ6 role: apply the windows' U, V outside the windows ----------------------------
// Windows [p_i, p_i + w_i) of one round are disjoint.  All left updates (rows of a window,
// columns to its right) are launched before all right updates (columns of a window, rows above,
// Q, Z): the only blocks both touch are (rows of a lower window) x (columns of an upper one),
// where the two products commute.
struct Win { int p, w; const double *U, *V; };

void slab_launch(Run &R, GqzSlabJobs &J, bool left, hipStream_t st = nullptr) {
    if (J.njobs == 0) return;
    if (!st) st = L.st;
    double fl, by;
    slab_work(J, fl, by);
    int ks = kt_begin(GQZ_KC_SLAB, st);
    if (left) gqz_slab_left_multi(st, J);
    else gqz_slab_right_multi(st, J);
    kt_end(ks, fl, by);
    R.st.slab_launches++;
    J.njobs = 0;
}

void apply_windows(Run &R, const View &v, const Win *ws, int nw) {
    GqzSlabJobs J;
    J.njobs = 0;
    for (int i = 0; i < nw; ++i) {
        const Win &x = ws[i];
        const int ncol = v.n - x.p - x.w;
        if (J.njobs + 2 > GQZ_SLAB_MAXJOBS) slab_launch(R, J, true);
        gqz_slab_add(J, v.H + (size_t)x.p + (size_t)(x.p + x.w) * v.ld, v.ld, ncol, x.U, x.w, x.w);
        gqz_slab_add(J, v.T + (size_t)x.p + (size_t)(x.p + x.w) * v.ld, v.ld, ncol, x.U, x.w, x.w);
    }
    slab_launch(R, J, true);
    for (int i = 0; i < nw; ++i) {
        const Win &x = ws[i];
        if (J.njobs + 4 > GQZ_SLAB_MAXJOBS) slab_launch(R, J, false);
        gqz_slab_add(J, v.H + (size_t)x.p * v.ld, v.ld, x.p, x.V, x.w, x.w);
        gqz_slab_add(J, v.T + (size_t)x.p * v.ld, v.ld, x.p, x.V, x.w, x.w);
        if (v.Q) gqz_slab_add(J, v.Q + (size_t)x.p * v.ldq, v.ldq, v.nq, x.U, x.w, x.w);
        if (v.Z) gqz_slab_add(J, v.Z + (size_t)x.p * v.ldq, v.ldq, v.nq, x.V, x.w, x.w);
    }
    slab_launch(R, J, false);
}

void apply_window(Run &R, const View &v, int p, int w, const double *U, const double *V) {
    Win x = { p, w, U, V };
    apply_windows(R, v, &x, 1);
}

// ---- small active block: LDS-resident base QZ, then the off-block updates ------------------
int small_block(Run &R, const View &v, int lo, int hi) {
    double t0 = now();
    int m = hi - lo + 1;
    size_t mk = L.ws.mark();
    double *U = L.ws.alloc((size_t)m * m), *V = L.ws.alloc((size_t)m * m);
    GqzSmallJob J = {};
    J.H = v.H + (size_t)lo + (size_t)lo * v.ld; J.T = v.T + (size_t)lo + (size_t)lo * v.ld;
    J.ld = v.ld; J.m = m; J.U = U; J.V = V; J.ldu = m;
    J.ar = v.ar + lo; J.ai = v.ai + lo; J.be = v.be + lo;
    J.info = L.d_out + 8; J.ttol = R.ttol; J.smlnum = v.smlnum;
    int kk = kt_begin(GQZ_KC_SMALL);
    gqz_launch_small1(L.st, J);
    kt_end(kk, 0.0, 64.0 * m * m);
    apply_window(R, v, lo, m, U, V);
    sync_read(L.h_out + 8, L.d_out + 8, sizeof(int));
    int info = L.h_out[8];
    L.ws.release(mk);
    R.st.small_calls++;
    R.st.t_small += now() - t0;
    return info ? lo + info : 0;
}

int qz_run(Run &R, const View &v, int ilo, int ihi, int level);

// ---- aggressive early deflation on the window [kbot-jw+1, kbot] ---------------------------
int aed(Run &R, const View &v, int ktop, int kbot, int jw, int level, std::vector<double> &sh, int &ls) {
    double t0 = now();
    int kwtop = kbot - jw + 1;
    size_t mk = L.ws.mark();
    size_t jj = (size_t)jw * jw;
    double *Hw = L.ws.alloc(jj), *Tw = L.ws.alloc(jj), *U = L.ws.alloc(jj), *V = L.ws.alloc(jj);
    double *weig = L.ws.alloc(3 * (size_t)jw), *scr = L.ws.alloc(4 * (size_t)jw);
    int nmin = R.prm.nmin > 0 ? std::min(R.prm.nmin, 64) : 64;
    if (jw <= std::min(nmin, gqz_aed_small_max())) {
        // the whole AED in one LDS-resident kernel (window QZ, reorder, spike, restore, write-back)
        double *pscr = L.ws.alloc((size_t)(jw / 2 + 1) * 65);
        int kk = kt_begin(GQZ_KC_SMALL);
        gqz_launch_aed_small(L.st, v.H, v.T, v.ld, kwtop, jw, kwtop > ktop ? 1 : 0, R.ttol,
                             kSafmin * ((double)jw / kUlp), U, V, weig, weig + jw, weig + 2 * jw, pscr, scr,
                             L.d_out + 16);
        kt_end(kk, 0.0, 48.0 * jw * jw);
        HIPCHK(hipMemcpyAsync(L.h_out + 16, L.d_out + 16, 5 * sizeof(int), hipMemcpyDeviceToHost, L.st));
        HIPCHK(hipMemcpyAsync(L.h_buf, weig, 3 * sizeof(double) * jw, hipMemcpyDeviceToHost, L.st));
        HIPCHK(hipStreamSynchronize(L.st));
        if (!KT.pend.empty()) kt_drain();
        R.st.aed_calls++;
        R.st.small_calls++;
        if (L.h_out[20] != 0) {     // window QZ failed: soft recovery, no deflation (PDHGEQZ0.f:944-949)
            L.ws.release(mk);
            ls = 0;
            if (level == 0) R.st.t_aed += now() - t0;
            return 0;
        }
        int kbc = L.h_out[16];
        R.st.swaps += L.h_out[17];
        R.st.swap_fail += L.h_out[18] ? 1 : 0;
        const int spike = L.h_out[19];
        if (L.h_out[18] == 1) {     // window not quasi-triangular: nothing written back, no shifts
            L.ws.release(mk);
            ls = 0;
            if (level == 0) R.st.t_aed += now() - t0;
            return 0;
        }
        if (kbc < jw || !spike) apply_window(R, v, kwtop, jw, U, V);
        sh.assign(3 * (size_t)kbc, 0.0);
        for (int i = 0; i < kbc; ++i) {
            sh[i] = L.h_buf[i];
            sh[kbc + i] = L.h_buf[jw + i];
            sh[2 * kbc + i] = L.h_buf[2 * jw + i];
        }
        ls = kbc;
        L.ws.release(mk);
        if (level == 0) R.st.t_aed += now() - t0;
        return jw - kbc;
    }
    gqz_launch_window_in(L.st, v.H, v.T, v.ld, kwtop, jw, Hw, Tw, U, V);
    View W;
    W.H = Hw; W.T = Tw; W.Q = U; W.Z = V; W.ld = jw; W.ldq = jw; W.n = jw; W.nq = jw;
    W.ar = weig; W.ai = weig + jw; W.be = weig + 2 * jw;
    W.smlnum = kSafmin * ((double)jw / kUlp);
    int info;
    if (jw <= nmin) {
        GqzSmallJob J = {};
        J.H = Hw; J.T = Tw; J.ld = jw; J.m = jw; J.U = U; J.V = V; J.ldu = jw;
        J.ar = W.ar; J.ai = W.ai; J.be = W.be; J.info = L.d_out + 9; J.ttol = R.ttol; J.smlnum = W.smlnum;
        int kk = kt_begin(GQZ_KC_SMALL);
        gqz_launch_small1(L.st, J);
        kt_end(kk, 0.0, 48.0 * jw * jw);
        sync_read(L.h_out + 9, L.d_out + 9, sizeof(int));
        info = L.h_out[9];
    } else {
        info = qz_run(R, W, 0, jw - 1, level + 1);
    }
    if (info != 0) {      // soft recovery: no deflation (PDHGEQZ0.f:944-949)
        L.ws.release(mk);
        ls = 0;
        return 0;
    }
    const double *sptr = (kwtop > ktop) ? v.H + (size_t)kwtop + (size_t)(kwtop - 1) * v.ld : nullptr;
    double *pscr = L.ws.alloc((size_t)(jw / 2 + 1) * 65);
    int kr = kt_begin(GQZ_KC_REORDER);
    gqz_launch_aed_reorder(L.st, Hw, Tw, U, V, jw, sptr, W.smlnum, pscr, L.d_out + 16);
    kt_end(kr, 0.0, 32.0 * jj);
    gqz_launch_aed_eigs(L.st, Hw, Tw, jw, W.ar, W.ai, W.be);
    HIPCHK(hipMemcpyAsync(L.h_out + 16, L.d_out + 16, 4 * sizeof(int), hipMemcpyDeviceToHost, L.st));
    HIPCHK(hipMemcpyAsync(L.h_buf, weig, 3 * sizeof(double) * jw, hipMemcpyDeviceToHost, L.st));
    HIPCHK(hipStreamSynchronize(L.st));
    if (!KT.pend.empty()) kt_drain();
    int kbc = L.h_out[16];
    R.st.swaps += L.h_out[17];
    R.st.swap_fail += L.h_out[18] ? 1 : 0;
    int spike = L.h_out[19];
    if (L.h_out[18] == 1) {
        // a window that is not quasi-triangular has no block structure to reorder: the AED is
        // discarded (the window copy is not written back); no deflation and no shifts, so the
        // sweep falls back to the eigenvalues of the bottom 2x2 block.  After a rejected swap
        // (status 2) the window is a consistent orthogonal equivalence and the deflations found
        // before it are kept (qzearly.f:318-326)
        L.ws.release(mk);
        ls = 0;
        R.st.aed_calls++;
        if (level == 0) R.st.t_aed += now() - t0;
        return 0;
    }
    // nothing deflated: the window is left as it was (PDHGEQZ3.f:840, :1555); its eigenvalues are
    // still returned as shifts
    const bool restore = (kbc < jw) || !spike;
    if (restore) {
        if (kbc > 0 && spike) {
            int kp = kt_begin(GQZ_KC_SPIKE);
            gqz_launch_aed_spike(L.st, Hw, Tw, U, V, jw, sptr, scr, L.d_out + 16);
            kt_end(kp, 0.0, 32.0 * jj);
        }
        if (kbc > 2 && spike) {
            int kg = kt_begin(GQZ_KC_HT);
            if (kbc > kHtBlockedMin) {
                double *rws = L.ws.alloc(gqz_gghrd_blk_ws_bytes(jw, jw) / sizeof(double) + 1);
                gqz_launch_gghrd_blk(L.st, Hw, Tw, jw, jw, 0, kbc - 1, U, V, jw, jw, rws);
            } else {
                gqz_launch_gghrd_wg(L.st, Hw, Tw, jw, jw, 0, kbc - 1, U, V, jw, jw);
            }
            kt_end(kg, 0.0, 32.0 * jj);
        }
        gqz_launch_window_out(L.st, v.H, v.T, v.ld, kwtop, jw, Hw, Tw, U, kwtop > ktop ? (kbc > 0 ? 1 : 2) : 0);
        apply_window(R, v, kwtop, jw, U, V);
    }
    sh.assign(3 * (size_t)kbc, 0.0);
    for (int i = 0; i < kbc; ++i) {
        sh[i] = L.h_buf[i];
        sh[kbc + i] = L.h_buf[jw + i];
        sh[2 * kbc + i] = L.h_buf[2 * jw + i];
    }
    ls = kbc;
    L.ws.release(mk);
    R.st.aed_calls++;
    if (level == 0) R.st.t_aed += now() - t0;
    return jw - kbc;
}

// ---- the chains of one sweep, chased concurrently (PDHGEQZ5 with several windows) -----------
// Chain c (shift pairs soff..) follows chain c-1 down the diagonal.  In every round each chain
// that can move chases its bulges through one window (one workgroup per chain, one launch),
// then all windows' U, V are applied.  A chain's window stays above the window (or, when that
// chain waits, above the bulges) of the chain ahead of it, so the windows of a round are
// disjoint.  Each chain sees exactly the window sequence of a lone chain (intro window at ktop,
// then windows starting where the previous one stopped), so the result does not depend on how
// many chains share a round.
//
// Overlap (L.overlap): the updates of round r run on the second stream while later rounds are
// chased.  The columns that a chain's next window reads ("near" columns [p + w, p' + w'),
// p' = p + steps) get their left update from the chain's next launch itself (k_chase applies the
// previous U to them before it loads its window); the rest of the left update ("far") and the
// right update (rows above, Q, Z) go to the second stream:
//     stream 1: chase(r) -> wait B(r-1) -> event A(r) -> chase(r+1) [near(previous) + windows] -> ...
//     stream 2: wait A(r) -> left-far(r) -> right(r) -> event B(r)
// B(r) and chase(r+1) are released by the same event, so the chase is dispatched first (it follows the
// event in its own queue) and gets its compute units before the update kernels fill the chip (its
// workgroup needs a whole unit's LDS).  chase(r+1) runs concurrently with B(r), so it must touch
// nothing that B(r) touches:
//   * windows: the scheduler keeps every window of round r+1 above the round-r window start of each
//     chain ahead (right(r) writes rows above p of the columns [p, p + w));
//   * a chain's pending near block N (rows [p, p + w), columns [p + w, p' + w')): N ends above the
//     window start of every chain ahead, now and later (checked when the chain is launched: chains
//     only move down), so no right update ever reaches N while it is pending; a right update that
//     did would mix columns with and without the chain's U;
//   * chains behind a chain with a pending N stay above its row p until it moves again (its prevp
//     persists), so their left updates, which do not commute with its U, never reach N's rows.
// chase(r+1) waits for B(r-1): the near block of a chain overlaps the far block of its previous
// window in rows.  U, V live in two slots per chain, alternating with the chain's launches: B(r)
// reads round r's while the chain's next launch writes the other.
struct Chain { int soff, nb, gtop; bool started, done; int prevp; int nl, pw, pnc; };

hipEvent_t sync_event(size_t i) {
    while (L.sev.size() <= i) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        L.sev.push_back(e);
    }
    return L.sev[i];
}

// blocks below L.overlap_min (the nested QZ of AED windows, the end of a top-level QZ) keep the updates
// on the chase stream: their slab updates are short and the stream ordering would cost more than it hides

// One round of the chain scheduler (host only, no device state).  A chain launched in a round gets a
// record: its window [p, p + w), the chase steps, whether it is the chain's last window, nc (the near
// columns its U leaves pending for its next launch, overlap only) and near/pp/pw (the pending block
// of its previous window that this launch applies first).
struct SchedWin { int chain, p, w, steps, last, intro, nc, near, pp, pw; };

int sched_select(const std::vector<Chain> &ch, int ktop, int kbot, int n, bool ovl, int wmax, int nslot,
                 SchedWin *out) {
    const int nch = (int)ch.size();
    int nw = 0;
    // rows from which the chains ahead forbid access; nlim: columns from which the windows of the
    // chains ahead start, this round and later (a pending near block must end above it)
    int limit = kbot + 1, nlim = kbot + 1;
    int active = 0;
    for (int c = 0; c < nch && nw < nslot; ++c) {
        const Chain &x = ch[c];
        const int pp = x.prevp;
        if (x.done) {
            if (ovl) limit = std::min(limit, pp);
            continue;
        }
        if (!x.started && active >= nslot) break;
        const int p = x.started ? x.gtop : ktop;
        const int w = std::min(wmax, kbot - p + 1);
        const int last = (p + w - 1 == kbot);
        const int steps = last ? w : (w - 3 * x.nb - 1);
        int nc = 0;
        if (!last) {
            const int np = p + steps, nw2 = std::min(wmax, kbot - np + 1);
            nc = std::max(0, std::min(np + nw2, n) - (p + w));
        }
        if (p + w > limit || (ovl && nc > 0 && p + w + nc > nlim)) {
            // wait: the chain ahead still occupies rows below; nothing behind may pass
            limit = std::min(limit, x.started ? x.gtop : ktop);
            if (ovl) limit = std::min(limit, pp);
            nlim = std::min(nlim, p);
            if (!x.started) break;
            ++active;
            continue;
        }
        SchedWin &o = out[nw++];
        o.chain = c; o.p = p; o.w = w; o.steps = steps; o.last = last; o.intro = x.started ? 0 : 1;
        o.nc = ovl ? nc : 0;
        o.near = ovl ? x.pnc : 0; o.pp = x.prevp; o.pw = x.pw;
        limit = p;
        if (ovl) limit = std::min(limit, pp);
        nlim = std::min(nlim, p);
        ++active;
    }
    return nw;
}

// after a launch: a finished chain's window start constrains the next round only; the launched
// chains move on
void sched_commit(std::vector<Chain> &ch, const SchedWin *sw, int nw) {
    for (auto &x : ch) if (x.done) x.prevp = INT32_MAX;
    for (int q = 0; q < nw; ++q) {
        Chain &x = ch[sw[q].chain];
        x.started = true;
        x.gtop = sw[q].p + sw[q].steps;
        x.prevp = sw[q].p;
        x.pw = sw[q].w;
        x.pnc = sw[q].nc;
        ++x.nl;
        if (sw[q].last) x.done = true;
    }
}

// no window could move: with overlap the only possible obstacle is the last right update of a finished
// chain; once stream 2 has finished, those constraints go.  Returns false when there is none (a stall).
bool sched_unblock(std::vector<Chain> &ch) {
    bool pend = false;
    for (auto &x : ch) pend = pend || (x.done && x.prevp != INT32_MAX);
    for (auto &x : ch) if (x.done) x.prevp = INT32_MAX;
    return pend;
}

void sched_reset(std::vector<Chain> &ch) {
    for (auto &x : ch) { x.prevp = INT32_MAX; x.nl = 0; x.pw = 0; x.pnc = 0; }
}

void chase_chains(Run &R, const View &v, int ktop, int kbot, std::vector<Chain> &ch, const double *d_sr,
                  const double *d_si, const double *d_sb) {
    const int WMAX = gqz_chase_max_window();
    const int nslot = GQZ_MAXCH;
    const bool ovl = L.overlap && L.st2 && kbot - ktop + 1 >= L.overlap_min;
    size_t mk = L.ws.mark();
    const size_t slot = (size_t)WMAX * WMAX;
    const int nch = (int)ch.size();
    // without overlap: slots by launch position and round parity; with overlap: two per chain
    const size_t nslots = ovl ? 2 * (size_t)nch : 2 * (size_t)nslot;
    double *U = L.ws.alloc(nslots * slot), *V = L.ws.alloc(nslots * slot);
    sched_reset(ch);
    int first = 0;                       // oldest chain not yet done
    long round = 0;
    bool b_pend[2] = { false, false };   // stream 2 holds updates recorded in B[parity]
    auto wait_b = [&](int par) {
        if (b_pend[par]) HIPCHK(hipStreamWaitEvent(L.st, sync_event(2 + par), 0));
    };
    // the previous round's windows, whose far-left and right updates are not launched yet
    Win pws[GQZ_MAXCH];
    int pnear[GQZ_MAXCH], pnw = 0, ppar = 0;
    auto launch_b = [&]() {              // B(previous round), released by A on stream 1
        if (pnw == 0) return;
        HIPCHK(hipEventRecord(sync_event(0), L.st));
        HIPCHK(hipStreamWaitEvent(L.st2, sync_event(0), 0));
        GqzSlabJobs J;
        J.njobs = 0;
        for (int q = 0; q < pnw; ++q) {
            const Win &x = pws[q];
            const int c0 = x.p + x.w + pnear[q], ncol = v.n - c0;
            if (J.njobs + 2 > GQZ_SLAB_MAXJOBS) slab_launch(R, J, true, L.st2);
            gqz_slab_add(J, v.H + (size_t)x.p + (size_t)c0 * v.ld, v.ld, ncol, x.U, x.w, x.w);
            gqz_slab_add(J, v.T + (size_t)x.p + (size_t)c0 * v.ld, v.ld, ncol, x.U, x.w, x.w);
        }
        slab_launch(R, J, true, L.st2);
        for (int q = 0; q < pnw; ++q) {
            const Win &x = pws[q];
            if (J.njobs + 4 > GQZ_SLAB_MAXJOBS) slab_launch(R, J, false, L.st2);
            gqz_slab_add(J, v.H + (size_t)x.p * v.ld, v.ld, x.p, x.V, x.w, x.w);
            gqz_slab_add(J, v.T + (size_t)x.p * v.ld, v.ld, x.p, x.V, x.w, x.w);
            if (v.Q) gqz_slab_add(J, v.Q + (size_t)x.p * v.ldq, v.ldq, v.nq, x.U, x.w, x.w);
            if (v.Z) gqz_slab_add(J, v.Z + (size_t)x.p * v.ldq, v.ldq, v.nq, x.V, x.w, x.w);
        }
        slab_launch(R, J, false, L.st2);
        HIPCHK(hipEventRecord(sync_event(2 + ppar), L.st2));
        b_pend[ppar] = true;
        pnw = 0;
    };
    while (first < nch) {
        const int par = (int)(round & 1);
        SchedWin sw[GQZ_MAXCH];
        const int nw = sched_select(ch, ktop, kbot, v.n, ovl, WMAX, nslot, sw);
        if (nw == 0) {
            if (!ovl || !sched_unblock(ch)) throw GqzError("chain scheduler made no progress");
            // let stream 2 finish (the finished chain's right update was the obstacle)
            launch_b();
            wait_b(0); wait_b(1);
            b_pend[0] = b_pend[1] = false;
            continue;
        }
        GqzChainSet S = {};
        S.H = v.H; S.T = v.T; S.ld = v.ld; S.U = U; S.V = V; S.ustride = WMAX * WMAX;
        S.sr = d_sr; S.si = d_si; S.sb = d_sb; S.smlnum = v.smlnum;
        S.nch = nw;
        Win ws[GQZ_MAXCH];
        for (int q = 0; q < nw; ++q) {
            const SchedWin &o = sw[q];
            const Chain &x = ch[o.chain];
            S.p[q] = o.p; S.w[q] = o.w; S.nb[q] = x.nb; S.soff[q] = x.soff;
            S.intro[q] = o.intro; S.t0[q] = o.intro ? -1 : 0; S.last[q] = o.last;
            S.steps[q] = o.steps;
            S.uslot[q] = ovl ? 2 * o.chain + (x.nl & 1) : par * nslot + q;
            if (o.near > 0) {
                S.near[q] = o.near; S.pslot[q] = 2 * o.chain + ((x.nl - 1) & 1); S.pp[q] = o.pp; S.pw[q] = o.pw;
            }
            ws[q] = Win{ o.p, o.w, U + (size_t)S.uslot[q] * slot, V + (size_t)S.uslot[q] * slot };
        }
        if (ovl) {
            wait_b(par);                 // B(r-2)
            launch_b();                  // B(r-1), released together with this chase
        }
        double by = 0.0;
        for (int q = 0; q < nw; ++q) by += 48.0 * S.w[q] * S.w[q] + 32.0 * S.pw[q] * S.near[q];
        int kc = kt_begin(GQZ_KC_CHASE);
        gqz_chase_set(L.st, S);
        kt_end(kc, 0.0, by);
        R.st.chase_windows += nw;
        if (!ovl) {
            apply_windows(R, v, ws, nw);
        } else {
            for (int q = 0; q < nw; ++q) { pws[q] = ws[q]; pnear[q] = sw[q].nc; }
            pnw = nw;
            ppar = par;
        }
        sched_commit(ch, sw, nw);
        while (first < nch && ch[first].done) ++first;
        ++round;
    }
    // everything after the sweep sees the updates of every round
    launch_b();
    wait_b(0); wait_b(1);
    L.ws.release(mk);
}

// shift preparation (PDHGEQZ0.f:1068-1135) and the chains of the sweep
void sweep(Run &R, const View &v, int ktop, int kbot, std::vector<double> &sh, int avail, int ns_cap,
           int level) {
    double t0 = now();
    std::vector<double> sr(sh.begin(), sh.begin() + avail), si(sh.begin() + avail, sh.begin() + 2 * avail),
        sb(sh.begin() + 2 * avail, sh.begin() + 3 * avail);
    int ns = std::min(ns_cap, std::max(2, kbot - ktop - 1));
    ns -= ns % 2;
    if (avail < 2) {
        // too few shifts: eigenvalues of the bottom 2x2 (exceptional fall-back)
        double hb[4], tb[4];
        for (int j = 0; j < 2; ++j) {
            HIPCHK(hipMemcpyAsync(hb + 2 * j, v.H + (size_t)(kbot - 1) + (size_t)(kbot - 1 + j) * v.ld, 2 * sizeof(double), hipMemcpyDeviceToHost, L.st));
            HIPCHK(hipMemcpyAsync(tb + 2 * j, v.T + (size_t)(kbot - 1) + (size_t)(kbot - 1 + j) * v.ld, 2 * sizeof(double), hipMemcpyDeviceToHost, L.st));
        }
        HIPCHK(hipStreamSynchronize(L.st));
        double m10 = hb[1] / tb[3], m11 = hb[3] / tb[3];
        double m00 = (hb[0] - tb[2] * m10) / tb[0], m01 = (hb[2] - tb[2] * m11) / tb[0];
        double tr = 0.5 * (m00 + m11), det = m00 * m11 - m01 * m10, disc = tr * tr - det;
        sr.assign(2, tr); si.assign(2, 0.0); sb.assign(2, 1.0);
        if (disc >= 0) { sr[0] = tr + std::sqrt(disc); sr[1] = tr - std::sqrt(disc); }
        else { si[0] = std::sqrt(-disc); si[1] = -si[0]; }
        avail = 2;
    }
    if (avail > ns) {        // bubble sort by decreasing |lambda|
        for (int k = avail - 1; k > 0; --k) {
            bool sorted = true;
            for (int i = 0; i < k; ++i) {
                double d1 = std::fabs((std::fabs(sr[i]) + std::fabs(si[i])) / sb[i]);
                double d2 = std::fabs((std::fabs(sr[i + 1]) + std::fabs(si[i + 1])) / sb[i + 1]);
                if (d1 < d2) { std::swap(sr[i], sr[i + 1]); std::swap(si[i], si[i + 1]); std::swap(sb[i], sb[i + 1]); sorted = false; }
            }
            if (sorted) break;
        }
    }
    for (int i = avail - 1; i >= 2; i -= 2) {   // keep conjugate pairs together
        if (si[i] == 0.0 && si[i - 1] != 0.0) {
            double t;
            t = sr[i]; sr[i] = sr[i - 1]; sr[i - 1] = sr[i - 2]; sr[i - 2] = t;
            t = si[i]; si[i] = si[i - 1]; si[i - 1] = si[i - 2]; si[i - 2] = t;
            t = sb[i]; sb[i] = sb[i - 1]; sb[i - 1] = sb[i - 2]; sb[i - 2] = t;
        }
    }
    ns = std::min(ns, avail);
    ns -= ns % 2;
    if (ns < 2) return;
    int ks = avail - ns;
    // non-finite shifts are replaced by a harmless real pair (the bottom-most finite one)
    for (int i = ks; i < avail; ++i)
        if (!std::isfinite(sr[i] / sb[i]) || !std::isfinite(si[i] / sb[i])) { sr[i] = 1.0; si[i] = 0.0; sb[i] = 1.0; }
    for (int i = ks; i + 1 < avail; i += 2)
        if ((si[i] == 0.0) != (si[i + 1] == 0.0)) { si[i] = si[i + 1] = 0.0; }
    // upload [sr | si | sb] of the ns shifts used
    size_t mk = L.ws.mark();
    double *d_sh = L.ws.alloc(3 * (size_t)ns);
    for (int i = 0; i < ns; ++i) {
        L.h_buf[i] = sr[ks + i]; L.h_buf[ns + i] = si[ks + i]; L.h_buf[2 * ns + i] = sb[ks + i];
    }
    HIPCHK(hipMemcpyAsync(d_sh, L.h_buf, 3 * sizeof(double) * ns, hipMemcpyHostToDevice, L.st));
    // default chain length: about six chains in flight, 10 to 20 shifts each (measured: with 60 shifts a
    // sweep, chains of 10 take 18.3 s at n = 16000 against 20.5 s for 20; with 128 shifts, 20 is 4 %
    // faster than 10)
    int chain = R.prm.chain > 0 ? R.prm.chain : std::max(10, std::min(20, ns / 6));
    chain = std::max(2, std::min(chain, 40));
    chain -= chain % 2;
    std::vector<Chain> chains;
    for (int off = 0; off < ns; off += chain) {
        int cnt = std::min(chain, ns - off);
        cnt -= cnt % 2;
        if (cnt < 2) break;
        chains.push_back(Chain{ off, cnt / 2, ktop, false, false, INT32_MAX });
    }
    chase_chains(R, v, ktop, kbot, chains, d_sh, d_sh + ns, d_sh + 2 * ns);
    // the staging buffer is reused by the next readback: make sure the upload has completed
    HIPCHK(hipStreamSynchronize(L.st));
    L.ws.release(mk);
    R.st.sweeps++;
    R.st.shifts_used += ns;
    if (level == 0) R.st.t_sweep += now() - t0;
}

// one inference of the learned policy on the library stream
void mlp_infer(const float feat[12], float out[3]) {
    std::memcpy(L.h_mlp_io, feat, 12 * sizeof(float));
    HIPCHK(hipMemcpyAsync(L.d_mlp_x, L.h_mlp_io, 12 * sizeof(float), hipMemcpyHostToDevice, L.st));
    gqz_launch_mlp_fwd(L.st, L.d_mlp, L.d_mlp_x, 1, L.d_mlp_y);
    HIPCHK(hipMemcpyAsync(L.h_mlp_io + 12, L.d_mlp_y, 3 * sizeof(float), hipMemcpyDeviceToHost, L.st));
    HIPCHK(hipStreamSynchronize(L.st));
    std::memcpy(out, L.h_mlp_io + 12, 3 * sizeof(float));
}

// deflation-history features for the parameter policy (12 values, all O(1))
void features(Run &R, int nh, float f[12]) {
    int n = std::max(1, R.n_top);
    auto last = [](const std::vector<double> &h, int k) {
        return (int)h.size() > k ? h[h.size() - 1 - k] : 0.0;
    };
    double m3 = 0.0;
    int c3 = 0;
    for (int k = 0; k < 3 && k < (int)R.hist_frac.size(); ++k) { m3 += last(R.hist_frac, k); ++c3; }
    if (c3) m3 /= c3;
    f[0] = (float)(std::log2((double)std::max(nh, 2)) / 16.0);
    f[1] = (float)((double)nh / n);
    f[2] = (float)last(R.hist_frac, 0);
    f[3] = (float)last(R.hist_frac, 1);
    f[4] = (float)m3;
    f[5] = (float)((double)R.last_ndinf / std::max(1, nh));
    f[6] = (float)((double)R.last_ns / std::max(1, nh));
    f[7] = (float)((double)R.last_nw / std::max(1, nh));
    f[8] = (float)(R.last_nibble / 100.0);
    f[9] = (float)R.last_swept;
    f[10] = (float)(std::min<double>(R.hist_frac.size(), 200) / 200.0);
    f[11] = (float)(std::log2((double)std::max(n, 2)) / 16.0);
}

// Interior zero T(z, z) of the active block [., kb] chased down to T(kb, kb) (the chase of
// DHGEQZ7 / PDHGEQZ8 in NB windows, updates as PDHGEQZ6): each window [z - 1, z - 1 + w) holds the
// zero in its second row and moves it to its last row, then U, V update the pencil outside it.
void zero_chase(Run &R, const View &v, int z, int kb) {
    const int WMAX = gqz_zchase_max_window();
    size_t mk = L.ws.mark();
    double *U = L.ws.alloc((size_t)WMAX * WMAX), *V = L.ws.alloc((size_t)WMAX * WMAX);
    while (z < kb) {
        const int p = z - 1, w = std::min(WMAX, kb - p + 1);
        int ki = kt_begin(GQZ_KC_INF);
        gqz_launch_zchase_window(L.st, v.H, v.T, v.ld, p, w, U, V);
        kt_end(ki, 0.0, 64.0 * w * w);
        apply_window(R, v, p, w, U, V);
        z = p + w - 1;
    }
    L.ws.release(mk);
}

int qz_run(Run &R, const View &v, int ilo, int ihi, int level) {
    KtLevel klev(level);
    int nmin = R.prm.nmin > 0 ? std::min(R.prm.nmin, 64) : 64;
    int nh0 = ihi - ilo + 1;
    if (nh0 <= 0) return 0;
    if (nh0 <= nmin) return small_block(R, v, ilo, ihi);
    const bool top = (level == 0);
    int ns_def = (top && R.prm.ns > 0) ? R.prm.ns : piparmq_ns(nh0);
    int nw_def = (top && R.prm.nw > 0) ? R.prm.nw : piparmq_nw(nh0);
    int nib_def = (top && R.prm.nibble > 0) ? R.prm.nibble : piparmq_nibble(nh0);
    long maxit = (long)(R.prm.maxit > 0 ? R.prm.maxit : 30) * nh0;
    std::vector<double> sh;
    int kbot = ihi;
    for (long it = 0;; ++it) {
        if (kbot < ilo) break;
        if (it >= maxit) return kbot + 1;
        double ti = now();
        R.st.iters++;
        // deflation scan and infinite eigenvalues; an interior zero of diag(T) is chased to the bottom
        // of the active block in LDS windows (zero_chase) and the scan resumes below the active top
        int ktop = -1, kt = 0, kb = 0, ninf = 0, kt0 = -1, kb_cur = kbot;
        for (;;) {
            int ki = kt_begin(GQZ_KC_INF);
            gqz_launch_inf_defl(L.st, v.H, v.T, v.Q, v.Z, v.ld, v.ldq, v.n, v.nq, ilo, kb_cur, R.ttol, v.smlnum,
                                v.ar, v.ai, v.be, L.d_out, kt0, 0);
            kt_end(ki, 0.0, 32.0 * (kb_cur - ilo + 1));
            sync_read(L.h_out, L.d_out, 5 * sizeof(int));
            if (ktop < 0) ktop = L.h_out[0];
            kt = L.h_out[1]; kb = L.h_out[2]; ninf += L.h_out[3];
            const int zr = L.h_out[4];
            if (zr < 0) break;
            zero_chase(R, v, zr, kb);
            kt0 = kt; kb_cur = kb;
        }
        R.st.t_inf += now() - ti;
        if (ninf > 0) {
            R.st.inf_defl += ninf;
            if (top) R.last_ndinf = ninf;
            if (kt > ktop && kb >= kt) {
                int r = qz_run(R, v, kt, kb, level);
                if (r) return r;
                kbot = ktop - 1;
            } else {
                kbot = (kb < kt) ? ktop - 1 : kb;
            }
            continue;
        }
        int nh = kbot - ktop + 1;
        if (nh <= nmin) {
            int r = small_block(R, v, ktop, kbot);
            if (r) return r;
            kbot = ktop - 1;
            continue;
        }
        int ns_it = ns_def, nw_it = nw_def, nib_it = nib_def;
        float feat[12];
        if (top) {
            features(R, nh, feat);
            if (R.prm.use_policy && L.mlp_on) {
                // learned policy: log2 multipliers of the PIPARMQ values from the MLP kernel
                float o[3];
                mlp_infer(feat, o);
                auto mult = [](float x, float lo, float hi) {
                    return std::exp2((double)std::min(hi, std::max(lo, std::isfinite(x) ? x : 0.0f)));
                };
                int a = (int)std::lround(ns_def * mult(o[0], -2.0f, 1.0f));
                int b = (int)std::lround(nw_def * mult(o[1], -5.0f, 1.0f));
                int c = (int)std::lround(nib_def * mult(o[2], -2.0f, 4.0f));
                ns_it = std::max(2, a - a % 2);
                nw_it = std::max(8, b);
                nib_it = std::min(100, std::max(1, c));
            } else if (R.prm.use_policy && L.pol) {
                int a = ns_def, b = nw_def, c = nib_def;
                L.pol(L.pol_user, feat, nh, &a, &b, &c);
                if (a > 0) ns_it = std::max(2, a - a % 2);
                if (b > 0) nw_it = std::max(4, b);
                if (c > 0) nib_it = c;
            }
        }
        nw_it = std::min(nw_it, 1024);
        int jw = std::min(nw_it, nh);
        int ls = 0;
        int nd = aed(R, v, ktop, kbot, jw, level, sh, ls);
        kbot -= nd;
        int swept = 0;
        if (kbot >= ktop && (nd == 0 || (100 * nd <= jw * nib_it && kbot - ktop + 1 > nmin))) {
            sweep(R, v, ktop, kbot, sh, ls, ns_it, level);
            swept = 1;
        }
        if (top && L.progress_s > 0.0) {
            const double t = now();
            if (t - R.t_progress >= L.progress_s) {
                R.t_progress = t;
                std::fprintf(stderr, "[gqz] n=%d iteration %ld: active block [%d, %d], %d rows left, %ld sweeps, "
                             "%ld AED calls, %.1f s\n", R.n_top, it, ktop, kbot, kbot - ilo + 1, R.st.sweeps,
                             R.st.aed_calls, t - R.t_start);
                std::fflush(stderr);
            }
        }
        if (top) {
            R.hist_frac.push_back((double)nd / std::max(1, jw));
            R.last_ns = ns_it; R.last_nw = nw_it; R.last_nibble = nib_it; R.last_swept = swept;
            if (L.trace_cap > 0 && (int)L.trace.size() < L.trace_cap) {
                gqz_trace_rec rec;
                rec.nh = nh; rec.ktop = ktop; rec.kbot = kbot + nd; rec.ns = ns_it; rec.nw = nw_it;
                rec.nibble = nib_it; rec.nd = nd; rec.ndinf = 0; rec.swept = swept;
                std::memcpy(rec.feat, feat, sizeof feat);
                rec.t_iter = now() - ti;
                L.trace.push_back(rec);
            }
            R.last_ndinf = 0;
        }
    }
    return 0;
}

size_t arena_bytes(int n, const gqz_params *prm) {
    int nw = piparmq_nw(std::max(n, 1));
    if (prm && prm->nw > nw) nw = prm->nw;
    nw = std::max(std::min(std::max(nw, 1024), n), 64);
    size_t d = 6 * (size_t)nw * nw + 80 * (size_t)nw + 4 * (size_t)1024 * 1024;
    return d * sizeof(double);
}

void ensure_ws(size_t bytes) {
    if (L.ws.cap >= bytes) return;
    if (L.ws.base) HIPCHK(hipFree(L.ws.base));
    L.ws.base = nullptr; L.ws.cap = 0; L.ws.top = 0;
    HIPCHK(hipMalloc((void **)&L.ws.base, bytes));
    L.ws.cap = bytes;
}

int check_init() {
    if (L.dev < 0) {
        if (gqz_init(0) != 0) return -1;
    }
    HIPCHK(hipSetDevice(L.dev));
    return 0;
}

int decode(char c, const char *opts) {
    char u = (char)toupper((unsigned char)c);
    for (int i = 0; opts[i]; ++i) if (opts[i] == u) return i;
    return -1;
}

}  // namespace

int gqz_init_current() { return check_init(); }
int gqz_ready_device() {
    if (L.dev < 0) return -1;
    return hipSetDevice(L.dev) == hipSuccess ? L.dev : -1;
}
void gqz_set_error(const char *msg) { L.err = msg; }


// ================================================================================================
// C ABI
// ================================================================================================
extern "C" {

const char *gqz_version(void) { return "gqz 0.1 (gfx950, fp64)"; }

// PIPARMQ / PILAENVX values (piparmq.f:300-531), the same functions the QZ uses for its defaults:
// ISPEC 50 NMIN1, 51 NW, 52 NIBBLE (PDHGEQZ0 form), 53 NS, 54 NMIN2, 55 NMIN3.
int gqz_piparmq(int ispec, int nh) {
    switch (ispec) {
    case 50: return 6000;
    case 51: return piparmq_nw(nh);
    case 52: return piparmq_nibble(nh);
    case 53: return piparmq_ns(nh);
    case 54: return 201;
    case 55: return 201;
    default: return -1;
    }
}
const char *gqz_last_error(void) { return L.err.c_str(); }
void *gqz_stream(void) { return (void *)L.st; }

int gqz_init(int device) {
    try {
        std::lock_guard<std::mutex> g(L.mu);
        int cnt = 0;
        if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0) {
            L.err = "no HIP device available: the MI355X QZ library requires a GPU";
            return -1;
        }
        if (device < 0 || device >= cnt) { L.err = "bad device index"; return -1; }
        if (L.dev == device) return 0;
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&L.st2, hipStreamNonBlocking));
        HIPCHK(hipMalloc((void **)&L.d_out, 64 * sizeof(int)));
        HIPCHK(hipMalloc((void **)&L.d_norm, 8 * sizeof(double)));
        HIPCHK(hipHostMalloc((void **)&L.h_out, 64 * sizeof(int), hipHostMallocDefault));
        L.h_buf_n = 3 * 8192 + 64;
        HIPCHK(hipHostMalloc((void **)&L.h_buf, L.h_buf_n * sizeof(double), hipHostMallocDefault));
        L.dev = device;
        if (gqz_batched_scratch_setup(device) != 0) return -1;
        return 0;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -1;
    }
}

// Releases what gqz_init and the calls since allocated.  Every release is attempted even after a
// failure; the first failure is reported (-1, text in gqz_last_error).
int gqz_finalize(void) {
    if (L.dev < 0) return 0;
    hipError_t first = hipSuccess;
    auto note = [&first](hipError_t e) { if (first == hipSuccess) first = e; };
    note(hipSetDevice(L.dev));
    note(hipStreamSynchronize(L.st));
    if (L.st2) note(hipStreamSynchronize(L.st2));
    if (L.ws.base) note(hipFree(L.ws.base));
    L.ws = Arena();
    note(hipFree(L.d_out)); note(hipFree(L.d_norm));
    note(hipHostFree(L.h_out)); note(hipHostFree(L.h_buf));
    if (L.d_mlp) {
        note(hipFree(L.d_mlp)); note(hipFree(L.d_mlp_x)); note(hipFree(L.d_mlp_y));
        note(hipHostFree(L.h_mlp_io));
    }
    L.d_out = nullptr; L.d_norm = nullptr; L.h_out = nullptr; L.h_buf = nullptr;
    L.d_mlp = L.d_mlp_x = L.d_mlp_y = L.h_mlp_io = nullptr;
    L.mlp_on = false;
    for (hipEvent_t e : L.sev) note(hipEventDestroy(e));
    L.sev.clear();
    note(hipStreamDestroy(L.st));
    if (L.st2) note(hipStreamDestroy(L.st2));
    L.st = nullptr; L.st2 = nullptr; L.dev = -1;
    if (first != hipSuccess) {
        L.err = std::string("gqz_finalize: ") + hipGetErrorString(first);
        return -1;
    }
    return 0;
}

int gqz_sync(void) {
    if (L.dev < 0) return 0;
    return hipStreamSynchronize(L.st) == hipSuccess ? 0 : -1;
}

int gqz_ktime_enable(unsigned mask) {
    if (L.dev >= 0) {
        if (hipStreamSynchronize(L.st) != hipSuccess) { L.err = "gqz_ktime_enable: stream error"; return -1; }
        kt_drain();
    }
    KT.mask = mask;
    KT.rng = 1;
    for (int c = 0; c < 2 * GQZ_KC_COUNT; ++c) {
        KT.launches[c] = KT.timed[c] = 0;
        KT.ms[c] = KT.flops[c] = KT.bytes[c] = 0.0;
    }
    return 0;
}

int gqz_ktime_sample(int every) {
    if (every < 1) return -1;
    KT.every = every;
    return 0;
}

int gqz_ktime_get(gqz_ktime *out, int cap) {
    static const char *names[2 * GQZ_KC_COUNT] = {
        "slab_update", "chase_window", "small_qz", "aed_reorder", "aed_spike", "ht_givens", "inf_deflation",
        "slab_update@aed", "chase_window@aed", "small_qz@aed", "aed_reorder@aed", "aed_spike@aed",
        "ht_givens@aed", "inf_deflation@aed" };
    if (L.dev >= 0) {
        if (hipStreamSynchronize(L.st) != hipSuccess) return -1;
        kt_drain();
    }
    int k = std::min(cap, (int)(2 * GQZ_KC_COUNT));
    for (int c = 0; c < k; ++c) {
        // summed time of the timed launches, scaled to all launches of the class
        out[c].name = names[c]; out[c].launches = KT.launches[c];
        out[c].ms = KT.timed[c] ? KT.ms[c] * (double)KT.launches[c] / (double)KT.timed[c] : 0.0;
        out[c].flops = KT.flops[c]; out[c].bytes = KT.bytes[c];
    }
    return k;
}

int gqz_mlp_nparams_c(void) { return gqz_mlp_nparams(); }

int gqz_mlp_forward_dev(const float *params, const float *x, int batch, float *y) {
    if (batch < 0 || !params || (batch > 0 && (!x || !y))) return -1;
    try {
        if (check_init()) return -100;
        gqz_launch_mlp_fwd(L.st, params, x, batch, y);
        HIPCHK(hipStreamSynchronize(L.st));
        return 0;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

int gqz_mlp_grad_dev(const float *params, const float *x, const float *target, int batch, float *grad,
                     float *loss) {
    if (batch <= 0 || !params || !x || !target || !grad) return -1;
    try {
        if (check_init()) return -100;
        gqz_launch_mlp_grad(L.st, params, x, target, batch, grad, loss);
        HIPCHK(hipStreamSynchronize(L.st));
        return 0;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

int gqz_mlp_train_step_dev(float *params, float *grad, float *m, float *v, int step, float lr, const float *x,
                           const float *target, int batch, float *loss) {
    if (batch <= 0 || step < 1 || !params || !grad || !m || !v || !x || !target) return -1;
    try {
        if (check_init()) return -100;
        gqz_launch_mlp_grad(L.st, params, x, target, batch, grad, loss);
        gqz_launch_mlp_adam(L.st, params, grad, m, v, step, lr);
        return 0;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

int gqz_mlp_policy(const float *params_host, int enable) {
    try {
        if (check_init()) return -100;
        if (!L.d_mlp) {
            HIPCHK(hipMalloc((void **)&L.d_mlp, sizeof(float) * (gqz_mlp_nparams() + 16)));
            HIPCHK(hipMalloc((void **)&L.d_mlp_x, sizeof(float) * 16));
            HIPCHK(hipMalloc((void **)&L.d_mlp_y, sizeof(float) * 16));
            HIPCHK(hipHostMalloc((void **)&L.h_mlp_io, sizeof(float) * 16, hipHostMallocDefault));
        }
        if (params_host) {
            HIPCHK(hipMemcpy(L.d_mlp, params_host, sizeof(float) * gqz_mlp_nparams(), hipMemcpyHostToDevice));
        }
        L.mlp_on = enable && (params_host || L.mlp_on);
        return 0;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

int gqz_set_policy(gqz_policy_fn fn, void *user) { L.pol = fn; L.pol_user = user; return 0; }
int gqz_trace_enable(int cap) { L.trace.clear(); L.trace_cap = std::max(0, cap); return 0; }
int gqz_progress(double seconds) { L.progress_s = seconds > 0.0 ? seconds : 0.0; return 0; }
int gqz_overlap(int on) {
    const int prev = L.overlap ? 1 : 0;
    if (on >= 0) L.overlap = on != 0;
    return prev;
}
int gqz_overlap_min(int n) {
    const int prev = L.overlap_min;
    if (n >= 0) L.overlap_min = n;
    return prev;
}
int gqz_sched_replay(int n, int ktop, int kbot, int nch, const int *nb, int ovl, int *out, int cap) {
    const int wmax = gqz_chase_max_window();
    if (n < 1 || ktop < 0 || kbot >= n || kbot - ktop < 2 || nch < 1 || !nb || cap < 0 || (cap > 0 && !out))
        return -1;
    std::vector<Chain> ch(nch);
    for (int c = 0; c < nch; ++c) {
        if (nb[c] < 1 || 3 * nb[c] + 2 > wmax) return -1;
        ch[c] = Chain{ 0, nb[c], ktop, false, false, INT32_MAX, 0, 0, 0 };
    }
    sched_reset(ch);
    int first = 0, rec = 0, round = 0;
    auto put = [&](const int *r) {
        if (rec < cap) for (int k = 0; k < 8; ++k) out[8 * rec + k] = r[k];
        ++rec;
    };
    while (first < nch) {
        SchedWin sw[GQZ_MAXCH];
        const int nw = sched_select(ch, ktop, kbot, n, ovl != 0, wmax, GQZ_MAXCH, sw);
        if (nw == 0) {
            if (!ovl || !sched_unblock(ch)) return -1;
            const int r[8] = { round, -1, 0, 0, 0, 0, 0, 0 };       // stream 2 drained
            put(r);
            continue;
        }
        for (int q = 0; q < nw; ++q) {
            const SchedWin &o = sw[q];
            const int r[8] = { round, o.chain, o.p, o.w, o.steps, o.last, o.nc, o.near };
            put(r);
        }
        sched_commit(ch, sw, nw);
        while (first < nch && ch[first].done) ++first;
        if (++round > 100000000) return -1;
    }
    return rec;
}
int gqz_trace_get(gqz_trace_rec *out, int cap) {
    int k = std::min<int>(cap, (int)L.trace.size());
    if (out) for (int i = 0; i < k; ++i) out[i] = L.trace[i];
    return k;
}

int gqz_hgeqz_dev(char job, char compq, char compz, int n, int ilo, int ihi, double *H, int ldh, double *T,
                  int ldt, double *alphar, double *alphai, double *beta, double *Q, int ldq, double *Z, int ldz,
                  const gqz_params *prm, gqz_stats *stats) {
    // argument checks in PDHGEQZ order (PDHGEQZ.f:344-405)
    int ijob = decode(job, "ES"), icq = decode(compq, "NIV"), icz = decode(compz, "NIV");
    if (ijob < 0) return -1;
    if (icq < 0) return -2;
    if (icz < 0) return -3;
    if (n < 0) return -4;
    if (ilo < 1) return -5;
    if (ihi > n || ihi < ilo - 1) return -6;
    if (ldh < std::max(1, n) || ldt != ldh) return -8;
    if (icq > 0 && (ldq < std::max(1, n) || !Q)) return -15;
    if (icz > 0 && (ldz < std::max(1, n) || !Z)) return -17;
    if (icq > 0 && icz > 0 && ldq != ldz) return -17;
    if (n == 0 || ihi < ilo) return 0;
    try {
        if (check_init()) return -100;
        double t0 = now();
        Run R;
        R.t_start = R.t_progress = t0;
        R.prm = prm ? *prm : gqz_params{};
        std::memset(&R.st, 0, sizeof R.st);
        R.n_top = n;
        ensure_ws(arena_bytes(n, prm));
        L.ws.top = 0;
        if (L.trace_cap > 0) L.trace.clear();
        // TTOL = ulp * ||T(1:NH, 1:NH)||_F, NH = IHI - ILO + 1: the reference takes the leading NH x NH
        // block (PDLANGE('F', NH, NH, T, 1, 1, ...), PDHGEQZ.f:453-458), not T(ILO:IHI, ILO:IHI)
        HIPCHK(hipMemsetAsync(L.d_norm, 0, sizeof(double), L.st));
        gqz_launch_sumsq(L.st, T, ldt, ihi - ilo + 1, ihi - ilo + 1, 1, L.d_norm);
        double tf2 = 0.0;
        sync_read(&tf2, L.d_norm, sizeof(double));
        R.ttol = std::max(kSafmin, kUlp * std::sqrt(tf2));
        if (icq == 1) gqz_launch_identity(L.st, Q, ldq, n, n);
        if (icz == 1) gqz_launch_identity(L.st, Z, ldz, n, n);
        gqz_launch_fill(L.st, alphar + (ilo - 1), ihi - ilo + 1, -1.0);
        gqz_launch_fill(L.st, alphai + (ilo - 1), ihi - ilo + 1, -1.0);
        gqz_launch_fill(L.st, beta + (ilo - 1), ihi - ilo + 1, -1.0);
        View v;
        v.H = H; v.T = T; v.Q = icq ? Q : nullptr; v.Z = icz ? Z : nullptr;
        v.ld = ldh; v.ldq = icq ? ldq : ldz; v.n = n; v.nq = n;
        v.ar = alphar; v.ai = alphai; v.be = beta;
        v.smlnum = kSafmin * ((double)(ihi - ilo + 1) / kUlp);
        int info = qz_run(R, v, ilo - 1, ihi - 1, 0);
        // eigenvalues (and beta >= 0) from the final Schur form, including the parts outside
        // [ilo, ihi] which are read off the diagonal
        gqz_launch_finalize(L.st, H, T, v.Z, ldh, v.ldq, n, 0, n - 1, alphar, alphai, beta);
        HIPCHK(hipStreamSynchronize(L.st));
        R.st.t_total = now() - t0;
        if (stats) *stats = R.st;
        return info;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

int gqz_geqrb_dev(char compq, int n, double *A, int lda, double *B, int ldb, double *Q, int ldq) {
    int icq = decode(compq, "NIV");
    if (icq < 0) return -1;
    if (n < 0) return -2;
    if (n == 0) return 0;
    try {
        if (check_init()) return -100;
        DevBuf tau_buf;
        double *tau = tau_buf.alloc_async(sizeof(double) * (size_t)n, L.st);
        if (icq == 1) gqz_launch_identity(L.st, Q, ldq, n, n);
        if (n > 64) {
            DevBuf wbuf;
            double *wq = wbuf.alloc_async(sizeof(double) * gqz_house_qr_ws_doubles(n), L.st);
            gqz_launch_house_qr_blk(L.st, n, A, lda, B, ldb, icq ? Q : nullptr, ldq, tau, wq);
            HIPCHK(hipStreamSynchronize(L.st));
        } else {
            gqz_launch_house_qr(L.st, n, A, lda, B, ldb, icq ? Q : nullptr, ldq, tau);
        }
        HIPCHK(hipStreamSynchronize(L.st));
        return 0;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

int gqz_gghrd_dev(char compq, char compz, int n, int ilo, int ihi, double *A, int lda, double *B, int ldb,
                  double *Q, int ldq, double *Z, int ldz) {
    int icq = decode(compq, "NIV"), icz = decode(compz, "NIV");
    if (icq < 0) return -1;
    if (icz < 0) return -2;
    if (n < 0) return -3;
    if (ilo < 1) return -4;
    if (ihi > n || ihi < ilo - 1) return -5;
    if (lda < std::max(1, n) || ldb != lda) return -7;
    if ((icq > 0 && icz > 0) && ldq != ldz) return -13;
    if (n == 0) return 0;
    try {
        if (check_init()) return -100;
        if (icq == 1) gqz_launch_identity(L.st, Q, ldq, n, n);
        if (icz == 1) gqz_launch_identity(L.st, Z, ldz, n, n);
        if (ihi - ilo >= 2) {
            if (ihi - ilo + 1 > kHtBlockedMin) {
                DevBuf rws;
                rws.alloc_async(gqz_gghrd_blk_ws_bytes(n, n), L.st);
                // with progress lines on, the column sweeps go in chunks with a line after each
                const int chunk = L.progress_s > 0.0 ? 256 : n;
                const double t0 = now();
                double tp = t0;
                for (int j0 = ilo - 1; j0 <= ihi - 3; j0 += chunk) {
                    gqz_launch_gghrd_blk(L.st, A, B, lda, n, j0, ihi - 1, icq ? Q : nullptr, icz ? Z : nullptr,
                                         icq ? ldq : ldz, n, rws.p, j0 + chunk - 1);
                    if (L.progress_s > 0.0) {
                        HIPCHK(hipStreamSynchronize(L.st));
                        const double t = now();
                        if (t - tp >= L.progress_s) {
                            tp = t;
                            std::fprintf(stderr, "[gqz] HT reduction n=%d: %d of %d column sweeps, %.1f s\n", n,
                                         std::min(j0 + chunk, ihi - 2) - (ilo - 1), ihi - ilo - 1, t - t0);
                            std::fflush(stderr);
                        }
                    }
                }
                HIPCHK(hipStreamSynchronize(L.st));
            } else {
                gqz_launch_gghrd_wg(L.st, A, B, lda, n, ilo - 1, ihi - 1, icq ? Q : nullptr, icz ? Z : nullptr,
                                    icq ? ldq : ldz, n);
            }
        }
        HIPCHK(hipStreamSynchronize(L.st));
        return 0;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

int gqz_compres_dev(int n, const double *A, const double *B, const double *S, const double *T, const double *Q,
                    const double *Z, int ld, double out[4]) {
    if (n <= 0) return -1;
    try {
        if (check_init()) return -100;
        size_t nn = (size_t)n * n;
        DevBuf w1_buf, w2_buf;
        double *W1 = w1_buf.alloc_async(nn * sizeof(double), L.st);
        double *W2 = w2_buf.alloc_async(nn * sizeof(double), L.st);
        HIPCHK(hipMemsetAsync(L.d_norm, 0, 8 * sizeof(double), L.st));
        const double *X[2] = { A, B }, *Y[2] = { S, T };
        for (int t = 0; t < 2; ++t) {
            gqz_gemm(L.st, 0, 0, n, n, n, X[t], ld, Z, ld, W1, n);        // X Z
            gqz_gemm(L.st, 1, 0, n, n, n, Q, ld, W1, n, W2, n);           // Q^T X Z
            gqz_launch_sub(L.st, W2, n, Y[t], ld, n);
            gqz_launch_sumsq(L.st, W2, n, n, n, 0, L.d_norm + 2 * t);
            gqz_launch_sumsq(L.st, X[t], ld, n, n, 0, L.d_norm + 2 * t + 1);
        }
        const double *O[2] = { Q, Z };
        for (int t = 0; t < 2; ++t) {
            gqz_gemm(L.st, 1, 0, n, n, n, O[t], ld, O[t], ld, W2, n);
            gqz_launch_sub(L.st, W2, n, nullptr, n, n);
            gqz_launch_sumsq(L.st, W2, n, n, n, 0, L.d_norm + 4 + t);
        }
        double h[8];
        HIPCHK(hipMemcpyAsync(h, L.d_norm, 8 * sizeof(double), hipMemcpyDeviceToHost, L.st));
        HIPCHK(hipStreamSynchronize(L.st));
        out[0] = std::sqrt(h[0]) / std::sqrt(h[1]);
        out[1] = std::sqrt(h[2]) / std::sqrt(h[3]);
        out[2] = std::sqrt(h[4]) / (n * kUlp);
        out[3] = std::sqrt(h[5]) / (n * kUlp);
        return 0;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

int gqz_slab_update_dev(char side, int m, int k, double *Cm, int ldc, const double *F, int ldf) {
    try {
        if (check_init()) return -100;
        char s = (char)toupper((unsigned char)side);
        if (s == 'R') gqz_slab_right(L.st, Cm, ldc, m, k, F, ldf, nullptr);
        else if (s == 'L') gqz_slab_left(L.st, Cm, ldc, k, m, F, ldf, nullptr);
        else return -1;
        HIPCHK(hipStreamSynchronize(L.st));
        return 0;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

int gqz_chase_window_dev(double *H, double *T, int ld, int p, int w, int nb, int t0, int steps, int intro,
                         int last, double *U, double *V, const double *sr, const double *si, const double *sb) {
    if (w < 4 || w > gqz_chase_max_window() || nb < 1 || nb > 20) return -1;
    try {
        if (check_init()) return -100;
        GqzChainArgs a = {};
        a.H = H; a.T = T; a.ld = ld; a.p = p; a.w = w; a.nb = nb; a.t0 = t0; a.steps = steps; a.intro = intro;
        a.last = last; a.U = U; a.V = V; a.sr = sr; a.si = si; a.sb = sb;
        a.smlnum = kSafmin * ((double)w / kUlp);
        gqz_chase_window(L.st, a);
        HIPCHK(hipStreamSynchronize(L.st));
        return 0;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

int gqz_small_qz_dev(int m, double *H, double *T, int ld, double *U, double *V, double *alphar, double *alphai,
                     double *beta, double ttol) {
    if (m < 1 || m > 64) return -1;
    try {
        if (check_init()) return -100;
        GqzSmallJob J = {};
        J.H = H; J.T = T; J.ld = ld; J.m = m; J.U = U; J.V = V; J.ldu = m;
        J.ar = alphar; J.ai = alphai; J.be = beta; J.info = L.d_out + 10; J.ttol = ttol;
        J.smlnum = kSafmin * ((double)m / kUlp);
        gqz_launch_small1(L.st, J);
        sync_read(L.h_out + 10, L.d_out + 10, sizeof(int));
        return L.h_out[10];
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

// ---- host-pointer entry points: copy in, run, copy out ---------------------------------------
static int host_roundtrip_hgeqz(char job, char compq, char compz, int n, int ilo, int ihi, double *H, int ldh,
                                double *T, int ldt, double *alphar, double *alphai, double *beta, double *Q,
                                int ldq, double *Z, int ldz, const gqz_params *prm, gqz_stats *stats) {
    if (n <= 0) return gqz_hgeqz_dev(job, compq, compz, n, ilo, ihi, H, ldh, T, ldt, alphar, alphai, beta, Q,
                                     ldq, Z, ldz, prm, stats);
    int icq = decode(compq, "NIV"), icz = decode(compz, "NIV");
    if (icq < 0) return -2;
    if (icz < 0) return -3;
    if (ldh < n) return -8;
    if (ldt < n) return -10;
    try {
        if (check_init()) return -100;
        size_t nn = (size_t)n * n;
        DevBuf buf;
        double *d = buf.alloc(sizeof(double) * (4 * nn + 3 * (size_t)n));
        double *dH = d, *dT = d + nn, *dQ = d + 2 * nn, *dZ = d + 3 * nn, *de = d + 4 * nn;
        HIPCHK(hipMemcpy2DAsync(dH, n * sizeof(double), H, ldh * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        HIPCHK(hipMemcpy2DAsync(dT, n * sizeof(double), T, ldt * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        if (icq == 2) HIPCHK(hipMemcpy2DAsync(dQ, n * sizeof(double), Q, ldq * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        if (icz == 2) HIPCHK(hipMemcpy2DAsync(dZ, n * sizeof(double), Z, ldz * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        int info = gqz_hgeqz_dev(job, compq, compz, n, ilo, ihi, dH, n, dT, n, de, de + n, de + 2 * n,
                                 icq ? dQ : nullptr, n, icz ? dZ : nullptr, n, prm, stats);
        if (info >= 0) {
            HIPCHK(hipMemcpy2DAsync(H, ldh * sizeof(double), dH, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
            HIPCHK(hipMemcpy2DAsync(T, ldt * sizeof(double), dT, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
            if (icq) HIPCHK(hipMemcpy2DAsync(Q, ldq * sizeof(double), dQ, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
            if (icz) HIPCHK(hipMemcpy2DAsync(Z, ldz * sizeof(double), dZ, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
            HIPCHK(hipMemcpyAsync(alphar, de, n * sizeof(double), hipMemcpyDeviceToHost, L.st));
            HIPCHK(hipMemcpyAsync(alphai, de + n, n * sizeof(double), hipMemcpyDeviceToHost, L.st));
            HIPCHK(hipMemcpyAsync(beta, de + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost, L.st));
        }
        HIPCHK(hipStreamSynchronize(L.st));
        return info;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

int gqz_hgeqz(char job, char compq, char compz, int n, int ilo, int ihi, double *H, int ldh, double *T, int ldt,
              double *alphar, double *alphai, double *beta, double *Q, int ldq, double *Z, int ldz,
              const gqz_params *prm, gqz_stats *stats) {
    return host_roundtrip_hgeqz(job, compq, compz, n, ilo, ihi, H, ldh, T, ldt, alphar, alphai, beta, Q, ldq, Z,
                                ldz, prm, stats);
}

int gqz_gghrd(char compq, char compz, int n, int ilo, int ihi, double *A, int lda, double *B, int ldb, double *Q,
              int ldq, double *Z, int ldz) {
    int icq = decode(compq, "NIV"), icz = decode(compz, "NIV");
    if (icq < 0) return -1;
    if (icz < 0) return -2;
    if (n <= 0) return n < 0 ? -3 : 0;
    try {
        if (check_init()) return -100;
        size_t nn = (size_t)n * n;
        DevBuf buf;
        double *d = buf.alloc(sizeof(double) * 4 * nn);
        double *dA = d, *dB = d + nn, *dQ = d + 2 * nn, *dZ = d + 3 * nn;
        HIPCHK(hipMemcpy2DAsync(dA, n * sizeof(double), A, lda * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        HIPCHK(hipMemcpy2DAsync(dB, n * sizeof(double), B, ldb * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        if (icq == 2) HIPCHK(hipMemcpy2DAsync(dQ, n * sizeof(double), Q, ldq * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        if (icz == 2) HIPCHK(hipMemcpy2DAsync(dZ, n * sizeof(double), Z, ldz * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        int info = gqz_gghrd_dev(compq, compz, n, ilo, ihi, dA, n, dB, n, dQ, n, dZ, n);
        if (info == 0) {
            HIPCHK(hipMemcpy2DAsync(A, lda * sizeof(double), dA, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
            HIPCHK(hipMemcpy2DAsync(B, ldb * sizeof(double), dB, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
            if (icq) HIPCHK(hipMemcpy2DAsync(Q, ldq * sizeof(double), dQ, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
            if (icz) HIPCHK(hipMemcpy2DAsync(Z, ldz * sizeof(double), dZ, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
        }
        HIPCHK(hipStreamSynchronize(L.st));
        return info;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

// PDGGHRD with the routine's own Q convention (host pointers): COMPQ='I' returns Q_ht^T,
// COMPQ='V' returns Q_ht^T * Q1 (see qz_desc.cpp); the product runs on the device.
int gqz_gghrd_rowq(char compq, char compz, int n, int ilo, int ihi, double *A, int lda, double *B, int ldb,
                   double *Q, int ldq, double *Z, int ldz) {
    int icq = decode(compq, "NIV"), icz = decode(compz, "NIV");
    if (icq < 0) return -1;
    if (icz < 0) return -2;
    if (n <= 0) return n < 0 ? -3 : 0;
    try {
        if (check_init()) return -100;
        size_t nn = (size_t)n * n;
        DevBuf buf;
        double *d = buf.alloc(sizeof(double) * 6 * nn);
        double *dA = d, *dB = d + nn, *dQh = d + 2 * nn, *dZ = d + 3 * nn, *dQ1 = d + 4 * nn, *dW = d + 5 * nn;
        HIPCHK(hipMemcpy2DAsync(dA, n * sizeof(double), A, lda * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        HIPCHK(hipMemcpy2DAsync(dB, n * sizeof(double), B, ldb * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        if (icz == 2) HIPCHK(hipMemcpy2DAsync(dZ, n * sizeof(double), Z, ldz * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        if (icq == 2) HIPCHK(hipMemcpy2DAsync(dQ1, n * sizeof(double), Q, ldq * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice, L.st));
        else if (icq == 1) gqz_launch_identity(L.st, dQ1, n, n, n);
        int info = gqz_gghrd_dev(icq ? 'I' : 'N', compz, n, ilo, ihi, dA, n, dB, n, dQh, n, dZ, n);
        if (info == 0) {
            if (icq) gqz_gemm(L.st, 1, 0, n, n, n, dQh, n, dQ1, n, dW, n);
            HIPCHK(hipMemcpy2DAsync(A, lda * sizeof(double), dA, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
            HIPCHK(hipMemcpy2DAsync(B, ldb * sizeof(double), dB, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
            if (icq) HIPCHK(hipMemcpy2DAsync(Q, ldq * sizeof(double), dW, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
            if (icz) HIPCHK(hipMemcpy2DAsync(Z, ldz * sizeof(double), dZ, n * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost, L.st));
        }
        HIPCHK(hipStreamSynchronize(L.st));
        return info;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

// ---- PDTGORD role: reorder a generalized Schur form so that selected eigenvalues lead ----------
// Windowed reordering (the delay-and-accumulate scheme of src/reorder/PDTGORD.f:600-1600 on a 1x1
// grid): each pass carries the next `carry` selected eigenvalues (whole blocks) below the sorted
// top region up to it through a chain of diagonal windows of order <= `wmax`, each window bottom-up
// from the carried blocks; inside a window the LDS kernel moves the selected blocks to its top with
// DTGEX2 swaps and accumulates U, V, and the MFMA slab kernels apply U to the rows right of the
// window (and Q), V to the columns above it (and Z).  Returns 0, or the 1-based row of the pair whose
// swap was rejected (the pencil is then partially reordered, consistently), as PDTGORD's INFO.
// sel (host, n) is updated to the final positions of the selected eigenvalues; *m is their number.
int gqz_tgord_dev(int wantq, int wantz, int *sel, const int *para, int n, double *S, int lds, double *T, int ldt,
                  double *Q, int ldq, double *Z, int ldz, double *alphar, double *alphai, double *beta, int *m) {
    if (!sel && n > 0) return -3;
    if (n < 0) return -5;
    if (lds < std::max(1, n)) return -7;
    if (ldt != lds) return -9;
    if (wantq && (!Q || ldq < std::max(1, n))) return -11;
    if (wantz && (!Z || ldz < std::max(1, n))) return -13;
    if (m) *m = 0;
    if (n == 0) return 0;
    try {
        if (check_init()) return -100;
        const int wcap = gqz_tgord_max_window();
        int wmax = para && para[2] > 0 ? std::min(para[2], wcap) : wcap;
        wmax = std::max(4, wmax);
        int carry = para && para[1] > 0 ? para[1] : wmax / 2;
        carry = std::max(2, std::min(carry, wmax / 2));
        // block structure: sub[i] = (S(i+1, i) != 0)
        std::vector<double> sd(std::max(1, n - 1));
        if (n > 1) {
            HIPCHK(hipMemcpy2DAsync(sd.data(), sizeof(double), S + 1, (size_t)(lds + 1) * sizeof(double),
                                    sizeof(double), n - 1, hipMemcpyDeviceToHost, L.st));
            HIPCHK(hipStreamSynchronize(L.st));
        }
        std::vector<char> sub(n, 0), sl(n, 0);
        for (int i = 0; i + 1 < n; ++i) sub[i] = sd[i] != 0.0;
        // a complex pair is selected when either of its rows is (PDTGORD.f:512-527)
        for (int k = 0; k < n; ++k) sl[k] = sel[k] != 0;
        for (int k = 0; k + 1 < n; ++k)
            if (sub[k] && (sl[k] || sl[k + 1])) sl[k] = sl[k + 1] = 1;
        int msel = 0;
        for (int k = 0; k < n; ++k) msel += sl[k];
        if (m) *m = msel;
        size_t mk = L.ws.mark();
        ensure_ws(arena_bytes(std::min(n, 64), nullptr));
        mk = L.ws.mark();
        double *U = L.ws.alloc((size_t)wcap * wcap), *V = L.ws.alloc((size_t)wcap * wcap);
        Run R;
        R.prm = gqz_params{};
        std::memset(&R.st, 0, sizeof R.st);
        int info = 0;
        int ilo = 0;
        while (ilo < n && sl[ilo]) ++ilo;
        auto bsize = [&](int k) { return (k + 1 < n && sub[k]) ? 2 : 1; };
        while (ilo < msel && info == 0) {
            // the carried set: the first selected blocks at or below ilo, at most `carry` rows
            int k = ilo, cnt = 0, last = -1;
            while (k < n) {
                const int bs = bsize(k);
                if (sl[k]) {
                    if (cnt + bs > carry && cnt > 0) break;
                    cnt += bs; last = k + bs - 1;
                    if (cnt >= carry) break;
                }
                k += bs;
            }
            if (cnt == 0) break;
            int bottom = last + 1;
            for (;;) {
                int p = std::max(ilo, bottom - wmax);
                if (p > ilo && sub[p - 1]) ++p;           // keep 2x2 blocks whole
                const int w = bottom - p;
                unsigned long long mask = 0ull;
                for (int i = 0; i < w; ++i) if (sl[p + i]) mask |= 1ull << i;
                gqz_launch_tgord_window(L.st, S, T, lds, p, w, mask, U, V, L.d_out + 24);
                // far updates: rows of the window right of it (S, T) and Q by U; columns above it
                // (S, T) and Z by V
                GqzSlabJobs J;
                J.njobs = 0;
                gqz_slab_add(J, S + (size_t)p + (size_t)bottom * lds, lds, n - bottom, U, w, w);
                gqz_slab_add(J, T + (size_t)p + (size_t)bottom * ldt, ldt, n - bottom, U, w, w);
                slab_launch(R, J, true);
                gqz_slab_add(J, S + (size_t)p * lds, lds, p, V, w, w);
                gqz_slab_add(J, T + (size_t)p * ldt, ldt, p, V, w, w);
                if (wantq) gqz_slab_add(J, Q + (size_t)p * ldq, ldq, n, U, w, w);
                if (wantz) gqz_slab_add(J, Z + (size_t)p * ldz, ldz, n, V, w, w);
                slab_launch(R, J, false);
                sync_read(L.h_out + 24, L.d_out + 24, 6 * sizeof(int));
                const int *o = L.h_out + 24;
                const unsigned long long ms = (unsigned long long)(unsigned)o[1] | ((unsigned long long)(unsigned)o[2] << 32);
                const unsigned long long mb = (unsigned long long)(unsigned)o[3] | ((unsigned long long)(unsigned)o[4] << 32);
                for (int i = 0; i < w; ++i) {
                    sl[p + i] = (ms >> i) & 1ull;
                    if (i + 1 < w) sub[p + i] = (mb >> i) & 1ull;
                }
                if (o[0]) { info = p + o[0]; break; }
                int top = 0;
                while (top < w && sl[p + top]) ++top;
                if (p == ilo) { ilo += top; while (ilo < n && sl[ilo]) ++ilo; break; }
                int nb = std::min(top, carry);
                if (nb < top && nb > 0 && sub[p + nb - 1]) --nb;     // do not cut a 2x2 block
                if (nb <= 0) break;
                bottom = p + nb;
            }
        }
        L.ws.release(mk);
        for (int k = 0; k < n; ++k) sel[k] = sl[k];
        gqz_launch_finalize(L.st, S, T, wantz ? Z : nullptr, lds, ldz, n, 0, n - 1, alphar, alphai, beta);
        HIPCHK(hipStreamSynchronize(L.st));
        return info;
    } catch (const std::exception &e) {
        L.err = e.what();
        return -100;
    }
}

}  // extern "C"
