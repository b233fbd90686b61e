// gqz_misc.h -- host launchers of the auxiliary kernels (k_misc.hip, k_wgqz.hip).
#pragma once
#include "gqz_kernels.h"

// kt0 >= 0: active block starts at kt0 (no search); chase = 0: an interior zero is reported in out[4]
// instead of being chased (see k_misc.hip)
void gqz_launch_inf_defl(hipStream_t st, double *H, double *T, double *Q, double *Z, int ld, int ldq,
                         int n, int nq, int ilo, int kbot, double ttol, double smlnum, double *ar,
                         double *ai, double *be, int *out, int kt0 = -1, int chase = 1);
// the interior-zero chase through the LDS window [p, p + w), w <= gqz_zchase_max_window(): the zero at
// T(p+1, p+1) ends at T(p+w-1, p+w-1); U, V (w x w, ld w) for the off-window updates
int gqz_zchase_max_window();
void gqz_launch_zchase_window(hipStream_t st, double *H, double *T, int ld, int p, int w, double *U, double *V);
void gqz_launch_window_in(hipStream_t st, const double *H, const double *T, int ld, int k0, int jw,
                          double *Hw, double *Tw, double *U, double *V);
void gqz_launch_window_out(hipStream_t st, double *H, double *T, int ld, int k0, int jw, const double *Hw,
                           const double *Tw, const double *U, int has_spike);
void gqz_launch_identity(hipStream_t st, double *A, int ld, int m, int n);
void gqz_launch_fill(hipStream_t st, double *x, int n, double v);
void gqz_launch_finalize(hipStream_t st, double *H, double *T, double *Z, int ld, int ldz, int nz, int ilo,
                         int ihi, double *ar, double *ai, double *be);
void gqz_launch_small(hipStream_t st, const GqzSmallJob *djobs, int njobs, int mmax);
void gqz_launch_small1(hipStream_t st, const GqzSmallJob &job);
void gqz_launch_gghrd_wg(hipStream_t st, double *A, double *B, int ld, int n, int ilo, int ihi, double *Q,
                         double *Z, int ldq, int nq);
// Givens HT reduction with the rotations of each column sweep applied in bulk (k_ht_blk.hip); ws holds
// gqz_gghrd_blk_ws_bytes(n, nq) bytes (rotation list, progress counters, per-block carries)
size_t gqz_gghrd_blk_ws_bytes(int n, int nq);
// (the column sweeps j = ilo .. min(jlast, ihi - 2); jlast < 0: all of them; a call for a later range
// continues the reduction, since a sweep depends only on j and ihi)
void gqz_launch_gghrd_blk(hipStream_t st, double *A, double *B, int ld, int n, int ilo, int ihi, double *Q,
                          double *Z, int ldq, int nq, void *ws, int jlast = -1);
void gqz_launch_house_qr(hipStream_t st, int n, double *A, int lda, double *B, int ldb, double *Q, int ldq,
                         double *tau);
// blocked QR of B (compact WY panels of 32 columns, GEMM updates); ws: gqz_house_qr_ws_doubles(n)
size_t gqz_house_qr_ws_doubles(int n);
void gqz_launch_house_qr_blk(hipStream_t st, int n, double *A, int lda, double *B, int ldb, double *Q, int ldq,
                             double *tau, double *ws);
void gqz_launch_sumsq(hipStream_t st, const double *A, int ld, int m, int n, int upper, double *out);
void gqz_launch_sub(hipStream_t st, double *R, int ldr, const double *S, int lds, int n);
// AED window post-processing (k_aed.hip); pscr holds (jw/2 + 1) * 65 doubles
void gqz_launch_aed_reorder(hipStream_t st, double *Hw, double *Tw, double *U, double *V, int jw,
                            const double *sptr, double smlnum, double *pscr, int *out);
// whole AED of a window of order jw <= gqz_aed_small_max() in one LDS-resident kernel; out: 5 ints
// (kbc, swaps, reorder failed, spike present, INFO of the window QZ); pscr: (jw/2+1)*65 doubles,
// scr: 4 jw doubles
int gqz_aed_small_max();
void gqz_launch_aed_small(hipStream_t st, double *H, double *T, int ldh, int kwtop, int jw, int has_spike,
                          double ttol, double smlnum, double *U, double *V, double *war, double *wai, double *wb,
                          double *pscr, double *scr, int *out);
void gqz_launch_aed_eigs(hipStream_t st, const double *Hw, const double *Tw, int jw, double *war, double *wai,
                         double *wb);
// eigenvalue reordering inside one diagonal window [p, p + w), w <= gqz_tgord_max_window() (k_aed.hip);
// out: 6 ints (status, selection mask lo/hi, 2x2-block mask lo/hi, swaps); U, V: w x w, ld w
int gqz_tgord_max_window();
void gqz_launch_tgord_window(hipStream_t st, double *S, double *T, int ld, int p, int w, unsigned long long sel,
                             double *U, double *V, int *out);
// spike reflector + RQ update of T (scr: 4 * jw doubles); reads kbc from out[0]
void gqz_launch_aed_spike(hipStream_t st, double *Hw, double *Tw, double *U, double *V, int jw,
                          const double *sptr, double *scr, const int *out);

// Raise a kernel's dynamic-LDS limit on the current device to at least `bytes`, never lowering it:
// the attribute is process-wide, so callers on several host threads go through one lock.
#include <mutex>
inline void gqz_need_lds(const void *fn, size_t bytes, size_t *cur_per_device) {
    static std::mutex mu;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    std::lock_guard<std::mutex> g(mu);
    if (bytes > cur_per_device[dev]) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        cur_per_device[dev] = bytes;
    }
}

// library state shared with entry points living beside their kernels (qz_host.cpp)
int gqz_init_current();                 // initialise on device 0 if needed and make it current; 0 = ok
int gqz_ready_device();                 // the initialised device made current, -1 before gqz_init (no init)
extern "C" int gqz_batched_scratch_setup(int device);   // per-device scratch address (k_batched.hip)
void gqz_set_error(const char *msg);    // text returned by gqz_last_error

#include <stdexcept>
#include <string>
#define GQZ_HIP(x)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess)                                                                   \
            throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_));          \
    } while (0)

// learned parameter policy (k_mlp.hip): FP32 perceptron 12 -> 64 -> 64 -> 3
int gqz_mlp_nparams();
void gqz_launch_mlp_fwd(hipStream_t st, const float *P, const float *X, int B, float *Y);
void gqz_launch_mlp_grad(hipStream_t st, const float *P, const float *X, const float *Tg, int B, float *G,
                         float *loss);
void gqz_launch_mlp_adam(hipStream_t st, float *P, float *G, float *M, float *V, int step, float lr);
void gqz_launch_fill_f(hipStream_t st, float *x, int n, float v);
