// k_ht_blk.hip -- Givens Hessenberg-triangular reduction (PDGGHRD, src/ab2ht/PDGGHRD.f:365-642) with
// the rotations of a column sweep delayed and applied in bulk, as the reference groups its rotations
// before applying them (ACCUMULATEROWROTATIONS.f:237-330, BLOCKSLIVERROWUPDATE.f90).
//
// The rotations are those of the unblocked algorithm, in the same order: for column j = ilo..ihi-2
// and position c = ihi down to j+2, a row rotation R_c on rows (c-1, c) zeroes A(c, j), and a column
// rotation G_c on columns (c-1, c) zeroes the fill B(c, c-1) it leaves.  Positions are grouped in
// blocks of HT_K: block b holds c in [c_lo(b), c_hi(b)], c_hi(b) = ihi - HT_K b, and its "region" is
// B(r0..c_hi, r0..c_hi) with r0 = c_lo - 1.  Row i in [r0(b), c_hi(b)) is a row of region b.
//
// One sweep is seven launches:
//   k_ht_lead  (cooperative) workgroup 0 forms the rotations: region b in LDS, one wave runs the
//              rotation chain (lane 0 forms R_c and G_c, the other lanes apply them to the region's
//              row and column pair), the other waves prefetch region b+1 meanwhile.  Region b+1 is
//              untouched by earlier blocks except its last column c_hi(b+1) = r0(b), which has
//              received every earlier G.  That column is formed without applying the G's to B:
//              with P_b = G_ihi ... G_c_lo(b) (B <- B P_b), column r0(b) of B P_b is B u_b,
//              u_b = P_b e_r0(b), and u_b = [w_b ; s_b u_(b-1)] where w_b (on r0(b)..c_hi(b)-1) and
//              s_b = prod(-gs) come from block b's G's alone.  So for a row i of region b+1,
//                y(i) = acc_b(i),  acc_b = B(i, r0(b)..c_hi(b)-1) . w_b + s_b acc_(b-1),
//                acc_(-1) = B(i, ihi),
//              a Horner recurrence over the untouched row of B.  The other workgroups ("helpers",
//              HT_HR rows each) run that recurrence for the rows of later regions as the lead
//              publishes w_b, so the lead adds only the last term.  The lead writes region b back,
//              except that column, which goes to a side buffer (the far updates still need the
//              untouched entries), and publishes w_b, s_b and the block's rotations.
//   k_ht_far1/2/3 <G>  the sweep's column rotations on the rows 0..ihi of A, the rows of Z, the rows
//              0..j of B and, for a row of region g >= 1, blocks 0..g-1 (then the side-buffer value
//              into column c_hi(g)).  Each (row, block) is one item: the rotations of a block are an
//              affine map of the carried entry, x_out = gamma + delta x_in, with delta = prod(-gs) or
//              prod(rs) known from the rotations, so pass 1 forms gamma for every item, pass 2 runs
//              the carry along each row block by block, and pass 3 applies the block with its carry.
//   k_ht_far1/2/3 <R>  the row rotations, the same way, on the columns j+1.. of A, the rows of Q
//              (as column rotations of Q) and the columns of B right of each rotation's region.
// The column rotations go first: B's rows then match the lead's side buffer, and A's row and column
// updates touch the same entries, so they must not run together.
#include "gqz_dev.h"
#include "gqz_misc.h"

#define HT_K 32
#define HT_R (HT_K + 1)
#define HT_NT 256
#define HT_HR 64            // rows per helper workgroup (4 parts of 8 columns per row)

// rotation list of one sweep: for position c (index c - j - 2): rc, rs (row rotation: x = row c,
// y = row c-1, x' = rc x - rs y, y' = rc y + rs x), gc, gs (column rotation: x = col c, y = col c-1,
// x' = gc x + gs y, y' = gc y - gs x)
struct HtRot { double rc, rs, gc, gs; };

// block geometry of sweep j
struct HtSweep {
    int j, ihi, nb;
    __host__ __device__ int chi(int b) const { return ihi - HT_K * b; }
    __host__ __device__ int r0(int b) const { const int r = ihi - HT_K * b - HT_K; return r > j + 1 ? r : j + 1; }
};

// broadcast of lane 0's double to the whole wave (two 32-bit readlanes)
__device__ __forceinline__ double bcast0(double x) {
    const unsigned long long u = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)(u & 0xffffffffull));
    const int hi = __builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
    return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
// lane l's double (l uniform across the wave)
__device__ __forceinline__ double rdlane(double x, int l) {
    const unsigned long long u = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)(u & 0xffffffffull), l);
    const int hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
// Givens rotation with d_lartgf's arithmetic and conventions (c >= 0; r = sign(f) sqrt(f^2 + g^2),
// f = 0 counting as positive) without branches: d = sqrt(f^2 + g^2) from d_sqrtf, 1/d from the
// hardware reciprocal and two Newton steps whose residual 1 - d r is exact (fma).  bad: operands whose
// squares leave the safe range; the caller then takes d_lartgf's scaled path.  d, rd are used only
// when f, g are both non-zero and not bad, where f^2 + g^2 lies in d_sqrtf's range.
__device__ __forceinline__ void ht_lartg(double f, double g, double &c, double &s, double &r, bool &bad) {
    const double m = fmax(fabs(f), fabs(g));
    bad = (g != 0.0) && (f != 0.0) && !(m > 1e-140 && m < 1e140);
    const double d = d_sqrtf(fma(f, f, g * g));
    double rd = __builtin_amdgcn_rcp(d);
    rd = fma(rd, fma(-d, rd, 1.0), rd);
    rd = fma(rd, fma(-d, rd, 1.0), rd);
    const bool gz = (g == 0.0), fz = (f == 0.0);
    c = gz ? 1.0 : (fz ? 0.0 : fabs(f) * rd);
    s = gz ? 0.0 : (fz ? (g > 0.0 ? 1.0 : -1.0) : (f > 0.0 ? g : -g) * rd);
    r = gz ? f : (fz ? fabs(g) : copysign(d, f));
}
__device__ __forceinline__ void ht_rot(double f, double g, double &c, double &s, double &r) {
    bool bad;
    ht_lartg(f, g, c, s, r, bad);
    if (bad) d_lartgf(f, g, c, s, r);
}
// LDS writes of one lane seen by the wave's other lanes in the next step: a wave's LDS operations are
// performed in order, so only the compiler must be kept from moving accesses across (no wait for
// outstanding global memory operations here)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
// progress counters: 64-bit, the sweep in the high half, so values only grow during a reduction.
// The polling loads are relaxed (an acquire load at device scope invalidates the L2 of the polling
// XCD every time, which the other workgroups there, the lead among them, would pay for); one
// acquire fence follows the load that meets the target.
__device__ __forceinline__ void ht_wait(unsigned long long *p, unsigned long long target) {
    while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
// the release store makes the posting wave's earlier stores visible first; every value published
// through a counter is written by the wave that posts it
__device__ __forceinline__ void ht_post(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Helper workgroup h: rows [j+1 + h HT_HR, ...) below c_hi(2) (rows of regions >= 2).  For a row of
// region g it runs acc over blocks 0 .. g-2 and leaves acc_(g-2) in accg; ctr[1+h] counts its passes.
// It waits only for blocks the lead publishes unconditionally, and the lead waits for a pass only
// after publishing the block that pass needs.
__device__ void ht_helper(const double *B, int ld, const HtSweep S, const double *Wg, const double *dGg,
                          double *accg, unsigned long long *ctr, unsigned long long tag) {
    __shared__ double red[4][HT_HR];
    const int h = blockIdx.x - 1, j = S.j, ihi = S.ihi;
    const int lim = ihi - 2 * HT_K;
    const int hr0 = j + 1 + h * HT_HR, hr1 = min(hr0 + HT_HR, lim);
    if (hr0 >= hr1) return;
    const int tid = threadIdx.x, rr = tid % HT_HR, part = tid / HT_HR;
    const int i = hr0 + rr;
    const bool act = i < hr1;
    const int gi = act ? (ihi - 1 - i) / HT_K : 0;
    double acc = (act && part == 0) ? B[GQZ_IX(i, ihi, ld)] : 0.0;
    const int maxpass = (ihi - 1 - hr0) / HT_K - 2;
    for (int bp = 0; bp <= maxpass; ++bp) {
        if (tid == 0) ht_wait(&ctr[0], tag | (unsigned long long)(bp + 1));
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const int r0 = S.r0(bp), w = S.chi(bp) - r0 + 1;
        const bool use = act && bp <= gi - 2;
        double ps = 0.0;
        if (use) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int t = part * 8 + u;
                if (t < w - 1) ps = fma(B[GQZ_IX(i, r0 + t, ld)], Wg[bp * HT_R + t], ps);
            }
        }
        red[part][rr] = ps;
        __syncthreads();
        if (part == 0 && use) {                   // wave 0, which also posts
            acc = fma(dGg[bp], acc, (red[0][rr] + red[1][rr]) + (red[2][rr] + red[3][rr]));
            if (bp == gi - 2) accg[i] = acc;
        }
        __syncthreads();
        if (tid == 0) ht_post(&ctr[1 + h], tag | (unsigned long long)(bp + 1));
    }
}

// R chain of one block on one wave (every lane computes it, uniform operands): for q = w-1 .. 1,
// R at position r0 + q zeroes A(r0+q, j) against A(r0+q-1, j) (colA, the block's untouched column
// entries in LDS) and the carried entry; carry ends as the new A(r0, j)
__device__ __forceinline__ void ht_rchain(const double *colA, int w, double &carry, double *sRc, double *sRs, int lane) {
    double f = colA[w - 2];
    for (int q = w - 1; q >= 1; --q) {
        const double fn = (q >= 2) ? colA[q - 2] : 0.0;
        double c, s, r;
        ht_rot(f, carry, c, s, r);
        carry = r;
        if (lane == 0) { sRc[q] = c; sRs[q] = s; }
        f = fn;
    }
}

extern "C" __global__ void __launch_bounds__(HT_NT) k_ht_lead(double *A, double *B, int ld, HtSweep S, HtRot *rot,
                                                              double *Wg, double *dRg, double *dGg, double *Sbuf,
                                                              double *accg, unsigned long long *ctr,
                                                              unsigned long long tag) {
    if (blockIdx.x > 0) { ht_helper(B, ld, S, Wg, dGg, accg, ctr, tag); return; }
    __shared__ double sReg[2][HT_R * HT_R];  // regions b, b+1 (ld HT_R)
    __shared__ double sAc[2][HT_R];           // A(r0 .. c_hi, j) of regions b, b+1 (untouched entries)
    __shared__ double sRc[2][HT_R], sRs[2][HT_R];   // R of regions b, b+1, by region column index
    __shared__ double sCar[2];                // new A(r0, j) after the R's of regions b, b+1
    __shared__ double sP[HT_K * HT_K];        // B(rows of region b+1, r0(b) .. c_hi(b)-1), ld HT_K
    __shared__ double sAcc[HT_K];             // acc_(b-1) of the rows of region b+1
    __shared__ double sW[HT_R];               // w_b (zero beyond w-1), s_b at index HT_K
    __shared__ HtRot srot[HT_R];              // block b's rotations, by region column index
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = S.j, ihi = S.ihi, nb = S.nb;
    {
        const int r0 = S.r0(0), w = ihi - r0 + 1;
        for (int idx = tid; idx < w * w; idx += HT_NT) {
            const int i = idx % w, k = idx / w;
            sReg[0][i + HT_R * k] = (i <= k) ? B[GQZ_IX(r0 + i, r0 + k, ld)] : 0.0;
        }
        for (int i = tid; i < w; i += HT_NT) sAc[0][i] = A[GQZ_IX(r0 + i, j, ld)];
        __syncthreads();
        if (wave == 0) {
            double carry = sAc[0][w - 1];
            ht_rchain(sAc[0], w, carry, sRc[0], sRs[0], lane);
            if (lane == 0) sCar[0] = carry;
        }
        __syncthreads();
    }
    double rcarry = sCar[0];                  // wave 1: the R chain's carry, one block ahead
    for (int b = 0; b < nb; ++b) {
        const int cur = b & 1, nxt = cur ^ 1;
        const int c_hi = S.chi(b), r0 = S.r0(b), w = c_hi - r0 + 1;
        double *sB = sReg[cur];
        if (wave == 0) {
            // The G chain.  Lane k keeps one carried entry: while k <= q-2 ("column lane") B(k, q),
            // the region's column q at row k; once k >= q+1 ("row lane") B(q, k), row q at column k.
            // At position c = r0 + q every lane forms G_c from the corner (uniform operands: R_c and
            // B(q-1, q-1) from LDS, B(q-1, q) and B(q, q) carried), and each lane applies the pair it
            // owns: a row lane R_c to (B(q-1, k), B(q, k)), a column lane G_c to (B(k, q), B(k, q-1)).
            // The corner's next B(q-2, q-1) is formed from lane q-2's operands read before that lane
            // updates them, so no cross-lane read waits on the chain; LDS holds the other operands
            // (prefetched one position ahead) and receives each entry once it is final for the chain.
            const int k = lane;
            const double *rcv = sRc[cur], *rsv = sRs[cur];
            double cv = (k <= w - 2) ? sB[k + HT_R * (w - 1)] : 0.0;
            double pf = (k <= w - 3) ? sB[k + HT_R * (w - 2)] : 0.0;
            double b11 = sB[(w - 1) * (HT_R + 1)];
            double b01 = rdlane(cv, w - 2);
            double rc = rcv[w - 1], rs = rsv[w - 1], b00 = sB[(w - 2) * (HT_R + 1)];
            double kcc = 0.0, kss = 0.0, kr2 = 0.0;                 // lane k: G and B(k, k) of position k
            for (int q = w - 1; q >= 1; --q) {
                double xq = 0.0, yq = 0.0, rcn = 0.0, rsn = 0.0, b00n = 0.0, pfn = pf;
                // the next position's operand of this lane: row q-2 (k >= q) or column q-2 (k <= q-3),
                // loaded first so that its latency hides behind this position's chain.  No store of
                // this position touches these entries except B(q-2, q) (lane q's), which lane q-2
                // finalises below and lane q then takes from its register.
                const bool pr = k >= q && k < w, pl = q >= 2 && (pr || k <= q - 3);
                if (q >= 2) {
                    xq = rdlane(cv, q - 2); yq = rdlane(pf, q - 2);
                    rcn = rcv[q - 1]; rsn = rsv[q - 1]; b00n = sB[(q - 2) * (HT_R + 1)];
                    if (pl && k != q) pfn = sB[pr ? (q - 2) + HT_R * k : k + HT_R * (q - 2)];
                }
                const double nb00 = rc * b00, nb10 = -rs * b00;
                const double nb01 = rc * b01 + rs * b11, nb11 = rc * b11 - rs * b01;
                double cc, ss, r2;
                ht_rot(nb11, nb10, cc, ss, r2);
                const double n01 = cc * nb01 + ss * nb00;           // B(q-1, q)
                b11 = cc * nb00 - ss * nb01;                        // B(q-1, q-1): the next corner's b11
                b01 = cc * yq - ss * xq;                            // B(q-2, q-1): the next corner's b01
                // the lane's pair: a row lane R_c on (B(q-1, k), B(q, k)), a column lane G_c on
                // (B(k, q), B(k, q-1)); branch-free, one store
                const bool rl = k >= q + 1 && k < w, gl = k <= q - 2;
                const double ca = rl ? rc : cc, sa = rl ? rs : ss;
                const double a = rl ? pf : cv, bb = rl ? cv : pf;
                const double u = ca * a + sa * bb, v = ca * bb - sa * a;
                if (rl || gl) sB[rl ? q + HT_R * k : k + HT_R * q] = rl ? v : u;   // row q / column q final
                cv = rl ? u : (gl ? v : (k == q ? n01 : cv));      // carried: row q-1 / column q-1
                if (k == q) { kcc = cc; kss = ss; kr2 = r2; }
                if (q >= 2) {
                    const double u2 = rdlane(u, q - 2);             // B(q-2, q) after G_c (lane q-2: column lane)
                    if (k == q && pl) pfn = u2;
                }
                pf = pfn;
                rc = rcn; rs = rsn; b00 = b00n;
                wave_sync();
            }
            if (k >= 1 && k < w) {
                sB[HT_R * k] = cv;                                  // row 0
                sB[k * (HT_R + 1)] = kr2;
                sB[k + HT_R * (k - 1)] = 0.0;
                HtRot t; t.rc = rcv[k]; t.rs = rsv[k]; t.gc = kcc; t.gs = kss;
                srot[k] = t;
            }
            wave_sync();
            {
                // w_b(t) = gc_{t+1} prod_{m <= t} (-gs_m) and s_b = prod(-gs), prod(rs) by a scan across
                // the lanes (lane t: position r0 + t + 1); then publish the block
                double gcv = 0.0, ngs = 1.0, rsk = 1.0;
                if (k < w - 1) {
                    const HtRot rt = srot[k + 1];
                    rot[r0 + k + 1 - j - 2] = rt;
                    gcv = rt.gc; ngs = -rt.gs; rsk = rt.rs;
                }
                double incl = ngs;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const double o = __shfl_up(incl, off);
                    if (k >= off) incl *= o;
                }
                double excl = __shfl_up(incl, 1);
                if (k == 0) excl = 1.0;
                const double wv = gcv * excl;
                if (k < w - 1) { sW[k] = wv; Wg[b * HT_R + k] = wv; }
                else if (k < HT_K) sW[k] = 0.0;
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) rsk *= __shfl_xor(rsk, off);
                const double pg = rdlane(incl, w - 2);
                if (lane == 0) {
                    sB[0] = b11;
                    sW[HT_K] = pg;
                    dGg[b] = pg;
                    dRg[b] = rsk;
                    ht_post(&ctr[0], tag | (unsigned long long)(b + 1));
                }
            }
        } else if (b + 1 < nb) {
            // region b+1 (rows/columns r0n .. r0): its entries are untouched so far, except the last
            // column, formed below; wave 1 also runs its R chain, waves 2-3 load the rest
            const int r0n = S.r0(b + 1), wn = r0 - r0n + 1;
            if (wave == 1) {
                for (int i = lane; i < wn - 1; i += 64) sAc[nxt][i] = A[GQZ_IX(r0n + i, j, ld)];
                if (lane == 0) sAc[nxt][wn - 1] = 0.0;
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                ht_rchain(sAc[nxt], wn, rcarry, sRc[nxt], sRs[nxt], lane);
                if (lane == 0) sCar[nxt] = rcarry;
            } else {
                const int t2 = tid - 128;
                double *sN = sReg[nxt];
                for (int idx = t2; idx < wn * (wn - 1); idx += HT_NT - 128) {
                    const int i = idx % wn, k = idx / wn;
                    sN[i + HT_R * k] = (i <= k) ? B[GQZ_IX(r0n + i, r0n + k, ld)] : 0.0;
                }
                for (int idx = t2; idx < (wn - 1) * (w - 1); idx += HT_NT - 128) {
                    const int i = idx % (wn - 1), t = idx / (wn - 1);
                    sP[i + HT_K * t] = B[GQZ_IX(r0n + i, r0 + t, ld)];
                }
                if (wave == 2) {
                    if (b == 0) {
                        for (int i = lane; i < wn - 1; i += 64) sAcc[i] = B[GQZ_IX(r0n + i, ihi, ld)];
                    } else {
                        // acc_(b-1) of these rows: helper pass b-1 done
                        if (lane == 0) {
                            const int h0 = (r0n - (j + 1)) / HT_HR, h1 = (r0 - 1 - (j + 1)) / HT_HR;
                            for (int h = h0; h <= h1; ++h) ht_wait(&ctr[1 + h], tag | (unsigned long long)b);
                        }
                        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                        for (int i = lane; i < wn - 1; i += 64) sAcc[i] = accg[r0n + i];
                    }
                }
            }
        }
        __syncthreads();
        // region b back to B; its last column above the diagonal to the side buffer (b >= 1)
        for (int idx = tid; idx < w * w; idx += HT_NT) {
            const int i = idx % w, k = idx / w;
            if (i > k + 1) continue;
            const double v = (i <= k) ? sB[i + HT_R * k] : 0.0;
            if (b > 0 && k == w - 1 && i < w - 1) Sbuf[r0 + i] = v;
            else B[GQZ_IX(r0 + i, r0 + k, ld)] = v;
        }
        for (int i = tid; i < w; i += HT_NT) A[GQZ_IX(r0 + i, j, ld)] = (i == 0) ? sCar[cur] : 0.0;
        if (b + 1 < nb) {
            const int r0n = S.r0(b + 1), wn = r0 - r0n + 1;
            double *sN = sReg[nxt];
            {
                // 8 threads per row, 4 terms each (sW is zero past w-2), reduced across the 8 lanes
                const int i = tid >> 3, p = tid & 7;
                double y = 0.0;
                if (i < wn - 1) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) y = fma(sP[i + HT_K * (4 * p + u)], sW[4 * p + u], y);
                }
                y += __shfl_xor(y, 1);
                y += __shfl_xor(y, 2);
                y += __shfl_xor(y, 4);
                if (i < wn - 1 && p == 0) sN[i + HT_R * (wn - 1)] = fma(sW[HT_K], sAcc[i], y);
            }
            if (tid == 0) sN[(wn - 1) + HT_R * (wn - 1)] = sB[0];
        }
        __syncthreads();
    }
}

// ---- far updates: items (vector, block) ----------------------------------------------------------

// one item: entries e[c * st] of a row or column, block b; the vector's last block bl stores its
// final carry at e[r0(bl)], or (fin) the value sval there instead
struct HtItem {
    double *e;
    long st;
    int b, bl, fin;
    double sval;
};
// one vector: its blocks bf .. bl; the item of block b is i0 + (b - bf) di, or for a column of B
// under the row rotations (k >= 0) found from the block layout
struct HtJob {
    double *e;
    long st;
    int bf, bl, k;
    long i0, di;
};

// Column-rotation items: [A rows 0..ihi][Z rows][B rows 0..j][B rows of region g, blocks 0..g-1],
// each part block-major (consecutive threads: consecutive rows, one block)
struct HtGeoG {
    double *A, *B, *Z, *Sbuf;
    int ld, ldz, nz;
    HtSweep S;
    __host__ __device__ long nA() const { return S.ihi + 1; }
    __host__ __device__ long seg0() const { return (long)S.nb * nA(); }
    __host__ __device__ long seg1() const { return (long)S.nb * nz; }
    __host__ __device__ long seg2() const { return (long)S.nb * (S.j + 1); }
    __host__ __device__ int mg(int g) const { return S.chi(g) - S.r0(g); }
    __host__ __device__ long base3(int g) const { return 16L * g * (g - 1); }
    __host__ __device__ long seg3() const { return S.nb > 1 ? base3(S.nb - 1) + (long)mg(S.nb - 1) * (S.nb - 1) : 0; }
    __host__ __device__ long items() const { return seg0() + seg1() + seg2() + seg3(); }
    // vectors: A rows, Z rows, B rows 0 .. r0(0)-1
    __host__ __device__ long jobs() const { return nA() + nz + S.r0(0); }
    __device__ bool item(long it, HtItem &o) const {
        const int j = S.j;
        o.fin = 0; o.sval = 0.0; o.bl = S.nb - 1; o.st = ld;
        if (it < seg0()) { o.e = A + it % nA(); o.b = (int)(it / nA()); return true; }
        it -= seg0();
        if (it < seg1()) { o.e = Z + it % nz; o.st = ldz; o.b = (int)(it / nz); return true; }
        it -= seg1();
        if (it < seg2()) { o.e = B + it % (j + 1); o.b = (int)(it / (j + 1)); return true; }
        it -= seg2();
        if (it >= seg3()) return false;
        int g = (int)(0.5 + sqrt(0.25 + (double)it / 16.0));
        if (g > S.nb - 1) g = S.nb - 1;
        if (g < 1) g = 1;
        while (g > 1 && base3(g) > it) --g;
        while (g < S.nb - 1 && base3(g + 1) <= it) ++g;
        const long off = it - base3(g);
        const int m = mg(g), i = S.r0(g) + (int)(off % m);
        o.e = B + i; o.b = (int)(off / m); o.bl = g - 1; o.fin = 1; o.sval = Sbuf[i];
        return true;
    }
    __device__ bool job(long q, HtJob &jb) const {
        const int j = S.j;
        jb.bf = 0; jb.bl = S.nb - 1; jb.k = -1; jb.st = ld;
        if (q < nA()) { jb.e = A + q; jb.i0 = q; jb.di = nA(); return true; }
        q -= nA();
        if (q < nz) { jb.e = Z + q; jb.st = ldz; jb.i0 = seg0() + q; jb.di = nz; return true; }
        q -= nz;
        if (q < j + 1) { jb.e = B + q; jb.i0 = seg0() + seg1() + q; jb.di = j + 1; return true; }
        const int i = (int)q;
        if (i >= S.r0(0)) return false;
        const int g = (S.ihi - 1 - i) / HT_K;
        jb.e = B + i; jb.bl = g - 1;
        jb.i0 = seg0() + seg1() + seg2() + base3(g) + (i - S.r0(g));
        jb.di = mg(g);
        return true;
    }
    __device__ long job_item(const HtJob &jb, int b) const { return jb.i0 + (long)(b - jb.bf) * jb.di; }
};

// Row-rotation items: [A cols j+1..n-1][Q rows][B cols right of block b's region, by block]
struct HtGeoR {
    double *A, *B, *Q;
    int ld, ldq, nq, n;
    HtSweep S;
    __host__ __device__ long nc() const { return n - S.j - 1; }
    __host__ __device__ long seg0() const { return (long)S.nb * nc(); }
    __host__ __device__ long seg1() const { return (long)S.nb * nq; }
    __host__ __device__ long E() const { return n - 1 - S.ihi; }
    __host__ __device__ long baseB(int b) const { return (long)b * E() + 16L * b * (b - 1); }
    __host__ __device__ long seg2() const { return baseB(S.nb); }
    __host__ __device__ long items() const { return seg0() + seg1() + seg2(); }
    // vectors: A cols, Q rows, B cols j+2 .. n-1
    __host__ __device__ long jobs() const { return nc() + nq + (n - S.j - 2 > 0 ? n - S.j - 2 : 0); }
    __device__ bool item(long it, HtItem &o) const {
        o.fin = 0; o.sval = 0.0; o.bl = S.nb - 1; o.st = 1;
        if (it < seg0()) {
            const int k = S.j + 1 + (int)(it % nc());
            o.e = A + (long)k * ld; o.b = (int)(it / nc());
            return true;
        }
        it -= seg0();
        if (it < seg1()) { o.e = Q + it % nq; o.st = ldq; o.b = (int)(it / nq); return true; }
        it -= seg1();
        if (it >= seg2()) return false;
        // largest b with baseB(b) = 16 b^2 + (E - 16) b <= it
        const double e16 = (double)(E() - 16);
        int b = (int)((-e16 + sqrt(e16 * e16 + 64.0 * (double)it)) / 32.0);
        if (b < 0) b = 0;
        if (b > S.nb - 1) b = S.nb - 1;
        while (b > 0 && baseB(b) > it) --b;
        while (b < S.nb - 1 && baseB(b + 1) <= it) ++b;
        const int k = S.chi(b) + 1 + (int)(it - baseB(b));
        o.e = B + (long)k * ld; o.b = b;
        return true;
    }
    __device__ bool job(long q, HtJob &jb) const {
        jb.bf = 0; jb.bl = S.nb - 1; jb.k = -1; jb.st = 1;
        if (q < nc()) { jb.e = A + (long)(S.j + 1 + q) * ld; jb.i0 = q; jb.di = nc(); return true; }
        q -= nc();
        if (q < nq) { jb.e = Q + q; jb.st = ldq; jb.i0 = seg0() + q; jb.di = nq; return true; }
        q -= nq;
        const int k = S.j + 2 + (int)q;
        const int bk = (k <= S.ihi) ? (S.ihi - k) / HT_K : -1;   // k in (r0(bk), c_hi(bk)]
        jb.bf = bk + 1;
        if (jb.bf > S.nb - 1) return false;
        jb.e = B + (long)k * ld; jb.k = k; jb.i0 = 0; jb.di = 0;
        return true;
    }
    __device__ long job_item(const HtJob &jb, int b) const {
        if (jb.k < 0) return jb.i0 + (long)(b - jb.bf) * jb.di;
        return seg0() + seg1() + baseB(b) + (jb.k - S.chi(b) - 1);
    }
};

template <bool GK>
__device__ __forceinline__ void ht_rot_carry(const HtRot &r, double x, double y, double &out, double &xn) {
    if (GK) { out = r.gc * x + r.gs * y; xn = r.gc * y - r.gs * x; }
    else { out = r.rc * x - r.rs * y; xn = r.rc * y + r.rs * x; }
}

// pass 1: gamma (carry out of the block for a zero carry in), and the block's entry r0, kept because
// the block above overwrites it in pass 3
template <bool GK, class Geo>
__global__ void __launch_bounds__(256) k_ht_far1(Geo geo, const HtRot *rot, double *gam, double *stash) {
    const long it = (long)blockIdx.x * blockDim.x + threadIdx.x;
    HtItem o;
    if (!geo.item(it, o)) return;
    const HtSweep S = geo.S;
    const int r0 = S.r0(o.b), w = S.chi(o.b) - r0 + 1;
    double y[HT_K];
#pragma unroll
    for (int t = 0; t < HT_K; ++t) y[t] = (t < w - 1) ? o.e[(long)(r0 + t) * o.st] : 0.0;
    double x = 0.0;
#pragma unroll
    for (int t = HT_K - 1; t >= 0; --t) {
        if (t >= w - 1) continue;
        double out;
        ht_rot_carry<GK>(rot[r0 + t + 1 - S.j - 2], x, y[t], out, x);
    }
    gam[it] = x;
    stash[it] = y[0];
}

// pass 2: the carry into each block of a vector, starting from the vector's entry c_hi(bf)
template <bool GK, class Geo>
__global__ void __launch_bounds__(256) k_ht_far2(Geo geo, const double *dg, const double *gam, double *xin) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= geo.jobs()) return;
    HtJob jb;
    if (!geo.job(q, jb)) return;
    double x = jb.e[(long)geo.S.chi(jb.bf) * jb.st];
    for (int b = jb.bf; b <= jb.bl; ++b) {
        const long it = geo.job_item(jb, b);
        xin[it] = x;
        x = fma(dg[b], x, gam[it]);
    }
}

// pass 3: the block with its carry: outputs e[c_lo .. c_hi], and at the vector's last block e[r0]
template <bool GK, class Geo>
__global__ void __launch_bounds__(256) k_ht_far3(Geo geo, const HtRot *rot, const double *xin, const double *stash) {
    const long it = (long)blockIdx.x * blockDim.x + threadIdx.x;
    HtItem o;
    if (!geo.item(it, o)) return;
    const HtSweep S = geo.S;
    const int r0 = S.r0(o.b), w = S.chi(o.b) - r0 + 1;
    double y[HT_K];
    y[0] = stash[it];
#pragma unroll
    for (int t = 1; t < HT_K; ++t) y[t] = (t < w - 1) ? o.e[(long)(r0 + t) * o.st] : 0.0;
    double x = xin[it];
#pragma unroll
    for (int t = HT_K - 1; t >= 0; --t) {
        if (t >= w - 1) continue;
        double out;
        ht_rot_carry<GK>(rot[r0 + t + 1 - S.j - 2], x, y[t], out, x);
        o.e[(long)(r0 + t + 1) * o.st] = out;
    }
    if (o.b == o.bl) o.e[(long)r0 * o.st] = o.fin ? o.sval : x;
}

// ---- host ----------------------------------------------------------------------------------------

namespace {
struct HtWs {
    HtRot *rot;
    double *W, *dR, *dG, *Sbuf, *acc, *gam, *stash, *xin;
    unsigned long long *ctr;
};
int ht_nbmax(int n) { return (n + HT_K - 1) / HT_K + 1; }
long ht_items_max(int n, int nq) {
    const long nb = ht_nbmax(n);
    return nb * (3L * n + nq + 2) + 16L * nb * nb;
}
int ht_nhmax(int n) { return n / HT_HR + 2; }
size_t ht_align(size_t x) { return (x + 255) & ~(size_t)255; }
HtWs ht_carve(void *ws, int n, int nq, size_t *total) {
    const size_t nb = ht_nbmax(n), items = ht_items_max(n, nq);
    char *p = (char *)ws;
    size_t off = 0;
    HtWs w;
    auto take = [&](size_t bytes) { char *q = p ? p + off : nullptr; off += ht_align(bytes); return q; };
    w.rot = (HtRot *)take(sizeof(HtRot) * (size_t)(n + 2));
    w.W = (double *)take(sizeof(double) * nb * HT_R);
    w.dR = (double *)take(sizeof(double) * nb);
    w.dG = (double *)take(sizeof(double) * nb);
    w.Sbuf = (double *)take(sizeof(double) * (size_t)(n + 1));
    w.acc = (double *)take(sizeof(double) * (size_t)(n + 1));
    w.ctr = (unsigned long long *)take(sizeof(unsigned long long) * (size_t)(1 + ht_nhmax(n)));
    w.gam = (double *)take(sizeof(double) * items);
    w.stash = (double *)take(sizeof(double) * items);
    w.xin = (double *)take(sizeof(double) * items);
    if (total) *total = off;
    return w;
}
// workgroups of k_ht_lead that can be resident at once (the cooperative launch needs all of them)
int ht_coop_capacity() {
    static int cap = -1;
    if (cap < 0) {
        int dev = 0, ncu = 0, per = 0;
        GQZ_HIP(hipGetDevice(&dev));
        GQZ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        GQZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)k_ht_lead, HT_NT, 0));
        cap = ncu * per;
    }
    return cap;
}
unsigned ht_grid(long n) { return (unsigned)((n + 255) / 256); }
}  // namespace

// bytes of the workspace of a reduction of order n with nq rows of Q and Z
size_t gqz_gghrd_blk_ws_bytes(int n, int nq) {
    size_t total = 0;
    ht_carve(nullptr, n, nq > n ? nq : n, &total);
    return total;
}

void gqz_launch_gghrd_blk(hipStream_t st, double *A, double *B, int ld, int n, int ilo, int ihi, double *Q,
                          double *Z, int ldq, int nq, void *ws, int jlast) {
    if (ihi - ilo < 2) return;
    const HtWs w = ht_carve(ws, n, nq > n ? nq : n, nullptr);
    GQZ_HIP(hipMemsetAsync(w.ctr, 0, sizeof(unsigned long long) * (size_t)(1 + ht_nhmax(n)), st));
    const int cap = ht_coop_capacity();
    const int jend = (jlast >= 0 && jlast < ihi - 2) ? jlast : ihi - 2;
    for (int j = ilo; j <= jend; ++j) {
        HtSweep S;
        S.j = j; S.ihi = ihi; S.nb = (ihi - j - 1 + HT_K - 1) / HT_K;
        unsigned long long tag = (unsigned long long)(j - ilo + 1) << 32;
        const int lim = ihi - 2 * HT_K;
        const int nh = lim > j + 1 ? (lim - (j + 1) + HT_HR - 1) / HT_HR : 0;
        if (1 + nh > cap) throw std::runtime_error("gghrd_blk: helper grid exceeds the co-resident capacity");
        HtRot *rot = w.rot;
        double *W = w.W, *dR = w.dR, *dG = w.dG, *Sbuf = w.Sbuf, *acc = w.acc;
        unsigned long long *ctr = w.ctr;
        void *args[] = { &A, &B, &ld, &S, &rot, &W, &dR, &dG, &Sbuf, &acc, &ctr, &tag };
        GQZ_HIP(hipLaunchCooperativeKernel((const void *)k_ht_lead, dim3(1 + nh), dim3(HT_NT), args, 0, st));
        HtGeoG gg;
        gg.A = A; gg.B = B; gg.Z = Z; gg.Sbuf = w.Sbuf; gg.ld = ld; gg.ldz = ldq; gg.nz = Z ? nq : 0; gg.S = S;
        const long ig = gg.items(), jg = gg.jobs();
        hipLaunchKernelGGL((k_ht_far1<true, HtGeoG>), dim3(ht_grid(ig)), dim3(256), 0, st, gg, w.rot, w.gam, w.stash);
        hipLaunchKernelGGL((k_ht_far2<true, HtGeoG>), dim3(ht_grid(jg)), dim3(256), 0, st, gg, w.dG, w.gam, w.xin);
        hipLaunchKernelGGL((k_ht_far3<true, HtGeoG>), dim3(ht_grid(ig)), dim3(256), 0, st, gg, w.rot, w.xin, w.stash);
        HtGeoR gr;
        gr.A = A; gr.B = B; gr.Q = Q; gr.ld = ld; gr.ldq = ldq; gr.nq = Q ? nq : 0; gr.n = n; gr.S = S;
        const long ir = gr.items(), jr = gr.jobs();
        hipLaunchKernelGGL((k_ht_far1<false, HtGeoR>), dim3(ht_grid(ir)), dim3(256), 0, st, gr, w.rot, w.gam, w.stash);
        hipLaunchKernelGGL((k_ht_far2<false, HtGeoR>), dim3(ht_grid(jr)), dim3(256), 0, st, gr, w.dR, w.gam, w.xin);
        hipLaunchKernelGGL((k_ht_far3<false, HtGeoR>), dim3(ht_grid(ir)), dim3(256), 0, st, gr, w.rot, w.xin, w.stash);
    }
    GQZ_HIP(hipGetLastError());
}
