// FP64 guard band of the mixed-precision forms: which K1 blocks to recompute
// in FP64 so that every sync-threshold decision is the FP64 one.
//
// The reference computes demod_05 in FP64 (fm_demod.cu:445-486) and decides
// x <= pulse_threshold_hz on it (vsync_discover.cu:17-20, sync_pulses.cu:
// 33-41, hsync_refine.cu:29-67).  The mixed forms compute K1 part A on the
// matrix cores and part B in FP32; a decision can only differ where the
// value lies within the form's error of the threshold, or where the phase is
// ill-conditioned (the analytic signal's envelope far below its neighbours',
// inside a dropout or at the zero-filled end of a chunk).  Part B flags each
// 256 samples holding an output within `gdelta` of the threshold; this pass
// lists the FP64 design's part-B blocks that hold a flagged sample or whose
// input support sees an envelope tile below rho times the largest envelope
// there, and the part-A blocks those read.  The FP64 K1 kernels then
// recompute exactly those blocks (block-list mode), overwriting demod,
// demod_05 and the threshold bits of each listed part-B block.
#include <hip/hip_runtime.h>

#include "common.h"
#include "launch.h"

namespace cuvhs {

// envelope tile of sample r of a segment (MFMA part A: 16 tiles of 1024
// outputs per block of mvalid samples; output o = r - b mvalid + 1)
__device__ __forceinline__ int env_tile(int r, int mvalid) {
    const int b = r / mvalid;
    return b * 16 + (r - b * mvalid + 1) / 1024;
}

// range [lo, hi] of the samples a window reads, as segment-relative sample
// pieces: the window q in [q_lo, q_lo + B) of segment f reads
// seg0 + (q mod N), clamped to [0, total - 1] (fm_demod.cu:144-152)
struct SampleSpan {
    long long lo[2], hi[2];
    int n;
};
__device__ __forceinline__ SampleSpan window_span(int f, int spfp, int N, int q_lo, int B, long long total) {
    SampleSpan sp;
    const long long seg0 = (long long)f * spfp - kOverlap;
    const int q_hi = q_lo + B - 1;
    auto clampv = [&](long long s) { return s < 0 ? 0 : (s > total - 1 ? total - 1 : s); };
    if (q_lo < 0) {
        sp.lo[0] = clampv(seg0 + q_lo + N);
        sp.hi[0] = clampv(seg0 + N - 1);
        sp.lo[1] = clampv(seg0);
        sp.hi[1] = clampv(seg0 + q_hi);
        sp.n = 2;
    } else if (q_hi >= N) {
        sp.lo[0] = clampv(seg0 + q_lo);
        sp.hi[0] = clampv(seg0 + N - 1);
        sp.lo[1] = clampv(seg0);
        sp.hi[1] = clampv(seg0 + q_hi - N);
        sp.n = 2;
    } else {
        sp.lo[0] = clampv(seg0 + q_lo);
        sp.hi[0] = clampv(seg0 + q_hi);
        sp.n = 1;
    }
    return sp;
}

// The chunk's values from a replayed graph's plan (GuardLaunch::plan).
__device__ __forceinline__ void plan_guard(GuardLaunch& g) {
    g.total = g.plan->total;
    g.nseg = g.plan->fields_k1;
    g.nsub = ((g.total - 1) >> kGuardSubShift) + 1;
}

__global__ __launch_bounds__(256) void k_guard_mark(GuardLaunch g) {
    if (g.plan) plan_guard(g);
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx == 0) {
        g.counts[0] = 0;
        g.counts[1] = 0;
    }
    if (idx >= (long long)g.nseg * g.nbB) return;
    const int f = (int)(idx / g.nbB), b = (int)(idx % g.nbB);
    const int r0 = b * g.VB;
    if (r0 >= g.spfp) return;
    const int r1 = min(r0 + g.VB, g.spfp);
    bool flag = false;
    const long long s0 = (long long)f * g.spfp + r0, s1 = (long long)f * g.spfp + r1;
    for (long long k = s0 >> kGuardSubShift; k <= (s1 - 1) >> kGuardSubShift && !flag; k++) flag = g.gflag[k] != 0;
    // the FP64 block's input window (its rawdemod reads)
    const SampleSpan sp = window_span(f, g.spfp, g.N, kOverlap + r0 - g.lpB, g.BB, g.total);
    if (!flag && g.blk_env) {
        float mn = 3.0e38f, mx = 0.0f;
        for (int p = 0; p < sp.n; p++) {
            // rawdemod at s uses the analytic signal at s and s - 1
            const long long a = sp.lo[p] > 0 ? sp.lo[p] - 1 : 0;
            for (long long s = a; s <= sp.hi[p];) {
                const int seg = (int)(s / g.spfp);
                const int r = (int)(s - (long long)seg * g.spfp);
                const int tl = env_tile(r, g.mvalid);
                const float2 e = g.blk_env[(size_t)seg * g.ntile + tl];
                mn = fminf(mn, e.x);
                mx = fmaxf(mx, e.y);
                // the next tile's first sample: the tile's end, or the MFMA
                // block's, or the segment's, whichever comes first
                const int bm = r / g.mvalid;
                const int o = r - bm * g.mvalid + 1;
                int r_next = bm * g.mvalid + (o / 1024 + 1) * 1024 - 1;
                r_next = min(r_next, (bm + 1) * g.mvalid);
                r_next = min(r_next, g.spfp);
                s = (long long)seg * g.spfp + (r_next > r ? r_next : r + 1);
            }
        }
        flag = mn < g.rho * mx;
    }
    if (!flag) return;
    g.markB[(size_t)f * g.nbB + b] = 1;
    // every part-A block that writes a sample of the window: walk the span
    // block by block (a segment's last block is shorter than VA, so the walk
    // stops at the segment's end before moving on)
    for (int p = 0; p < sp.n; p++) {
        for (long long s = sp.lo[p]; s <= sp.hi[p];) {
            const int seg = (int)(s / g.spfp);
            const int r = (int)(s - (long long)seg * g.spfp);
            const int ba = r / g.VA;
            g.markA[(size_t)seg * g.nbA + ba] = 1;
            s = (long long)seg * g.spfp + min((ba + 1) * g.VA, g.spfp);
        }
    }
}

__global__ __launch_bounds__(256) void k_guard_compact(GuardLaunch g) {
    if (g.plan) plan_guard(g);
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    if (idx < g.nsub) g.gflag[idx] = 0;   // read by the marking pass; clear for the next chunk
    if (idx < (long long)g.nseg * g.nbA && g.markA[idx]) {
        g.markA[idx] = 0;
        const int pos = atomicAdd(&g.counts[0], 1);
        g.listA[pos] = make_int2((int)(idx / g.nbA), (int)(idx % g.nbA));
    }
    if (idx < (long long)g.nseg * g.nbB && g.markB[idx]) {
        g.markB[idx] = 0;
        const int pos = atomicAdd(&g.counts[1], 1);
        g.listB[pos] = make_int2((int)(idx / g.nbB), (int)(idx % g.nbB));
    }
}

// the counts into the chunk's control block (for reporting)
__global__ void k_guard_report(const int* counts, ChunkCtl* ctl) {
    ctl->guard_a = counts[0];
    ctl->guard_b = counts[1];
}

hipError_t run_guard_lists(const GuardLaunch& g, hipStream_t st) {
    const long long nb = (long long)g.nseg * g.nbB;
    const long long na = (long long)g.nseg * g.nbA;
    long long n = na > nb ? na : nb;
    if (g.nsub > n) n = g.nsub;
    hipLaunchKernelGGL(k_guard_mark, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, g);
    hipLaunchKernelGGL(k_guard_compact, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g);
    if (g.ctl) hipLaunchKernelGGL(k_guard_report, dim3(1), dim3(1), 0, st, g.counts, g.ctl);
    return hipGetLastError();
}

}  // namespace cuvhs
