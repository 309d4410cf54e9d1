// Sync threshold searches: run boundaries of demod_05 <= threshold, VSYNC
// discovery, chunk scheduling (VSYNC clustering) and K2 pulse lists.
//
// Reference: vsync_discover.cu:5-72 (one thread per sample, serial width
// count, atomic append), pipeline.cu:1567-1651 (host sort + clustering),
// sync_pulses.cu:10-82 (one thread per field scanning ~5e5 samples).
//
// Here the signal is scanned once: every workgroup turns 8192 samples into
// ordered run-boundary events (start = first sample <= thr, end = first
// sample > thr after a run) with wavefront ballots and a workgroup scan, and
// all three consumers work on that short, sorted event list.  Both the
// reference's serial loops are functions of these boundaries alone:
//   VSYNC width = min(end - start, max_w + 1), end = chunk end if no end;
//   K2 keeps runs with off < start and end < off + spfp, length end - start.
#include <hip/hip_runtime.h>

#include "common.h"
#include "launch.h"

namespace cuvhs {

constexpr int kEvThreads = 256;
constexpr int kEvPerThread = 32;
constexpr int kEvSpan = kEvThreads * kEvPerThread;  // samples per workgroup

// events of sample i: bit0 = start, bit1 = end
__device__ __forceinline__ int event_at(const double* x, long long i, double thr) {
    bool b = x[i] <= thr;
    if (i == 0) return b ? 1 : 0;
    bool p = x[i - 1] <= thr;
    if (b && !p) return 1;
    if (!b && p) return 2;
    return 0;
}

// Sample i of workgroup g, iteration k, thread t: g*kEvSpan + k*256 + t, so
// each load instruction of a wavefront reads 64 consecutive doubles.  Each
// thread's start/end events are kept as two 32-bit masks (bit k: iteration
// k) for k_ev_write, which so never reads the signal again.
__global__ __launch_bounds__(kEvThreads) void k_ev_count(const double* x, long long total, double thr, int* wg_count,
                                                         uint2* masks) {
    __shared__ int s_total;
    if (threadIdx.x == 0) s_total = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kEvSpan + threadIdx.x;
    unsigned int fs = 0, fe = 0;
#pragma unroll 8
    for (int k = 0; k < kEvPerThread; k++) {
        long long i = base + (long long)k * kEvThreads;
        int e = i < total ? event_at(x, i, thr) : 0;
        if (e == 1) fs |= 1u << k;
        if (e == 2) fe |= 1u << k;
    }
    masks[(size_t)blockIdx.x * kEvThreads + threadIdx.x] = make_uint2(fs, fe);
    int c = __popc(fs | fe);
    // wavefront sum, then one LDS atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(&s_total, c);
    __syncthreads();
    if (threadIdx.x == 0) wg_count[blockIdx.x] = s_total;
}

// single workgroup: exclusive scan of per-workgroup counts
__global__ __launch_bounds__(1024) void k_ev_scan(const int* wg_count, int nwg, int* wg_off, ChunkCtl* ctl) {
    __shared__ int sh[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nwg; base += 1024) {
        int i = base + threadIdx.x;
        int v = i < nwg ? wg_count[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int add = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < nwg) wg_off[i] = carry + sh[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += sh[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) ctl->nevents = carry;
}

// events encoded as 2*pos + is_end, ascending.  Order inside a workgroup is
// (iteration k, wavefront w, lane) = ascending sample index; the offsets come
// from per-(k, w) ballot counts.
__global__ __launch_bounds__(kEvThreads) void k_ev_write(const uint2* masks, const int* wg_off, int* events) {
    constexpr int kWaves = kEvThreads / 64;
    __shared__ int s_cnt[kEvPerThread][kWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long base = (long long)blockIdx.x * kEvSpan + threadIdx.x;
    const uint2 msk = masks[(size_t)blockIdx.x * kEvThreads + threadIdx.x];
    const unsigned int fs = msk.x, fe = msk.y;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int k = 0; k < kEvPerThread; k++) {
        unsigned long long m = __ballot(((fs | fe) >> k) & 1u);
        if (lane == 0) s_cnt[k][wave] = __popcll(m);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = wg_off[blockIdx.x];
        for (int k = 0; k < kEvPerThread; k++)
            for (int w = 0; w < kWaves; w++) { int c = s_cnt[k][w]; s_cnt[k][w] = acc; acc += c; }
    }
    __syncthreads();
    for (int k = 0; k < kEvPerThread; k++) {
        unsigned int bit = ((fs | fe) >> k) & 1u;
        unsigned long long m = __ballot(bit);
        if (bit) {
            long long i = base + (long long)k * kEvThreads;
            int pos = s_cnt[k][wave] + __popcll(m & lt);
            events[pos] = (int)(2 * i + ((fe >> k) & 1u));
        }
    }
}

// VSYNC discovery over the event list + chunk clustering (one workgroup).
// offsets[] and ctl->{nf, next_raw_offset, vsync_count, status} are written.
// The candidate compaction takes 4096 events per step, four consecutive ones
// per thread (events alternate start/end, so at most two candidates each):
// two ballots give each candidate its rank in the wavefront, the 16
// wavefront counts are exchanged through LDS (double-buffered, one barrier
// per step), and the next step's events are loaded before the barrier.  The
// first kCandLds candidates are also kept in LDS for the clustering loop.
constexpr int kCandLds = 4096;
__global__ __launch_bounds__(1024) void k_vsync_cluster(const int* events, ChunkCtl* ctl, long long total,
                                                        int min_w, int max_w, int cap, int* cand,
                                                        unsigned long long* offsets, int num_fields, int spf,
                                                        int spl, int fields_k1, int do_cluster) {
    constexpr int kWaves = 1024 / 64;
    __shared__ int wcnt[2][kWaves];
    __shared__ int lcand[kCandLds];
    const int nev = ctl->nevents;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int carry_all = 0;
    // event k of the step, or 0 past the end; e[4] is the next group's first
    auto load = [&](int k0, int* e) {
#pragma unroll
        for (int q = 0; q < 5; q++) e[q] = k0 + q < nev ? events[k0 + q] : 0;
    };
    int k0 = 4 * (int)threadIdx.x;
    int e[5], nx[5];
    load(k0, e);
    for (int base = 0, it = 0; base < nev; base += 4096, it ^= 1) {
        int pos[2] = {0, 0};
        int c = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = k0 + q;
            if (k < nev && (e[q] & 1) == 0) {
                const int p = e[q] >> 1;
                long long end = total;
                if (k + 1 < nev) end = e[q + 1] >> 1;  // next event is this run's end
                long long w = end - p;
                if (w > max_w + 1) w = max_w + 1;
                if (w >= min_w && w <= max_w) pos[c++ & 1] = p;
            }
        }
        load(k0 + 4096, nx);
        const unsigned long long m0 = __ballot(c >= 1), m1 = __ballot(c >= 2);
        if (lane == 0) wcnt[it][wave] = __popcll(m0) + __popcll(m1);
        __syncthreads();
        int woff = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            int v = wcnt[it][w];
            woff += w < wave ? v : 0;
            tot += v;
        }
        const int idx = carry_all + woff + __popcll(m0 & lt) + __popcll(m1 & lt);
        for (int q = 0; q < c; q++) {
            if (idx + q < cap) cand[idx + q] = pos[q];
            if (idx + q < kCandLds) lcand[idx + q] = pos[q];
        }
        carry_all += tot;
        k0 += 4096;
#pragma unroll
        for (int q = 0; q < 5; q++) e[q] = nx[q];
    }
    __syncthreads();
    if (threadIdx.x != 0 || !do_cluster) {
        if (threadIdx.x == 0) ctl->vsync_count = carry_all;
        return;
    }
    // pipeline.cu:1595-1651 (sequential, <= 20 * fields candidates)
    int count = carry_all;
    ctl->vsync_count = count;
    int status = 0;
    int stored = count < cap ? count : cap;
    if (count > cap) status |= 1;
    int n = 0;
    unsigned long long last = 0;
    for (int i = 0; i < stored; i++) {
        int pos = i < kCandLds ? lcand[i] : cand[i];
        if (n == 0 || ((unsigned long long)pos - last > (unsigned long long)(spf / 2))) {
            unsigned long long safe = (pos > 5 * spl) ? (unsigned long long)(pos - 5 * spl) : 0ull;
            if (n < num_fields) offsets[n] = safe;
            n++;
            last = safe;
        }
    }
    if (n == 0) {
        status |= 2;
        for (int i = 0; i < fields_k1; i++) offsets[i] = (unsigned long long)i * (unsigned long long)spf;
        n = fields_k1;
    }
    if (n > num_fields) n = num_fields;
    unsigned long long actual = (unsigned long long)spf;
    if (n >= 2) actual = (offsets[n - 1] - offsets[0]) / (unsigned long long)(n - 1);
    unsigned long long predicted = offsets[n - 1] + actual;
    unsigned long long margin = 10ull * (unsigned long long)spl;
    ctl->next_raw_offset = predicted > margin ? ctl->raw_offset + predicted - margin : ctl->raw_offset + predicted;
    ctl->nf = n;
    ctl->status |= status;
}

// K2: one wavefront per field.  Events alternate start/end, so after an
// optional leading end (a run already open at the window start, whose pulse
// the reference drops) pulse j is the event pair (k0 + 2j, k0 + 2j + 1), and
// the pulses the reference's serial scan keeps are the prefix of pairs whose
// end lies inside the window (end positions increase with j).
__global__ __launch_bounds__(64) void k_sync_pulses(const int* events, const ChunkCtl* ctl,
                                                    const unsigned long long* offsets, int spfp, int* starts,
                                                    int* lens, int* count, int nf_override) {
    const int f = blockIdx.x;
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    if (f >= nf) return;
    const int lane = threadIdx.x;
    const int nev = ctl->nevents;
    const long long off = (long long)offsets[f];
    const long long wend = off + spfp;
    // first event with position > off: 64-way search, the lanes probing
    // evenly spaced events of the remaining range [lo, hi]
    int lo = 0, hi = nev;
    while (hi > lo) {
        const int step = (hi - lo + 63) / 64;
        const int idx = lo + lane * step;
        const bool le = idx < hi && (long long)(events[idx] >> 1) <= off;
        const int t = __popcll(__ballot(le));  // the probes at or before off form a prefix
        if (t == 0) break;
        const int nlo = lo + (t - 1) * step + 1;
        hi = min(lo + t * step, hi);
        lo = nlo;
    }
    int k0 = lo;
    if (k0 < nev && (events[k0] & 1)) k0++;
    int* st = starts + (size_t)f * kMaxPulses;
    int* ln = lens + (size_t)f * kMaxPulses;
    int c = 0;
    for (int base = 0; base < kMaxPulses; base += 64) {
        const int j = base + lane;
        const long long ke = (long long)k0 + 2LL * j + 1;
        bool ok = false;
        int ps = 0, pe = 0;
        if (ke < nev) {
            ps = events[ke - 1] >> 1;
            pe = events[ke] >> 1;
            ok = (long long)pe < wend;
        }
        const unsigned long long inv = ~__ballot(ok);
        const int run = inv ? __builtin_ctzll(inv) : 64;
        if (lane < run && j < kMaxPulses) {
            st[j] = ps;
            ln[j] = pe - ps;
        }
        c = base + run;
        if (run < 64) break;
    }
    if (lane == 0) count[f] = c < kMaxPulses ? c : kMaxPulses;
}

// ---- host launchers ----
int events_nwg(long long total) { return (int)((total + kEvSpan - 1) / kEvSpan); }

size_t events_mask_count(long long total) { return (size_t)events_nwg(total) * kEvThreads; }

hipError_t run_events(const double* x, long long total, double thr, int* wg_count, int* wg_off, uint2* masks,
                      int* events, ChunkCtl* ctl, hipStream_t st) {
    int nwg = events_nwg(total);
    if (nwg == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ev_count, dim3(nwg), dim3(kEvThreads), 0, st, x, total, thr, wg_count, masks);
    hipLaunchKernelGGL(k_ev_scan, dim3(1), dim3(1024), 0, st, wg_count, nwg, wg_off, ctl);
    hipLaunchKernelGGL(k_ev_write, dim3(nwg), dim3(kEvThreads), 0, st, masks, wg_off, events);
    return hipGetLastError();
}

hipError_t run_vsync_cluster(const int* events, ChunkCtl* ctl, long long total, int min_w, int max_w, int cap,
                             int* cand, unsigned long long* offsets, int num_fields, int spf, int spl, int fields_k1,
                             int do_cluster, hipStream_t st) {
    hipLaunchKernelGGL(k_vsync_cluster, dim3(1), dim3(1024), 0, st, events, ctl, total, min_w, max_w, cap, cand,
                       offsets, num_fields, spf, spl, fields_k1, do_cluster);
    return hipGetLastError();
}

hipError_t run_sync_pulses(const int* events, const ChunkCtl* ctl, const unsigned long long* offsets, int spfp,
                           int* starts, int* lens, int* count, int max_fields, int nf_override, hipStream_t st) {
    if (max_fields < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sync_pulses, dim3(max_fields), dim3(64), 0, st, events, ctl, offsets, spfp, starts, lens,
                       count, nf_override);
    return hipGetLastError();
}

}  // namespace cuvhs
