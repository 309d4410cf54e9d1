// Sync threshold searches: run boundaries of demod_05 <= threshold, VSYNC
// discovery, chunk scheduling (VSYNC clustering) and K2 pulse lists.
//
// Reference: vsync_discover.cu:5-72 (one thread per sample, serial width
// count, atomic append), pipeline.cu:1567-1651 (host sort + clustering),
// sync_pulses.cu:10-82 (one thread per field scanning ~5e5 samples).
//
// Here the signal is reduced once to a bitmask (bit s = demod_05[s] <= thr),
// written by K1 part B as it produces demod_05 (or by k_thr_bits from an
// array), and every workgroup turns 8192 bits into ordered run-boundary
// events (start = first sample <= thr, end = first sample > thr after a run)
// with a workgroup scan; all three consumers work on that short, sorted
// event list.  Both the reference's serial loops are functions of these
// boundaries alone:
//   VSYNC width = min(end - start, max_w + 1), end = chunk end if no end;
//   K2 keeps runs with off < start and end < off + spfp, length end - start.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "launch.h"

namespace cuvhs {

constexpr int kEvThreads = 256;
constexpr int kEvPerThread = 32;
constexpr int kEvSpan = kEvThreads * kEvPerThread;  // samples per workgroup

// Threshold bitmask of an array: one ballot per 64 samples (64-aligned, so
// each wavefront writes two whole words; the buffer holds thr_bits_words).
__global__ __launch_bounds__(256) void k_thr_bits(const double* x, long long total, double thr, unsigned int* bits) {
    const int lane = threadIdx.x & 63;
    const long long stride = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i - lane < total; i += stride) {
        const unsigned long long m = __ballot(i < total && x[i] <= thr);
        if (lane == 0) {
            bits[i >> 5] = (unsigned int)m;
            bits[(i >> 5) + 1] = (unsigned int)(m >> 32);
        }
    }
}

// Thread j owns samples 32 j .. 32 j + 31 (bitmask word j).  Start events:
// bit set, previous sample's clear (sample 0: set); end events: bit clear,
// previous set; none at or past total.  The (start, end) masks are kept for
// k_ev_write, which so never reads the bitmask again.
__global__ __launch_bounds__(kEvThreads) void k_ev_count(const unsigned int* bits, long long total, int* wg_count,
                                                         uint2* masks, const ChunkPlan* plan) {
    __shared__ int s_total;
    if (plan) total = plan->total;   // a replayed graph: sized for the batch
    if (threadIdx.x == 0) s_total = 0;
    __syncthreads();
    const long long j = (long long)blockIdx.x * kEvThreads + threadIdx.x;
    const long long s = j * 32;
    unsigned int fs = 0, fe = 0;
    if (s < total) {
        const unsigned int b = bits[j];
        const unsigned int p = (b << 1) | (j > 0 ? bits[j - 1] >> 31 : 0u);  // bit k: sample s + k - 1
        fs = b & ~p;
        fe = ~b & p;
        if (total - s < 32) {
            const unsigned int valid = (1u << (unsigned int)(total - s)) - 1u;
            fs &= valid;
            fe &= valid;
        }
    }
    masks[j] = make_uint2(fs, fe);
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

// events encoded as 2*pos + is_end, ascending: threads own consecutive
// samples, so a thread's slot is its workgroup's offset plus the exclusive
// scan of the per-thread counts.
__global__ __launch_bounds__(kEvThreads) void k_ev_write(const uint2* masks, const int* wg_off, int* events) {
    constexpr int kWaves = kEvThreads / 64;
    __shared__ int s_wave[kWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long j = (long long)blockIdx.x * kEvThreads + threadIdx.x;
    const uint2 msk = masks[j];
    unsigned int all = msk.x | msk.y;
    const int c = __popc(all);
    int x = c;  // inclusive scan over the wavefront
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    int pos = wg_off[blockIdx.x] + x - c;
    for (int w = 0; w < wave; w++) pos += s_wave[w];
    const long long s0 = j * 32;
    while (all) {
        const int k = __ffs(all) - 1;
        events[pos++] = (int)(2 * (s0 + k) + ((msk.y >> k) & 1u));
        all &= all - 1u;
    }
}

// VSYNC discovery over the event list + chunk clustering (one workgroup).
// offsets[] and ctl->{nf, next_raw_offset, vsync_count, status} are written.
// The candidate compaction takes 16384 events per step, sixteen consecutive
// ones per thread (events alternate start/end, so at most eight candidates
// each; the seventeenth event closes the last run): a wavefront scan of the
// per-thread counts and the 16 wavefront totals exchanged through LDS
// (double-buffered, one barrier per step) give each candidate its slot.
// The first kCandLds candidates are also kept in LDS for the clustering.
constexpr int kCandLds = 4096;
constexpr int kVsEvPerThread = 16;
constexpr int kVsStep = 1024 * kVsEvPerThread;

// pipeline.cu:1598-1610 on sorted candidates: candidate j is accepted iff
// it lies more than spf/2 past the last accepted one's safe offset (its
// position less 5 lines), the first one always.  Positions ascend, so the
// candidate accepted after i is the first j with pos[j] > safe(i) + spf/2
// (upper_bound; j > i since spf/2 > 5 lines), and the accepted set is the
// chain 0, next(0), next(next(0)), ...
__device__ __forceinline__ long long vs_safe(int pos, int spl) {
    return pos > 5 * spl ? (long long)pos - 5LL * spl : 0LL;
}

// host_out (optional): pinned host memory that receives the finished control
// block (visible to the host once the kernel has completed), in place of a
// separate device-to-host copy.
__global__ __launch_bounds__(1024) void k_vsync_cluster(const int* events, ChunkCtl* ctl, long long total,
                                                        int min_w, int max_w, int cap, int* cand,
                                                        unsigned long long* offsets, int num_fields, int spf,
                                                        int spl, int fields_k1, int do_cluster,
                                                        ChunkCtl* host_out, const ChunkPlan* plan) {
    constexpr int kWaves = 1024 / 64;
    if (plan) {   // a replayed graph: the chunk's values from its plan
        total = plan->total;
        cap = plan->vsync_cap;
        num_fields = plan->max_fields;
        fields_k1 = plan->fields_k1;
    }
    __shared__ int wcnt[2][kWaves];
    __shared__ int lcand[kCandLds];
    __shared__ int lnext[kCandLds];
    const int nev = ctl->nevents;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int carry_all = 0;
    for (int base = 0, it = 0; base < nev; base += kVsStep, it ^= 1) {
        const int k0 = base + kVsEvPerThread * (int)threadIdx.x;
        int e[kVsEvPerThread + 1];
        if (k0 + kVsEvPerThread < nev) {
            const int4* e4 = reinterpret_cast<const int4*>(events + k0);
#pragma unroll
            for (int q = 0; q < kVsEvPerThread / 4; q++) {
                const int4 x = e4[q];
                e[4 * q] = x.x;
                e[4 * q + 1] = x.y;
                e[4 * q + 2] = x.z;
                e[4 * q + 3] = x.w;
            }
            e[kVsEvPerThread] = events[k0 + kVsEvPerThread];
        } else {
#pragma unroll
            for (int q = 0; q <= kVsEvPerThread; q++) e[q] = k0 + q < nev ? events[k0 + q] : 0;
        }
        unsigned int cm = 0;  // bit q: event k0 + q starts a VSYNC-width run
#pragma unroll
        for (int q = 0; q < kVsEvPerThread; q++) {
            const int k = k0 + q;
            if (k < nev && (e[q] & 1) == 0) {
                const long long p = e[q] >> 1;
                long long end = total;
                if (k + 1 < nev) end = e[q + 1] >> 1;  // next event is this run's end
                long long w = end - p;
                if (w > max_w + 1) w = max_w + 1;
                if (w >= min_w && w <= max_w) cm |= 1u << q;
            }
        }
        const int c = __popc(cm);
        int x = c;  // inclusive scan over the wavefront
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wcnt[it][wave] = x;
        __syncthreads();
        int woff = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            const int v = wcnt[it][w];
            woff += w < wave ? v : 0;
            tot += v;
        }
        int idx = carry_all + woff + x - c;
#pragma unroll
        for (int q = 0; q < kVsEvPerThread; q++) {
            if ((cm >> q) & 1u) {
                const int p = e[q] >> 1;
                if (idx < cap) cand[idx] = p;
                if (idx < kCandLds) lcand[idx] = p;
                idx++;
            }
        }
        carry_all += tot;
    }
    __syncthreads();
    if (!do_cluster) {
        if (threadIdx.x == 0) {
            ctl->vsync_count = carry_all;
            if (host_out) *host_out = *ctl;
        }
        return;
    }
    const int count = carry_all;
    const int stored = count < cap ? count : cap;
    const bool in_lds = stored <= kCandLds;
    if (in_lds) {
        const long long half = spf / 2;
        for (int i = threadIdx.x; i < stored; i += 1024) {
            const long long thr = vs_safe(lcand[i], spl) + half;
            int lo = i + 1, hi = stored;  // first j in (i, stored) with lcand[j] > thr
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if ((long long)lcand[mid] > thr) hi = mid;
                else lo = mid + 1;
            }
            lnext[i] = lo;
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    ctl->vsync_count = count;
    int status = 0;
    if (count > cap) status |= 1;
    int n = 0;
    if (in_lds) {
        for (int i = 0; i < stored && n < num_fields; i = lnext[i]) offsets[n++] = (unsigned long long)vs_safe(lcand[i], spl);
    } else {
        unsigned long long last = 0;
        for (int i = 0; i < stored && n < num_fields; i++) {
            const int pos = i < kCandLds ? lcand[i] : cand[i];
            if (n == 0 || ((unsigned long long)pos - last > (unsigned long long)(spf / 2))) {
                last = (unsigned long long)vs_safe(pos, spl);
                offsets[n++] = last;
            }
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
    if (host_out) *host_out = *ctl;
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

size_t thr_bits_words(long long total) { return (size_t)((total + 63) / 64) * 2; }

// plan: total is the batch's (the grids' size); the kernels count the
// chunk's samples from the plan
hipError_t run_events_bits(const unsigned int* bits, long long total, int* wg_count, int* wg_off, uint2* masks,
                           int* events, ChunkCtl* ctl, hipStream_t st, const ChunkPlan* plan) {
    int nwg = events_nwg(total);
    if (nwg == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_ev_count, dim3(nwg), dim3(kEvThreads), 0, st, bits, total, wg_count, masks, plan);
    hipLaunchKernelGGL(k_ev_scan, dim3(1), dim3(1024), 0, st, wg_count, nwg, wg_off, ctl);
    hipLaunchKernelGGL(k_ev_write, dim3(nwg), dim3(kEvThreads), 0, st, masks, wg_off, events);
    return hipGetLastError();
}

hipError_t run_events(const double* x, long long total, double thr, unsigned int* bits, int* wg_count, int* wg_off,
                      uint2* masks, int* events, ChunkCtl* ctl, hipStream_t st) {
    if (total <= 0) return hipErrorInvalidValue;
    long long waves = (total + 63) / 64;
    int nb = (int)std::min<long long>((waves + 3) / 4, 4096);
    hipLaunchKernelGGL(k_thr_bits, dim3(nb), dim3(256), 0, st, x, total, thr, bits);
    return run_events_bits(bits, total, wg_count, wg_off, masks, events, ctl, st);
}

hipError_t run_vsync_cluster(const int* events, ChunkCtl* ctl, long long total, int min_w, int max_w, int cap,
                             int* cand, unsigned long long* offsets, int num_fields, int spf, int spl, int fields_k1,
                             int do_cluster, ChunkCtl* host_out, hipStream_t st, const ChunkPlan* plan) {
    hipLaunchKernelGGL(k_vsync_cluster, dim3(1), dim3(1024), 0, st, events, ctl, total, min_w, max_w, cap, cand,
                       offsets, num_fields, spf, spl, fields_k1, do_cluster, host_out, plan);
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
