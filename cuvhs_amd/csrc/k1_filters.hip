// K1 FM demodulation and the chroma capture-rate pre-band-pass as blocked
// overlap-save convolutions in LDS.
//
// Reference: fm_demod.cu:396-486 (whole-segment D2Z -> filter -> Z2Z/Z2D over
// N = next_7smooth(spfp + 2048) points per field segment) and
// chroma_decode.cu:594-633 (per-segment circular pre-band-pass over
// N2 = next_7smooth(spfp)).
//
// Equivalence: the reference computes, for segment f and window position q,
//   y_f(q) = sum_{m} g_N[m] * x_f((q - m) mod N)
// with x_f(p) = src[clamp(f*spfp - 2048 + p, 0, T-1)] (fm_demod.cu:133-153).
// Each workgroup here computes V consecutive outputs of one segment as a
// B-point circular convolution of B loaded inputs with the taps of g_N kept
// in [-lm, lp]; the loaded input indices wrap modulo N exactly like the
// reference's window, so the only difference is the dropped taps
// (|g| < eps * max) plus rounding.  The B-point convolution itself is one of
// the wave-local blocks of fft_lds.h (conv4096, conv_rows): a radix-16 stage
// across the workgroup, then per-wave sub-transforms, the spectrum product
// and the inverse, with two workgroup barriers per block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <set>
#include <utility>

#include "atan_f64.h"
#include "common.h"
#include "fft_lds.h"
#include "launch.h"

namespace cuvhs {

// RawReader::convert (raw_reader.cpp:71-91): u8 (s - 128) * 256, s16 as is,
// u16 s - 32768.
__device__ __forceinline__ double norm_sample(const void* cap, int fmt, long long s) {
    if (fmt == 0) return ((double)static_cast<const uint8_t*>(cap)[s] - 128.0) * 256.0;
    if (fmt == 1) return (double)static_cast<const int16_t*>(cap)[s];
    return (double)static_cast<const uint16_t*>(cap)[s] - 32768.0;
}

// Sample s of the chunk as the reference's chunk buffer holds it: normalised,
// minus the read's mean under --dc-correct (raw_reader.cpp:93-100), and zero
// past the read (pipeline.cu:1539-1541).
__device__ __forceinline__ double raw_sample(const void* cap, int fmt, long long s, long long n_valid, double dc) {
    return s < n_valid ? norm_sample(cap, fmt, s) - dc : 0.0;
}

constexpr int kMeanBlocks = 256;

// Per-read mean for --dc-correct: fixed slices, fixed-order tree sums (the
// reference sums sequentially on the host; the orders differ at rounding).
__global__ __launch_bounds__(256) void k_input_sum(const void* cap, int fmt, long long n, double* partial,
                                                   const ChunkPlan* plan) {
    __shared__ double red[256];
    if (plan) {
        cap = plan->chunk;
        n = plan->n_read;
    }
    const long long per = (n + kMeanBlocks - 1) / kMeanBlocks;
    const long long a = (long long)blockIdx.x * per;
    const long long b = a + per < n ? a + per : n;
    double s = 0.0;
    for (long long i = a + threadIdx.x; i < b; i += 256) s += norm_sample(cap, fmt, i);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_input_mean(const double* partial, long long n, double* mean,
                                                    const ChunkPlan* plan) {
    __shared__ double red[kMeanBlocks];
    if (plan) n = plan->n_read;
    red[threadIdx.x] = partial[threadIdx.x];
    __syncthreads();
    for (int o = kMeanBlocks / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *mean = red[0] / (double)n;
}

// q mod n for q in (-n, 2n): every window index of these kernels lies there
// (the block start is within one segment and a block is shorter than n).
__device__ __forceinline__ int wrap_n(int q, int n) { return q < 0 ? q + n : (q >= n ? q - n : q); }

// 4096-point blocks exchange real and imaginary parts through LDS
// separately: 35 KB instead of 70 KB per FP64 workgroup, 17 KB instead of 35
// KB per FP32 one (five FP32 workgroups per CU instead of four; the
// 600-field decode's FP32 form 73.5k against 72.4k fields/s).
template <int B, typename T>
constexpr bool k1_whole() { return B != 4096; }
template <int B, typename T>
constexpr size_t k1_lds_bytes() { return (size_t)lds_elems(B) * (k1_whole<B, T>() ? 2 * sizeof(T) : sizeof(T)); }

// Workgroups per CU the register allocation must allow: five FP32
// 4096-point blocks (17 KB of LDS each) at 102 VGPRs; FP64 4096-point blocks
// (35 KB) to 168 VGPRs, three per CU (held to 128 VGPRs for four per CU they
// spill and run slower: part B 3.03 against 2.79 ms per 600 fields).
template <int B, typename T>
constexpr int k1_min_blocks() { return B == 4096 ? (sizeof(T) == 4 ? 5 : 3) : 1; }

// Phase of z = x + iy in the transform's precision (atan_f64.h for FP64).
__device__ __forceinline__ double phase_of(double y, double x) { return atan2_f64(y, x); }
__device__ __forceinline__ float phase_of(float y, float x) { return atan2f(y, x); }

// Stores of K1 outputs in the storage type the launch selects.
__device__ __forceinline__ void store_sig(void* p, int f32, long long i, double v) {
    if (f32) static_cast<float*>(p)[i] = (float)v;
    else static_cast<double*>(p)[i] = v;
}
__device__ __forceinline__ double load_sig(const void* p, int f32, long long i) {
    return f32 ? (double)static_cast<const float*>(p)[i] : static_cast<const double*>(p)[i];
}

// Part A: analytic signal, envelope and instantaneous frequency of block b of
// segment f.  T is the transform precision: double (the reference's FP64
// arithmetic) or float (the FP32 form); outputs are stored as the launch
// selects (double or float).
template <int B, typename T>
__device__ __forceinline__ void k1_part_a_block(const K1ALaunch& a, int f, int b) {
    using V2 = typename C2<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    V2* buf = reinterpret_cast<V2*>(smem);
    const V2* G = static_cast<const V2*>(a.G);
    const V2* tw = static_cast<const V2*>(a.tw);
    constexpr int NT = B / 16;
    if (a.ctl && f == 0 && b == 0 && threadIdx.x == 0) {
        // the chunk's control block starts here: no host-to-device copy
        ChunkCtl z{};
        z.fields_k1 = a.ctl_fields_k1;
        z.raw_offset = a.ctl_raw_offset;
        *a.ctl = z;
    }
    const int qs = (kOverlap - 1) + b * a.V;  // first analytic output
    if (qs + 1 >= kOverlap + a.spfp) return;
    if (a.thr_bits) {
        // clear the threshold bitmask words of this block's samples
        // [f spfp + b V, f spfp + min((b + 1) V, spfp)) for part B's ORs
        // (words shared with a neighbouring block are cleared by both)
        const long long s0 = (long long)f * a.spfp + (long long)b * a.V;
        const long long s1 = (long long)f * a.spfp + min((long long)(b + 1) * a.V, (long long)a.spfp);
        for (long long w = (s0 >> 5) + threadIdx.x; w <= ((s1 - 1) >> 5); w += NT) a.thr_bits[w] = 0u;
    }
    const int seg0 = f * a.spfp - kOverlap;
    const int last = (int)a.total - 1;
    const double dcv = a.dc ? *a.dc : 0.0;
    T xin[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int s = seg0 + wrap_n(qs - a.lp + (int)threadIdx.x + i * NT, a.N);
        s = s < 0 ? 0 : (s > last ? last : s);
        xin[i] = (T)raw_sample(a.cap, a.in_fmt, s, a.n_valid, dcv);
    }
    V2 v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        v[i].x = xin[i];
        v[i].y = (T)0;
    }
    const int qend = kOverlap + a.spfp;  // one past the last demod output
    const T PI = (T)3.14159265358979323846;
    const T TAU = (T)6.28318530717958647692;
    static_assert(B == 4096 || B == 8192, "K1 blocks are wave-local 4096- or 8192-point blocks");
    {
        // G is in the wave-local block's bin order; v[r] becomes a(qs + o),
        // o = NT r + t - lp
        if constexpr (B == 4096) conv4096<T, k1_whole<B, T>()>(smem, v, tw, G);
        else conv_rows<T, 512>(buf, v, tw, G);
        constexpr int NW = NT / 64;
        __shared__ T edge[NW][16];
        // the thread index anew: indices derived from it before the
        // transform are not kept live across it
        int t = (int)threadIdx.x;
        asm volatile("" : "+v"(t));
        const int lane = t & 63, wv = t >> 6;
        T ang[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            // only outputs 0..V are used (the rest of the block is the
            // convolution's wrap-around)
            int o = r * NT + t - a.lp;
            ang[r] = (T)0;
            if (o >= 0 && o <= a.V) {
                ang[r] = phase_of(v[r].y, v[r].x);
                int q = qs + o;
                if (a.env && q >= kOverlap && q < qend)
                    store_sig(a.env, a.out_f32, (long long)f * a.spfp + (q - kOverlap),
                              (double)sqrt(v[r].x * v[r].x + v[r].y * v[r].y));
            }
        }
        // the phase at o - 1 is held by lane - 1 of this wave, or for lane 0
        // by lane 63 of the previous wave (of the previous row for t = 0)
        if (lane == 63) {
#pragma unroll
            for (int r = 0; r < 16; r++) edge[wv][r] = ang[r];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 16; r++) {
            T prev = __shfl_up(ang[r], 1);
            if (lane == 0) prev = wv > 0 ? edge[wv - 1][r] : (r > 0 ? edge[NW - 1][r - 1] : (T)0);
            int o = r * NT + t - a.lp;
            int q = qs + o;
            if (o >= 1 && o <= a.V && q < qend) {
                T delta = ang[r] - prev;
                if (delta > PI) delta -= TAU;
                if (delta < -PI) delta += TAU;
                store_sig(a.rawdemod, a.out_f32, (long long)f * a.spfp + (q - kOverlap), (double)delta * a.scale);
            }
        }
    }
}

// The chunk's values from a replayed graph's plan (K1ALaunch::plan).
__device__ __forceinline__ void plan_chunk_a(K1ALaunch& a) {
    const ChunkPlan& p = *a.plan;
    a.cap = p.chunk;
    a.n_valid = p.n_read;
    a.total = p.total;
    a.ctl_raw_offset = (long long)p.raw_offset;
    a.ctl_fields_k1 = p.fields_k1;
}

template <int B, typename T>
__global__ __launch_bounds__(B / 16, (k1_min_blocks<B, T>())) void k1_part_a(K1ALaunch a) {
    if (a.plan) {
        if ((int)blockIdx.y >= a.plan->fields_k1) return;
        plan_chunk_a(a);
    }
    k1_part_a_block<B, T>(a, blockIdx.y, blockIdx.x);
}

// Block-list mode (the guard band's FP64 recompute): workgroup i takes the
// listed block i; the grid covers the longest list (every block of the
// chunk) and workgroups past the list's length return.  One block per
// workgroup keeps the full launch's registers and occupancy (a persistent
// loop over the list took 216 VGPRs, two workgroups per CU, and two rounds
// of block latency; measured 0.9% / 1.6% more fields/s mixed-fp16 / bf16
// this way).
template <int B, typename T>
__global__ __launch_bounds__(B / 16, (k1_min_blocks<B, T>())) void k1_part_a_list(K1ALaunch a) {
    if (a.plan) plan_chunk_a(a);
    const int i = blockIdx.x;
    if (i >= *a.bcount) return;
    const int2 fb = a.blist[i];
    k1_part_a_block<B, T>(a, fb.x, fb.y);
}

// Part B: video and sync filters in one complex transform
// (real input, kernel g_video + i * g_sync -> real part demod, imag demod_05).
template <int B, typename T>
__device__ __forceinline__ void k1_part_b_block(const K1BLaunch& a, int f, int b) {
    using V2 = typename C2<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    V2* buf = reinterpret_cast<V2*>(smem);
    const V2* G = static_cast<const V2*>(a.G);
    const V2* tw = static_cast<const V2*>(a.tw);
    constexpr int NT = B / 16;
    const int qs = kOverlap + b * a.V;
    const int qend = kOverlap + a.spfp;
    if (qs >= qend) return;
    const int seg0 = f * a.spfp - kOverlap;
    const int last = (int)a.total - 1;
    T xin[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int s = seg0 + wrap_n(qs - a.lp + (int)threadIdx.x + i * NT, a.N);
        s = s < 0 ? 0 : (s > last ? last : s);
        xin[i] = (T)load_sig(a.rawdemod, a.in_f32, s);
    }
    V2 v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        v[i].x = xin[i];
        v[i].y = (T)0;
    }
    static_assert(B == 4096 || B == 8192, "K1 blocks are wave-local 4096- or 8192-point blocks");
    {
        if constexpr (B == 4096) conv4096<T, k1_whole<B, T>()>(smem, v, tw, G);
        else conv_rows<T, 512>(buf, v, tw, G);
#pragma unroll
        for (int r = 0; r < 16; r++) {
            int o = r * NT + (int)threadIdx.x - a.lp;
            int q = qs + o;
            if (o >= 0 && o < a.V && q < qend) {
                long long s = (long long)f * a.spfp + (q - kOverlap);
                store_sig(a.demod, a.demod_f32, s, (double)v[r].x);
                store_sig(a.demod05, a.d05_f32, s, (double)v[r].y);
                // the guard band's flag of the output's 256 samples: close
                // enough to the sync threshold that its decision needs FP64
                if (a.gflag && fabs((double)v[r].y - a.thr) < a.gdelta) a.gflag[s >> kGuardSubShift] = 1;
            }
        }
        if (a.thr_bits) {
            // sync-threshold bitmask of the stored demod_05 values over the
            // block's samples [s_lo, s_end), assembled in LDS: the
            // wavefront's 64 consecutive samples of row r as one ballot,
            // ORed into the (up to three) words they straddle.  Words inside
            // the range are then stored whole; the first and last, shared
            // with the neighbouring blocks when partly covered, are ORed
            // into the bitmask part A cleared (in block-list mode, where the
            // words hold earlier bits, the block's own bits are cleared
            // first and the neighbours' kept).
            constexpr int kWords = B / 32 + 4;
            __shared__ unsigned int s_bits[kWords];
            const int lane = (int)threadIdx.x & 63;
            const long long s_lo = (long long)f * a.spfp + (qs - kOverlap);
            const long long s_end = (long long)f * a.spfp + (min(qs + a.V, qend) - kOverlap);
            const long long w_lo = s_lo >> 5, w_hi = (s_end - 1) >> 5;
            const int nw = (int)(w_hi - w_lo) + 1;
            const int off = (int)(s_lo & 31);
            __syncthreads();
            for (int i = (int)threadIdx.x; i < kWords; i += NT) s_bits[i] = 0u;
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int o = r * NT + (int)threadIdx.x - a.lp;
                const int q = qs + o;
                const bool in = o >= 0 && o < a.V && q < qend;
                const unsigned long long m = __ballot(in && (double)v[r].y <= a.thr);
                if (m != 0ull && lane < 3) {
                    const int rel = off + (o - lane);  // lane 0's sample - 32 w_lo
                    const int sh = rel & 31;
                    unsigned int part;
                    if (lane == 0) part = (unsigned int)(m << sh);
                    else if (lane == 1) part = (unsigned int)((m << sh) >> 32);
                    else part = sh ? (unsigned int)(m >> (64 - sh)) : 0u;
                    if (part) atomicOr(&s_bits[(rel >> 5) + lane], part);
                }
            }
            __syncthreads();
            for (int i = (int)threadIdx.x; i < nw; i += NT) {
                const unsigned int word = s_bits[i];
                const int lo_bit = i == 0 ? off : 0;
                const int hi_bit = (i == nw - 1 && (s_end & 31) != 0) ? (int)(s_end & 31) : 32;
                const bool partial = lo_bit != 0 || hi_bit != 32;
                if (!partial) {
                    a.thr_bits[w_lo + i] = word;
                } else if (a.blist) {
                    const unsigned int own = (hi_bit == 32 ? ~0u : ((1u << hi_bit) - 1u)) & ~((1u << lo_bit) - 1u);
                    atomicAnd(a.thr_bits + (w_lo + i), ~own);
                    if (word) atomicOr(a.thr_bits + (w_lo + i), word & own);
                } else if (word) {
                    atomicOr(a.thr_bits + (w_lo + i), word);
                }
            }
        }
    }
}

template <int B, typename T>
__global__ __launch_bounds__(B / 16, (k1_min_blocks<B, T>())) void k1_part_b(K1BLaunch a) {
    if (a.plan) {
        if ((int)blockIdx.y >= a.plan->fields_k1) return;
        a.total = a.plan->total;
    }
    k1_part_b_block<B, T>(a, blockIdx.y, blockIdx.x);
}

// Block-list mode of part B: workgroup i takes the listed block i (as
// k1_part_a_list).
template <int B, typename T>
__global__ __launch_bounds__(B / 16, (k1_min_blocks<B, T>())) void k1_part_b_list(K1BLaunch a) {
    if (a.plan) a.total = a.plan->total;
    const int i = blockIdx.x;
    if (i >= *a.bcount) return;
    const int2 fb = a.blist[i];
    k1_part_b_block<B, T>(a, fb.x, fb.y);
}

// Chroma pre-band-pass: segment f is raw[f*spfp, (f+1)*spfp) zero-padded to N2
// and filtered circularly (chroma_decode.cu:606-627).  The filter is real,
// so each workgroup convolves two adjacent V-output blocks in one complex
// transform: block 2p in the real part, block 2p+1 in the imaginary part
// (G is the spectrum of a real kernel, so the two never mix beyond rounding).
// T is the transform precision: double in the FP64 form (the reference's
// FP64 cuFFT band-pass; real and imaginary parts exchanged through LDS in
// turn, 139 KB for 16384 points), float in the FP32 form.  The output is
// stored as float either way: it holds only the 60 kHz - 1.2 MHz band, so
// its FP32 rounding (6e-8 of the band's own level) is far below a chroma LSB.
template <int B, typename T>
__global__ __launch_bounds__(B / 16) void k_prebpf(PreLaunch a) {
    using V2 = typename C2<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NT = B / 16;
    const int f = blockIdx.y;
    if (a.plan) {
        if (f >= a.plan->max_fields) return;
        a.cap = a.plan->chunk;
        a.n_valid = a.plan->n_read;
    }
    const int ns0 = 2 * (int)blockIdx.x * a.V;
    if (ns0 >= a.spfp) return;
    const int ns1 = ns0 + a.V;
    const bool two = ns1 < a.spfp;
    const int seg0 = f * a.spfp;
    const double dcv = a.dc ? *a.dc : 0.0;
    V2 v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        int n0 = wrap_n(ns0 - a.lp + (int)threadIdx.x + i * NT, a.N2);
        int n1 = wrap_n(ns1 - a.lp + (int)threadIdx.x + i * NT, a.N2);
        v[i].x = n0 < a.spfp ? (T)raw_sample(a.cap, a.in_fmt, seg0 + n0, a.n_valid, dcv) : (T)0;
        v[i].y = two && n1 < a.spfp ? (T)raw_sample(a.cap, a.in_fmt, seg0 + n1, a.n_valid, dcv) : (T)0;
    }
    // G is in the wave-local block's bin order; v[r] becomes the outputs at
    // o = NT r + t - lp of both blocks
    conv_rows<T, NT, sizeof(T) == 4>(smem, v, static_cast<const V2*>(a.tw), static_cast<const V2*>(a.G));
#pragma unroll
    for (int r = 0; r < 16; r++) {
        int o = r * NT + (int)threadIdx.x - a.lp;
        if (o >= 0 && o < a.V) {
            if (ns0 + o < a.spfp) a.out[seg0 + ns0 + o] = (float)v[r].x;
            if (ns1 + o < a.spfp) a.out[seg0 + ns1 + o] = (float)v[r].y;
        }
    }
}

__global__ void k_fill_zero(unsigned int* p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long st = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = 0u;
}

// K1 outputs of [plan total, plan clear_end) to zero in three buffers of
// w0 / w1 / w2 32-bit words per sample (what an earlier, longer chunk left
// there); nothing to do for most chunks.
__global__ __launch_bounds__(256) void k_clear_past(unsigned int* p0, int w0, unsigned int* p1, int w1,
                                                    unsigned int* p2, int w2, const ChunkPlan* plan) {
    const long long t0 = plan->total, t1 = plan->clear_end;
    if (t1 <= t0) return;
    const long long st = (long long)gridDim.x * blockDim.x;
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (long long i = t0 * w0 + i0; i < t1 * w0; i += st) p0[i] = 0u;
    for (long long i = t0 * w1 + i0; i < t1 * w1; i += st) p1[i] = 0u;
    for (long long i = t0 * w2 + i0; i < t1 * w2; i += st) p2[i] = 0u;
}

// The chunk's plan to the device (a kernel argument carries it: cheaper on
// the head stream than a copy from pinned memory).
__global__ void k_put_plan(ChunkPlan* dst, ChunkPlan v) {
    if (threadIdx.x == 0) *dst = v;
}

__global__ void k_fill_f64(double* p, long long n, double v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long st = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += st) p[i] = v;
}

// ---- host launchers ----
// Dynamic LDS above 64 KB needs a per-kernel attribute; set it once per
// device rather than before every launch.
// Keyed by the kernel itself: kernels of one signature share a function
// type, so a static per template instance would cover all of them at once.
template <typename K>
static void ensure_lds(K* kern, size_t sh) {
    static std::mutex mu;
    static std::set<std::pair<const void*, int>> done;
    int d = 0;
    (void)hipGetDevice(&d);
    const std::pair<const void*, int> key{(const void*)kern, d};
    std::lock_guard<std::mutex> lock(mu);
    if (done.count(key)) return;
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    done.insert(key);
}
template <int B, typename T>
static hipError_t launch_a(const K1ALaunch& a, int nf, hipStream_t st) {
    int nb = (int)((a.spfp + a.V - 1) / a.V);
    size_t sh = k1_lds_bytes<B, T>();
    if (a.blist) {
        ensure_lds(k1_part_a_list<B, T>, sh);
        hipLaunchKernelGGL((k1_part_a_list<B, T>), dim3(nb * nf), dim3(B / 16), sh, st, a);
    } else {
        ensure_lds(k1_part_a<B, T>, sh);
        hipLaunchKernelGGL((k1_part_a<B, T>), dim3(nb, nf), dim3(B / 16), sh, st, a);
    }
    return hipGetLastError();
}
template <int B, typename T>
static hipError_t launch_b(const K1BLaunch& a, int nf, hipStream_t st) {
    int nb = (int)((a.spfp + a.V - 1) / a.V);
    size_t sh = k1_lds_bytes<B, T>();
    if (a.blist) {
        ensure_lds(k1_part_b_list<B, T>, sh);
        hipLaunchKernelGGL((k1_part_b_list<B, T>), dim3(nb * nf), dim3(B / 16), sh, st, a);
    } else {
        ensure_lds(k1_part_b<B, T>, sh);
        hipLaunchKernelGGL((k1_part_b<B, T>), dim3(nb, nf), dim3(B / 16), sh, st, a);
    }
    return hipGetLastError();
}
template <int B, typename T>
static hipError_t launch_pre(const PreLaunch& a, int nf, hipStream_t st) {
    int nb = (int)((a.spfp + 2 * a.V - 1) / (2 * a.V));
    // whole complex elements in FP32, real and imaginary parts in turn in FP64
    size_t sh = (size_t)lds_elems(B) * (sizeof(T) == 4 ? 2 * sizeof(T) : sizeof(T));
    ensure_lds(k_prebpf<B, T>, sh);
    hipLaunchKernelGGL((k_prebpf<B, T>), dim3(nb, nf), dim3(B / 16), sh, st, a);
    return hipGetLastError();
}

hipError_t run_k1_part_a(int B, int f32, const K1ALaunch& a, int nf, hipStream_t st) {
    if (!f32 && B == 4096) return launch_a<4096, double>(a, nf, st);
    if (!f32 && B == 8192) return launch_a<8192, double>(a, nf, st);
    if (f32 && B == 4096) return launch_a<4096, float>(a, nf, st);
    if (f32 && B == 8192) return launch_a<8192, float>(a, nf, st);
    return hipErrorInvalidValue;
}
hipError_t run_k1_part_b(int B, int f32, const K1BLaunch& a, int nf, hipStream_t st) {
    if (!f32 && B == 4096) return launch_b<4096, double>(a, nf, st);
    if (!f32 && B == 8192) return launch_b<8192, double>(a, nf, st);
    if (f32 && B == 4096) return launch_b<4096, float>(a, nf, st);
    if (f32 && B == 8192) return launch_b<8192, float>(a, nf, st);
    return hipErrorInvalidValue;
}
hipError_t run_prebpf(int B, int f32, const PreLaunch& a, int nf, hipStream_t st) {
    if (f32 && B == 8192) return launch_pre<8192, float>(a, nf, st);
    if (f32 && B == 16384) return launch_pre<16384, float>(a, nf, st);
    if (!f32 && B == 8192) return launch_pre<8192, double>(a, nf, st);
    if (!f32 && B == 16384) return launch_pre<16384, double>(a, nf, st);
    return hipErrorInvalidValue;
}
hipError_t run_input_mean(const void* cap, int in_fmt, long long n, double* partial, double* mean, hipStream_t st,
                          const ChunkPlan* plan) {
    if (n <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_input_sum, dim3(kMeanBlocks), dim3(256), 0, st, cap, in_fmt, n, partial, plan);
    hipLaunchKernelGGL(k_input_mean, dim3(1), dim3(kMeanBlocks), 0, st, partial, n, mean, plan);
    return hipGetLastError();
}
__global__ void k_to_f64(const float* src, double* dst, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long st = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += st) dst[i] = (double)src[i];
}

hipError_t run_to_f64(const void* src, int f32, double* dst, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!f32) return hipMemcpyAsync(dst, src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st);
    hipLaunchKernelGGL(k_to_f64, dim3(1024), dim3(256), 0, st, static_cast<const float*>(src), dst, n);
    return hipGetLastError();
}

hipError_t run_fill_zero(void* p, long long words, hipStream_t st) {
    if (words <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_zero, dim3(1024), dim3(256), 0, st, static_cast<unsigned int*>(p), words);
    return hipGetLastError();
}

hipError_t run_clear_past(void* p0, int w0, void* p1, int w1, void* p2, int w2, const ChunkPlan* plan,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_clear_past, dim3(512), dim3(256), 0, st, static_cast<unsigned int*>(p0), w0,
                       static_cast<unsigned int*>(p1), w1, static_cast<unsigned int*>(p2), w2, plan);
    return hipGetLastError();
}

hipError_t run_put_plan(ChunkPlan* dst, const ChunkPlan& v, hipStream_t st) {
    hipLaunchKernelGGL(k_put_plan, dim3(1), dim3(64), 0, st, dst, v);
    return hipGetLastError();
}

hipError_t run_fill_f64(double* p, long long n, double v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3(1024), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

// LDS layout of the self-tested blocks as the decoder runs them: 4096-point
// blocks as K1's, FP64 1024- and 16384-point blocks exchange real and
// imaginary parts in turn.
template <int B, typename T>
constexpr bool selftest_whole() {
    return B == 4096 ? k1_whole<B, T>() : !(sizeof(T) == 8 && (B == 1024 || B == 16384));
}

// Self-test of the convolution blocks the decoder runs: y = B IFFT(FFT(x) G)
// for one block, in the register layout the kernels use (v[r] = x[NT r + t]),
// with G in the block's bin order.  B = 4096: conv4096 (FP64 with split
// exchanges as in K1, FP32 whole); 8192 / 16384: conv_rows<512 / 1024>
// (FP64 16384 with split exchanges as in the pre-band-pass);
// 1024: filter1024_wave (one wavefront, G applied per bin as in the chroma
// kernel: FP64 with split exchanges, or FP32 whole as in the reduced forms).
template <int B, typename T>
__global__ __launch_bounds__(B == 1024 ? 64 : B / 16) void k_block_selftest(const typename C2<T>::type* x,
                                                                          const typename C2<T>::type* gp,
                                                                          typename C2<T>::type* y,
                                                                          const typename C2<T>::type* tw) {
    using V = typename C2<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int NT = B == 1024 ? 64 : B / 16;
    V v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) v[r] = x[NT * r + threadIdx.x];
    if constexpr (B == 1024) {
        filter1024_wave<T, selftest_whole<B, T>()>(smem, v, tw, [&](V z, int k) { return cmul(z, gp[k]); });
    } else if constexpr (B == 4096) {
        conv4096<T, k1_whole<B, T>()>(smem, v, tw, gp);
    } else {
        conv_rows<T, B / 16, selftest_whole<B, T>()>(smem, v, tw, gp);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) y[NT * r + threadIdx.x] = v[r];
}

template <int B, typename T>
static hipError_t block_st_launch(const void* x, const void* gp, void* y, const void* tw, hipStream_t st) {
    using V = typename C2<T>::type;
    size_t sh = (size_t)lds_elems(B) * (selftest_whole<B, T>() ? sizeof(V) : sizeof(T));
    (void)hipFuncSetAttribute((const void*)k_block_selftest<B, T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    hipLaunchKernelGGL((k_block_selftest<B, T>), dim3(1), dim3(B == 1024 ? 64 : B / 16), sh, st, (const V*)x,
                       (const V*)gp, (V*)y, (const V*)tw);
    return hipGetLastError();
}

hipError_t run_block_selftest(int B, int is_f32, const void* x, const void* gp, void* y, const void* tw,
                              hipStream_t st) {
    if (!is_f32) {
        if (B == 1024) return block_st_launch<1024, double>(x, gp, y, tw, st);
        if (B == 4096) return block_st_launch<4096, double>(x, gp, y, tw, st);
        if (B == 8192) return block_st_launch<8192, double>(x, gp, y, tw, st);
        if (B == 16384) return block_st_launch<16384, double>(x, gp, y, tw, st);
    } else {
        if (B == 1024) return block_st_launch<1024, float>(x, gp, y, tw, st);
        if (B == 4096) return block_st_launch<4096, float>(x, gp, y, tw, st);
        if (B == 8192) return block_st_launch<8192, float>(x, gp, y, tw, st);
        if (B == 16384) return block_st_launch<16384, float>(x, gp, y, tw, st);
    }
    return hipErrorInvalidValue;
}

// FP64 arctangent self-test (atan_f64.h) over device arrays.
__global__ __launch_bounds__(256) void k_atan2_selftest(const double* y, const double* x, double* out, long long n) {
    const long long stride = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) out[i] = atan2_f64(y[i], x[i]);
}

hipError_t run_atan2_selftest(const double* y, const double* x, double* out, long long n, hipStream_t st) {
    if (n <= 0) return hipErrorInvalidValue;
    const int nb = (int)std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_atan2_selftest, dim3(nb), dim3(256), 0, st, y, x, out, n);
    return hipGetLastError();
}

}  // namespace cuvhs
