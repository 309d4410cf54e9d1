// K1 part A on the matrix cores: the analytic signal of the RF band-pass x
// Hilbert filter as a Toeplitz product, for the mixed-precision forms.
//
// Reference: fm_demod.cu:396-443 (D2Z, x Hrf, one-sided Z2Z inverse, |a|,
// atan2 and the wrapped phase difference, all FP64 cuFFT).  Here, for
// outputs q of segment f (window coordinates as in k1_part_a),
//   a(q) = sum_m h[m] x_f(q - m),  m = -lm .. lp,
// with h the filter's impulse response truncated to taps above 1e-8 of the
// largest.  Written as y[o] = sum_u w[u] xw[o + u] (w[u] = h[lp - u], xw the
// input window from q0 - lp), a tile of 1024 outputs o = 32 j + m is the
// matrix product C[m][j] = sum_c sum_r A_c[m][r] B_c[r][j] with
//   A_c[m][r] = w[16 c + r - m]   (32 x 16, the same for every tile)
//   B_c[r][j] = xw[32 j + 16 c + r]   (16 x 32, read from LDS)
// one v_mfma_f32_32x32x16 per (c, operand).  The capture's u8 samples minus
// 128 are small integers, exact in f16 and bf16, so only the taps need
// splitting: w = hi + lo (each f16 or bf16, scaled by a power of two into
// range), two products per step, accumulated in FP32.  The taps decay by
// orders of magnitude away from the kernel's centre; where every tap a group
// of four steps holds is below 2^-16 (bf16: 2^-19) of the largest, its lo
// part (at most 2^-11 of the tap in f16, 2^-8 in bf16) is below the FP32
// accumulation's own rounding and is dropped: those groups, at both ends of
// the kernel, take one product per step, the groups [glo0, glo1) between
// them two.  A workgroup of four
// waves computes 16,384 consecutive outputs (the first only as the phase
// predecessor of the second), each wave four tiles that share every A
// operand load.  Lane layout of the 32x32x16 operations (gfx950): lane l
// holds A[l % 32][8 (l / 32) + k] and B[8 (l / 32) + k][l % 32], k < 8, and
// accumulator i holds C[8 (i / 4) + 4 (l / 32) + i % 4][l % 32].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <type_traits>
#include <vector>

#include "common.h"
#include "launch.h"

namespace cuvhs {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kMWaves = 4, kMTiles = 4;
constexpr int kMOut = kMWaves * kMTiles * 1024;   // outputs computed per workgroup
constexpr int kMValid = kMOut - 1;                // the first is only a phase predecessor
constexpr int kMGroup = 4;                        // tap steps per LDS group
constexpr int kEpiRow = 1024 + 32;                // floats per padded output row of the epilogue
constexpr int kStagePasses = 6;                   // interior staging: 16 elements x 256 threads per pass

// LDS index of input element i: 8 halves of padding after every 32, so the
// 64-byte stride between the columns j of a B operand read spreads over the
// banks; a read of 8 elements starting at a multiple of 8 stays in one run.
__device__ __forceinline__ int xpad(int i) { return i + ((i >> 5) << 3); }

// atan2 in FP32 for the phase of the analytic signal: octant reduction to
// t = min(|x|, |y|) / max(|x|, |y|) in [0, 1], atan t = t Q(t^2) with Q of
// degree 7 fitted (iteratively reweighted least squares on Chebyshev nodes)
// for the smallest largest absolute error on [0, 1]: 3.8e-8 rad in exact
// arithmetic, 1.2e-7 rad evaluated in FP32 (atan2f's own error near pi is
// 2.4e-7).  Then the octant's reflections; atan2(0, 0) = 0.
__device__ __forceinline__ float phase_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mn = fminf(ax, ay), mx = fmaxf(ax, ay);
    const float t = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
    const float s = t * t;
    float q = -4.054525867e-03f;
    q = fmaf(q, s, 2.186283097e-02f);
    q = fmaf(q, s, -5.591218546e-02f);
    q = fmaf(q, s, 9.642191231e-02f);
    q = fmaf(q, s, -1.390863061e-01f);
    q = fmaf(q, s, 1.994656622e-01f);
    q = fmaf(q, s, -3.332986236e-01f);
    q = fmaf(q, s, 9.999993443e-01f);
    float r = t * q;
    if (ay > ax) r = 1.57079632679489662f - r;
    if (x < 0.0f) r = 3.14159265358979324f - r;
    return copysignf(r, y);
}

template <typename E> struct MfmaOps;
template <> struct MfmaOps<_Float16> {
    using V8 = f16x8;
    static __device__ __forceinline__ f32x16 mma(V8 a, V8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct MfmaOps<__bf16> {
    using V8 = bf16x8;
    static __device__ __forceinline__ f32x16 mma(V8 a, V8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ int wrap_m(int q, int n) { return q < 0 ? q + n : (q >= n ? q - n : q); }

template <typename E>
__global__ __launch_bounds__(256, 2) __attribute__((amdgpu_waves_per_eu(2, 2))) void k1_part_a_mfma(K1MLaunch a) {
    using Ops = MfmaOps<E>;
    using V8 = typename Ops::V8;
    extern __shared__ __align__(16) unsigned char smem[];
    E* xw = reinterpret_cast<E*>(smem);
    __shared__ float s_edge[kMWaves];
    const int f = blockIdx.y, b = blockIdx.x;
    const int t = threadIdx.x;
    if (a.plan) {   // a replayed graph: the chunk's values from its plan
        const ChunkPlan& p = *a.plan;
        if (f >= p.fields_k1) return;
        a.cap = p.chunk;
        a.n_valid = p.n_read;
        a.total = p.total;
        a.ctl_raw_offset = (long long)p.raw_offset;
        a.ctl_fields_k1 = p.fields_k1;
    }
    if (a.ctl && f == 0 && b == 0 && t == 0) {
        ChunkCtl z{};
        z.fields_k1 = a.ctl_fields_k1;
        z.raw_offset = a.ctl_raw_offset;
        *a.ctl = z;
    }
    const int qs = (kOverlap - 1) + b * kMValid;   // output o = 0 of this workgroup
    const int qend = kOverlap + a.spfp;
    if (qs + 1 >= qend) return;
    if (a.thr_bits) {
        const long long s0 = (long long)f * a.spfp + (long long)b * kMValid;
        const long long s1 = (long long)f * a.spfp + min((long long)(b + 1) * kMValid, (long long)a.spfp);
        for (long long w = (s0 >> 5) + t; w <= ((s1 - 1) >> 5); w += 256) a.thr_bits[w] = 0u;
    }
    // input window xw[i] = x_f(qs - lp + i): u8 sample minus 128 (the 256
    // scale is applied to the product), zero past the read, the segment's
    // window rules (replicate below the chunk, clamp at its end, wrap mod N)
    const int seg0 = f * a.spfp - kOverlap;
    const long long last = a.total - 1;
    const uint8_t* cap = static_cast<const uint8_t*>(a.cap);
    const long long base = (long long)seg0 + qs - a.lp;   // sample of xw[0] when the window does not wrap
    if (qs - a.lp >= 0 && qs - a.lp + a.nx <= a.N && base >= 4 && base + a.nx + 4 <= a.n_valid) {
        // interior: 16 elements per thread and pass from five aligned dwords
        // (the window is 16-element aligned: nx is a multiple of 16)
        const unsigned int sh = (unsigned int)(reinterpret_cast<uintptr_t>(cap + base) & 3u);
        const unsigned int* wp = reinterpret_cast<const unsigned int*>(cap + (base - sh));
        const int nk = a.nx >> 4;   // at most kStagePasses * 256 (checked by the launcher)
        // every load first (one memory latency for the whole window), then
        // the conversions and LDS stores
        unsigned int d[kStagePasses][5];
#pragma unroll
        for (int r = 0; r < kStagePasses; r++) {
            const int k = min(t + 256 * r, nk - 1);   // past the window: a repeat, not stored
#pragma unroll
            for (int m = 0; m < 5; m++) d[r][m] = wp[4 * k + m];
        }
#pragma unroll
        for (int r = 0; r < kStagePasses; r++) {
            const int k = t + 256 * r;
            if (k >= nk) continue;
            E e[16];
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const unsigned int v = __builtin_amdgcn_alignbyte(d[r][m + 1], d[r][m], sh);
#pragma unroll
                for (int z = 0; z < 4; z++) e[4 * m + z] = (E)((float)((v >> (8 * z)) & 0xffu) - 128.0f);
            }
            uint4* dst = reinterpret_cast<uint4*>(xw + xpad(16 * k));
            dst[0] = *reinterpret_cast<const uint4*>(&e[0]);
            dst[1] = *reinterpret_cast<const uint4*>(&e[8]);
        }
    } else {
        // xw[i] = x_f(qs - lp + i): u8 sample minus 128 (the 256 scale is
        // applied to the product), zero past the read, the segment's window
        // rules (replicate below the chunk, clamp at its end, wrap mod N);
        // eight loads in flight per thread (every load in bounds: an index
        // past the read loads sample 0 and is replaced by zero)
        const long long nv = a.n_valid;
        for (int i0 = 0; i0 < a.nx; i0 += 8 * 256) {
            if (nv <= 0) {   // nothing read: the window is zero
                for (int i = i0 + t; i < min(i0 + 8 * 256, a.nx); i += 256) xw[xpad(i)] = (E)0.0f;
                continue;
            }
            unsigned int u[8];
            bool in[8];
#pragma unroll
            for (int z = 0; z < 8; z++) {
                const int i = i0 + z * 256 + t;
                long long s = seg0 + wrap_m(qs - a.lp + (i < a.nx ? i : 0), a.N);
                s = s < 0 ? 0 : (s > last ? last : s);
                in[z] = s < nv;
                u[z] = cap[in[z] ? s : 0];
            }
            float v[8];
#pragma unroll
            for (int z = 0; z < 8; z++) v[z] = in[z] ? (float)u[z] - 128.0f : 0.0f;
#pragma unroll
            for (int z = 0; z < 8; z++) {
                const int i = i0 + z * 256 + t;
                if (i < a.nx) xw[xpad(i)] = (E)v[z];
            }
        }
    }
    __syncthreads();

    const int w = t >> 6, l = t & 63, j = l & 31, h = l >> 5;
    f32x16 acc[2][kMTiles];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
        for (int k = 0; k < kMTiles; k++)
#pragma unroll
            for (int i = 0; i < 16; i++) acc[p][k][i] = 0.0f;
    // Tap operands through LDS in groups of kMGroup steps, double-buffered,
    // and the operands of each step loaded into registers during the step
    // before.  Group g: the workgroup loads group g + 1's taps from global
    // memory at its start, stores them to the idle buffer after step 2's
    // products (every wave's reads of that buffer ended before the previous
    // group's barrier), then one barrier, so step 3 can prefetch the next
    // group's first step.  Within a step, the eight accumulators take their
    // hi products first and their lo products after, so no product waits on
    // the one before it.
    uint4* sA = reinterpret_cast<uint4*>(smem + a.a_lds_offset);   // [2][kMGroup * 4 * 64]
    constexpr int kGroupEntries = kMGroup * 4 * 64;
    const int ngroups = a.nc / kMGroup;   // the table is padded to whole groups
#pragma unroll
    for (int q = 0; q < kGroupEntries / 256; q++) sA[q * 256 + t] = a.A[q * 256 + t];
    __syncthreads();
    // B operand of step c, tile k: xw + xpad(xb + 1024 k + 16 c).  xb mod 32
    // is 0 or 8, so that is xpad(xb) + 1280 k + 40 (c / 2) + 16 (c mod 2):
    // per group of four steps, a base plus constants
    const E* xlane = xw + xpad(w * (kMTiles * 1024) + 32 * j + 8 * h);
    static_assert(kMGroup == 4, "step offsets below are for groups of four");
    constexpr int kStepOff[4] = {0, 16, 40, 56};
    uint4 ac[4], an[4];
    V8 bc[kMTiles], bn[kMTiles];
#pragma unroll
    for (int p = 0; p < 4; p++) ac[p] = sA[p * 64 + l];
    // B operand offset of group g: 80 elements per group of four steps
    auto gofs = [&](int g) { return 80 * g; };
#pragma unroll
    for (int k = 0; k < kMTiles; k++) bc[k] = *reinterpret_cast<const V8*>(xlane + gofs(0) + 1280 * k);
    auto run_group = [&](int g, auto lo_tag) {
        constexpr bool LO = decltype(lo_tag)::value;
        const int buf = g & 1;
        // the last group reloads itself into the idle buffer and prefetches
        // its own first step again (no branch; the values go unused)
        const int gn = g + 1 < ngroups ? g + 1 : g;
        static_assert(kGroupEntries == 4 * 256, "four staged entries per thread");
        const uint4* An = a.A + (size_t)gn * kGroupEntries + t;
        const uint4 st0 = An[0], st1 = An[256], st2 = An[512], st3 = An[768];
        const uint4* sa = sA + buf * kGroupEntries;
        uint4* sn = sA + (buf ^ 1) * kGroupEntries;
#pragma unroll
        for (int cc = 0; cc < kMGroup; cc++) {
            if (cc + 1 < kMGroup) {
#pragma unroll
                for (int p = 0; p < 4; p++) an[p] = sa[((cc + 1) * 4 + p) * 64 + l];
#pragma unroll
                for (int k = 0; k < kMTiles; k++)
                    bn[k] = *reinterpret_cast<const V8*>(xlane + gofs(g) + kStepOff[cc + 1] + 1280 * k);
            } else {
#pragma unroll
                for (int p = 0; p < 4; p++) an[p] = sn[p * 64 + l];
#pragma unroll
                for (int k = 0; k < kMTiles; k++) bn[k] = *reinterpret_cast<const V8*>(xlane + gofs(gn) + 1280 * k);
            }
#pragma unroll
            for (int k = 0; k < kMTiles; k++) {
                acc[0][k] = Ops::mma(__builtin_bit_cast(V8, ac[0]), bc[k], acc[0][k]);
                acc[1][k] = Ops::mma(__builtin_bit_cast(V8, ac[2]), bc[k], acc[1][k]);
            }
            if constexpr (LO) {
#pragma unroll
                for (int k = 0; k < kMTiles; k++) {
                    acc[0][k] = Ops::mma(__builtin_bit_cast(V8, ac[1]), bc[k], acc[0][k]);
                    acc[1][k] = Ops::mma(__builtin_bit_cast(V8, ac[3]), bc[k], acc[1][k]);
                }
            }
            // issue order: (at a group's start, the four global tap loads),
            // then two products (one without lo) and one read of the next
            // step's operands, eight times: the reads complete under the
            // products
            if (cc == 0) __builtin_amdgcn_sched_group_barrier(0x020, 4, 0);
#pragma unroll
            for (int z = 0; z < 8; z++) {
                __builtin_amdgcn_sched_group_barrier(0x008, LO ? 2 : 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            }
            if (cc == kMGroup - 2) {
                sn[t] = st0;
                sn[256 + t] = st1;
                sn[512 + t] = st2;
                sn[768 + t] = st3;
                __builtin_amdgcn_sched_group_barrier(0x200, 4, 0);
                __syncthreads();
            }
#pragma unroll
            for (int p = 0; p < 4; p++) ac[p] = an[p];
#pragma unroll
            for (int k = 0; k < kMTiles; k++) bc[k] = bn[k];
        }
    };
    for (int g = 0; g < a.glo0; g++) run_group(g, std::false_type{});
    for (int g = a.glo0; g < a.glo1; g++) run_group(g, std::true_type{});
    for (int g = a.glo1; g < ngroups; g++) run_group(g, std::false_type{});
    __syncthreads();   // every wave's reads of xw end before the epilogue reuses it

    // epilogue: envelope |acc| scale, phase arg(acc) (the scale is a
    // positive power of two), and the wrapped phase difference to the
    // previous output
    const float scale = a.scale;
    const float PI = 3.14159265358979323846f, TAU = 6.28318530717958647692f;
    // phase of this wave's last output (tile 3, row 31, column 31: lane 63,
    // accumulator 15) for the next wave's first output
    if (l == 63) s_edge[w] = phase_atan2(acc[1][kMTiles - 1][15], acc[0][kMTiles - 1][15]);
    __syncthreads();
    float carry = w > 0 ? s_edge[w - 1] : 0.0f;   // phase before this tile's first output
    // per wave two LDS rows of 1024 (+ 32 padding) floats, envelope then
    // phase difference: output o at o + o / 32.  Lane (j, h) writes rows
    // 8 (i / 4) + 4 h + i % 4 of column j, i.e. o + o / 32 = 33 j + 4 h +
    // 8 (i / 4) + i % 4; reads back o = 64 r + l at 66 r + l + l / 32.  Only
    // this wave touches its rows, and a wave's LDS operations complete in
    // order, so a wave-scope fence (against compiler reordering) suffices.
    float* se = reinterpret_cast<float*>(smem) + w * (2 * kEpiRow);
    float* sd = se + kEpiRow;
    float* const sew = se + 33 * j + 4 * h;
    float* const sdw = sd + 33 * j + 4 * h;
    const float* const ser = se + l + (l >> 5);
    const float* const sdr = sd + l + (l >> 5);
#pragma unroll
    for (int k = 0; k < kMTiles; k++) {
        float ph[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const float re = acc[0][k][i], im = acc[1][k][i];
            sew[8 * (i >> 2) + (i & 3)] = __builtin_amdgcn_sqrtf(re * re + im * im) * scale;
            ph[i] = phase_atan2(im, re);
        }
        // predecessor phases: row m - 1 of the same column, or row 31 of
        // column j - 1 for row 0 (lane j - 1 + 32), or the carry for column 0
        float pv[16];
#pragma unroll
        for (int i4 = 0; i4 < 16; i4 += 4) {
            // rows 4h + 8 (i4 / 4) .. + 3 of this lane; its row - 1 for the
            // first is in the partner lane l ^ 32: accumulator i4 + 3 there
            // (this lane h = 1), or i4 - 1 (h = 0)
            const float send = h == 0 ? ph[i4 + 3] : (i4 >= 4 ? ph[i4 - 1] : 0.0f);
            const float recv = __shfl_xor(send, 32);
            pv[i4] = recv;
#pragma unroll
            for (int r = 1; r < 4; r++) pv[i4 + r] = ph[i4 + r - 1];
        }
        {
            const float r31 = __shfl(ph[15], (l + 31) & 63);   // row 31 of column j - 1 (lane j - 1 + 32)
            if (h == 0) pv[0] = j > 0 ? r31 : carry;
        }
        carry = __shfl(ph[15], 63);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            float d = ph[i] - pv[i];
            if (d > PI) d -= TAU;
            if (d < -PI) d += TAU;
            sdw[8 * (i >> 2) + (i & 3)] = d * a.fs_over_tau;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // outputs o = 64 r + l of the tile: sample q = qs + obase + o
        const int obase = w * (kMTiles * 1024) + k * 1024;
        const int q0 = qs + obase;
        float emin = 3.0e38f, emax = 0.0f;
        if (q0 >= kOverlap && q0 + 1023 < qend) {
            // the whole tile is inside the segment's outputs; every output
            // but the block's first (o = 0, only a phase predecessor) has
            // its phase difference
            const long long s0 = (long long)f * a.spfp + (q0 - kOverlap) + l;
            float* envp = a.env + s0;
            float* rawp = a.rawdemod + s0;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const float e = ser[66 * r];
                envp[64 * r] = e;
                emin = fminf(emin, e);
                emax = fmaxf(emax, e);
                if (r > 0 || obase + l >= 1) rawp[64 * r] = sdr[66 * r];
            }
        } else {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int o = 64 * r + l;
                const int q = q0 + o;
                if (q < kOverlap || q >= qend) continue;
                const long long s = (long long)f * a.spfp + (q - kOverlap);
                const float e = ser[66 * r];
                a.env[s] = e;
                emin = fminf(emin, e);
                emax = fmaxf(emax, e);
                if (obase + o >= 1) a.rawdemod[s] = sdr[66 * r];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (a.blk_env) {
            // the tile's envelope range, for the guard band
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                emin = fminf(emin, __shfl_xor(emin, o));
                emax = fmaxf(emax, __shfl_xor(emax, o));
            }
            if (l == 0) a.blk_env[((size_t)f * a.nblk + b) * 16 + w * kMTiles + k] = make_float2(emin, emax);
        }
    }
}

int mfma_blocks_per_segment(int spfp) { return (spfp + kMValid - 1) / kMValid; }
int mfma_valid_per_block() { return kMValid; }

// A operands of every step c: [c][re_hi, re_lo, im_hi, im_lo][lane] x 8
// elements, lane l element k = w[16 c + 8 (l / 32) + k - l % 32] (zero
// outside the kernel), w[u] = h[lp - u] scaled by 2^e so that the largest
// magnitude is at most 2^14 (f16 range), split into hi = E(w) and lo =
// E(w - hi).  [*glo0, *glo1): the groups of kMGroup steps from the first to
// the last one holding a tap above lo_rel of the largest (their lo parts
// matter).  Returns the step count; *scale receives 2^-e.
template <typename E>
static int build_taps(const std::vector<std::complex<double>>& taps, int lm, int lp, std::vector<uint4>& out,
                      double* scale, double lo_rel, int* glo0, int* glo1) {
    const int L = lm + lp + 1;
    const int nc = (L + 31 + 15) / 16;
    const int ncp = (nc + kMGroup - 1) / kMGroup * kMGroup;   // padded with zero steps to whole groups
    double mx = 0.0;
    for (const auto& v : taps) mx = std::max(mx, std::max(std::fabs(v.real()), std::fabs(v.imag())));
    int e = 0;
    if (mx > 0.0) e = 14 - (int)std::ceil(std::log2(mx));
    const double S = std::ldexp(1.0, e);
    *scale = std::ldexp(1.0, -e);
    *glo0 = ncp / kMGroup;
    *glo1 = 0;
    for (int c = 0; c < nc; c++)
        for (int u = 16 * c - 31; u <= 16 * c + 15; u++) {
            if (u < 0 || u >= L) continue;
            const std::complex<double> w = taps[(size_t)(L - 1 - u)];
            if (std::max(std::fabs(w.real()), std::fabs(w.imag())) > lo_rel * mx) {
                *glo0 = std::min(*glo0, c / kMGroup);
                *glo1 = std::max(*glo1, c / kMGroup + 1);
            }
        }
    if (*glo1 <= *glo0) *glo0 = *glo1 = 0;
    out.assign((size_t)ncp * 4 * 64, uint4{});
    for (int c = 0; c < nc; c++)
        for (int l = 0; l < 64; l++) {
            E v[4][8];
            for (int k = 0; k < 8; k++) {
                const int u = 16 * c + 8 * (l / 32) + k - (l % 32);
                std::complex<double> w(0.0, 0.0);
                if (u >= 0 && u < L) w = taps[(size_t)(L - 1 - u)] * S;   // taps[m + lm], m = lp - u
                const E rh = (E)w.real(), ih = (E)w.imag();
                v[0][k] = rh;
                v[1][k] = (E)(w.real() - (double)rh);
                v[2][k] = ih;
                v[3][k] = (E)(w.imag() - (double)ih);
            }
            for (int p = 0; p < 4; p++) {
                uint4 q;
                static_assert(sizeof(v[p]) == sizeof(uint4), "8 elements of 16 bits");
                memcpy(&q, v[p], sizeof(q));
                out[(size_t)(c * 4 + p) * 64 + l] = q;
            }
        }
    return ncp;
}

int build_mfma_taps(const std::vector<std::complex<double>>& taps, int lm, int lp, int bf16, std::vector<uint4>& out,
                    double* scale, int* glo0, int* glo1) {
    // lo parts dropped below 2^-16 (f16) / 2^-19 (bf16) of the largest tap:
    // there they are at most 2^-27 of it
    return bf16 ? build_taps<__bf16>(taps, lm, lp, out, scale, std::ldexp(1.0, -19), glo0, glo1)
                : build_taps<_Float16>(taps, lm, lp, out, scale, std::ldexp(1.0, -16), glo0, glo1);
}

// LDS elements of the staged input: outputs of the workgroup plus the
// kernel's span, as read by the last tile of the last wave.
int mfma_input_elems(int nc) { return kMOut - 32 + 16 * nc; }

// dynamic LDS: the staged input (padded), then two groups of tap operands
size_t mfma_lds_bytes(int nx) { return ((size_t)(nx + (nx >> 5) * 8 + 8) * 2 + 15) / 16 * 16 + 2 * kMGroup * 4 * 64 * 16; }
int mfma_a_lds_offset(int nx) { return (int)(((size_t)(nx + (nx >> 5) * 8 + 8) * 2 + 15) / 16 * 16); }

hipError_t run_k1_part_a_mfma(int bf16, const K1MLaunch& a, int nf, hipStream_t st) {
    const int nb = mfma_blocks_per_segment(a.spfp);
    const size_t sh = mfma_lds_bytes(a.nx);
    if (sh > 160 * 1024 || sh < (size_t)kMWaves * 2 * kEpiRow * 4 || a.a_lds_offset != mfma_a_lds_offset(a.nx) ||
        a.nc % kMGroup || a.nx % 16 || a.nx > kStagePasses * 256 * 16 || a.glo0 < 0 || a.glo0 > a.glo1 ||
        a.glo1 > a.nc / kMGroup)
        return hipErrorInvalidValue;
    const void* kern = bf16 ? (const void*)k1_part_a_mfma<__bf16> : (const void*)k1_part_a_mfma<_Float16>;
    (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (bf16) hipLaunchKernelGGL(k1_part_a_mfma<__bf16>, dim3(nb, nf), dim3(256), sh, st, a);
    else hipLaunchKernelGGL(k1_part_a_mfma<_Float16>, dim3(nb, nf), dim3(256), sh, st, a);
    return hipGetLastError();
}

}  // namespace cuvhs
