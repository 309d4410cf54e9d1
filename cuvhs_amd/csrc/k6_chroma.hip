// K6 colour-under chroma decode.
//
// Reference: chroma_decode.cu:526-1002.  The reference resolves head-track
// flips with a host loop: per 15-field sub-chunk it re-runs resample ->
// 1024-point FFT band-pass -> inverse FFT -> burst metric up to three
// times, copying one int to the host per pass, and detects the first
// track and the NTSC phase from field 0 with two more host round trips.
// PAL runs the same logic on PAL's constants (1135-sample lines, 313 lines
// per field, a 2048-point line transform, 4.43 MHz subcarrier).
//
// Here the decision logic needs only per-field burst metrics for both track
// hypotheses.  For even output lines the heterodyne phase is the same for
// both tracks; for odd lines switching track adds pi, i.e. negates the
// line.  The band-pass is linear, so one pass over all lines with track 0
// gives every band-passed line for either track up to its sign, the
// track-0 metric (|a + b|) and the track-1 metric (|a - b|) of every field.
// A single-thread kernel then replays the reference's detection, retry and
// state logic exactly, and the output kernel applies the resolved track's
// sign, the ACC gain and the u16 conversion to the stored lines.
#include <hip/hip_runtime.h>

#include "common.h"
#include "fft_lds.h"
#include "launch.h"
#include "wave.h"

namespace cuvhs {

constexpr int kSkip = 16;          // chroma_decode.cu:255

// The geometry the chroma kernels are compiled for (the launcher checks it
// against the launch's): samples per output line, burst window samples,
// output lines per field, line locations per frame, line transform points.
template <int LL_, int BL_, int FL_, int LPF_, int NF_>
struct ChromaGeo {
    static constexpr int LL = LL_, BL = BL_, FL = FL_, LPF = LPF_, NF = NF_;
    // line pairs of the burst metric (chroma_decode.cu:383-411)
    static constexpr int PAIRS = (FL_ - 2 * kSkip - 1) / 2;
};
using NtscChroma = ChromaGeo<910, 36, 263, 525, 1024>;     // 115 pairs
using PalChroma = ChromaGeo<1135, 40, 313, 625, 2048>;     // 140 pairs

// Resampled sample of one line of LL samples (chroma_decode.cu:75-137),
// before the heterodyne.
template <int LL>
__device__ __forceinline__ double chroma_resample(const ChromaLaunch& a, double ls, double le, int col,
                                                 long long total) {
    if (col >= LL) return 0.0;
    double frac = (double)col / (double)LL;
    double coord = ls + frac * (le - ls);
    int ci = (int)coord;
    double x = coord - (double)ci;
    if (ci < 1 || (long long)ci + 2 >= total) return 0.0;
    double p0 = a.craw[ci - 1], p1 = a.craw[ci], p2 = a.craw[ci + 1], p3 = a.craw[ci + 2];
    double aa = p2 - p0;
    double bb = 2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3;
    double cc = 3.0 * (p1 - p2) + p3 - p0;
    return p1 + 0.5 * x * (aa + x * (bb + x * cc));
}

// Start of the heterodyne -cos(2 pi het_scale (line * L + col) + phase) of
// chroma_decode.cu:143-148 (L the output line length) at column col = t,
// the first of the columns t + 64 r that thread t owns; het_step moves (s, c) on to the
// next of them by the fixed rotation exp(i 2 pi het_scale 64) (het_dc,
// het_ds).  The rotation differs from a direct cos by ~1e-14 relative; the
// reference's own FP64 argument, up to 4.4e5 rad, is rounded at ~5e-11 rad.
template <int LL>
__device__ __forceinline__ void het_start(const ChromaLaunch& a, int line, double phase_offset, double* s,
                                          double* c) {
    sincos(2.0 * M_PI * a.het_scale * (double)(line * LL + (int)threadIdx.x) + phase_offset, s, c);
}
__device__ __forceinline__ void het_step(const ChromaLaunch& a, double* s, double* c) {
    const double cn = *c * a.het_dc - *s * a.het_ds;
    *s = *s * a.het_dc + *c * a.het_ds;
    *c = cn;
}

// One wavefront per pair of lines (la, la + 1): resample + heterodyne with
// track 0 (chroma_decode.cu:75-157), NF-point FFT (NF = 1024 for NTSC's
// 910-sample lines, 2048 for PAL's 1135), band-pass (:163-182), inverse
// FFT; the burst window (FP64) and the line (FP32, ~1e-3 LSB of the u16
// output) are stored for the metric and output kernels.  The two real lines
// go in as the real and imaginary parts of one complex transform; the
// band-pass keeps bins by |frequency|, so the band-passed line la comes back
// as the real part and la + 1 as the imaginary part.  (The reference
// transforms each line on its own; the results agree to rounding.)  Thread t
// computes columns t + 64 r, the register layout
// filter1024_wave takes and returns, so the lines never pass through LDS
// outside the transform; rows past the line's last sample are zero and are
// neither computed nor stored.  NF = 2048 splits into even and odd bins
// first (x[n] +- x[n + 1024], the odd half times W2048^n: a radix-2 step
// inside each lane), filters each half as a 1024-point block with the bins
// it holds, and joins them the same way.  T: the transform's type.  The
// resampling coordinate and the heterodyne stay FP64 in both (positions
// reach 46 M samples, the heterodyne argument 4.4e5 rad); T = float is the
// reduced-precision forms' line filter, with whole complex elements in LDS
// (the same 8.7 KB as the FP64 filter's real and imaginary parts in turn).
template <typename T, typename G>
__device__ __forceinline__ void chroma_line(const ChromaLaunch& a, void* buf, const typename C2<T>::type* tw,
                                            const typename C2<T>::type* tw2) {
    using V = typename C2<T>::type;
    constexpr bool whole = sizeof(T) == 4;
    constexpr int NF = G::NF, LL = G::LL, BL = G::BL;
    static_assert(NF == 1024 || NF == 2048, "line transform size");
    // rows r (columns t + 64 r) that hold samples: 15 of NTSC's 16, 18 of
    // PAL's 32; the rest of the transform's input is zero
    constexpr int RD = (LL + 63) / 64;
    constexpr int RX = RD > 16 ? RD - 16 : 0;   // data rows past the first 1024 points
    static_assert(RD <= 16 || NF == 2048, "a line fits its transform");
    const int f = blockIdx.y;
    const int nf = a.nf_override > 0 ? a.nf_override : a.ctl->nf;
    if (f >= nf) return;
    const int la = a.line0 + 2 * (int)blockIdx.x;
    if (la >= a.line_end) return;
    const int lb = la + 1;
    const bool has_b = lb < a.line_end;
    const long long total = (long long)nf * a.spfp;
    const double* ll = a.linelocs + (size_t)f * G::LPF;
    const int rot = 3;  // track 0
    const int l = (int)threadIdx.x & 63;
    V v[16];        // rows 0..15
    V x2[RX > 0 ? RX : 1];   // rows 16 .. RD - 1
#pragma unroll
    for (int r = 0; r < 16; r++) {
        v[r].x = (T)0;
        v[r].y = (T)0;
    }
#pragma unroll
    for (int r = 0; r < (RX > 0 ? RX : 1); r++) {
        x2[r].x = (T)0;
        x2[r].y = (T)0;
    }
    {
        const double ls = ll[la], le = ll[la + 1];
        double s, c;
        het_start<LL>(a, la, (double)((la * rot + 0) & 3) * (M_PI * 0.5), &s, &c);
#pragma unroll
        for (int r = 0; r < RD; r++) {
            const T x = (T)(chroma_resample<LL>(a, ls, le, l + 64 * r, total) * -c);
            if (r < 16) v[r].x = x;
            else x2[r - 16].x = x;
            het_step(a, &s, &c);
        }
    }
    if (has_b) {
        const double ls = ll[lb], le = ll[lb + 1];
        double s, c;
        het_start<LL>(a, lb, (double)((lb * rot + 0) & 3) * (M_PI * 0.5), &s, &c);
#pragma unroll
        for (int r = 0; r < RD; r++) {
            const T x = (T)(chroma_resample<LL>(a, ls, le, l + 64 * r, total) * -c);
            if (r < 16) v[r].y = x;
            else x2[r - 16].y = x;
            het_step(a, &s, &c);
        }
    }
    // band-pass: bins with lo <= |k| <= hi kept (chroma_decode.cu:163-182)
    const int lo = a.lo, hi = a.hi;
    auto keep = [&](V z, int k) {
        const int kk = k <= NF / 2 ? k : NF - k;
        if (kk < lo || kk > hi) {
            z.x = (T)0;
            z.y = (T)0;
        }
        return z;
    };
    if constexpr (NF == 1024) {
        filter1024_wave<T, whole>(buf, v, tw, keep);
    } else {
        // even bins: a[n] = x[n] + x[n + 1024]; odd bins: b[n] = (x[n] -
        // x[n + 1024]) W2048^n, n = 64 r + l; x[n + 1024] is zero past the
        // RX data rows
        V w[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            V x0 = v[r], d = v[r];
            if (r < RX) {
                x0 = cadd(v[r], x2[r]);
                d = csub(v[r], x2[r]);
            }
            v[r] = x0;
            w[r] = cmul(d, tw2[64 * r + l]);
        }
        filter1024_wave<T, whole>(buf, v, tw, [&](V z, int k) { return keep(z, 2 * k); });
        filter1024_wave<T, whole>(buf, w, tw, [&](V z, int k) { return keep(z, 2 * k + 1); });
        // y[n] = a'[n] + W2048^-n b'[n], y[n + 1024] = a'[n] - W2048^-n b'[n]
        // (the latter only where the line has samples)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const V o = cmul(w[r], cconj(tw2[64 * r + l]));
            if (r < RX) x2[r] = csub(v[r], o);
            v[r] = cadd(v[r], o);
        }
    }
    const T scale = (T)a.fft_scale;
    double* da = a.burst + ((size_t)f * G::FL + la) * BL;
    float* fa = a.filt + ((size_t)f * G::FL + la) * LL;
#pragma unroll
    for (int r = 0; r < RD; r++) {
        const int col = l + 64 * r;
        const V y = r < 16 ? v[r] : x2[r - 16];
        const T xa = y.x * scale, xb = y.y * scale;
        if (col < LL) {
            fa[col] = (float)xa;
            if (has_b) fa[LL + col] = (float)xb;
        }
        const int bi = col - a.burst_start;
        if (bi >= 0 && bi < BL) {
            da[bi] = (double)xa;
            if (has_b) da[BL + bi] = (double)xb;
        }
    }
}

__global__ __launch_bounds__(64, 3) void k_chroma_line(ChromaLaunch a) {
    __shared__ __align__(16) double buf[lds_elems(1024)];
    chroma_line<double, NtscChroma>(a, buf, a.tw, nullptr);
}
__global__ __launch_bounds__(64, 4) void k_chroma_line_f32(ChromaLaunch a) {
    __shared__ __align__(16) float2 buf[lds_elems(1024)];
    chroma_line<float, NtscChroma>(a, buf, a.tw_f, nullptr);
}
// PAL: 2048-point line transforms (twice the registers per lane)
__global__ __launch_bounds__(64) void k_chroma_line2048(ChromaLaunch a) {
    __shared__ __align__(16) double buf[lds_elems(1024)];
    chroma_line<double, PalChroma>(a, buf, a.tw, a.tw2);
}
__global__ __launch_bounds__(64, 2) void k_chroma_line2048_f32(ChromaLaunch a) {
    __shared__ __align__(16) float2 buf[lds_elems(1024)];
    chroma_line<float, PalChroma>(a, buf, a.tw_f, a.tw2_f);
}

// ACC + u16 of every line (chroma_decode.cu:188-231) with the resolved track:
// an odd line of track 1 is the negated track-0 line.  One workgroup per
// (line, field).
template <typename G>
__global__ __launch_bounds__(256) void k_chroma_out(ChromaLaunch a) {
    const int f = blockIdx.y, line = blockIdx.x;
    const int nf = a.nf_override > 0 ? a.nf_override : a.ctl->nf;
    if (f >= nf) return;
    const size_t row = (size_t)f * G::FL + line;
    // burst RMS (chroma_decode.cu:205-221, whose atomic sum has no fixed
    // order; here, as in the oracle, sample order): every wavefront loads the
    // 36 (NTSC) or 40 (PAL) burst samples at once and sums their squares
    // through lane reads, so no lane waits on a workgroup barrier
    const int lane = (int)threadIdx.x & 63;
    constexpr int bl = G::BL;
    const double* da = a.burst + row * bl;
    const double bv = lane < bl ? da[lane] : 0.0;
    double ssq = 0.0;
    for (int i = 0; i < bl; i++) {
        const double d = lane_value(bv, i);
        ssq += d * d;
    }
    const double rms = sqrt(ssq / (double)bl);
    const double s_acc = (rms > 1.0) ? (a.burst_abs_ref / rms) : 0.0;
    const double acc = ((line & 1) && a.track[f]) ? -s_acc : s_acc;
    const float* fa = a.filt + row * G::LL;
    uint16_t* o = (a.plan ? a.plan->chroma : a.out) + row * G::LL;
    for (int col = threadIdx.x; col < G::LL; col += 256) {
        double ov = (double)fa[col] * acc + 32768.0;
        if (ov < 0.0) ov = 0.0;
        if (ov > 65535.0) ov = 65535.0;
        o[col] = (uint16_t)(ov + 0.5);
    }
}

// Per-field burst-cancellation metrics for both tracks and, for field 0,
// per-line burst I/Q.  One wavefront per field: lanes evaluate pairs, lane 0
// accumulates them in pair order (chroma_decode.cu:383-411, :415-450).
template <typename G>
__global__ __launch_bounds__(64) void k_chroma_metrics(const double* burst, const ChunkCtl* ctl, int nf_override,
                                                       double* m0, double* m1, const double* cs_tab, double* iq0) {
    constexpr int field_lines = G::FL, bl = G::BL, pairs = G::PAIRS;
    __shared__ double s0[pairs], s1[pairs];
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    const int f = blockIdx.x;
    if (f >= nf) return;
    for (int p = threadIdx.x; p < pairs; p += 64) {
        const double* ra = burst + ((size_t)f * field_lines + kSkip + 2 * p) * bl;
        const double* rb = ra + bl;
        double a0 = 0.0, a1 = 0.0;
        for (int i = 0; i < bl; i++) {
            a0 += fabs(ra[i] + rb[i]);
            a1 += fabs(ra[i] + (-rb[i]));
        }
        s0[p] = a0 / bl;
        s1[p] = a1 / bl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
        for (int p = 0; p < pairs; p++) { t0 += s0[p]; t1 += s1[p]; }
        m0[f] = t0 / pairs;
        m1[f] = t1 / pairs;
    }
    if (f == 0 && iq0) {
        for (int line = kSkip + threadIdx.x; line < field_lines - kSkip; line += 64) {
            const double* r = burst + (size_t)line * bl;
            double I = 0.0, Q = 0.0;
            for (int i = 0; i < bl; i++) {
                I += r[i] * cs_tab[2 * i];
                Q += r[i] * cs_tab[2 * i + 1];
            }
            iq0[2 * line] = I;
            iq0[2 * line + 1] = Q;
        }
    }
}

__device__ int ntsc_phase_id_first(int quadrant) {
    // lookup_ntsc_phase(1, quadrant, -1): chroma_decode.cu:337-377
    switch (quadrant) {
        case 2: return 3;
        case 1: return 1;
        case 0: return 3;
        case 3: return 1;
    }
    return 1;
}

// Detection, retry replay and state update (chroma_decode.cu:674-1001).  The
// logic is sequential (one thread); for chunks of up to kResolveLds fields
// the metrics are staged in LDS and the tracks built there, so its many
// dependent reads and rewrites wait on LDS rather than on L2.
constexpr int kResolveLds = 1024;
__device__ void resolve_tracks(DecodeState* dst, int nf, const double* m0, const double* m1, const double* iq0,
                               int* out_track, int* phase_ids, int field_lines);
__global__ void k_chroma_resolve(const ChunkCtl* ctl, DecodeState* dst, int nf_override, const double* g_m0,
                                 const double* g_m1, const double* iq0, int* g_track, int* phase_ids,
                                 const ChunkPlan* plan, int field_lines) {
    __shared__ double s_m0[kResolveLds], s_m1[kResolveLds];
    if (plan) phase_ids = plan->phase_id;   // a replayed graph: the chunk's place in the outputs
    __shared__ int s_track[kResolveLds];
    if (blockIdx.x != 0) return;
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    if (nf <= 0) return;
    const bool staged = nf <= kResolveLds;
    if (staged) {
        for (int i = threadIdx.x; i < nf; i += blockDim.x) {
            s_m0[i] = g_m0[i];
            s_m1[i] = g_m1[i];
        }
    }
    __syncthreads();
    const double* m0 = staged ? s_m0 : g_m0;
    const double* m1 = staged ? s_m1 : g_m1;
    int* out_track = staged ? s_track : g_track;
    if (threadIdx.x == 0) resolve_tracks(dst, nf, m0, m1, iq0, out_track, phase_ids, field_lines);
    __syncthreads();
    if (staged)
        for (int i = threadIdx.x; i < nf; i += blockDim.x) g_track[i] = s_track[i];
}

__device__ void resolve_tracks(DecodeState* dst, int nf, const double* m0, const double* m1, const double* iq0,
                               int* out_track, int* phase_ids, int field_lines) {
    ChromaState cs = dst->chroma;
    const int cycle[8] = {3, 2, 1, 4, 3, 2, 1, 4};
    int det, cycle_start;
    double thr;
    if (cs.valid) {
        det = cs.current_track;
        thr = cs.good_metric_threshold;
        cycle_start = cs.cycle_start;
    } else {
        double metric[2] = {m0[0], m1[0]};
        det = (metric[1] < metric[0]) ? 1 : 0;
        thr = metric[det] * 4.0;
        double It = 0.0, Qt = 0.0;
        for (int line = kSkip; line < field_lines - kSkip; line++) {
            double I = iq0[2 * line], Q = iq0[2 * line + 1];
            if (det == 1 && (line & 1)) { I = -I; Q = -Q; }
            double mag = sqrt(I * I + Q * Q);
            if (mag > 0.0) { It += I / mag; Qt += Q / mag; }
        }
        double ph = atan2(Qt, It);
        double deg = fmod(ph * 180.0 / M_PI + 360.0, 360.0);
        int q0 = ((int)(deg / 90.0 + 0.5)) % 4;
        int pid = ntsc_phase_id_first(q0);
        cycle_start = 0;
        for (int i = 0; i < 8; i += 2)
            if (cycle[i] == pid) { cycle_start = i; break; }
    }
    int current = det;
    for (int f = 0; f < nf; f++) {
        out_track[f] = (f & 1) ? (1 - current) : current;
        phase_ids[f] = cycle[(cycle_start + f) % 8];
    }
    // out_track doubles as h_track; the tracks of the last computed pass of
    // each sub-chunk are what the output uses.
    int fpc = (4096 / field_lines) * field_lines;
    int tl = nf * field_lines;
    int cl = tl < fpc ? tl : fpc;
    fpc = cl / field_lines;
    for (int fs = 0; fs < nf; fs += fpc) {
        int fthis = min(fpc, nf - fs);
        int used[16];
        int retries = 3;
        bool ok = false;
        while (!ok && retries-- > 0) {
            for (int i = 0; i < fthis; i++) used[i] = out_track[fs + i];
            int first_bad = -1;
            for (int i = 0; i < fthis; i++) {
                double m = used[i] ? m1[fs + i] : m0[fs + i];
                if (m > thr) { first_bad = i; break; }
            }
            if (first_bad < 0) {
                ok = true;
            } else {
                current = 1 - current;
                for (int ff = fs + first_bad; ff < nf; ff++) out_track[ff] = (ff & 1) ? (1 - current) : current;
            }
        }
        // output uses the tracks of the last pass; later fields keep the flip
        for (int i = 0; i < fthis; i++) out_track[fs + i] = used[i];
        if (!ok) {
            // the flip of the final pass applies to the following sub-chunks only
            for (int ff = fs + fthis; ff < nf; ff++) out_track[ff] = (ff & 1) ? (1 - current) : current;
        }
    }
    cs.valid = 1;
    cs.current_track = (nf & 1) ? (1 - current) : current;
    cs.good_metric_threshold = thr;
    cs.cycle_start = (cycle_start + nf) % 8;
    dst->chroma = cs;
}

hipError_t run_chroma(ChromaLaunch a, DecodeState* dst, int max_fields, double* m0, double* m1,
                      const double* cs_tab, double* iq0, int* out_track, int* phase_ids, int need_iq,
                      hipStream_t st) {
    // band-pass of every line with track 0; metrics from lines 16 .. field
    // lines - 17.  The kernels are compiled for NTSC's and PAL's geometry.
    auto is = [&](auto g) {
        using G = decltype(g);
        return a.nfft == G::NF && a.line_len == G::LL && a.burst_len == G::BL && a.field_lines == G::FL &&
               a.lpf == G::LPF;
    };
    const bool pal = is(PalChroma{});
    if (!pal && !is(NtscChroma{})) return hipErrorInvalidValue;
    ChromaLaunch m = a;
    m.line0 = 0;
    m.line_end = a.field_lines;
    const dim3 lg((m.line_end - m.line0 + 1) / 2, max_fields);
    if (pal) {
        if (a.f32) hipLaunchKernelGGL(k_chroma_line2048_f32, lg, dim3(64), 0, st, m);
        else hipLaunchKernelGGL(k_chroma_line2048, lg, dim3(64), 0, st, m);
        hipLaunchKernelGGL(k_chroma_metrics<PalChroma>, dim3(max_fields), dim3(64), 0, st, a.burst, a.ctl,
                           a.nf_override, m0, m1, cs_tab, need_iq ? iq0 : nullptr);
    } else {
        if (a.f32) hipLaunchKernelGGL(k_chroma_line_f32, lg, dim3(64), 0, st, m);
        else hipLaunchKernelGGL(k_chroma_line, lg, dim3(64), 0, st, m);
        hipLaunchKernelGGL(k_chroma_metrics<NtscChroma>, dim3(max_fields), dim3(64), 0, st, a.burst, a.ctl,
                           a.nf_override, m0, m1, cs_tab, need_iq ? iq0 : nullptr);
    }
    hipLaunchKernelGGL(k_chroma_resolve, dim3(1), dim3(64), 0, st, a.ctl, dst, a.nf_override, m0, m1, iq0, out_track,
                       phase_ids, a.plan, a.field_lines);
    ChromaLaunch o = a;
    o.track = out_track;
    if (pal) hipLaunchKernelGGL(k_chroma_out<PalChroma>, dim3(a.field_lines, max_fields), dim3(256), 0, st, o);
    else hipLaunchKernelGGL(k_chroma_out<NtscChroma>, dim3(a.field_lines, max_fields), dim3(256), 0, st, o);
    return hipGetLastError();
}

}  // namespace cuvhs
