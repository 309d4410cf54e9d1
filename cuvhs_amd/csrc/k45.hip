// K4 hsync refinement and K5 TBC luma resampling.
//
// Reference: hsync_refine.cu:21-271, tbc_resample.cu:19-186.  Both are
// gather kernels bounded by HBM/L2 traffic; the window and rounding rules
// (C truncation of the window starts, (int)(x + 0.5) rounding of u16) are
// kept exactly because they decide which samples are read.
#include <hip/hip_runtime.h>

#include "common.h"
#include "launch.h"
#include "wave.h"

namespace cuvhs {

// Wavefront helpers: all 64 lanes of a wavefront cooperate on one line; every
// lane returns the same value.
template <typename S>
__device__ __forceinline__ double wv_find_falling(const S* sig, int start, int count, double target, int total,
                                                  int lane) {
    if (start < 0 || start + count > total || count < 2) return -1.0;
    for (int b = start + 1; b < start + count; b += 64) {
        int i = b + lane;
        bool c = false;
        if (i < start + count) c = (double)sig[i - 1] >= target && (double)sig[i] < target;
        unsigned long long m = __ballot(c);
        if (m) {
            int f = b + __ffsll((long long)m) - 1;   // first crossing (hsync_refine.cu:29-39)
            double a = (double)sig[f - 1] - target, bb = (double)sig[f] - target;
            double frac = (a - bb != 0.0) ? (a / (a - bb)) : 0.0;
            return (double)(f - 1) + frac;
        }
    }
    return -1.0;
}

// quirk: the reference's interpolation guard (a + b != 0, hsync_refine.cu:
// 63-64), which returns the sample itself when the two samples lie
// symmetrically about the target.  In FP64 that has measure zero; on the
// reduced-precision forms' FP32-quantized demod_05 it happens whenever the
// target is representable (PAL's 3,960,000 Hz threshold) and the samples
// straddle it by equal steps, a half-sample jump the FP64 values never take.
// Those forms therefore interpolate without it (quirk = false), the
// continuous value the FP64 form gives.
template <typename S>
__device__ __forceinline__ double wv_find_rising(const S* sig, int start, int count, double target, int total,
                                                 int lane, bool quirk = true) {
    if (start < 0 || start + count > total || count < 2) return -1.0;
    if ((double)sig[start] > target) return -1.0;
    for (int b = start + 1; b < start + count; b += 64) {
        int i = b + lane;
        bool c = false;
        if (i < start + count) c = (double)sig[i - 1] < target && (double)sig[i] >= target;
        unsigned long long m = __ballot(c);
        if (m) {
            int f = b + __ffsll((long long)m) - 1;   // hsync_refine.cu:57-67
            double a = target - (double)sig[f - 1], bb = target - (double)sig[f];
            double frac = (quirk ? (a + bb != 0.0) : (a - bb != 0.0)) ? (a / (a - bb)) : 0.0;
            return (double)(f - 1) + frac;
        }
    }
    return -1.0;
}

// The n/2-th smallest of up to 64 samples (hsync_refine.cu:72-97): lane j
// holds sample j and counts its stable rank against the window read back
// from the wavefront's LDS slots as broadcasts, two samples per read.
template <typename S>
__device__ __forceinline__ double wv_window_median(const S* sig, int start, int count, int total, int lane,
                                                   double* wl) {
    if (start < 0 || start + count > total || count <= 0) return 0.0;
    const int n = count < 64 ? count : 64;
    const int mid = n / 2;
    double v = lane < n ? (double)sig[start + lane] : 0.0;
    wl[lane] = v;
    wave_sync();
    int r = 0;
    for (int k = 0; k < n; k += 2) {
        const double2 p = *reinterpret_cast<const double2*>(wl + k);
        r += (p.x < v) || (p.x == v && k < lane);
        if (k + 1 < n) r += (p.y < v) || (p.y == v && k + 1 < lane);
    }
    unsigned long long m = __ballot(lane < n && r == mid);
    const double res = lane_value(v, __ffsll((long long)m) - 1);
    wave_sync();  // the reads finish before the wavefront's next window is stored
    return res;
}

// One wavefront per line (hsync_refine.cu:99-239).  S: storage type of
// demod_05 (double, or float in the reduced-precision forms); the crossings,
// levels and positions are computed in FP64.
template <typename S>
__global__ __launch_bounds__(256) void k_hsync_refine(const S* __restrict__ d05, double* __restrict__ ll,
                                                      const ChunkCtl* ctl, int nf_override, int total_per_field,
                                                      double thr, double one_usec_samples, int als, double hw,
                                                      double right_edge_offset, int lpf, int quirk) {
    __shared__ __align__(16) double s_win[4][64];
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    const int lane = threadIdx.x & 63;
    double* wl = s_win[threadIdx.x >> 6];
    const int idx = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (idx >= nf * lpf) return;
    const int line = idx % lpf;
    if (line < als) return;
    const int total = nf * total_per_field;
    const double original = ll[idx];
    const int one_usec = (int)(one_usec_samples + 0.5);
    const int search_start = (int)(original - one_usec);
    const int search_count = one_usec * 2;
    double zc = wv_find_falling(d05, search_start, search_count, thr, total, lane);
    double left = -1.0;
    if (zc >= 0.0) {
        int zci = (int)(zc + 0.5);
        double sync_level = wv_window_median(d05, zci + (int)(1.0 * one_usec_samples), (int)(1.5 * one_usec_samples),
                                             total, lane, wl);
        double porch_level = wv_window_median(d05, zci + (int)(8.0 * one_usec_samples), (int)(1.0 * one_usec_samples),
                                              total, lane, wl);
        if (porch_level > sync_level) {
            double mid = (sync_level + porch_level) / 2.0;
            int rs = (int)(original - one_usec * 2);
            int rc = (int)(one_usec_samples * 4);
            double zc2 = wv_find_falling(d05, rs, rc, mid, total, lane);
            if (zc2 >= 0.0 && fabs(zc2 - zc) <= one_usec_samples * 0.5) left = zc2;
        }
    }
    double right = -1.0;
    const int rs0 = (int)(original + hw - one_usec_samples);
    const int rc0 = (int)(hw * 2);
    double rcross = wv_find_rising(d05, rs0, rc0, thr, total, lane, quirk != 0);
    if (rcross >= 0.0) {
        double zc_fr = rcross - hw;
        int zfi = (int)(zc_fr + 0.5);
        double rsync = wv_window_median(d05, zfi + (int)(1.0 * one_usec_samples), (int)(1.5 * one_usec_samples),
                                        total, lane, wl);
        double rporch = wv_window_median(d05, zfi + (int)(hw + 1.0 * one_usec_samples), (int)(1.0 * one_usec_samples),
                                         total, lane, wl);
        if (rporch > rsync) {
            double rmid = (rsync + rporch) / 2.0;
            double zc2r = wv_find_rising(d05, rs0, rc0, rmid, total, lane, quirk != 0);
            if (zc2r >= 0.0 && fabs(zc2r - rcross) <= one_usec_samples * 0.5) {
                double cand = rcross - hw + right_edge_offset;
                double ref = (left >= 0.0) ? left : original;
                if (fabs(cand - ref) <= one_usec_samples * 2.0) right = cand;
            }
        }
    }
    if (lane == 0) {
        if (right >= 0.0) ll[idx] = right;
        else if (left >= 0.0) ll[idx] = left;
    }
}

// One thread per output pixel; consecutive threads walk a TBC row, so the
// four demod gathers of neighbouring pixels overlap in cache.  S: storage
// and interpolation type of demod (double, or float in the reduced-precision
// forms); the sampling coordinate is FP64 (positions reach 46 M samples).
template <typename S>
__global__ __launch_bounds__(256) void k_tbc_resample(const S* __restrict__ demod, const double* __restrict__ ll,
                                                      uint16_t* __restrict__ luma, const ChunkCtl* ctl,
                                                      int nf_override, int total_per_field, double ire0,
                                                      double hz_ire, double vsync_ire, double out_zero,
                                                      double out_scale, const ChunkPlan* plan, int lpf,
                                                      int field_lines, int line_len) {
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    if (plan) luma = plan->luma;   // a replayed graph: the chunk's place in the outputs
    const int ppf = field_lines * line_len;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nf * ppf) return;
    const int total = nf * total_per_field;
    int field = idx / ppf;
    int rem = idx - field * ppf;
    int row = rem / line_len;
    int col = rem - row * line_len;
    // row + 1 < lines per frame always holds (263 of 525 rows, 313 of 625)
    double ls = ll[field * lpf + row];
    double le = ll[field * lpf + row + 1];
    double len = le - ls;
    double frac = (double)col / (double)line_len;
    double coord = ls + frac * len;
    int ci = (int)coord;
    double x = coord - (double)ci;
    if (ci < 1 || ci + 2 >= total) {
        luma[idx] = (uint16_t)out_zero;
        return;
    }
    const S p0 = demod[ci - 1], p1 = demod[ci], p2 = demod[ci + 1], p3 = demod[ci + 2];
    const S xs = (S)x;
    const S a = p2 - p0;
    const S b = (S)2 * p0 - (S)5 * p1 + (S)4 * p2 - p3;
    const S c = (S)3 * (p1 - p2) + p3 - p0;
    const S hz = p1 + (S)0.5 * xs * (a + xs * (b + xs * c));
    const S ire = (hz - (S)ire0) / (S)hz_ire - (S)vsync_ire;
    S v = ire * (S)out_scale + (S)out_zero;
    if (v < (S)0) v = (S)0;
    if (v > (S)65535) v = (S)65535;
    luma[idx] = (uint16_t)(v + (S)0.5);
}

hipError_t run_hsync_refine(const void* d05, int f32, double* ll, const ChunkCtl* ctl, int nf_override,
                            int max_fields, const Fmt& fmt, hipStream_t st, int reduced) {
    const int quirk = reduced ? 0 : 1;
    int n = max_fields * fmt.lines_per_frame;
    double ous = 1.0e-6 * fmt.sample_rate;
    double reo = 2.25 * ((fmt.sample_rate / 1.0e6) / 40.0);
    if (f32)
        hipLaunchKernelGGL(k_hsync_refine<float>, dim3((n + 3) / 4), dim3(256), 0, st, static_cast<const float*>(d05),
                           ll, ctl, nf_override, fmt.spfp, fmt.pulse_threshold_hz, ous, fmt.active_line_start,
                           fmt.hsync_width, reo, fmt.lines_per_frame, quirk);
    else
        hipLaunchKernelGGL(k_hsync_refine<double>, dim3((n + 3) / 4), dim3(256), 0, st,
                           static_cast<const double*>(d05), ll, ctl, nf_override, fmt.spfp, fmt.pulse_threshold_hz,
                           ous, fmt.active_line_start, fmt.hsync_width, reo, fmt.lines_per_frame, quirk);
    return hipGetLastError();
}

hipError_t run_tbc_resample(const void* demod, int f32, const double* ll, uint16_t* luma, const ChunkCtl* ctl,
                            int nf_override, int max_fields, const Fmt& fmt, hipStream_t st, const ChunkPlan* plan) {
    long long n = (long long)max_fields * fmt.field_lines * fmt.line_len;
    const dim3 grid((unsigned)((n + 255) / 256));
    if (f32)
        hipLaunchKernelGGL(k_tbc_resample<float>, grid, dim3(256), 0, st, static_cast<const float*>(demod), ll, luma,
                           ctl, nf_override, fmt.spfp, fmt.ire0, fmt.hz_ire, fmt.vsync_ire, fmt.output_zero,
                           fmt.output_scale, plan, fmt.lines_per_frame, fmt.field_lines, fmt.line_len);
    else
        hipLaunchKernelGGL(k_tbc_resample<double>, grid, dim3(256), 0, st, static_cast<const double*>(demod), ll,
                           luma, ctl, nf_override, fmt.spfp, fmt.ire0, fmt.hz_ire, fmt.vsync_ire, fmt.output_zero,
                           fmt.output_scale, plan, fmt.lines_per_frame, fmt.field_lines, fmt.line_len);
    return hipGetLastError();
}

}  // namespace cuvhs
