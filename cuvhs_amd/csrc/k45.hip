// K4 hsync refinement and K5 TBC luma resampling.
//
// Reference: hsync_refine.cu:21-271, tbc_resample.cu:19-186.  Both are
// gather kernels bounded by HBM/L2 traffic; the window and rounding rules
// (C truncation of the window starts, (int)(x + 0.5) rounding of u16) are
// kept exactly because they decide which samples are read.
#include <hip/hip_runtime.h>

#include "common.h"
#include "launch.h"

namespace cuvhs {

__device__ __forceinline__ double find_falling(const double* sig, int start, int count, double target, int total) {
    if (start < 0 || start + count > total || count < 2) return -1.0;
    double prev = sig[start];
    for (int i = start + 1; i < start + count; i++) {
        double curr = sig[i];
        if (prev >= target && curr < target) {
            double a = prev - target, b = curr - target;
            double frac = (a - b != 0.0) ? (a / (a - b)) : 0.0;
            return (double)(i - 1) + frac;
        }
        prev = curr;
    }
    return -1.0;
}

__device__ __forceinline__ double find_rising(const double* sig, int start, int count, double target, int total) {
    if (start < 0 || start + count > total || count < 2) return -1.0;
    if (sig[start] > target) return -1.0;
    double prev = sig[start];
    for (int i = start + 1; i < start + count; i++) {
        double curr = sig[i];
        if (prev < target && curr >= target) {
            double a = target - prev, b = target - curr;
            double frac = (a + b != 0.0) ? (a / (a - b)) : 0.0;
            return (double)(i - 1) + frac;
        }
        prev = curr;
    }
    return -1.0;
}

// Median of up to 64 samples by partial selection (hsync_refine.cu:72-97).
__device__ __forceinline__ double window_median(const double* sig, int start, int count, int total) {
    if (start < 0 || start + count > total || count <= 0) return 0.0;
    double buf[64];
    int n = count < 64 ? count : 64;
    for (int i = 0; i < n; i++) buf[i] = sig[start + i];
    int mid = n / 2;
    for (int i = 0; i <= mid; i++)
        for (int j = i + 1; j < n; j++)
            if (buf[j] < buf[i]) { double t = buf[i]; buf[i] = buf[j]; buf[j] = t; }
    return buf[mid];
}

__global__ __launch_bounds__(256) void k_hsync_refine(const double* __restrict__ d05, double* __restrict__ ll,
                                                      const ChunkCtl* ctl, int nf_override, int total_per_field,
                                                      double thr, double one_usec_samples, int als, double hw,
                                                      double right_edge_offset) {
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nf * kLinesPerFrame) return;
    int line = idx % kLinesPerFrame;
    if (line < als) return;
    const int total = nf * total_per_field;
    double original = ll[idx];
    int one_usec = (int)(one_usec_samples + 0.5);
    int search_start = (int)(original - one_usec);
    int search_count = one_usec * 2;
    double zc = find_falling(d05, search_start, search_count, thr, total);
    double left = -1.0;
    if (zc >= 0.0) {
        int zci = (int)(zc + 0.5);
        double sync_level = window_median(d05, zci + (int)(1.0 * one_usec_samples), (int)(1.5 * one_usec_samples), total);
        double porch_level = window_median(d05, zci + (int)(8.0 * one_usec_samples), (int)(1.0 * one_usec_samples), total);
        if (porch_level > sync_level) {
            double mid = (sync_level + porch_level) / 2.0;
            int rs = (int)(original - one_usec * 2);
            int rc = (int)(one_usec_samples * 4);
            double zc2 = find_falling(d05, rs, rc, mid, total);
            if (zc2 >= 0.0 && fabs(zc2 - zc) <= one_usec_samples * 0.5) left = zc2;
        }
    }
    double right = -1.0;
    int rs0 = (int)(original + hw - one_usec_samples);
    int rc0 = (int)(hw * 2);
    double rcross = find_rising(d05, rs0, rc0, thr, total);
    if (rcross >= 0.0) {
        double zc_fr = rcross - hw;
        int zfi = (int)(zc_fr + 0.5);
        double rsync = window_median(d05, zfi + (int)(1.0 * one_usec_samples), (int)(1.5 * one_usec_samples), total);
        double rporch = window_median(d05, zfi + (int)(hw + 1.0 * one_usec_samples), (int)(1.0 * one_usec_samples), total);
        if (rporch > rsync) {
            double rmid = (rsync + rporch) / 2.0;
            double zc2r = find_rising(d05, rs0, rc0, rmid, total);
            if (zc2r >= 0.0 && fabs(zc2r - rcross) <= one_usec_samples * 0.5) {
                double cand = rcross - hw + right_edge_offset;
                double ref = (left >= 0.0) ? left : original;
                if (fabs(cand - ref) <= one_usec_samples * 2.0) right = cand;
            }
        }
    }
    if (right >= 0.0) ll[idx] = right;
    else if (left >= 0.0) ll[idx] = left;
}

// One thread per output pixel; consecutive threads walk a TBC row, so the
// four demod gathers of neighbouring pixels overlap in cache.
__global__ __launch_bounds__(256) void k_tbc_resample(const double* __restrict__ demod, const double* __restrict__ ll,
                                                      uint16_t* __restrict__ luma, const ChunkCtl* ctl,
                                                      int nf_override, int total_per_field, double ire0,
                                                      double hz_ire, double vsync_ire, double out_zero,
                                                      double out_scale) {
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    const int ppf = kFieldLines * kLineLen;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nf * ppf) return;
    const int total = nf * total_per_field;
    int field = idx / ppf;
    int rem = idx - field * ppf;
    int row = rem / kLineLen;
    int col = rem - row * kLineLen;
    // row + 1 < 525 always holds for 263 rows
    double ls = ll[field * kLinesPerFrame + row];
    double le = ll[field * kLinesPerFrame + row + 1];
    double len = le - ls;
    double frac = (double)col / (double)kLineLen;
    double coord = ls + frac * len;
    int ci = (int)coord;
    double x = coord - (double)ci;
    if (ci < 1 || ci + 2 >= total) {
        luma[idx] = (uint16_t)out_zero;
        return;
    }
    double p0 = demod[ci - 1], p1 = demod[ci], p2 = demod[ci + 1], p3 = demod[ci + 2];
    double a = p2 - p0;
    double b = 2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3;
    double c = 3.0 * (p1 - p2) + p3 - p0;
    double hz = p1 + 0.5 * x * (a + x * (b + x * c));
    double ire = (hz - ire0) / hz_ire - vsync_ire;
    double v = ire * out_scale + out_zero;
    if (v < 0.0) v = 0.0;
    if (v > 65535.0) v = 65535.0;
    luma[idx] = (uint16_t)(v + 0.5);
}

hipError_t run_hsync_refine(const double* d05, double* ll, const ChunkCtl* ctl, int nf_override, int max_fields,
                            const Fmt& fmt, hipStream_t st) {
    int n = max_fields * kLinesPerFrame;
    double ous = 1.0e-6 * fmt.sample_rate;
    double reo = 2.25 * ((fmt.sample_rate / 1.0e6) / 40.0);
    hipLaunchKernelGGL(k_hsync_refine, dim3((n + 255) / 256), dim3(256), 0, st, d05, ll, ctl, nf_override, fmt.spfp,
                       fmt.pulse_threshold_hz, ous, fmt.active_line_start, fmt.hsync_width, reo);
    return hipGetLastError();
}

hipError_t run_tbc_resample(const double* demod, const double* ll, uint16_t* luma, const ChunkCtl* ctl,
                            int nf_override, int max_fields, const Fmt& fmt, hipStream_t st) {
    long long n = (long long)max_fields * kFieldLines * kLineLen;
    hipLaunchKernelGGL(k_tbc_resample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, demod, ll, luma, ctl,
                       nf_override, fmt.spfp, fmt.ire0, fmt.hz_ire, fmt.vsync_ire, fmt.output_zero, fmt.output_scale);
    return hipGetLastError();
}

}  // namespace cuvhs
