// K3 line location (the reference's alternative to K2b, chosen there with
// CUVHS_K2B_DISABLE) and the K3/K4/K5 diagnostic counters of its debug logs.
//
// Reference: line_locs.cu:18-476 (K3), line_locs.cu:585-653 (K3 spacing
// counters), hsync_refine.cu:273-349 (K4 counters), tbc_resample.cu:75-116
// (K5 counters).  K3 is a per-field sequential scan in the reference (one
// thread per field); here one wavefront per field: the pulse classification
// and the width sum run across the lanes, the state machine, marker
// clustering and line snapping run on lane 0 over LDS copies of the field's
// pulses.  Every floating-point expression keeps the reference's operation
// order; the width and spacing sums add integers, so their order is free.
#include <hip/hip_runtime.h>

#include "common.h"
#include "launch.h"

namespace cuvhs {

struct K3Params {
    double tol, hsync_nom, eq_nom, vsync_nom, inlinelen, spf;
    double hsync_min, hsync_max, eq_min, eq_max, vsync_min, vsync_max;  // nominal classes
    int lpf, field_lines, line_len;     // line locations per frame, output lines and samples
    int ntsc;                           // field parity rule (line_locs.cu:219-235)
    int avs, ave;                       // active output columns (tbc_resample.cu:156-157)
};

K3Params k3_params(const Fmt& f) {
    K3Params p;
    p.tol = 0.5e-6 * f.sample_rate;
    p.hsync_nom = f.hsync_width;
    p.eq_nom = f.eq_pulse_width;
    p.vsync_nom = f.vsync_width;
    p.inlinelen = (double)f.samples_per_line;
    p.spf = (double)f.samples_per_field;
    p.hsync_min = f.hsync_width - p.tol;
    p.hsync_max = f.hsync_width + p.tol;
    p.eq_min = f.eq_pulse_width - p.tol;
    p.eq_max = f.eq_pulse_width + p.tol;
    p.vsync_min = f.vsync_width * 0.5;
    p.vsync_max = f.vsync_width + 1.0e-6 * f.sample_rate;
    p.lpf = f.lines_per_frame;
    p.field_lines = f.field_lines;
    p.line_len = f.line_len;
    p.ntsc = f.system == kSystemNtsc;
    p.avs = f.active_video_start;
    p.ave = f.active_video_end;
    return p;
}

__device__ __forceinline__ int k3_class(double len, double vmin, double vmax, double hmin, double hmax, double emin,
                                        double emax) {
    if (len >= vmin && len <= vmax) return kPulseVsync;
    if (len >= hmin && len <= hmax) return kPulseHsync;
    if (len >= emin && len <= emax) return kPulseEq1;
    return kPulseHsync;
}

// Phase 1 (k_classify_pulses, line_locs.cu:18-48): nominal width classes.
__global__ void k_k3_classify(const int* lens, const int* count, const ChunkCtl* ctl, int nf_override, K3Params p,
                              int* types) {
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    const int f = blockIdx.y;
    if (f >= nf) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count[f]; i += gridDim.x * blockDim.x)
        types[(size_t)f * kMaxPulses + i] = k3_class((double)lens[(size_t)f * kMaxPulses + i], p.vsync_min,
                                                     p.vsync_max, p.hsync_min, p.hsync_max, p.eq_min, p.eq_max);
}

// Phases 2 and 3 (k_compute_linelocs, line_locs.cu:60-476).
__global__ __launch_bounds__(64) void k_k3_field(const int* starts_g, const int* lens_g, const int* count,
                                                 const ChunkCtl* ctl, int nf_override, K3Params p, double* linelocs,
                                                 int* is_first) {
    __shared__ int st[kMaxPulses], ln[kMaxPulses], ty[kMaxPulses];
    __shared__ int ms[kMaxPulses], ml[kMaxPulses], hs[kMaxPulses];
    __shared__ long long s_sum;
    __shared__ int s_cnt;
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    const int f = blockIdx.x;
    if (f >= nf) return;
    const int lane = threadIdx.x;
    const int npc = count[f];
    const int n = npc < kMaxPulses ? npc : kMaxPulses;
    // Step 0: adaptive classification (:96-137)
    const double generous_hsync_min = p.hsync_nom - 3.5 * p.tol;
    const double generous_hsync_max = p.hsync_nom + 4.0 * p.tol;
    long long sum = 0;
    int cnt = 0;
    for (int i = lane; i < n; i += 64) {
        const int s = starts_g[(size_t)f * kMaxPulses + i], l = lens_g[(size_t)f * kMaxPulses + i];
        st[i] = s;
        ln[i] = l;
        const double len = (double)l;
        if (len >= generous_hsync_min && len <= generous_hsync_max) {
            sum += l;
            cnt++;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_down(sum, off, 64);
        cnt += __shfl_down(cnt, off, 64);
    }
    if (lane == 0) {
        s_sum = sum;
        s_cnt = cnt;
    }
    __syncthreads();
    const double hsync_median_est = (s_cnt > 0) ? ((double)s_sum / s_cnt) : p.hsync_nom;
    const double hsync_offset = hsync_median_est - p.hsync_nom;
    const double a_hsync_min = p.hsync_nom + hsync_offset - p.tol;
    const double a_hsync_max = p.hsync_nom + hsync_offset + p.tol;
    const double a_eq_min = p.eq_nom + hsync_offset - p.tol;
    const double a_eq_max = p.eq_nom + hsync_offset + p.tol;
    const double a_vsync_min = (p.vsync_nom + hsync_offset) * 0.5;
    const double a_vsync_max = p.vsync_nom + hsync_offset + 2.0 * p.tol;
    for (int i = lane; i < n; i += 64)
        ty[i] = k3_class((double)ln[i], a_vsync_min, a_vsync_max, a_hsync_min, a_hsync_max, a_eq_min, a_eq_max);
    __syncthreads();
    if (lane != 0) return;
    const double inlinelen = p.inlinelen;
    // Step 1: vblank state machine (:139-196)
    int ref_pulse_idx = -1, last_hsync_idx = -1, first_eq1_idx = -1, last_eq2_idx = -1, state = -1;
    for (int i = 0; i < n; i++) {
        const int t = ty[i];
        if (state == -1) {
            if (t == kPulseHsync) { state = kPulseHsync; last_hsync_idx = i; }
        } else if (state == kPulseHsync) {
            if (t == kPulseHsync) last_hsync_idx = i;
            else if (t == kPulseEq1) { state = kPulseEq1; first_eq1_idx = i; }
            else if (t == kPulseVsync) state = kPulseVsync;
        } else if (state == kPulseEq1) {
            if (t == kPulseVsync) state = kPulseVsync;
            else if (t == kPulseHsync) { state = kPulseHsync; last_hsync_idx = i; first_eq1_idx = -1; }
        } else if (state == kPulseVsync) {
            if (t == kPulseEq1) { state = kPulseEq2; last_eq2_idx = i; }
            else if (t == kPulseHsync) { ref_pulse_idx = i; break; }
        } else if (state == kPulseEq2) {
            if (t == kPulseEq1) last_eq2_idx = i;
            else if (t == kPulseHsync) {
                if (last_eq2_idx < 0) last_eq2_idx = i - 1;
                ref_pulse_idx = i;
                break;
            }
        }
    }
    // Step 1b: field parity (:198-235): a long entry or exit spacing is a
    // first field in NTSC, a second field in PAL
    int field_parity = -1;
    if (ref_pulse_idx >= 0 && last_hsync_idx >= 0 && first_eq1_idx >= 0) {
        double entry_lines = (double)(st[first_eq1_idx] - st[last_hsync_idx]) / inlinelen;
        field_parity = ((entry_lines > 0.75) == (p.ntsc != 0)) ? 1 : 0;
    }
    if (field_parity < 0 && ref_pulse_idx >= 0 && last_eq2_idx >= 0) {
        double exit_lines = (double)(st[ref_pulse_idx] - st[last_eq2_idx]) / inlinelen;
        field_parity = ((exit_lines > 0.75) == (p.ntsc != 0)) ? 1 : 0;
    }
    is_first[f] = field_parity;
    // Step 2: mean line length from the longest HSYNC run (:231-277)
    int best_run_start = -1, best_run_len = 0, cur_run_start = -1, cur_run_len = 0;
    for (int i = 0; i < n; i++) {
        if (ty[i] == kPulseHsync) {
            if (cur_run_start == -1) { cur_run_start = i; cur_run_len = 1; }
            else cur_run_len++;
        } else {
            if (cur_run_len > best_run_len) { best_run_start = cur_run_start; best_run_len = cur_run_len; }
            cur_run_start = -1;
            cur_run_len = 0;
        }
    }
    if (cur_run_len > best_run_len) { best_run_start = cur_run_start; best_run_len = cur_run_len; }
    double meanlinelen = inlinelen;
    if (best_run_len >= 2) {
        double ssum = 0.0;
        int count2 = 0;
        for (int i = best_run_start + 1; i < best_run_start + best_run_len; i++) {
            double spacing = (double)(st[i] - st[i - 1]);
            if (spacing >= inlinelen * 0.95 && spacing <= inlinelen * 1.05) { ssum += spacing; count2++; }
        }
        if (count2 > 0) meanlinelen = ssum / count2;
    }
    // Step 3: reference pulse (:279-324)
    double ref_position, ref_line;
    if (ref_pulse_idx >= 0) {
        ref_position = (double)st[ref_pulse_idx];
        ref_line = 10.0;
        if (ref_pulse_idx > n / 2) {
            double corrected = ref_position - p.spf;
            if (corrected > 0.0) ref_position = corrected;
        }
    } else {
        ref_pulse_idx = 0;
        for (int i = 0; i < n; i++)
            if (ty[i] == kPulseHsync) { ref_pulse_idx = i; break; }
        // a field without pulses reads position 0 (the reference reads a
        // stale pulse slot there)
        ref_position = n > 0 ? (double)st[ref_pulse_idx] : 0.0;
        ref_line = 0.0;
    }
    // Step 4a: line markers (:346-387)
    const double tiny_hsync_max = p.hsync_nom * 0.75;
    const double surrogate_gap_max = inlinelen * 0.25;
    int marker_count = 0;
    for (int i = 0; i < n && marker_count < kMaxPulses; i++) {
        const int ti = ty[i];
        const bool surrogate_eq = (ti == kPulseEq1 || ti == kPulseEq2) && i + 1 < n && ty[i + 1] == kPulseHsync &&
                                  (double)ln[i + 1] <= tiny_hsync_max &&
                                  (double)(st[i + 1] - st[i]) <= surrogate_gap_max;
        const bool suppress = ti == kPulseHsync && (double)ln[i] <= tiny_hsync_max && i > 0 &&
                              (ty[i - 1] == kPulseEq1 || ty[i - 1] == kPulseEq2) &&
                              (double)(st[i] - st[i - 1]) <= surrogate_gap_max;
        if (surrogate_eq || (ti == kPulseHsync && !suppress)) {
            ms[marker_count] = st[i];
            ml[marker_count] = ln[i];
            marker_count++;
        }
    }
    // Step 4b: collapse near-duplicate markers (:389-433)
    const double duplicate_gap_max = inlinelen * 0.75;
    int num_hsyncs = 0, cluster_start = 0;
    while (cluster_start < marker_count && num_hsyncs < kMaxPulses) {
        int best_idx = cluster_start;
        double best_score = fabs((double)ml[cluster_start] - p.hsync_nom);
        int cluster_end = cluster_start + 1;
        while (cluster_end < marker_count && (double)(ms[cluster_end] - ms[cluster_start]) < duplicate_gap_max) {
            double score = fabs((double)ml[cluster_end] - p.hsync_nom);
            if (score < best_score || (score == best_score && ms[cluster_end] < ms[best_idx])) {
                best_idx = cluster_end;
                best_score = score;
            }
            cluster_end++;
        }
        hs[num_hsyncs++] = ms[best_idx];
        cluster_start = cluster_end;
    }
    // Step 4c: greedy nearest-marker snapping (:437-476)
    double* locs = linelocs + (size_t)f * p.lpf;
    const double max_allowed_distance = meanlinelen / 1.5;
    int cur_hsync = 0;
    for (int line = 0; line < p.lpf; line++) {
        double expected = ref_position + meanlinelen * (line - ref_line);
        double loc = expected;
        if (cur_hsync < num_hsyncs) {
            double best_dist = max_allowed_distance;
            int best_pos = -1;
            int search = cur_hsync;
            while (search < num_hsyncs) {
                double dist = fabs((double)hs[search] - expected);
                if (dist <= best_dist) {
                    best_dist = dist;
                    best_pos = hs[search];
                    cur_hsync = search;
                }
                if (search + 1 < num_hsyncs) {
                    double next_dist = fabs((double)hs[search + 1] - expected);
                    if (next_dist > dist) break;
                } else {
                    break;
                }
                search++;
            }
            if (best_pos >= 0) {
                loc = (double)best_pos;
                cur_hsync++;
            }
        }
        locs[line] = loc;
    }
}

// ---- diagnostic counters (atomics into ChunkCtl::diag) ----

// K3 spacing counters (k_line_locs_debug_analyze, line_locs.cu:585-626):
// diag[0] bad spacing, diag[1] isolated spacing, diag[2] large jump.
__global__ void k_diag_spacing(const double* ll, ChunkCtl* ctl, int als, double nominal, int lpf) {
    const int nf = ctl->nf;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nf * (lpf - 1)) return;
    const int field = idx / (lpf - 1);
    const int line = idx % (lpf - 1);
    if (line < als || line + 1 >= lpf) return;
    const double* l = ll + (size_t)field * lpf;
    const double spacing = l[line + 1] - l[line];
    if (spacing < 0.9 * nominal || spacing > 1.1 * nominal) atomicAdd(&ctl->diag[0], 1);
    if (fabs(spacing - nominal) > 0.25 * nominal) atomicAdd(&ctl->diag[2], 1);
    if (line > als && line + 2 < lpf) {
        const double prev_spacing = l[line] - l[line - 1];
        const double next_spacing = l[line + 2] - l[line + 1];
        const double neighbor_avg = 0.5 * (prev_spacing + next_spacing);
        if (fabs(spacing - neighbor_avg) > 0.15 * nominal) atomicAdd(&ctl->diag[1], 1);
    }
}

// K4 counters (k_hsync_refine_debug_analyze, hsync_refine.cu:273-316):
// diag[3] large delta, diag[4] isolated jump, diag[5] refined sync-like.
__global__ void k_diag_k4(const double* d05, const double* before, const double* after, ChunkCtl* ctl, int spfp,
                          int als, double thr, double large, double isolated, int lpf) {
    const int nf = ctl->nf;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nf * lpf) return;
    const int field = idx / lpf;
    const int line = idx % lpf;
    if (line < als || line + 1 >= lpf) return;
    const int base = field * lpf;
    const double after_loc = after[base + line];
    const double delta = after_loc - before[base + line];
    if (fabs(delta) > large) atomicAdd(&ctl->diag[3], 1);
    if (line > als && line + 1 < lpf) {
        const double prev_delta = after[base + line - 1] - before[base + line - 1];
        const double next_delta = after[base + line + 1] - before[base + line + 1];
        const double neighbor_avg = 0.5 * (prev_delta + next_delta);
        if (fabs(delta - neighbor_avg) > isolated) atomicAdd(&ctl->diag[4], 1);
    }
    const int sample = (int)(after_loc + 0.5);
    if (sample >= 0 && sample < nf * spfp && d05[sample] <= thr) atomicAdd(&ctl->diag[5], 1);
}

// K5 counters (k_tbc_resample, tbc_resample.cu:75-116): diag[6] pixels out
// of bounds, diag[7] rows of suspicious length, diag[8] active-picture pixels
// at or below the sync threshold.  One wavefront count per atomic.
__global__ __launch_bounds__(256) void k_diag_k5(const double* demod, const double* ll, ChunkCtl* ctl, int spfp,
                                                 double spl, double thr, K3Params p) {
    const int nf = ctl->nf;
    const int ppf = p.field_lines * p.line_len;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    bool oob = false, geom = false, sync_like = false;
    if (idx < nf * ppf) {
        const int total = nf * spfp;
        const int field = idx / ppf;
        const int rem = idx - field * ppf;
        const int row = rem / p.line_len;
        const int col = rem - row * p.line_len;
        const double ls = ll[field * p.lpf + row];
        const double le = ll[field * p.lpf + row + 1];
        const double len = le - ls;
        geom = col == 0 && (len <= 0.0 || len < 0.8 * spl || len > 1.2 * spl);
        const double frac = (double)col / (double)p.line_len;
        const double coord = ls + frac * len;
        const int ci = (int)coord;
        const double x = coord - (double)ci;
        if (ci < 1 || ci + 2 >= total) {
            oob = true;
        } else {
            const double p0 = demod[ci - 1], p1 = demod[ci], p2 = demod[ci + 1], p3 = demod[ci + 2];
            const double a = p2 - p0;
            const double b = 2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3;
            const double c = 3.0 * (p1 - p2) + p3 - p0;
            const double hz = p1 + 0.5 * x * (a + x * (b + x * c));
            sync_like = col >= p.avs && col < p.ave && hz <= thr;
        }
    }
    const int c_oob = __popcll(__ballot(oob)), c_geom = __popcll(__ballot(geom)), c_sync = __popcll(__ballot(sync_like));
    if ((threadIdx.x & 63) == 0) {
        if (c_oob) atomicAdd(&ctl->diag[6], c_oob);
        if (c_geom) atomicAdd(&ctl->diag[7], c_geom);
        if (c_sync) atomicAdd(&ctl->diag[8], c_sync);
    }
}

// ---- host launchers ----

hipError_t run_k3(const ChunkCtl* ctl, int nf_override, int max_fields, const Fmt& fmt, const int* starts,
                  const int* lens, const int* count, int* types, double* linelocs, int* is_first, hipStream_t st) {
    const K3Params p = k3_params(fmt);
    if (types)
        hipLaunchKernelGGL(k_k3_classify, dim3((kMaxPulses + 255) / 256, max_fields), dim3(256), 0, st, lens, count,
                           ctl, nf_override, p, types);
    hipLaunchKernelGGL(k_k3_field, dim3(max_fields), dim3(64), 0, st, starts, lens, count, ctl, nf_override, p,
                       linelocs, is_first);
    return hipGetLastError();
}

hipError_t run_diag_spacing(const double* ll, ChunkCtl* ctl, int max_fields, const Fmt& fmt, hipStream_t st) {
    const int n = max_fields * (fmt.lines_per_frame - 1);
    hipLaunchKernelGGL(k_diag_spacing, dim3((n + 255) / 256), dim3(256), 0, st, ll, ctl, fmt.active_line_start,
                       (double)fmt.samples_per_line, fmt.lines_per_frame);
    return hipGetLastError();
}

hipError_t run_diag_k4(const double* d05, const double* before, const double* after, ChunkCtl* ctl, int max_fields,
                       const Fmt& fmt, hipStream_t st) {
    const int n = max_fields * fmt.lines_per_frame;
    const double us = 1.0e-6 * fmt.sample_rate;
    hipLaunchKernelGGL(k_diag_k4, dim3((n + 255) / 256), dim3(256), 0, st, d05, before, after, ctl, fmt.spfp,
                       fmt.active_line_start, fmt.pulse_threshold_hz, 1.5 * us, 1.0 * us, fmt.lines_per_frame);
    return hipGetLastError();
}

hipError_t run_diag_k5(const double* demod, const double* ll, ChunkCtl* ctl, int max_fields, const Fmt& fmt,
                       hipStream_t st) {
    const long long n = (long long)max_fields * fmt.field_lines * fmt.line_len;
    hipLaunchKernelGGL(k_diag_k5, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, demod, ll, ctl, fmt.spfp,
                       (double)fmt.samples_per_line, fmt.pulse_threshold_hz, k3_params(fmt));
    return hipGetLastError();
}

}  // namespace cuvhs
