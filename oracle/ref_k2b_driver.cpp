// Driver around the reference's own host-side K2b line location and chroma
// phase helpers.  TEST INFRASTRUCTURE ONLY: it exists to pin the oracle's
// restatement of them, and nothing in the product links it.
//
// The reference keeps these functions as plain C++ inside two CUDA files:
//   src/pipeline/pipeline.cu:133-148, 410-893  host_pulse_qualitycheck and
//       the K2b functions (HostPulse ... host_valid_pulses_to_linelocs_exactish)
//   src/pipeline/chroma_decode.cu:329-450  NTSC phase table,
//       lookup_ntsc_phase, measure_burst_cancellation, measure_burst_phase
// They need nothing but the standard library and the reference's own
// headers (format/video_format.h, pipeline/line_locs.h,
// pipeline/sync_pulses.h).  The Makefile target `refk2b` cuts those line
// ranges out of the mounted reference into the git-ignored oracle/_ref/ and
// compiles this file, which includes them, together with the reference's
// video_format.cpp.  No stand-in of any reference header or library is used.
//
// The per-chunk loop below is this project's restatement of the caller,
// Pipeline::process_chunk (pipeline.cu:1717-1778), with the cross-field
// cadence state (pipeline.h:92-97) held by the caller of ref_k2b_chunk.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "format/video_format.h"
#include "pipeline/line_locs.h"
#include "pipeline/sync_pulses.h"

#include "k2b_host.inc"
#include "chroma_host.inc"

extern "C" {

// Cross-field K2b cadence state (pipeline.h:92-97).
typedef struct k2b_state_io {
    int prev_first_field;
    int prev_progressive_field;
    long long prev_first_hsync_readloc;
    double prev_first_hsync_loc;
    double prev_first_hsync_diff;
} k2b_state_io;

enum { kAnchorVals = 9 };

// One chunk of K2b as process_chunk runs it (pipeline.cu:1717-1778).
// starts/lens/count ([nf][800], count[nf]) hold the K2 pulses on entry and the
// validated pulses on return; types [nf][800]; good [nf][800] (the validated
// pulses' good flags, 0/1); linelocs [nf][525] (PAL [nf][625]); is_first
// [nf]; anchor [nf][9]: valid, line0loc, first_hsync_loc, first_hsync_line,
// next_field, first_field, progressive_field, prev_hsync_diff, meanlinelen.
// Any of good / anchor may be NULL.  system: 0 NTSC, 1 PAL.
int ref_k2b_chunk_sys(int system, double mhz, k2b_state_io* st, uint64_t raw_offset, const uint64_t* offsets, int nf,
                      int* starts, int* lens, int* count, int* types, unsigned char* good, double* linelocs,
                      int* is_first, double* anchor) {
    if (!st || nf < 0 || system < 0 || system > 1) return -1;
    VideoFormat fmt(system == 1 ? VideoSystem::PAL : VideoSystem::NTSC, mhz);
    const int LPF = fmt.lines_per_frame;
    for (int field = 0; field < nf; field++) {
        int* fs = starts + (size_t)field * MAX_PULSES;
        int* fl = lens + (size_t)field * MAX_PULSES;
        const int before = std::max(0, std::min(count[field], MAX_PULSES));
        const double first_start = before > 0 ? (double)fs[0] : 0.0;
        std::vector<HostValidPulse> vp = host_refine_pulses_exactish(fs, fl, before, fmt);
        const double mll = host_compute_meanlinelen_from_valid_pulses(vp, (double)fmt.samples_per_line);
        double prev_offset_lines = 0.0;
        if (st->prev_first_hsync_readloc != -1 && mll > 0.0)
            prev_offset_lines =
                ((double)st->prev_first_hsync_readloc - (double)(raw_offset + offsets[field])) / mll;
        HostK2bAnchor an = host_get_first_hsync_loc_exactish(vp, mll, fmt, st->prev_first_field, prev_offset_lines,
                                                             st->prev_first_hsync_loc, st->prev_first_hsync_diff, 75);
        const int after = std::min((int)vp.size(), MAX_PULSES);
        for (int i = 0; i < MAX_PULSES; i++) {
            const bool in = i < after;
            fs[i] = in ? vp[(size_t)i].pulse.start : 0;
            fl[i] = in ? vp[(size_t)i].pulse.len : 0;
            types[(size_t)field * MAX_PULSES + i] = in ? vp[(size_t)i].type : PULSE_HSYNC;
            if (good) good[(size_t)field * MAX_PULSES + i] = in && vp[(size_t)i].good ? 1 : 0;
        }
        count[field] = after;
        double* ll = linelocs + (size_t)field * LPF;
        if (an.valid) {
            std::vector<double> locs =
                host_valid_pulses_to_linelocs_exactish(vp, an.first_hsync_loc, an.first_hsync_line, mll, LPF);
            std::copy(locs.begin(), locs.end(), ll);
            is_first[field] = an.first_field;
            st->prev_first_field = an.first_field;
            st->prev_progressive_field = an.progressive_field;
            st->prev_first_hsync_readloc = (long long)(raw_offset + offsets[field]);
            st->prev_first_hsync_loc = an.first_hsync_loc;
            st->prev_first_hsync_diff = an.prev_hsync_diff;
        } else {
            for (int line = 0; line < LPF; line++) ll[line] = first_start + mll * (double)line;
            is_first[field] = (st->prev_first_field >= 0) ? (1 - st->prev_first_field) : 1;
        }
        if (anchor) {
            double* a = anchor + (size_t)field * kAnchorVals;
            a[0] = an.valid ? 1.0 : 0.0;
            a[1] = an.line0loc;
            a[2] = an.first_hsync_loc;
            a[3] = an.first_hsync_line;
            a[4] = an.next_field;
            a[5] = (double)an.first_field;
            a[6] = (double)an.progressive_field;
            a[7] = an.prev_hsync_diff;
            a[8] = mll;
        }
    }
    return 0;
}

int ref_k2b_chunk(double mhz, k2b_state_io* st, uint64_t raw_offset, const uint64_t* offsets, int nf, int* starts,
                  int* lens, int* count, int* types, unsigned char* good, double* linelocs, int* is_first,
                  double* anchor) {
    return ref_k2b_chunk_sys(0, mhz, st, raw_offset, offsets, nf, starts, lens, count, types, good, linelocs, is_first,
                             anchor);
}

int ref_chroma_params_sys(int system, double mhz, double* out);

// Chroma constants of the reference's VideoFormat for a sample rate: fsc,
// output_rate, output_line_len, output_field_lines, and the burst window in
// output samples as chroma_decode() derives it (chroma_decode.cu:548-549).
int ref_chroma_params(double mhz, double* out) { return ref_chroma_params_sys(0, mhz, out); }

int ref_chroma_params_sys(int system, double mhz, double* out) {
    VideoFormat fmt(system == 1 ? VideoSystem::PAL : VideoSystem::NTSC, mhz);
    out[0] = fmt.fsc;
    out[1] = fmt.output_rate;
    out[2] = (double)fmt.output_line_len;
    out[3] = (double)fmt.output_field_lines;
    out[4] = (double)(int)(fmt.burst_start_us * 1e-6 * fmt.output_rate + 0.5);
    out[5] = (double)(int)(fmt.burst_end_us * 1e-6 * fmt.output_rate + 0.5);
    return 0;
}

// lookup_ntsc_phase (chroma_decode.cu:359-381): returns the phase ID and
// writes the phase offset.
int ref_lookup_ntsc_phase(int is_first, int quadrant, int delta, int* offset) {
    NTSCPhaseResult r = lookup_ntsc_phase(is_first, quadrant, delta);
    if (offset) *offset = r.offset;
    return r.phase_id;
}

// measure_burst_cancellation (chroma_decode.cu:383-411) on one field's
// band-passed lines, [lines][fft_size] unnormalised inverse transforms.
double ref_burst_cancellation(const double* h, int lines, int fft_size, int line_len, int burst_start, int burst_end,
                              double fft_scale) {
    return measure_burst_cancellation(h, lines, fft_size, line_len, burst_start, burst_end, fft_scale);
}

// measure_burst_phase (chroma_decode.cu:415-450), degrees in [0, 360).
double ref_burst_phase(const double* h, int lines, int fft_size, int line_len, int burst_start, int burst_end,
                       double fft_scale, double fsc, double output_rate) {
    return measure_burst_phase(h, lines, fft_size, line_len, burst_start, burst_end, fft_scale, fsc, output_rate);
}

}  // extern "C"
