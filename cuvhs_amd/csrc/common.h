// Shared definitions of the MI355X decoder: video format constants and the
// device-side layout of the per-chunk control block.
#pragma once
#include <stdint.h>

#include <cmath>

namespace cuvhs {

constexpr int kMaxPulses = 800;         // sync_pulses.h:6
constexpr int kMaxDropouts = 512;       // dropout_detect.h:5
constexpr int kMaxRfDropouts = 128;     // dropout_detect.cu:123
constexpr int kOverlap = 2048;          // fm_demod.cu:17
constexpr int kFieldMarginLines = 20;   // pipeline.cu:1339
constexpr int kVpCap = 2048;            // K2b validated-pulse capacity per field
constexpr int kPulseHsync = 0, kPulseEq1 = 1, kPulseVsync = 2, kPulseEq2 = 3;  // line_locs.h:6

// Video systems (video_format.h:4) and the largest of their geometries, for
// arrays sized at compile time: 625 line locations per frame, 313 output
// lines of 1135 samples per field, a 2048-point chroma line transform.
constexpr int kSystemNtsc = 0, kSystemPal = 1;
constexpr int kMaxLinesPerFrame = 625;
constexpr int kMaxFieldLines = 313;
constexpr int kMaxLineLen = 1135;
constexpr int kMaxBurstLen = 40;

// VideoFormat (video_format.cpp:23-104) plus the chunk geometry the pipeline
// derives from it.  The reference's PAL format keeps its NTSC pipeline and
// K1 filters and changes these constants (and the few branches on the
// system in K2b and K3).
struct Fmt {
    int system;                                 // kSystemNtsc / kSystemPal
    double sample_rate;
    double line_rate, fsc, chroma_under, burst_abs_ref, burst_start_us, burst_end_us;
    double hz_ire, ire0, vsync_ire, sync_tip_hz, pulse_threshold_hz;
    double output_zero, output_scale, output_rate;
    double hsync_width, eq_pulse_width, vsync_width;
    int lines_per_frame, lines_per_field;       // 525 / 263, 625 / 313
    int line_len, field_lines;                  // output_line_len, output_field_lines
    int chroma_fft;                             // next_pow2(line_len) (chroma_decode.cu:538)
    int samples_per_line, samples_per_field, spfp;
    int num_eq_pulses, field_lines_first, field_lines_second, active_line_start;
    int burst_start, burst_end;                 // TBC columns (tbc_writer.cpp:100-101)
    int active_video_start, active_video_end;   // TBC columns (tbc_resample.cu:156-157)
    int vsync_min_w, vsync_max_w;               // vsync_discover.cu:59-60
};

inline Fmt make_fmt(double mhz, int system = kSystemNtsc) {
    Fmt f{};
    f.system = system;
    f.sample_rate = mhz * 1e6;
    if (system == kSystemPal) {
        f.lines_per_frame = 625;
        f.lines_per_field = 313;
        f.line_rate = 15625.0;
        f.fsc = 4433618.75;
        f.chroma_under = ((625.0 * 25.0) * 40.0) + 1953.0;
        f.burst_abs_ref = 5000.0;
        f.burst_start_us = 5.6;
        f.burst_end_us = 7.85;
        f.line_len = 1135;
        f.field_lines = 313;
        f.num_eq_pulses = 5;
        f.field_lines_first = 312;
        f.field_lines_second = 313;
        f.vsync_ire = -0.3 * (100.0 / 0.7);
        f.hz_ire = 1e6 / (100.0 + (-f.vsync_ire));
        f.ire0 = 4.8e6 - (f.hz_ire * 100.0);
        f.output_zero = 256.0;
        f.output_scale = (54016.0 - 256.0) / (100.0 - f.vsync_ire);
        f.active_line_start = 8;
        f.vsync_width = 27.3e-6 * f.sample_rate;
        f.active_video_start = 185;
        f.active_video_end = 1107;
    } else {
        f.lines_per_frame = 525;
        f.lines_per_field = 263;
        f.line_rate = 15734.264;
        f.fsc = 3579545.0;
        f.chroma_under = (525.0 * (30.0 / 1.001)) * 40.0;
        f.burst_abs_ref = 4416.0;
        f.burst_start_us = 5.3;
        f.burst_end_us = 7.8;
        f.line_len = 910;
        f.field_lines = 263;
        f.num_eq_pulses = 6;
        f.field_lines_first = 263;
        f.field_lines_second = 262;
        f.hz_ire = 1e6 / 140.0;
        f.ire0 = 4.4e6 - (f.hz_ire * 100.0);
        f.vsync_ire = -40.0;
        f.output_zero = 1024.0;
        f.output_scale = (51200.0 - 1024.0) / (100.0 - f.vsync_ire);
        f.active_line_start = 10;
        f.vsync_width = 27.1e-6 * f.sample_rate;
        f.active_video_start = 134;
        f.active_video_end = 894;
    }
    f.chroma_fft = 1;
    while (f.chroma_fft < f.line_len) f.chroma_fft <<= 1;
    f.sync_tip_hz = f.ire0 + f.hz_ire * f.vsync_ire;
    f.pulse_threshold_hz = f.ire0 + f.hz_ire * (-20.0);
    f.output_rate = 4.0 * f.fsc;
    f.samples_per_line = static_cast<int>(std::round(f.sample_rate / f.line_rate));
    f.samples_per_field = f.samples_per_line * f.lines_per_field;
    f.hsync_width = 4.7e-6 * f.sample_rate;
    f.eq_pulse_width = 2.3e-6 * f.sample_rate;
    f.spfp = f.samples_per_field + kFieldMarginLines * f.samples_per_line;
    f.burst_start = (int)(f.burst_start_us * 1e-6 * f.output_rate + 0.5);
    f.burst_end = (int)(f.burst_end_us * 1e-6 * f.output_rate + 0.5);
    f.vsync_min_w = (int)(20.0e-6 * f.sample_rate);
    f.vsync_max_w = (int)(35.0e-6 * f.sample_rate);
    return f;
}

// K2b cadence state carried across fields (pipeline.h:92-97)
struct K2bState {
    int prev_first_field;
    int prev_progressive_field;
    long long prev_first_hsync_readloc;
    double prev_first_hsync_loc;
    double prev_first_hsync_diff;
};

// ChromaState (chroma_decode.h:9-15)
struct ChromaState {
    int valid;
    int current_track;
    double good_metric_threshold;
    int cycle_start;
};

// Per-chunk control block in device memory: written by the scheduling
// kernels, read by later kernels of the same chunk (so no host round trip is
// needed inside a chunk) and copied to the host.
struct ChunkCtl {
    int fields_k1;            // segments K1 demodulated
    int nf;                   // fields after VSYNC clustering
    int vsync_count;          // raw candidate count
    int status;               // warning bits
    int vp_overflow;          // K2b capacity exceeded in some field
    int nevents;              // threshold-crossing events found
    int guard_a, guard_b;     // FP64 guard-band blocks recomputed (K1 part A, part B)
    uint64_t raw_offset;
    uint64_t next_raw_offset;
    // diagnostic counters of the reference's debug logs, when enabled:
    // K3 bad/isolated spacing, large jump; K4 large delta, isolated jump,
    // refined sync-like; K5 out-of-bounds pixels, bad rows, sync-like pixels
    int diag[9];
};

// State carried from chunk to chunk; only the line-location, chroma and
// output-metadata stages touch it, in chunk order.
struct DecodeState {
    K2bState k2b;
    ChromaState chroma;
    int last_parity;          // pipeline.cu:2575 (function-static in the reference)
    int pad_;
};

}  // namespace cuvhs
