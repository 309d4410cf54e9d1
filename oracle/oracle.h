// CPU FP64 oracle for the cuVHS decode hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library restates the reference's stage
// logic (hunterschuler/cuVHS, src/pipeline/*.cu, src/io/*.cpp,
// src/format/*.cpp) statement by statement on the CPU in FP64.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
// only as the checker.  The product (cuvhs_amd/, include/cuvhs.h) never links
// or calls it.
//
// Parity pins (see tests/): FFT vs numpy.fft + brute-force DFT; filter
// design vs scipy.signal; VideoFormat known answers; u8 conversion
// exhaustive; K2/VSYNC vs brute-force run scans; K4 vs closed-form linear
// edges; K5 vs closed-form Catmull-Rom on polynomials; JSON vs golden text;
// K2b line location (pipeline.cu:133-148, 410-893) and the chroma burst
// metrics, burst phase and NTSC phase table (chroma_decode.cu:329-450) bit
// for bit against the reference's own host code compiled from its sources
// (oracle/ref_k2b_driver.cpp, vectors in tests/golden/ref_*.npz).  Pinned
// by invariants: the chunk/VSYNC clustering (pipeline.cu:1595-1651), the
// chroma retry (chroma_decode.cu:774-975).
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_PULSES 800          /* sync_pulses.h:6 */
#define OR_MAX_DROPOUTS 512        /* dropout_detect.h:5 */
#define OR_LINES_PER_FRAME 525     /* NTSC; PAL 625 */
#define OR_FIELD_LINES 263         /* NTSC; PAL 313 */
#define OR_LINE_LEN 910            /* NTSC; PAL 1135 */

typedef struct or_format {
    double sample_rate;
    int samples_per_line;
    int samples_per_field;
    int spfp;                 /* samples_per_field + 20 lines (pipeline.cu:1339-1342) */
    int k1_fft_size;          /* next 7-smooth >= spfp + 2048 */
    int chroma_fft_size;      /* next 7-smooth >= spfp */
    double hsync_width, eq_pulse_width, vsync_width;
    double ire0, hz_ire, vsync_ire, sync_tip_hz, pulse_threshold_hz;
    double output_zero, output_scale, output_rate, fsc, chroma_under;
    double burst_abs_ref, black16b, white16b;
    int burst_start, burst_end;
} or_format;

int or_format_init(double mhz, or_format* out);
/* system: 0 NTSC, 1 PAL (video_format.cpp:48-87); the _sys variants below
 * take it, the others are NTSC */
int or_format_init_sys(int system, double mhz, or_format* out);

/* --- filter tables (for design checks) --------------------------------- */
/* which: 0 = K1 RF (real, fb bins), 1 = K1 Fvideo (complex, interleaved),
 *        2 = K1 Fvideo05 (complex), 3 = chroma pre-BPF |H|^2/N (real)
 * returns number of bins written (or required when out == NULL) */
int or_filter_table(double mhz, int which, double* out);

/* --- FFT (for tests) ----------------------------------------------------- */
int or_fft(int n, const double* in_interleaved, double* out_interleaved, int sign);

/* --- stage functions on host buffers (mirroring the product's C ABI) ----- */
/* K1: raw[nfields*spfp] (already converted to s16 scale) */
int or_fm_demod(double mhz, const double* raw, int nfields,
                double* demod, double* demod05, double* env);
/* VSYNC discovery: returns raw count (may exceed cap); idx holds the first
 * min(count,cap) candidates in ascending order */
int or_vsync_discover(double mhz, const double* d05, int total, int* idx, int cap);
/* K2 */
int or_sync_pulses(double mhz, const double* d05, const uint64_t* offsets, int nfields,
                   int* starts, int* lens, int* count);
/* K4 (in place) */
int or_hsync_refine(double mhz, const double* d05, int total, double* linelocs, int nfields);
/* K5 */
int or_tbc_resample(double mhz, const double* demod, int total, const double* linelocs,
                    int nfields, uint16_t* luma);

/* --- whole decode ------------------------------------------------------- */
typedef struct or_dec or_dec;
or_dec* or_create(double mhz, int batch, int record_trace);
or_dec* or_create_sys(int system, double mhz, int batch, int record_trace);
void or_destroy(or_dec* d);
/* Input sample format (0 u8, 1 s16, 2 u16; 16-bit samples little-endian as
 * the host stores them) and per-read DC removal (raw_reader.cpp:71-100).
 * Default u8 without DC removal. */
int or_set_input(or_dec* d, int format, int dc_correct);
/* Line locator: 0 = K2b (default), 1 = K3 (the reference with
 * CUVHS_K2B_DISABLE set, line_locs.cu); diagnostics != 0 records per chunk
 * the reference's K3/K4/K5 debug counters (line_locs.cu:585-653,
 * hsync_refine.cu:273-349, tbc_resample.cu:75-116) in this order:
 * K3 bad_spacing, isolated_spacing, large_jump; K4 large_delta,
 * isolated_jump, refined_sync_like; K5 oob_pixels, bad_geom_lines,
 * sync_like_pixels. */
int or_set_line_locator(or_dec* d, int mode, int diagnostics);
int or_chunk_diagnostics(const or_dec* d, int c, int* out9);
/* One diagnostic counter set: which = 3 (a = linelocs), 4 (a = demod_05 of
 * `total` samples, b = linelocs before K4, c = after), 5 (a = demod of
 * `total` samples, b = linelocs); nf fields; three counts out. */
int or_diag_counts(double mhz, int which, const double* a, const double* b, const double* c, int total, int nf,
                   int* out3);
/* K3 on K2 pulse lists ([nf][800] each): nominal types (line_locs.cu:18-48),
 * line locations and field parity (1 first, 0 second, -1 unknown). */
int or_k3_linelocs(double mhz, const int* starts, const int* lens, const int* count, int nf, int* types,
                   double* linelocs, int* is_first);
int or_k3_linelocs_sys(int system, double mhz, const int* starts, const int* lens, const int* count, int nf,
                       int* types, double* linelocs, int* is_first);
/* The conversion alone: n samples (n or 2n bytes) -> out[n] doubles. */
int or_convert(int format, int dc_correct, const uint8_t* data, size_t n, double* out);
/* Decode a whole capture held in memory: n samples (n or 2n bytes).  Returns
 * field count. */
int or_decode(or_dec* d, const uint8_t* data, size_t n);
int or_num_fields(const or_dec* d);
const uint16_t* or_luma(const or_dec* d);
const uint16_t* or_chroma(const or_dec* d);
/* per-field metadata arrays (length num_fields each); any pointer may be NULL */
void or_field_meta(const or_dec* d, uint64_t* file_loc, int* is_first, int* phase_id,
                   int* do_count);
/* dropouts of one field; returns count */
int or_field_dropouts(const or_dec* d, int field, int* lines, int* startx, int* endx);
int or_write_outputs(const or_dec* d, const char* out_base);
/* trace of chunk c (record_trace=1); returns fields in chunk or -1 */
int or_trace_chunk(const or_dec* d, int c, uint64_t* raw_offset, int* fields_k1,
                   uint64_t* next_offset, uint64_t* offsets, int* pulse_count,
                   int* pulse_starts, int* pulse_lens, int* vpulse_count, int* vpulse_starts,
                   int* vpulse_lens, int* vpulse_types, double* ll_coarse, double* ll_refined,
                   int* is_first_k2b);
int or_num_chunks(const or_dec* d);
/* VSYNC candidates chunk c found (may exceed the cap of 20 per field); -1 if no such chunk */
int or_trace_candidates(const or_dec* d, int c);
/* burst metrics of the field-0 track trial (track 0, track 1) and the track
 * chosen; -1 before any chroma was decoded */
int or_chroma_trial(const or_dec* d, double* metrics, int* track);

/* The field-0 track trial of the first chunk's chroma (record_trace=1):
 * which = 0 / 1: band-passed lines [263][1024] (unnormalised inverse
 * transforms) with track 0 / 1; 2: re-run with the detected track, from
 * which the burst phase is measured (chroma_decode.cu:688-731).  Returns the
 * element count, -1 if none was kept. */
int or_chroma_trial_lines(const or_dec* d, int which, double* out);

/* K2b on one chunk's K2 pulses, stand-alone (pipeline.cu:1717-1778), with
 * the cross-field cadence state (pipeline.h:92-97) passed in and out.
 * starts/lens/count ([nf][800]) are replaced by the validated pulses; good
 * [nf][800] gets their good flags, anchor [nf][9] per field: valid,
 * line0loc, first_hsync_loc, first_hsync_line, next_field, first_field,
 * progressive_field, prev_hsync_diff, meanlinelen (good, anchor optional). */
typedef struct or_k2b_state {
    int prev_first_field;
    int prev_progressive_field;
    long long prev_first_hsync_readloc;
    double prev_first_hsync_loc;
    double prev_first_hsync_diff;
} or_k2b_state;
int or_k2b_chunk(double mhz, or_k2b_state* st, uint64_t raw_offset, const uint64_t* offsets, int nf, int* starts,
                 int* lens, int* count, int* types, unsigned char* good, double* linelocs, int* is_first,
                 double* anchor);
int or_k2b_chunk_sys(int system, double mhz, or_k2b_state* st, uint64_t raw_offset, const uint64_t* offsets, int nf,
                     int* starts, int* lens, int* count, int* types, unsigned char* good, double* linelocs,
                     int* is_first, double* anchor);
/* lookup_ntsc_phase (chroma_decode.cu:359-381): the phase ID */
int or_lookup_ntsc_phase(int is_first, int quadrant, int delta);
/* measure_burst_cancellation / measure_burst_phase (chroma_decode.cu:383-450)
 * on one field's band-passed lines [263][1024] (NTSC) */
double or_burst_metric(double mhz, const double* lines);
double or_burst_phase(double mhz, const double* lines);
/* the same with the system's geometry (PAL: [313][2048] lines) */
double or_burst_metric_sys(int system, double mhz, const double* lines);
double or_burst_phase_sys(int system, double mhz, const double* lines);

/* Seconds spent per stage over the decodes of d (9 values: conversion, K1,
 * VSYNC discovery + clustering, K2, K2b or K3, K4, K5, K6, K7); returns 9. */
int or_stage_seconds(const or_dec* d, double* out);
/* OpenMP threads for later calls (n > 0); returns the current maximum. */
int or_set_threads(int n);

/* .tbc.json from explicit metadata (do_* arrays concatenated over fields) */
int or_write_json_meta(double mhz, const char* path, int n, const uint64_t* file_loc, const int* is_first,
                       const int* phase_id, const int* do_count, const int* do_line, const int* do_startx,
                       const int* do_endx);
int or_write_json_meta_sys(int system, double mhz, const char* path, int n, const uint64_t* file_loc,
                           const int* is_first, const int* phase_id, const int* do_count, const int* do_line,
                           const int* do_startx, const int* do_endx);

/* CLI-equivalent: decode file -> out_base.{tbc,_chroma.tbc,tbc.json} */
int or_decode_file(const char* in_path, const char* out_base, double mhz, int batch);
int or_decode_file_fmt(const char* in_path, const char* out_base, double mhz, int batch, int format, int dc_correct);

#ifdef __cplusplus
}
#endif
