/*
 * cuvhs-mi355x: C ABI of the MI355X-native VHS RF decoder.
 *
 * The drop-in boundary is the reference's stage API (hunterschuler/cuVHS,
 * src/pipeline/ headers) plus its chunk loop (pipeline.cu:1525-2713).  Every
 * entry point takes plain pointers and sizes; device pointers are HIP device
 * memory on the context's device.  Every call returns an int status:
 *   0  OK
 *   <0 error (CUVHS_E_*)
 *   >0 data warning bits (CUVHS_W_*), the reference's soft paths kept
 *      observable (pipeline.cu:1584-1587, 1614-1620).
 *
 * Stage entry points enqueue on the context's stream and synchronise it.
 * Chunk decodes run a chunk's head (K1, sync events, VSYNC clustering) on
 * the context's stream and its tail (pulses through metadata) on a second
 * stream owned by the context, with the chroma pre-band-pass on a third and,
 * when the process has more than four hardware queues (GPU_MAX_HW_QUEUES,
 * read at cuvhs_create), part of the tail on a fourth, so that consecutive
 * chunks overlap; they return once all are idle.  All streams have the
 * default priority.  A context is not thread-safe; distinct contexts are
 * independent.
 */
#ifndef CUVHS_H
#define CUVHS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUVHS_MAX_PULSES 800          /* sync_pulses.h:6 */
#define CUVHS_MAX_DROPOUTS 512        /* dropout_detect.h:5 */
#define CUVHS_LINES_PER_FRAME 525     /* linelocs per field (video_format.cpp:26) */
#define CUVHS_FIELD_LINES 263         /* TBC rows per field */
#define CUVHS_LINE_LEN 910            /* TBC samples per row */
/* The same for PAL (video_format.cpp:49-61); cuvhs_format reports the
 * context's own. */
#define CUVHS_PAL_LINES_PER_FRAME 625
#define CUVHS_PAL_FIELD_LINES 313
#define CUVHS_PAL_LINE_LEN 1135

enum {
    CUVHS_OK = 0,
    CUVHS_E_ARG = -1,
    CUVHS_E_HIP = -2,
    CUVHS_E_ALLOC = -3,
    CUVHS_E_UNSUPPORTED = -4,
    CUVHS_E_CAPACITY = -5
};

enum {
    CUVHS_W_VSYNC_OVERFLOW = 1,       /* pipeline.cu:1584: candidates truncated */
    CUVHS_W_NO_VSYNC = 2,             /* pipeline.cu:1614: nominal spacing fallback */
    CUVHS_W_EOF = 4                   /* pipeline.cu:1535: fewer than one field left */
};

/* Video system (VideoSystem, video_format.h:4; --system, main.cpp:54).  PAL
 * is the reference's PAL format: its NTSC pipeline and K1 filters with PAL's
 * constants (video_format.cpp:48-87): 625 lines, 313 output lines of 1135
 * samples per field, 5 EQ pulses, PAL IRE levels, 4.43 MHz subcarrier and
 * 626,953 Hz colour-under, a 2048-point chroma line transform, and PAL's
 * branches of K2b (pipeline.cu:707-717, 811) and K3 (line_locs.cu:219-235). */
enum { CUVHS_SYSTEM_NTSC = 0, CUVHS_SYSTEM_PAL = 1 };
/* Arithmetic of the decode (the reference computes in FP64 with cuFFT,
 * fm_demod.cu:396-486, chroma_decode.cu:568-635).
 * FP64: the reference's precision, the default.
 * FP32: the plain FP32 form: K1's transforms, phase and envelope, the chroma
 *   pre-band-pass, K5 and the chroma line filter in FP32, demod / demod_05 /
 *   envelope stored as float; FP64 only in line locations and K4's
 *   positions (and the heterodyne and resampling coordinates).
 * MIXED_F16 / MIXED_BF16: K1's RF band-pass and Hilbert filter as Toeplitz
 *   products on the matrix cores (u8 samples exact in f16 / bf16, taps split
 *   hi + lo, FP32 accumulation), K1's video filters and the rest as the FP32
 *   form except demod_05, kept in FP64; an FP64 guard band recomputes the K1
 *   blocks whose sync-threshold decisions the reduced precision could change
 *   (outputs near the threshold, envelope dips).  u8 input without
 *   --dc-correct only (else CUVHS_E_UNSUPPORTED). */
enum { CUVHS_PRECISION_FP64 = 0, CUVHS_PRECISION_FP32 = 1, CUVHS_PRECISION_MIXED_F16 = 2,
       CUVHS_PRECISION_MIXED_BF16 = 3 };
/* Capture sample formats (InputFormat, video_format.h; RawReader::convert,
 * raw_reader.cpp:71-91): u8 -> (s - 128) * 256, s16 as is, u16 -> s - 32768.
 * 16-bit samples are in host byte order. */
enum { CUVHS_INPUT_U8 = 0, CUVHS_INPUT_S16 = 1, CUVHS_INPUT_U16 = 2 };
/* Line locator: K2b (the reference's default, pipeline.cu:410-893) or K3
 * (line_locs.cu, what the reference runs with CUVHS_K2B_DISABLE set,
 * pipeline.cu:1694). */
enum { CUVHS_LINELOC_K2B = 0, CUVHS_LINELOC_K3 = 1 };

typedef struct cuvhs_ctx cuvhs_ctx;

typedef struct cuvhs_config {
    int system;               /* CUVHS_SYSTEM_NTSC or CUVHS_SYSTEM_PAL */
    double sample_rate_mhz;   /* capture rate, e.g. 28.0 */
    int batch_fields;         /* F: fields per chunk; results depend on F */
    int precision;            /* CUVHS_PRECISION_* */
    void* stream;             /* hipStream_t, NULL = a stream owned by the context */
    int device;               /* HIP device ordinal; -1 = the device with the
                                 most memory (GPUDevice::init(-1),
                                 device.cu:13-24) */
    int input_format;         /* CUVHS_INPUT_* (0 = u8) */
    int dc_correct;           /* subtract each chunk read's mean (--dc-correct,
                                 InputConditioning, raw_reader.h:21-23) */
    int line_locator;         /* CUVHS_LINELOC_* (0 = K2b) */
    /* Block sizes of the overlap-save convolutions (no counterpart in the
     * reference, whose transforms are one cuFFT per field segment): K1 part
     * A and part B (4096 or 8192 points; part A is ignored by the mixed
     * forms, whose part A is a matrix-core product) and the chroma
     * pre-band-pass (8192 or 16384).  0 = the size the autotuner measured
     * fastest on this device for this sample rate and precision
     * (cuvhs_amd/autotune.py, table cuvhs_amd/csrc/tuning_gfx950.h), or, for
     * a rate or form it did not measure, the library's cost model.  A size
     * the filter's support does not fit is refused (CUVHS_E_ARG). */
    int block_k1a, block_k1b, block_pre;
    /* 1: cuvhs_decode_capture replays HIP graphs of its stages, captured by
     * cuvhs_create once per buffer slot, with each chunk's values read on the
     * device from a per-slot plan (no counterpart in the reference's host
     * loop, pipeline.cu:2626-2713); 0: it launches the kernels one by one.
     * Results are identical.  Profiling and diagnostics always launch one by
     * one.  cuvhs_create captures one context at a time in the process; a
     * capture must not overlap another thread's HIP calls, so contexts with
     * graphs are created before decoding starts in other threads. */
    int graphs;
} cuvhs_config;

/* Derived timing constants (VideoFormat, video_format.cpp:23-104). */
typedef struct cuvhs_format {
    double sample_rate;
    int samples_per_line;     /* round(fs / line rate): 15734.264 Hz NTSC, 15625 Hz PAL */
    int samples_per_field;    /* samples_per_line * 263 (PAL 313) */
    int spfp;                 /* samples_per_field + 20 lines (pipeline.cu:1339-1342) */
    double pulse_threshold_hz, ire0, hz_ire, output_scale, output_zero, output_rate;
    double black16b, white16b;
    int burst_start, burst_end;
    int system;               /* CUVHS_SYSTEM_* */
    int lines_per_frame;      /* line locations per field: 525 / 625 */
    int field_lines;          /* TBC rows per field: 263 / 313 */
    int line_len;             /* TBC samples per row: 910 / 1135 */
    double fsc;               /* colour subcarrier (Hz) */
} cuvhs_format;

/* Create/destroy.  Replaces Pipeline::allocate_buffers + FMDemodState::init
 * (pipeline.cu:1383-1475, fm_demod.cu:273-374). */
int cuvhs_create(const cuvhs_config* cfg, cuvhs_ctx** out);
int cuvhs_destroy(cuvhs_ctx* ctx);
int cuvhs_get_format(const cuvhs_ctx* ctx, cuvhs_format* out);
/* VideoFormat::VideoFormat (video_format.cpp:23-104) for NTSC, and for a
 * given system (PAL: video_format.cpp:48-87) */
int cuvhs_format_for(double sample_rate_mhz, cuvhs_format* out);
int cuvhs_format_for_system(int system, double sample_rate_mhz, cuvhs_format* out);
/* Clears the state carried across chunks (K2b cadence, ChromaState, field
 * parity: pipeline.h:93-97, chroma_decode.h:9-15, pipeline.cu:2575). */
int cuvhs_reset(cuvhs_ctx* ctx);

/* ---- chunk-level decode (Pipeline::process_chunk, pipeline.cu:1525-2624) ---- */
typedef struct cuvhs_chunk_info {
    int fields;               /* fields decoded in this chunk (0 = stop) */
    int fields_k1;            /* segments demodulated (n_read / spf) */
    int vsync_candidates;     /* raw candidate count before truncation */
    int status;               /* warning bits for this chunk */
    uint64_t raw_offset;      /* chunk start in the capture (samples) */
    uint64_t next_raw_offset; /* start of the next chunk */
    int guard_blocks_a;       /* mixed forms: K1 part-A / part-B blocks the FP64 */
    int guard_blocks_b;       /*   guard band recomputed (0 in the other forms) */
} cuvhs_chunk_info;

/* Device pointers to the last chunk's outputs, owned by the context and valid
 * until the next decode call.  Field i of the chunk:
 *   luma/chroma: [fields][263][910] u16 (PAL: [fields][313][1135])
 *   file_loc[i] (u64), is_first[i], phase_id[i], do_count[i],
 *   do_line/do_startx/do_endx: [fields][512] */
typedef struct cuvhs_chunk_outputs {
    const uint16_t* luma;
    const uint16_t* chroma;
    const uint64_t* file_loc;
    const int32_t* is_first;
    const int32_t* phase_id;
    const int32_t* do_count;
    const int32_t* do_line;
    const int32_t* do_startx;
    const int32_t* do_endx;
} cuvhs_chunk_outputs;

/* Decode one chunk (Pipeline::process_chunk).  d_chunk points at capture
 * sample raw_offset in device memory (samples of the context's input
 * format; all counts and offsets are in samples), with `avail` samples readable from
 * there (the rest of the capture, or at least max_fields * spfp); samples
 * past `avail` read as zero (pipeline.cu:1539-1541).  raw_offset is the
 * chunk's position in the capture (fileLoc, K2b cadence).  At most
 * max_fields (<= batch_fields) fields are decoded; info->next_raw_offset is
 * where the next chunk starts (pipeline.cu:1636-1651).  Fewer than one field
 * of samples (avail == 0 with a null pointer included) decodes nothing and
 * returns CUVHS_W_EOF in info->status. */
int cuvhs_decode_chunk(cuvhs_ctx* ctx, const void* d_chunk, uint64_t avail, uint64_t raw_offset,
                       int max_fields, cuvhs_chunk_info* info);
int cuvhs_chunk_outputs_get(const cuvhs_ctx* ctx, cuvhs_chunk_outputs* out);

/* Decode a whole device-resident capture (Pipeline::run, pipeline.cu:2626-2713)
 * into caller device buffers sized for max_out_fields fields.  Returns the
 * field count through *fields: 0 for a capture shorter than one field
 * (capture_len == 0 with a null pointer included).  More fields than
 * max_out_fields: CUVHS_E_CAPACITY. */
int cuvhs_decode_capture(cuvhs_ctx* ctx, const void* d_capture, uint64_t capture_len, int max_out_fields,
                         uint16_t* d_luma, uint16_t* d_chroma, uint64_t* d_file_loc, int32_t* d_is_first,
                         int32_t* d_phase_id, int32_t* d_do_count, int32_t* d_do_line, int32_t* d_do_startx,
                         int32_t* d_do_endx, int* fields);

/* ---- stage entry points on caller device buffers (parity tests) ----
 * All FP64 signal buffers hold nfields*spfp samples, laid out as the
 * reference's chunk buffers. */

/* fm_demod (fm_demod.h:46-53): one chunk read of the context's input format
 * (n_valid samples, zero past it), in the context's precision form (outputs
 * widened to FP64). */
int cuvhs_fm_demod(cuvhs_ctx* ctx, const void* d_raw, uint64_t n_valid, int nfields, double* d_demod,
                   double* d_demod05, double* d_env);
/* discover_vsyncs (vsync_discover.h:6-11): ascending candidates, returns the
 * raw count through *count (may exceed cap). */
int cuvhs_vsync_discover(cuvhs_ctx* ctx, const double* d_demod05, int total, int* d_idx, int cap, int* count);
/* sync_pulses (sync_pulses.h:20-27) with per-field offsets. */
int cuvhs_sync_pulses(cuvhs_ctx* ctx, const double* d_demod05, int total, const uint64_t* d_offsets, int nfields,
                      int* d_starts, int* d_lens, int* d_count);
/* K2b line location (pipeline.cu:1694-1791) on one chunk's K2 pulses, using
 * and updating the context's carried cadence state. */
int cuvhs_linelocs(cuvhs_ctx* ctx, uint64_t raw_offset, const uint64_t* d_offsets, int nfields, int* d_starts,
                   int* d_lens, int* d_count, int* d_types, double* d_linelocs, int* d_is_first);
/* line_locs (line_locs.h:48-56): K3 on K2 pulse lists ([nfields][800]);
 * d_types (optional) receives the nominal width classes, d_is_first the
 * field parity (1 first, 0 second, -1 unknown). */
int cuvhs_k3_linelocs(cuvhs_ctx* ctx, const int* d_starts, const int* d_lens, const int* d_count, int nfields,
                      int* d_types, double* d_linelocs, int* d_is_first);
/* hsync_refine (hsync_refine.h:18-22), in place. */
int cuvhs_hsync_refine(cuvhs_ctx* ctx, const double* d_demod05, int total, double* d_linelocs, int nfields);
/* tbc_resample (tbc_resample.h:16-27), debug counters omitted. */
int cuvhs_tbc_resample(cuvhs_ctx* ctx, const double* d_demod, int total, const double* d_linelocs, int nfields,
                       uint16_t* d_luma);
/* chroma_decode (chroma_decode.h:38-46): chunk read in, phase IDs out
 * (host array of nfields), using and updating the carried ChromaState. */
int cuvhs_chroma(cuvhs_ctx* ctx, const void* d_raw, uint64_t n_valid, const double* d_linelocs,
                 int nfields, uint16_t* d_chroma, int* h_phase_ids);
/* dropout_detect (dropout_detect.h:22-32), concealment in place. */
int cuvhs_dropouts(cuvhs_ctx* ctx, const double* d_env, const double* d_linelocs, int nfields, uint16_t* d_luma,
                   uint16_t* d_chroma, int* d_lines, int* d_starts, int* d_ends, int* d_count);

/* dropout_detect without a context and without allocating (no counterpart
 * in the reference: for callers that cannot allocate device memory, such as
 * code under HIP graph capture or tight serving loops).  Same inputs,
 * outputs and results as cuvhs_dropouts for a capture at sample_rate_mhz;
 * the kernels' scratch is a file-scope device array of 7.0 MB (held on each
 * device until the process ends) that holds at most 128 fields: more return
 * CUVHS_E_CAPACITY, with the reason in cuvhs_last_error(NULL).  Enqueues on
 * `stream` (hipStream_t, NULL = the current device's default stream) and
 * returns without waiting.  Calls that overlap on one device share the
 * scratch and race; serialise them (one stream, or events). */
int cuvhs_dropouts_noalloc(double sample_rate_mhz, const double* d_env, const double* d_linelocs, int nfields,
                           uint16_t* d_luma, uint16_t* d_chroma, int* d_lines, int* d_starts, int* d_ends,
                           int* d_count, void* stream);

/* Self-test of the convolution blocks the decoder runs (no counterpart in
 * the reference, whose transforms are cuFFT's, fm_demod.cu:396-486):
 * d_y = B * IFFT(FFT(d_x) * d_g) for one B-point block of complex values
 * (double2, or float2 with is_f32), d_g in natural bin order.  B = 4096
 * (K1 blocks, FP64 or FP32), 8192 (FP64 or FP32), 16384 (the chroma
 * pre-band-pass, FP64 or FP32), 1024 (the chroma line filter, FP64 or
 * FP32); other shapes return CUVHS_E_ARG. */
int cuvhs_selftest_block(int B, int is_f32, const void* d_x, const void* d_g, void* d_y);

/* Self-test of the FP64 arctangent of K1's phase (no counterpart in the
 * reference; replaces the atan2 of fm_demod.cu:203):
 * d_out[i] = atan2(d_y[i], d_x[i]) for i < n, device arrays of doubles. */
int cuvhs_selftest_atan2(const double* d_y, const double* d_x, double* d_out, int64_t n);

/* Per-stage device time (HIP events on the stream each stage runs on)
 * accumulated over chunk decodes while enabled; enabling clears the totals.
 * enable = 1 keeps the overlapped schedule (stages that share the GPU with
 * another chunk's stages read longer); enable = 2 runs every stage of every
 * chunk in sequence on the context's stream, so each total is the stage's
 * own time.  cuvhs_profile_read fills up to n stage totals (ms) and names
 * and returns the number of stages; the last one is the chroma
 * pre-band-pass.  (Counterpart of nothing in the reference, which only times
 * the whole chunk loop: pipeline.cu:2642-2708.) */
int cuvhs_profile(cuvhs_ctx* ctx, int enable);
int cuvhs_profile_read(const cuvhs_ctx* ctx, double* ms, const char** names, int n);
/* Work done by the chunk decodes counted in the same window: capture samples
 * demodulated by K1 (fields_k1 * spfp per chunk), chunks, and fields decoded.
 * Any pointer may be NULL.  (No counterpart in the reference.) */
int cuvhs_profile_work(const cuvhs_ctx* ctx, uint64_t* samples, int* chunks, int* fields);
/* K1's block shapes as designed for this context: block size (points) and
 * valid outputs per block of part A (analytic) and part B (video and sync
 * filters).  Any pointer may be NULL.  (No counterpart in the reference,
 * which transforms whole segments, fm_demod.cu:396-486.) */
int cuvhs_k1_blocks(const cuvhs_ctx* ctx, int* block_a, int* valid_a, int* block_b, int* valid_b);
/* Block size and outputs per block of the chroma pre-band-pass, and where
 * the context's three block sizes came from: 0 = the caller's config, 1 =
 * the autotuned table, 2 = the cost model (no counterpart in the reference). */
int cuvhs_prebpf_block(const cuvhs_ctx* ctx, int* block, int* valid, int* source);
/* Floating-point operations the matrix cores execute in K1 part A per
 * demodulated segment (mixed forms: every 32x32x16 product of the hi/lo tap
 * operands, padding steps included; 0 in the FFT forms), for the MFMA
 * utilisation figure.  (No counterpart in the reference.) */
int cuvhs_k1_matrix_flops(const cuvhs_ctx* ctx, double* per_segment);

/* The reference's debug-log counters (CUVHS_DEBUG_K3_LOG, _K4_LOG, _K5_LOG:
 * line_locs.cu:585-653, hsync_refine.cu:273-349, tbc_resample.cu:75-116),
 * collected per chunk while enabled; enabling clears the totals.  Order:
 * K3 bad_spacing, isolated_spacing, large_jump; K4 large_delta,
 * isolated_jump, refined_sync_like; K5 oob_pixels, bad_geom_lines,
 * sync_like_pixels.  cuvhs_diagnostics_read fills totals since enabling and
 * the last chunk's values (either may be NULL) and returns 9.  FP64 form
 * only (CUVHS_E_UNSUPPORTED otherwise). */
int cuvhs_diagnostics(cuvhs_ctx* ctx, int enable);
int cuvhs_diagnostics_read(const cuvhs_ctx* ctx, int64_t* totals, int32_t* last_chunk);

/* Last error text for the context; with NULL, of the last failed
 * context-free call (cuvhs_dropouts_noalloc). */
const char* cuvhs_last_error(const cuvhs_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* CUVHS_H */
