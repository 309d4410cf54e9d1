// Kernel launch parameter blocks and host launchers shared by the decoder's
// translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <vector>

#include "common.h"

namespace cuvhs {

struct K2bField;
struct K2bAnchorOut;

// Per-chunk values of the stage graphs of cuvhs_decode_capture, which are
// captured once per buffer slot and replayed for every chunk that slot
// takes: the host fills them in, a kernel writes them to device memory ahead
// of the chunk's head stage, and each kernel given a plan reads them in
// place of the values it was captured with (grids are then sized for the
// context's batch, and workgroups past the chunk's fields return).  Kernels
// launched directly get no plan.
struct ChunkPlan {
    const void* chunk;              // the chunk's first sample in the capture
    long long n_read;               // samples read from it (zero past them)
    long long total;                // fields_k1 * spfp: K1's outputs
    long long clear_end;            // K1 outputs of [total, clear_end) are cleared (a longer chunk's)
    unsigned long long raw_offset;  // the chunk's first sample's index in the capture
    int fields_k1;                  // fields K1 demodulates
    int max_fields;                 // fields the chunk may hold
    int vsync_cap;                  // VSYNC candidates kept
    // the chunk's first field in the output arrays (the caller's, or the
    // context's when the caller passed none)
    uint16_t* luma;
    uint16_t* chroma;
    unsigned long long* file_loc;
    int *is_first, *phase_id, *do_count, *do_line, *do_startx, *do_endx;
};

struct K1ALaunch {
    const void* cap;            // capture samples (format in_fmt)
    int in_fmt;                 // 0 u8, 1 s16, 2 u16
    const double* dc;           // per-read mean to subtract, or null
    long long n_valid, total;
    int spfp, N, lm, lp, V;
    double scale;
    const void* G;              // block spectrum, double2 or float2 (the transform precision)
    const void* tw;             // twiddles W_B^m, same type
    void* rawdemod;             // double, or float with out_f32
    void* env;                  // same type as rawdemod; may be null (not stored)
    int out_f32;
    ChunkCtl* ctl;              // if set, initialised by the first workgroup:
    long long ctl_raw_offset;   //   zeros, fields_k1 and raw_offset
    int ctl_fields_k1;
    unsigned int* thr_bits;     // if set, zeroed over [0, total) for part B
    const int2* blist;          // if set: only the blocks (segment, block) blist[i], i < *bcount
    const int* bcount;
    const ChunkPlan* plan;      // if set: cap, n_valid, total, ctl_* from the plan
};
// K1 part A on the matrix cores (mixed-precision forms, u8 input): see
// k1_mfma.hip.  Outputs are FP32.
struct K1MLaunch {
    const void* cap;            // u8 capture samples
    long long n_valid, total;
    int spfp, N, lp;
    int nc, nx;                 // steps of 16 taps (whole groups), staged input elements
    int glo0, glo1;             // groups of four steps run with their lo parts
    int a_lds_offset;           // byte offset of the tap operands in LDS (mfma_a_lds_offset)
    float scale;                // 256 (u8 scale) / tap scale
    float fs_over_tau;          // sample rate / 2 pi
    const uint4* A;             // tap operands [nc][4][64] x 16 B
    float* rawdemod;
    float* env;
    ChunkCtl* ctl;              // as in K1ALaunch
    long long ctl_raw_offset;
    int ctl_fields_k1;
    unsigned int* thr_bits;
    float2* blk_env;            // if set: (min, max) envelope of each tile of 1024
    int nblk;                   //   outputs, [field][nblk blocks][16 tiles]
    const ChunkPlan* plan;      // as in K1ALaunch
};
struct K1BLaunch {
    const void* rawdemod;       // double, or float with in_f32
    int in_f32;
    long long total;
    int spfp, N, lm, lp, V;
    const void* G;              // as in K1ALaunch
    const void* tw;
    void* demod;                // double, or float with demod_f32
    int demod_f32;
    void* demod05;              // double, or float with d05_f32
    int d05_f32;
    unsigned int* thr_bits;     // if set: bit s = (demod05[s] <= thr), s < total
    double thr;
    const int2* blist;          // as in K1ALaunch; the block's bitmask bits are
    const int* bcount;          // rewritten, its neighbours' bits kept
    unsigned char* gflag;       // if set: byte s >> kGuardSubShift of the chunk
    double gdelta;              // set to 1 when output s lies within gdelta of
                                // thr (never cleared here: the guard's
                                // compaction clears the flags it read)
    const ChunkPlan* plan;      // if set: total (and the fields) from the plan
};
// The guard band's block selection (guard.hip).
// Samples per guard flag of the reduced-precision part B: 256.
constexpr int kGuardSubShift = 8;
struct GuardLaunch {
    unsigned char* gflag;       // part-B threshold flags, one per 256 samples of
    long long nsub;             //   the chunk (nsub of them), cleared once read
    const float2* blk_env;      // MFMA part-A envelope tiles [segment][ntile], or null
    int ntile, mvalid;          // tiles per segment; samples per MFMA block
    float rho;                  // a tile below rho x the largest around flags the block
    int spfp, nseg, N;
    long long total;
    int VA, nbA;                // FP64 part A: outputs per block, blocks per segment
    int VB, nbB, BB, lpB;       // FP64 part B: outputs, blocks, block points, causal taps
    unsigned char* markA;       // [segment][nbA], [segment][nbB]: cleared by the compaction
    unsigned char* markB;
    int2* listA;                // listed blocks (segment, block)
    int2* listB;
    int* counts;                // [0] part A, [1] part B
    ChunkCtl* ctl;              // if set: receives the counts (guard_a, guard_b)
    const ChunkPlan* plan;      // if set: nsub, nseg, total from the plan
};
struct PreLaunch {
    const void* cap;
    int in_fmt;
    const double* dc;
    long long n_valid;
    int spfp, N2, lm, lp, V;
    const void* G;      // block spectrum: double2 (FP64 form) or float2 (FP32 form)
    const void* tw;     // twiddles of the same type
    float* out;
    const ChunkPlan* plan;  // if set: cap, n_valid and the field count from the plan
};
struct ChromaLaunch {
    const float* craw;
    const double* linelocs;
    const ChunkCtl* ctl;
    int nf_override;
    int spfp;
    double het_scale, fft_scale, burst_abs_ref;
    double het_dc, het_ds;      // cos / sin of 2 pi het_scale 64 (heterodyne step of 64 columns)
    int burst_start, lo, hi;
    int burst_len;              // burst window samples (36 NTSC, 40 PAL)
    int lpf, field_lines, line_len;   // line locations per frame; output lines, samples
    int nfft;                   // line transform points (1024 NTSC, 2048 PAL)
    const int* track;           // resolved track per field (output kernel)
    int line0, line_end;        // lines [line0, line_end) of each field
    double* burst;              // burst window of every line, track 0, scaled
    float* filt;                // band-passed lines, track 0, scaled: [field][lines][line_len]
    uint16_t* out;
    const double2* tw;          // W1024^m, m < 1024
    const float2* tw_f;         // the same twiddles in FP32 (f32)
    const double2* tw2;         // W2048^m, m < 2048 (nfft 2048)
    const float2* tw2_f;
    const double2* mask;        // band-pass bins: (1, 0) kept, (0, 0) removed
    int f32;                    // line filter in FP32 (the reduced-precision forms)
    const ChunkPlan* plan;      // if set: out (and the phase IDs) from the plan
};


// f32: transform in FP32 (float2 spectra and twiddles) instead of FP64
hipError_t run_k1_part_a(int B, int f32, const K1ALaunch& a, int nf, hipStream_t st);
hipError_t run_k1_part_b(int B, int f32, const K1BLaunch& a, int nf, hipStream_t st);
hipError_t run_prebpf(int B, int f32, const PreLaunch& a, int nf, hipStream_t st);
// K1 part A on the matrix cores (bf16: bf16 operands, else f16)
hipError_t run_k1_part_a_mfma(int bf16, const K1MLaunch& a, int nf, hipStream_t st);
int mfma_blocks_per_segment(int spfp);
int mfma_valid_per_block();
int mfma_input_elems(int nc);
int mfma_a_lds_offset(int nx);
// A operands of the MFMA part A for an impulse response taps[m + lm], m =
// -lm..lp; returns the step count, *scale the tap scale to undo and
// [*glo0, *glo1) the groups of four steps whose lo parts are kept
int build_mfma_taps(const std::vector<std::complex<double>>& taps, int lm, int lp, int bf16, std::vector<uint4>& out,
                    double* scale, int* glo0, int* glo1);
// n samples of a float (f32) or double signal into a double array
hipError_t run_to_f64(const void* src, int f32, double* dst, long long n, hipStream_t st);
// the guard band's FP64 block lists (marks, compaction, counts)
hipError_t run_guard_lists(const GuardLaunch& g, hipStream_t st);
hipError_t run_fill_f64(double* p, long long n, double v, hipStream_t st);
// zero n 32-bit words (a fill kernel, not a memset: safe under graph capture)
hipError_t run_fill_zero(void* p, long long words, hipStream_t st);
hipError_t run_k3(const ChunkCtl* ctl, int nf_override, int max_fields, const Fmt& fmt, const int* starts,
                  const int* lens, const int* count, int* types, double* linelocs, int* is_first, hipStream_t st);
hipError_t run_diag_spacing(const double* ll, ChunkCtl* ctl, int max_fields, const Fmt& fmt, hipStream_t st);
hipError_t run_diag_k4(const double* d05, const double* before, const double* after, ChunkCtl* ctl, int max_fields,
                       const Fmt& fmt, hipStream_t st);
hipError_t run_diag_k5(const double* demod, const double* ll, ChunkCtl* ctl, int max_fields, const Fmt& fmt,
                       hipStream_t st);
hipError_t run_input_mean(const void* cap, int in_fmt, long long n, double* partial, double* mean, hipStream_t st,
                          const ChunkPlan* plan = nullptr);
// [plan total, plan clear_end) of three K1 output buffers with w0 / w1 / w2
// 32-bit words per sample to zero
hipError_t run_clear_past(void* p0, int w0, void* p1, int w1, void* p2, int w2, const ChunkPlan* plan,
                          hipStream_t st);
// a chunk's plan into device memory, by value
hipError_t run_put_plan(ChunkPlan* dst, const ChunkPlan& v, hipStream_t st);
// one convolution block of the kinds the decoder runs (gp in the block's bin order)
hipError_t run_block_selftest(int B, int is_f32, const void* x, const void* gp, void* y, const void* tw,
                              hipStream_t st);
hipError_t run_atan2_selftest(const double* y, const double* x, double* out, long long n, hipStream_t st);
int events_nwg(long long total);
size_t events_mask_count(long long total);  // uint2 entries of the masks buffer
size_t thr_bits_words(long long total);     // 32-bit words of a threshold bitmask
// events from the threshold bitmask (bit s = x[s] <= thr), as K1 part B writes it
hipError_t run_events_bits(const unsigned int* bits, long long total, int* wg_count, int* wg_off, uint2* masks,
                           int* events, ChunkCtl* ctl, hipStream_t st, const ChunkPlan* plan = nullptr);
// the same from a demod_05 array: bitmask into bits, then run_events_bits
hipError_t run_events(const double* x, long long total, double thr, unsigned int* bits, int* wg_count, int* wg_off,
                      uint2* masks, int* events, ChunkCtl* ctl, hipStream_t st);
// host_out (optional): pinned host copy of the finished control block
hipError_t run_vsync_cluster(const int* events, ChunkCtl* ctl, long long total, int min_w, int max_w, int cap,
                             int* cand, unsigned long long* offsets, int num_fields, int spf, int spl, int fields_k1,
                             int do_cluster, ChunkCtl* host_out, hipStream_t st, const ChunkPlan* plan = nullptr);
hipError_t run_sync_pulses(const int* events, const ChunkCtl* ctl, const unsigned long long* offsets, int spfp,
                           int* starts, int* lens, int* count, int max_fields, int nf_override, hipStream_t st);
hipError_t run_k2b(const ChunkCtl* ctl_dev, ChunkCtl* ctl_mut, DecodeState* dst, int nf_override, int max_fields, const Fmt& fmt,
                   const unsigned long long* offsets, const int* k2_starts, const int* k2_lens, const int* k2_count,
                   int* vp_start, int* vp_len, unsigned char* vp_tg, K2bField* summ, K2bAnchorOut* anchors,
                   int* out_starts, int* out_lens, int* out_count, int* out_types, double* linelocs, int* is_first,
                   hipStream_t st);
// the two halves of run_k2b: everything per field (needs only the chunk's
// K2 pulses), then the cross-field cadence and the line locations
hipError_t run_k2b_fields(const ChunkCtl* ctl_dev, int nf_override, int max_fields, const Fmt& fmt,
                          const int* k2_starts, const int* k2_lens, const int* k2_count, int* vp_start, int* vp_len,
                          unsigned char* vp_tg, K2bField* summ, hipStream_t st);
hipError_t run_k2b_cadence(const ChunkCtl* ctl_dev, ChunkCtl* ctl_mut, DecodeState* dst, int nf_override,
                           int max_fields, const Fmt& fmt, const unsigned long long* offsets, const int* vp_start,
                           const int* vp_len, const unsigned char* vp_tg, const K2bField* summ, K2bAnchorOut* anchors,
                           int* out_starts, int* out_lens, int* out_count, int* out_types, double* linelocs,
                           int* is_first, hipStream_t st);
size_t k2b_field_summary_bytes();
size_t k2b_anchor_bytes();
// f32: the signal is stored as float (reduced-precision forms), else double;
// reduced: demod_05 comes from a reduced-precision K1 (FP32-quantized), so
// the rising-edge interpolation drops the reference's symmetric-sample guard
hipError_t run_hsync_refine(const void* d05, int f32, double* ll, const ChunkCtl* ctl, int nf_override,
                            int max_fields, const Fmt& fmt, hipStream_t st, int reduced = 0);
hipError_t run_tbc_resample(const void* demod, int f32, const double* ll, uint16_t* luma, const ChunkCtl* ctl,
                            int nf_override, int max_fields, const Fmt& fmt, hipStream_t st,
                            const ChunkPlan* plan = nullptr);
hipError_t run_chroma(ChromaLaunch a, DecodeState* dst, int max_fields, double* m0, double* m1,
                      const double* cs_tab, double* iq0, int* out_track, int* phase_ids, int need_iq,
                      hipStream_t st);
size_t dropout_class_bytes(int max_fields);
size_t dropout_partial_count(int max_fields);
hipError_t run_dropouts(const double* env, const double* ll, const ChunkCtl* ctl, int nf_override, int max_fields,
                        const Fmt& fmt, double* partial, unsigned char* cls, uint16_t* luma, uint16_t* chroma,
                        int* lines, int* starts, int* ends, int* count, hipStream_t st);
// the two halves of run_dropouts: envelope sums and block classes (need only
// env and line locations), then the scan and concealment (luma and chroma
// written first)
hipError_t run_dropout_classes(const void* env, int f32, const double* ll, const ChunkCtl* ctl, int nf_override,
                               int max_fields, const Fmt& fmt, double* partial, unsigned char* cls, hipStream_t st);
hipError_t run_dropout_scan(const unsigned char* cls, const double* ll, const ChunkCtl* ctl, int nf_override,
                            int max_fields, const Fmt& fmt, uint16_t* luma, uint16_t* chroma, int* lines, int* starts,
                            int* ends, int* count, hipStream_t st, const ChunkPlan* plan = nullptr);
hipError_t run_field_meta(const ChunkCtl* ctl, DecodeState* dst, const unsigned long long* offsets,
                          const int* is_first_k2b, unsigned long long* file_loc, int* is_first_out, hipStream_t st,
                          const ChunkPlan* plan = nullptr);

// launch parameter blocks (layouts shared with the kernel files)
}  // namespace cuvhs
