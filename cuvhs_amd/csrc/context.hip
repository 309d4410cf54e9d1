// Decoder context, chunk driver and the C ABI (include/cuvhs.h).
//
// Reference: Pipeline::allocate_buffers / process_chunk / run
// (pipeline.cu:1383-1475, 1525-2624, 2626-2713).  A chunk is enqueued as a
// fixed sequence of kernels; the field count that VSYNC clustering produces
// stays in a device control block that every later kernel reads.  A chunk has
// a head stage (K1, sync events, VSYNC clustering) and a tail stage (pulses,
// line locations, refinement, resampling, chroma, dropouts, metadata) on two
// streams with two sets of per-chunk buffers, so the head of chunk c+1 runs
// while the tail of chunk c does.  The host waits once per chunk, for the
// head's control block (next chunk offset, field count).
#include <hip/hip_runtime.h>

#include <sched.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cuvhs.h"
#include "common.h"
#include "design.h"
#include "fft_lds.h"
#include "launch.h"
#include "tuning_gfx950.h"

using namespace cuvhs;

constexpr int kStages = 13;
// Chunk buffer sets in flight: the head of chunk c+1 may start once the tail
// of chunk c+1-kSlots has finished, so with three the heads never wait for
// the tail launched just before them.
constexpr int kSlots = 3;
// Stage graphs of a slot (graph mode): the per-read mean (--dc-correct), the
// head (K1, guard band, events, VSYNC), the pre-band-pass, the pulse stage
// (K2, K2b per field), and the tail in four pieces between its cross-stream
// waits (line locations and K4; K5 and the dropout classes on the fork
// stream; chroma; the dropout scan, metadata and the control block's copy).
enum { kGMean, kGHead, kGPre, kGPulse, kGTail1, kGFork, kGTail2, kGTail3, kGraphKinds };
constexpr int kMarks = 18;   // 0-5 head, 6-14 tail, 15-17 the early K2 / K2b-field work
// Guard band of the mixed forms.  A part-B block of the FP64 design is
// recomputed in FP64 when a demod_05 output of the FP32 part B in its 256-
// sample neighbourhood lies within delta Hz of the sync threshold, or when an
// envelope tile in its input is below rho times the largest there.  Measured
// without the guard on the 600-field 28 MHz capture (3.0e8 samples, binned
// by the envelope over the largest envelope within two tiles of 1024): the
// f16 form's demod_05 error is at most 1.55 Hz in every bin, 1.02 Hz among
// the samples below 0.005 (dropouts); the bf16 form's is at most 1.55 Hz
// there but 28.7 Hz next to dropouts, where part B's filter spreads the
// larger bf16 phase error of the dip (16-bit taps) onto full-envelope
// samples.  delta: 8 Hz for f16 (five times its largest error), 32 Hz for
// bf16 (eight times its 4.0 Hz outside the recomputed blocks).  rho keeps
// dropouts in FP64 in both: a dip flattens the sync edges K4 interpolates,
// so even f16's 1.5 Hz moves a refined line start there enough to change
// chroma (a 3.58 MHz carrier) by up to 3 LSB (71 samples over 1 LSB in 140
// fields with rho = 0.002, none with 0.1); it also covers the zero-filled
// end of a capture, where the phase is ill-conditioned in any precision.
constexpr double kGuardDeltaF16 = 8.0;
constexpr float kGuardRhoF16 = 0.1f;
constexpr double kGuardDeltaBf16 = 32.0;
constexpr float kGuardRhoBf16 = 0.5f;

// Buffers and control block of one chunk in flight.  Two slots alternate so
// that the head stage of chunk c+1 (K1, sync events, VSYNC clustering) can
// run while the tail stage of chunk c (pulses, line locations, refinement,
// resampling, chroma, dropouts) still reads chunk c's signals.
struct Slot {
    double* demod = nullptr;
    double* d05 = nullptr;
    double* env = nullptr;
    int* events = nullptr;
    int* wg_count = nullptr;
    uint2* ev_masks = nullptr;
    int* wg_off = nullptr;
    unsigned long long* offsets = nullptr;
    ChunkCtl* ctl = nullptr;      // device
    ChunkCtl* hctl = nullptr;     // pinned host copy
    float* craw = nullptr;        // chroma pre-band-pass output
    double* dc = nullptr;         // per-read mean (--dc-correct)
    double* dc_partial = nullptr; // its partial sums
    hipEvent_t dc_done = nullptr;
    hipEvent_t head_done = nullptr, tail_done = nullptr, pre0 = nullptr, pre1 = nullptr;
    hipEvent_t k4_done = nullptr, aux_done = nullptr;  // fork / join of the tail's side-stream work
    hipEvent_t k2_done = nullptr;  // K2 pulses and K2b's per-field work (side stream)
    // K2 pulses and K2b's per-field results: written on the side stream as
    // soon as the chunk's head is done, while the tail of the chunk before
    // still reads its own
    int *k2s = nullptr, *k2l = nullptr, *k2c = nullptr;
    int *vps = nullptr, *vpl = nullptr;
    unsigned char* vptg = nullptr;
    void* summ = nullptr;
    hipEvent_t mark[kMarks] = {};
    // host view of the chunk
    const void* chunk = nullptr;
    uint64_t n_read = 0, raw_offset = 0;
    int max_fields = 0, fk1 = 0;
    // demod/d05/env hold K1 output in [0, k1_extent) and zeros beyond
    long long k1_extent = 0;
    // graph mode of cuvhs_decode_capture: the chunk's plan (host copy, filled
    // when the chunk is planned, and its device copy, written ahead of the
    // chunk's head) and the slot's stage graphs, captured at first use and
    // replayed for every chunk
    ChunkPlan hplan{};
    ChunkPlan* plan = nullptr;
    hipGraphExec_t graph[kGraphKinds] = {};
};

struct cuvhs_ctx {
    cuvhs_config cfg{};
    Fmt fmt{};
    int F = 0;
    hipStream_t stream = nullptr;   // head stage (and the caller's stream, if given)
    bool own_stream = false;
    hipStream_t tail = nullptr;     // tail stage
    hipStream_t side = nullptr;     // chroma pre-band-pass, next to the tail's sync stages
    hipStream_t aux = nullptr;      // the tail's forked work, when the process has a queue for it
    bool aux_queue = false;         // more hardware queues than the default four
    std::string err;
    K1Design k1;
    PreDesign pre;
    int k1_f32 = 0;            // K1's FFT blocks in FP32 (FP32 form; part B of the mixed forms)
    int mixed = 0, mixed_bf16 = 0;  // mixed forms: K1 part A on the matrix cores (f16 or bf16)
    int sig_f32 = 0;           // demod, envelope (and rawdemod) stored as float
    int d05_f32 = 0;           // demod_05 stored as float (FP32 form)
    // mixed forms: part A tap operands, and the FP64 guard band
    uint4* m_taps = nullptr;
    int m_nc = 0, m_nx = 0, m_nblk = 0;
    int m_glo0 = 0, m_glo1 = 0;        // groups of four tap steps run with their lo parts
    double m_scale = 0.0;
    K1Design k1g;              // the FP64 design the guard band recomputes with
    void *G_ga = nullptr, *G_gb = nullptr, *tw_ga = nullptr, *tw_gb = nullptr;
    unsigned char* g_flag = nullptr;   // 256-sample flags of part-B outputs near the threshold
    float2* g_env = nullptr;           // envelope range per tile of 1024 outputs
    unsigned char *g_markA = nullptr, *g_markB = nullptr;
    int2 *g_listA = nullptr, *g_listB = nullptr;
    int* g_counts = nullptr;
    double* g_raw64 = nullptr;         // FP64 rawdemod of the listed part-A blocks
    double g_delta = 0.0;              // threshold band (Hz) of the FP32 part B
    float g_rho = 0.0f;                // envelope ratio that marks an ill-conditioned phase
    void* G_a = nullptr;       // K1 block spectra and twiddles: double2, or float2 when k1_f32
    void* G_b = nullptr;
    void* G_p = nullptr;       // pre-band-pass spectrum: double2, or float2 when pre_f32
    void* tw_a = nullptr;
    void* tw_b = nullptr;
    void* tw_p = nullptr;
    int pre_f32 = 0;           // chroma pre-band-pass transforms in FP32 (CUVHS_PRECISION_FP32)
    double2* tw_1024 = nullptr;
    float2* tw_1024f = nullptr;     // the chroma line filter's twiddles in FP32 (chroma_f32)
    double2* tw_2048 = nullptr;     // PAL's 2048-point line transform: W2048^m
    float2* tw_2048f = nullptr;
    int chroma_f32 = 0;             // chroma line filter in FP32 (the reduced-precision forms)
    int block_source = 2;           // block sizes from: 0 the config, 1 the tuned table, 2 the cost model
    int graphs = 0;                 // cuvhs_decode_capture replays captured stage graphs (cuvhs_config.graphs)
    // the dropout sums and classes on the tail stream (after chroma) instead
    // of forked with K5: in the mixed forms, whose head stream, where the
    // fork lands with four hardware queues, also carries the guard band
    // (measured on the 600-field decode: mixed-fp16 70.1k / 69.9k against
    // 68.2k / 68.4k fields/s forked; FP32 72.1k against 73.0k / 72.3k)
    bool classes_on_tail = false;
    double* cs_tab = nullptr;       // cos/sin of the burst phase detector
    double2* chroma_mask = nullptr; // per-line band-pass: (1, 0) kept bins, (0, 0) others
    float* chroma_filt = nullptr;   // band-passed chroma lines of a chunk (track 0)
    long long buf_n = 0;            // F * spfp + one segment of zero padding
    Slot slot[kSlots];
    double* rawdemod = nullptr;     // head only
    int* cand = nullptr;            // head only
    unsigned int* thr_bits = nullptr;  // head only: sync-threshold bitmask of d05
    int *ps = nullptr, *pl = nullptr, *pc = nullptr, *pt = nullptr;
    void* anchors = nullptr;
    double* ll = nullptr;
    int* isf_k2b = nullptr;
    uint16_t* luma = nullptr;
    uint16_t* chroma = nullptr;
    double* burst = nullptr;
    double *m0 = nullptr, *m1 = nullptr, *iq0 = nullptr;
    int* track = nullptr;
    int* phase = nullptr;
    int *dol = nullptr, *dos = nullptr, *doe = nullptr, *doc = nullptr;
    double* do_partial = nullptr;
    unsigned char* do_cls = nullptr;
    unsigned long long* file_loc = nullptr;
    int* isf_out = nullptr;
    DecodeState* state = nullptr;   // device
    DecodeState* hstate = nullptr;  // pinned, for resets
    std::vector<void*> allocs;
    bool prof = false;
    bool serial = false;            // profiling: every stage on the head stream, no overlap
    double stage_ms[kStages + 1] = {};
    bool diag = false;              // K3/K4/K5 diagnostic counters
    long long diag_total[9] = {};
    int diag_last[9] = {};
    double* ll_coarse = nullptr;    // line locations before K4 (diagnostics)
    uint64_t work_samples = 0;      // profiling: K1 samples demodulated
    int work_chunks = 0, work_fields = 0;
};

namespace {

int fail(cuvhs_ctx* c, hipError_t e, const char* what) {
    if (c) c->err = std::string(what) + ": " + hipGetErrorString(e);
    return CUVHS_E_HIP;
}

#define HIPCHK(c, x)                                   \
    do {                                               \
        hipError_t e_ = (x);                           \
        if (e_ != hipSuccess) return fail(c, e_, #x);  \
    } while (0)

template <typename T>
hipError_t dalloc(cuvhs_ctx* c, T** p, size_t n) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    e = hipMemset(v, 0, std::max<size_t>(n, 1) * sizeof(T));
    c->allocs.push_back(v);
    *p = (T*)v;
    return e;
}

template <typename V, typename S>
std::vector<V> twiddles(int B) {
    std::vector<V> t((size_t)B);
    const long double two_pi = 6.283185307179586476925286766559L;
    for (int m = 0; m < B; m++) {
        long double a = two_pi * (long double)m / (long double)B;
        t[(size_t)m].x = (S)cosl(a);
        t[(size_t)m].y = (S)(-sinl(a));
    }
    return t;
}

// Bin order of the wave-local convolution blocks (fft_lds.h) of size n.
size_t block_bin(size_t n, size_t k) {
    if (n == 4096) return (size_t)conv4096_gperm((int)k);
    if (n == 8192) return (size_t)conv_rows_gperm<512>((int)k);
    return (size_t)conv_rows_gperm<1024>((int)k);
}

// perm: store bin k at block_bin(size, k)
template <typename V, typename S>
hipError_t upload_spectrum(cuvhs_ctx* c, V** dst, const std::vector<cdbl>& s, bool perm = false) {
    std::vector<V> h(s.size());
    for (size_t i = 0; i < s.size(); i++) {
        size_t j = perm ? block_bin(s.size(), i) : i;
        h[j].x = (S)s[i].real();
        h[j].y = (S)s[i].imag();
    }
    hipError_t e = dalloc(c, dst, h.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice);
}

hipError_t upload_twiddles(cuvhs_ctx* c, void** dst, int B, int f32) {
    hipError_t e;
    if (f32) {
        auto t = twiddles<float2, float>(B);
        if ((e = dalloc(c, (float2**)dst, t.size())) != hipSuccess) return e;
        return hipMemcpy(*dst, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice);
    }
    auto t = twiddles<double2, double>(B);
    if ((e = dalloc(c, (double2**)dst, t.size())) != hipSuccess) return e;
    return hipMemcpy(*dst, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice);
}

// The autotuned block sizes for a sample rate and precision form, if the
// table has a row for them.
const TunedBlocks* tuned_blocks(double mhz, int precision) {
    for (const TunedBlocks& t : kTunedBlocks)
        if (t.precision == precision && std::fabs(t.mhz - mhz) <= 0.05) return &t;
    return nullptr;
}

int sample_bytes(const cuvhs_ctx* c) { return c->cfg.input_format == CUVHS_INPUT_U8 ? 1 : 2; }

void reset_state(DecodeState* h) {
    memset(h, 0, sizeof(*h));
    h->k2b.prev_first_field = -1;
    h->k2b.prev_progressive_field = -1;
    h->k2b.prev_first_hsync_readloc = -1;
    h->k2b.prev_first_hsync_loc = -1.0;
    h->k2b.prev_first_hsync_diff = -1.0;
    h->chroma.valid = 0;
    h->last_parity = 1;
}

// Hardware queues HIP gives this process (GPU_MAX_HW_QUEUES, default 4).
// Streams beyond them share queues, and work on a shared queue runs in
// submission order whatever its stream.
int hw_queues() {
    const char* v = getenv("GPU_MAX_HW_QUEUES");
    const int n = v ? atoi(v) : 0;
    return n > 0 ? n : 4;
}

void mark(cuvhs_ctx* c, Slot& s, int i, hipStream_t st) {
    if (c->prof) (void)hipEventRecord(s.mark[i], st);
}

// Per-line chroma band-pass bins (chroma_decode.cu:538-545).
void chroma_band(const Fmt& f, int* lo, int* hi) {
    int fsc_bin = (int)(f.fsc / f.output_rate * f.chroma_fft + 0.5);
    int bw = (int)(500000.0 / f.output_rate * f.chroma_fft + 0.5);
    if (bw < 10) bw = 10;
    *lo = fsc_bin - bw;
    *hi = fsc_bin + bw;
}

ChromaLaunch chroma_launch(cuvhs_ctx* c, const ChunkCtl* ctl, int nf_override, const float* craw) {
    ChromaLaunch a{};
    a.craw = craw;
    a.linelocs = c->ll;
    a.ctl = ctl;
    a.nf_override = nf_override;
    a.spfp = c->fmt.spfp;
    a.het_scale = (c->fmt.fsc + c->fmt.chroma_under) / c->fmt.output_rate;
    a.het_dc = cos(2.0 * M_PI * a.het_scale * 64.0);
    a.het_ds = sin(2.0 * M_PI * a.het_scale * 64.0);
    a.fft_scale = 1.0 / (double)c->fmt.chroma_fft;
    a.burst_abs_ref = c->fmt.burst_abs_ref;
    a.burst_start = c->fmt.burst_start;
    a.burst_len = c->fmt.burst_end - c->fmt.burst_start;
    a.lpf = c->fmt.lines_per_frame;
    a.field_lines = c->fmt.field_lines;
    a.line_len = c->fmt.line_len;
    a.nfft = c->fmt.chroma_fft;
    chroma_band(c->fmt, &a.lo, &a.hi);
    a.mask = c->chroma_mask;
    a.burst = c->burst;
    a.filt = c->chroma_filt;
    a.out = c->chroma;
    a.tw = c->tw_1024;
    a.tw_f = c->tw_1024f;
    a.tw2 = c->tw_2048;
    a.tw2_f = c->tw_2048f;
    a.f32 = c->chroma_f32;
    return a;
}

// ctl (optional): the chunk's control block, initialised by K1's first
// workgroup.  thr_bits (optional): receives the sync-threshold bitmask of d05.
// Outputs go to the slot's demod / d05 / env in the form's storage types.
// plan (graph mode): the kernels read the chunk's values from it; nfk1 is then
// the batch (the grids' size) and K1's outputs past the chunk are cleared on
// the device from the plan's extent.
int k1_launch(cuvhs_ctx* c, Slot& s, const void* chunk, long long n_valid, int nfk1, hipStream_t st,
              ChunkCtl* ctl = nullptr, unsigned int* thr_bits = nullptr, const ChunkPlan* plan = nullptr) {
    const Fmt& f = c->fmt;
    long long T = (long long)nfk1 * f.spfp;
    const double fs_tau = f.sample_rate / (2.0 * 3.14159265358979323846);
    if (c->mixed) {
        // part A on the matrix cores: FP32 rawdemod and envelope, and the
        // envelope of each tile of 1024 outputs for the guard band
        K1MLaunch m{};
        m.cap = chunk;
        m.n_valid = n_valid;
        m.total = T;
        m.spfp = f.spfp;
        m.N = c->k1.N;
        m.lp = c->k1.rf.lp;
        m.nc = c->m_nc;
        m.nx = c->m_nx;
        m.glo0 = c->m_glo0;
        m.glo1 = c->m_glo1;
        m.a_lds_offset = mfma_a_lds_offset(c->m_nx);
        m.scale = (float)c->m_scale;
        m.fs_over_tau = (float)fs_tau;
        m.A = c->m_taps;
        m.rawdemod = reinterpret_cast<float*>(c->rawdemod);
        m.env = reinterpret_cast<float*>(s.env);
        m.ctl = ctl;
        m.ctl_raw_offset = (long long)s.raw_offset;
        m.ctl_fields_k1 = s.fk1;
        m.thr_bits = thr_bits;
        m.blk_env = c->g_env;
        m.nblk = c->m_nblk;
        m.plan = plan;
        HIPCHK(c, run_k1_part_a_mfma(c->mixed_bf16, m, nfk1, st));
    } else {
        K1ALaunch a{};
        a.cap = chunk;
        a.in_fmt = c->cfg.input_format;
        a.dc = c->cfg.dc_correct ? s.dc : nullptr;
        a.n_valid = n_valid;
        a.total = T;
        a.spfp = f.spfp;
        a.N = c->k1.N;
        a.lm = c->k1.rf.lm;
        a.lp = c->k1.rf.lp;
        a.V = c->k1.rf.valid;
        a.scale = fs_tau;
        a.G = c->G_a;
        a.tw = c->tw_a;
        a.rawdemod = c->rawdemod;
        a.env = s.env;
        a.out_f32 = c->sig_f32;
        a.ctl = ctl;
        a.ctl_raw_offset = (long long)s.raw_offset;
        a.ctl_fields_k1 = s.fk1;
        a.thr_bits = thr_bits;
        a.plan = plan;
        HIPCHK(c, run_k1_part_a(c->k1.rf.B, c->k1_f32, a, nfk1, st));
    }
    mark(c, s, 1, st);
    K1BLaunch b{};
    b.rawdemod = c->rawdemod;
    b.in_f32 = c->sig_f32;
    b.total = T;
    b.spfp = f.spfp;
    b.N = c->k1.N;
    b.lm = c->k1.video.lm;
    b.lp = c->k1.video.lp;
    b.V = c->k1.video.valid;
    b.G = c->G_b;
    b.tw = c->tw_b;
    b.demod = s.demod;
    b.demod_f32 = c->sig_f32;
    b.demod05 = s.d05;
    b.d05_f32 = c->d05_f32;
    b.thr_bits = thr_bits;
    b.thr = f.pulse_threshold_hz;
    b.plan = plan;
    if (c->mixed) {
        b.gflag = c->g_flag;
        b.gdelta = c->g_delta;
    }
    HIPCHK(c, run_k1_part_b(c->k1.video.B, c->k1_f32, b, nfk1, st));
    mark(c, s, 2, st);
    if (c->mixed) {
        // FP64 guard band: the blocks whose sync decisions the reduced
        // precision could change, recomputed with the FP64 design
        GuardLaunch g{};
        g.gflag = c->g_flag;
        g.nsub = ((T - 1) >> kGuardSubShift) + 1;
        g.blk_env = c->g_env;
        g.ntile = c->m_nblk * 16;
        g.mvalid = mfma_valid_per_block();
        g.rho = c->g_rho;
        g.spfp = f.spfp;
        g.nseg = nfk1;
        g.N = c->k1g.N;
        g.total = T;
        g.VA = c->k1g.rf.valid;
        g.nbA = (f.spfp + g.VA - 1) / g.VA;
        g.VB = c->k1g.video.valid;
        g.nbB = (f.spfp + g.VB - 1) / g.VB;
        g.BB = c->k1g.video.B;
        g.lpB = c->k1g.video.lp;
        g.markA = c->g_markA;
        g.markB = c->g_markB;
        g.listA = c->g_listA;
        g.listB = c->g_listB;
        g.counts = c->g_counts;
        g.ctl = ctl;
        g.plan = plan;
        HIPCHK(c, run_guard_lists(g, st));
        K1ALaunch a{};
        a.cap = chunk;
        a.in_fmt = c->cfg.input_format;
        a.n_valid = n_valid;
        a.total = T;
        a.spfp = f.spfp;
        a.N = c->k1g.N;
        a.lm = c->k1g.rf.lm;
        a.lp = c->k1g.rf.lp;
        a.V = c->k1g.rf.valid;
        a.scale = fs_tau;
        a.G = c->G_ga;
        a.tw = c->tw_ga;
        a.rawdemod = c->g_raw64;
        a.blist = c->g_listA;
        a.bcount = c->g_counts;
        a.plan = plan;
        HIPCHK(c, run_k1_part_a(c->k1g.rf.B, 0, a, nfk1, st));
        K1BLaunch gb{};
        gb.rawdemod = c->g_raw64;
        gb.total = T;
        gb.spfp = f.spfp;
        gb.N = c->k1g.N;
        gb.lm = c->k1g.video.lm;
        gb.lp = c->k1g.video.lp;
        gb.V = c->k1g.video.valid;
        gb.G = c->G_gb;
        gb.tw = c->tw_gb;
        gb.demod = s.demod;
        gb.demod_f32 = c->sig_f32;
        gb.demod05 = s.d05;
        gb.d05_f32 = c->d05_f32;
        gb.thr_bits = thr_bits;
        gb.thr = f.pulse_threshold_hz;
        gb.blist = c->g_listB;
        gb.bcount = c->g_counts + 1;
        gb.plan = plan;
        HIPCHK(c, run_k1_part_b(c->k1g.video.B, 0, gb, nfk1, st));
    }
    mark(c, s, 3, st);
    // beyond the demodulated region the buffers read as zero: only what an
    // earlier, longer chunk left there needs clearing
    const int ws = c->sig_f32 ? 1 : 2, wd = c->d05_f32 ? 1 : 2;   // 32-bit words per sample
    if (plan) {
        HIPCHK(c, run_clear_past(s.demod, ws, s.d05, wd, s.env, ws, plan, st));
        return 0;
    }
    if (s.k1_extent > T) {
        const long long n = s.k1_extent - T;
        HIPCHK(c, run_fill_zero(reinterpret_cast<unsigned int*>(s.demod) + T * ws, n * ws, st));
        HIPCHK(c, run_fill_zero(reinterpret_cast<unsigned int*>(s.d05) + T * wd, n * wd, st));
        HIPCHK(c, run_fill_zero(reinterpret_cast<unsigned int*>(s.env) + T * ws, n * ws, st));
    }
    s.k1_extent = T;
    return 0;
}

int prebpf_launch(cuvhs_ctx* c, const void* chunk, long long n_valid, int nf_max, float* out, const double* dc,
                  hipStream_t st, const ChunkPlan* plan = nullptr) {
    PreLaunch p{};
    p.cap = chunk;
    p.in_fmt = c->cfg.input_format;
    p.dc = c->cfg.dc_correct ? dc : nullptr;
    p.n_valid = n_valid;
    p.spfp = c->fmt.spfp;
    p.N2 = c->pre.N2;
    p.lm = c->pre.bpf.lm;
    p.lp = c->pre.bpf.lp;
    p.V = c->pre.bpf.valid;
    p.G = c->G_p;
    p.tw = c->tw_p;
    p.out = out;
    p.plan = plan;
    HIPCHK(c, run_prebpf(c->pre.bpf.B, c->pre_f32, p, nf_max, st));
    return 0;
}

// RawReader::read_at + the chunk sizing of Pipeline::run / process_chunk
// (pipeline.cu:1530-1548, 2651-2658).  Returns false when no chunk starts here.
bool plan_chunk(cuvhs_ctx* c, Slot& s, const void* chunk, uint64_t avail, uint64_t raw_offset, int max_fields) {
    const uint64_t spf = (uint64_t)c->fmt.samples_per_field;
    uint64_t n_read = std::min<uint64_t>((uint64_t)max_fields * (uint64_t)c->fmt.spfp, avail);
    if (n_read < spf) return false;
    s.chunk = chunk;
    s.n_read = n_read;
    s.raw_offset = raw_offset;
    s.max_fields = max_fields;
    s.fk1 = (int)std::min<uint64_t>(n_read / spf, (uint64_t)max_fields);
    return true;
}

// The chroma pre-band-pass of a chunk, which needs only the capture, on the
// side stream (after the head stage's input mean under --dc-correct).
int pre_stage(cuvhs_ctx* c, Slot& s) {
    hipStream_t side = c->serial ? c->stream : c->side;
    if (c->cfg.dc_correct) HIPCHK(c, hipStreamWaitEvent(side, s.dc_done, 0));
    HIPCHK(c, hipEventRecord(s.pre0, side));
    int r = prebpf_launch(c, s.chunk, (long long)s.n_read, s.max_fields, s.craw, s.dc, side);
    if (r) return r;
    HIPCHK(c, hipEventRecord(s.pre1, side));
    return 0;
}

// K2 pulses and K2b's per-field work (the vblank state machine, mean line
// length, vblank pulses and first-HSYNC anchors: everything that does not
// depend on the field before) on the side stream, once the chunk's head is
// done.  They need only the head's outputs, and their few wavefronts
// (latency-bound, one per field) run beside the previous chunk's tail
// instead of in it; the tail waits for k2_done.
int pulse_stage(cuvhs_ctx* c, Slot& s) {
    const Fmt& f = c->fmt;
    hipStream_t side = c->serial ? c->stream : c->side;
    HIPCHK(c, hipStreamWaitEvent(side, s.head_done, 0));
    mark(c, s, 15, side);
    HIPCHK(c, run_sync_pulses(s.events, s.ctl, s.offsets, f.spfp, s.k2s, s.k2l, s.k2c, s.max_fields, 0, side));
    mark(c, s, 16, side);
    if (c->cfg.line_locator != CUVHS_LINELOC_K3)
        HIPCHK(c, run_k2b_fields(s.ctl, 0, s.max_fields, f, s.k2s, s.k2l, s.k2c, s.vps, s.vpl, s.vptg,
                                 (K2bField*)s.summ, side));
    mark(c, s, 17, side);
    HIPCHK(c, hipEventRecord(s.k2_done, side));
    return 0;
}

// Head stage on the head stream: K1, sync events, VSYNC clustering; with
// pre, the chunk's pre-band-pass goes in first on the side stream; then the
// chunk's pulse stage on the side stream.
int head_stage(cuvhs_ctx* c, Slot& s, bool pre = true) {
    const Fmt& f = c->fmt;
    hipStream_t st = c->stream;
    if (c->cfg.dc_correct) {
        // per-read mean of the chunk's samples, needed by both K1 and the
        // pre-band-pass (raw_reader.cpp:93-100)
        HIPCHK(c, run_input_mean(s.chunk, c->cfg.input_format, (long long)s.n_read, s.dc_partial, s.dc, st));
        HIPCHK(c, hipEventRecord(s.dc_done, st));
    }
    ChunkCtl* h = s.hctl;
    mark(c, s, 0, st);
    int r = k1_launch(c, s, s.chunk, (long long)s.n_read, s.fk1, st, s.ctl, c->thr_bits);
    if (r) return r;
    const long long T = (long long)s.fk1 * f.spfp;
    HIPCHK(c, run_events_bits(c->thr_bits, T, s.wg_count, s.wg_off, s.ev_masks, s.events, s.ctl, st));
    mark(c, s, 4, st);
    // the kernel also writes the control block to the pinned host copy
    HIPCHK(c, run_vsync_cluster(s.events, s.ctl, T, f.vsync_min_w, f.vsync_max_w, s.fk1 * 20, c->cand, s.offsets,
                                s.max_fields, f.samples_per_field, f.samples_per_line, s.fk1, 1, h, st));
    mark(c, s, 5, st);
    HIPCHK(c, hipEventRecord(s.head_done, st));
    // the pre-band-pass goes in after the head's launches: the head chain is
    // what the next chunk waits for
    if (pre) {
        int r = pre_stage(c, s);
        if (r) return r;
    }
    return pulse_stage(c, s);
}

struct OutArrays {
    uint16_t* luma;
    uint16_t* chroma;
    uint64_t* file_loc;
    int32_t *is_first, *phase_id, *do_count, *do_line, *do_startx, *do_endx;
};

// Tail stage on the tail stream.  out (optional): caller arrays that receive
// the chunk's outputs in place of the context's, with room for the field
// count the head stage produced.
// fork: the stream that takes K5 and the dropout classes beside chroma (the
// tail stream itself in serial profiling and diagnostics).
int tail_stage(cuvhs_ctx* c, Slot& s, const OutArrays* out, hipStream_t fork_stream) {
    const Fmt& f = c->fmt;
    hipStream_t st = c->serial ? c->stream : c->tail;
    HIPCHK(c, hipStreamWaitEvent(st, s.head_done, 0));
    mark(c, s, 6, st);
    // the chunk's K2 pulses and K2b per-field results (pulse_stage)
    HIPCHK(c, hipStreamWaitEvent(st, s.k2_done, 0));
    mark(c, s, 7, st);
    if (c->cfg.line_locator == CUVHS_LINELOC_K3) {
        // line_locs on the K2 pulses (pipeline.cu:1848-1856); K2b's cadence
        // state is left untouched, as in the reference
        HIPCHK(c, run_k3(s.ctl, 0, s.max_fields, f, s.k2s, s.k2l, s.k2c, c->pt, c->ll, c->isf_k2b, st));
    } else {
        HIPCHK(c, run_k2b_cadence(s.ctl, s.ctl, c->state, 0, s.max_fields, f, s.offsets, s.vps, s.vpl, s.vptg,
                                  (const K2bField*)s.summ, (K2bAnchorOut*)c->anchors, c->ps, c->pl, c->pc, c->pt,
                                  c->ll, c->isf_k2b, st));
    }
    if (c->diag) {
        HIPCHK(c, run_diag_spacing(c->ll, s.ctl, s.max_fields, f, st));
        HIPCHK(c, hipMemcpyAsync(c->ll_coarse, c->ll, (size_t)s.max_fields * f.lines_per_frame * sizeof(double),
                                 hipMemcpyDeviceToDevice, st));
    }
    mark(c, s, 8, st);
    HIPCHK(c, run_hsync_refine(s.d05, c->d05_f32, c->ll, s.ctl, 0, s.max_fields, f, st,
                               c->cfg.precision != CUVHS_PRECISION_FP64));
    if (c->diag) HIPCHK(c, run_diag_k4(s.d05, c->ll_coarse, c->ll, s.ctl, s.max_fields, f, st));
    mark(c, s, 9, st);
    // outputs go straight to the caller's arrays where given (the field count
    // n was checked against their capacity), else to the context's
    uint16_t* luma = out && out->luma ? out->luma : c->luma;
    uint16_t* chroma = out && out->chroma ? out->chroma : c->chroma;
    int* phase = out && out->phase_id ? out->phase_id : c->phase;
    int* doc = out && out->do_count ? out->do_count : c->doc;
    int* dol = out && out->do_line ? out->do_line : c->dol;
    int* dos = out && out->do_startx ? out->do_startx : c->dos;
    int* doe = out && out->do_endx ? out->do_endx : c->doe;
    unsigned long long* floc = out && out->file_loc ? reinterpret_cast<unsigned long long*>(out->file_loc) : c->file_loc;
    int* isf = out && out->is_first ? out->is_first : c->isf_out;
    // K5 and the dropout classes need only the line locations: they run on
    // the fork stream beside chroma, and the dropout scan, which conceals
    // into both outputs, joins them
    const bool fork = !c->serial && !c->diag && fork_stream != st;
    hipStream_t ax = fork ? fork_stream : st;
    if (fork) {
        HIPCHK(c, hipEventRecord(s.k4_done, st));
        HIPCHK(c, hipStreamWaitEvent(ax, s.k4_done, 0));
    }
    HIPCHK(c, run_tbc_resample(s.demod, c->sig_f32, c->ll, luma, s.ctl, 0, s.max_fields, f, ax));
    if (c->diag) HIPCHK(c, run_diag_k5(s.demod, c->ll, s.ctl, s.max_fields, f, st));
    mark(c, s, 10, st);
    if (fork) {
        if (!c->classes_on_tail)
            HIPCHK(c, run_dropout_classes(s.env, c->sig_f32, c->ll, s.ctl, 0, s.max_fields, f, c->do_partial,
                                          c->do_cls, ax));
        HIPCHK(c, hipEventRecord(s.aux_done, ax));
    }
    HIPCHK(c, hipStreamWaitEvent(st, s.pre1, 0));
    mark(c, s, 11, st);
    ChromaLaunch cl = chroma_launch(c, s.ctl, 0, s.craw);
    cl.out = chroma;
    HIPCHK(c, run_chroma(cl, c->state, s.max_fields, c->m0, c->m1, c->cs_tab, c->iq0, c->track, phase, 1, st));
    mark(c, s, 12, st);
    if (!fork || c->classes_on_tail)
        HIPCHK(c, run_dropout_classes(s.env, c->sig_f32, c->ll, s.ctl, 0, s.max_fields, f, c->do_partial, c->do_cls,
                                      st));
    if (fork) HIPCHK(c, hipStreamWaitEvent(st, s.aux_done, 0));
    HIPCHK(c, run_dropout_scan(c->do_cls, c->ll, s.ctl, 0, s.max_fields, f, luma, chroma, dol, dos, doe, doc, st));
    mark(c, s, 13, st);
    HIPCHK(c, run_field_meta(s.ctl, c->state, s.offsets, c->isf_k2b, floc, isf, st));
    mark(c, s, 14, st);
    HIPCHK(c, hipMemcpyAsync(s.hctl, s.ctl, sizeof(ChunkCtl), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipEventRecord(s.tail_done, st));
    return 0;
}

// after tail_done of slot s: accumulate its stage times
void collect_profile(cuvhs_ctx* c, Slot& s) {
    if (!c->prof) return;
    c->work_samples += (uint64_t)s.fk1 * (uint64_t)c->fmt.spfp;
    c->work_chunks++;
    c->work_fields += s.hctl->nf;
    auto span = [&](int a, int b) {
        float ms = 0.0f;
        return hipEventElapsedTime(&ms, s.mark[a], s.mark[b]) == hipSuccess ? (double)ms : 0.0;
    };
    for (int i = 0; i < kStages; i++) {
        // marks 0-5 head, 6-14 tail, 15-17 the pulse stage: k2_pulses is
        // 15-16, k2b_linelocs the per-field part 16-17 plus the tail's
        // cadence and line locations 7-8 (6-7 is the tail's wait for them)
        if (i == 5) c->stage_ms[i] += span(15, 16);
        else if (i == 6) c->stage_ms[i] += span(16, 17) + span(7, 8);
        else {
            const int a = i < 5 ? i : i + 1;
            c->stage_ms[i] += span(a, a + 1);
        }
    }
    float pre = 0.0f;
    if (hipEventElapsedTime(&pre, s.pre0, s.pre1) == hipSuccess) c->stage_ms[kStages] += pre;
}

int sync_all(cuvhs_ctx* c) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipStreamSynchronize(c->tail));
    HIPCHK(c, hipStreamSynchronize(c->side));
    HIPCHK(c, hipStreamSynchronize(c->aux));
    return 0;
}

// Host wait for an event by polling: the decode loop's waits gate the next
// launches, and a blocking wait's wake-up latency would leave the GPU idle
// between the head of one chunk and the tail launched after it.  The poll
// spins for the first 50 us, then yields the core between polls, and after
// 2 ms sleeps 50 us between polls, so a long wait does not hold a host core.
hipError_t spin_wait(hipEvent_t ev) {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    hipError_t e;
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
        const auto dt = clk::now() - t0;
        if (dt < std::chrono::microseconds(50)) continue;
        if (dt < std::chrono::milliseconds(2)) sched_yield();
        else std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    return e;
}

int finish_tail(cuvhs_ctx* c, Slot& s) {
    HIPCHK(c, spin_wait(s.tail_done));
    collect_profile(c, s);
    if (c->diag) {
        for (int i = 0; i < 9; i++) {
            c->diag_last[i] = s.hctl->diag[i];
            c->diag_total[i] += s.hctl->diag[i];
        }
    }
    if (s.hctl->vp_overflow) {
        c->err = "K2b validated-pulse capacity exceeded";
        return CUVHS_E_CAPACITY;
    }
    return 0;
}

// ---- graph mode of cuvhs_decode_capture ----
//
// Each stage's kernel sequence is captured once per slot, on the one stream
// it runs on and between the cross-stream waits (which stay outside the
// graphs, as in the launch-by-launch schedule), and replayed for every chunk
// the slot takes.  Every varying value comes from the slot's device plan: the
// host fills it when it plans the chunk and a one-thread kernel writes it to
// the device (as its argument) ahead of the chunk's head, after the tail of
// the chunk that last used the slot has finished.  The grids are sized for
// the batch, and the kernels return past the chunk's fields.
//
// The graphs are captured by cuvhs_create, one capture at a time in the
// process (g_capture_mutex) and in relaxed mode: HIP's stream capture is not
// isolated between threads (measured: a capture in one thread made another
// thread's hipStreamSynchronize and synchronous hipMemcpy fail with "not
// permitted when stream is capturing"), so no capture may overlap a decode.
std::mutex g_capture_mutex;

// The kernel sequence of each stage graph, as launched on st.
int seq_mean(cuvhs_ctx* c, Slot& s, hipStream_t st) {
    const long long S = (long long)c->F * c->fmt.spfp;
    HIPCHK(c, run_input_mean(nullptr, c->cfg.input_format, S, s.dc_partial, s.dc, st, s.plan));
    return 0;
}
int seq_head(cuvhs_ctx* c, Slot& s, hipStream_t st) {
    const Fmt& f = c->fmt;
    const int F = c->F;
    const long long S = (long long)F * f.spfp;   // the batch: the grids' size
    const int r = k1_launch(c, s, nullptr, S, F, st, s.ctl, c->thr_bits, s.plan);
    if (r) return r;
    HIPCHK(c, run_events_bits(c->thr_bits, S, s.wg_count, s.wg_off, s.ev_masks, s.events, s.ctl, st, s.plan));
    HIPCHK(c, run_vsync_cluster(s.events, s.ctl, S, f.vsync_min_w, f.vsync_max_w, F * 20, c->cand, s.offsets, F,
                                f.samples_per_field, f.samples_per_line, F, 1, nullptr, st, s.plan));
    return 0;
}
int seq_pre(cuvhs_ctx* c, Slot& s, hipStream_t st) {
    return prebpf_launch(c, nullptr, (long long)c->F * c->fmt.spfp, c->F, s.craw, s.dc, st, s.plan);
}
int seq_pulse(cuvhs_ctx* c, Slot& s, hipStream_t st) {
    const Fmt& f = c->fmt;
    HIPCHK(c, run_sync_pulses(s.events, s.ctl, s.offsets, f.spfp, s.k2s, s.k2l, s.k2c, c->F, 0, st));
    if (c->cfg.line_locator != CUVHS_LINELOC_K3)
        HIPCHK(c, run_k2b_fields(s.ctl, 0, c->F, f, s.k2s, s.k2l, s.k2c, s.vps, s.vpl, s.vptg, (K2bField*)s.summ, st));
    return 0;
}
int seq_tail1(cuvhs_ctx* c, Slot& s, hipStream_t st) {
    const Fmt& f = c->fmt;
    if (c->cfg.line_locator == CUVHS_LINELOC_K3) {
        HIPCHK(c, run_k3(s.ctl, 0, c->F, f, s.k2s, s.k2l, s.k2c, c->pt, c->ll, c->isf_k2b, st));
    } else {
        HIPCHK(c, run_k2b_cadence(s.ctl, s.ctl, c->state, 0, c->F, f, s.offsets, s.vps, s.vpl, s.vptg,
                                  (const K2bField*)s.summ, (K2bAnchorOut*)c->anchors, c->ps, c->pl, c->pc, c->pt,
                                  c->ll, c->isf_k2b, st));
    }
    HIPCHK(c, run_hsync_refine(s.d05, c->d05_f32, c->ll, s.ctl, 0, c->F, f, st,
                               c->cfg.precision != CUVHS_PRECISION_FP64));
    return 0;
}
int seq_fork(cuvhs_ctx* c, Slot& s, hipStream_t st) {
    HIPCHK(c, run_tbc_resample(s.demod, c->sig_f32, c->ll, nullptr, s.ctl, 0, c->F, c->fmt, st, s.plan));
    if (!c->classes_on_tail)
        HIPCHK(c, run_dropout_classes(s.env, c->sig_f32, c->ll, s.ctl, 0, c->F, c->fmt, c->do_partial, c->do_cls,
                                      st));
    return 0;
}
int seq_tail2(cuvhs_ctx* c, Slot& s, hipStream_t st) {
    ChromaLaunch cl = chroma_launch(c, s.ctl, 0, s.craw);
    cl.plan = s.plan;
    HIPCHK(c, run_chroma(cl, c->state, c->F, c->m0, c->m1, c->cs_tab, c->iq0, c->track, c->phase, 1, st));
    return 0;
}
int seq_tail3(cuvhs_ctx* c, Slot& s, hipStream_t st) {
    if (c->classes_on_tail)
        HIPCHK(c, run_dropout_classes(s.env, c->sig_f32, c->ll, s.ctl, 0, c->F, c->fmt, c->do_partial, c->do_cls,
                                      st));
    HIPCHK(c, run_dropout_scan(c->do_cls, c->ll, s.ctl, 0, c->F, c->fmt, c->luma, c->chroma, c->dol, c->dos, c->doe,
                               c->doc, st, s.plan));
    HIPCHK(c, run_field_meta(s.ctl, c->state, s.offsets, c->isf_k2b, c->file_loc, c->isf_out, st, s.plan));
    HIPCHK(c, hipMemcpyAsync(s.hctl, s.ctl, sizeof(ChunkCtl), hipMemcpyDeviceToHost, st));
    return 0;
}

// The fork stream of cuvhs_decode_capture's tails: the aux stream when the
// process has a hardware queue for it, else the head stream.
hipStream_t fork_stream(const cuvhs_ctx* c) { return c->aux_queue ? c->aux : c->stream; }

// Capture seq's launches on st into *g.
int capture_graph(cuvhs_ctx* c, Slot& s, hipGraphExec_t* g, hipStream_t st,
                  int (*seq)(cuvhs_ctx*, Slot&, hipStream_t)) {
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return fail(c, e, "hipStreamBeginCapture");
    const int r = seq(c, s, st);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(st, &graph);
    if (r != 0) {
        if (graph) (void)hipGraphDestroy(graph);
        return r;
    }
    if (e != hipSuccess) return fail(c, e, "hipStreamEndCapture");
    e = hipGraphInstantiate(g, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        *g = nullptr;
        return fail(c, e, "hipGraphInstantiate");
    }
    return 0;
}

// Every slot's stage graphs (cuvhs_create, graph mode).
int capture_graphs(cuvhs_ctx* c) {
    std::lock_guard<std::mutex> lock(g_capture_mutex);
    for (Slot& s : c->slot) {
        int r = 0;
        if (c->cfg.dc_correct && (r = capture_graph(c, s, &s.graph[kGMean], c->stream, seq_mean)) != 0) return r;
        if ((r = capture_graph(c, s, &s.graph[kGHead], c->stream, seq_head)) != 0) return r;
        if ((r = capture_graph(c, s, &s.graph[kGPre], c->side, seq_pre)) != 0) return r;
        if ((r = capture_graph(c, s, &s.graph[kGPulse], c->side, seq_pulse)) != 0) return r;
        if ((r = capture_graph(c, s, &s.graph[kGTail1], c->tail, seq_tail1)) != 0) return r;
        if ((r = capture_graph(c, s, &s.graph[kGFork], fork_stream(c), seq_fork)) != 0) return r;
        if ((r = capture_graph(c, s, &s.graph[kGTail2], c->tail, seq_tail2)) != 0) return r;
        if ((r = capture_graph(c, s, &s.graph[kGTail3], c->tail, seq_tail3)) != 0) return r;
    }
    return 0;
}

// The outputs of cuvhs_decode_capture: the caller's arrays, where given.
struct OutBases {
    uint16_t *luma, *chroma;
    uint64_t* file_loc;
    int32_t *is_first, *phase_id, *do_count, *do_line, *do_startx, *do_endx;
};

// The planned chunk of slot s into its host plan: its read, K1's extent
// (and what an earlier, longer chunk left past it), its VSYNC candidate cap,
// and where its fields go (out_fields fields precede it in the outputs).
void fill_plan(cuvhs_ctx* c, Slot& s, const OutBases& ob, long long out_fields) {
    ChunkPlan& p = s.hplan;
    const long long T = (long long)s.fk1 * c->fmt.spfp;
    const size_t tbc = (size_t)c->fmt.field_lines * c->fmt.line_len;
    p.chunk = s.chunk;
    p.n_read = (long long)s.n_read;
    p.total = T;
    p.clear_end = std::max(s.k1_extent, T);
    s.k1_extent = T;
    p.raw_offset = s.raw_offset;
    p.fields_k1 = s.fk1;
    p.max_fields = s.max_fields;
    p.vsync_cap = s.fk1 * 20;
    p.luma = ob.luma ? ob.luma + (size_t)out_fields * tbc : c->luma;
    p.chroma = ob.chroma ? ob.chroma + (size_t)out_fields * tbc : c->chroma;
    p.file_loc = ob.file_loc ? reinterpret_cast<unsigned long long*>(ob.file_loc) + out_fields : c->file_loc;
    p.is_first = ob.is_first ? ob.is_first + out_fields : c->isf_out;
    p.phase_id = ob.phase_id ? ob.phase_id + out_fields : c->phase;
    p.do_count = ob.do_count ? ob.do_count + out_fields : c->doc;
    p.do_line = ob.do_line ? ob.do_line + (size_t)out_fields * kMaxDropouts : c->dol;
    p.do_startx = ob.do_startx ? ob.do_startx + (size_t)out_fields * kMaxDropouts : c->dos;
    p.do_endx = ob.do_endx ? ob.do_endx + (size_t)out_fields * kMaxDropouts : c->doe;
}

// Head stage in graph mode: the plan to the device, the per-read mean, the
// head (K1, guard band, events, VSYNC clustering) on the head stream, then
// the pre-band-pass and the pulse stage on the side stream.
int head_stage_graph(cuvhs_ctx* c, Slot& s) {
    hipStream_t st = c->stream, side = c->side;
    HIPCHK(c, run_put_plan(s.plan, s.hplan, st));
    if (c->cfg.dc_correct) HIPCHK(c, hipGraphLaunch(s.graph[kGMean], st));
    HIPCHK(c, hipEventRecord(s.dc_done, st));   // the plan (and the mean) are on the device
    HIPCHK(c, hipGraphLaunch(s.graph[kGHead], st));
    // the control block (field count, next chunk offset) to the host by a
    // copy: a graph's kernel nodes need not make their stores to host memory
    // visible when they end, as a kernel launched on its own does
    HIPCHK(c, hipMemcpyAsync(s.hctl, s.ctl, sizeof(ChunkCtl), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipEventRecord(s.head_done, st));
    HIPCHK(c, hipStreamWaitEvent(side, s.dc_done, 0));
    HIPCHK(c, hipEventRecord(s.pre0, side));
    HIPCHK(c, hipGraphLaunch(s.graph[kGPre], side));
    HIPCHK(c, hipEventRecord(s.pre1, side));
    HIPCHK(c, hipStreamWaitEvent(side, s.head_done, 0));
    HIPCHK(c, hipGraphLaunch(s.graph[kGPulse], side));
    HIPCHK(c, hipEventRecord(s.k2_done, side));
    return 0;
}

// Tail stage in graph mode, with the launch-by-launch schedule's waits:
// line locations and K4; K5 (with the dropout classes outside the mixed
// forms) on the fork stream; chroma once the pre-band-pass is done; the
// dropout scan (joining the fork), metadata and the control block's copy to
// the host.
int tail_stage_graph(cuvhs_ctx* c, Slot& s) {
    hipStream_t st = c->tail, fork = fork_stream(c);
    HIPCHK(c, hipStreamWaitEvent(st, s.head_done, 0));
    HIPCHK(c, hipStreamWaitEvent(st, s.k2_done, 0));
    HIPCHK(c, hipGraphLaunch(s.graph[kGTail1], st));
    HIPCHK(c, hipEventRecord(s.k4_done, st));
    HIPCHK(c, hipStreamWaitEvent(fork, s.k4_done, 0));
    HIPCHK(c, hipGraphLaunch(s.graph[kGFork], fork));
    HIPCHK(c, hipEventRecord(s.aux_done, fork));
    HIPCHK(c, hipStreamWaitEvent(st, s.pre1, 0));
    HIPCHK(c, hipGraphLaunch(s.graph[kGTail2], st));
    HIPCHK(c, hipStreamWaitEvent(st, s.aux_done, 0));
    HIPCHK(c, hipGraphLaunch(s.graph[kGTail3], st));
    HIPCHK(c, hipEventRecord(s.tail_done, st));
    return 0;
}

// ---- scratch of the allocation-free dropout entry point ----
//
// cuvhs_dropouts_noalloc takes its device scratch (the per-slice envelope
// sums and the 16-sample block classes of K7, the buffers a context
// allocates for cuvhs_dropouts) from this file-scope array instead of
// hipMalloc: kNoAllocMaxFields fields, 7.0 MB, held on every device the
// library runs on until the process ends.  Host code never uses the array's
// name as an address: the device address is taken once per device with
// hipGetSymbolAddress (with that device current) and kept in
// g_noalloc_scratch.  Every call on a device shares its array, so calls that
// overlap on one device race; the kernels write every scratch element they
// read in the same call, so nothing needs clearing between calls.
constexpr int kNoAllocMaxFields = 128;
constexpr size_t kNoAllocBytes = 6979584;  // 128 fields x (53,504 class bytes + 128 partial sums x 8 B)
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
std::mutex g_noalloc_mu;
void* g_noalloc_scratch[kNoAllocDevices] = {};
std::string g_free_err;  // last error of a context-free call (cuvhs_last_error(NULL))

int noalloc_fail(int rc, const std::string& what) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    g_free_err = what;
    return rc;
}

// The scratch array's device address on the current device.
int noalloc_scratch(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return noalloc_fail(CUVHS_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kNoAllocDevices) return noalloc_fail(CUVHS_E_UNSUPPORTED, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_free_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return CUVHS_E_HIP;
        }
        g_noalloc_scratch[dev] = p;
    }
    *out = g_noalloc_scratch[dev];
    return 0;
}

}  // namespace

template <typename V, typename T>
static int block_selftest(int B, int is_f32, const void* d_x, const void* d_g, void* d_y) {
    std::vector<V> g((size_t)B), gp((size_t)B);
    if (hipMemcpy(g.data(), d_g, (size_t)B * sizeof(V), hipMemcpyDeviceToHost) != hipSuccess) return CUVHS_E_HIP;
    for (size_t k = 0; k < (size_t)B; k++) gp[B == 1024 ? k : block_bin((size_t)B, k)] = g[k];
    auto t = twiddles<V, T>(B);
    void* tw = nullptr;
    void* dg = nullptr;
    if (hipMalloc(&tw, (size_t)B * sizeof(V)) != hipSuccess) return CUVHS_E_ALLOC;
    if (hipMalloc(&dg, (size_t)B * sizeof(V)) != hipSuccess) { (void)hipFree(tw); return CUVHS_E_ALLOC; }
    hipError_t e = hipMemcpy(tw, t.data(), (size_t)B * sizeof(V), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dg, gp.data(), (size_t)B * sizeof(V), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = run_block_selftest(B, is_f32, d_x, dg, d_y, tw, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(dg);
    (void)hipFree(tw);
    return e == hipSuccess ? 0 : (e == hipErrorInvalidValue ? CUVHS_E_ARG : CUVHS_E_HIP);
}

extern "C" {

int cuvhs_format_for_system(int system, double mhz, cuvhs_format* o) {
    if (!o || !(mhz > 0.0) || (system != CUVHS_SYSTEM_NTSC && system != CUVHS_SYSTEM_PAL)) return CUVHS_E_ARG;
    Fmt f = make_fmt(mhz, system == CUVHS_SYSTEM_PAL ? kSystemPal : kSystemNtsc);
    o->sample_rate = f.sample_rate;
    o->samples_per_line = f.samples_per_line;
    o->samples_per_field = f.samples_per_field;
    o->spfp = f.spfp;
    o->pulse_threshold_hz = f.pulse_threshold_hz;
    o->ire0 = f.ire0;
    o->hz_ire = f.hz_ire;
    o->output_scale = f.output_scale;
    o->output_zero = f.output_zero;
    o->output_rate = f.output_rate;
    o->black16b = (0.0 - f.vsync_ire) * f.output_scale + f.output_zero;
    o->white16b = (100.0 - f.vsync_ire) * f.output_scale + f.output_zero;
    o->burst_start = f.burst_start;
    o->burst_end = f.burst_end;
    o->system = system;
    o->lines_per_frame = f.lines_per_frame;
    o->field_lines = f.field_lines;
    o->line_len = f.line_len;
    o->fsc = f.fsc;
    return 0;
}

int cuvhs_format_for(double mhz, cuvhs_format* o) { return cuvhs_format_for_system(CUVHS_SYSTEM_NTSC, mhz, o); }

int cuvhs_create(const cuvhs_config* cfg, cuvhs_ctx** out) {
    if (!cfg || !out) return CUVHS_E_ARG;
    *out = nullptr;
    if ((cfg->system != CUVHS_SYSTEM_NTSC && cfg->system != CUVHS_SYSTEM_PAL) || cfg->precision < CUVHS_PRECISION_FP64 ||
        cfg->precision > CUVHS_PRECISION_MIXED_BF16)
        return CUVHS_E_UNSUPPORTED;
    // the matrix-core part A takes u8 samples, exact in f16 / bf16 after
    // removing 128 (a per-read DC or 16-bit samples would need splitting too)
    if (cfg->precision >= CUVHS_PRECISION_MIXED_F16 && (cfg->input_format != CUVHS_INPUT_U8 || cfg->dc_correct))
        return CUVHS_E_UNSUPPORTED;
    if (!(cfg->sample_rate_mhz >= 16.0 && cfg->sample_rate_mhz <= 40.0)) return CUVHS_E_UNSUPPORTED;
    if (cfg->input_format < CUVHS_INPUT_U8 || cfg->input_format > CUVHS_INPUT_U16) return CUVHS_E_UNSUPPORTED;
    if (cfg->line_locator != CUVHS_LINELOC_K2B && cfg->line_locator != CUVHS_LINELOC_K3) return CUVHS_E_UNSUPPORTED;
    if (cfg->batch_fields < 1 || cfg->batch_fields > 4096) return CUVHS_E_ARG;
    if ((cfg->block_k1a && cfg->block_k1a != 4096 && cfg->block_k1a != 8192) ||
        (cfg->block_k1b && cfg->block_k1b != 4096 && cfg->block_k1b != 8192) ||
        (cfg->block_pre && cfg->block_pre != 8192 && cfg->block_pre != 16384))
        return CUVHS_E_ARG;
    cuvhs_ctx* c = new cuvhs_ctx();
    c->cfg = *cfg;
    c->graphs = cfg->graphs != 0;
    c->fmt = make_fmt(cfg->sample_rate_mhz, cfg->system == CUVHS_SYSTEM_PAL ? kSystemPal : kSystemNtsc);
    c->F = cfg->batch_fields;
    const Fmt& f = c->fmt;
    if ((long long)c->F * f.spfp > 0x3fffffffLL) { delete c; return CUVHS_E_CAPACITY; }
    auto bad = [&](hipError_t e, const char* w) { fail(c, e, w); cuvhs_destroy(c); return CUVHS_E_HIP; };
    if (cfg->device < 0) {
        // GPUDevice::init(-1) (device.cu:13-24): the device with the most memory
        int n = 0;
        hipError_t ce = hipGetDeviceCount(&n);
        if (ce != hipSuccess || n < 1) return bad(ce != hipSuccess ? ce : hipErrorNoDevice, "hipGetDeviceCount");
        size_t best = 0;
        for (int i = 0; i < n; i++) {
            size_t mem = 0;
            if (hipDeviceTotalMem(&mem, i) == hipSuccess && mem > best) {
                best = mem;
                c->cfg.device = i;
            }
        }
    }
    hipError_t e = hipSetDevice(c->cfg.device);
    if (e != hipSuccess) return bad(e, "hipSetDevice");
    if (cfg->stream) {
        c->stream = (hipStream_t)cfg->stream;
    } else {
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) return bad(e, "hipStreamCreate");
        c->own_stream = true;
    }
    e = hipStreamCreateWithFlags(&c->tail, hipStreamNonBlocking);
    if (e != hipSuccess) return bad(e, "hipStreamCreate tail");
    e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e != hipSuccess) return bad(e, "hipStreamCreate side");
    e = hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking);
    if (e != hipSuccess) return bad(e, "hipStreamCreate aux");
    c->aux_queue = hw_queues() >= 5;
    for (Slot& s : c->slot) {
        for (hipEvent_t* ev : {&s.head_done, &s.tail_done, &s.dc_done, &s.k4_done, &s.aux_done, &s.k2_done}) {
            e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
            if (e != hipSuccess) return bad(e, "hipEventCreate");
        }
        for (hipEvent_t* ev : {&s.pre0, &s.pre1}) {
            e = hipEventCreate(ev);
            if (e != hipSuccess) return bad(e, "hipEventCreate");
        }
    }
    try {
        // taps below 1e-13 (FP64) or 1e-8 (FP32) of the largest are dropped:
        // both far below the rounding of the transform they feed
        const int bk1[2] = {4096, 8192};
        c->mixed = cfg->precision >= CUVHS_PRECISION_MIXED_F16;
        c->mixed_bf16 = cfg->precision == CUVHS_PRECISION_MIXED_BF16;
        c->classes_on_tail = c->mixed;
        c->k1_f32 = cfg->precision != CUVHS_PRECISION_FP64;
        c->sig_f32 = cfg->precision != CUVHS_PRECISION_FP64;
        c->d05_f32 = cfg->precision == CUVHS_PRECISION_FP32;
        // block sizes: the caller's, else the autotuned table's, else the
        // cost model's choice among the supported ones
        const TunedBlocks* tb = tuned_blocks(cfg->sample_rate_mhz, cfg->precision);
        const int ua = cfg->block_k1a ? cfg->block_k1a : (tb ? tb->k1a : 0);
        const int ub = cfg->block_k1b ? cfg->block_k1b : (tb ? tb->k1b : 0);
        const int up = cfg->block_pre ? cfg->block_pre : (tb ? tb->pre : 0);
        c->block_source = (cfg->block_k1a || cfg->block_k1b || cfg->block_pre) ? 0 : (tb ? 1 : 2);
        const int* cha = ua ? &ua : bk1;
        const int* chb = ub ? &ub : bk1;
        const int na = ua ? 1 : 2, nb = ub ? 1 : 2;
        c->k1 = c->k1_f32 ? design_k1(f, 1e-8, cha, na, chb, nb, 1, 8) : design_k1(f, 1e-13, cha, na, chb, nb, 1, 16);
        if (c->mixed) {
            c->k1g = design_k1(f, 1e-13, bk1, 2, bk1, 2, 1, 16);
            // sync-threshold band of the FP32 part B and the envelope ratio
            // below which a tile's phase is treated as ill-conditioned
            c->g_delta = c->mixed_bf16 ? kGuardDeltaBf16 : kGuardDeltaF16;
            c->g_rho = c->mixed_bf16 ? kGuardRhoBf16 : kGuardRhoF16;
        }
        const int bpre[2] = {8192, 16384};
        // the pre-band-pass: FP64 with taps to 1e-9 of the largest (its
        // truncation is then ~1e-6 of the chroma band's level even when the
        // band holds only quantisation noise), or FP32 with taps to 1e-7
        c->pre_f32 = cfg->precision != CUVHS_PRECISION_FP64;
        c->chroma_f32 = cfg->precision != CUVHS_PRECISION_FP64;
        const int* chp = up ? &up : bpre;
        const int np = up ? 1 : 2;
        c->pre = c->pre_f32 ? design_prebpf(f, 1e-7, chp, np) : design_prebpf(f, 1e-9, chp, np);
    } catch (...) {
        // no supported block size fits a filter: a size the caller forced,
        // or a sample rate whose filters are too long
        const bool forced = cfg->block_k1a || cfg->block_k1b || cfg->block_pre;
        cuvhs_destroy(c);
        return forced ? CUVHS_E_ARG : CUVHS_E_UNSUPPORTED;
    }
#define ALLOC(p, n)                                    \
    do {                                               \
        hipError_t e2 = dalloc(c, &(p), (size_t)(n));  \
        if (e2 != hipSuccess) return bad(e2, #p);      \
    } while (0)
    {
        const bool pa = true, pb = true;  // every K1 block size is a wave-local block
        if (c->k1_f32) {
            if ((e = upload_spectrum<float2, float>(c, (float2**)&c->G_a, c->k1.rf.spectrum, pa)) != hipSuccess) return bad(e, "G_a");
            if ((e = upload_spectrum<float2, float>(c, (float2**)&c->G_b, c->k1.video.spectrum, pb)) != hipSuccess) return bad(e, "G_b");
        } else {
            if ((e = upload_spectrum<double2, double>(c, (double2**)&c->G_a, c->k1.rf.spectrum, pa)) != hipSuccess) return bad(e, "G_a");
            if ((e = upload_spectrum<double2, double>(c, (double2**)&c->G_b, c->k1.video.spectrum, pb)) != hipSuccess) return bad(e, "G_b");
        }
    }
    if (c->mixed) {
        if ((e = upload_spectrum<double2, double>(c, (double2**)&c->G_ga, c->k1g.rf.spectrum, true)) != hipSuccess)
            return bad(e, "G_ga");
        if ((e = upload_spectrum<double2, double>(c, (double2**)&c->G_gb, c->k1g.video.spectrum, true)) != hipSuccess)
            return bad(e, "G_gb");
        if ((e = upload_twiddles(c, &c->tw_ga, c->k1g.rf.B, 0)) != hipSuccess) return bad(e, "twiddles");
        if ((e = upload_twiddles(c, &c->tw_gb, c->k1g.video.B, 0)) != hipSuccess) return bad(e, "twiddles");
        std::vector<uint4> taps;
        c->m_nc = build_mfma_taps(c->k1.rf.taps, c->k1.rf.lm, c->k1.rf.lp, c->mixed_bf16, taps, &c->m_scale,
                                  &c->m_glo0, &c->m_glo1);
        c->m_scale *= 256.0;   // u8 sample minus 128, times 256 (raw_reader.cpp:71-76)
        c->m_nx = mfma_input_elems(c->m_nc);
        c->m_nblk = mfma_blocks_per_segment(f.spfp);
        // the window never wraps past the segment's FFT length for kept outputs
        if (c->k1.rf.lm > c->k1.N - f.spfp - kOverlap) return bad(hipErrorInvalidValue, "MFMA part A span");
        if ((e = dalloc(c, &c->m_taps, taps.size())) != hipSuccess) return bad(e, "taps");
        if ((e = hipMemcpy(c->m_taps, taps.data(), taps.size() * sizeof(uint4), hipMemcpyHostToDevice)) != hipSuccess)
            return bad(e, "taps");
    }
    if (c->pre_f32) e = upload_spectrum<float2, float>(c, (float2**)&c->G_p, c->pre.bpf.spectrum, true);
    else e = upload_spectrum<double2, double>(c, (double2**)&c->G_p, c->pre.bpf.spectrum, true);
    if (e != hipSuccess) return bad(e, "G_p");
    {
        if ((e = upload_twiddles(c, &c->tw_a, c->k1.rf.B, c->k1_f32)) != hipSuccess) return bad(e, "twiddles");
        if ((e = upload_twiddles(c, &c->tw_b, c->k1.video.B, c->k1_f32)) != hipSuccess) return bad(e, "twiddles");
        if ((e = upload_twiddles(c, &c->tw_p, c->pre.bpf.B, c->pre_f32)) != hipSuccess) return bad(e, "twiddles");
        auto t1 = twiddles<double2, double>(1024);
        ALLOC(c->tw_1024, t1.size());
        if ((e = hipMemcpy(c->tw_1024, t1.data(), t1.size() * sizeof(double2), hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "twiddles");
        if (c->chroma_f32) {
            auto t2 = twiddles<float2, float>(1024);
            ALLOC(c->tw_1024f, t2.size());
            if ((e = hipMemcpy(c->tw_1024f, t2.data(), t2.size() * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess)
                return bad(e, "twiddles");
        }
        if (f.chroma_fft == 2048) {
            auto t3 = twiddles<double2, double>(2048);
            ALLOC(c->tw_2048, t3.size());
            if ((e = hipMemcpy(c->tw_2048, t3.data(), t3.size() * sizeof(double2), hipMemcpyHostToDevice)) != hipSuccess)
                return bad(e, "twiddles");
            if (c->chroma_f32) {
                auto t4 = twiddles<float2, float>(2048);
                ALLOC(c->tw_2048f, t4.size());
                if ((e = hipMemcpy(c->tw_2048f, t4.data(), t4.size() * sizeof(float2), hipMemcpyHostToDevice)) !=
                    hipSuccess)
                    return bad(e, "twiddles");
            }
        }
        // burst I/Q reference (chroma_decode.cu:435-437), computed on the host
        const int blen = f.burst_end - f.burst_start;
        std::vector<double> cs((size_t)2 * blen);
        for (int i = 0; i < blen; i++) {
            double t = 2.0 * M_PI * f.fsc / f.output_rate * (double)(f.burst_start + i);
            cs[(size_t)2 * i] = cos(t);
            cs[(size_t)2 * i + 1] = sin(t);
        }
        ALLOC(c->cs_tab, cs.size());
        if ((e = hipMemcpy(c->cs_tab, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "burst table");
        // band-pass bins kept by the per-line chroma filter (chroma_decode.cu:163-182)
        int lo = 0, hi = 0;
        chroma_band(f, &lo, &hi);
        std::vector<double2> mk((size_t)f.chroma_fft);
        for (int k = 0; k < f.chroma_fft; k++) {
            int kk = k <= f.chroma_fft / 2 ? k : f.chroma_fft - k;
            mk[(size_t)k].x = (kk >= lo && kk <= hi) ? 1.0 : 0.0;
            mk[(size_t)k].y = 0.0;
        }
        ALLOC(c->chroma_mask, mk.size());
        if ((e = hipMemcpy(c->chroma_mask, mk.data(), mk.size() * sizeof(double2), hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "chroma mask");
    }
    const long long S = (long long)c->F * f.spfp;
    const size_t Fz = (size_t)c->F;
    c->buf_n = S + f.spfp;  // one segment of zero padding past the chunk
    ALLOC(c->rawdemod, S);
    if (c->mixed) {
        const int nbA = (f.spfp + c->k1g.rf.valid - 1) / c->k1g.rf.valid;
        const int nbB = (f.spfp + c->k1g.video.valid - 1) / c->k1g.video.valid;
        ALLOC(c->g_flag, ((S - 1) >> kGuardSubShift) + 1);   // zeroed here, then by each compaction
        ALLOC(c->g_env, Fz * c->m_nblk * 16);
        ALLOC(c->g_markA, Fz * nbA);
        ALLOC(c->g_markB, Fz * nbB);
        ALLOC(c->g_listA, Fz * nbA);
        ALLOC(c->g_listB, Fz * nbB);
        ALLOC(c->g_counts, 4);
        ALLOC(c->g_raw64, S);
    }
    ALLOC(c->cand, 20 * c->F);
    ALLOC(c->thr_bits, thr_bits_words(S));
    for (Slot& s : c->slot) {
        ALLOC(s.demod, c->buf_n);
        ALLOC(s.d05, c->buf_n);
        ALLOC(s.env, c->buf_n);
        ALLOC(s.events, S + 2);
        ALLOC(s.wg_count, events_nwg(S) + 1);
        ALLOC(s.wg_off, events_nwg(S) + 1);
        ALLOC(s.ev_masks, events_mask_count(S));
        ALLOC(s.offsets, Fz);
        ALLOC(s.craw, S);
        ALLOC(s.dc, 1);
        ALLOC(s.dc_partial, 256);
        ALLOC(s.ctl, 1);
        e = hipHostMalloc((void**)&s.hctl, sizeof(ChunkCtl), hipHostMallocDefault);
        if (e != hipSuccess) return bad(e, "hipHostMalloc");
        memset(s.hctl, 0, sizeof(ChunkCtl));
        ALLOC(s.plan, 1);
    }
    for (Slot& s : c->slot) {
        ALLOC(s.k2s, Fz * kMaxPulses);
        ALLOC(s.k2l, Fz * kMaxPulses);
        ALLOC(s.k2c, Fz);
        ALLOC(s.vps, Fz * kVpCap);
        ALLOC(s.vpl, Fz * kVpCap);
        ALLOC(s.vptg, Fz * kVpCap);
        unsigned char* p1 = nullptr;
        ALLOC(p1, Fz * k2b_field_summary_bytes());
        s.summ = p1;
    }
    ALLOC(c->ps, Fz * kMaxPulses);
    ALLOC(c->pl, Fz * kMaxPulses);
    ALLOC(c->pc, Fz);
    ALLOC(c->pt, Fz * kMaxPulses);
    {
        unsigned char* p2 = nullptr;
        ALLOC(p2, Fz * k2b_anchor_bytes());
        c->anchors = p2;
    }
    ALLOC(c->ll, Fz * f.lines_per_frame);
    ALLOC(c->ll_coarse, Fz * f.lines_per_frame);
    ALLOC(c->isf_k2b, Fz);
    ALLOC(c->luma, Fz * f.field_lines * f.line_len);
    ALLOC(c->chroma, Fz * f.field_lines * f.line_len);
    ALLOC(c->burst, Fz * f.field_lines * (f.burst_end - f.burst_start));
    ALLOC(c->chroma_filt, Fz * f.field_lines * f.line_len);
    ALLOC(c->m0, Fz);
    ALLOC(c->m1, Fz);
    ALLOC(c->iq0, 2 * f.field_lines);
    ALLOC(c->track, Fz);
    ALLOC(c->phase, Fz);
    ALLOC(c->dol, Fz * kMaxDropouts);
    ALLOC(c->dos, Fz * kMaxDropouts);
    ALLOC(c->doe, Fz * kMaxDropouts);
    ALLOC(c->doc, Fz);
    ALLOC(c->do_partial, dropout_partial_count(c->F));
    ALLOC(c->do_cls, dropout_class_bytes(c->F));
    ALLOC(c->file_loc, Fz);
    ALLOC(c->isf_out, Fz);
    ALLOC(c->state, 1);
#undef ALLOC
    e = hipHostMalloc((void**)&c->hstate, sizeof(DecodeState), hipHostMallocDefault);
    if (e != hipSuccess) return bad(e, "hipHostMalloc");
    reset_state(c->hstate);
    e = hipMemcpy(c->state, c->hstate, sizeof(DecodeState), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bad(e, "state upload");
    if (c->graphs) {
        const int r = capture_graphs(c);
        if (r != 0) {
            cuvhs_destroy(c);
            return r;
        }
    }
    *out = c;
    return 0;
}

int cuvhs_destroy(cuvhs_ctx* c) {
    if (!c) return CUVHS_E_ARG;
    for (hipStream_t st : {c->stream, c->tail, c->side, c->aux})
        if (st) (void)hipStreamSynchronize(st);
    for (Slot& s : c->slot) {
        for (hipEvent_t e : s.mark)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : {s.head_done, s.tail_done, s.pre0, s.pre1, s.dc_done, s.k4_done, s.aux_done, s.k2_done})
            if (e) (void)hipEventDestroy(e);
        if (s.hctl) (void)hipHostFree(s.hctl);
        for (hipGraphExec_t g : s.graph)
            if (g) (void)hipGraphExecDestroy(g);
    }
    if (c->aux) (void)hipStreamDestroy(c->aux);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->tail) (void)hipStreamDestroy(c->tail);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->hstate) (void)hipHostFree(c->hstate);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int cuvhs_get_format(const cuvhs_ctx* c, cuvhs_format* out) {
    if (!c) return CUVHS_E_ARG;
    return cuvhs_format_for_system(c->cfg.system, c->cfg.sample_rate_mhz, out);
}


int cuvhs_reset(cuvhs_ctx* c) {
    if (!c) return CUVHS_E_ARG;
    int r = sync_all(c);
    if (r) return r;
    reset_state(c->hstate);
    HIPCHK(c, hipMemcpy(c->state, c->hstate, sizeof(DecodeState), hipMemcpyHostToDevice));
    return 0;
}

const char* cuvhs_last_error(const cuvhs_ctx* c) {
    if (c) return c->err.c_str();
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return g_free_err.empty() ? "no context" : g_free_err.c_str();
}

int cuvhs_decode_chunk(cuvhs_ctx* c, const void* chunk, uint64_t avail, uint64_t raw_offset, int max_fields,
                       cuvhs_chunk_info* info) {
    if (!c || (!chunk && avail) || !info || max_fields < 1 || max_fields > c->F) return CUVHS_E_ARG;
    memset(info, 0, sizeof(*info));
    info->raw_offset = raw_offset;
    Slot& s = c->slot[0];
    if (!plan_chunk(c, s, chunk, avail, raw_offset, max_fields)) {
        info->status = CUVHS_W_EOF;
        return 0;
    }
    int r = head_stage(c, s);
    if (r) { sync_all(c); return r; }
    HIPCHK(c, hipEventSynchronize(s.head_done));
    const ChunkCtl h = *s.hctl;
    info->fields = h.nf;
    info->fields_k1 = s.fk1;
    info->vsync_candidates = h.vsync_count;
    info->status = h.status;
    info->next_raw_offset = h.next_raw_offset;
    info->guard_blocks_a = h.guard_a;
    info->guard_blocks_b = h.guard_b;
    if (h.nf > 0) {
        r = tail_stage(c, s, nullptr, c->side);
        if (r) { sync_all(c); return r; }
        r = finish_tail(c, s);
        if (r) return r;
    }
    return info->status;
}

int cuvhs_chunk_outputs_get(const cuvhs_ctx* c, cuvhs_chunk_outputs* o) {
    if (!c || !o) return CUVHS_E_ARG;
    o->luma = c->luma;
    o->chroma = c->chroma;
    o->file_loc = (const uint64_t*)c->file_loc;
    o->is_first = c->isf_out;
    o->phase_id = c->phase;
    o->do_count = c->doc;
    o->do_line = c->dol;
    o->do_startx = c->dos;
    o->do_endx = c->doe;
    return 0;
}

int cuvhs_decode_capture(cuvhs_ctx* c, const void* cap, uint64_t cap_len, int max_out, uint16_t* d_luma,
                         uint16_t* d_chroma, uint64_t* d_file_loc, int32_t* d_is_first, int32_t* d_phase_id,
                         int32_t* d_do_count, int32_t* d_do_line, int32_t* d_do_startx, int32_t* d_do_endx,
                         int* fields) {
    if (!c || (!cap && cap_len) || !fields || max_out < 0) return CUVHS_E_ARG;
    *fields = 0;
    const uint64_t spf = (uint64_t)c->fmt.samples_per_field;
    const size_t tbc = (size_t)c->fmt.field_lines * c->fmt.line_len;
    // the stages replay their captured graphs unless profiling or diagnostics
    // need the launch-by-launch schedule
    const bool graphs = c->graphs && !c->prof && !c->diag && !c->serial;
    const OutBases ob{d_luma, d_chroma, d_file_loc, d_is_first, d_phase_id, d_do_count, d_do_line, d_do_startx,
                      d_do_endx};
    // chunk sizing of Pipeline::run (pipeline.cu:2651-2658); out_fields:
    // the fields of the chunks before it
    auto plan = [&](Slot& s, uint64_t off, long long out_fields) -> bool {
        if (off >= cap_len) return false;
        uint64_t remaining = (cap_len - off) / spf;
        if (remaining == 0) return false;
        int this_chunk = (int)std::min<uint64_t>((uint64_t)c->F, remaining);
        if (!plan_chunk(c, s, static_cast<const uint8_t*>(cap) + off * sample_bytes(c), cap_len - off, off,
                        this_chunk))
            return false;
        if (graphs) fill_plan(c, s, ob, out_fields);
        return true;
    };
    auto head = [&](Slot& s) { return graphs ? head_stage_graph(c, s) : head_stage(c, s); };
    if (!plan(c->slot[0], 0, 0)) return 0;
    int rc = head(c->slot[0]);
    if (rc) { sync_all(c); return rc; }
    int total = 0, warn = 0;
    long long chunk = 0;        // the chunk whose head is in flight
    long long finished = 0;     // tails of chunks < finished have been collected
    while (true) {
        Slot& s = c->slot[chunk % kSlots];
        const hipError_t he = spin_wait(s.head_done);
        if (he != hipSuccess) { rc = fail(c, he, "head stage"); break; }
        const int nf = s.hctl->nf;
        const uint64_t next = s.hctl->next_raw_offset;
        warn |= s.hctl->status & ~CUVHS_W_EOF;
        if (nf == 0) break;
        if (total + nf > max_out) { rc = CUVHS_E_CAPACITY; break; }
        OutArrays o{};
        o.luma = d_luma ? d_luma + (size_t)total * tbc : nullptr;
        o.chroma = d_chroma ? d_chroma + (size_t)total * tbc : nullptr;
        o.file_loc = d_file_loc ? d_file_loc + total : nullptr;
        o.is_first = d_is_first ? d_is_first + total : nullptr;
        o.phase_id = d_phase_id ? d_phase_id + total : nullptr;
        o.do_count = d_do_count ? d_do_count + total : nullptr;
        o.do_line = d_do_line ? d_do_line + (size_t)total * kMaxDropouts : nullptr;
        o.do_startx = d_do_startx ? d_do_startx + (size_t)total * kMaxDropouts : nullptr;
        o.do_endx = d_do_endx ? d_do_endx + (size_t)total * kMaxDropouts : nullptr;
        // the next chunk's buffer set is free once the tail of the chunk that
        // last used it has finished; the next head (with its pre-band-pass on
        // the side stream) goes in before this chunk's tail, so the head
        // chain starts as early as it can.  The tail's forked work goes on
        // the aux stream when the process has a hardware queue for it, else
        // on the head stream behind that head (a queue shared with the side
        // stream would hold it behind the next pre-band-pass).
        while (finished <= chunk + 1 - kSlots && rc == 0) rc = finish_tail(c, c->slot[finished++ % kSlots]);
        if (rc) break;
        const bool more = plan(c->slot[(chunk + 1) % kSlots], next, (long long)total + nf);
        if (more && (rc = head(c->slot[(chunk + 1) % kSlots])) != 0) break;
        if ((rc = graphs ? tail_stage_graph(c, s) : tail_stage(c, s, &o, fork_stream(c))) != 0) break;
        total += nf;
        chunk++;
        if (!more) break;
    }
    // tails launched and not yet collected: chunks [finished, chunk)
    for (; finished < chunk; finished++) {
        int r = finish_tail(c, c->slot[finished % kSlots]);
        if (!rc) rc = r;
    }
    int r = sync_all(c);
    if (!rc) rc = r;
    if (rc) return rc;
    *fields = total;
    return warn;
}

static const char* kStageNames[kStages + 1] = {
    "k1_analytic", "k1_video", "k1_guard", "sync_events", "vsync_cluster", "k2_pulses", "k2b_linelocs",
    "k4_hsync_refine", "k5_tbc_resample", "k6_prebpf_wait", "k6_chroma", "k7_dropouts", "field_meta",
    "k6_prebpf"};

int cuvhs_profile(cuvhs_ctx* c, int enable) {
    if (!c) return CUVHS_E_ARG;
    if (enable && !c->slot[0].mark[0]) {
        for (Slot& s : c->slot)
            for (hipEvent_t& e : s.mark) HIPCHK(c, hipEventCreate(&e));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipStreamSynchronize(c->tail));
    HIPCHK(c, hipStreamSynchronize(c->side));
    HIPCHK(c, hipStreamSynchronize(c->aux));
    c->prof = enable != 0;
    c->serial = enable == 2;
    for (double& v : c->stage_ms) v = 0.0;
    c->work_samples = 0;
    c->work_chunks = 0;
    c->work_fields = 0;
    return 0;
}

int cuvhs_diagnostics(cuvhs_ctx* c, int enable) {
    if (!c) return CUVHS_E_ARG;
    // the counters read the FP64 signals of the reference's arithmetic
    if (enable && c->cfg.precision != CUVHS_PRECISION_FP64) return CUVHS_E_UNSUPPORTED;
    int r = sync_all(c);
    if (r) return r;
    c->diag = enable != 0;
    for (int i = 0; i < 9; i++) {
        c->diag_total[i] = 0;
        c->diag_last[i] = 0;
    }
    return 0;
}

int cuvhs_diagnostics_read(const cuvhs_ctx* c, int64_t* totals, int32_t* last_chunk) {
    if (!c) return CUVHS_E_ARG;
    for (int i = 0; i < 9; i++) {
        if (totals) totals[i] = c->diag_total[i];
        if (last_chunk) last_chunk[i] = c->diag_last[i];
    }
    return 9;
}

int cuvhs_k3_linelocs(cuvhs_ctx* c, const int* starts, const int* lens, const int* count, int nf, int* types,
                      double* linelocs, int* is_first) {
    if (!c || !starts || !lens || !count || !linelocs || !is_first || nf < 1 || nf > c->F) return CUVHS_E_ARG;
    HIPCHK(c, run_k3(c->slot[0].ctl, nf, nf, c->fmt, starts, lens, count, types, linelocs, is_first, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cuvhs_profile_work(const cuvhs_ctx* c, uint64_t* samples, int* chunks, int* fields) {
    if (!c) return CUVHS_E_ARG;
    if (samples) *samples = c->work_samples;
    if (chunks) *chunks = c->work_chunks;
    if (fields) *fields = c->work_fields;
    return 0;
}

int cuvhs_k1_matrix_flops(const cuvhs_ctx* c, double* per_segment) {
    if (!c || !per_segment) return CUVHS_E_ARG;
    // workgroups per segment x 4 waves x (steps of the groups with lo x 16
    // products: 4 tiles x real/imaginary x hi/lo; the other steps x 8) x
    // 2 x 32 x 32 x 16
    const double nlo = 4.0 * (c->m_glo1 - c->m_glo0);   // steps with lo parts
    const double prods = 16.0 * nlo + 8.0 * (c->m_nc - nlo);
    *per_segment = c->mixed ? (double)mfma_blocks_per_segment(c->fmt.spfp) * 4.0 * prods * 32768.0 : 0.0;
    return 0;
}

int cuvhs_k1_blocks(const cuvhs_ctx* c, int* block_a, int* valid_a, int* block_b, int* valid_b) {
    if (!c) return CUVHS_E_ARG;
    if (block_a) *block_a = c->k1.rf.B;
    if (valid_a) *valid_a = c->k1.rf.valid;
    if (block_b) *block_b = c->k1.video.B;
    if (valid_b) *valid_b = c->k1.video.valid;
    return 0;
}

int cuvhs_prebpf_block(const cuvhs_ctx* c, int* block, int* valid, int* source) {
    if (!c) return CUVHS_E_ARG;
    if (block) *block = c->pre.bpf.B;
    if (valid) *valid = c->pre.bpf.valid;
    if (source) *source = c->block_source;
    return 0;
}

int cuvhs_profile_read(const cuvhs_ctx* c, double* ms, const char** names, int n) {
    if (!c || n < 0) return CUVHS_E_ARG;
    for (int i = 0; i < n && i <= kStages; i++) {
        if (ms) ms[i] = c->stage_ms[i];
        if (names) names[i] = kStageNames[i];
    }
    return kStages + 1;
}

// ---- stage entry points (synchronous, on the context's stream, slot 0) ----

int cuvhs_fm_demod(cuvhs_ctx* c, const void* raw, uint64_t n_valid, int nf, double* demod, double* d05,
                   double* env) {
    if (!c || !raw || nf < 1 || nf > c->F || !demod || !d05 || !env) return CUVHS_E_ARG;
    Slot& s = c->slot[0];
    if (c->cfg.dc_correct && n_valid > 0)
        HIPCHK(c, run_input_mean(raw, c->cfg.input_format, (long long)n_valid, s.dc_partial, s.dc, c->stream));
    const bool prof = c->prof;
    c->prof = false;
    // the mixed forms' guard band needs the threshold bitmask (its FP64
    // blocks rewrite their bits); the context's bitmask buffer serves
    int r = k1_launch(c, s, raw, (long long)n_valid, nf, c->stream, nullptr, c->mixed ? c->thr_bits : nullptr);
    c->prof = prof;
    if (r) return r;
    const long long n = (long long)nf * c->fmt.spfp;
    HIPCHK(c, run_to_f64(s.demod, c->sig_f32, demod, n, c->stream));
    HIPCHK(c, run_to_f64(s.d05, c->d05_f32, d05, n, c->stream));
    HIPCHK(c, run_to_f64(s.env, c->sig_f32, env, n, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cuvhs_vsync_discover(cuvhs_ctx* c, const double* d05, int total, int* idx, int cap, int* count) {
    if (!c || !d05 || !idx || !count || total < 1 || cap < 0 || total > (long long)c->F * c->fmt.spfp) return CUVHS_E_ARG;
    const Fmt& f = c->fmt;
    Slot& s = c->slot[0];
    HIPCHK(c, run_events(d05, total, f.pulse_threshold_hz, c->thr_bits, s.wg_count, s.wg_off, s.ev_masks, s.events,
                         s.ctl, c->stream));
    HIPCHK(c, run_vsync_cluster(s.events, s.ctl, total, f.vsync_min_w, f.vsync_max_w, cap, idx, s.offsets, 1,
                                f.samples_per_field, f.samples_per_line, 1, 0, nullptr, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.hctl, s.ctl, sizeof(ChunkCtl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *count = s.hctl->vsync_count;
    return 0;
}

int cuvhs_sync_pulses(cuvhs_ctx* c, const double* d05, int total, const uint64_t* offsets, int nf, int* starts,
                      int* lens, int* count) {
    if (!c || !d05 || !offsets || nf < 1 || total < 1 || total > (long long)c->F * c->fmt.spfp) return CUVHS_E_ARG;
    const Fmt& f = c->fmt;
    Slot& s = c->slot[0];
    HIPCHK(c, run_events(d05, total, f.pulse_threshold_hz, c->thr_bits, s.wg_count, s.wg_off, s.ev_masks, s.events,
                         s.ctl, c->stream));
    HIPCHK(c, run_sync_pulses(s.events, s.ctl, (const unsigned long long*)offsets, f.spfp, starts, lens, count, nf,
                              nf, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cuvhs_linelocs(cuvhs_ctx* c, uint64_t raw_offset, const uint64_t* offsets, int nf, int* starts, int* lens,
                   int* count, int* types, double* linelocs, int* is_first) {
    if (!c || !offsets || nf < 1 || nf > c->F) return CUVHS_E_ARG;
    Slot& s = c->slot[0];
    memset(s.hctl, 0, sizeof(ChunkCtl));
    s.hctl->raw_offset = raw_offset;
    s.hctl->nf = nf;
    HIPCHK(c, hipMemcpyAsync(s.ctl, s.hctl, sizeof(ChunkCtl), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.k2s, starts, (size_t)nf * kMaxPulses * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.k2l, lens, (size_t)nf * kMaxPulses * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.k2c, count, (size_t)nf * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, run_k2b(s.ctl, s.ctl, c->state, nf, nf, c->fmt, (const unsigned long long*)offsets, s.k2s, s.k2l,
                      s.k2c, s.vps, s.vpl, s.vptg, (K2bField*)s.summ, (K2bAnchorOut*)c->anchors, starts, lens,
                      count, types, linelocs, is_first, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.hctl, s.ctl, sizeof(ChunkCtl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (s.hctl->vp_overflow) {
        c->err = "K2b validated-pulse capacity exceeded";
        return CUVHS_E_CAPACITY;
    }
    return 0;
}

int cuvhs_hsync_refine(cuvhs_ctx* c, const double* d05, int total, double* ll, int nf) {
    if (!c || !d05 || !ll || nf < 1 || total != nf * c->fmt.spfp) return CUVHS_E_ARG;
    HIPCHK(c, run_hsync_refine(d05, 0, ll, c->slot[0].ctl, nf, nf, c->fmt, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cuvhs_tbc_resample(cuvhs_ctx* c, const double* demod, int total, const double* ll, int nf, uint16_t* luma) {
    if (!c || !demod || !ll || !luma || nf < 1 || total != nf * c->fmt.spfp) return CUVHS_E_ARG;
    HIPCHK(c, run_tbc_resample(demod, 0, ll, luma, c->slot[0].ctl, nf, nf, c->fmt, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cuvhs_chroma(cuvhs_ctx* c, const void* raw, uint64_t n_valid, const double* ll, int nf, uint16_t* out,
                 int* h_phase) {
    if (!c || !raw || !ll || !out || nf < 1 || nf > c->F) return CUVHS_E_ARG;
    Slot& s = c->slot[0];
    if (c->cfg.dc_correct && n_valid > 0)
        HIPCHK(c, run_input_mean(raw, c->cfg.input_format, (long long)n_valid, s.dc_partial, s.dc, c->stream));
    int r = prebpf_launch(c, raw, (long long)n_valid, nf, s.craw, s.dc, c->stream);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(c->ll, ll, (size_t)nf * c->fmt.lines_per_frame * 8, hipMemcpyDeviceToDevice, c->stream));
    ChromaLaunch a = chroma_launch(c, s.ctl, nf, s.craw);
    a.out = out;
    HIPCHK(c, run_chroma(a, c->state, nf, c->m0, c->m1, c->cs_tab, c->iq0, c->track, c->phase, 1, c->stream));
    if (h_phase) HIPCHK(c, hipMemcpyAsync(h_phase, c->phase, (size_t)nf * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cuvhs_dropouts(cuvhs_ctx* c, const double* env, const double* ll, int nf, uint16_t* luma, uint16_t* chroma,
                   int* lines, int* starts, int* ends, int* count) {
    if (!c || !env || !ll || nf < 1 || nf > c->F) return CUVHS_E_ARG;
    HIPCHK(c, run_dropouts(env, ll, c->slot[0].ctl, nf, nf, c->fmt, c->do_partial, c->do_cls, luma, chroma, lines,
                           starts, ends, count, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// dropout_detect on caller buffers without a context and without any
// allocation: the scratch comes from the file-scope array above.  Enqueues on
// `stream` (NULL: the default stream of the current device) and returns.
int cuvhs_dropouts_noalloc(double sample_rate_mhz, const double* env, const double* ll, int nf, uint16_t* luma,
                           uint16_t* chroma, int* lines, int* starts, int* ends, int* count, void* stream) {
    if (!env || !ll || !luma || !chroma || !lines || !starts || !ends || !count || nf < 1)
        return noalloc_fail(CUVHS_E_ARG, "cuvhs_dropouts_noalloc: null buffer or no field");
    if (!(sample_rate_mhz >= 16.0 && sample_rate_mhz <= 40.0))
        return noalloc_fail(CUVHS_E_UNSUPPORTED, "cuvhs_dropouts_noalloc: sample rate outside 16-40 MHz");
    if (nf > kNoAllocMaxFields)
        return noalloc_fail(CUVHS_E_CAPACITY, "cuvhs_dropouts_noalloc: " + std::to_string(nf) +
                                                  " fields exceed the scratch limit of 128 fields per call");
    const size_t cls_bytes = dropout_class_bytes(kNoAllocMaxFields);
    if (cls_bytes + dropout_partial_count(kNoAllocMaxFields) * sizeof(double) > kNoAllocBytes)
        return noalloc_fail(CUVHS_E_UNSUPPORTED, "cuvhs_dropouts_noalloc: scratch layout larger than the array");
    void* base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    unsigned char* cls = static_cast<unsigned char*>(base);
    double* partial = reinterpret_cast<double*>(cls + cls_bytes);
    const Fmt f = make_fmt(sample_rate_mhz);
    hipError_t e = run_dropouts(env, ll, nullptr, nf, nf, f, partial, cls, luma, chroma, lines, starts, ends, count,
                                static_cast<hipStream_t>(stream));
    if (e != hipSuccess)
        return noalloc_fail(CUVHS_E_HIP, std::string("cuvhs_dropouts_noalloc: ") + hipGetErrorString(e));
    return 0;
}

// Convolution-block self test (not a reference stage): one block of the
// kind the decoder runs, with the spectrum given in natural bin order and
// permuted here exactly as the context uploads its filters.
int cuvhs_selftest_block(int B, int is_f32, const void* d_x, const void* d_g, void* d_y) {
    if (!d_x || !d_g || !d_y) return CUVHS_E_ARG;
    // the shapes the decoder runs; anything else is refused before any copy
    const bool ok = B == 1024 || B == 4096 || B == 8192 || B == 16384;
    if (!ok) return CUVHS_E_ARG;
    return is_f32 ? block_selftest<float2, float>(B, 1, d_x, d_g, d_y)
                  : block_selftest<double2, double>(B, 0, d_x, d_g, d_y);
}

// FP64 arctangent self test (not a reference stage).
int cuvhs_selftest_atan2(const double* d_y, const double* d_x, double* d_out, int64_t n) {
    if (!d_y || !d_x || !d_out || n < 1) return CUVHS_E_ARG;
    hipError_t e = run_atan2_selftest(d_y, d_x, d_out, (long long)n, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? 0 : CUVHS_E_HIP;
}

}  // extern "C"
