// Decoder context, chunk driver and the C ABI (include/cuvhs.h).
//
// Reference: Pipeline::allocate_buffers / process_chunk / run
// (pipeline.cu:1383-1475, 1525-2624, 2626-2713).  A chunk is enqueued as a
// fixed sequence of kernels on the context's stream; the field count that
// VSYNC clustering produces stays in a device control block that every later
// kernel reads, so the only host synchronisation per chunk is the final copy
// of that control block (next chunk offset, field count).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cuvhs.h"
#include "common.h"
#include "design.h"
#include "launch.h"

using namespace cuvhs;

struct cuvhs_ctx {
    cuvhs_config cfg{};
    Fmt fmt{};
    int F = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // the chroma pre-band-pass needs only the raw chunk: it runs on a second
    // stream (default priority) next to the sync/line-location stages
    hipStream_t side = nullptr;
    hipEvent_t ev_k1 = nullptr, ev_pre = nullptr, ev_pre0 = nullptr;
    std::string err;
    // designs
    K1Design k1;
    PreDesign pre;
    // device tables
    double2* G_a = nullptr;
    double2* G_b = nullptr;
    float2* G_p = nullptr;
    double2* tw_a = nullptr;   // W_B for part A's B
    double2* tw_b = nullptr;
    float2* tw_p = nullptr;
    double2* tw_1024 = nullptr;
    double* cs_tab = nullptr;  // cos/sin of the burst detector
    // signal buffers
    long long buf_n = 0;       // F * spfp + pad
    double* rawdemod = nullptr;
    double* demod = nullptr;
    double* d05 = nullptr;
    double* env = nullptr;
    float* craw = nullptr;
    // events
    int* events = nullptr;
    int* wg_count = nullptr;
    int* wg_off = nullptr;
    int* cand = nullptr;
    // per field
    unsigned long long* offsets = nullptr;
    int *k2s = nullptr, *k2l = nullptr, *k2c = nullptr;
    int *ps = nullptr, *pl = nullptr, *pc = nullptr, *pt = nullptr;
    int *vps = nullptr, *vpl = nullptr;
    unsigned char* vptg = nullptr;
    void* summ = nullptr;
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
    ChunkCtl* ctl = nullptr;       // device
    ChunkCtl* hctl = nullptr;      // pinned host shadow
    std::vector<void*> allocs;
    // per-stage timing with HIP events on the context stream
    static constexpr int kStages = 12;
    bool prof = false;
    hipEvent_t ev[kStages + 1] = {};
    double stage_ms[kStages + 1] = {};  // + the side-stream pre-band-pass
    long long stage_calls = 0;
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

template <typename V, typename S>
hipError_t upload_spectrum(cuvhs_ctx* c, V** dst, const std::vector<cdbl>& s) {
    std::vector<V> h(s.size());
    for (size_t i = 0; i < s.size(); i++) { h[i].x = (S)s[i].real(); h[i].y = (S)s[i].imag(); }
    hipError_t e = dalloc(c, dst, h.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice);
}

void reset_state(ChunkCtl* h) {
    memset(h, 0, sizeof(*h));
    h->k2b.prev_first_field = -1;
    h->k2b.prev_progressive_field = -1;
    h->k2b.prev_first_hsync_readloc = -1;
    h->k2b.prev_first_hsync_loc = -1.0;
    h->k2b.prev_first_hsync_diff = -1.0;
    h->chroma.valid = 0;
    h->last_parity = 1;
}

ChromaLaunch chroma_launch(cuvhs_ctx* c, int nf_override) {
    ChromaLaunch a{};
    a.craw = c->craw;
    a.linelocs = c->ll;
    a.ctl = c->ctl;
    a.nf_override = nf_override;
    a.spfp = c->fmt.spfp;
    a.het_scale = (c->fmt.fsc + c->fmt.chroma_under) / c->fmt.output_rate;
    a.fft_scale = 1.0 / (double)kChromaFFT;
    a.burst_abs_ref = c->fmt.burst_abs_ref;
    a.burst_start = c->fmt.burst_start;
    int fsc_bin = (int)(c->fmt.fsc / c->fmt.output_rate * kChromaFFT + 0.5);
    int bw = (int)(500000.0 / c->fmt.output_rate * kChromaFFT + 0.5);
    if (bw < 10) bw = 10;
    a.lo = fsc_bin - bw;
    a.hi = fsc_bin + bw;
    a.burst = c->burst;
    a.out = c->chroma;
    a.tw = c->tw_1024;
    return a;
}

void mark(cuvhs_ctx* c, int i) {
    if (c->prof) (void)hipEventRecord(c->ev[i], c->stream);
}

int k1_launch(cuvhs_ctx* c, const uint8_t* chunk, long long n_valid, int nfk1) {
    const Fmt& f = c->fmt;
    long long T = (long long)nfk1 * f.spfp;
    K1ALaunch a{};
    a.cap = chunk;
    a.n_valid = n_valid;
    a.total = T;
    a.spfp = f.spfp;
    a.N = c->k1.N;
    a.lm = c->k1.rf.lm;
    a.lp = c->k1.rf.lp;
    a.V = c->k1.rf.valid;
    a.scale = f.sample_rate / (2.0 * 3.14159265358979323846);
    a.G = c->G_a;
    a.tw = c->tw_a;
    a.rawdemod = c->rawdemod;
    a.env = c->env;
    HIPCHK(c, run_k1_part_a(c->k1.rf.B, a, nfk1, c->stream));
    mark(c, 1);
    K1BLaunch b{};
    b.rawdemod = c->rawdemod;
    b.total = T;
    b.spfp = f.spfp;
    b.N = c->k1.N;
    b.lm = c->k1.video.lm;
    b.lp = c->k1.video.lp;
    b.V = c->k1.video.valid;
    b.G = c->G_b;
    b.tw = c->tw_b;
    b.demod = c->demod;
    b.demod05 = c->d05;
    HIPCHK(c, run_k1_part_b(c->k1.video.B, b, nfk1, c->stream));
    // beyond the demodulated region the buffers read as zero
    HIPCHK(c, run_fill_f64(c->demod + T, c->buf_n - T, 0.0, c->stream));
    HIPCHK(c, run_fill_f64(c->d05 + T, c->buf_n - T, 0.0, c->stream));
    HIPCHK(c, run_fill_f64(c->env + T, c->buf_n - T, 0.0, c->stream));
    return 0;
}

int prebpf_launch(cuvhs_ctx* c, const uint8_t* chunk, long long n_valid, int nf_max, hipStream_t st) {
    PreLaunch p{};
    p.cap = chunk;
    p.n_valid = n_valid;
    p.spfp = c->fmt.spfp;
    p.N2 = c->pre.N2;
    p.lm = c->pre.bpf.lm;
    p.lp = c->pre.bpf.lp;
    p.V = c->pre.bpf.valid;
    p.G = c->G_p;
    p.tw = c->tw_p;
    p.out = c->craw;
    HIPCHK(c, run_prebpf(c->pre.bpf.B, p, nf_max, st));
    return 0;
}

}  // namespace

extern "C" {

int cuvhs_format_for(double mhz, cuvhs_format* o) {
    if (!o || !(mhz > 0.0)) return CUVHS_E_ARG;
    Fmt f = make_fmt(mhz);
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
    return 0;
}

int cuvhs_create(const cuvhs_config* cfg, cuvhs_ctx** out) {
    if (!cfg || !out) return CUVHS_E_ARG;
    *out = nullptr;
    if (cfg->system != CUVHS_SYSTEM_NTSC || cfg->precision != CUVHS_PRECISION_FP64) return CUVHS_E_UNSUPPORTED;
    if (!(cfg->sample_rate_mhz >= 16.0 && cfg->sample_rate_mhz <= 40.0)) return CUVHS_E_UNSUPPORTED;
    if (cfg->batch_fields < 1 || cfg->batch_fields > 4096) return CUVHS_E_ARG;
    cuvhs_ctx* c = new cuvhs_ctx();
    c->cfg = *cfg;
    c->fmt = make_fmt(cfg->sample_rate_mhz);
    c->F = cfg->batch_fields;
    const Fmt& f = c->fmt;
    if ((long long)c->F * f.spfp > 0x3fffffffLL) { delete c; return CUVHS_E_CAPACITY; }
    auto bad = [&](hipError_t e, const char* w) { fail(c, e, w); cuvhs_destroy(c); return CUVHS_E_HIP; };
    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) return bad(e, "hipSetDevice");
    if (cfg->stream) {
        c->stream = (hipStream_t)cfg->stream;
    } else {
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) return bad(e, "hipStreamCreate");
        c->own_stream = true;
    }
    e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e != hipSuccess) return bad(e, "hipStreamCreate side");
    for (hipEvent_t* ev : {&c->ev_k1, &c->ev_pre, &c->ev_pre0}) {
        e = hipEventCreate(ev);
        if (e != hipSuccess) return bad(e, "hipEventCreate");
    }
    try {
        const int bk1[2] = {4096, 8192};
        c->k1 = design_k1(f, 1e-13, bk1, 2, 1);
        const int bpre[2] = {8192, 16384};
        c->pre = design_prebpf(f, 1e-9, bpre, 2);
    } catch (...) {
        cuvhs_destroy(c);
        return CUVHS_E_UNSUPPORTED;
    }
#define ALLOC(p, n)                                    \
    do {                                               \
        hipError_t e2 = dalloc(c, &(p), (size_t)(n));  \
        if (e2 != hipSuccess) return bad(e2, #p);      \
    } while (0)
    if ((e = upload_spectrum<double2, double>(c, &c->G_a, c->k1.rf.spectrum)) != hipSuccess) return bad(e, "G_a");
    if ((e = upload_spectrum<double2, double>(c, &c->G_b, c->k1.video.spectrum)) != hipSuccess) return bad(e, "G_b");
    if ((e = upload_spectrum<float2, float>(c, &c->G_p, c->pre.bpf.spectrum)) != hipSuccess) return bad(e, "G_p");
    {
        auto ta = twiddles<double2, double>(c->k1.rf.B);
        ALLOC(c->tw_a, ta.size());
        if ((e = hipMemcpy(c->tw_a, ta.data(), ta.size() * sizeof(double2), hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "twiddles");
        auto tb = twiddles<double2, double>(c->k1.video.B);
        ALLOC(c->tw_b, tb.size());
        if ((e = hipMemcpy(c->tw_b, tb.data(), tb.size() * sizeof(double2), hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "twiddles");
        auto tp = twiddles<float2, float>(c->pre.bpf.B);
        ALLOC(c->tw_p, tp.size());
        if ((e = hipMemcpy(c->tw_p, tp.data(), tp.size() * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "twiddles");
        auto t1 = twiddles<double2, double>(kChromaFFT);
        ALLOC(c->tw_1024, t1.size());
        if ((e = hipMemcpy(c->tw_1024, t1.data(), t1.size() * sizeof(double2), hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "twiddles");
        // burst I/Q reference (chroma_decode.cu:435-437), computed on the host
        std::vector<double> cs(2 * 36);
        for (int i = 0; i < 36; i++) {
            double t = 2.0 * M_PI * f.fsc / f.output_rate * (double)(f.burst_start + i);
            cs[(size_t)2 * i] = cos(t);
            cs[(size_t)2 * i + 1] = sin(t);
        }
        ALLOC(c->cs_tab, cs.size());
        if ((e = hipMemcpy(c->cs_tab, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) return bad(e, "burst table");
    }
    const long long S = (long long)c->F * f.spfp;
    c->buf_n = S + f.spfp;  // one segment of zero padding past the chunk
    ALLOC(c->rawdemod, S);
    ALLOC(c->demod, c->buf_n);
    ALLOC(c->d05, c->buf_n);
    ALLOC(c->env, c->buf_n);
    ALLOC(c->craw, S);
    ALLOC(c->events, S + 2);
    ALLOC(c->wg_count, events_nwg(S) + 1);
    ALLOC(c->wg_off, events_nwg(S) + 1);
    ALLOC(c->cand, 20 * c->F);
    const size_t Fz = (size_t)c->F;
    ALLOC(c->offsets, Fz);
    ALLOC(c->k2s, Fz * kMaxPulses);
    ALLOC(c->k2l, Fz * kMaxPulses);
    ALLOC(c->k2c, Fz);
    ALLOC(c->ps, Fz * kMaxPulses);
    ALLOC(c->pl, Fz * kMaxPulses);
    ALLOC(c->pc, Fz);
    ALLOC(c->pt, Fz * kMaxPulses);
    ALLOC(c->vps, Fz * kVpCap);
    ALLOC(c->vpl, Fz * kVpCap);
    ALLOC(c->vptg, Fz * kVpCap);
    {
        unsigned char* p1 = nullptr;
        ALLOC(p1, Fz * k2b_field_summary_bytes());
        c->summ = p1;
        unsigned char* p2 = nullptr;
        ALLOC(p2, Fz * k2b_anchor_bytes());
        c->anchors = p2;
    }
    ALLOC(c->ll, Fz * kLinesPerFrame);
    ALLOC(c->isf_k2b, Fz);
    ALLOC(c->luma, Fz * kFieldLines * kLineLen);
    ALLOC(c->chroma, Fz * kFieldLines * kLineLen);
    ALLOC(c->burst, Fz * kFieldLines * 36);
    ALLOC(c->m0, Fz);
    ALLOC(c->m1, Fz);
    ALLOC(c->iq0, 2 * kFieldLines);
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
    ALLOC(c->ctl, 1);
#undef ALLOC
    e = hipHostMalloc((void**)&c->hctl, sizeof(ChunkCtl), hipHostMallocDefault);
    if (e != hipSuccess) return bad(e, "hipHostMalloc");
    reset_state(c->hctl);
    e = hipMemcpy(c->ctl, c->hctl, sizeof(ChunkCtl), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bad(e, "ctl upload");
    *out = c;
    return 0;
}

int cuvhs_destroy(cuvhs_ctx* c) {
    if (!c) return CUVHS_E_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    for (hipEvent_t e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {c->ev_k1, c->ev_pre, c->ev_pre0})
        if (e) (void)hipEventDestroy(e);
    if (c->side) (void)hipStreamDestroy(c->side);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->hctl) (void)hipHostFree(c->hctl);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int cuvhs_get_format(const cuvhs_ctx* c, cuvhs_format* out) {
    if (!c) return CUVHS_E_ARG;
    return cuvhs_format_for(c->cfg.sample_rate_mhz, out);
}

int cuvhs_reset(cuvhs_ctx* c) {
    if (!c) return CUVHS_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    reset_state(c->hctl);
    HIPCHK(c, hipMemcpy(c->ctl, c->hctl, sizeof(ChunkCtl), hipMemcpyHostToDevice));
    return 0;
}

const char* cuvhs_last_error(const cuvhs_ctx* c) { return c ? c->err.c_str() : "no context"; }

int cuvhs_decode_chunk(cuvhs_ctx* c, const uint8_t* chunk, uint64_t avail, uint64_t raw_offset, int max_fields,
                       cuvhs_chunk_info* info) {
    if (!c || !chunk || !info || max_fields < 1 || max_fields > c->F) return CUVHS_E_ARG;
    memset(info, 0, sizeof(*info));
    info->raw_offset = raw_offset;
    const Fmt& f = c->fmt;
    const uint64_t spf = (uint64_t)f.samples_per_field;
    const uint64_t total_samples = (uint64_t)max_fields * (uint64_t)f.spfp;
    // RawReader::read_at (raw_reader.cpp:128-152) + pipeline.cu:1534-1548
    uint64_t n_read = std::min<uint64_t>(total_samples, avail);
    if (n_read < spf) {
        info->status = CUVHS_W_EOF;
        return 0;
    }
    int fk1 = (int)(n_read / spf);
    if (fk1 > max_fields) fk1 = max_fields;
    const long long T = (long long)fk1 * f.spfp;
    // control block: carried state from the host shadow, per-chunk fields reset
    c->hctl->fields_k1 = fk1;
    c->hctl->nf = 0;
    c->hctl->vsync_count = 0;
    c->hctl->status = 0;
    c->hctl->vp_overflow = 0;
    c->hctl->nevents = 0;
    c->hctl->raw_offset = raw_offset;
    c->hctl->next_raw_offset = 0;
    HIPCHK(c, hipMemcpyAsync(c->ctl, c->hctl, sizeof(ChunkCtl), hipMemcpyHostToDevice, c->stream));
    mark(c, 0);
    int r = k1_launch(c, chunk, (long long)n_read, fk1);
    if (r) return r;
    mark(c, 2);
    HIPCHK(c, hipEventRecord(c->ev_k1, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_k1, 0));
    HIPCHK(c, hipEventRecord(c->ev_pre0, c->side));
    r = prebpf_launch(c, chunk, (long long)n_read, max_fields, c->side);
    if (r) return r;
    HIPCHK(c, hipEventRecord(c->ev_pre, c->side));
    HIPCHK(c, run_events(c->d05, T, f.pulse_threshold_hz, c->wg_count, c->wg_off, c->events, c->ctl, c->stream));
    mark(c, 3);
    HIPCHK(c, run_vsync_cluster(c->events, c->ctl, T, f.vsync_min_w, f.vsync_max_w, fk1 * 20, c->cand, c->offsets,
                                max_fields, f.samples_per_field, f.samples_per_line, fk1, 1, c->stream));
    mark(c, 4);
    HIPCHK(c, run_sync_pulses(c->events, c->ctl, c->offsets, f.spfp, c->k2s, c->k2l, c->k2c, max_fields, 0,
                              c->stream));
    mark(c, 5);
    HIPCHK(c, run_k2b(c->ctl, c->ctl, 0, max_fields, f, c->offsets, c->k2s, c->k2l, c->k2c, c->vps, c->vpl, c->vptg,
                      (K2bField*)c->summ, (K2bAnchorOut*)c->anchors, c->ps, c->pl, c->pc, c->pt, c->ll, c->isf_k2b,
                      c->stream));
    mark(c, 6);
    HIPCHK(c, run_hsync_refine(c->d05, c->ll, c->ctl, 0, max_fields, f, c->stream));
    mark(c, 7);
    HIPCHK(c, run_tbc_resample(c->demod, c->ll, c->luma, c->ctl, 0, max_fields, f, c->stream));
    mark(c, 8);
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_pre, 0));
    mark(c, 9);
    HIPCHK(c, run_chroma(chroma_launch(c, 0), c->ctl, max_fields, c->m0, c->m1, c->cs_tab, c->iq0, c->track, c->phase,
                         c->hctl->chroma.valid ? 0 : 1, c->stream));
    mark(c, 10);
    HIPCHK(c, run_dropouts(c->env, c->ll, c->ctl, 0, max_fields, f, c->do_partial, c->do_cls, c->luma, c->chroma, c->dol, c->dos, c->doe,
                           c->doc, c->stream));
    mark(c, 11);
    HIPCHK(c, run_field_meta(c->ctl, c->offsets, c->isf_k2b, c->file_loc, c->isf_out, c->stream));
    mark(c, 12);
    HIPCHK(c, hipMemcpyAsync(c->hctl, c->ctl, sizeof(ChunkCtl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->prof) {
        for (int i = 0; i < cuvhs_ctx::kStages; i++) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]) == hipSuccess) c->stage_ms[i] += ms;
        }
        // the pre-band-pass runs on the side stream: its own duration, and
        // the main-stream wait for it stays in its slot
        float pre = 0.0f;
        if (hipEventElapsedTime(&pre, c->ev_pre0, c->ev_pre) == hipSuccess) c->stage_ms[cuvhs_ctx::kStages] += pre;
        c->stage_calls++;
    }
    info->fields = c->hctl->nf;
    info->fields_k1 = fk1;
    info->vsync_candidates = c->hctl->vsync_count;
    info->status = c->hctl->status;
    info->next_raw_offset = c->hctl->next_raw_offset;
    if (c->hctl->vp_overflow) {
        c->err = "K2b validated-pulse capacity exceeded";
        return CUVHS_E_CAPACITY;
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

int cuvhs_decode_capture(cuvhs_ctx* c, const uint8_t* cap, uint64_t cap_len, int max_out, uint16_t* d_luma,
                         uint16_t* d_chroma, uint64_t* d_file_loc, int32_t* d_is_first, int32_t* d_phase_id,
                         int32_t* d_do_count, int32_t* d_do_line, int32_t* d_do_startx, int32_t* d_do_endx,
                         int* fields) {
    if (!c || !cap || !fields || max_out < 0) return CUVHS_E_ARG;
    *fields = 0;
    const uint64_t spf = (uint64_t)c->fmt.samples_per_field;
    const size_t tbc = (size_t)kFieldLines * kLineLen;
    uint64_t raw_offset = 0;
    int total = 0;
    int warn = 0;
    while (true) {
        if (raw_offset >= cap_len) break;
        uint64_t remaining = (cap_len - raw_offset) / spf;
        if (remaining == 0) break;
        int this_chunk = (int)std::min<uint64_t>((uint64_t)c->F, remaining);
        cuvhs_chunk_info info;
        int r = cuvhs_decode_chunk(c, cap + raw_offset, cap_len - raw_offset, raw_offset, this_chunk, &info);
        if (r < 0) return r;
        warn |= info.status & ~CUVHS_W_EOF;
        if (info.fields == 0) break;
        int n = info.fields;
        if (total + n > max_out) return CUVHS_E_CAPACITY;
        hipStream_t st = c->stream;
        if (d_luma) HIPCHK(c, hipMemcpyAsync(d_luma + (size_t)total * tbc, c->luma, n * tbc * 2, hipMemcpyDeviceToDevice, st));
        if (d_chroma) HIPCHK(c, hipMemcpyAsync(d_chroma + (size_t)total * tbc, c->chroma, n * tbc * 2, hipMemcpyDeviceToDevice, st));
        if (d_file_loc) HIPCHK(c, hipMemcpyAsync(d_file_loc + total, c->file_loc, n * 8, hipMemcpyDeviceToDevice, st));
        if (d_is_first) HIPCHK(c, hipMemcpyAsync(d_is_first + total, c->isf_out, n * 4, hipMemcpyDeviceToDevice, st));
        if (d_phase_id) HIPCHK(c, hipMemcpyAsync(d_phase_id + total, c->phase, n * 4, hipMemcpyDeviceToDevice, st));
        if (d_do_count) HIPCHK(c, hipMemcpyAsync(d_do_count + total, c->doc, n * 4, hipMemcpyDeviceToDevice, st));
        size_t dsz = (size_t)n * kMaxDropouts * 4;
        if (d_do_line) HIPCHK(c, hipMemcpyAsync(d_do_line + (size_t)total * kMaxDropouts, c->dol, dsz, hipMemcpyDeviceToDevice, st));
        if (d_do_startx) HIPCHK(c, hipMemcpyAsync(d_do_startx + (size_t)total * kMaxDropouts, c->dos, dsz, hipMemcpyDeviceToDevice, st));
        if (d_do_endx) HIPCHK(c, hipMemcpyAsync(d_do_endx + (size_t)total * kMaxDropouts, c->doe, dsz, hipMemcpyDeviceToDevice, st));
        total += n;
        raw_offset = info.next_raw_offset;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *fields = total;
    return warn;
}

static const char* kStageNames[cuvhs_ctx::kStages + 1] = {
    "k1_analytic", "k1_video", "sync_events", "vsync_cluster", "k2_pulses", "k2b_linelocs",
    "k4_hsync_refine", "k5_tbc_resample", "k6_prebpf_wait", "k6_chroma", "k7_dropouts", "field_meta",
    "k6_prebpf"};

int cuvhs_profile(cuvhs_ctx* c, int enable) {
    if (!c) return CUVHS_E_ARG;
    if (enable && !c->ev[0]) {
        for (hipEvent_t& e : c->ev) HIPCHK(c, hipEventCreate(&e));
    }
    c->prof = enable != 0;
    for (double& v : c->stage_ms) v = 0.0;
    c->stage_calls = 0;
    return 0;
}

int cuvhs_profile_read(const cuvhs_ctx* c, double* ms, const char** names, int n) {
    if (!c || n < 0) return CUVHS_E_ARG;
    for (int i = 0; i < n && i <= cuvhs_ctx::kStages; i++) {
        if (ms) ms[i] = c->stage_ms[i];
        if (names) names[i] = kStageNames[i];
    }
    return cuvhs_ctx::kStages + 1;
}

// ---- stage entry points ----

int cuvhs_fm_demod(cuvhs_ctx* c, const uint8_t* raw, uint64_t n_valid, int nf, double* demod, double* d05,
                   double* env) {
    if (!c || !raw || nf < 1 || nf > c->F || !demod || !d05 || !env) return CUVHS_E_ARG;
    int r = k1_launch(c, raw, (long long)n_valid, nf);
    if (r) return r;
    size_t n = (size_t)nf * c->fmt.spfp * sizeof(double);
    HIPCHK(c, hipMemcpyAsync(demod, c->demod, n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d05, c->d05, n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(env, c->env, n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cuvhs_vsync_discover(cuvhs_ctx* c, const double* d05, int total, int* idx, int cap, int* count) {
    if (!c || !d05 || !idx || !count || total < 1 || cap < 0 || total > (long long)c->F * c->fmt.spfp) return CUVHS_E_ARG;
    const Fmt& f = c->fmt;
    HIPCHK(c, run_events(d05, total, f.pulse_threshold_hz, c->wg_count, c->wg_off, c->events, c->ctl, c->stream));
    HIPCHK(c, run_vsync_cluster(c->events, c->ctl, total, f.vsync_min_w, f.vsync_max_w, cap, idx, c->offsets, 1,
                                f.samples_per_field, f.samples_per_line, 1, 0, c->stream));
    ChunkCtl h;
    HIPCHK(c, hipMemcpyAsync(&h, c->ctl, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *count = h.vsync_count;
    return 0;
}

int cuvhs_sync_pulses(cuvhs_ctx* c, const double* d05, int total, const uint64_t* offsets, int nf, int* starts,
                      int* lens, int* count) {
    if (!c || !d05 || !offsets || nf < 1 || total < 1 || total > (long long)c->F * c->fmt.spfp) return CUVHS_E_ARG;
    const Fmt& f = c->fmt;
    HIPCHK(c, run_events(d05, total, f.pulse_threshold_hz, c->wg_count, c->wg_off, c->events, c->ctl, c->stream));
    HIPCHK(c, run_sync_pulses(c->events, c->ctl, (const unsigned long long*)offsets, f.spfp, starts, lens, count, nf,
                              nf, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cuvhs_linelocs(cuvhs_ctx* c, uint64_t raw_offset, const uint64_t* offsets, int nf, int* starts, int* lens,
                   int* count, int* types, double* linelocs, int* is_first) {
    if (!c || !offsets || nf < 1 || nf > c->F) return CUVHS_E_ARG;
    c->hctl->raw_offset = raw_offset;
    c->hctl->nf = nf;
    c->hctl->vp_overflow = 0;
    HIPCHK(c, hipMemcpyAsync(c->ctl, c->hctl, sizeof(ChunkCtl), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->k2s, starts, (size_t)nf * kMaxPulses * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->k2l, lens, (size_t)nf * kMaxPulses * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->k2c, count, (size_t)nf * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, run_k2b(c->ctl, c->ctl, nf, nf, c->fmt, (const unsigned long long*)offsets, c->k2s, c->k2l, c->k2c,
                      c->vps, c->vpl, c->vptg, (K2bField*)c->summ, (K2bAnchorOut*)c->anchors, starts, lens, count,
                      types, linelocs, is_first, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->hctl, c->ctl, sizeof(ChunkCtl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return c->hctl->vp_overflow ? CUVHS_E_CAPACITY : 0;
}

int cuvhs_hsync_refine(cuvhs_ctx* c, const double* d05, int total, double* ll, int nf) {
    if (!c || !d05 || !ll || nf < 1 || total != nf * c->fmt.spfp) return CUVHS_E_ARG;
    HIPCHK(c, run_hsync_refine(d05, ll, c->ctl, nf, nf, c->fmt, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cuvhs_tbc_resample(cuvhs_ctx* c, const double* demod, int total, const double* ll, int nf, uint16_t* luma) {
    if (!c || !demod || !ll || !luma || nf < 1 || total != nf * c->fmt.spfp) return CUVHS_E_ARG;
    HIPCHK(c, run_tbc_resample(demod, ll, luma, c->ctl, nf, nf, c->fmt, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cuvhs_chroma(cuvhs_ctx* c, const uint8_t* raw, uint64_t n_valid, const double* ll, int nf, uint16_t* out,
                 int* h_phase) {
    if (!c || !raw || !ll || !out || nf < 1 || nf > c->F) return CUVHS_E_ARG;
    int r = prebpf_launch(c, raw, (long long)n_valid, nf, c->stream);
    if (r) return r;
    HIPCHK(c, hipMemcpyAsync(c->ll, ll, (size_t)nf * kLinesPerFrame * 8, hipMemcpyDeviceToDevice, c->stream));
    ChromaLaunch a = chroma_launch(c, nf);
    a.out = out;
    HIPCHK(c, run_chroma(a, c->ctl, nf, c->m0, c->m1, c->cs_tab, c->iq0, c->track, c->phase,
                         c->hctl->chroma.valid ? 0 : 1, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->hctl, c->ctl, sizeof(ChunkCtl), hipMemcpyDeviceToHost, c->stream));
    if (h_phase) HIPCHK(c, hipMemcpyAsync(h_phase, c->phase, (size_t)nf * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int cuvhs_dropouts(cuvhs_ctx* c, const double* env, const double* ll, int nf, uint16_t* luma, uint16_t* chroma,
                   int* lines, int* starts, int* ends, int* count) {
    if (!c || !env || !ll || nf < 1 || nf > c->F) return CUVHS_E_ARG;
    HIPCHK(c, run_dropouts(env, ll, c->ctl, nf, nf, c->fmt, c->do_partial, c->do_cls, luma, chroma, lines, starts, ends, count, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// FFT building-block self test (not a reference stage): B-point LDS FFT of
// device data with the context-independent twiddle table.
int cuvhs_selftest_fft(int B, int is_f32, const void* d_in, void* d_out, int sign) {
    void* tw = nullptr;
    size_t es = is_f32 ? sizeof(float2) : sizeof(double2);
    if (hipMalloc(&tw, (size_t)B * es) != hipSuccess) return CUVHS_E_ALLOC;
    if (is_f32) {
        auto t = twiddles<float2, float>(B);
        if (hipMemcpy(tw, t.data(), t.size() * es, hipMemcpyHostToDevice) != hipSuccess) { (void)hipFree(tw); return CUVHS_E_HIP; }
    } else {
        auto t = twiddles<double2, double>(B);
        if (hipMemcpy(tw, t.data(), t.size() * es, hipMemcpyHostToDevice) != hipSuccess) { (void)hipFree(tw); return CUVHS_E_HIP; }
    }
    hipError_t e = run_fft_selftest(B, is_f32, d_in, d_out, tw, sign, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(tw);
    return e == hipSuccess ? 0 : CUVHS_E_HIP;
}

}  // extern "C"
