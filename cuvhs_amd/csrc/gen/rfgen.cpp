// Synthetic VHS RF generator (u8 captures), NTSC or PAL.
//
// Produces deterministic captures for tests and benchmarks: NTSC 525/59.94
// sync structure (6 EQ, 6 broad, 6 EQ pulses at half-line spacing, 4.7 us
// HSYNC) or PAL 625/50 (5 EQ, 5 broad, 5 EQ), active lines in bands of
// horizontal ramps and 75% colour bars, the luma FM-modulated on the
// system's VHS deviation (NTSC: sync tip 3.4 MHz, peak white 4.4 MHz; PAL:
// 3.8 and 4.8 MHz), the colour-under QAM (NTSC 40 fH, PAL 40 fH + 1953 Hz)
// carrying the VHS per-line +-90 degree rotation of an alternating head
// track, Gaussian noise at a stated SNR, RF-envelope dropouts, wow/flutter
// time warp and head-switch timing steps.  The line structure of the two
// fields follows what the decoder's field-order logic expects (the spacing
// between the last HSYNC and the first EQ pulse, and between the last EQ
// pulse and the first HSYNC: NTSC first field 1 and 0.5 lines, second 0.5
// and 1; PAL first field 0.5 and 0.5, second 1 and 1).
//
// Determinism: the output depends only on the parameters, never on the
// thread count.  Work is split into fixed 64 Ki-sample blocks; the FM phase
// at each block start comes from a prefix sum of per-block phase increments.
#include "rfgen.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

const double kPi = 3.14159265358979323846;
const int kBlock = 1 << 16;

// The line structure and levels of a system.
struct Sys {
    double line_rate;     // Hz
    double H;             // line period
    double field_lines;   // 262.5 / 312.5
    double fsc_gen;       // line-locked subcarrier: 227.5 / 283.75 fH
    double tbc_len;       // TBC samples per line: 4 fsc_gen / fH
    double hz_ire, ire0, sync_ire;
    double cu_per_sample; // colour-under cycles per TBC sample (decoder's subcarrier)
    double vblank_end;    // lines of half-line pulses from the first pre-EQ: 9 / 7.5
    double broad0, broad1;// broad pulses between these positions (lines)
    double eq_w, broad_w; // pulse widths (s)
    double burst0, burst1;// burst gate (s from the HSYNC edge)
    double act0, act1;    // active picture (s)
    double act_w;         // its width (s)
    int first_active;     // first line of the picture bands
    bool pal;
};

Sys make_sys(int system) {
    const double us = 1e-6;
    Sys y;
    y.pal = system == 1;
    if (y.pal) {
        y.line_rate = 15625.0;
        y.field_lines = 312.5;
        y.fsc_gen = 283.75 * y.line_rate;
        y.tbc_len = 1135.0;
        y.sync_ire = -0.3 * (100.0 / 0.7);
        y.hz_ire = 1e6 / (100.0 - y.sync_ire);
        y.ire0 = 4.8e6 - y.hz_ire * 100.0;
        y.cu_per_sample = (625.0 * 25.0 * 40.0 + 1953.0) / (4.0 * 4433618.75);
        y.vblank_end = 7.5;
        y.broad0 = 2.5;
        y.broad1 = 5.0;
        y.eq_w = 2.35 * us;
        y.broad_w = 27.3 * us;
        y.burst0 = 5.6 * us;
        y.burst1 = 7.85 * us;
        y.act0 = 10.5 * us;
        y.act1 = 62.5 * us;
        y.act_w = 52.0 * us;
        y.first_active = 23;
    } else {
        y.line_rate = 15734.264;
        y.field_lines = 262.5;
        y.fsc_gen = 227.5 * y.line_rate;
        y.tbc_len = 910.0;
        y.sync_ire = -40.0;
        y.hz_ire = 1e6 / 140.0;
        y.ire0 = 4.4e6 - y.hz_ire * 100.0;
        y.cu_per_sample = 629370.6293706293 / (4.0 * 3579545.0);
        y.vblank_end = 9.0;
        y.broad0 = 3.0;
        y.broad1 = 6.0;
        y.eq_w = 2.3 * us;
        y.broad_w = 27.1 * us;
        y.burst0 = 5.3 * us;
        y.burst1 = 7.8 * us;
        y.act0 = 10.9 * us;
        y.act1 = 62.0 * us;
        y.act_w = 51.1 * us;
        y.first_active = 20;
    }
    y.H = 1.0 / y.line_rate;
    return y;
}

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {  // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double gauss() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
    }
    int poisson(double lam) {
        double L = exp(-lam), p = 1.0;
        int k = 0;
        do { k++; p *= uniform(); } while (p > L);
        return k - 1;
    }
};

uint64_t mix(uint64_t a, uint64_t b) {
    Rng r(a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull));
    r.next();
    return r.next();
}

struct Dropout { double t0, t1; };

struct FieldParams {
    double snr_db = 35.0;
    double hs_offset = 0.0;  // head-switch timing offset (s)
    std::vector<Dropout> dropouts;
};

struct Model {
    cuvhs_gen_params p;
    Sys y;
    double fs;
    double t_lead;       // ideal time of field 0's first pre-EQ pulse
    int nfield;          // fields with parameters: index k+1 for k in [-1, nfield-2]
    std::vector<FieldParams> fp;
    double wow_a = 0, wow_f = 0, wow_ph = 0;
    double lsb_carrier = 90.0;
    double lsb_per_chroma_ire = 0.0;

    const FieldParams& field(long k) const {
        long i = k + 1;
        if (i < 0) i = 0;
        if (i >= (long)fp.size()) i = (long)fp.size() - 1;
        return fp[(size_t)i];
    }
};

// smoothed unit step, raised-cosine over width w centred on 0
inline double edge(double x, double w) {
    if (x <= -0.5 * w) return 0.0;
    if (x >= 0.5 * w) return 1.0;
    return 0.5 - 0.5 * cos(kPi * (x + 0.5 * w) / w);
}

inline double pulse(double x, double a, double b, double w) { return edge(x - a, w) - edge(x - b, w); }

// 75% colour bars: luma IRE, chroma amplitude (IRE, peak) and phase (deg from +U)
const double kBarY[7] = {77.0, 69.0, 56.0, 48.0, 36.0, 28.0, 15.0};
const double kBarA[7] = {0.0, 31.0, 44.0, 41.0, 41.0, 44.0, 31.0};
const double kBarP[7] = {0.0, 167.1, 283.5, 240.7, 60.7, 103.5, 347.1};

struct Sample {
    double ire;        // luma level
    double cu;         // colour-under waveform (IRE units of chroma amplitude)
};

// Signal at ideal time tau (already warped); field bookkeeping included.
Sample eval(const Model& m, double tau) {
    const Sys& y = m.y;
    const double kH = y.H;
    const double fl = y.field_lines * kH;
    long k = (long)floor((tau - m.t_lead) / fl);
    const FieldParams& f = m.field(k);
    double t = tau - f.hs_offset;
    k = (long)floor((t - m.t_lead) / fl);
    double fstart = m.t_lead + (double)k * fl;
    double v = (t - fstart) / kH;  // position in lines from the first pre-EQ
    bool type1 = (k % 2 + 2) % 2 == 0;   // the first field of the frame
    // HSYNCs on whole lines from the first pre-EQ pulse (NTSC first field,
    // PAL second field) or half a line later
    const bool whole = type1 != y.pal;
    const double us = 1e-6;
    const double w = 0.14 * us;    // edge width
    double ire = 0.0;              // blanking
    double ca = 0.0, cph = 0.0;    // chroma amplitude (IRE) and phase (rad, from +U)
    if (v < y.vblank_end) {
        // vertical interval: half-line spaced pulses
        double hv = floor(v * 2.0) * 0.5;
        double x = (v - hv) * kH;  // seconds into the half line
        bool broad = (v >= y.broad0 && v < y.broad1);
        double pw = broad ? y.broad_w : y.eq_w;
        ire = y.sync_ire * pulse(x, 0.0, pw, w);
    } else {
        double line_origin = whole ? floor(v) : floor(v - 0.5) + 0.5;
        double x = (v - line_origin) * kH;
        int line_in_field = (int)floor(line_origin);
        ire = y.sync_ire * pulse(x, 0.0, 4.7 * us, w);
        // burst along -U (180 degrees)
        double bg = pulse(x, y.burst0, y.burst1, 0.3 * us);
        ca = 20.0 * bg;
        cph = kPi;
        if (line_in_field >= y.first_active && x > y.act0 && x < y.act1) {
            double pos = (x - y.act0) / y.act_w;   // 0..1 across active video
            double gate = pulse(x, y.act0, y.act1, w);
            int band = (line_in_field / 16) & 1;
            if (band == 0) {
                ire += gate * 100.0 * pos;
            } else {
                int bar = std::min(6, (int)(pos * 7.0));
                double bx = pos * 7.0 - (double)bar;
                // soften bar edges in luma
                double yb = kBarY[bar];
                if (bar > 0 && bx < 0.02) yb = kBarY[bar - 1] + (kBarY[bar] - kBarY[bar - 1]) * (0.5 + bx / 0.04);
                ire += gate * yb;
                ca = gate * kBarA[bar];
                cph = kBarP[bar] * kPi / 180.0;
            }
        }
    }
    Sample s;
    s.ire = ire;
    // colour-under (decoder line grid: line 0 starts one entry spacing, 1 H
    // or 0.5 H, before the first pre-EQ pulse)
    double tau0 = fstart - (whole ? 1.0 : 0.5) * kH;
    double u = 4.0 * y.fsc_gen * (t - tau0);     // TBC-sample coordinate
    long r = (long)floor(u / y.tbc_len);
    int track = (int)((k % 2 + 2) % 2);
    int rot = track ? 1 : 3;
    // PAL's 1135-sample lines advance the subcarrier by 3/4 cycle per line
    // (NTSC's 910: 1/2); a further quarter cycle per line on PAL keeps the
    // burst alternating by half a cycle from line to line, the relation the
    // decoder's track detection measures (chroma_decode.cu:383-411)
    const long pal_q = y.pal ? r : 0;
    double phi_r = (double)((((r * rot) + pal_q) % 4 + 4) % 4) * (kPi * 0.5);
    double psi = cph + 2.0 * kPi * fmod(y.fsc_gen * tau0, 1.0);
    double ucyc = fmod(y.cu_per_sample * u, 1.0);
    s.cu = ca * cos(2.0 * kPi * ucyc + phi_r + kPi - psi);
    return s;
}

double tau_of(const Model& m, size_t s) {
    double t = (double)s / m.fs;
    if (m.wow_a != 0.0) t -= m.wow_a * sin(2.0 * kPi * m.wow_f * t + m.wow_ph);
    return t;
}

double gain_at(const Model& m, double tau) {
    const double fl = m.y.field_lines * m.y.H;
    long k = (long)floor((tau - m.t_lead) / fl);
    double g = 1.0;
    for (long kk = k - 1; kk <= k; kk++) {
        const FieldParams& f = m.field(kk);
        for (const Dropout& d : f.dropouts) {
            if (tau < d.t0 - 0.25e-6 || tau > d.t1 + 0.25e-6) continue;
            double inside = edge(tau - d.t0, 0.5e-6) - edge(tau - d.t1, 0.5e-6);
            g = std::min(g, 1.0 - 0.95 * inside);
        }
    }
    return g;
}

Model build_model(const cuvhs_gen_params* p) {
    Model m;
    m.p = *p;
    m.y = make_sys(p->system);
    const double kH = m.y.H, kFieldLines = m.y.field_lines;
    m.fs = p->sample_rate_mhz * 1e6;
    m.t_lead = p->lead_lines * kH;
    double dur = (double)p->n_samples / m.fs;
    m.nfield = (int)ceil((dur - m.t_lead) / (kFieldLines * kH)) + 3;
    m.fp.resize((size_t)m.nfield + 1);
    Rng cap(mix(p->seed, 0xC0FFEEull));
    if (p->wow_max_us > 0.0) {
        m.wow_a = cap.uniform() * p->wow_max_us * 1e-6;
        m.wow_f = p->wow_f_lo_hz + cap.uniform() * (p->wow_f_hi_hz - p->wow_f_lo_hz);
        m.wow_ph = cap.uniform() * 2.0 * kPi;
    }
    m.lsb_per_chroma_ire = m.lsb_carrier * pow(10.0, p->chroma_db / 20.0) / 40.0;
    // head-switch groups
    double hs = 0.0;
    int left = 0;
    Rng hsr(mix(p->seed, 0x4853ull));
    for (size_t i = 0; i < m.fp.size(); i++) {
        if (p->headswitch_max_us > 0.0) {
            if (left == 0) {
                left = 1 + (int)(hsr.uniform() * std::max(1, p->headswitch_group_max));
                hs = hsr.uniform() * p->headswitch_max_us * 1e-6;
            }
            left--;
        }
        FieldParams& f = m.fp[i];
        f.hs_offset = hs;
        Rng fr(mix(p->seed, 0x1000ull + i));
        f.snr_db = p->snr_lo_db + fr.uniform() * (p->snr_hi_db - p->snr_lo_db);
        if (p->dropouts_per_field > 0.0) {
            int nd = fr.poisson(p->dropouts_per_field);
            long k = (long)i - 1;
            double fstart = m.t_lead + (double)k * kFieldLines * kH;
            for (int j = 0; j < nd; j++) {
                double t0 = fstart + fr.uniform() * kFieldLines * kH;
                double len = (p->dropout_len_lo_us + fr.uniform() * (p->dropout_len_hi_us - p->dropout_len_lo_us)) * 1e-6;
                f.dropouts.push_back({t0, t0 + len});
            }
        }
    }
    return m;
}

// RBJ high-shelf biquad used as the recording pre-emphasis: the inverse of
// the decoder's de-emphasis (273,755.82 Hz, 13.9794 dB, Q 0.462088186).
void preemph_coeffs(double fs, double b[3], double a[3]) {
    const double f0 = 273755.82, dbgain = 13.9794, Q = 0.462088186;
    double A = pow(10.0, dbgain / 40.0);
    double w0 = 2.0 * kPi * f0 / fs;
    double alpha = sin(w0) / (2.0 * Q);
    double c = cos(w0), sA = sqrt(A);
    b[0] = A * ((A + 1) + (A - 1) * c + 2 * sA * alpha);
    b[1] = -2 * A * ((A - 1) + (A + 1) * c);
    b[2] = A * ((A + 1) + (A - 1) * c - 2 * sA * alpha);
    a[0] = (A + 1) - (A - 1) * c + 2 * sA * alpha;
    a[1] = 2 * ((A - 1) - (A + 1) * c);
    a[2] = (A + 1) - (A - 1) * c - 2 * sA * alpha;
}

const int kWarm = 768;  // IIR warm-up samples before each block

// Luma level -> 5-tap binomial smoothing -> pre-emphasis -> FM; plus the
// colour-under term, RF gain (dropouts) and noise.
void gen_block(const Model& m, size_t b, double phase0, uint8_t* out, double* dphase_out) {
    size_t s0 = b * (size_t)kBlock;
    size_t s1 = std::min(m.p.n_samples, s0 + (size_t)kBlock);
    if (s0 >= s1) { if (dphase_out) *dphase_out = 0.0; return; }
    size_t n = s1 - s0;
    size_t nh = n + kWarm;
    std::vector<double> ire(nh + 4), cu(n), tau(n + 1), y(nh);
    for (size_t i = 0; i < nh + 4; i++) {
        long s = (long)s0 - kWarm + (long)i - 2;
        double t = tau_of(m, s < 0 ? 0 : (size_t)s);
        if (s < 0) t += (double)s / m.fs;
        Sample sm = eval(m, t);
        ire[i] = sm.ire;
        if (i >= (size_t)kWarm + 2 && i < (size_t)kWarm + n + 2) cu[i - kWarm - 2] = sm.cu;
    }
    double bq[3], aq[3];
    preemph_coeffs(m.fs, bq, aq);
    double x1 = 0, x2 = 0, y1 = 0, y2 = 0;
    for (size_t i = 0; i < nh; i++) {
        double x = (ire[i] + 4.0 * ire[i + 1] + 6.0 * ire[i + 2] + 4.0 * ire[i + 3] + ire[i + 4]) / 16.0;
        if (i == 0) { x1 = x2 = y1 = y2 = x; }
        double v = (bq[0] * x + bq[1] * x1 + bq[2] * x2 - aq[1] * y1 - aq[2] * y2) / aq[0];
        x2 = x1; x1 = x; y2 = y1; y1 = v;
        y[i] = v;
    }
    for (size_t i = 0; i <= n; i++) tau[i] = tau_of(m, s0 + i);
    double ph = phase0;
    Rng nr(mix(m.p.seed, 0xABCDEF00ull + b));
    for (size_t i = 0; i < n; i++) {
        double f_inst = m.y.ire0 + m.y.hz_ire * y[i + kWarm];
        if (out) {
            double t = tau[i];
            double g = gain_at(m, t);
            const double fl = m.y.field_lines * m.y.H;
            long k = (long)floor((t - m.t_lead) / fl);
            double snr = m.field(k).snr_db;
            double sigma = sqrt(m.lsb_carrier * m.lsb_carrier * 0.5 / pow(10.0, snr / 10.0));
            double x = 128.0 + g * (m.lsb_carrier * cos(ph) + m.lsb_per_chroma_ire * cu[i]) + sigma * nr.gauss();
            double r = floor(x + 0.5);
            if (r < 0.0) r = 0.0;
            if (r > 255.0) r = 255.0;
            out[i] = (uint8_t)r;
        }
        ph += 2.0 * kPi * f_inst * (tau[i + 1] - tau[i]);
        ph = fmod(ph, 2.0 * kPi);
    }
    if (dphase_out) *dphase_out = ph - phase0;
}

}  // namespace

extern "C" {

void cuvhs_gen_defaults_system(cuvhs_gen_params* p, int config, int system) {
    memset(p, 0, sizeof(*p));
    p->system = system == 1 ? 1 : 0;
    p->sample_rate_mhz = 28.0;
    p->lead_lines = 20.0;
    p->snr_lo_db = 35.0;
    p->snr_hi_db = 35.0;
    p->chroma_db = -12.0;
    p->dropout_len_lo_us = 5.0;
    p->dropout_len_hi_us = 50.0;
    p->wow_f_lo_hz = 0.5;
    p->wow_f_hi_hz = 10.0;
    p->headswitch_group_max = 8;
    double spf_lines = p->system ? 313.0 : 263.0;
    int nfields = 4;
    if (config == 0) {
        p->seed = 0;
    } else if (config == 1 || config == 2) {
        p->seed = (uint64_t)config;
        p->snr_lo_db = 25.0;
        p->snr_hi_db = 40.0;
        p->dropouts_per_field = 2.0;
        p->wow_max_us = 2.0;
        nfields = 600;
        if (config == 2) {
            p->sample_rate_mhz = 40.0;
            p->headswitch_max_us = 4.0;
            nfields = 6000;
        }
    }
    int spl = (int)round(p->sample_rate_mhz * 1e6 / make_sys(p->system).line_rate);
    p->n_samples = (size_t)nfields * (size_t)(spl * (int)spf_lines);
}

void cuvhs_gen_defaults(cuvhs_gen_params* p, int config) { cuvhs_gen_defaults_system(p, config, 0); }

int cuvhs_gen(const cuvhs_gen_params* p, uint8_t* out) {
    if (!p || !out || !(p->sample_rate_mhz > 0.0) || p->n_samples == 0 || p->system < 0 || p->system > 1) return -1;
    Model m = build_model(p);
    size_t nb = (p->n_samples + kBlock - 1) / kBlock;
    std::vector<double> dph(nb, 0.0);
#pragma omp parallel for schedule(dynamic, 4)
    for (long b = 0; b < (long)nb; b++) gen_block(m, (size_t)b, 0.0, nullptr, &dph[(size_t)b]);
    std::vector<double> start(nb, 0.0);
    double acc = 0.0;
    for (size_t b = 0; b < nb; b++) {
        start[b] = acc;
        acc = fmod(acc + dph[b], 2.0 * kPi);
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (long b = 0; b < (long)nb; b++)
        gen_block(m, (size_t)b, start[(size_t)b], out + (size_t)b * kBlock, nullptr);
    return 0;
}

}  // extern "C"
