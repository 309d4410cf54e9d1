// CPU FP64 oracle: statement-by-statement restatement of the reference's
// decode path.  TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Every function cites the reference file:line it follows.  Where the
// reference is nondeterministic or reads memory it never wrote, the oracle
// fixes one well-defined behaviour and says so at that point:
//   * VSYNC candidates keep the first `cap` in ascending sample order
//     (reference: atomic order, vsync_discover.cu:38).
//   * demod/demod_05/envelope beyond the region K1 computed in this chunk are
//     zero (reference: stale or uninitialised device memory).
//   * dropout concealment applies entries in order (reference: racy parallel
//     copy, dropout_detect.cu:274-290).
//   * per-field mean envelope and burst sums are summed sequentially
//     (reference: parallel reductions in an unspecified order).
//
// How each part is pinned (tests/): format, convert and writer against the
// reference's own host sources built in oracle/_ref; K2b line location
// (pipeline.cu:133-148, 410-893: validated pulses, good flags, anchors, line
// locations, is_first) and the chroma burst metric, burst phase and NTSC
// phase table (chroma_decode.cu:329-450) bit for bit against the reference's
// own host functions, compiled as they lie (oracle/ref_k2b_driver.cpp), on
// seven degraded NTSC captures and four PAL captures, the PAL format and
// writer included (tests/test_reference_k2b.py, tests/test_reference_host.py:
// system 1 runs the reference's PAL branches, video_format.cpp:48-87,
// pipeline.cu:707-717, 811); K3 (NTSC and PAL parity rules,
// line_locs.cu:219-235) on ideal pulse trains of both field orders
// (tests/test_oracle_k3.py); FFT against numpy
// and brute-force DFTs; filter design against scipy.signal; demod, VSYNC, K2,
// K4, K5 by closed forms and brute force; K7 by RF dips placed at known
// samples.  Pinned by invariants: the VSYNC clustering (field offsets one
// field period apart), the phase-ID cycle (3, 2, 1, 4 field after field,
// across chunks), the chroma retry (track re-acquired after a field cut out
// of the tape).
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fft.h"

#ifdef _OPENMP
#include <omp.h>
#endif

using orc::cd;
using orc::FFTPlan;

namespace {

const double PI = 3.14159265358979323846;  // fm_demod.cu:13
const double TAU = 2.0 * PI;
const int OVERLAP = 2048;                  // fm_demod.cu:17
const int FIELD_MARGIN_LINES = 20;         // pipeline.cu:1339
const int MAX_PULSES = OR_MAX_PULSES;
const int MAX_DO = OR_MAX_DROPOUTS;
enum { PULSE_HSYNC = 0, PULSE_EQ1 = 1, PULSE_VSYNC = 2, PULSE_EQ2 = 3 };  // line_locs.h:6

// ---------------------------------------------------------------------------
// VideoFormat (video_format.cpp:23-104), NTSC (system 0) or PAL (system 1)
// ---------------------------------------------------------------------------
struct Fmt {
    int system;
    double sample_rate;
    int lines_per_frame, lines_per_field;
    double line_rate, fsc, chroma_under;
    double burst_abs_ref, burst_start_us, burst_end_us;
    int output_line_len, output_field_lines, num_eq_pulses, field_lines_first, field_lines_second;
    double hz_ire, ire0, vsync_ire, sync_tip_hz, pulse_threshold_hz;
    double output_zero, output_scale;
    int active_line_start;
    double output_rate;
    int samples_per_line, samples_per_field;
    double hsync_width, eq_pulse_width, bp_width, vsync_width;
    // derived (pipeline.cu:1342, fm_demod.cu:331)
    int spfp;
};

Fmt make_fmt(double mhz, int system = 0) {
    Fmt f;
    f.system = system;
    f.sample_rate = mhz * 1e6;
    if (system == 1) {
        // PAL branch (video_format.cpp:48-70, :83-87)
        f.lines_per_frame = 625;
        f.lines_per_field = 313;
        f.line_rate = 15625.0;
        f.fsc = 4433618.75;
        f.chroma_under = ((625.0 * 25.0) * 40.0) + 1953.0;
        f.burst_abs_ref = 5000.0;
        f.burst_start_us = 5.6;
        f.burst_end_us = 7.85;
        f.output_line_len = 1135;
        f.output_field_lines = 313;
        f.num_eq_pulses = 5;
        f.field_lines_first = 312;
        f.field_lines_second = 313;
        f.vsync_ire = -0.3 * (100.0 / 0.7);
        f.hz_ire = 1e6 / (100.0 + (-f.vsync_ire));
        f.ire0 = 4.8e6 - (f.hz_ire * 100.0);
    } else {
        f.lines_per_frame = 525;
        f.lines_per_field = 263;
        f.line_rate = 15734.264;
        f.fsc = 3579545.0;
        f.chroma_under = (525.0 * (30.0 / 1.001)) * 40.0;
        f.burst_abs_ref = 4416.0;
        f.burst_start_us = 5.3;
        f.burst_end_us = 7.8;
        f.output_line_len = 910;
        f.output_field_lines = 263;
        f.num_eq_pulses = 6;
        f.field_lines_first = 263;
        f.field_lines_second = 262;
        f.hz_ire = 1e6 / 140.0;
        f.ire0 = 4.4e6 - (f.hz_ire * 100.0);
        f.vsync_ire = -40.0;
    }
    f.sync_tip_hz = f.ire0 + f.hz_ire * f.vsync_ire;
    f.pulse_threshold_hz = f.ire0 + f.hz_ire * (-20.0);
    if (system == 1) {
        f.output_zero = 256.0;
        f.output_scale = (54016.0 - 256.0) / (100.0 - f.vsync_ire);
        f.active_line_start = 8;
    } else {
        f.output_zero = 1024.0;
        f.output_scale = (51200.0 - 1024.0) / (100.0 - f.vsync_ire);
        f.active_line_start = 10;
    }
    f.output_rate = 4.0 * f.fsc;
    f.samples_per_line = static_cast<int>(round(f.sample_rate / f.line_rate));
    f.samples_per_field = f.samples_per_line * f.lines_per_field;
    f.hsync_width = 4.7e-6 * f.sample_rate;
    f.eq_pulse_width = 2.3e-6 * f.sample_rate;
    f.bp_width = 1.5e-6 * f.sample_rate;
    f.vsync_width = (system == 1 ? 27.3e-6 : 27.1e-6) * f.sample_rate;
    f.spfp = f.samples_per_field + FIELD_MARGIN_LINES * f.samples_per_line;
    return f;
}

// active output columns of the K5 counters and the JSON (tbc_resample.cu:156-157,
// tbc_writer.cpp:106-107)
int active_video_start(const Fmt& f) { return f.system == 1 ? 185 : 134; }
int active_video_end(const Fmt& f) { return f.system == 1 ? 1107 : 894; }

// ---------------------------------------------------------------------------
// K1 filter design helpers (fm_demod.cu:35-127)
// ---------------------------------------------------------------------------
double butter_lpf_mag(double omega, double omega_c, int order) {
    if (omega_c <= 0.0 || omega_c >= PI) return 1.0;
    double tc = tan(omega_c / 2.0);
    if (tc == 0.0) return 1.0;
    double t = tan(omega / 2.0);
    double ratio = t / tc;
    return 1.0 / sqrt(1.0 + pow(ratio, 2.0 * order));
}

double butter_hpf_mag(double omega, double omega_c, int order) {
    if (omega <= 0.0) return 0.0;
    if (omega_c <= 0.0 || omega_c >= PI) return 1.0;
    double tc = tan(omega_c / 2.0);
    double t = tan(omega / 2.0);
    if (t == 0.0) return 0.0;
    double ratio = tc / t;
    return 1.0 / sqrt(1.0 + pow(ratio, 2.0 * order));
}

double butter_bpf_mag(double omega, double omega_l, double omega_h, int order) {
    if (omega <= 0.0 || omega >= PI) return 0.0;
    double tl = tan(omega_l / 2.0);
    double th = tan(omega_h / 2.0);
    double t = tan(omega / 2.0);
    double BW = th - tl;
    if (BW <= 0.0 || t == 0.0) return 0.0;
    double S = (t * t - tl * th) / (BW * t);
    return 1.0 / sqrt(1.0 + pow(S, 2.0 * order));
}

double supergauss_mag(double f_hz, double corner_freq, int order) {
    double ln2_half = log(2.0) / 2.0;
    double scale = pow(ln2_half, 1.0 / (2.0 * order));
    double x = 2.0 * f_hz * scale / corner_freq;
    return exp(-2.0 * pow(x, 2.0 * order));
}

void gen_shelf_high(double f0, double dbgain, double fs, double Q, double b[3], double a[3]) {
    double A = pow(10.0, dbgain / 40.0);
    double w0 = TAU * f0 / fs;
    double alpha = sin(w0) / (2.0 * Q);
    double cosw0 = cos(w0);
    double sqA = sqrt(A);
    b[0] = A * ((A + 1) + (A - 1) * cosw0 + 2 * sqA * alpha);
    b[1] = -2 * A * ((A - 1) + (A + 1) * cosw0);
    b[2] = A * ((A + 1) + (A - 1) * cosw0 - 2 * sqA * alpha);
    a[0] = (A + 1) - (A - 1) * cosw0 + 2 * sqA * alpha;
    a[1] = 2 * ((A - 1) - (A + 1) * cosw0);
    a[2] = (A + 1) - (A - 1) * cosw0 - 2 * sqA * alpha;
}

cd freqz_biquad(const double num[3], const double den[3], double omega) {
    cd z1 = std::exp(cd(0.0, -omega));
    cd z2 = z1 * z1;
    cd N = num[0] + num[1] * z1 + num[2] * z2;
    cd D = den[0] + den[1] * z1 + den[2] * z2;
    return N / D;
}

std::vector<double> firwin_lpf(int num_taps, double cutoff_normalized) {
    std::vector<double> h((size_t)num_taps);
    double alpha = (num_taps - 1) / 2.0;
    double sum = 0.0;
    for (int n = 0; n < num_taps; n++) {
        double m = n - alpha;
        double sinc_val;
        if (fabs(m) < 1e-12)
            sinc_val = cutoff_normalized;
        else
            sinc_val = sin(PI * cutoff_normalized * m) / (PI * m);
        double window = 0.54 - 0.46 * cos(TAU * n / (num_taps - 1));
        h[(size_t)n] = sinc_val * window;
        sum += h[(size_t)n];
    }
    for (int n = 0; n < num_taps; n++) h[(size_t)n] /= sum;
    return h;
}

cd fir_freqz(const std::vector<double>& h, double omega, int delay_compensate) {
    cd H(0.0, 0.0);
    for (size_t n = 0; n < h.size(); n++) H += h[n] * std::exp(cd(0.0, -(double)n * omega));
    if (delay_compensate > 0) H *= std::exp(cd(0.0, omega * delay_compensate));
    return H;
}

struct K1Filters {
    int N = 0, fb = 0;
    std::vector<double> rf;
    std::vector<cd> fv, fv05;
};

// FMDemodState::init (fm_demod.cu:273-374), with spf = spfp (pipeline.cu:1455)
K1Filters make_k1_filters(const Fmt& fmt) {
    K1Filters K;
    K.N = orc::next_7smooth(fmt.spfp + OVERLAP);
    K.fb = K.N / 2 + 1;
    double fs = fmt.sample_rate;
    double nyquist = fs / 2.0;
    double w_bpf_lo = PI * 1300000.0 / nyquist;
    double w_bpf_hi = PI * 7500000.0 / nyquist;
    double w_lpf = PI * 8000000.0 / nyquist;
    double w_hpf = PI * 1000000.0 / nyquist;
    K.rf.resize((size_t)K.fb);
    for (int k = 0; k < K.fb; k++) {
        double omega = PI * (double)k / (double)(K.fb - 1);
        double bpf = butter_bpf_mag(omega, w_bpf_lo, w_bpf_hi, 8);
        double lpf = butter_lpf_mag(omega, w_lpf, 8);
        double hpf = butter_hpf_mag(omega, w_hpf, 8);
        double hil = (k == 0 || k == K.fb - 1) ? 1.0 : 2.0;
        K.rf[(size_t)k] = bpf * lpf * hpf * hil;
    }
    double b_shelf[3], a_shelf[3];
    gen_shelf_high(273755.82, 13.9794, fs, 0.462088186, b_shelf, a_shelf);
    double sync_cutoff_norm = 500000.0 / nyquist;
    std::vector<double> sync_fir = firwin_lpf(65, sync_cutoff_norm);
    const int f05_offset = 32;
    K.fv.resize((size_t)K.fb);
    K.fv05.resize((size_t)K.fb);
    double inv_n = 1.0 / (double)K.N;
    for (int k = 0; k < K.fb; k++) {
        double omega = PI * (double)k / (double)(K.fb - 1);
        double f_hz = (double)k * nyquist / (double)(K.fb - 1);
        cd deemp = freqz_biquad(a_shelf, b_shelf, omega);
        double vlpf = supergauss_mag(f_hz, 3000000.0, 9);
        cd sync = fir_freqz(sync_fir, omega, f05_offset);
        K.fv[(size_t)k] = deemp * vlpf * inv_n;
        K.fv05[(size_t)k] = deemp * vlpf * sync * inv_n;
    }
    return K;
}

// chroma_decode.cu:568-587 — pre-BPF |H|^2 / N at the capture rate
std::vector<double> make_chroma_prebpf(const Fmt& fmt, int field_stride, int* n_out) {
    int raw_fft_size = orc::next_7smooth(field_stride);
    int raw_freq_bins = raw_fft_size / 2 + 1;
    double omega_lo = 2.0 * M_PI * 60000.0 / fmt.sample_rate;
    double omega_hi = 2.0 * M_PI * 1200000.0 / fmt.sample_rate;
    double raw_fft_norm = 1.0 / (double)raw_fft_size;
    std::vector<double> h((size_t)raw_freq_bins);
    for (int i = 0; i < raw_freq_bins; i++) {
        double omega = M_PI * (double)i / (double)(raw_freq_bins - 1);
        double mag = butter_bpf_mag(omega, omega_lo, omega_hi, 4);
        h[(size_t)i] = mag * mag * raw_fft_norm;
    }
    *n_out = raw_fft_size;
    return h;
}

// ---------------------------------------------------------------------------
// K1 fm_demod (fm_demod.cu:396-486 with kernels :133-257)
// raw: nfields*spfp samples; outputs likewise
// ---------------------------------------------------------------------------
void fm_demod(const Fmt& fmt, const K1Filters& K, const double* raw, int nf, double* demod,
              double* demod05, double* env) {
    const int spf = fmt.spfp;
    const int N = K.N, fb = K.fb;
    const long long total_src = (long long)nf * spf;
    FFTPlan plan(N);
    const double scale = fmt.sample_rate / TAU;
    std::vector<double> rawdemod((size_t)total_src);

    // k_copy_and_pad (fm_demod.cu:133-153)
    auto window = [&](const double* src, int f, std::vector<double>& w) {
        for (int p = 0; p < N; p++) {
            long long s = (long long)f * spf + p - OVERLAP;
            if (s < 0) w[(size_t)p] = src[0];
            else if (s < total_src) w[(size_t)p] = src[s];
            else w[(size_t)p] = src[total_src - 1];
        }
    };

    // Part A
#pragma omp parallel for schedule(dynamic, 1)
    for (int f = 0; f < nf; f++) {
        std::vector<double> w((size_t)N);
        std::vector<cd> half((size_t)fb), full((size_t)N), a((size_t)N);
        std::vector<double> ang((size_t)N);
        window(raw, f, w);
        orc::fft_r2c(plan, w.data(), half.data());
        for (int k = 0; k < fb; k++) {  // k_apply_rf_filter (:224-233)
            double rf = K.rf[(size_t)k];
            half[(size_t)k] = cd(half[(size_t)k].real() * rf, half[(size_t)k].imag() * rf);
        }
        for (int k = 0; k < N; k++) full[(size_t)k] = k < fb ? half[(size_t)k] : cd(0.0, 0.0);  // :235-249
        plan.exec(full.data(), a.data(), +1);  // Z2Z inverse, unnormalised
        for (int i = 0; i < spf; i++) {        // k_compute_envelope (:259-273)
            double re = a[(size_t)i + OVERLAP].real(), im = a[(size_t)i + OVERLAP].imag();
            env[(size_t)f * spf + i] = sqrt(re * re + im * im);
        }
        for (int p = 0; p < N; p++) ang[(size_t)p] = atan2(a[(size_t)p].imag(), a[(size_t)p].real());
        for (int i = 0; i < spf; i++) {        // k_unwrap_to_hz (:275-297)
            int ai = OVERLAP + i;
            double delta = ang[(size_t)ai] - ang[(size_t)ai - 1];
            if (delta > PI) delta -= TAU;
            if (delta < -PI) delta += TAU;
            rawdemod[(size_t)f * spf + i] = delta * scale;
        }
    }

    // Part B
#pragma omp parallel for schedule(dynamic, 1)
    for (int f = 0; f < nf; f++) {
        std::vector<double> w((size_t)N), y((size_t)N);
        std::vector<cd> X((size_t)fb), Y((size_t)fb);
        window(rawdemod.data(), f, w);
        orc::fft_r2c(plan, w.data(), X.data());
        for (int pass = 0; pass < 2; pass++) {
            const std::vector<cd>& F = pass == 0 ? K.fv : K.fv05;
            for (int k = 0; k < fb; k++) {  // k_apply_complex_filter (:299-310)
                cd v = X[(size_t)k], g = F[(size_t)k];
                Y[(size_t)k] = cd(v.real() * g.real() - v.imag() * g.imag(),
                                  v.real() * g.imag() + v.imag() * g.real());
            }
            orc::fft_c2r(plan, Y.data(), y.data());
            double* dst = pass == 0 ? demod : demod05;
            for (int i = 0; i < spf; i++) {  // k_trim (:208-222)
                int s = i + OVERLAP;
                if (s >= N) s = N - 1;
                dst[(size_t)f * spf + i] = y[(size_t)s];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// VSYNC discovery (vsync_discover.cu:5-72)
// ---------------------------------------------------------------------------
int vsync_discover(const Fmt& fmt, const double* x, int total, std::vector<int>& out, int cap) {
    double thr = fmt.pulse_threshold_hz;
    int min_w = (int)(20.0e-6 * fmt.sample_rate);
    int max_w = (int)(35.0e-6 * fmt.sample_rate);
    int count = 0;
    out.clear();
    for (int idx = 0; idx < total; idx++) {
        bool is_start = idx == 0 ? (x[idx] <= thr) : (x[idx] <= thr && x[idx - 1] > thr);
        if (!is_start) continue;
        int width = 0;
        while (idx + width < total && x[idx + width] <= thr && width <= max_w) width++;
        if (width >= min_w && width <= max_w) {
            if (count < cap) out.push_back(idx);
            count++;
        }
    }
    return count;
}

// ---------------------------------------------------------------------------
// K2 sync pulses (sync_pulses.cu:10-58)
// ---------------------------------------------------------------------------
void sync_pulses(const Fmt& fmt, const double* d05, const size_t* offsets, int nf, int* starts,
                 int* lens, int* count) {
    double thr = fmt.pulse_threshold_hz;
    int spf = fmt.spfp;
    for (int field = 0; field < nf; field++) {
        size_t off = offsets[field];
        const double* sig = d05 + off;
        int* st = starts + (size_t)field * MAX_PULSES;
        int* ln = lens + (size_t)field * MAX_PULSES;
        bool in_pulse = (sig[0] <= thr);
        int cur_start = 0, c = 0;
        for (int i = 0; i < spf; i++) {
            double val = sig[i];
            if (in_pulse) {
                if (val > thr) {
                    int length = i - cur_start;
                    if (length > 0 && cur_start != 0 && c < MAX_PULSES) {
                        st[c] = (int)(off + cur_start);
                        ln[c] = length;
                        c++;
                    }
                    in_pulse = false;
                }
            } else if (val <= thr) {
                cur_start = i;
                in_pulse = true;
            }
        }
        count[field] = c;
    }
}

// ---------------------------------------------------------------------------
// K2b host line location (pipeline.cu:133-148, 410-893)
// ---------------------------------------------------------------------------
struct HostPulse { int start = 0; int len = 0; };
struct HostValidPulse { int type = PULSE_HSYNC; HostPulse pulse; bool good = false; };
struct HostK2bAnchor {
    bool valid = false;
    double line0loc = 0.0, first_hsync_loc = 0.0, first_hsync_line = 0.0, next_field = 0.0;
    int first_field = -1, progressive_field = 0;
    double prev_hsync_diff = -1.0;
};
struct HostLineTimings {
    double hsync_min = 0, hsync_max = 0, eq_min = 0, eq_max = 0, vsync_min = 0, vsync_max = 0,
           hsync_nominal = 0;
};

bool host_pulse_qualitycheck(int prev_type, int pulse_type, int start_delta, int in_line_len) {
    double min_lines, max_lines;
    if (prev_type > 0 && pulse_type > 0) { min_lines = 0.4; max_lines = 0.6; }
    else if (prev_type == 0 && pulse_type == 0) { min_lines = 0.9; max_lines = 1.1; }
    else { min_lines = 0.4; max_lines = 1.1; }
    double line_span = (double)start_delta / (double)in_line_len;
    return line_span >= min_lines && line_span <= max_lines;
}

bool inrange_host(double v, double lo, double hi) { return v >= lo && v <= hi; }

double round_nearest_line_loc_host(double line_number) {
    return std::round(0.5 * std::round(line_number / 0.5) * 10.0) / 10.0;
}

HostLineTimings host_get_line_timings(const int* lens, int count, const Fmt& fmt) {
    HostLineTimings lt;
    double tol = fmt.sample_rate * 0.5e-6;
    double gmin = fmt.hsync_width - 3.5 * tol;
    double gmax = fmt.hsync_width + 4.0 * tol;
    double sum = 0.0;
    int n = 0;
    int npc = std::max(0, std::min(count, MAX_PULSES));
    for (int i = 0; i < npc; i++) {
        double len = (double)lens[i];
        if (len >= gmin && len <= gmax) { sum += len; n++; }
    }
    double est = n > 0 ? (sum / n) : fmt.hsync_width;
    double off = est - fmt.hsync_width;
    lt.hsync_nominal = fmt.hsync_width + off;
    lt.hsync_min = lt.hsync_nominal - tol;
    lt.hsync_max = lt.hsync_nominal + tol;
    lt.eq_min = fmt.eq_pulse_width + off - tol;
    lt.eq_max = fmt.eq_pulse_width + off + tol;
    lt.vsync_min = (fmt.vsync_width + off) * 0.5;
    lt.vsync_max = fmt.vsync_width + off + 2.0 * tol;
    return lt;
}

double host_meanlinelen(const std::vector<HostValidPulse>& vp, double nominal) {
    int longrun_start = -1, longrun_len = -1, currun_start = -1, currun_len = 0;
    for (int i = 0; i < (int)vp.size(); i++) {
        if (vp[(size_t)i].type != PULSE_HSYNC) {
            if (currun_start >= 0 && currun_len > longrun_len) { longrun_start = currun_start; longrun_len = currun_len; }
            currun_start = -1;
            currun_len = 0;
        } else if (currun_start < 0) {
            currun_start = i;
            currun_len = 0;
        } else {
            currun_len++;
        }
    }
    if (currun_start >= 0 && currun_len > longrun_len) { longrun_start = currun_start; longrun_len = currun_len; }
    std::vector<double> linelens;
    if (longrun_start >= 1 && longrun_len >= 1) {
        for (int i = longrun_start + 1; i < longrun_start + longrun_len; i++) {
            double linelen = (double)(vp[(size_t)i].pulse.start - vp[(size_t)i - 1].pulse.start);
            double ratio = linelen / nominal;
            if (ratio >= 0.95 && ratio <= 1.05) linelens.push_back(linelen);
        }
    }
    if (linelens.empty()) return nominal;
    double sum = 0.0;
    for (double x : linelens) sum += x;
    return sum / (double)linelens.size();
}

bool host_run_vblank_state_machine(const std::vector<HostPulse>& raw, size_t start_idx,
                                   const HostLineTimings& lt, int num_pulses, int in_line_len,
                                   std::vector<HostValidPulse>& out) {
    bool done = false;
    double num_pulses_half = num_pulses / 2.0;
    int vsync_start = -1;
    double state_end = 0.0;
    double state_length = -1.0;
    for (size_t i = start_idx; i < raw.size() && i < start_idx + 26; i++) {
        const HostPulse& p = raw[i];
        HostValidPulse sp;
        bool have = false;
        int state = out.empty() ? -1 : out.back().type;
        double len = (double)p.len;
        if (state == -1) {
            if (inrange_host(len, lt.hsync_min, lt.hsync_max)) { sp = {PULSE_HSYNC, p, false}; have = true; }
        } else if (state == PULSE_HSYNC) {
            if (inrange_host(len, lt.hsync_min, lt.hsync_max)) { sp = {PULSE_HSYNC, p, false}; have = true; }
            else if (inrange_host(len, lt.eq_min, lt.eq_max)) { sp = {PULSE_EQ1, p, false}; state_length = num_pulses_half; have = true; }
            else if (inrange_host(len, lt.vsync_min, lt.vsync_max)) { vsync_start = (int)out.size() - 1; sp = {PULSE_VSYNC, p, false}; have = true; }
        } else if (state == PULSE_EQ1) {
            if (inrange_host(len, lt.eq_min, lt.eq_max)) { sp = {PULSE_EQ1, p, false}; have = true; }
            else if (inrange_host(len, lt.vsync_min, lt.vsync_max)) { vsync_start = (int)out.size() - 1; sp = {PULSE_VSYNC, p, false}; state_length = num_pulses_half; have = true; }
            else if (inrange_host(len, lt.hsync_min, lt.hsync_max)) { sp = {PULSE_HSYNC, p, false}; have = true; }
        } else if (state == PULSE_VSYNC) {
            if (inrange_host(len, lt.eq_min, lt.eq_max)) { sp = {PULSE_EQ2, p, false}; state_length = num_pulses_half; have = true; }
            else if (inrange_host(len, lt.vsync_min, lt.vsync_max)) { sp = {PULSE_VSYNC, p, false}; have = true; }
            else if ((double)p.start > state_end && inrange_host(len, lt.hsync_min, lt.hsync_max)) { sp = {PULSE_HSYNC, p, false}; have = true; }
        } else if (state == PULSE_EQ2) {
            if (inrange_host(len, lt.eq_min, lt.eq_max)) { sp = {PULSE_EQ2, p, false}; have = true; }
            else if (inrange_host(len, lt.hsync_min, lt.hsync_max)) { sp = {PULSE_HSYNC, p, false}; have = true; done = true; }
        }
        if (have && sp.type != state) {
            if ((double)sp.pulse.start < state_end) {
                have = false;
            } else if (state_length > 0.0) {
                state_end = (double)sp.pulse.start + ((state_length - 0.1) * (double)in_line_len);
                state_length = -1.0;
            }
        }
        if (have) {
            bool good = out.empty() ? false
                                    : host_pulse_qualitycheck(out.back().type, sp.type,
                                                              sp.pulse.start - out.back().pulse.start, in_line_len);
            sp.good = good;
            out.push_back(sp);
        }
        if (done && vsync_start >= 0) return true;
    }
    return done;
}

std::vector<HostValidPulse> host_refine_pulses_exactish(const int* starts, const int* lens, int count,
                                                        const Fmt& fmt) {
    std::vector<HostValidPulse> vp;
    int npc = std::max(0, std::min(count, MAX_PULSES));
    if (npc <= 0) return vp;
    HostLineTimings lt = host_get_line_timings(lens, npc, fmt);
    std::vector<HostPulse> raw;
    raw.reserve((size_t)npc);
    for (int i = 0; i < npc; i++) raw.push_back({starts[i], lens[i]});
    for (int i = 0; i < npc;) {
        HostPulse cur = raw[(size_t)i];
        if (inrange_host((double)cur.len, lt.hsync_min, lt.hsync_max)) {
            bool good = !vp.empty() &&
                        host_pulse_qualitycheck(vp.back().type, PULSE_HSYNC, cur.start - vp.back().pulse.start,
                                                fmt.samples_per_line);
            vp.push_back({PULSE_HSYNC, cur, good});
            i++;
        } else if (i > 2 && inrange_host((double)cur.len, lt.eq_min, lt.eq_max) && !vp.empty() &&
                   vp.back().type == PULSE_HSYNC) {
            std::vector<HostValidPulse> seq = vp;
            bool done = host_run_vblank_state_machine(raw, (size_t)(i - 2), lt, fmt.num_eq_pulses,
                                                      fmt.samples_per_line, seq);
            if (done && seq.size() > vp.size()) {
                vp = seq;
                // pipeline.cu:637-638: the size difference is taken after the
                // assignment, so this always advances by exactly one.
                i += (int)(seq.size() - vp.size()) + 1;
                i = std::min(i, npc);
            } else {
                i++;
            }
        } else {
            i++;
        }
    }
    return vp;
}

HostK2bAnchor host_get_first_hsync_loc_exactish(const std::vector<HostValidPulse>& vp, double meanlinelen,
                                                const Fmt& fmt, int prev_first_field,
                                                double last_field_offset_lines, double prev_first_hsync_loc,
                                                double prev_hsync_diff, int field_order_confidence) {
    HostK2bAnchor out;
    if (vp.empty() || meanlinelen <= 0.0) return out;
    const int field_lines[2] = {fmt.field_lines_first, fmt.field_lines_second};
    const double TOL = 0.5;
    enum { F_EQ1 = 0, F_VS = 1, F_VE = 2, F_EQ2E = 3, L_EQ1 = 4, L_VS = 5, L_VE = 6, L_EQ2E = 7 };
    double field_order_lengths[4] = {-1, -1, -1, -1};
    int vblank_pulses[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    double vblank_lines[8] = {-1, -1, -1, -1, -1, -1, -1, -1};

    int last_pulse = -1, group = 0, field_group = 0;
    for (int i = 0; i < (int)vp.size(); i++) {
        if (last_pulse != -1 && vp[(size_t)i].good) {
            if (group == 0 && vp[(size_t)i].pulse.start > vp[0].pulse.start + field_lines[0] * meanlinelen) {
                group = 4;
                field_group = 2;
            }
            int last_type = vp[(size_t)last_pulse].type;
            int type = vp[(size_t)i].type;
            if (last_type == PULSE_HSYNC && type > PULSE_HSYNC) {
                vblank_pulses[0 + group] = vp[(size_t)i].pulse.start;
                field_order_lengths[0 + field_group] = round_nearest_line_loc_host(
                    (vp[(size_t)i].pulse.start - vp[(size_t)last_pulse].pulse.start) / meanlinelen);
            } else if (last_type == PULSE_EQ1 && type == PULSE_VSYNC) {
                vblank_pulses[1 + group] = vp[(size_t)i].pulse.start;
            } else if (last_type == PULSE_VSYNC && type == PULSE_EQ2) {
                vblank_pulses[2 + group] = vp[(size_t)i].pulse.start;
            } else if (last_type > PULSE_HSYNC && type == PULSE_HSYNC) {
                vblank_pulses[3 + group] = vp[(size_t)last_pulse].pulse.start;
                field_order_lengths[1 + field_group] = round_nearest_line_loc_host(
                    (vp[(size_t)i].pulse.start - vp[(size_t)last_pulse].pulse.start) / meanlinelen);
            }
        }
        last_pulse = i;
    }

    // pipeline.cu:707-717
    const bool pal = fmt.system == 1;
    double first_field_lengths[4] = {1, 0.5, 0.5, 1};
    double second_field_lengths[4] = {0.5, 1, 1, 0.5};
    double progressive_field_lengths[4] = {1, 0.5, 1, 0.5};
    if (pal) {
        const double pf[4] = {0.5, 0.5, 1, 1}, ps[4] = {1, 1, 0.5, 0.5}, pp[4] = {0.5, 0.5, 0.5, 0.5};
        for (int i = 0; i < 4; i++) {
            first_field_lengths[i] = pf[i];
            second_field_lengths[i] = ps[i];
            progressive_field_lengths[i] = pp[i];
        }
    }
    double ic = 0.0, id = 0.0, pc = 0.0, pd = 0.0;
    for (int i = 0; i < 4; i++) {
        double fl = field_order_lengths[i];
        if (fl == first_field_lengths[i]) { ic += 1.0; id += 1.0; }
        if (fl == second_field_lengths[i]) id += 1.0;
        if (fl == progressive_field_lengths[i]) { pc += 1.0; pd += 1.0; }
        if (fl != -1) pd += 1.0;
    }
    (void)pc;
    (void)pd;
    bool first_field;
    if (prev_first_field == -1)
        first_field = id == 0.0 || std::round(ic / std::max(1.0, id)) == 1.0;
    else
        first_field = !prev_first_field;

    if (id > 0.0) {
        if (prev_first_hsync_loc < 0.0) field_order_confidence = std::min(field_order_confidence, 50);
        double weighting = id / 4.0;
        int ffc = (int)std::llround((ic / id) * weighting * 100.0);
        int sfc = (int)std::llround(((id - ic) / id) * weighting * 100.0);
        if (ffc >= field_order_confidence && ffc > sfc) first_field = true;
        else if (sfc >= field_order_confidence && sfc > ffc) first_field = false;
    }

    double vsync_section_lines = fmt.num_eq_pulses / 2.0;
    double current_field_lines = first_field ? field_lines[0] : field_lines[1];
    double previous_field_lines = first_field ? field_lines[1] : field_lines[0];
    double* cfl = first_field ? first_field_lengths : second_field_lengths;
    double hsync_start_line;
    vblank_lines[F_EQ1] = 0.0 + cfl[0];
    vblank_lines[F_VS] = vblank_lines[F_EQ1] + vsync_section_lines;
    vblank_lines[F_VE] = vblank_lines[F_VS] + vsync_section_lines;
    vblank_lines[F_EQ2E] = vblank_lines[F_VE] + vsync_section_lines - 0.5;
    hsync_start_line = vblank_lines[F_EQ2E] + cfl[1];
    vblank_lines[L_EQ1] = current_field_lines + cfl[2];
    vblank_lines[L_VS] = vblank_lines[L_EQ1] + vsync_section_lines;
    vblank_lines[L_VE] = vblank_lines[L_VS] + vsync_section_lines;
    vblank_lines[L_EQ2E] = vblank_lines[L_VE] + vsync_section_lines - 0.5;

    auto calc_sync = [&](int a, int b, double& acc_loc, int& acc_count, double& acc_offset) {
        int p1 = vblank_pulses[a], p2 = vblank_pulses[b];
        if (p1 == -1 || p2 == -1 || meanlinelen == 0.0) return;
        double actual = ((double)p1 - (double)p2) / meanlinelen;
        double expected = vblank_lines[a] - vblank_lines[b];
        if (actual < expected + TOL && actual > expected - TOL) {
            acc_offset += actual - expected;
            acc_loc += (double)p2 + meanlinelen * (hsync_start_line - vblank_lines[b]);
            acc_count += 1;
        }
    };

    double fv_loc = 0.0, fv_off = 0.0;
    int fv_n = 0;
    int fi[4] = {F_EQ1, F_VS, F_VE, F_EQ2E};
    for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++) calc_sync(fi[a], fi[b], fv_loc, fv_n, fv_off);
    double lv_loc = 0.0, lv_off = 0.0;
    int lv_n = 0;
    int li[4] = {L_EQ1, L_VS, L_VE, L_EQ2E};
    for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++) calc_sync(li[a], li[b], lv_loc, lv_n, lv_off);

    double first_hsync_loc = 0.0;
    int valid_location_count = 0;
    double offset = 0.0;
    double first_est = fv_n ? fv_loc / fv_n : 0.0;
    double last_est = lv_n ? lv_loc / lv_n : 0.0;
    if (fv_n && lv_n && first_est < last_est + TOL * meanlinelen && first_est > last_est - TOL * meanlinelen) {
        first_hsync_loc = fv_loc + lv_loc;
        valid_location_count = fv_n + lv_n;
        offset = fv_off + lv_off;
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) calc_sync(fi[a], li[b], first_hsync_loc, valid_location_count, offset);
    } else if (fv_n == 6 || (prev_first_hsync_loc <= 0.0 && fv_n > lv_n)) {
        first_hsync_loc = fv_loc;
        valid_location_count = fv_n;
        offset = fv_off;
    } else if (lv_n == 6 || (prev_first_hsync_loc <= 0.0 && lv_n > fv_n)) {
        first_hsync_loc = lv_loc;
        valid_location_count = lv_n;
        offset = lv_off;
    }

    double est_field_lines = pal ? current_field_lines : previous_field_lines;  // pipeline.cu:811
    int estimated_hsync_loc = (int)std::llround(
        (last_field_offset_lines + est_field_lines + prev_first_hsync_loc / meanlinelen) * meanlinelen);
    bool used_estimated = false;
    if (valid_location_count == 0 && prev_first_hsync_loc > 0.0) {
        double ewo = (prev_hsync_diff <= 0.5 && prev_hsync_diff >= -0.5)
                         ? estimated_hsync_loc + meanlinelen * prev_hsync_diff
                         : (double)estimated_hsync_loc;
        if (ewo <= 0.0) ewo = vp.empty() ? 0.0 : (double)vp.front().pulse.start;
        first_hsync_loc += ewo;
        valid_location_count += 1;
        used_estimated = true;
    }
    if (valid_location_count <= 0) return out;
    offset /= valid_location_count;
    first_hsync_loc = std::round((first_hsync_loc + offset) / valid_location_count);
    if (!used_estimated) prev_hsync_diff = (first_hsync_loc - estimated_hsync_loc) / meanlinelen;

    double hsync_offset = 0.0;
    int hsync_count = 0;
    for (const HostValidPulse& p : vp) {
        if (p.type != PULSE_HSYNC || !p.good) continue;
        double lineloc = ((double)p.pulse.start - first_hsync_loc) / meanlinelen + hsync_start_line;
        int rlineloc = (int)std::llround(lineloc);
        if (rlineloc > current_field_lines) break;
        if (rlineloc >= hsync_start_line) {
            hsync_offset += first_hsync_loc + meanlinelen * ((double)rlineloc - hsync_start_line) - (double)p.pulse.start;
            hsync_count++;
        }
    }
    if (hsync_count > 0) {
        hsync_offset /= hsync_count;
        first_hsync_loc -= hsync_offset;
    }
    out.valid = true;
    out.line0loc = first_hsync_loc - meanlinelen * hsync_start_line;
    out.first_hsync_loc = first_hsync_loc;
    out.first_hsync_line = hsync_start_line;
    out.next_field = first_hsync_loc + meanlinelen * (vblank_lines[L_EQ1] - hsync_start_line);
    out.first_field = first_field ? 1 : 0;
    out.progressive_field = 0;
    out.prev_hsync_diff = prev_hsync_diff;
    return out;
}

std::vector<double> host_valid_pulses_to_linelocs_exactish(const std::vector<HostValidPulse>& vp, double ref_pulse,
                                                           double ref_line, double meanlinelen, int proclines) {
    std::vector<double> ll((size_t)proclines, 0.0);
    std::vector<double> ps;
    ps.reserve(vp.size());
    for (const HostValidPulse& p : vp) ps.push_back((double)p.pulse.start);
    std::sort(ps.begin(), ps.end());
    double max_allowed = meanlinelen / 1.5;
    size_t cur = 0;
    for (int line = 0; line < proclines; line++) {
        ll[(size_t)line] = ref_pulse + meanlinelen * ((double)line - ref_line);
        if (cur >= ps.size()) continue;
        double cd_ = fabs(ps[cur] - ll[(size_t)line]);
        double smallest = max_allowed;
        double loc = -1.0;
        size_t si = cur;
        while (si + 1 < ps.size()) {
            if (cd_ <= smallest) {
                smallest = cd_;
                cur = si;
                loc = ps[si];
            }
            double nd = fabs(ps[si + 1] - ll[(size_t)line]);
            if (nd > cd_) break;
            cd_ = nd;
            si++;
        }
        if (loc >= 0.0) {
            ll[(size_t)line] = loc;
            cur++;
        }
    }
    return ll;
}

struct K2bState {
    int prev_first_field = -1;
    int prev_progressive_field = -1;
    long long prev_first_hsync_readloc = -1;
    double prev_first_hsync_loc = -1.0;
    double prev_first_hsync_diff = -1.0;
};

// K2b driver (pipeline.cu:1696-1791)
// good (optional): the validated pulses' good flags [nf][800]; anchor
// (optional): per field valid, line0loc, first_hsync_loc, first_hsync_line,
// next_field, first_field, progressive_field, prev_hsync_diff, meanlinelen.
void k2b_chunk(const Fmt& fmt, K2bState& st, size_t raw_offset, const std::vector<size_t>& offs, int nf,
               int* starts, int* lens, int* count, int* types, double* linelocs, int* is_first,
               unsigned char* good = nullptr, double* anchor = nullptr) {
    const int LPF = fmt.lines_per_frame;
    for (int field = 0; field < nf; field++) {
        int before = std::max(0, std::min(count[field], MAX_PULSES));
        int* st_f = starts + (size_t)field * MAX_PULSES;
        int* ln_f = lens + (size_t)field * MAX_PULSES;
        std::vector<HostValidPulse> vp = host_refine_pulses_exactish(st_f, ln_f, before, fmt);
        double mll = host_meanlinelen(vp, (double)fmt.samples_per_line);
        double prev_offset_lines = 0.0;
        if (st.prev_first_hsync_readloc != -1 && mll > 0.0)
            prev_offset_lines = ((double)st.prev_first_hsync_readloc - (double)(raw_offset + offs[(size_t)field])) / mll;
        HostK2bAnchor an = host_get_first_hsync_loc_exactish(vp, mll, fmt, st.prev_first_field, prev_offset_lines,
                                                             st.prev_first_hsync_loc, st.prev_first_hsync_diff, 75);
        int after = std::min((int)vp.size(), MAX_PULSES);
        double first_start = before > 0 ? (double)st_f[0] : 0.0;
        for (int i = 0; i < after; i++) {
            st_f[i] = vp[(size_t)i].pulse.start;
            ln_f[i] = vp[(size_t)i].pulse.len;
            types[(size_t)field * MAX_PULSES + i] = vp[(size_t)i].type;
            if (good) good[(size_t)field * MAX_PULSES + i] = vp[(size_t)i].good ? 1 : 0;
        }
        for (int i = after; i < MAX_PULSES; i++) {
            st_f[i] = 0;
            ln_f[i] = 0;
            types[(size_t)field * MAX_PULSES + i] = PULSE_HSYNC;
            if (good) good[(size_t)field * MAX_PULSES + i] = 0;
        }
        if (anchor) {
            double* a = anchor + (size_t)field * 9;
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
        count[field] = after;
        double* ll = linelocs + (size_t)field * LPF;
        if (an.valid) {
            std::vector<double> locs =
                host_valid_pulses_to_linelocs_exactish(vp, an.first_hsync_loc, an.first_hsync_line, mll, LPF);
            std::copy(locs.begin(), locs.end(), ll);
            is_first[field] = an.first_field;
            st.prev_first_field = an.first_field;
            st.prev_progressive_field = an.progressive_field;
            st.prev_first_hsync_readloc = (long long)(raw_offset + offs[(size_t)field]);
            st.prev_first_hsync_loc = an.first_hsync_loc;
            st.prev_first_hsync_diff = an.prev_hsync_diff;
        } else {
            for (int line = 0; line < LPF; line++) ll[line] = first_start + mll * (double)line;
            is_first[field] = (st.prev_first_field >= 0) ? (1 - st.prev_first_field) : 1;
        }
    }
}

// ---------------------------------------------------------------------------
// K4 hsync refine (hsync_refine.cu:21-271)
// ---------------------------------------------------------------------------
double find_falling_crossing(const double* sig, int start, int count, double target, int total) {
    if (start < 0 || start + count > total || count < 2) return -1.0;
    for (int i = start + 1; i < start + count; i++) {
        double prev = sig[i - 1], curr = sig[i];
        if (prev >= target && curr < target) {
            double a = prev - target, b = curr - target;
            double frac = (a - b != 0.0) ? (a / (a - b)) : 0.0;
            return (double)(i - 1) + frac;
        }
    }
    return -1.0;
}

double find_rising_crossing(const double* sig, int start, int count, double target, int total) {
    if (start < 0 || start + count > total || count < 2) return -1.0;
    if (sig[start] > target) return -1.0;
    for (int i = start + 1; i < start + count; i++) {
        double prev = sig[i - 1], curr = sig[i];
        if (prev < target && curr >= target) {
            double a = target - prev, b = target - curr;
            double frac = (a + b != 0.0) ? (a / (a - b)) : 0.0;
            return (double)(i - 1) + frac;
        }
    }
    return -1.0;
}

double window_median(const double* sig, int start, int count, int total) {
    if (start < 0 || start + count > total || count <= 0) return 0.0;
    double buf[64];
    int n = std::min(count, 64);
    for (int i = 0; i < n; i++) buf[i] = sig[start + i];
    int mid = n / 2;
    for (int i = 0; i <= mid; i++)
        for (int j = i + 1; j < n; j++)
            if (buf[j] < buf[i]) std::swap(buf[i], buf[j]);
    return buf[mid];
}

void hsync_refine(const Fmt& fmt, const double* d05, double* linelocs, int nf, int total) {
    double one_usec_samples = 1.0e-6 * fmt.sample_rate;
    double right_edge_offset = 2.25 * ((fmt.sample_rate / 1.0e6) / 40.0);
    double thr = fmt.pulse_threshold_hz;
    double hw = fmt.hsync_width;
    int LPF = fmt.lines_per_frame;
    for (int idx = 0; idx < nf * LPF; idx++) {
        int line = idx % LPF;
        if (line < fmt.active_line_start) continue;
        double* loc = &linelocs[idx];
        double original = *loc;
        int one_usec = (int)(one_usec_samples + 0.5);
        int search_start = (int)(original - one_usec);
        int search_count = one_usec * 2;
        double zc = find_falling_crossing(d05, search_start, search_count, thr, total);
        double left_result = -1.0;
        if (zc >= 0.0) {
            int zc_int = (int)(zc + 0.5);
            double sync_level = window_median(d05, zc_int + (int)(1.0 * one_usec_samples),
                                              (int)(1.5 * one_usec_samples), total);
            double porch_level = window_median(d05, zc_int + (int)(8.0 * one_usec_samples),
                                               (int)(1.0 * one_usec_samples), total);
            if (porch_level > sync_level) {
                double midpoint = (sync_level + porch_level) / 2.0;
                int refine_start = (int)(original - one_usec * 2);
                int refine_count = (int)(one_usec_samples * 4);
                double zc2 = find_falling_crossing(d05, refine_start, refine_count, midpoint, total);
                if (zc2 >= 0.0 && fabs(zc2 - zc) <= one_usec_samples * 0.5) left_result = zc2;
            }
        }
        double right_result = -1.0;
        int rs = (int)(original + hw - one_usec_samples);
        int rc = (int)(hw * 2);
        double right_cross = find_rising_crossing(d05, rs, rc, thr, total);
        if (right_cross >= 0.0) {
            double zc_fr = right_cross - hw;
            int zc_fr_int = (int)(zc_fr + 0.5);
            double r_sync = window_median(d05, zc_fr_int + (int)(1.0 * one_usec_samples),
                                          (int)(1.5 * one_usec_samples), total);
            double r_porch = window_median(d05, zc_fr_int + (int)(hw + 1.0 * one_usec_samples),
                                           (int)(1.0 * one_usec_samples), total);
            if (r_porch > r_sync) {
                double r_mid = (r_sync + r_porch) / 2.0;
                int rrs = (int)(original + hw - one_usec_samples);
                int rrc = (int)(hw * 2);
                double zc2_r = find_rising_crossing(d05, rrs, rrc, r_mid, total);
                if (zc2_r >= 0.0 && fabs(zc2_r - right_cross) <= one_usec_samples * 0.5) {
                    double candidate = right_cross - hw + right_edge_offset;
                    double reference = (left_result >= 0.0) ? left_result : original;
                    if (fabs(candidate - reference) <= one_usec_samples * 2.0) right_result = candidate;
                }
            }
        }
        if (right_result >= 0.0) *loc = right_result;
        else if (left_result >= 0.0) *loc = left_result;
    }
}

// ---------------------------------------------------------------------------
// K5 TBC resample (tbc_resample.cu:19-133; debug counters omitted)
// ---------------------------------------------------------------------------
void tbc_resample(const Fmt& fmt, const double* demod, const double* linelocs, uint16_t* luma, int nf,
                  int total) {
    const int L = fmt.output_line_len, R = fmt.output_field_lines, LPF = fmt.lines_per_frame;
    for (int field = 0; field < nf; field++) {
        for (int out_line = 0; out_line < R; out_line++) {
            for (int out_col = 0; out_col < L; out_col++) {
                size_t idx = ((size_t)field * R + out_line) * L + out_col;
                int ll_line = out_line, ll_next = ll_line + 1;
                if (ll_next >= LPF) { luma[idx] = (uint16_t)fmt.output_zero; continue; }
                double line_start = linelocs[(size_t)field * LPF + ll_line];
                double line_end = linelocs[(size_t)field * LPF + ll_next];
                double line_len = line_end - line_start;
                double frac = (double)out_col / (double)L;
                double coord = line_start + frac * line_len;
                int ci = (int)coord;
                double x = coord - (double)ci;
                if (ci < 1 || ci + 2 >= total) { luma[idx] = (uint16_t)fmt.output_zero; continue; }
                double p0 = demod[ci - 1], p1 = demod[ci], p2 = demod[ci + 1], p3 = demod[ci + 2];
                double a = p2 - p0;
                double b = 2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3;
                double c = 3.0 * (p1 - p2) + p3 - p0;
                double hz = p1 + 0.5 * x * (a + x * (b + x * c));
                double ire = (hz - fmt.ire0) / fmt.hz_ire - fmt.vsync_ire;
                double out_val = ire * fmt.output_scale + fmt.output_zero;
                if (out_val < 0.0) out_val = 0.0;
                if (out_val > 65535.0) out_val = 65535.0;
                luma[idx] = (uint16_t)(out_val + 0.5);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// K6 chroma decode (chroma_decode.cu:13-1002)
// ---------------------------------------------------------------------------
struct ChromaState {
    bool valid = false;
    int current_track = 0;
    double good_metric_threshold = 0.0;
    int cycle_start = 0;
    double trial_metric[2] = {0.0, 0.0};   // field-0 trial of the first chunk, per track (test access)
    bool keep_trial = false;                // keep the trial's band-passed lines (test access)
    std::vector<double> trial_buf[3];       // track 0, track 1, detected track (re-run for the phase)
};

struct NTSCPhaseEntry { int is_first, quadrant, delta, phase_id, offset; };
const NTSCPhaseEntry ntsc_phase_table[] = {
    {1, 2, 0, 3, 0}, {0, 3, 1, 2, 1}, {1, 1, 2, 1, 1}, {0, 0, 3, 4, 0},
    {1, 0, 0, 3, 2}, {0, 1, 1, 2, 3}, {1, 3, 2, 1, 3}, {0, 2, 3, 4, 2},
    {1, 2, -1, 3, 0}, {0, 3, -1, 2, 1}, {1, 1, -1, 1, 1}, {0, 0, -1, 4, 0},
    {1, 0, -1, 3, 2}, {0, 1, -1, 2, 3}, {1, 3, -1, 1, 3}, {0, 2, -1, 4, 2},
};

int lookup_ntsc_phase_id(int is_first, int quadrant, int delta) {
    for (const auto& e : ntsc_phase_table)
        if (e.is_first == is_first && e.quadrant == quadrant && e.delta == delta) return e.phase_id;
    for (const auto& e : ntsc_phase_table)
        if (e.is_first == is_first && e.quadrant == quadrant && e.delta == -1) return e.phase_id;
    return 1;
}

struct ChromaConst {
    int fft_size = 1024, freq_bins = 513;
    double het_scale = 0, fft_scale = 0;
    int fsc_bin = 0, bandwidth_bins = 0, burst_start = 0, burst_end = 0;
};

ChromaConst chroma_consts(const Fmt& fmt) {
    ChromaConst c;
    int v = 1;
    while (v < fmt.output_line_len) v <<= 1;
    c.fft_size = v;
    c.freq_bins = v / 2 + 1;
    double het_freq = fmt.fsc + fmt.chroma_under;
    c.het_scale = het_freq / fmt.output_rate;
    c.fsc_bin = (int)(fmt.fsc / fmt.output_rate * c.fft_size + 0.5);
    c.bandwidth_bins = (int)(500000.0 / fmt.output_rate * c.fft_size + 0.5);
    if (c.bandwidth_bins < 10) c.bandwidth_bins = 10;
    c.burst_start = (int)(fmt.burst_start_us * 1e-6 * fmt.output_rate + 0.5);
    c.burst_end = (int)(fmt.burst_end_us * 1e-6 * fmt.output_rate + 0.5);
    c.fft_scale = 1.0 / (double)c.fft_size;
    return c;
}

// k_resample_raw_het + D2Z + k_chroma_bandpass + Z2D for fields
// [f0, f0+nfl) with the given per-field tracks (phase offsets all 0,
// chroma_decode.cu:754) -> out[(f-f0)*263 + line][fft_size]
void chroma_lines(const Fmt& fmt, const ChromaConst& cc, const FFTPlan& plan, const double* craw,
                  const double* linelocs, int total, int f0, int nfl, const std::vector<int>& track,
                  std::vector<double>& out) {
    const int R = fmt.output_field_lines, L = fmt.output_line_len, LPF = fmt.lines_per_frame;
    const int F = cc.fft_size;
    out.assign((size_t)nfl * R * F, 0.0);
#pragma omp parallel for schedule(static)
    for (int li = 0; li < nfl * R; li++) {
        std::vector<double> row((size_t)F);
        std::vector<cd> spec((size_t)cc.freq_bins);
        int abs_line = f0 * R + li;
        int field = abs_line / R;
        int out_line = abs_line % R;
        for (int col = 0; col < F; col++) {
            double v = 0.0;
            int ll_line = out_line, ll_next = out_line + 1;
            if (col < L && ll_next < LPF) {
                double ls = linelocs[(size_t)field * LPF + ll_line];
                double le = linelocs[(size_t)field * LPF + ll_next];
                double frac = (double)col / (double)L;
                double coord = ls + frac * (le - ls);
                int ci = (int)coord;
                double x = coord - (double)ci;
                if (!(ci < 1 || ci + 2 >= total)) {
                    double p0 = craw[ci - 1], p1 = craw[ci], p2 = craw[ci + 1], p3 = craw[ci + 2];
                    double a = p2 - p0;
                    double b = 2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3;
                    double c = 3.0 * (p1 - p2) + p3 - p0;
                    double raw_val = p1 + 0.5 * x * (a + x * (b + x * c));
                    int rot = track[(size_t)field] ? 1 : 3;
                    int line_phase = (out_line * rot + 0) & 3;
                    double phase_offset = (double)line_phase * (M_PI * 0.5);
                    double abs_sample = (double)(out_line * L + col);
                    double het = -cos(2.0 * M_PI * cc.het_scale * abs_sample + phase_offset);
                    v = raw_val * het;
                }
            }
            row[(size_t)col] = v;
        }
        orc::fft_r2c(plan, row.data(), spec.data());
        int lo = cc.fsc_bin - cc.bandwidth_bins, hi = cc.fsc_bin + cc.bandwidth_bins;
        for (int k = 0; k < cc.freq_bins; k++)
            if (k < lo || k > hi) spec[(size_t)k] = cd(0.0, 0.0);
        orc::fft_c2r(plan, spec.data(), &out[((size_t)li) * F]);
    }
}

// measure_burst_cancellation (chroma_decode.cu:383-411) / k_burst_cancellation (:242-296)
double burst_metric(const double* fieldbuf, const Fmt& fmt, const ChromaConst& cc) {
    const int SKIP = 16, R = fmt.output_field_lines, F = cc.fft_size;
    int burst_len = cc.burst_end - cc.burst_start;
    double total = 0.0;
    int count = 0;
    for (int line = SKIP; line < R - SKIP - 1; line += 2) {
        const double* a = fieldbuf + (size_t)line * F;
        const double* b = fieldbuf + (size_t)(line + 1) * F;
        double sum = 0.0;
        for (int i = cc.burst_start; i < cc.burst_end && i < fmt.output_line_len; i++)
            sum += fabs(a[i] * cc.fft_scale + b[i] * cc.fft_scale);
        total += sum / burst_len;
        count++;
    }
    return count > 0 ? total / count : 1e9;
}

// measure_burst_phase (chroma_decode.cu:415-450)
double burst_phase(const double* fieldbuf, const Fmt& fmt, const ChromaConst& cc) {
    const int SKIP = 16, R = fmt.output_field_lines, F = cc.fft_size;
    double It = 0.0, Qt = 0.0;
    for (int line = SKIP; line < R - SKIP; line++) {
        const double* d = fieldbuf + (size_t)line * F;
        double I = 0.0, Q = 0.0;
        for (int i = cc.burst_start; i < cc.burst_end && i < fmt.output_line_len; i++) {
            double val = d[i] * cc.fft_scale;
            double t = 2.0 * M_PI * fmt.fsc / fmt.output_rate * (double)i;
            I += val * cos(t);
            Q += val * sin(t);
        }
        double mag = sqrt(I * I + Q * Q);
        if (mag > 0.0) { It += I / mag; Qt += Q / mag; }
    }
    double ph = atan2(Qt, It);
    return fmod(ph * 180.0 / M_PI + 360.0, 360.0);
}

void chroma_decode(const Fmt& fmt, const double* raw, const double* linelocs, double* scratch, uint16_t* out,
                   int nf, int total, std::vector<int>& phase_ids, ChromaState* state) {
    const int R = fmt.output_field_lines, L = fmt.output_line_len;
    ChromaConst cc = chroma_consts(fmt);
    FFTPlan plan(cc.fft_size);
    // pre-BPF per field segment, circular (chroma_decode.cu:568-638)
    int field_stride = total / nf;
    int rn = 0;
    std::vector<double> hbpf = make_chroma_prebpf(fmt, field_stride, &rn);
    {
        FFTPlan bplan(rn);
        int rb = rn / 2 + 1;
#pragma omp parallel for schedule(dynamic, 1)
        for (int f = 0; f < nf; f++) {
            std::vector<double> pad((size_t)rn, 0.0), y((size_t)rn);
            std::vector<cd> X((size_t)rb);
            size_t fo = (size_t)f * field_stride;
            std::copy(raw + fo, raw + fo + field_stride, pad.begin());
            orc::fft_r2c(bplan, pad.data(), X.data());
            for (int k = 0; k < rb; k++) {
                double m = hbpf[(size_t)k];
                X[(size_t)k] = cd(X[(size_t)k].real() * m, X[(size_t)k].imag() * m);
            }
            orc::fft_c2r(bplan, X.data(), y.data());
            std::copy(y.begin(), y.begin() + field_stride, scratch + fo);
        }
    }
    const double* craw = scratch;
    static const int cycle[8] = {3, 2, 1, 4, 3, 2, 1, 4};
    int detected_track, cycle_start;
    double thr;
    std::vector<double> buf;
    if (state && state->valid) {
        detected_track = state->current_track;
        thr = state->good_metric_threshold;
        cycle_start = state->cycle_start;
    } else {
        double metric[2];
        std::vector<int> tr((size_t)nf, 0);
        for (int t = 0; t < 2; t++) {
            tr[0] = t;
            chroma_lines(fmt, cc, plan, craw, linelocs, total, 0, 1, tr, buf);
            metric[t] = burst_metric(buf.data(), fmt, cc);
            if (state && state->keep_trial) state->trial_buf[t] = buf;
        }
        detected_track = (metric[1] < metric[0]) ? 1 : 0;
        thr = metric[detected_track] * 4.0;
        if (state) { state->trial_metric[0] = metric[0]; state->trial_metric[1] = metric[1]; }
        tr[0] = detected_track;
        chroma_lines(fmt, cc, plan, craw, linelocs, total, 0, 1, tr, buf);
        if (state && state->keep_trial) state->trial_buf[2] = buf;
        double ph0 = burst_phase(buf.data(), fmt, cc);
        int q0 = ((int)(ph0 / 90.0 + 0.5)) % 4;
        int pid0 = lookup_ntsc_phase_id(1, q0, -1);
        cycle_start = 0;
        for (int i = 0; i < 8; i += 2)
            if (cycle[i] == pid0) { cycle_start = i; break; }
    }
    std::vector<int> h_track((size_t)nf);
    int current_track = detected_track;
    phase_ids.assign((size_t)nf, 0);
    for (int f = 0; f < nf; f++) {
        h_track[(size_t)f] = (f & 1) ? (1 - current_track) : current_track;
        phase_ids[(size_t)f] = cycle[(cycle_start + f) % 8];
    }
    // sub-chunks of whole fields, at most 4096 lines (chroma_decode.cu:783-801)
    int max_chunk_lines = (4096 / R) * R;
    int total_lines = nf * R;
    int chunk_lines = std::min(max_chunk_lines, total_lines);
    chunk_lines = (chunk_lines / R) * R;
    int fpc = chunk_lines / R;
    for (int fs = 0; fs < nf; fs += fpc) {
        int fthis = std::min(fpc, nf - fs);
        int retries = 3;
        bool ok = false;
        while (!ok && retries-- > 0) {
            chroma_lines(fmt, cc, plan, craw, linelocs, total, fs, fthis, h_track, buf);
            int first_bad = -1;
            for (int f = 0; f < fthis; f++) {
                double m = burst_metric(&buf[(size_t)f * R * cc.fft_size], fmt, cc);
                if (m > thr) { first_bad = f; break; }
            }
            if (first_bad < 0) {
                ok = true;
            } else {
                int abs_f = fs + first_bad;
                current_track = 1 - current_track;
                for (int ff = abs_f; ff < nf; ff++) h_track[(size_t)ff] = (ff & 1) ? (1 - current_track) : current_track;
            }
        }
        // k_chroma_acc_output (chroma_decode.cu:188-231)
        for (int li = 0; li < fthis * R; li++) {
            const double* d = &buf[(size_t)li * cc.fft_size];
            uint16_t* o = out + ((size_t)fs * R + li) * L;
            int blen = cc.burst_end - cc.burst_start;
            double ssq = 0.0;
            for (int i = 0; i < blen; i++) {
                double v = d[cc.burst_start + i] * cc.fft_scale;
                ssq += v * v;
            }
            double rms = sqrt(ssq / (double)blen);
            double acc = (rms > 1.0) ? (fmt.burst_abs_ref / rms) : 0.0;
            for (int col = 0; col < L; col++) {
                double v = d[col] * cc.fft_scale * acc;
                double ov = v + 32768.0;
                if (ov < 0.0) ov = 0.0;
                if (ov > 65535.0) ov = 65535.0;
                o[col] = (uint16_t)(ov + 0.5);
            }
        }
    }
    if (state) {
        state->valid = true;
        state->current_track = (nf & 1) ? (1 - current_track) : current_track;
        state->good_metric_threshold = thr;
        state->cycle_start = (cycle_start + nf) % 8;
    }
}

// ---------------------------------------------------------------------------
// K7 dropout detect + conceal (dropout_detect.cu:17-339)
// ---------------------------------------------------------------------------
int find_tbc_line(const double* ll, int als, int n_lines, double rf_pos) {
    if (rf_pos < ll[als] || rf_pos >= ll[als + n_lines]) return -1;
    int lo = 0, hi = n_lines - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (ll[als + mid] <= rf_pos) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

void dropout_detect(const Fmt& fmt, const double* env, const double* linelocs, uint16_t* luma, uint16_t* chroma,
                    int* do_lines, int* do_starts, int* do_ends, int* do_count, int nf) {
    const int spf = fmt.spfp, LPF = fmt.lines_per_frame, R = fmt.output_field_lines, L = fmt.output_line_len;
    const int als = fmt.active_line_start;
    for (int field = 0; field < nf; field++) {
        double sum = 0.0;
        for (int i = 0; i < spf; i++) sum += env[(size_t)field * spf + i];
        double mean_env = sum / (double)spf;
        const double* fll = linelocs + (size_t)field * LPF;
        int* ml = do_lines + (size_t)field * MAX_DO;
        int* ms = do_starts + (size_t)field * MAX_DO;
        int* me = do_ends + (size_t)field * MAX_DO;
        double down = mean_env * 0.18;
        double up = down * 1.25;
        int field_start = field * spf;
        int scan_start = std::max((int)fll[als], field_start);
        int scan_end = std::min((int)fll[als + R], field_start + spf);
        do_count[field] = 0;
        if (scan_end <= scan_start) continue;
        std::vector<int> rs, re;
        int rf_do_start = -1;
        const int MAX_RF_DO = 128;
        int nblk = (scan_end - scan_start + 16 - 1) / 16;
        for (int blk = 0; blk < nblk; blk++) {
            int bs = scan_start + blk * 16;
            int be = std::min(bs + 16, scan_end);
            double bsum = 0.0;
            for (int i = bs; i < be; i++) bsum += env[i];
            double bm = bsum / (double)(be - bs);
            if (bm <= down) {
                if (rf_do_start < 0) rf_do_start = bs;
            } else if (bm >= up) {
                if (rf_do_start >= 0) {
                    if (!rs.empty() && rf_do_start - re.back() <= 30) re.back() = bs;
                    else if ((int)rs.size() < MAX_RF_DO) { rs.push_back(rf_do_start); re.push_back(bs); }
                    rf_do_start = -1;
                }
            }
        }
        if (rf_do_start >= 0) {
            if (!rs.empty() && rf_do_start - re.back() <= 30) re.back() = scan_end;
            else if ((int)rs.size() < MAX_RF_DO) { rs.push_back(rf_do_start); re.push_back(scan_end); }
        }
        std::vector<int> fs2, fe2;
        for (size_t i = 0; i < rs.size(); i++)
            if (re[i] - rs[i] >= 10) { fs2.push_back(rs[i]); fe2.push_back(re[i]); }
        int tc = 0;
        for (size_t i = 0; i < fs2.size() && tc < MAX_DO; i++) {
            double srf = (double)fs2[i], erf = (double)fe2[i];
            int sl = find_tbc_line(fll, als, R, srf);
            if (sl < 0) continue;
            int el = find_tbc_line(fll, als, R, erf - 1.0);
            if (el < 0) el = R - 1;
            for (int l = sl; l <= el && tc < MAX_DO; l++) {
                double ls = fll[als + l], le = fll[als + l + 1];
                double llen = le - ls;
                if (llen <= 0.0) continue;
                int sx, ex;
                if (l == sl) { sx = (int)((srf - ls) / llen * (double)L); sx = std::max(0, sx); }
                else sx = 0;
                if (l == el) { ex = (int)((erf - ls) / llen * (double)L + 0.999); ex = std::min(L, ex); }
                else ex = L;
                if (ex > sx) { ml[tc] = l; ms[tc] = sx; me[tc] = ex; tc++; }
            }
        }
        do_count[field] = tc;
        // k_conceal (dropout_detect.cu:248-291), entries applied in order
        uint16_t* lu = luma + (size_t)field * R * L;
        uint16_t* ch = chroma + (size_t)field * R * L;
        for (int d = 0; d < tc; d++) {
            int line = ml[d], sx = ms[d], ex = me[d];
            int src = (line > 0) ? line - 1 : line + 1;
            if (src >= R) continue;
            for (int col = sx; col < ex; col++) {
                lu[line * L + col] = lu[src * L + col];
                ch[line * L + col] = ch[src * L + col];
            }
        }
    }
}

}  // namespace

// ===========================================================================
// Whole-capture decoder (pipeline.cu:1525-2713)
// ===========================================================================
// Wall time of each stage of a chunk, accumulated per decode (the CPU
// baseline's per-stage figures): convert, K1, VSYNC + clustering, K2, K2b,
// K4, K5, K6, K7.
constexpr int kOrStages = 9;
struct StageClock {
    double* acc;
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    explicit StageClock(double* a) : acc(a) {}
    void lap(int i) {
        auto n = std::chrono::steady_clock::now();
        acc[i] += std::chrono::duration<double>(n - t).count();
        t = n;
    }
};

struct ChunkTrace {
    size_t raw_offset = 0, next_offset = 0;
    int fields_k1 = 0, nf = 0;
    int candidates = 0;                   // VSYNC candidates found, before the cap
    std::vector<size_t> offsets;
    std::vector<int> pc, ps, pl;          // K2 output
    std::vector<int> vc, vs, vl, vt;      // K2b output
    std::vector<double> llc, llr;
    std::vector<int> is_first;
};

struct or_dec {
    Fmt fmt;
    int batch = 64;
    bool trace = false;
    K1Filters k1;
    K2bState k2b;
    ChromaState chroma;
    bool last_parity = true;
    int in_fmt = 0;       // 0 u8, 1 s16, 2 u16 (InputFormat, video_format.h)
    bool dc_correct = false;
    int line_locator = 0; // 0 K2b (default), 1 K3 (CUVHS_K2B_DISABLE)
    bool diag = false;    // collect the K3/K4/K5 diagnostic counters
    double stage_s[kOrStages] = {};
    std::vector<std::vector<int>> diag_counts;  // per chunk: K3 x3, K4 x3, K5 x3
    // persistent device-like buffers
    std::vector<double> raw, demod, d05, env;
    std::vector<uint16_t> luma, chroma_out;
    struct Meta { size_t file_loc; int is_first; int phase_id; std::vector<int> dl, ds, de; };
    std::vector<Meta> meta;
    std::vector<ChunkTrace> traces;
};

// ---------------------------------------------------------------------------
// K3 line location, the reference's alternative GPU path selected by
// CUVHS_K2B_DISABLE (line_locs.cu:18-476; pipeline.cu:1694, 1848-1856).
// ---------------------------------------------------------------------------
struct K3Thresholds {
    double tol, hsync_min, hsync_max, eq_min, eq_max, vsync_min, vsync_max;
};

// line_locs host entry (line_locs.cu:491-506)
static K3Thresholds k3_thresholds(const Fmt& fmt) {
    K3Thresholds t;
    t.tol = 0.5e-6 * fmt.sample_rate;
    t.hsync_min = fmt.hsync_width - t.tol;
    t.hsync_max = fmt.hsync_width + t.tol;
    t.eq_min = fmt.eq_pulse_width - t.tol;
    t.eq_max = fmt.eq_pulse_width + t.tol;
    t.vsync_min = fmt.vsync_width * 0.5;
    t.vsync_max = fmt.vsync_width + 1.0e-6 * fmt.sample_rate;
    return t;
}

// Phase 1: nominal width classes (k_classify_pulses, line_locs.cu:18-48)
static int k3_nominal_type(const K3Thresholds& t, int length) {
    double len = (double)length;
    if (len >= t.vsync_min && len <= t.vsync_max) return PULSE_VSYNC;
    if (len >= t.hsync_min && len <= t.hsync_max) return PULSE_HSYNC;
    if (len >= t.eq_min && len <= t.eq_max) return PULSE_EQ1;
    return PULSE_HSYNC;
}

// Phases 2 and 3 for one field (k_compute_linelocs, line_locs.cu:60-476).
static void k3_field(const Fmt& fmt, const int* starts, const int* lengths, int npc, double* locs, int* is_first) {
    const double tol_samples = 0.5e-6 * fmt.sample_rate;
    const double hsync_nominal = fmt.hsync_width, eq_nominal = fmt.eq_pulse_width, vsync_nominal = fmt.vsync_width;
    const int lines_per_frame = fmt.lines_per_frame;
    const double inlinelen = (double)fmt.samples_per_line;
    // Step 0: adaptive classification (:96-137)
    double generous_hsync_min = hsync_nominal - 3.5 * tol_samples;
    double generous_hsync_max = hsync_nominal + 4.0 * tol_samples;
    double hsync_sum = 0.0;
    int hsync_count = 0;
    for (int i = 0; i < npc; i++) {
        double len = (double)lengths[i];
        if (len >= generous_hsync_min && len <= generous_hsync_max) {
            hsync_sum += len;
            hsync_count++;
        }
    }
    double hsync_median_est = (hsync_count > 0) ? (hsync_sum / hsync_count) : hsync_nominal;
    double hsync_offset = hsync_median_est - hsync_nominal;
    double a_hsync_min = hsync_nominal + hsync_offset - tol_samples;
    double a_hsync_max = hsync_nominal + hsync_offset + tol_samples;
    double a_eq_min = eq_nominal + hsync_offset - tol_samples;
    double a_eq_max = eq_nominal + hsync_offset + tol_samples;
    double a_vsync_min = (vsync_nominal + hsync_offset) * 0.5;
    double a_vsync_max = vsync_nominal + hsync_offset + 2.0 * tol_samples;
    const int n = npc < 800 ? npc : 800;
    std::vector<int> types((size_t)n);
    for (int i = 0; i < n; i++) {
        double len = (double)lengths[i];
        if (len >= a_vsync_min && len <= a_vsync_max) types[(size_t)i] = PULSE_VSYNC;
        else if (len >= a_hsync_min && len <= a_hsync_max) types[(size_t)i] = PULSE_HSYNC;
        else if (len >= a_eq_min && len <= a_eq_max) types[(size_t)i] = PULSE_EQ1;
        else types[(size_t)i] = PULSE_HSYNC;
    }
    // Step 1: vblank state machine (:139-196)
    int ref_pulse_idx = -1, last_hsync_idx = -1, first_eq1_idx = -1, last_eq2_idx = -1, state = -1;
    for (int i = 0; i < n; i++) {
        int t = types[(size_t)i];
        if (state == -1) {
            if (t == PULSE_HSYNC) { state = PULSE_HSYNC; last_hsync_idx = i; }
        } else if (state == PULSE_HSYNC) {
            if (t == PULSE_HSYNC) last_hsync_idx = i;
            else if (t == PULSE_EQ1) { state = PULSE_EQ1; first_eq1_idx = i; }
            else if (t == PULSE_VSYNC) state = PULSE_VSYNC;
        } else if (state == PULSE_EQ1) {
            if (t == PULSE_VSYNC) state = PULSE_VSYNC;
            else if (t == PULSE_HSYNC) { state = PULSE_HSYNC; last_hsync_idx = i; first_eq1_idx = -1; }
        } else if (state == PULSE_VSYNC) {
            if (t == PULSE_EQ1) { state = PULSE_EQ2; last_eq2_idx = i; }
            else if (t == PULSE_HSYNC) { ref_pulse_idx = i; break; }
        } else if (state == PULSE_EQ2) {
            if (t == PULSE_EQ1) last_eq2_idx = i;
            else if (t == PULSE_HSYNC) {
                if (last_eq2_idx < 0) last_eq2_idx = i - 1;
                ref_pulse_idx = i;
                break;
            }
        }
    }
    // Step 1b: field parity (:198-235)
    const bool ntsc = fmt.system != 1;
    int field_parity = -1;
    if (ref_pulse_idx >= 0 && last_hsync_idx >= 0 && first_eq1_idx >= 0) {
        double entry_lines = (double)(starts[first_eq1_idx] - starts[last_hsync_idx]) / inlinelen;
        if (ntsc) field_parity = (entry_lines > 0.75) ? 1 : 0;
        else field_parity = (entry_lines > 0.75) ? 0 : 1;
    }
    if (field_parity < 0 && ref_pulse_idx >= 0 && last_eq2_idx >= 0) {
        double exit_lines = (double)(starts[ref_pulse_idx] - starts[last_eq2_idx]) / inlinelen;
        if (ntsc) field_parity = (exit_lines > 0.75) ? 1 : 0;
        else field_parity = (exit_lines > 0.75) ? 0 : 1;
    }
    *is_first = field_parity;
    // Step 2: mean line length from the longest HSYNC run (:231-277)
    int best_run_start = -1, best_run_len = 0, cur_run_start = -1, cur_run_len = 0;
    for (int i = 0; i < n; i++) {
        if (types[(size_t)i] == PULSE_HSYNC) {
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
        double sum = 0.0;
        int count = 0;
        for (int i = best_run_start + 1; i < best_run_start + best_run_len; i++) {
            double spacing = (double)(starts[i] - starts[i - 1]);
            if (spacing >= inlinelen * 0.95 && spacing <= inlinelen * 1.05) { sum += spacing; count++; }
        }
        if (count > 0) meanlinelen = sum / count;
    }
    // Step 3: reference pulse (:279-324)
    double ref_position, ref_line;
    if (ref_pulse_idx >= 0) {
        ref_position = (double)starts[ref_pulse_idx];
        ref_line = 10.0;
        if (ref_pulse_idx > n / 2) {
            double corrected = ref_position - (double)fmt.samples_per_field;
            if (corrected > 0.0) ref_position = corrected;
        }
    } else {
        ref_pulse_idx = 0;
        for (int i = 0; i < n; i++)
            if (types[(size_t)i] == PULSE_HSYNC) { ref_pulse_idx = i; break; }
        ref_position = (double)starts[ref_pulse_idx];
        ref_line = 0.0;
    }
    // Step 4a: line markers (:346-387)
    std::vector<int> ms, ml;
    double tiny_hsync_max = hsync_nominal * 0.75;
    double surrogate_gap_max = inlinelen * 0.25;
    for (int i = 0; i < n && (int)ms.size() < 800; i++) {
        int ti = types[(size_t)i];
        bool surrogate_eq = (ti == PULSE_EQ1 || ti == PULSE_EQ2) && i + 1 < n && types[(size_t)i + 1] == PULSE_HSYNC &&
                            (double)lengths[i + 1] <= tiny_hsync_max &&
                            (double)(starts[i + 1] - starts[i]) <= surrogate_gap_max;
        bool suppress = ti == PULSE_HSYNC && (double)lengths[i] <= tiny_hsync_max && i > 0 &&
                        (types[(size_t)i - 1] == PULSE_EQ1 || types[(size_t)i - 1] == PULSE_EQ2) &&
                        (double)(starts[i] - starts[i - 1]) <= surrogate_gap_max;
        if (surrogate_eq || (ti == PULSE_HSYNC && !suppress)) {
            ms.push_back(starts[i]);
            ml.push_back(lengths[i]);
        }
    }
    // Step 4b: collapse near-duplicate markers (:389-433)
    std::vector<int> hs;
    const int marker_count = (int)ms.size();
    double duplicate_gap_max = inlinelen * 0.75;
    int cluster_start = 0;
    while (cluster_start < marker_count && (int)hs.size() < 800) {
        int best_idx = cluster_start;
        double best_score = fabs((double)ml[(size_t)cluster_start] - hsync_nominal);
        int cluster_end = cluster_start + 1;
        while (cluster_end < marker_count &&
               (double)(ms[(size_t)cluster_end] - ms[(size_t)cluster_start]) < duplicate_gap_max) {
            double score = fabs((double)ml[(size_t)cluster_end] - hsync_nominal);
            if (score < best_score || (score == best_score && ms[(size_t)cluster_end] < ms[(size_t)best_idx])) {
                best_idx = cluster_end;
                best_score = score;
            }
            cluster_end++;
        }
        hs.push_back(ms[(size_t)best_idx]);
        cluster_start = cluster_end;
    }
    // Step 4c: greedy nearest-marker snapping (:437-476)
    const int num_hsyncs = (int)hs.size();
    double max_allowed_distance = meanlinelen / 1.5;
    int cur_hsync = 0;
    for (int line = 0; line < lines_per_frame; line++) {
        double expected = ref_position + meanlinelen * (line - ref_line);
        locs[line] = expected;
        if (cur_hsync >= num_hsyncs) continue;
        double best_dist = max_allowed_distance;
        int best_pos = -1;
        int search = cur_hsync;
        while (search < num_hsyncs) {
            double dist = fabs((double)hs[(size_t)search] - expected);
            if (dist <= best_dist) {
                best_dist = dist;
                best_pos = hs[(size_t)search];
                cur_hsync = search;
            }
            if (search + 1 < num_hsyncs) {
                double next_dist = fabs((double)hs[(size_t)search + 1] - expected);
                if (next_dist > dist) break;
            } else {
                break;
            }
            search++;
        }
        if (best_pos >= 0) {
            locs[line] = (double)best_pos;
            cur_hsync++;
        }
    }
}

// Diagnostic counters of the reference (K3: line_locs.cu:585-653, K4:
// hsync_refine.cu:273-349, K5: tbc_resample.cu:75-116), over nf fields.
static void k3_spacing_counts(const Fmt& fmt, const double* ll, int nf, int* out) {
    const int L = fmt.lines_per_frame, als = fmt.active_line_start;
    const double nominal = (double)fmt.samples_per_line;
    for (int f = 0; f < nf; f++) {
        const double* l = ll + (size_t)f * L;
        for (int line = als; line + 1 < L; line++) {
            double spacing = l[line + 1] - l[line];
            if (spacing < 0.9 * nominal || spacing > 1.1 * nominal) out[0]++;
            if (fabs(spacing - nominal) > 0.25 * nominal) out[2]++;
            if (line > als && line + 2 < L) {
                double avg = 0.5 * ((l[line] - l[line - 1]) + (l[line + 2] - l[line + 1]));
                if (fabs(spacing - avg) > 0.15 * nominal) out[1]++;
            }
        }
    }
}

static void k4_counts(const Fmt& fmt, const double* d05, int total, const double* before, const double* after, int nf,
                      int* out) {
    const int L = fmt.lines_per_frame, als = fmt.active_line_start;
    const double us = 1.0e-6 * fmt.sample_rate;
    for (int f = 0; f < nf; f++) {
        const double* b = before + (size_t)f * L;
        const double* a = after + (size_t)f * L;
        for (int line = als; line + 1 < L; line++) {
            double delta = a[line] - b[line];
            if (fabs(delta) > 1.5 * us) out[0]++;
            if (line > als) {
                double avg = 0.5 * ((a[line - 1] - b[line - 1]) + (a[line + 1] - b[line + 1]));
                if (fabs(delta - avg) > 1.0 * us) out[1]++;
            }
            int sample = (int)(a[line] + 0.5);
            if (sample >= 0 && sample < total && d05[sample] <= fmt.pulse_threshold_hz) out[2]++;
        }
    }
}

static void k5_counts(const Fmt& fmt, const double* demod, int total, const double* ll, int nf, int* out) {
    const int L = fmt.lines_per_frame;
    const double spl = (double)fmt.samples_per_line;
    for (int f = 0; f < nf; f++) {
        for (int row = 0; row < fmt.output_field_lines; row++) {
            double ls = ll[(size_t)f * L + row], le = ll[(size_t)f * L + row + 1];
            double len = le - ls;
            if (len <= 0.0 || len < 0.8 * spl || len > 1.2 * spl) out[1]++;
            for (int col = 0; col < fmt.output_line_len; col++) {
                double coord = ls + ((double)col / (double)fmt.output_line_len) * len;
                int ci = (int)coord;
                double x = coord - (double)ci;
                if (ci < 1 || ci + 2 >= total) { out[0]++; continue; }
                double p0 = demod[ci - 1], p1 = demod[ci], p2 = demod[ci + 1], p3 = demod[ci + 2];
                double a = p2 - p0, bb = 2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3, c = 3.0 * (p1 - p2) + p3 - p0;
                double hz = p1 + 0.5 * x * (a + x * (bb + x * c));
                if (col >= active_video_start(fmt) && col < active_video_end(fmt) && hz <= fmt.pulse_threshold_hz)
                    out[2]++;
            }
        }
    }
}

// RawReader::convert (raw_reader.cpp:71-100): bit-depth normalisation to the
// s16 range, then the optional per-read DC removal (sequential mean of the
// samples of this read).
extern "C" int or_convert(int format, int dc_correct, const uint8_t* data, size_t n, double* out) {
    if (format < 0 || format > 2) return -1;
    for (size_t i = 0; i < n; i++) {
        if (format == 0) {
            out[i] = (static_cast<double>(data[i]) - 128.0) * 256.0;
        } else if (format == 1) {
            int16_t v;
            memcpy(&v, data + 2 * i, 2);
            out[i] = static_cast<double>(v);
        } else {
            uint16_t v;
            memcpy(&v, data + 2 * i, 2);
            out[i] = static_cast<double>(v) - 32768.0;
        }
    }
    if (dc_correct && n > 0) {
        double sum = 0.0;
        for (size_t i = 0; i < n; i++) sum += out[i];
        double mean = sum / static_cast<double>(n);
        for (size_t i = 0; i < n; i++) out[i] -= mean;
    }
    return 0;
}

static int process_chunk(or_dec* d, const uint8_t* data, size_t total_n, size_t raw_offset, int num_fields,
                         size_t& next_raw_offset) {
    const Fmt& fmt = d->fmt;
    size_t spf = (size_t)fmt.samples_per_field;
    size_t spfp = (size_t)fmt.spfp;
    size_t tbc = (size_t)fmt.output_line_len * fmt.output_field_lines;
    size_t total_samples = (size_t)num_fields * spfp;
    // RawReader::read_at (raw_reader.cpp:128-152)
    size_t n_read = 0;
    if (raw_offset < total_n) n_read = std::min(total_samples, total_n - raw_offset);
    if (n_read < spf) return 0;
    double* h_raw = d->raw.data();
    StageClock clk(d->stage_s);
    or_convert(d->in_fmt, d->dc_correct ? 1 : 0, data + (d->in_fmt == 0 ? 1 : 2) * raw_offset, n_read, h_raw);
    for (size_t i = n_read; i < total_samples; i++) h_raw[i] = 0.0;
    clk.lap(0);
    int fields_loaded = (int)(n_read / spf);
    if (fields_loaded > num_fields) fields_loaded = num_fields;
    if (fields_loaded == 0) return 0;
    ChunkTrace tr;
    tr.raw_offset = raw_offset;
    tr.fields_k1 = fields_loaded;

    // K1 over fields_loaded segments; the rest of the buffers is zero
    std::fill(d->demod.begin(), d->demod.end(), 0.0);
    std::fill(d->d05.begin(), d->d05.end(), 0.0);
    std::fill(d->env.begin(), d->env.end(), 0.0);
    fm_demod(fmt, d->k1, h_raw, fields_loaded, d->demod.data(), d->d05.data(), d->env.data());
    clk.lap(1);

    // VSYNC discovery + clustering (pipeline.cu:1567-1651)
    int total_chunk = fields_loaded * (int)spfp;
    int cap = fields_loaded * 20;
    std::vector<int> cand;
    int num_candidates = vsync_discover(fmt, d->d05.data(), total_chunk, cand, cap);
    tr.candidates = num_candidates;
    if (num_candidates > cap)
        fprintf(stderr, "Warning: VSYNC candidate buffer overflow (%d > %d), truncating chunk results\n",
                num_candidates, cap);
    std::sort(cand.begin(), cand.end());
    std::vector<size_t> offs;
    for (int pos : cand) {
        if (offs.empty() || ((size_t)pos - offs.back() > (size_t)(fmt.samples_per_field / 2))) {
            size_t safe = (pos > (5 * fmt.samples_per_line)) ? (size_t)(pos - 5 * fmt.samples_per_line) : 0;
            offs.push_back(safe);
        }
    }
    if (offs.empty()) {
        fprintf(stderr, "Warning: no VSYNC candidates found in chunk at raw offset %zu, using nominal field spacing fallback\n",
                raw_offset);
        for (int i = 0; i < fields_loaded; i++) offs.push_back((size_t)i * spf);
    }
    fields_loaded = (int)offs.size();
    if (fields_loaded > num_fields) {
        fields_loaded = num_fields;
        offs.resize((size_t)fields_loaded);
    }
    size_t actual_spf = spf;
    if (offs.size() >= 2) actual_spf = (offs.back() - offs.front()) / (offs.size() - 1);
    size_t predicted = offs.back() + actual_spf;
    size_t margin = 10 * (size_t)fmt.samples_per_line;
    next_raw_offset = predicted > margin ? raw_offset + predicted - margin : raw_offset + predicted;
    const int nf = fields_loaded;
    clk.lap(2);

    // K2
    std::vector<int> ps((size_t)nf * MAX_PULSES, 0), pl((size_t)nf * MAX_PULSES, 0), pc((size_t)nf, 0);
    std::vector<int> pt((size_t)nf * MAX_PULSES, 0);
    sync_pulses(fmt, d->d05.data(), offs.data(), nf, ps.data(), pl.data(), pc.data());
    clk.lap(3);
    if (d->trace) { tr.pc = pc; tr.ps = ps; tr.pl = pl; tr.offsets = offs; }
    // K2b, or K3 when selected (pipeline.cu:1694-1856)
    std::vector<double> ll((size_t)nf * fmt.lines_per_frame);
    std::vector<int> is_first((size_t)nf);
    if (d->line_locator == 0) {
        k2b_chunk(fmt, d->k2b, raw_offset, offs, nf, ps.data(), pl.data(), pc.data(), pt.data(), ll.data(),
                  is_first.data());
    } else {
        K3Thresholds t3 = k3_thresholds(fmt);
        for (int f = 0; f < nf; f++) {
            for (int i = 0; i < pc[(size_t)f]; i++)
                pt[(size_t)f * MAX_PULSES + i] = k3_nominal_type(t3, pl[(size_t)f * MAX_PULSES + i]);
            k3_field(fmt, ps.data() + (size_t)f * MAX_PULSES, pl.data() + (size_t)f * MAX_PULSES, pc[(size_t)f],
                     ll.data() + (size_t)f * fmt.lines_per_frame, &is_first[(size_t)f]);
        }
    }
    if (d->trace) { tr.vc = pc; tr.vs = ps; tr.vl = pl; tr.vt = pt; tr.llc = ll; tr.is_first = is_first; }
    clk.lap(4);
    std::vector<int> dg(9, 0);
    if (d->diag) k3_spacing_counts(fmt, ll.data(), nf, dg.data());
    // K4
    int total = nf * (int)spfp;
    std::vector<double> coarse;
    if (d->diag) coarse = ll;
    hsync_refine(fmt, d->d05.data(), ll.data(), nf, total);
    clk.lap(5);
    if (d->trace) tr.llr = ll;
    if (d->diag) k4_counts(fmt, d->d05.data(), total, coarse.data(), ll.data(), nf, dg.data() + 3);
    // K5
    std::vector<uint16_t> luma((size_t)nf * tbc), chroma((size_t)nf * tbc);
    tbc_resample(fmt, d->demod.data(), ll.data(), luma.data(), nf, total);
    clk.lap(6);
    if (d->diag) {
        k5_counts(fmt, d->demod.data(), total, ll.data(), nf, dg.data() + 6);
        d->diag_counts.push_back(dg);
    }
    // K6 (bandpassed raw overwrites demod, pipeline.cu:2531)
    std::vector<int> phase_ids;
    chroma_decode(fmt, h_raw, ll.data(), d->demod.data(), chroma.data(), nf, total, phase_ids, &d->chroma);
    clk.lap(7);
    // K7
    std::vector<int> dl((size_t)nf * MAX_DO), dsx((size_t)nf * MAX_DO), dex((size_t)nf * MAX_DO), dc((size_t)nf);
    dropout_detect(fmt, d->env.data(), ll.data(), luma.data(), chroma.data(), dl.data(), dsx.data(), dex.data(),
                   dc.data(), nf);
    clk.lap(8);
    // output assembly (pipeline.cu:2575-2610)
    for (int i = 0; i < nf; i++) {
        or_dec::Meta m;
        m.file_loc = raw_offset + offs[(size_t)i];
        bool first_field;
        if (is_first[(size_t)i] >= 0 && (is_first[(size_t)i] == 1) != d->last_parity) first_field = (is_first[(size_t)i] == 1);
        else first_field = !d->last_parity;
        d->last_parity = first_field;
        m.is_first = first_field ? 1 : 0;
        m.phase_id = (i < (int)phase_ids.size()) ? phase_ids[(size_t)i] : 0;
        for (int k = 0; k < dc[(size_t)i]; k++) {
            m.dl.push_back(dl[(size_t)i * MAX_DO + k]);
            m.ds.push_back(dsx[(size_t)i * MAX_DO + k]);
            m.de.push_back(dex[(size_t)i * MAX_DO + k]);
        }
        d->meta.push_back(std::move(m));
    }
    d->luma.insert(d->luma.end(), luma.begin(), luma.end());
    d->chroma_out.insert(d->chroma_out.end(), chroma.begin(), chroma.end());
    if (d->trace) {
        tr.nf = nf;
        tr.next_offset = next_raw_offset;
        d->traces.push_back(std::move(tr));
    }
    return nf;
}

static int write_json_meta(const char* path, const std::vector<or_dec::Meta>& meta, const Fmt& fmt);

extern "C" {

int or_format_init(double mhz, or_format* o) { return or_format_init_sys(0, mhz, o); }

int or_format_init_sys(int system, double mhz, or_format* o) {
    if (!o || !(mhz > 0.0) || system < 0 || system > 1) return -1;
    Fmt f = make_fmt(mhz, system);
    o->sample_rate = f.sample_rate;
    o->samples_per_line = f.samples_per_line;
    o->samples_per_field = f.samples_per_field;
    o->spfp = f.spfp;
    o->k1_fft_size = orc::next_7smooth(f.spfp + OVERLAP);
    o->chroma_fft_size = orc::next_7smooth(f.spfp);
    o->hsync_width = f.hsync_width;
    o->eq_pulse_width = f.eq_pulse_width;
    o->vsync_width = f.vsync_width;
    o->ire0 = f.ire0;
    o->hz_ire = f.hz_ire;
    o->vsync_ire = f.vsync_ire;
    o->sync_tip_hz = f.sync_tip_hz;
    o->pulse_threshold_hz = f.pulse_threshold_hz;
    o->output_zero = f.output_zero;
    o->output_scale = f.output_scale;
    o->output_rate = f.output_rate;
    o->fsc = f.fsc;
    o->chroma_under = f.chroma_under;
    o->burst_abs_ref = f.burst_abs_ref;
    o->black16b = (0.0 - f.vsync_ire) * f.output_scale + f.output_zero;
    o->white16b = (100.0 - f.vsync_ire) * f.output_scale + f.output_zero;
    o->burst_start = (int)(f.burst_start_us * 1e-6 * f.output_rate + 0.5);
    o->burst_end = (int)(f.burst_end_us * 1e-6 * f.output_rate + 0.5);
    return 0;
}

int or_filter_table(double mhz, int which, double* out) {
    Fmt f = make_fmt(mhz);
    if (which == 3) {
        int n = 0;
        std::vector<double> h = make_chroma_prebpf(f, f.spfp, &n);
        if (out) std::copy(h.begin(), h.end(), out);
        return (int)h.size();
    }
    K1Filters K = make_k1_filters(f);
    if (which == 0) {
        if (out) std::copy(K.rf.begin(), K.rf.end(), out);
        return K.fb;
    }
    const std::vector<cd>& v = which == 1 ? K.fv : K.fv05;
    if (out)
        for (int k = 0; k < K.fb; k++) { out[2 * k] = v[(size_t)k].real(); out[2 * k + 1] = v[(size_t)k].imag(); }
    return K.fb;
}

int or_fft(int n, const double* in, double* out, int sign) {
    if (!orc::is_7smooth(n)) return -1;
    FFTPlan p(n);
    p.exec(reinterpret_cast<const cd*>(in), reinterpret_cast<cd*>(out), sign);
    return 0;
}

int or_fm_demod(double mhz, const double* raw, int nfields, double* demod, double* demod05, double* env) {
    Fmt f = make_fmt(mhz);
    K1Filters K = make_k1_filters(f);
    fm_demod(f, K, raw, nfields, demod, demod05, env);
    return 0;
}

int or_vsync_discover(double mhz, const double* d05, int total, int* idx, int cap) {
    Fmt f = make_fmt(mhz);
    std::vector<int> v;
    int c = vsync_discover(f, d05, total, v, cap);
    std::copy(v.begin(), v.end(), idx);
    return c;
}

int or_sync_pulses(double mhz, const double* d05, const uint64_t* offsets, int nfields, int* starts, int* lens,
                   int* count) {
    Fmt f = make_fmt(mhz);
    std::vector<size_t> o(offsets, offsets + nfields);
    sync_pulses(f, d05, o.data(), nfields, starts, lens, count);
    return 0;
}

int or_hsync_refine(double mhz, const double* d05, int total, double* linelocs, int nfields) {
    Fmt f = make_fmt(mhz);
    hsync_refine(f, d05, linelocs, nfields, total);
    return 0;
}

int or_tbc_resample(double mhz, const double* demod, int total, const double* linelocs, int nfields,
                    uint16_t* luma) {
    Fmt f = make_fmt(mhz);
    tbc_resample(f, demod, linelocs, luma, nfields, total);
    return 0;
}

or_dec* or_create(double mhz, int batch, int record_trace) { return or_create_sys(0, mhz, batch, record_trace); }

or_dec* or_create_sys(int system, double mhz, int batch, int record_trace) {
    if (!(mhz > 0.0) || batch < 1 || system < 0 || system > 1) return nullptr;
    or_dec* d = new or_dec();
    d->fmt = make_fmt(mhz, system);
    d->batch = batch;
    d->trace = record_trace != 0;
    d->chroma.keep_trial = d->trace;
    d->k1 = make_k1_filters(d->fmt);
    size_t n = (size_t)batch * d->fmt.spfp;
    // one segment of zero padding: a K2 window may start in the last segment
    size_t np = n + (size_t)d->fmt.spfp;
    d->raw.assign(n, 0.0);
    d->demod.assign(np, 0.0);
    d->d05.assign(np, 0.0);
    d->env.assign(np, 0.0);
    return d;
}

void or_destroy(or_dec* d) { delete d; }

int or_set_line_locator(or_dec* d, int mode, int diagnostics) {
    if (!d || mode < 0 || mode > 1) return -1;
    d->line_locator = mode;
    d->diag = diagnostics != 0;
    return 0;
}

int or_chunk_diagnostics(const or_dec* d, int c, int* out9) {
    if (!d || c < 0 || c >= (int)d->diag_counts.size()) return -1;
    for (int i = 0; i < 9; i++) out9[i] = d->diag_counts[(size_t)c][(size_t)i];
    return 0;
}

int or_k3_linelocs(double mhz, const int* starts, const int* lens, const int* count, int nf, int* types,
                   double* linelocs, int* is_first) {
    return or_k3_linelocs_sys(0, mhz, starts, lens, count, nf, types, linelocs, is_first);
}

int or_k3_linelocs_sys(int system, double mhz, const int* starts, const int* lens, const int* count, int nf,
                       int* types, double* linelocs, int* is_first) {
    if (system < 0 || system > 1) return -1;
    Fmt fmt = make_fmt(mhz, system);
    K3Thresholds t3 = k3_thresholds(fmt);
    for (int f = 0; f < nf; f++) {
        const int* st = starts + (size_t)f * MAX_PULSES;
        const int* ln = lens + (size_t)f * MAX_PULSES;
        if (types)
            for (int i = 0; i < count[f]; i++) types[(size_t)f * MAX_PULSES + i] = k3_nominal_type(t3, ln[i]);
        k3_field(fmt, st, ln, count[f], linelocs + (size_t)f * fmt.lines_per_frame, &is_first[f]);
    }
    return 0;
}

int or_diag_counts(double mhz, int which, const double* a, const double* b, const double* c, int total, int nf,
                   int* out3) {
    Fmt fmt = make_fmt(mhz);
    out3[0] = out3[1] = out3[2] = 0;
    if (which == 3) k3_spacing_counts(fmt, a, nf, out3);
    else if (which == 4) k4_counts(fmt, a, total, b, c, nf, out3);
    else if (which == 5) k5_counts(fmt, a, total, b, nf, out3);
    else return -1;
    return 0;
}

int or_set_input(or_dec* d, int format, int dc_correct) {
    if (!d || format < 0 || format > 2) return -1;
    d->in_fmt = format;
    d->dc_correct = dc_correct != 0;
    return 0;
}

int or_decode(or_dec* d, const uint8_t* data, size_t n) {
    size_t raw_offset = 0;
    int total_fields = 0;
    const size_t spf = (size_t)d->fmt.samples_per_field;
    while (true) {
        if (raw_offset >= n) break;  // read_at returns 0 past EOF (raw_reader.cpp:138)
        int remaining_fields = (int)((n - raw_offset) / spf);
        if (remaining_fields <= 0) break;
        int this_chunk = std::min(d->batch, remaining_fields);
        size_t next = 0;
        int processed = process_chunk(d, data, n, raw_offset, this_chunk, next);
        if (processed == 0) break;
        raw_offset = next;
        total_fields += processed;
    }
    return total_fields;
}

int or_num_fields(const or_dec* d) { return (int)d->meta.size(); }
const uint16_t* or_luma(const or_dec* d) { return d->luma.data(); }
const uint16_t* or_chroma(const or_dec* d) { return d->chroma_out.data(); }

void or_field_meta(const or_dec* d, uint64_t* file_loc, int* is_first, int* phase_id, int* do_count) {
    for (size_t i = 0; i < d->meta.size(); i++) {
        if (file_loc) file_loc[i] = d->meta[i].file_loc;
        if (is_first) is_first[i] = d->meta[i].is_first;
        if (phase_id) phase_id[i] = d->meta[i].phase_id;
        if (do_count) do_count[i] = (int)d->meta[i].dl.size();
    }
}

int or_field_dropouts(const or_dec* d, int field, int* lines, int* startx, int* endx) {
    if (field < 0 || field >= (int)d->meta.size()) return -1;
    const auto& m = d->meta[(size_t)field];
    for (size_t k = 0; k < m.dl.size(); k++) {
        if (lines) lines[k] = m.dl[k];
        if (startx) startx[k] = m.ds[k];
        if (endx) endx[k] = m.de[k];
    }
    return (int)m.dl.size();
}

int or_num_chunks(const or_dec* d) { return (int)d->traces.size(); }

int or_chroma_trial(const or_dec* d, double* metrics, int* track) {
    if (!d->chroma.valid) return -1;
    metrics[0] = d->chroma.trial_metric[0];
    metrics[1] = d->chroma.trial_metric[1];
    *track = metrics[1] < metrics[0] ? 1 : 0;
    return 0;
}

int or_chroma_trial_lines(const or_dec* d, int which, double* out) {
    if (which < 0 || which > 2) return -1;
    const std::vector<double>& b = d->chroma.trial_buf[which];
    if (b.empty()) return -1;
    if (out) std::copy(b.begin(), b.end(), out);
    return (int)b.size();
}

int or_k2b_chunk(double mhz, or_k2b_state* st, uint64_t raw_offset, const uint64_t* offsets, int nf, int* starts,
                 int* lens, int* count, int* types, unsigned char* good, double* linelocs, int* is_first,
                 double* anchor) {
    return or_k2b_chunk_sys(0, mhz, st, raw_offset, offsets, nf, starts, lens, count, types, good, linelocs, is_first,
                            anchor);
}

int or_k2b_chunk_sys(int system, double mhz, or_k2b_state* st, uint64_t raw_offset, const uint64_t* offsets, int nf,
                     int* starts, int* lens, int* count, int* types, unsigned char* good, double* linelocs,
                     int* is_first, double* anchor) {
    if (!st || nf < 0 || !(mhz > 0.0) || system < 0 || system > 1) return -1;
    Fmt fmt = make_fmt(mhz, system);
    K2bState k;
    k.prev_first_field = st->prev_first_field;
    k.prev_progressive_field = st->prev_progressive_field;
    k.prev_first_hsync_readloc = st->prev_first_hsync_readloc;
    k.prev_first_hsync_loc = st->prev_first_hsync_loc;
    k.prev_first_hsync_diff = st->prev_first_hsync_diff;
    std::vector<size_t> offs((size_t)nf);
    for (int i = 0; i < nf; i++) offs[(size_t)i] = (size_t)offsets[i];
    k2b_chunk(fmt, k, (size_t)raw_offset, offs, nf, starts, lens, count, types, linelocs, is_first, good, anchor);
    st->prev_first_field = k.prev_first_field;
    st->prev_progressive_field = k.prev_progressive_field;
    st->prev_first_hsync_readloc = k.prev_first_hsync_readloc;
    st->prev_first_hsync_loc = k.prev_first_hsync_loc;
    st->prev_first_hsync_diff = k.prev_first_hsync_diff;
    return 0;
}

int or_lookup_ntsc_phase(int is_first, int quadrant, int delta) { return lookup_ntsc_phase_id(is_first, quadrant, delta); }

double or_burst_metric(double mhz, const double* lines) { return or_burst_metric_sys(0, mhz, lines); }

double or_burst_phase(double mhz, const double* lines) { return or_burst_phase_sys(0, mhz, lines); }

double or_burst_metric_sys(int system, double mhz, const double* lines) {
    Fmt fmt = make_fmt(mhz, system == 1 ? 1 : 0);
    return burst_metric(lines, fmt, chroma_consts(fmt));
}

double or_burst_phase_sys(int system, double mhz, const double* lines) {
    Fmt fmt = make_fmt(mhz, system == 1 ? 1 : 0);
    return burst_phase(lines, fmt, chroma_consts(fmt));
}

int or_stage_seconds(const or_dec* d, double* out) {
    for (int i = 0; i < kOrStages; i++) out[i] = d->stage_s[i];
    return kOrStages;
}

int or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int or_trace_candidates(const or_dec* d, int c) {
    if (c < 0 || c >= (int)d->traces.size()) return -1;
    return d->traces[(size_t)c].candidates;
}

int or_trace_chunk(const or_dec* d, int c, uint64_t* raw_offset, int* fields_k1, uint64_t* next_offset,
                   uint64_t* offsets, int* pulse_count, int* pulse_starts, int* pulse_lens, int* vpulse_count,
                   int* vpulse_starts, int* vpulse_lens, int* vpulse_types, double* ll_coarse,
                   double* ll_refined, int* is_first_k2b) {
    if (c < 0 || c >= (int)d->traces.size()) return -1;
    const ChunkTrace& t = d->traces[(size_t)c];
    if (raw_offset) *raw_offset = t.raw_offset;
    if (fields_k1) *fields_k1 = t.fields_k1;
    if (next_offset) *next_offset = t.next_offset;
    if (offsets) for (int i = 0; i < t.nf; i++) offsets[i] = t.offsets[(size_t)i];
    if (pulse_count) std::copy(t.pc.begin(), t.pc.end(), pulse_count);
    if (pulse_starts) std::copy(t.ps.begin(), t.ps.end(), pulse_starts);
    if (pulse_lens) std::copy(t.pl.begin(), t.pl.end(), pulse_lens);
    if (vpulse_count) std::copy(t.vc.begin(), t.vc.end(), vpulse_count);
    if (vpulse_starts) std::copy(t.vs.begin(), t.vs.end(), vpulse_starts);
    if (vpulse_lens) std::copy(t.vl.begin(), t.vl.end(), vpulse_lens);
    if (vpulse_types) std::copy(t.vt.begin(), t.vt.end(), vpulse_types);
    if (ll_coarse) std::copy(t.llc.begin(), t.llc.end(), ll_coarse);
    if (ll_refined) std::copy(t.llr.begin(), t.llr.end(), ll_refined);
    if (is_first_k2b) std::copy(t.is_first.begin(), t.is_first.end(), is_first_k2b);
    return t.nf;
}

// TBCWriter::write_json (tbc_writer.cpp:89-175) + raw TBC streams (:55-65)
int or_write_outputs(const or_dec* d, const char* out_base) {
    std::string base(out_base);
    const Fmt& fmt = d->fmt;
    size_t tbc = (size_t)fmt.output_line_len * fmt.output_field_lines;
    FILE* fl = fopen((base + ".tbc").c_str(), "wb");
    FILE* fc = fopen((base + "_chroma.tbc").c_str(), "wb");
    if (!fl || !fc) { if (fl) fclose(fl); if (fc) fclose(fc); return -1; }
    fwrite(d->luma.data(), sizeof(uint16_t), d->meta.size() * tbc, fl);
    fwrite(d->chroma_out.data(), sizeof(uint16_t), d->meta.size() * tbc, fc);
    fclose(fl);
    fclose(fc);
    return write_json_meta((base + ".tbc.json").c_str(), d->meta, fmt);
}

int or_write_json_meta(double mhz, const char* path, int n, const uint64_t* file_loc, const int* is_first,
                       const int* phase_id, const int* do_count, const int* do_line, const int* do_startx,
                       const int* do_endx) {
    return or_write_json_meta_sys(0, mhz, path, n, file_loc, is_first, phase_id, do_count, do_line, do_startx,
                                  do_endx);
}

int or_write_json_meta_sys(int system, double mhz, const char* path, int n, const uint64_t* file_loc,
                           const int* is_first, const int* phase_id, const int* do_count, const int* do_line,
                           const int* do_startx, const int* do_endx) {
    if (system < 0 || system > 1) return -1;
    std::vector<or_dec::Meta> meta((size_t)n);
    int k = 0;
    for (int i = 0; i < n; i++) {
        meta[(size_t)i].file_loc = file_loc[i];
        meta[(size_t)i].is_first = is_first[i];
        meta[(size_t)i].phase_id = phase_id[i];
        for (int j = 0; j < do_count[i]; j++, k++) {
            meta[(size_t)i].dl.push_back(do_line[k]);
            meta[(size_t)i].ds.push_back(do_startx[k]);
            meta[(size_t)i].de.push_back(do_endx[k]);
        }
    }
    return write_json_meta(path, meta, make_fmt(mhz, system));
}

int or_decode_file(const char* in_path, const char* out_base, double mhz, int batch) {
    return or_decode_file_fmt(in_path, out_base, mhz, batch, 0, 0);
}

int or_decode_file_fmt(const char* in_path, const char* out_base, double mhz, int batch, int format, int dc_correct) {
    FILE* f = fopen(in_path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<uint8_t> buf((size_t)std::max(0L, sz));
    size_t got = fread(buf.data(), 1, buf.size(), f);
    fclose(f);
    or_dec* d = or_create(mhz, batch, 0);
    if (!d) return -1;
    if (or_set_input(d, format, dc_correct) != 0) {
        or_destroy(d);
        return -1;
    }
    or_decode(d, buf.data(), got / (format == 0 ? 1 : 2));
    int rc = or_write_outputs(d, out_base);
    int n = or_num_fields(d);
    or_destroy(d);
    return rc < 0 ? rc : n;
}

}  // extern "C"

// TBCWriter::write_json (tbc_writer.cpp:89-175)
static int write_json_meta(const char* path, const std::vector<or_dec::Meta>& meta, const Fmt& fmt) {
    FILE* fp = fopen(path, "w");
    if (!fp) return -1;
    double black16b = (0.0 - fmt.vsync_ire) * fmt.output_scale + fmt.output_zero;
    double white16b = (100.0 - fmt.vsync_ire) * fmt.output_scale + fmt.output_zero;
    int burst_start = (int)(fmt.burst_start_us * 1e-6 * fmt.output_rate + 0.5);
    int burst_end = (int)(fmt.burst_end_us * 1e-6 * fmt.output_rate + 0.5);
    int n = (int)meta.size();
    fprintf(fp, "{\n  \"videoParameters\": {\n");
    const bool pal = fmt.system == 1;
    fprintf(fp, "    \"system\": \"%s\",\n    \"isSubcarrierLocked\": false,\n    \"isSourcePal\": %s,\n",
            pal ? "PAL" : "NTSC", pal ? "true" : "false");
    fprintf(fp, "    \"numberOfSequentialFields\": %d,\n", n);
    fprintf(fp, "    \"black16bIre\": %.1f,\n", black16b);
    fprintf(fp, "    \"white16bIre\": %.1f,\n", white16b);
    fprintf(fp, "    \"sampleRate\": %.0f,\n", fmt.output_rate);
    fprintf(fp, "    \"fieldWidth\": %d,\n", fmt.output_line_len);
    fprintf(fp, "    \"fieldHeight\": %d,\n", fmt.output_field_lines);
    fprintf(fp, "    \"colourBurstStart\": %d,\n", burst_start);
    fprintf(fp, "    \"colourBurstEnd\": %d,\n", burst_end);
    fprintf(fp, "    \"activeVideoStart\": %d,\n", active_video_start(fmt));
    fprintf(fp, "    \"activeVideoEnd\": %d,\n", active_video_end(fmt));
    fprintf(fp, "    \"tapeFormat\": \"VHS\",\n    \"isMapped\": false\n  },\n");
    fprintf(fp, "  \"fields\": [\n");
    for (int i = 0; i < n; i++) {
        const auto& m = meta[(size_t)i];
        fprintf(fp, "    {\n");
        fprintf(fp, "      \"isFirstField\": %s,\n", m.is_first ? "true" : "false");
        fprintf(fp, "      \"seqNo\": %d,\n", i + 1);
        fprintf(fp, "      \"fileLoc\": %zu,\n", m.file_loc);
        if (m.phase_id > 0) fprintf(fp, "      \"fieldPhaseID\": %d,\n", m.phase_id);
        fprintf(fp, "      \"dropOuts\": {\n        \"fieldLine\": [");
        for (size_t j = 0; j < m.dl.size(); j++) fprintf(fp, j ? ", %d" : "%d", m.dl[j]);
        fprintf(fp, "],\n        \"startx\": [");
        for (size_t j = 0; j < m.ds.size(); j++) fprintf(fp, j ? ", %d" : "%d", m.ds[j]);
        fprintf(fp, "],\n        \"endx\": [");
        for (size_t j = 0; j < m.de.size(); j++) fprintf(fp, j ? ", %d" : "%d", m.de[j]);
        fprintf(fp, "]\n      }\n");
        fprintf(fp, "    }%s\n", (i < n - 1) ? "," : "");
    }
    fprintf(fp, "  ]\n}\n");
    fclose(fp);
    return 0;
}

