// Host-side filter design (see design.h).
//
// The magnitude/phase formulas are the reference's (fm_demod.cu:35-127,
// :348-395; chroma_decode.cu:40-49, :574-587): they define the decoder's
// frequency responses, so they are restated with the same constants and the
// same operation order.  What differs is how they are applied: the reference
// multiplies whole-segment spectra (N ~ 5e5 bins); here the responses are
// turned into their circular impulse responses on the same N-point grid,
// truncated where the taps fall below eps * max, and applied by short
// overlap-save blocks whose input indices wrap modulo N exactly as the
// reference's circular segment transform does.
#include "design.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <utility>

namespace cuvhs {

namespace {

const double PI = 3.14159265358979323846;
const double TAU = 2.0 * PI;

double butter_lpf_mag(double omega, double omega_c, int order) {
    if (omega_c <= 0.0 || omega_c >= PI) return 1.0;
    double tc = tan(omega_c / 2.0);
    if (tc == 0.0) return 1.0;
    double t = tan(omega / 2.0);
    return 1.0 / sqrt(1.0 + pow(t / tc, 2.0 * order));
}

double butter_hpf_mag(double omega, double omega_c, int order) {
    if (omega <= 0.0) return 0.0;
    if (omega_c <= 0.0 || omega_c >= PI) return 1.0;
    double tc = tan(omega_c / 2.0);
    double t = tan(omega / 2.0);
    if (t == 0.0) return 0.0;
    return 1.0 / sqrt(1.0 + pow(tc / t, 2.0 * order));
}

double butter_bpf_mag(double omega, double omega_l, double omega_h, int order) {
    if (omega <= 0.0 || omega >= PI) return 0.0;
    double tl = tan(omega_l / 2.0), th = tan(omega_h / 2.0), t = tan(omega / 2.0);
    double bw = th - tl;
    if (bw <= 0.0 || t == 0.0) return 0.0;
    double s = (t * t - tl * th) / (bw * t);
    return 1.0 / sqrt(1.0 + pow(s, 2.0 * order));
}

double supergauss_mag(double f_hz, double corner, int order) {
    double ln2_half = log(2.0) / 2.0;
    double scale = pow(ln2_half, 1.0 / (2.0 * order));
    double x = 2.0 * f_hz * scale / corner;
    return exp(-2.0 * pow(x, 2.0 * order));
}

// RBJ high shelf; the decoder uses it inverted (numerator/denominator swapped)
void shelf_high(double f0, double dbgain, double fs, double q, double b[3], double a[3]) {
    double A = pow(10.0, dbgain / 40.0);
    double w0 = TAU * f0 / fs;
    double alpha = sin(w0) / (2.0 * q);
    double c = cos(w0), sA = sqrt(A);
    b[0] = A * ((A + 1) + (A - 1) * c + 2 * sA * alpha);
    b[1] = -2 * A * ((A - 1) + (A + 1) * c);
    b[2] = A * ((A + 1) + (A - 1) * c - 2 * sA * alpha);
    a[0] = (A + 1) - (A - 1) * c + 2 * sA * alpha;
    a[1] = 2 * ((A - 1) - (A + 1) * c);
    a[2] = (A + 1) - (A - 1) * c - 2 * sA * alpha;
}

cdbl biquad_response(const double num[3], const double den[3], double omega) {
    cdbl z1 = std::exp(cdbl(0.0, -omega));
    cdbl z2 = z1 * z1;
    cdbl n = num[0] + num[1] * z1 + num[2] * z2;
    cdbl d = den[0] + den[1] * z1 + den[2] * z2;
    return n / d;
}

std::vector<double> hamming_lowpass(int taps, double cutoff) {
    std::vector<double> h((size_t)taps);
    double alpha = (taps - 1) / 2.0, sum = 0.0;
    for (int n = 0; n < taps; n++) {
        double m = n - alpha;
        double s = fabs(m) < 1e-12 ? cutoff : sin(PI * cutoff * m) / (PI * m);
        double w = 0.54 - 0.46 * cos(TAU * n / (taps - 1));
        h[(size_t)n] = s * w;
        sum += h[(size_t)n];
    }
    for (double& v : h) v /= sum;
    return h;
}

cdbl fir_response(const std::vector<double>& h, double omega, int delay) {
    cdbl acc(0.0, 0.0);
    for (size_t n = 0; n < h.size(); n++) acc += h[n] * std::exp(cdbl(0.0, -(double)n * omega));
    if (delay > 0) acc *= std::exp(cdbl(0.0, omega * delay));
    return acc;
}

std::vector<int> factorize(int n) {
    std::vector<int> f;
    while (n % 4 == 0) { f.push_back(4); n /= 4; }
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) { f.push_back(p); n /= p; }
    if (n != 1) throw std::runtime_error("size is not 7-smooth");
    return f;
}

// Every loop below over bins, taps or transform columns computes each element
// on its own, so they run in parallel (OpenMP, the host's cores) with the
// same results as run serially.

// Hermitian extension of a half spectrum (DC/Nyquist imaginary parts
// ignored, the cuFFT C2R convention) followed by the unnormalised inverse:
// the real circular kernel the reference's Z2D applies.
std::vector<double> real_kernel_from_half(const std::vector<cdbl>& half, int n) {
    std::vector<cdbl> full((size_t)n);
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= n / 2; k++) full[(size_t)k] = half[(size_t)k];
#pragma omp parallel for schedule(static)
    for (int k = n / 2 + 1; k < n; k++) full[(size_t)k] = std::conj(half[(size_t)(n - k)]);
    full[0] = cdbl(full[0].real(), 0.0);
    if (n % 2 == 0) full[(size_t)n / 2] = cdbl(full[(size_t)n / 2].real(), 0.0);
    host_fft(full, +1);
    std::vector<double> g((size_t)n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) g[(size_t)i] = full[(size_t)i].real();
    return g;
}

// truncation lags of a circular kernel (index m mod n)
void kernel_support(const std::vector<cdbl>& g, double eps, int* lm, int* lp) {
    int n = (int)g.size();
    double mx = 0.0;
    for (const cdbl& v : g) mx = std::max(mx, std::abs(v));
    double thr = eps * mx;
    *lp = 0;
    *lm = 0;
    for (int m = 0; m < n / 2; m++) {
        if (std::abs(g[(size_t)m]) > thr) *lp = m;
        if (m > 0 && std::abs(g[(size_t)(n - m)]) > thr) *lm = m;
    }
}

// Block size choice: FFT work per valid output, credited by how many
// workgroups fit in one CU's 160 KB of LDS.  The transforms are latency-bound
// (a barrier per stage), and more independent workgroups per CU hide more of
// it: measured on MI355X, FP64 4096-point K1 blocks (two per CU) run ~1.3x
// more work per unit time than 8192-point blocks (one per CU), and FP32
// 4096-point blocks (four per CU) ~1.45x more than FP32 8192-point blocks
// (two per CU) with the same waves per CU.  The FP64 4096-point kernels now
// exchange real and imaginary parts separately and run three per CU (limited
// by registers); that only widens their lead, and the table keeps the
// whole-complex figures, under which the choices are the same (4096 points
// for K1 part A at 28-40 MHz; 4096 / 8192 / 8192 for part B at
// 28 / 35.8 / 40 MHz).
BlockFilter make_block(const std::vector<cdbl>& g, double eps, const int* bch, int nb, int extra, int elem_bytes) {
    int n = (int)g.size();
    int lm = 0, lp = 0;
    kernel_support(g, eps, &lm, &lp);
    BlockFilter bf;
    double best = 1e300;
    for (int i = 0; i < nb; i++) {
        int B = bch[i];
        int v = B - lm - lp - extra;
        if (v < 256) continue;
        double lds = (double)(B + B / 16) * (double)elem_bytes;
        int per_cu = (int)(160.0 * 1024.0 / lds);
        double occ = per_cu >= 4 ? 1.9 : (per_cu == 3 ? 1.6 : (per_cu == 2 ? 1.3 : 1.0));
        double cost = (double)B * log2((double)B) / (double)v / occ;
        if (cost < best) { best = cost; bf.B = B; bf.valid = v; }
    }
    if (bf.B == 0) throw std::runtime_error("no block size fits the filter support");
    bf.lm = lm;
    bf.lp = lp;
    bf.taps.resize((size_t)(lm + lp + 1));
    std::vector<cdbl> placed((size_t)bf.B, cdbl(0.0, 0.0));
    for (int m = -lm; m <= lp; m++) {
        cdbl v = g[(size_t)((m % n + n) % n)];
        bf.taps[(size_t)(m + lm)] = v;
        placed[(size_t)((m % bf.B + bf.B) % bf.B)] = v;
    }
    host_fft(placed, -1);
    double inv = 1.0 / (double)bf.B;
    for (cdbl& v : placed) v *= inv;
    bf.spectrum = std::move(placed);
    return bf;
}

}  // namespace

bool is_7smooth(int n) {
    if (n <= 0) return false;
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

int next_7smooth(int n) {
    while (!is_7smooth(n)) n++;
    return n;
}

// Twiddles exp(sign 2 pi i k / n), k < n, rounded from long double; kept
// for the sizes already used (a context's designs reuse one size).
static const std::vector<cdbl>& twiddle_table(int n, int sign) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::vector<cdbl>> cache;
    std::lock_guard<std::mutex> lock(mu);
    std::vector<cdbl>& tw = cache[{n, sign}];
    if (tw.empty()) {
        tw.resize((size_t)n);
        const long double two_pi = 6.283185307179586476925286766559L;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < n; k++) {
            long double a = two_pi * (long double)k / (long double)n;
            tw[(size_t)k] = cdbl((double)cosl(a), (double)(sign * sinl(a)));
        }
    }
    return tw;
}

void host_fft(std::vector<cdbl>& x, int sign) {
    const int n = (int)x.size();
    if (n <= 1) return;
    std::vector<int> fac = factorize(n);
    const std::vector<cdbl>& tw = twiddle_table(n, sign);
    std::vector<cdbl> y((size_t)n);
    cdbl* a = x.data();
    cdbl* b = y.data();
    long long ns = 1;
    for (int p : fac) {
        int m = n / p;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < m; j++) {
            long long k = j % ns;
            cdbl v[7], o[7];
            for (int r = 0; r < p; r++) {
                long long e = (long long)r * k * (n / (ns * p));
                v[r] = a[j + (size_t)r * m] * tw[(size_t)(e % n)];
            }
            for (int s = 0; s < p; s++) {
                cdbl acc(0.0, 0.0);
                for (int r = 0; r < p; r++) acc += v[r] * tw[(size_t)(((long long)r * s % p) * (n / p))];
                o[s] = acc;
            }
            size_t idx = (size_t)((j / ns) * ns * p + k);
            for (int s = 0; s < p; s++) b[idx + (size_t)s * ns] = o[s];
        }
        std::swap(a, b);
        ns *= p;
    }
    if (a != x.data()) std::copy(a, a + n, x.data());
}

namespace {

// The two circular kernels of K1 on the N-point grid, which depend only on
// the sample rate and N: the analytic RF band-pass and video + i * sync.
struct K1Kernels {
    std::vector<cdbl> rf, video;
};

K1Kernels k1_kernels(double fs, int N) {
    const int fb = N / 2 + 1;
    const double ny = fs / 2.0;
    const double w_lo = PI * 1300000.0 / ny, w_hi = PI * 7500000.0 / ny;
    const double w_lpf = PI * 8000000.0 / ny, w_hpf = PI * 1000000.0 / ny;
    K1Kernels kk;
    // analytic RF response: bins 0..N/2 kept, Hilbert weights (1,2,...,2,1)
    std::vector<cdbl>& spec = kk.rf;
    spec.assign((size_t)N, cdbl(0.0, 0.0));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < fb; k++) {
        double om = PI * (double)k / (double)(fb - 1);
        double hil = (k == 0 || k == fb - 1) ? 1.0 : 2.0;
        spec[(size_t)k] = butter_bpf_mag(om, w_lo, w_hi, 8) * butter_lpf_mag(om, w_lpf, 8) *
                          butter_hpf_mag(om, w_hpf, 8) * hil;
    }
    host_fft(spec, +1);  // unnormalised inverse: the analytic signal carries the N scale

    double bs[3], as[3];
    shelf_high(273755.82, 13.9794, fs, 0.462088186, bs, as);
    std::vector<double> sync = hamming_lowpass(65, 500000.0 / ny);
    std::vector<cdbl> fv((size_t)fb), f05((size_t)fb);
    double inv_n = 1.0 / (double)N;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < fb; k++) {
        double om = PI * (double)k / (double)(fb - 1);
        double f_hz = (double)k * ny / (double)(fb - 1);
        cdbl de = biquad_response(as, bs, om);
        double vl = supergauss_mag(f_hz, 3000000.0, 9);
        cdbl sy = fir_response(sync, om, 32);
        fv[(size_t)k] = de * vl * inv_n;
        f05[(size_t)k] = de * vl * sy * inv_n;
    }
    std::vector<double> gv = real_kernel_from_half(fv, N);
    std::vector<double> g5 = real_kernel_from_half(f05, N);
    kk.video.resize((size_t)N);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) kk.video[(size_t)i] = cdbl(gv[(size_t)i], g5[(size_t)i]);
    return kk;
}

// The kernels of the last few (rate, N) pairs designed: a mixed-precision
// context designs K1 twice (its own blocks and the FP64 guard band's) and a
// server creates many contexts at one rate, all from the same kernels.
std::shared_ptr<const K1Kernels> cached_k1_kernels(double fs, int N) {
    static std::mutex mu;
    static std::map<std::pair<double, int>, std::shared_ptr<const K1Kernels>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({fs, N});
    if (it != cache.end()) return it->second;
    if (cache.size() >= 4) cache.clear();
    auto kk = std::make_shared<const K1Kernels>(k1_kernels(fs, N));
    cache[{fs, N}] = kk;
    return kk;
}

}  // namespace

K1Design design_k1(const Fmt& fmt, double eps, const int* ach, int na, const int* bch, int nb, int extra_a,
                   int elem_bytes) {
    K1Design d;
    d.N = next_7smooth(fmt.spfp + kOverlap);
    std::shared_ptr<const K1Kernels> kk = cached_k1_kernels(fmt.sample_rate, d.N);
    d.rf = make_block(kk->rf, eps, ach, na, extra_a, elem_bytes);
    d.video = make_block(kk->video, eps, bch, nb, 0, elem_bytes);
    return d;
}

PreDesign design_prebpf(const Fmt& fmt, double eps, const int* bch, int nb) {
    PreDesign d;
    d.N2 = next_7smooth(fmt.spfp);
    const int N2 = d.N2, fb = N2 / 2 + 1;
    double om_lo = 2.0 * M_PI * 60000.0 / fmt.sample_rate;
    double om_hi = 2.0 * M_PI * 1200000.0 / fmt.sample_rate;
    double norm = 1.0 / (double)N2;
    std::vector<cdbl> half((size_t)fb);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < fb; i++) {
        double om = M_PI * (double)i / (double)(fb - 1);
        double mag = butter_bpf_mag(om, om_lo, om_hi, 4);
        half[(size_t)i] = cdbl(mag * mag * norm, 0.0);
    }
    std::vector<double> g = real_kernel_from_half(half, N2);
    std::vector<cdbl> gc(g.begin(), g.end());
    d.bpf = make_block(gc, eps, bch, nb, 0, 8);
    return d;
}

}  // namespace cuvhs
