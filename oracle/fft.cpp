// Oracle-only mixed-radix FFT; see fft.h.  Test infrastructure.
#include "fft.h"

#include <cmath>
#include <stdexcept>

namespace orc {

bool is_7smooth(int n) {
    if (n <= 0) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    while (n % 5 == 0) n /= 5;
    while (n % 7 == 0) n /= 7;
    return n == 1;
}

int next_7smooth(int n) {
    while (!is_7smooth(n)) n++;
    return n;
}

FFTPlan::FFTPlan(int n_) : n(n_) {
    if (!is_7smooth(n)) throw std::runtime_error("FFTPlan: size is not 7-smooth");
    int m = n;
    while (m % 4 == 0) { factors.push_back(4); m /= 4; }
    for (int p : {2, 3, 5, 7})
        while (m % p == 0) { factors.push_back(p); m /= p; }
    tw.resize(n);
    const long double two_pi = 6.283185307179586476925286766559L;
    for (int k = 0; k < n; k++) {
        long double a = two_pi * (long double)k / (long double)n;
        tw[k] = cd((double)cosl(a), (double)-sinl(a));
    }
}

namespace {

struct Ctx {
    const FFTPlan* p;
    int sign;
    cd twid(long long j) const {  // exp(sign * 2 pi i j / N)
        cd w = p->tw[(size_t)(j % p->n)];
        return sign < 0 ? w : std::conj(w);
    }
};

void rec(const Ctx& c, const cd* x, long long xs, cd* y, int n, int fi, long long twstep) {
    if (n == 1) { y[0] = x[0]; return; }
    int p = c.p->factors[(size_t)fi];
    int m = n / p;
    for (int q = 0; q < p; q++) rec(c, x + q * xs, xs * p, y + (size_t)q * m, m, fi + 1, twstep * p);
    cd tmp[7];
    cd out[7];
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) tmp[q] = y[(size_t)q * m + k] * c.twid(twstep * (long long)q * k);
        for (int r = 0; r < p; r++) {
            cd s(0.0, 0.0);
            for (int q = 0; q < p; q++) s += tmp[q] * c.twid((long long)((q * r) % p) * (c.p->n / p));
            out[r] = s;
        }
        for (int r = 0; r < p; r++) y[(size_t)k + (size_t)r * m] = out[r];
    }
}

}  // namespace

void FFTPlan::exec(const cd* in, cd* out, int sign) const {
    Ctx c{this, sign};
    rec(c, in, 1, out, n, 0, 1);
}

void fft_r2c(const FFTPlan& p, const double* in, cd* out_half) {
    std::vector<cd> a((size_t)p.n), b((size_t)p.n);
    for (int i = 0; i < p.n; i++) a[(size_t)i] = cd(in[i], 0.0);
    p.exec(a.data(), b.data(), -1);
    for (int k = 0; k <= p.n / 2; k++) out_half[k] = b[(size_t)k];
}

void fft_c2r(const FFTPlan& p, const cd* in_half, double* out) {
    int n = p.n;
    std::vector<cd> a((size_t)n), b((size_t)n);
    for (int k = 0; k <= n / 2; k++) a[(size_t)k] = in_half[k];
    for (int k = n / 2 + 1; k < n; k++) a[(size_t)k] = std::conj(in_half[n - k]);
    // DC (and Nyquist for even n) are treated as real
    a[0] = cd(a[0].real(), 0.0);
    if (n % 2 == 0) a[(size_t)n / 2] = cd(a[(size_t)n / 2].real(), 0.0);
    p.exec(a.data(), b.data(), +1);
    for (int i = 0; i < n; i++) out[i] = b[(size_t)i].real();
}

}  // namespace orc
