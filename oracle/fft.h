// Oracle-only FFT (test infrastructure; never linked into the product).
//
// Textbook mixed-radix Cooley-Tukey DFT, FP64, sizes whose prime factors are
// 2, 3, 5 and 7 (the reference only plans 7-smooth sizes:
// fm_demod.cu:75-86, chroma_decode.cu:23-34, and 1024 for the per-line
// chroma transform, chroma_decode.cu:538).  Conventions follow cuFFT, the
// library the reference calls:
//   forward  X[k] = sum_n x[n] exp(-2 pi i k n / N)
//   inverse  x[n] = sum_k X[k] exp(+2 pi i k n / N)   (unnormalised)
// Parity of this routine is pinned by tests/test_oracle_fft.py against
// numpy.fft and a brute-force DFT.
#pragma once
#include <complex>
#include <vector>

namespace orc {

using cd = std::complex<double>;

struct FFTPlan {
    int n = 0;
    std::vector<int> factors;     // radix sequence, product == n
    std::vector<cd> tw;           // exp(-2 pi i k / n), k in [0, n)
    explicit FFTPlan(int n);
    // out-of-place complex transform; sign -1 = forward, +1 = inverse
    void exec(const cd* in, cd* out, int sign) const;
};

// D2Z: real -> half spectrum (n/2+1 bins)
void fft_r2c(const FFTPlan& p, const double* in, cd* out_half);
// Z2D: half spectrum -> real, Hermitian extension, imaginary parts of the
// DC / Nyquist bins ignored, unnormalised (cuFFT C2R convention)
void fft_c2r(const FFTPlan& p, const cd* in_half, double* out);

bool is_7smooth(int n);
int next_7smooth(int n);

}  // namespace orc
