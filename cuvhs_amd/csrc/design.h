// Host-side filter design: the reference's frequency responses, turned into
// truncated impulse responses and block spectra for overlap-save on the GPU.
#pragma once
#include <complex>
#include <vector>

#include "common.h"

namespace cuvhs {

using cdbl = std::complex<double>;

// Iterative Stockham mixed-radix DFT for sizes whose prime factors are 2, 3,
// 5 and 7.  sign = -1 forward, +1 inverse (unnormalised).
void host_fft(std::vector<cdbl>& x, int sign);
bool is_7smooth(int n);
int next_7smooth(int n);

// A linear-convolution kernel y[q] = sum_{m=-lm..lp} g[m] x[q-m], and its
// FFT_B spectrum (pre-scaled by 1/B) for blocks of B points.
struct BlockFilter {
    int B = 0;                    // block FFT size (power of two)
    int lm = 0, lp = 0;           // non-causal / causal taps kept
    int valid = 0;                // outputs per block = B - lm - lp - extra
    std::vector<cdbl> taps;       // g[m], m = -lm..lp
    std::vector<cdbl> spectrum;   // B bins
};

struct K1Design {
    int N = 0;                    // the reference's per-segment FFT length
    BlockFilter rf;               // analytic RF band-pass (complex kernel, x N scale)
    BlockFilter video;            // real video kernel + i * real sync kernel
};

// fm_demod.cu:273-374 with spf = spfp (pipeline.cu:1455).
// eps: relative tap magnitude below which the kernel is truncated.
// elem_bytes: bytes per complex LDS element of the transform (16 FP64, 8 FP32).
// a_choices / b_choices: the block sizes part A / part B may take (the
// cheapest under the cost model, or the only one given).
K1Design design_k1(const Fmt& fmt, double eps, const int* a_choices, int na, const int* b_choices, int nb,
                   int extra_a, int elem_bytes);

// chroma_decode.cu:568-587: |H|^2/N pre-band-pass at the capture rate, as a
// real kernel on the N2 = next_7smooth(spfp) grid.
struct PreDesign {
    int N2 = 0;
    BlockFilter bpf;
};
PreDesign design_prebpf(const Fmt& fmt, double eps, const int* b_choices, int nb_choices);

}  // namespace cuvhs
