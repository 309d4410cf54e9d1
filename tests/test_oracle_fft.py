"""Oracle FFT (oracle/fft.cpp) against numpy.fft fixtures and a brute-force DFT."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def run_fft(oracle, x, sign):
    n = x.size
    a = np.ascontiguousarray(np.stack([x.real, x.imag], 1))
    b = np.zeros_like(a)
    assert oracle.lib().or_fft(n, a.ctypes.data, b.ctypes.data, sign) == 0
    return b[:, 0] + 1j * b[:, 1]


def test_fft_golden_vectors(oracle):
    g = np.load(os.path.join(GOLDEN, "fft_vectors.npz"))
    sizes = sorted(int(k[1:]) for k in g.files if k.startswith("x"))
    assert len(sizes) >= 10
    for n in sizes:
        x = g[f"x{n}"]
        for sign, key in ((-1, "fwd"), (1, "inv")):
            y = run_fft(oracle, x, sign)
            ref = g[f"{key}{n}"]
            assert np.abs(y - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), (n, sign)


@pytest.mark.parametrize("n", [2, 3, 5, 7, 12, 35, 64])
def test_fft_brute_force(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert np.allclose(run_fft(oracle, x, -1), W @ x, atol=1e-12)
    assert np.allclose(run_fft(oracle, x, 1), W.conj() @ x, atol=1e-12)


def test_fft_reference_sizes(oracle):
    # the reference's per-segment sizes at 28 MHz (K1) and its chroma size
    rng = np.random.default_rng(7)
    for n in (506250, 504000, 1024):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        y = run_fft(oracle, x, -1)
        ref = np.fft.fft(x)
        assert np.abs(y - ref).max() / np.abs(ref).max() < 1e-13


def test_fft_rejects_non_smooth(oracle):
    a = np.zeros(22)
    assert oracle.lib().or_fft(11, a.ctypes.data, a.ctypes.data, -1) == -1
