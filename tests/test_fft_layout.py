"""The index algebra of the wave-local convolution blocks (fft_lds.h:
conv4096 and conv_rows<R>), emulated lane by lane with numpy and compared
with numpy.fft: stage 1 across the workgroup, the per-wave passes with
their transposes, the spectrum in the blocks' permuted bin order, and the
inverse.  This pins the decomposition the HIP code implements; the GPU
parity tests pin the kernels themselves."""
import numpy as np
import pytest


def dft(v, s):
    n = v.shape[-1]
    k = np.arange(n)
    return v @ np.exp(s * 2j * np.pi * np.outer(k, k) / n).T


def conv_rows_emulated(x, g, R):
    NB, G, L = 16 * R, R // 256, R // 16
    Q = 16 // G
    W = lambda e: np.exp(-2j * np.pi * (np.asarray(e) % NB) / NB)

    def gperm(k):  # conv_rows_gperm<R>
        k1, k2 = k & 15, k >> 4
        c, e1, e2 = k2 & 15, (k2 >> 4) & 15, k2 >> 8
        return k1 * R + (G * (e1 % Q) + e2) * L + (G * c + e1 // Q)

    gp = np.zeros(NB, complex)
    gp[gperm(np.arange(NB))] = g
    buf = np.zeros(NB, complex)
    t = np.arange(R)
    v = dft(np.stack([x[R * r + t] for r in range(16)], 1), -1) * W(np.outer(t, np.arange(16)))
    for r in range(16):
        buf[r * R + t] = v[:, r]
    l = np.arange(L)
    hi, lo = l // G, l % G
    for row in range(0, NB, R):
        v = dft(np.stack([buf[row + L * a + l] for a in range(16)], 1), -1) * W(16 * np.outer(l, np.arange(16)))
        for c in range(16):
            buf[row + L * c + l] = v[:, c]
        v = dft(np.stack([buf[row + L * hi + G * b + lo] for b in range(16)], 1), -1) * W(256 * np.outer(lo, np.arange(16)))
        for e1 in range(16):
            buf[row + L * hi + G * e1 + lo] = v[:, e1]
        v = np.stack([buf[row + L * hi + 16 * lo + j] for j in range(16)], 1)
        for e in range(0, 16, G):
            v[:, e:e + G] = dft(v[:, e:e + G], -1)
        v = v * np.stack([gp[row + j * L + l] for j in range(16)], 1)
        for e in range(0, 16, G):
            v[:, e:e + G] = dft(v[:, e:e + G], +1)
        for j in range(16):
            buf[row + L * hi + 16 * lo + j] = v[:, j]
        v = np.stack([buf[row + L * hi + G * e1 + lo] for e1 in range(16)], 1)
        v = dft(v * np.conj(W(256 * np.outer(lo, np.arange(16)))), +1)
        for b in range(16):
            buf[row + L * hi + G * b + lo] = v[:, b]
        v = np.stack([buf[row + L * c + l] for c in range(16)], 1)
        v = dft(v * np.conj(W(16 * np.outer(l, np.arange(16)))), +1)
        for a in range(16):
            buf[row + L * a + l] = v[:, a]
    v = np.stack([buf[r * R + t] for r in range(16)], 1)
    v = dft(v * np.conj(W(np.outer(t, np.arange(16)))), +1)
    y = np.zeros(NB, complex)
    for r in range(16):
        y[R * r + t] = v[:, r]
    return y


def conv4096_emulated(x, g):
    NB = 4096
    W = lambda e: np.exp(-2j * np.pi * (np.asarray(e) % NB) / NB)

    def gperm(k):  # conv4096_gperm
        k1, c, d = k & 15, (k >> 4) & 15, k >> 8
        return (k1 >> 2) * 1024 + d * 64 + (k1 & 3) * 16 + c

    gp = np.zeros(NB, complex)
    gp[gperm(np.arange(NB))] = g
    buf = np.zeros(NB, complex)
    t = np.arange(256)
    v = dft(np.stack([x[256 * r + t] for r in range(16)], 1), -1) * W(np.outer(t, np.arange(16)))
    for r in range(16):
        buf[(r << 8) + t] = v[:, r]
    lane = np.arange(64)
    q = lane & 15
    for w in range(4):
        base = ((w << 2) + (lane >> 4)) << 8
        v = dft(np.stack([buf[base + 16 * a + q] for a in range(16)], 1), -1) * W(16 * np.outer(q, np.arange(16)))
        for c in range(16):
            buf[base + 16 * c + q] = v[:, c]
        v = dft(np.stack([buf[base + 16 * q + b] for b in range(16)], 1), -1)
        v = v * np.stack([gp[(w << 10) + (d << 6) + lane] for d in range(16)], 1)
        v = dft(v, +1) * np.conj(W(16 * np.outer(q, np.arange(16))))
        for b in range(16):
            buf[base + 16 * q + b] = v[:, b]
        v = dft(np.stack([buf[base + 16 * c + q] for c in range(16)], 1), +1)
        for a in range(16):
            buf[base + 16 * a + q] = v[:, a]
    v = np.stack([buf[(r << 8) + t] for r in range(16)], 1)
    v = dft(v * np.conj(W(np.outer(t, np.arange(16)))), +1)
    y = np.zeros(NB, complex)
    for r in range(16):
        y[256 * r + t] = v[:, r]
    return y


@pytest.mark.parametrize("n", [4096, 8192, 16384])
def test_block_layout_is_circular_convolution(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    g = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = conv4096_emulated(x, g) if n == 4096 else conv_rows_emulated(x, g, n // 16)
    ref = np.fft.ifft(np.fft.fft(x) * g) * n
    assert np.abs(y - ref).max() < 1e-11 * np.abs(ref).max()
    # a transposed bin order (a plausible mistake) does not pass
    if n == 4096:
        g2 = g.reshape(16, 256).T.reshape(-1)
        assert np.abs(conv4096_emulated(x, g2) - ref).max() > 1e-3 * np.abs(ref).max()


def filter1024_emulated(x, keep):
    """fft_lds.h filter1024_wave: lane l holds x[64 n1 + l]; rows of 64 to
    lanes 4 j .. 4 j + 3; bin k = j + 16 (4 q + i) + 256 d at v[4 i + d]."""
    NB = 1024
    W = lambda e: np.exp(-2j * np.pi * (np.asarray(e) % NB) / NB)
    buf = np.zeros(NB, complex)
    l = np.arange(64)
    j, q = l >> 2, l & 3
    v = dft(np.stack([x[64 * r + l] for r in range(16)], 1), -1) * W(np.outer(l, np.arange(16)))
    for k1 in range(16):
        buf[64 * k1 + l] = v[:, k1]
    v = dft(np.stack([buf[64 * j + 4 * a + q] for a in range(16)], 1), -1) * W(16 * np.outer(q, np.arange(16)))
    for c in range(16):
        buf[64 * j + 4 * c + q] = v[:, c]
    v = np.stack([buf[64 * j + 16 * q + m] for m in range(16)], 1)
    for i in range(4):
        v[:, 4 * i:4 * i + 4] = dft(v[:, 4 * i:4 * i + 4], -1)
    for i in range(4):
        for d in range(4):
            k = j + 16 * (4 * q + i) + 256 * d
            v[:, 4 * i + d] *= keep[k]
    for i in range(4):
        v[:, 4 * i:4 * i + 4] = dft(v[:, 4 * i:4 * i + 4], +1)
    for m in range(16):
        buf[64 * j + 16 * q + m] = v[:, m]
    v = np.stack([buf[64 * j + 4 * c + q] for c in range(16)], 1)
    v = dft(v * np.conj(W(16 * np.outer(q, np.arange(16)))), +1)
    for a in range(16):
        buf[64 * j + 4 * a + q] = v[:, a]
    v = np.stack([buf[64 * k1 + l] for k1 in range(16)], 1)
    v = dft(v * np.conj(W(np.outer(l, np.arange(16)))), +1)
    y = np.zeros(NB, complex)
    for r in range(16):
        y[64 * r + l] = v[:, r]
    return y


def test_filter1024_wave_layout():
    rng = np.random.default_rng(1024)
    x = rng.standard_normal(1024) + 1j * rng.standard_normal(1024)
    kk = np.minimum(np.arange(1024), 1024 - np.arange(1024))
    keep = ((kk >= 180) & (kk <= 300)).astype(float)  # a chroma-like band by |k|
    ref = np.fft.ifft(np.fft.fft(x) * keep) * 1024
    assert np.abs(filter1024_emulated(x, keep) - ref).max() < 1e-11 * np.abs(ref).max()
    g = rng.standard_normal(1024) + 1j * rng.standard_normal(1024)  # any per-bin map
    ref = np.fft.ifft(np.fft.fft(x) * g) * 1024
    assert np.abs(filter1024_emulated(x, g) - ref).max() < 1e-11 * np.abs(ref).max()
