"""The FP64 arctangent of K1's phase (cuvhs_amd/csrc/atan_f64.h), restated in
numpy with the coefficients read from the header, against numpy's arctan2:
the range reduction, the polynomial and the octant/sign fix-ups must hold
the result within 4 ulp (the FMAs emulated in extended precision), and the
signed-zero and axis cases exactly.  The generator's interpolation bound
(gen/atan_coeffs.py) is checked too.  The GPU test runs the kernel itself."""
import os
import re
import subprocess
import sys

import numpy as np

HDR = os.path.join(os.path.dirname(__file__), "..", "cuvhs_amd", "csrc", "atan_f64.h")
GEN = os.path.join(os.path.dirname(__file__), "..", "cuvhs_amd", "csrc", "gen", "atan_coeffs.py")


def header_constants():
    src = open(HDR).read()
    body = src[src.index("double atan2_f64"):]
    p0 = float(re.search(r"double p = ([-0-9.e]+);", body).group(1))
    coef = [float(v) for v in re.findall(r"fma_sgpr\(p, u, ([-0-9.e]+)\)", body)]
    thr = float(re.search(r"a > ([0-9.]+) \* b", body).group(1))
    return p0, coef, thr


def atan2_method(y, x):
    p0, coef, thr = header_constants()
    L = np.longdouble

    def fma(a, b, c):
        return (L(a) * L(b) + L(c)).astype(np.float64)

    ax, ay = np.abs(x), np.abs(y)
    sw = ay > ax
    a = np.where(sw, ax, ay)
    b = np.where(sw, ay, ax)
    red = a > thr * b
    num = np.where(red, a - b, a)
    den = np.where(red, a + b, b)
    den = np.where(den > 0, den, 1.0)
    t = num / den
    u = t * t
    p = np.full_like(t, p0)
    for c in coef:
        p = fma(p, u, c)
    r = fma(t * u, p, t)
    r = np.where(red, r + np.pi / 4, r)
    r = np.where(sw, np.pi / 2 - r, r)
    r = np.where(np.signbit(x), np.pi - r, r)
    return np.copysign(r, y)


def test_header_matches_generator():
    out = subprocess.run([sys.executable, GEN], capture_output=True, text=True, check=True).stdout
    gen = [float(l.split(",")[0]) for l in out.splitlines() if l.strip().endswith(tuple(f"c{k}" for k in range(20)))]
    p0, coef, _ = header_constants()
    assert [p0] + coef == gen[::-1]
    bound = float(re.search(r"relative to atan\(t\): ([0-9.e+-]+)", out).group(1))
    assert bound < 2e-17


def test_method_accuracy():
    rng = np.random.default_rng(11)
    n = 1 << 20
    y = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 6, n)
    x = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 6, n)
    k = n // 4
    x[:k] = np.abs(x[:k])
    y[:k] = x[:k] * 0.41421356237309503 * (1 + rng.uniform(-1e-12, 1e-12, k))
    x[k:2 * k] = y[k:2 * k] * (1 + rng.uniform(-1e-9, 1e-9, k))
    got, want = atan2_method(y, x), np.arctan2(y, x)
    assert (np.abs(got - want) / np.spacing(np.abs(want))).max() <= 4.0
    sy = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 0.0, 0.0, 1.0, 1.0, -1.0, 3.0])
    sx = np.array([0.0, 0.0, -0.0, -0.0, 0.0, 0.0, 1.0, -1.0, 1.0, -1.0, -1.0, -3.0])
    g, w = atan2_method(sy, sx), np.arctan2(sy, sx)
    assert np.array_equal(g, w) and np.array_equal(np.signbit(g), np.signbit(w))
