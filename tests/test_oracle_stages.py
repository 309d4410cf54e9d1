"""Oracle stage functions against brute-force scans, closed forms and
whole-decode invariants of the 4-field 28 MHz capture."""
import ctypes

import numpy as np
import pytest

MHZ = 28.0
SPL = 1780
SPF = SPL * 263
SPFP = SPF + 20 * SPL
THR = 3542857.1428571427
HW = 4.7e-6 * 28e6


def runs(x, thr, lo, hi):
    """(start, end) of every maximal run x <= thr inside [lo, hi) (numpy)."""
    b = np.zeros(hi - lo + 2, bool)
    b[1:-1] = x[lo:hi] <= thr
    d = np.diff(b.astype(np.int8))
    return np.flatnonzero(d == 1) + lo, np.flatnonzero(d == -1) + lo


def synth_pulses(seed, n, npulse):
    rng = np.random.default_rng(seed)
    x = np.full(n, 4.0e6) + rng.uniform(-1e5, 1e5, n)
    pos = np.sort(rng.choice(n - 2000, npulse, replace=False))
    for p in pos:
        w = int(rng.choice([0, 30, 64, 131, 560, 759, 980, 1200]) + rng.integers(0, 20))
        x[p:p + w] = 3.45e6
    x[0:50] = 3.4e6      # a run starting at sample 0
    x[-40:] = 3.4e6      # a run open at the end
    return x


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_vsync_discover_brute_force(oracle, seed):
    n = 3 * SPFP
    x = synth_pulses(seed, n, 900)
    cap = 2000
    idx = np.zeros(cap, np.int32)
    cnt = oracle.lib().or_vsync_discover(MHZ, x.ctypes.data, n, idx.ctypes.data, cap)
    s, e = runs(x, THR, 0, n)
    w = np.minimum(e - s, 979 + 1)
    exp = s[(w >= 560) & (w <= 979)]
    assert cnt == exp.size
    assert np.array_equal(idx[:min(cnt, cap)], exp[:cap])


@pytest.mark.parametrize("seed", [3, 4])
def test_sync_pulses_brute_force(oracle, seed):
    n = 4 * SPFP
    x = synth_pulses(seed, n + SPFP, 1500)
    offs = np.array([0, 123457, SPFP + 17, 2 * SPFP + 5000], np.uint64)
    st = np.zeros(4 * 800, np.int32)
    ln = np.zeros_like(st)
    ct = np.zeros(4, np.int32)
    oracle.lib().or_sync_pulses(MHZ, x.ctypes.data, offs.ctypes.data, 4, st.ctypes.data, ln.ctypes.data,
                                ct.ctypes.data)
    for f, o in enumerate(offs.astype(np.int64)):
        s, e = runs(x, THR, int(o), int(o) + SPFP)
        keep = (s > o) & (e < o + SPFP)   # a run at the window start is dropped, an open run is not closed
        es, el = s[keep][:800], (e - s)[keep][:800]
        assert ct[f] == es.size
        assert np.array_equal(st[f * 800:f * 800 + ct[f]], es)
        assert np.array_equal(ln[f * 800:f * 800 + ct[f]], el)


def trapezoid_field():
    """Flat porch P with, per line, a 4-sample linear falling edge at a0, a flat
    sync tip S and a 4-sample rising edge at b0 = a0 + 132."""
    P, S = 4.2e6, 3.4e6
    x = np.full(SPFP, P)
    ll = np.zeros(525)
    b0s = np.zeros(525)
    for l in range(525):
        a0 = 3000 + l * SPL + (l % 7)
        ll[l] = a0 + 2
        if a0 + 400 >= SPFP:
            continue
        b0 = a0 + 132
        for j in range(5):
            x[a0 + j] = P + (S - P) * j / 4
            x[b0 + j] = S + (P - S) * j / 4
        x[a0 + 5:b0] = S
        b0s[l] = b0
    return x, ll, b0s, P, S


def test_hsync_refine_closed_form(oracle):
    x, ll, b0s, P, S = trapezoid_field()
    orig = ll.copy()
    oracle.lib().or_hsync_refine(MHZ, x.ctypes.data, SPFP, ll.ctypes.data, 1)
    right_edge_offset = 2.25 * (28.0 / 40.0)
    for l in range(525):
        if l < 10:
            assert ll[l] == orig[l]  # lines before active_line_start are untouched
        elif b0s[l] > 0:
            exp = b0s[l] + 4 * (THR - S) / (P - S) - HW + right_edge_offset
            assert abs(ll[l] - exp) < 1e-9, l
        else:
            assert ll[l] == orig[l]  # windows past the buffer keep the estimate


def test_tbc_resample_closed_form(oracle):
    rng = np.random.default_rng(5)
    n = SPFP
    c0, c1 = 3.7e6, 0.6e6 / n
    idx = np.arange(n, dtype=np.float64)
    for quad in (0.0, 1e-6):
        demod = c0 + c1 * idx + quad * (idx - n / 2) ** 2 * 1e-3
        ll = 1000.37 + np.arange(525) * 1779.91 + rng.uniform(-0.4, 0.4, 525)
        out = np.zeros(263 * 910, np.uint16)
        oracle.lib().or_tbc_resample(MHZ, demod.ctypes.data, n, ll.ctypes.data, 1, out.ctypes.data)
        col = np.arange(910) / 910.0
        coord = ll[:263, None] + col[None, :] * (ll[1:264] - ll[:263])[:, None]
        hz = c0 + c1 * coord + quad * (coord - n / 2) ** 2 * 1e-3
        v = ((hz - 3685714.2857142854) / 7142.8571428571431 + 40.0) * 358.4 + 1024.0
        exp = np.floor(np.clip(v, 0, 65535) + 0.5).astype(np.int64).reshape(-1)
        d = np.abs(out.astype(np.int64) - exp)
        assert d.max() <= 1 and (d == 0).mean() > 0.9999  # Catmull-Rom is exact on quadratics


def test_decode_capture0_invariants(ref0):
    r = ref0
    assert r["fields"] == 4
    fl = r["file_loc"].astype(np.int64)
    # 20-line lead-in, first broad pulse 3 lines into the vblank, offset 5 lines earlier
    line = 28e6 / 15734.264
    assert abs(fl[0] - 18 * line) < 10
    assert np.all(np.abs(np.diff(fl) - 262.5 * line) < 3)   # one field period of the generator
    assert list(r["is_first"]) == [0, 1, 0, 1]   # forced alternation from false (pipeline.cu:2575-2593)
    c = r["chunks"][0]
    sp = np.diff(c["ll_refined"][:, 10:263], axis=1)
    assert np.all(np.abs(sp - line) < 3)
    # ACC normalises the burst: RMS of (chroma - 32768) over the burst window is 4416
    ch = r["chroma"].astype(np.float64) - 32768.0
    rms = np.sqrt((ch[:, 30:240, 76:112] ** 2).mean(axis=2))
    assert np.all(np.abs(rms - 4416.0) < 1.5)
    # ramp lines rise monotonically across active video
    lu = r["luma"].astype(np.int64)
    assert np.all(np.diff(lu[0, 40, 200:850:25]) > 0)


def test_decode_is_deterministic(oracle, cap0, ref0):
    r2 = oracle.decode(cap0, 28.0, 64, trace=False)
    for k in ("luma", "chroma", "file_loc", "is_first", "phase_id", "do_count"):
        assert np.array_equal(r2[k], ref0[k]), k
