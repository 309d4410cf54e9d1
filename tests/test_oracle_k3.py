"""The oracle's K3 line location (line_locs.cu) and the reference's
K3/K4/K5 debug counters, pinned on constructed inputs whose answers follow
from the algorithm's definitions: ideal NTSC and PAL pulse trains of both
field parities, and line grids with planted spacing faults counted by hand."""
import numpy as np
import pytest

MHZ = 28.0
SPL = 1780
SPF = 468140
HW, EQW, VW = 132, 64, 759        # 4.7 / 2.3 / 27.1 us at 28 MHz, rounded


def ideal_field(first, lead_lines=5, start=3000):
    """K2 pulse list of one NTSC field window: lead_lines HSYNCs, then 6 EQ,
    6 broad and 6 EQ pulses at half-line spacing, then HSYNCs.  The first
    EQ follows the last lead HSYNC by 1 line (first field) or 1/2 line."""
    starts, lens = [], []
    p = start
    for _ in range(lead_lines):
        starts.append(p)
        lens.append(HW)
        p += SPL
    p += 0 if first else -SPL // 2
    for w in [EQW] * 6 + [VW] * 6 + [EQW] * 6:
        starts.append(p)
        lens.append(w)
        p += SPL // 2
    first_h = len(starts)
    while p < start + SPF:
        starts.append(p)
        lens.append(HW)
        p += SPL
    st = np.zeros((1, 800), np.int32)
    ln = np.zeros((1, 800), np.int32)
    st[0, :len(starts)] = starts
    ln[0, :len(lens)] = lens
    return st, ln, np.array([len(starts)], np.int32), first_h


@pytest.mark.parametrize("first", [True, False])
def test_k3_ideal_field(oracle, first):
    st, ln, ct, first_h = ideal_field(first)
    types, ll, isf = oracle.k3_linelocs(st, ln, ct, MHZ)
    assert isf[0] == (1 if first else 0)
    n = ct[0]
    # nominal classes: 132 -> H, 64 -> EQ, 759 -> VSYNC
    exp_types = np.where(ln[0, :n] == VW, 2, np.where(ln[0, :n] == EQW, 1, 0))
    assert np.array_equal(types[0, :n], exp_types)
    # the first HSYNC after the vblank is line 10; every later line snaps to
    # the next HSYNC of the exact 1780-sample cadence
    hs = st[0, first_h:n]
    m = min(hs.size, 525 - 10)
    assert np.array_equal(ll[0, 10:10 + m], hs[:m].astype(float))
    # lines past the pulse list continue the cadence
    if 10 + m < 525:
        assert np.allclose(np.diff(ll[0, 10 + m - 1:]), SPL)


PAL_SPL, PAL_SPF = 1792, 560896
PAL_VW = 764                      # 27.3 us at 28 MHz, rounded


def ideal_pal_field(first, lead_lines=5, start=3000):
    """K2 pulse list of one PAL field window: lead_lines HSYNCs, then 5 EQ,
    5 broad and 5 EQ pulses at half-line spacing, then HSYNCs.  The first EQ
    follows the last lead HSYNC by 1/2 line (first field) or 1 line, and the
    first HSYNC follows the last EQ by the same (PAL's field order,
    pipeline.cu:707-717)."""
    step = 0 if first else PAL_SPL // 2
    starts, lens = [], []
    p = start
    for _ in range(lead_lines):
        starts.append(p)
        lens.append(HW)
        p += PAL_SPL
    p += step - PAL_SPL // 2
    for w in [EQW] * 5 + [PAL_VW] * 5 + [EQW] * 5:
        starts.append(p)
        lens.append(w)
        p += PAL_SPL // 2
    p += step
    first_h = len(starts)
    while p < start + PAL_SPF:
        starts.append(p)
        lens.append(HW)
        p += PAL_SPL
    st = np.zeros((1, 800), np.int32)
    ln = np.zeros((1, 800), np.int32)
    st[0, :len(starts)] = starts
    ln[0, :len(lens)] = lens
    return st, ln, np.array([len(starts)], np.int32), first_h


@pytest.mark.parametrize("first", [True, False])
def test_k3_ideal_pal_field(oracle, first):
    """PAL's parity rule is NTSC's inverted (line_locs.cu:219-235): a short
    entry spacing (1/2 line) is a first field; line 10 is the first HSYNC
    after the vblank and the 625 line locations follow the cadence."""
    st, ln, ct, first_h = ideal_pal_field(first)
    types, ll, isf = oracle.k3_linelocs(st, ln, ct, MHZ, system="pal")
    assert isf[0] == (1 if first else 0)
    n = ct[0]
    exp_types = np.where(ln[0, :n] == PAL_VW, 2, np.where(ln[0, :n] == EQW, 1, 0))
    assert np.array_equal(types[0, :n], exp_types)
    hs = st[0, first_h:n]
    m = min(hs.size, 625 - 10)
    assert ll.shape == (1, 625)
    assert np.array_equal(ll[0, 10:10 + m], hs[:m].astype(float))
    if 10 + m < 625:
        assert np.allclose(np.diff(ll[0, 10 + m - 1:]), PAL_SPL)
    # the same pulses read as NTSC take NTSC's rule: the parity flips
    _, _, isf_n = oracle.k3_linelocs(st, ln, ct, MHZ)
    assert isf_n[0] == (0 if first else 1)


def test_k3_without_vblank_uses_first_hsync(oracle):
    """No VSYNC: reference line 0 at the first HSYNC, mean spacing."""
    starts = np.arange(50) * SPL + 777
    st = np.zeros((1, 800), np.int32)
    ln = np.zeros((1, 800), np.int32)
    st[0, :50] = starts
    ln[0, :50] = HW
    _, ll, isf = oracle.k3_linelocs(st, ln, np.array([50], np.int32), MHZ)
    assert isf[0] == -1
    assert np.array_equal(ll[0, :50], starts.astype(float))
    assert np.allclose(ll[0, 50:], 777 + SPL * np.arange(50, 525))


def grid(n_lines=525, start=1000.0):
    return start + SPL * np.arange(n_lines, dtype=float)


def test_k3_spacing_counters(oracle):
    ll = grid()
    assert oracle.diag_counts(3, MHZ, ll) == (0, 0, 0)
    ll[100] += 300    # two spacings 2080 / 1480 out of +-10 %, both unlike their neighbours
    assert oracle.diag_counts(3, MHZ, ll) == (2, 2, 0)
    ll[200] += 500    # two more, also beyond 25 %: large jumps
    assert oracle.diag_counts(3, MHZ, ll) == (4, 4, 2)
    ll2 = grid()
    ll2[5] += 500     # below the active line start: not counted
    assert oracle.diag_counts(3, MHZ, ll2) == (0, 0, 0)


def test_k4_counters(oracle):
    before = grid()
    after = before.copy()
    d05 = np.full(600000, 4.0e6)
    assert oracle.diag_counts(4, MHZ, d05, before, after) == (0, 0, 0)
    after[50] += 50                  # > 1.5 us (42 samples) and > 1 us from its neighbours' mean
    d05[int(after[60] + 0.5)] = 3.4e6  # refined location on a sync-level sample
    assert oracle.diag_counts(4, MHZ, d05, before, after) == (1, 1, 1)
    after[51] += 30                  # 30 samples: not large; |30 - (50 + 0) / 2| = 5: not isolated;
    # line 50 now sees neighbours (0 + 30) / 2 = 15 -> |50 - 15| = 35 > 28: still isolated
    assert oracle.diag_counts(4, MHZ, d05, before, after) == (1, 1, 1)


def test_k5_counters(oracle):
    ll = grid(start=2000.0)
    demod = np.full(SPF + 20 * SPL, 4.0e6)
    assert oracle.diag_counts(5, MHZ, demod, ll) == (0, 0, 0)
    # a whole row at sync level: its 760 active-picture pixels are sync-like
    a, b = int(ll[5]) - 10, int(ll[6]) + 10
    demod[a:b] = 3.4e6
    assert oracle.diag_counts(5, MHZ, demod, ll)[2] == 760
    # one short row (70 % of a line), the rest shifted to keep their length
    ll2 = grid(start=2000.0)
    ll2[8:] -= 0.3 * SPL
    assert oracle.diag_counts(5, MHZ, np.full(SPF + 20 * SPL, 4.0e6), ll2) == (0, 1, 0)
    # row 0 starting before the buffer: columns with coordinate < 1 are out
    # of bounds; coord = -5 + col * 1780 / 910 < 1 for col 0..3
    ll3 = grid(start=-5.0)
    assert oracle.diag_counts(5, MHZ, np.full(SPF + 20 * SPL, 4.0e6), ll3) == (4, 0, 0)


def test_decode_k3_mode(oracle, cap0):
    """The K3 decode of the 4-field capture: same field schedule and
    metadata sequence as K2b, different line locations."""
    r2 = oracle.decode(cap0, MHZ, 64, trace=True, diagnostics=True)
    r3 = oracle.decode(cap0, MHZ, 64, trace=True, line_locator=1, diagnostics=True)
    assert r2["fields"] == r3["fields"] == 4
    assert np.array_equal(r2["file_loc"], r3["file_loc"])
    assert np.array_equal(r2["is_first"], r3["is_first"])
    assert r2["diag"].shape == r3["diag"].shape == (1, 9)
    assert not np.array_equal(r2["chunks"][0]["ll_coarse"], r3["chunks"][0]["ll_coarse"])
