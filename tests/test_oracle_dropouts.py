"""K7 dropout detection in the oracle pinned by placed RF dips
(dropout_detect.cu:17-339).  The RF of capture 0 is attenuated to 3% over
15% of a line (30% to 45% of the line's span from the refined line
locations) in three places:

* A: field 1, line location 100 — inside the field's scan window
  [max(ll[10], f*spfp), min(ll[273], (f+1)*spfp));
* B: field 2, line location 15 — before 2*spfp, so outside the window (the
  window quirk: the window starts at f*spfp, which runs ahead of the
  fields' real positions by f*(spfp - spf) samples);
* C: field 2, line location 100 — inside.

A dip is a run of 16-sample envelope blocks at or below 0.18 of the field's
mean envelope, so A and C each give one entry on TBC row 90 (line location
100 counted from line location 10), its columns the dip's span scaled to
910 per line, to within a 16-sample block plus the envelope filter's edge
(16 columns, about 31 samples); B gives none.  Concealment copies row 89
into row 90 over those columns.  The other fields' entries are unchanged
(field 3's come from the zeros past the end of the capture)."""
import numpy as np
import pytest

SPFP = 503740


def dipped_capture(cap, ll):
    """cap with the three dips; returns (capture, {name: (field, line, s0, s1)})."""
    out = cap.copy()
    spans = {}
    for name, f, line in (("A", 1, 100), ("B", 2, 15), ("C", 2, 100)):
        a, b = ll[f][line], ll[f][line + 1]
        s0, s1 = int(a + 0.30 * (b - a)), int(a + 0.45 * (b - a))
        seg = out[s0:s1].astype(np.int64) - 128
        out[s0:s1] = (128 + np.round(seg * 0.03)).astype(np.uint8)
        spans[name] = (f, line, s0, s1)
    return out, spans


@pytest.fixture(scope="module")
def dipped(oracle, cap0, ref0):
    cap, spans = dipped_capture(cap0, ref0["chunks"][0]["ll_refined"])
    return cap, spans, oracle.decode(cap, 28.0, 64, trace=True)


def test_window_bounds_of_the_dips(ref0):
    ll = ref0["chunks"][0]["ll_refined"]
    # A and C lie inside their fields' scan windows, B before its window
    for f, line, inside in ((1, 100, True), (2, 15, False), (2, 100, True)):
        lo = max(int(ll[f][10]), f * SPFP)
        hi = min(int(ll[f][273]), (f + 1) * SPFP)
        assert (lo <= ll[f][line] < hi) == inside
        assert ll[f][10] <= ll[f][line]        # every one is past line location 10


def test_placed_dips_give_their_entries(dipped, ref0):
    cap, spans, r = dipped
    assert r["fields"] == 4
    assert list(r["do_count"][:3]) == [0, 1, 1]
    ll = r["chunks"][0]["ll_refined"]
    for name in ("A", "C"):
        f, line, s0, s1 = spans[name]
        a, b = ll[f][line], ll[f][line + 1]
        want_sx = (s0 - a) / (b - a) * 910
        want_ex = (s1 - a) / (b - a) * 910
        assert r["do_line"][f][0] == line - 10
        assert abs(r["do_startx"][f][0] - want_sx) <= 16, (name, r["do_startx"][f][0], want_sx)
        assert abs(r["do_endx"][f][0] - want_ex) <= 16, (name, r["do_endx"][f][0], want_ex)
    # field 3: the entries past the end of the capture, as without the dips
    n3 = ref0["do_count"][3]
    assert r["do_count"][3] == n3
    for k in ("do_line", "do_startx", "do_endx"):
        assert np.array_equal(r[k][3][:n3], ref0[k][3][:n3])


def test_concealment_copies_the_row_above(dipped, ref0):
    cap, spans, r = dipped
    for name in ("A", "C"):
        f = spans[name][0]
        row, sx, ex = r["do_line"][f][0], r["do_startx"][f][0], r["do_endx"][f][0]
        for k in ("luma", "chroma"):
            assert np.array_equal(r[k][f][row, sx:ex], r[k][f][row - 1, sx:ex])
        # the rows on either side differ there, so the copy's direction shows
        assert (ref0["luma"][f][row - 1, sx:ex] != ref0["luma"][f][row + 1, sx:ex]).any()
