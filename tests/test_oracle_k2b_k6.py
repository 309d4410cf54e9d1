"""Invariants that pin the oracle's K2b line location and K6 track choice,
which have no closed form (pipeline.cu:450-893, chroma_decode.cu:674-1001).

K2b: on a clean capture (60 dB SNR, no dropouts) every active line's coarse
location (line locations 10..262) is the start of a K2 pulse, one line
apart; where the field runs past the end of the capture, the pulses stop
and the locations continue at a constant spacing (the mean line length).

K6: the track trial on field 0 (burst cancellation metric per head track)
picks the track whose metric is the smaller, by far (right track about
150, wrong track about 1650 on the synthetic captures); a capture that
starts one recorded field later starts on the other head, so the chosen
track alternates with the starting field, and the metrics swap."""
import numpy as np
import pytest

SPF = 468140
LINE = 28e6 / 15734.264


@pytest.fixture(scope="module")
def clean(oracle):
    from cuvhs_amd import native
    cap = native.generate(0, snr_lo_db=60.0, snr_hi_db=60.0, dropouts_per_field=0.0)
    return cap, oracle.decode(cap, 28.0, 64, trace=True)


def test_k2b_active_lines_on_pulse_starts(clean):
    cap, r = clean
    c = r["chunks"][0]
    assert c["nf"] == 4
    for f in range(4):
        starts = set(int(x) for x in c["k2_starts"][f][:c["k2_count"][f]])
        llc = c["ll_coarse"][f]
        inside = [l for l in range(10, 263) if llc[l + 1] < cap.size - 2 * LINE]
        assert len(inside) > 200
        for l in inside:
            assert llc[l] == round(llc[l]) and int(llc[l]) in starts, (f, l, llc[l])
        sp = np.diff(llc[inside[0]:inside[-1] + 1])
        assert np.all(np.abs(sp - LINE) < 3), (f, sp.min(), sp.max())
        # lines not on a pulse: only at the end of the last field, past the
        # capture, one contiguous run at the constant fallback spacing
        past = [l for l in range(10, 264) if llc[l] != round(llc[l])]
        if f < 3:
            assert not past
        else:
            assert len(past) > 5 and past == list(range(past[0], 264))
            assert llc[past[0]] > cap.size - 2 * LINE
            d = np.diff(llc[past[0]:264])
            assert np.all(np.abs(d - d[0]) < 1e-6) and abs(d[0] - LINE) < 3


@pytest.mark.parametrize("start", range(4))
def test_chroma_track_follows_the_recording(oracle, start):
    from cuvhs_amd import native
    g = native.generate(1, n_samples=(start + 3) * SPF)
    r = oracle.decode(np.ascontiguousarray(g[start * SPF:]), 28.0, 64, trace=False)
    m = r["track_trial"]["metrics"]
    track = r["track_trial"]["track"]
    assert track == start % 2
    assert m[1 - track] > 5.0 * m[track], m


def test_phase_ids_follow_the_cycle_across_chunks(oracle):
    """Phase IDs step through the 8-field cycle 3, 2, 1, 4, 3, 2, 1, 4
    (chroma_decode.cu:671-672, 761) field after field, also across chunk
    boundaries, where the cycle position is carried in the chroma state
    (chroma_decode.cu:999-1000):
    70 fields in chunks of 16 (five chunks)."""
    from cuvhs_amd import native
    cap = native.generate(1, n_samples=70 * SPF)
    r = oracle.decode(cap, 28.0, 16, trace=True)
    assert len(r["chunks"]) == 5
    ids = [int(x) for x in r["phase_id"]]
    assert len(ids) == r["fields"] >= 65
    order = {3: 2, 2: 1, 1: 4, 4: 3}   # successor in the cycle
    assert all(order[a] == b for a, b in zip(ids, ids[1:])), ids


@pytest.mark.parametrize("start", [0, 1])
def test_k2b_vblank_anchor_and_field_order(oracle, start):
    """The vertical interval of a clean capture (the generator writes 6
    equalising, 6 broad and 6 equalising pulses at half-line spacing) comes
    out of K2b's vblank state machine (pipeline.cu:519-609) as 6 EQ1, 6 VSYNC
    and 6 EQ2 pulses;
    line location 10 is the first HSYNC after them (the first-HSYNC anchor,
    pipeline.cu:650-855); K2b's field order matches the generator's field
    type (even recorded fields start their lines on whole-line origins),
    also when the capture starts one recorded field later."""
    from cuvhs_amd import native
    g = native.generate(0, snr_lo_db=60.0, snr_hi_db=60.0, dropouts_per_field=0.0, n_samples=(start + 4) * SPF)
    r = oracle.decode(np.ascontiguousarray(g[start * SPF:]), 28.0, 64, trace=True)
    c = r["chunks"][0]
    assert c["nf"] >= 3
    for f in range(c["nf"]):
        n = c["vp_count"][f]
        ty = [int(x) for x in c["vp_types"][f][:n]]
        st = c["vp_starts"][f][:n]
        i = ty.index(1)                               # first EQ1
        assert ty[i:i + 18] == [1] * 6 + [2] * 6 + [3] * 6, ty[:30]
        h = i + 18 + ty[i + 18:].index(0)             # first HSYNC after the interval
        assert c["ll_coarse"][f][10] == st[h]
        assert c["is_first_k2b"][f] == int((start + f) % 2 == 0)


def test_chroma_track_reacquired_after_a_dropped_field(oracle):
    """A capture with one recorded field cut out (samples 6 spf to 7 spf of
    the tape) breaks the head alternation: the field after the cut is read
    by the same head as the one before it.  The per-sub-chunk retry
    (chroma_decode.cu:774-801, 900-975) flips the track at the first bad
    field, so each recorded field after the cut decodes as it does without
    the cut: the same fileLoc (less one field period), identical luma, and
    chroma within 1 LSB on over 98% of the pixels (the circular pre-band-pass
    of the chunk's segments sees different neighbours at segment edges).  On
    the wrong track every other line's chroma would be inverted."""
    from cuvhs_amd import native
    g = native.generate(1, n_samples=14 * SPF)
    plain = oracle.decode(np.ascontiguousarray(g[:13 * SPF]), 28.0, 64, trace=False)
    cut = oracle.decode(np.ascontiguousarray(np.concatenate([g[:6 * SPF], g[7 * SPF:]])), 28.0, 64, trace=False)
    for j in range(6, 11):   # recorded field j + 1 (the last field is cut by the capture end)
        assert int(cut["file_loc"][j]) + SPF == int(plain["file_loc"][j + 1])
        assert np.array_equal(cut["luma"][j], plain["luma"][j + 1]) or cut["do_count"][j] != plain["do_count"][j + 1]
        d = np.abs(cut["chroma"][j].astype(np.int64) - plain["chroma"][j + 1].astype(np.int64))
        assert (d > 1).mean() < 0.02, (j, float((d > 1).mean()))
