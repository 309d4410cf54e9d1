"""The oracle's K2b line location and chroma phase-ID choice against the
reference's own host code.

The reference's host K2b (pipeline.cu:133-148, 410-893) and its NTSC phase
table and burst metrics (chroma_decode.cu:329-450) are plain C++; they are
compiled as they lie into oracle/_ref/libref_k2b.so (oracle/Makefile target
`refk2b`, which needs the reference mounted) and their outputs on degraded
captures are committed as vectors (tests/golden/make_ref_k2b_fixtures.py):

* K2b: the K2 pulse lists the oracle finds on seven degraded captures (the
  full 601-field 28 MHz benchmark capture with SNR 25-40 dB, dropouts and
  wow/flutter; 640 fields of the 40 MHz head-switch capture; 25 dB; 18-20 dB;
  eight dropouts per field; a capture with a recorded field cut out; 35.8
  MHz), and four PAL captures (200 fields at 28 MHz, 192 fields of the 40
  MHz head-switch model, 18-20 dB, a field cut out), which take the
  reference's PAL branches (pipeline.cu:707-717, 811); chunks of 64 fields
  with the cadence state carried across chunks (the two long NTSC captures
  are committed as their first three chunks; all of their chunks were
  compared when the vectors were made).
  The oracle's K2b must give the same validated pulses, types, good flags,
  line locations, is_first and anchor values (valid, line0loc,
  first_hsync_loc, first_hsync_line, next_field, first_field,
  progressive_field, prev_hsync_diff, meanlinelen), bit for bit.
* Chroma: on the oracle's band-passed lines of the field-0 track trial of
  fourteen small captures (both tracks, 28 / 35.8 / 40 MHz, 18-20 dB, NTSC
  and PAL: 313 lines of 2048-point transforms), the
  reference's measure_burst_cancellation must equal the oracle's metric for
  each track, and the reference's measure_burst_phase, quantised to a
  quadrant and looked up in the reference's own table, must give the phase ID
  the oracle assigns to field 0 (chroma_decode.cu:720-742).
* The lookup table for every (is_first, quadrant, delta) key.

Where the compiled reference library is present (this container), the
oracle is also compared with it directly on the same inputs, which checks
that the committed vectors still come out of the reference.
"""
import glob
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

sys.path.insert(0, GOLDEN)
import make_ref_k2b_fixtures as FX  # noqa: E402

K2B_FILES = sorted(glob.glob(os.path.join(GOLDEN, "ref_k2b_*.npz")))
KEYS = ("count", "starts", "lens", "types", "good", "linelocs", "is_first", "anchor")


def _case(path):
    return os.path.basename(path)[len("ref_k2b_"):-len(".npz")]


def test_every_k2b_case_has_vectors():
    assert sorted(_case(p) for p in K2B_FILES) == sorted(FX.K2B_CASES)


def _compare(got, exp, where):
    for k in KEYS:
        a, b = got[k], exp[k]
        assert a.shape == b.shape, (where, k, a.shape, b.shape)
        if not np.array_equal(a, b):
            idx = np.argwhere(a != b)[0]
            pytest.fail(f"{where}: {k} differs first at {tuple(idx)}: {a[tuple(idx)]} vs {b[tuple(idx)]}")


@pytest.mark.parametrize("path", K2B_FILES, ids=_case)
def test_oracle_k2b_equals_reference(oracle, path):
    z = np.load(path)
    mhz = float(z["mhz"])
    system = "pal" if "system" in z.files and int(z["system"]) == 1 else "ntsc"
    chunks = FX.unpack_k2b(z)
    assert len(chunks) >= 1
    st = oracle.K2bState.initial()
    R = oracle.ref_k2b_lib()
    sr = oracle.K2bState.initial()
    fields = 0
    for ci, c in enumerate(chunks):
        # the cadence state the reference carried into this chunk
        assert (st.prev_first_field, st.prev_progressive_field, st.prev_first_hsync_readloc,
                st.prev_first_hsync_loc, st.prev_first_hsync_diff) == c["state_in"], ci
        got = oracle.k2b_chunk(mhz, st, c["raw_offset"], c["offsets"], c["k2_count"], c["k2_starts"], c["k2_lens"],
                               system)
        _compare(got, c["expected"], f"chunk {ci}")
        if R is not None:
            live = oracle.k2b_chunk_with(R.ref_k2b_chunk_sys, mhz, sr, c["raw_offset"], c["offsets"], c["k2_count"],
                                         c["k2_starts"], c["k2_lens"], system)
            _compare(live, c["expected"], f"chunk {ci} (reference library)")
        fields += c["nf"]
    assert fields == len(z["is_first"])


def test_vectors_cover_the_whole_captures():
    """When the vectors were made, the oracle's K2b was run beside the
    reference's on every chunk of every capture (2,352 fields, 558 of them PAL; the two long
    captures keep only their first three chunks in the committed vectors) and
    no chunk differed.  The committed fields hold the validation quirks of
    noisy vertical intervals: validated pulses flagged not good (partial
    re-appended vblank copies, pipeline.cu:611-648) in every capture."""
    total = 0
    for path in K2B_FILES:
        z = np.load(path)
        assert int(z["chunks_differing"]) == 0, path
        total += int(z["fields_total"])
        assert int((z["vp_good"] == 0).sum()) > 5 * len(z["is_first"]), path
    assert total > 1700


@pytest.mark.parametrize("i", range(len(FX.CHROMA_CASES)))
def test_chroma_trial_and_first_phase_id_equal_reference(oracle, i):
    z = np.load(os.path.join(GOLDEN, "ref_chroma_phase.npz"))
    case = FX.CHROMA_CASES[i]
    mhz, config, over, fields, start = case[:5]
    system = FX.case_system(case, 5)
    row = z["cases"][i]
    assert (row[0], row[1], row[2], row[3], row[7]) == (mhz, config, fields, start, oracle.SYSTEMS[system])
    r = FX.chroma_trial(mhz, config, over, fields, start, system)
    lines = r["track_trial"]["lines"]
    assert lines[0].shape == (oracle.GEOMETRY[system][1], oracle.GEOMETRY[system][3])
    m = r["track_trial"]["metrics"]
    # the metric of each track, bit for bit (same operation order)
    assert m[0] == row[4] and m[1] == row[5], (m, row[4:6])
    assert oracle.burst_metric(mhz, lines[0], system) == row[4]
    # the reference's burst phase -> quadrant -> its own lookup table -> the
    # phase ID of field 0 (chroma_decode.cu:728-734)
    assert oracle.burst_phase(mhz, lines[2], system) == row[6]
    quadrant = int(row[6] / 90.0 + 0.5) % 4
    pid0 = int(z["lookup"][1, quadrant, 0, 0])
    assert int(r["phase_id"][0]) == pid0, (r["phase_id"][:4], quadrant)
    R = oracle.ref_k2b_lib()
    if R is not None:
        live = oracle.ref_chroma_trial(R, mhz, lines, system)
        assert live == (row[4], row[5], row[6])


def test_chroma_cases_reach_every_quadrant_parity():
    """The trial captures start on both tracks and at four field offsets, so
    field 0's phase ID takes more than one value."""
    z = np.load(os.path.join(GOLDEN, "ref_chroma_phase.npz"))
    quads = {int(p / 90.0 + 0.5) % 4 for p in z["cases"][:, 6]}
    assert len(quads) >= 2, quads


def test_ntsc_phase_lookup_equals_reference(oracle):
    z = np.load(os.path.join(GOLDEN, "ref_chroma_phase.npz"))
    R = oracle.ref_k2b_lib()
    import ctypes
    for isf in range(2):
        for q in range(4):
            for d in range(-1, 4):
                assert oracle.lookup_ntsc_phase(isf, q, d) == z["lookup"][isf, q, d + 1, 0], (isf, q, d)
                if R is not None:
                    off = ctypes.c_int(0)
                    assert R.ref_lookup_ntsc_phase(isf, q, d, ctypes.byref(off)) == z["lookup"][isf, q, d + 1, 0]
                    assert off.value == z["lookup"][isf, q, d + 1, 1]
