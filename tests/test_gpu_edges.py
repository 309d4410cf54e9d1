"""Captures at the edges of the chunk loop and of the decoder's soft paths,
each decoded by the whole-capture and the per-chunk entry points and
compared with the oracle:

* lengths (pipeline.cu:2626-2713 sizes each chunk from the whole fields
  left; a read shorter than one field ends the decode, pipeline.cu:1535):
  empty, one sample short of a field, exactly one field, one field and a
  fraction, two and a half fields, also at 40 MHz and for s16 / u16 input;
* a silent capture (the nominal-spacing fallback, pipeline.cu:1614-1626);
* a VSYNC candidate overflow, and a capture with nothing in the chroma band;
* RF dips placed inside and outside the dropout scan windows;
* captures starting on either chroma track, and one with a recorded field
  cut out (the track retry);
* an output buffer too small, and the CLI on short and silent files, from a
  file and from stdin.

A field that reaches past the end of the capture demodulates zeros there
(the chunk buffer is zero past the read, pipeline.cu:1539-1541): the
band-passed signal falls to the transform's rounding floor (an envelope
near 1e-6 against 1e10 inside the capture) and the instantaneous frequency
is the phase of rounding noise, which depends on the transform's summation
order.  Those pixels are not determined by the capture, so the u16 planes
are compared on the rows whose span, plus two lines for the hsync search
windows and the chroma filter, lies inside the read; the integer outputs
are compared on every field.  The silent capture is exact arithmetic end to
end (all-zero samples, no sync), so its outputs are compared exactly."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

MHZ = 28.0
SPF = 468140
SPFP = 503740
W_VSYNC_OVERFLOW = 1
W_NO_VSYNC = 2


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def N():
    from cuvhs_amd import native
    native.lib()
    return native


@pytest.fixture(scope="module")
def dec(N):
    d = N.Decoder(MHZ, 64)
    yield d
    d.close()


SPL = 1780
LPF = 525


def rows_inside(ref, n, batch=64, spfp=503740, spl=SPL):
    """[fields, 263] mask of the output rows whose samples (the row's span
    from the oracle's refined line locations, plus two lines) lie inside
    their chunk's read.  Line locations are chunk positions; a chunk reads
    min(batch * spfp, n - raw_offset) samples."""
    masks = []
    for c in ref["chunks"]:
        n_read = min(batch * spfp, n - c["raw_offset"])
        ll = np.asarray(c["ll_refined"]).reshape(c["nf"], LPF)
        masks.append(ll[:, 1:264] + 2 * spl < n_read)
    return np.concatenate(masks)


def compare(out, ref, exact=False, rows=None):
    n = out["fields"]
    assert n == ref["fields"]
    for k in ("file_loc", "is_first", "phase_id", "do_count"):
        g = out[k].cpu().numpy() if out[k] is not None else np.zeros(0)
        assert np.array_equal(g.astype(np.int64), np.asarray(ref[k]).astype(np.int64)), k
    dc = ref["do_count"]
    for k in ("do_line", "do_startx", "do_endx"):
        if n == 0:
            continue
        g = out[k].cpu().numpy()
        for i in range(n):
            assert np.array_equal(g[i, :dc[i]], ref[k][i, :dc[i]]), (k, i)
    for k in ("luma", "chroma"):
        if n == 0:
            assert out[k] is None or out[k].numel() == 0
            continue
        d = np.abs(out[k].cpu().numpy().astype(np.int64) - ref[k].astype(np.int64))
        if rows is not None:
            d = d[rows]
        assert int(d.max()) <= (0 if exact else 1), (k, int(d.max()), int((d > 1).sum()))


def silent(n):
    return np.full(n, 128, np.uint8)


CASES = {
    "empty": lambda c: c[:0],
    "one_short": lambda c: c[:SPF - 1],
    "one_field": lambda c: c[:SPF],
    "one_and_a_bit": lambda c: c[:SPF + 100],
    "two_and_a_half": lambda c: c[:int(2.5 * SPF)],
}


@pytest.mark.parametrize("case", list(CASES))
def test_capture_lengths_vs_oracle(torch, N, dec, oracle, cap0, case):
    cap = np.ascontiguousarray(CASES[case](cap0))
    print(f"case {case}: {cap.size} samples", flush=True)
    ref = oracle.decode(cap, MHZ, 64, trace=True)
    want = {"empty": 0, "one_short": 0, "one_field": 1, "one_and_a_bit": 1, "two_and_a_half": 2}[case]
    assert ref["fields"] == want
    rows = None
    if want:
        rows = rows_inside(ref, cap.size)
        # the fields start in the first 40 lines, so well over 200 rows of each lie inside
        assert (rows.sum(axis=1) > 200).all()
    d = torch.from_numpy(cap).to("cuda:0")
    dec.reset()
    out = dec.decode_capture(d)
    assert out["status"] >= 0
    compare(out, ref, rows=rows)
    dec.reset()
    outc = dec.decode_chunked(d)
    compare(outc, ref, rows=rows)
    # candidate counts where the chunk's K1 segments lie inside the read (past
    # it, VSYNC-width runs of the rounding noise count too: one field of this
    # capture has 6 inside and, in the oracle, 3 more past the end)
    for g, o in zip(outc["chunks"], ref["chunks"]):
        if o["fields_k1"] * SPFP <= cap.size - o["raw_offset"]:
            assert g["vsync_candidates"] == o["candidates"]
    if want == 0:
        assert outc["chunks"] == []


@pytest.mark.parametrize("mhz,fmt_name,dc,fields", [(40.0, "u8", False, 1.5), (28.0, "u16", True, 1.2),
                                                     (28.0, "s16", False, 2.5)])
def test_short_captures_other_rates_and_formats(torch, N, oracle, mhz, fmt_name, dc, fields):
    """Partial-field captures through the 40 MHz decoder (8192-point K1
    blocks) and the s16 / u16 input paths (u16 with the per-read mean
    removed, raw_reader.cpp:93-100)."""
    f = N.format_for(mhz)
    n = int(fields * f.samples_per_field)
    g = N.generate(2 if mhz == 40.0 else 1, sample_rate_mhz=mhz, n_samples=n + f.samples_per_field)[:n]
    if fmt_name == "u16":
        cap = (g.astype(np.uint16) << 8) | np.uint16(0x55)
    elif fmt_name == "s16":
        cap = ((g.astype(np.int16) - 128) * 256).astype(np.int16)
    else:
        cap = np.ascontiguousarray(g)
    print(f"{mhz} MHz {fmt_name} dc={dc}: {n} samples", flush=True)
    ref = oracle.decode(cap, mhz, 64, trace=True, dc_correct=dc)
    assert ref["fields"] == int(fields)
    rows = rows_inside(ref, n, spfp=f.spfp, spl=f.samples_per_line)
    assert (rows.sum(axis=1) > 200).all()
    d = N.Decoder(mhz, 64, input_format=fmt_name, dc_correct=dc)
    try:
        dd = torch.from_numpy(cap).to("cuda:0")
        out = d.decode_capture(dd)
        compare(out, ref, rows=rows)
        d.reset()
        compare(d.decode_chunked(dd), ref, rows=rows)
    finally:
        d.close()


def test_silent_capture_nominal_fallback(torch, N, dec, oracle):
    cap = silent(5 * SPF + 777)
    ref = oracle.decode(cap, MHZ, 64, trace=True)
    assert ref["fields"] == 5
    assert [int(x) for x in ref["file_loc"]] == [k * SPF for k in range(5)]
    d = torch.from_numpy(cap).to("cuda:0")
    dec.reset()
    out = dec.decode_capture(d)
    assert out["status"] & W_NO_VSYNC
    compare(out, ref, exact=True)
    dec.reset()
    outc = dec.decode_chunked(d)
    assert [c["status"] & W_NO_VSYNC for c in outc["chunks"]] == [W_NO_VSYNC]
    assert outc["chunks"][0]["vsync_candidates"] == 0
    assert outc["chunks"][0]["next_raw_offset"] == ref["chunks"][0]["next_offset"]
    compare(outc, ref, exact=True)


def square_fm(n, colour_under):
    """u8 RF: FM alternating 800 samples at the sync tip (-40 IRE, a
    VSYNC-width run below the pulse threshold) with 1200 at 44 IRE, plus,
    if given, a colour-under carrier (40 fH) of that amplitude in u8 steps."""
    from cuvhs_amd import native
    fmt = native.format_for(MHZ)
    fs = fmt.sample_rate
    k = np.arange(n)
    f = np.where((k % 2000) < 800, fmt.ire0 - 40 * fmt.hz_ire, fmt.ire0 + 44 * fmt.hz_ire)
    rf = 128 + 90 * np.cos(np.cumsum(2 * np.pi * f / fs))
    if colour_under:
        rf = rf + colour_under * np.cos(2 * np.pi * 40 * 15734.264 / fs * k)
    return np.round(rf).astype(np.uint8)


def decode_both(dec, torch, cap):
    d = torch.from_numpy(cap).to("cuda:0")
    dec.reset()
    out = dec.decode_capture(d)
    dec.reset()
    return out, dec.decode_chunked(d)


def test_vsync_candidate_overflow(torch, N, dec, oracle):
    """More VSYNC-width runs than the candidate cap (20 per field,
    pipeline.cu:1578-1590), square-wave FM with a colour-under carrier, two
    padded segments long so that no chunk's K1 segments reach past its read
    (every candidate is then a run of the capture): the first chunk finds
    504 candidates for a cap of 40, keeps the first 40 in ascending order
    and clusters them into one field; the second chunk (one field)
    overflows again."""
    n = 2 * SPFP
    cap = square_fm(n, 20)
    ref = oracle.decode(cap, MHZ, 64, trace=True)
    want_cands = [c["candidates"] for c in ref["chunks"]]
    assert want_cands == [504, 251] and ref["fields"] == 2
    assert all(c["fields_k1"] * SPFP <= n - c["raw_offset"] for c in ref["chunks"])
    rows = rows_inside(ref, n)
    assert (rows.sum(axis=1) > 200).all()
    out, outc = decode_both(dec, torch, cap)
    assert out["status"] & W_VSYNC_OVERFLOW
    compare(out, ref, rows=rows)
    assert [c["vsync_candidates"] for c in outc["chunks"]] == want_cands
    assert all(c["status"] & W_VSYNC_OVERFLOW for c in outc["chunks"])
    assert [c["raw_offset"] for c in outc["chunks"]] == [c["raw_offset"] for c in ref["chunks"]]
    assert [c["fields_k1"] for c in outc["chunks"]] == [c["fields_k1"] for c in ref["chunks"]]
    compare(outc, ref, rows=rows)


def test_chroma_band_empty(torch, N, dec, oracle):
    """The same square-wave FM without a colour-under carrier: the chroma
    band (60 kHz - 1.2 MHz) holds only the u8 quantisation, its RMS 1/443 of
    the input's (1/4.6 with the carrier, about 1/7 for the synthetic
    captures).  The burst gain scales the band-passed signal to a fixed
    burst level, so in u16 steps any pre-band-pass error relative to the
    input's RMS grows by that ratio.  The FP64 form runs the pre-band-pass
    in FP64 with taps down to 1e-9 of the largest (as the reference's FP64
    cuFFT band-pass, chroma_decode.cu:568-635), so chroma stays within one
    step here too (an FP32 band-pass with taps to 1e-7 was up to 7 steps off
    on about 2% of the pixels)."""
    n = 2 * SPFP
    cap = square_fm(n, 0)
    ref = oracle.decode(cap, MHZ, 64, trace=True)
    assert [c["candidates"] for c in ref["chunks"]] == [504, 251]
    rows = rows_inside(ref, n)
    out, _ = decode_both(dec, torch, cap)
    assert out["fields"] == ref["fields"]
    for k in ("file_loc", "is_first", "phase_id", "do_count"):
        assert np.array_equal(out[k].cpu().numpy().astype(np.int64), np.asarray(ref[k]).astype(np.int64)), k
    d = np.abs(out["luma"].cpu().numpy().astype(np.int64) - ref["luma"].astype(np.int64))[rows]
    assert int(d.max()) <= 1
    d = np.abs(out["chroma"].cpu().numpy().astype(np.int64) - ref["chroma"].astype(np.int64))[rows]
    assert int(d.max()) <= 1, (int(d.max()), float((d > 1).mean()))


def test_placed_dropouts(torch, N, dec, oracle, cap0, ref0):
    """Capture 0 with RF dips placed inside and outside the K7 scan windows
    (the oracle's entries for it are pinned in test_oracle_dropouts)."""
    from test_oracle_dropouts import dipped_capture
    cap, _ = dipped_capture(cap0, ref0["chunks"][0]["ll_refined"])
    ref = oracle.decode(cap, MHZ, 64, trace=True)
    assert list(ref["do_count"][:3]) == [0, 1, 1]
    out, outc = decode_both(dec, torch, cap)
    compare(out, ref)
    compare(outc, ref)


@pytest.mark.parametrize("start", [1, 2])
def test_capture_starting_on_either_track(torch, N, dec, oracle, start):
    """A capture that starts `start` recorded fields into the tape: the
    chroma track trial picks track start % 2 (the other heterodyne rotation
    for odd starts), as the oracle does."""
    g = N.generate(1, n_samples=(start + 4) * SPF)
    cap = np.ascontiguousarray(g[start * SPF:])
    ref = oracle.decode(cap, MHZ, 64, trace=True)
    assert ref["track_trial"]["track"] == start % 2 and ref["fields"] == 4
    rows = rows_inside(ref, cap.size)
    out, outc = decode_both(dec, torch, cap)
    compare(out, ref, rows=rows)
    compare(outc, ref, rows=rows)


def test_capture_with_a_dropped_field(torch, N, dec, oracle):
    """One recorded field cut out of the tape: the chroma track flips at the
    first field after the cut (the retry pass), on the GPU as in the oracle."""
    g = N.generate(1, n_samples=14 * SPF)
    cap = np.ascontiguousarray(np.concatenate([g[:6 * SPF], g[7 * SPF:]]))
    ref = oracle.decode(cap, MHZ, 64, trace=True)
    assert ref["fields"] == 13
    rows = rows_inside(ref, cap.size)
    out, outc = decode_both(dec, torch, cap)
    compare(out, ref, rows=rows)
    compare(outc, ref, rows=rows)


def test_output_capacity_too_small(torch, N, dec, cap0):
    d = torch.from_numpy(cap0).to("cuda:0")
    dec.reset()
    with pytest.raises(N.CuvhsError, match="decode_capture"):
        dec.decode_capture(d, max_fields=3)
    # the context stays usable after the refused call
    dec.reset()
    out = dec.decode_capture(d)
    assert out["fields"] == 4


@pytest.mark.parametrize("case", ["one_short", "silent"])
def test_cli_edge_files_match_oracle(tmp_path, oracle, cap0, case):
    exe = os.path.join(ROOT, "cuvhs_amd", "lib", "cuvhs")
    cap = cap0[:SPF - 1] if case == "one_short" else silent(2 * SPF + 5)
    src = str(tmp_path / "cap.u8")
    cap.tofile(src)
    print(f"cli case {case}: {cap.size} samples", flush=True)
    r = subprocess.run([exe, src, str(tmp_path / "g")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    nf = oracle.lib().or_decode_file(src.encode(), str(tmp_path / "o").encode(), MHZ, 64)
    assert nf == (0 if case == "one_short" else 2)
    if case == "silent":
        assert "no VSYNC candidates found in chunk at raw offset 0" in r.stderr
    with open(tmp_path / "g.tbc.json") as a, open(tmp_path / "o.tbc.json") as b:
        assert a.read() == b.read()
    for s in (".tbc", "_chroma.tbc"):
        a = np.fromfile(str(tmp_path / ("g" + s)), np.uint16)
        b = np.fromfile(str(tmp_path / ("o" + s)), np.uint16)
        assert a.size == b.size == nf * 263 * 910 and np.array_equal(a, b)
    # the same capture streamed on stdin (length unknown to the reader)
    with open(src, "rb") as f:
        r = subprocess.run([exe, "-", str(tmp_path / "s")], stdin=f, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "streaming" in r.stderr
    for s in (".tbc", "_chroma.tbc", ".tbc.json"):
        with open(tmp_path / ("g" + s), "rb") as a, open(tmp_path / ("s" + s), "rb") as b:
            assert a.read() == b.read(), s
