"""The oracle's and the decoder's restatements of the reference's host-only
code (video_format.cpp, raw_reader.cpp, tbc_writer.cpp) against output of the
reference's own compiled sources: committed fixtures, plus a live run of
oracle/_ref/ref_host where the reference is mounted."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

REF = os.path.join(ROOT, "oracle", "_ref", "ref_host")


class OrFormat(ctypes.Structure):
    _fields_ = [("sample_rate", ctypes.c_double), ("samples_per_line", ctypes.c_int),
                ("samples_per_field", ctypes.c_int), ("spfp", ctypes.c_int), ("k1_fft_size", ctypes.c_int),
                ("chroma_fft_size", ctypes.c_int)] + [(n, ctypes.c_double) for n in (
                    "hsync_width", "eq_pulse_width", "vsync_width", "ire0", "hz_ire", "vsync_ire", "sync_tip_hz",
                    "pulse_threshold_hz", "output_zero", "output_scale", "output_rate", "fsc", "chroma_under",
                    "burst_abs_ref", "black16b", "white16b")] + [("burst_start", ctypes.c_int),
                                                                  ("burst_end", ctypes.c_int)]


def ref_formats():
    with open(os.path.join(GOLDEN, "ref_format.json")) as f:
        return json.load(f)


FORMAT_KEYS = [(s, m) for s in ("ntsc", "pal") for m in ("28", "35.8", "40")]


def ref_key(system, mhz):
    return ("pal_" if system == "pal" else "") + mhz


@pytest.mark.parametrize("system,mhz", FORMAT_KEYS)
def test_oracle_format_matches_reference(oracle, system, mhz):
    ref = ref_formats()[ref_key(system, mhz)]
    o = OrFormat()
    L = oracle.lib()
    assert L.or_format_init_sys(oracle.SYSTEMS[system], float(mhz), ctypes.byref(o)) == 0
    names = {n for n, _ in OrFormat._fields_}
    checked = 0
    for k, v in ref.items():
        if k in names:
            assert getattr(o, k) == v, k
            checked += 1
    assert checked >= 17
    # the burst window in output samples as chroma_decode() derives it
    assert o.burst_start == int(ref["burst_start_us"] * 1e-6 * ref["output_rate"] + 0.5)
    assert o.burst_end == int(ref["burst_end_us"] * 1e-6 * ref["output_rate"] + 0.5)


@pytest.mark.parametrize("system,mhz", FORMAT_KEYS)
def test_decoder_format_matches_reference(system, mhz):
    from cuvhs_amd import native
    ref = ref_formats()[ref_key(system, mhz)]
    f = native.format_for(float(mhz), system)
    for k in ("sample_rate", "samples_per_line", "samples_per_field", "pulse_threshold_hz", "ire0", "hz_ire",
              "output_scale", "output_zero", "output_rate", "fsc"):
        assert getattr(f, k) == ref[k], k
    assert f.spfp == ref["samples_per_field"] + 20 * ref["samples_per_line"]
    assert (f.lines_per_frame, f.line_len, f.field_lines) == (
        ref["lines_per_frame"], ref["output_line_len"], ref["output_field_lines"])
    assert f.black16b == (0.0 - ref["vsync_ire"]) * ref["output_scale"] + ref["output_zero"]
    assert f.white16b == (100.0 - ref["vsync_ire"]) * ref["output_scale"] + ref["output_zero"]
    assert f.burst_start == int(ref["burst_start_us"] * 1e-6 * ref["output_rate"] + 0.5)
    assert f.burst_end == int(ref["burst_end_us"] * 1e-6 * ref["output_rate"] + 0.5)
    if system == "ntsc":
        assert (f.black16b, f.white16b, f.burst_start, f.burst_end) == (15360.0, 51200.0, 76, 112)


def test_u8_conversion_exhaustive():
    with open(os.path.join(GOLDEN, "ref_convert_u8.txt")) as f:
        vals = np.array([float(x) for x in f.read().split()])
    assert vals.size == 256
    assert np.array_equal(vals, (np.arange(256) - 128.0) * 256.0)


META = dict(file_loc=[32032, 499165, 966298], is_first=[0, 1, 0], phase_id=[1, 4, 0], do_count=[0, 2, 0],
            do_line=[12, 13], do_startx=[100, 0], do_endx=[910, 57])


WRITER_FIXTURE = {"ntsc": "ref_writer_3fields.tbc.json", "pal": "ref_writer_3fields_pal.tbc.json"}


def oracle_json(oracle, path, system="ntsc"):
    L = oracle.lib()
    L.or_write_json_meta_sys.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_int] + \
        [ctypes.c_void_p] * 7
    arrs = [np.array(META["file_loc"], np.uint64)] + [np.array(META[k], np.int32) for k in (
        "is_first", "phase_id", "do_count", "do_line", "do_startx", "do_endx")]
    assert L.or_write_json_meta_sys(oracle.SYSTEMS[system], 28.0, path.encode(), 3,
                                    *[a.ctypes.data for a in arrs]) == 0
    with open(path) as f:
        return f.read()


@pytest.mark.parametrize("system", ["ntsc", "pal"])
def test_oracle_json_matches_reference_writer(oracle, tmp_path, system):
    with open(os.path.join(GOLDEN, WRITER_FIXTURE[system])) as f:
        golden = f.read()
    assert oracle_json(oracle, str(tmp_path / "o.tbc.json"), system) == golden
    if system == "pal":
        assert '"system": "PAL"' in golden and '"isSourcePal": true' in golden and '"fieldWidth": 1135' in golden


@pytest.mark.skipif(not os.path.exists(REF), reason="reference host build not present")
@pytest.mark.parametrize("system", ["ntsc", "pal"])
def test_live_reference_build_matches_fixtures(tmp_path, system):
    extra = ["pal"] if system == "pal" else []
    out = subprocess.run([REF, "format", "28"] + extra, capture_output=True, text=True, check=True).stdout
    live = {k: float(v) for k, v in (l.split("=") for l in out.split())}
    assert live == ref_formats()[ref_key(system, "28")]
    base = str(tmp_path / "r")
    subprocess.run([REF, "json", base] + extra, check=True, capture_output=True)
    with open(base + ".tbc.json") as f, open(os.path.join(GOLDEN, WRITER_FIXTURE[system])) as g:
        assert f.read() == g.read()


@pytest.mark.parametrize("fmt", ["u8", "s16", "u16"])
@pytest.mark.parametrize("dc", [False, True])
def test_oracle_conversion_matches_reference(oracle, fmt, dc):
    """RawReader::convert for every input format, with and without the
    per-read DC removal: bit-identical to the reference's compiled code."""
    g = np.load(os.path.join(GOLDEN, "ref_convert_formats.npz"))
    out = oracle.convert(g[fmt + "_in"], dc_correct=dc)
    ref = g[fmt + ("_dc" if dc else "") + "_out"]
    assert np.array_equal(out, ref)


def test_conversion_closed_forms(oracle):
    """The normalisations are exact affine maps, and the DC removal leaves a
    zero-mean sequence (to rounding)."""
    x = np.array([0, 1, 127, 128, 255], np.uint8)
    assert np.array_equal(oracle.convert(x), (x.astype(float) - 128.0) * 256.0)
    y = np.array([-32768, -1, 0, 1, 32767], np.int16)
    assert np.array_equal(oracle.convert(y), y.astype(float))
    z = np.array([0, 1, 32768, 65535], np.uint16)
    assert np.array_equal(oracle.convert(z), z.astype(float) - 32768.0)
    w = oracle.convert(np.arange(1000, dtype=np.int16), dc_correct=True)
    assert abs(w.mean()) < 1e-9 and w[0] == -499.5
