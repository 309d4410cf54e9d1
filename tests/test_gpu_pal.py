"""PAL decodes on the GPU against the oracle.

The reference's PAL format is its NTSC pipeline and K1 filters run on PAL's
constants (video_format.cpp:48-87: 625 lines, 313 output lines of 1135
samples per field, 5 EQ pulses, PAL levels, a 2048-point chroma line
transform) with PAL's branches of K2b (pipeline.cu:707-717, 811) and K3
(line_locs.cu:219-235).  The oracle's PAL restatement is pinned to the
reference's compiled VideoFormat, TBCWriter and K2b (tests/test_reference_
host.py, tests/test_reference_k2b.py); these tests hold the GPU to it on
synthetic PAL captures (the generator's signal model on PAL's line
structure and levels), with the same bar as NTSC: identical field-level
outputs, u16 luma and chroma within 1 LSB."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


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


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def pal_capture(N, mhz, config, fields, **over):
    spf = N.format_for(mhz, "pal").samples_per_field
    kw = dict(over)
    if mhz != 28.0 and config != 2:
        kw["sample_rate_mhz"] = mhz
    return N.generate(config, system="pal", n_samples=fields * spf, **kw)


def compare(out, ref, lim=1):
    n = out["fields"]
    assert n == ref["fields"]
    for k in ("file_loc", "is_first", "phase_id", "do_count"):
        assert np.array_equal(out[k].cpu().numpy().astype(np.int64), np.asarray(ref[k]).astype(np.int64)), k
    dc = ref["do_count"]
    for k in ("do_line", "do_startx", "do_endx"):
        g = out[k].cpu().numpy()
        for i in range(n):
            assert np.array_equal(g[i, :dc[i]], ref[k][i, :dc[i]]), (k, i)
    for k in ("luma", "chroma"):
        a = out[k].cpu().numpy()
        assert a.shape[1:] == (313, 1135), a.shape
        d = np.abs(a.astype(np.int64) - ref[k].astype(np.int64))
        assert int(d.max()) <= lim, (k, int(d.max()), int((d > 1).sum()))


@pytest.fixture(scope="module")
def pal28(N, oracle):
    """100 PAL fields at 28 MHz (the 600-field workload's degraded signal
    model: SNR 25-40 dB, dropouts, wow/flutter) and the oracle's decode."""
    cap = pal_capture(N, 28.0, 1, 100)
    return cap, oracle.decode(cap, 28.0, 32, trace=True, system="pal")


def test_pal_decode_matches_oracle(torch, N, pal28):
    """F = 32: three full chunks and a short one."""
    cap, ref = pal28
    assert ref["fields"] >= 98 and ref["do_count"].sum() > 0
    # the capture exercises both field parities and the whole phase cycle
    assert set(ref["is_first"].tolist()) == {0, 1} and len(set(ref["phase_id"].tolist())) == 4
    dec = N.Decoder(28.0, 32, system="pal")
    assert (dec.fmt.system, dec.fmt.field_lines, dec.fmt.line_len) == (1, 313, 1135)
    out = dec.decode_capture(dev(torch, cap))
    compare(out, ref)
    # the chunk-by-chunk entry point (the CLI's loop) gives the same
    dec.reset()
    ch = dec.decode_chunked(dev(torch, cap))
    compare(ch, ref)
    assert [c["raw_offset"] for c in ch["chunks"] if c["fields"]] == [c["raw_offset"] for c in ref["chunks"]]
    dec.close()


def test_pal_40mhz_headswitch_matches_oracle(torch, N, oracle):
    """70 PAL fields of the 40 MHz head-switch model (0-4 us offsets in
    field groups), F = 32: 8192-point video blocks at 40 MHz."""
    cap = pal_capture(N, 40.0, 2, 70)
    ref = oracle.decode(cap, 40.0, 32, trace=False, system="pal")
    assert ref["fields"] >= 68
    dec = N.Decoder(40.0, 32, system="pal")
    out = dec.decode_capture(dev(torch, cap))
    compare(out, ref)
    dec.close()


def test_pal_k3_matches_oracle(torch, N, oracle, pal28):
    """K3 line location (CUVHS_K2B_DISABLE) with PAL's field-parity rule."""
    cap, _ = pal28
    cap = cap[: 60 * N.format_for(28.0, "pal").samples_per_field]
    ref = oracle.decode(cap, 28.0, 32, trace=False, line_locator=1, system="pal")
    dec = N.Decoder(28.0, 32, line_locator="k3", system="pal")
    out = dec.decode_capture(dev(torch, cap))
    compare(out, ref)
    dec.close()


@pytest.mark.parametrize("form", ["mixed-fp16", "fp32"])
def test_pal_reduced_forms(torch, N, pal28, form):
    """The mixed form keeps every sync-threshold decision of the FP64 form
    (guard band) and its outputs within 1 (luma) / 2 (chroma) LSB; the plain
    FP32 form keeps the field-level decisions."""
    from cuvhs_amd import precision
    cap, ref = pal28
    cap = cap[: 70 * N.format_for(28.0, "pal").samples_per_field]
    rep = precision.compare(dev(torch, cap), 28.0, 32, form, system="pal")
    print(rep)
    m = rep["mismatches"]
    for k in ("fields", "file_loc", "is_first", "phase_id", "dropout_fields"):
        assert m[k] == 0, (k, m)
    lu, ch = rep["output_error_lsb"]["luma"], rep["output_error_lsb"]["chroma"]
    n_out = rep["fields"] * 313 * 1135
    if form == "mixed-fp16":
        assert m["sync_threshold_flips"] == 0, m
        assert lu["max"] <= 1 and ch["max"] <= 2 and ch["samples_over_1lsb"] <= 1e-6 * n_out, (lu, ch)
    else:
        assert lu["max"] <= 8 and ch["max"] <= 8 and ch["samples_over_1lsb"] <= 1e-5 * n_out, (lu, ch)


def test_pal_graphs_equal_launches(torch, N, pal28):
    """Graph mode of the capture decode gives the launch-by-launch outputs."""
    cap, _ = pal28
    cap = cap[: 70 * N.format_for(28.0, "pal").samples_per_field]
    d = dev(torch, cap)
    a = N.Decoder(28.0, 32, precision="mixed-fp16", system="pal")
    want = a.decode_capture(d)
    a.close()
    g = N.Decoder(28.0, 32, precision="mixed-fp16", system="pal", graphs=True)
    got = g.decode_capture(d)
    g.close()
    assert got["fields"] == want["fields"]
    for k in ("luma", "chroma", "file_loc", "is_first", "phase_id", "do_count"):
        assert np.array_equal(got[k].cpu().numpy(), want[k].cpu().numpy()), k


def test_pal_cli_matches_oracle_files(tmp_path, N, oracle, pal28):
    """`cuvhs --system PAL`: the .tbc.json byte-identical to the oracle's
    writer (the reference's TBCWriter for PAL: system, isSourcePal, 1135 x
    313, PAL levels and burst window), TBC planes within 1 LSB."""
    cap, _ = pal28
    cap = cap[: 40 * N.format_for(28.0, "pal").samples_per_field]
    src = str(tmp_path / "pal.u8")
    cap.tofile(src)
    exe = os.path.join(ROOT, "cuvhs_amd", "lib", "cuvhs")
    r = subprocess.run([exe, "--system", "PAL", "--batch", "32", src, str(tmp_path / "g")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Format: PAL" in r.stderr
    L = oracle.lib()
    d = L.or_create_sys(1, 28.0, 32, 0)
    try:
        n = L.or_decode(d, cap.ctypes.data, cap.size)
        assert L.or_write_outputs(d, str(tmp_path / "o").encode()) == 0
    finally:
        L.or_destroy(d)
    with open(tmp_path / "g.tbc.json") as a, open(tmp_path / "o.tbc.json") as b:
        ja = a.read()
        assert ja == b.read()
    assert '"system": "PAL"' in ja and '"fieldWidth": 1135' in ja and '"fieldHeight": 313' in ja
    for s in (".tbc", "_chroma.tbc"):
        a = np.fromfile(str(tmp_path / ("g" + s)), np.uint16).astype(np.int64)
        b = np.fromfile(str(tmp_path / ("o" + s)), np.uint16).astype(np.int64)
        assert a.size == b.size == n * 313 * 1135 and np.abs(a - b).max() <= 1
