"""The reduced-precision forms against the FP64 form and the oracle on the
140-field prefix of the 600-field 28 MHz workload.

The FP64 form reproduces the reference's arithmetic; the other forms are
alternatives whose error is measured, not assumed away.

FP32 (K1 transforms, phase and envelope in FP32): K1 errors of a few tens of
Hz at most (the sync threshold decisions sit ~1 MHz below the blanking level,
and the u16 luma step is ~20 Hz), every field-level decision (field count,
fileLoc, isFirstField, fieldPhaseID, dropout entries) equal, and u16 outputs
within 1 LSB except at a few samples per million.  Those come from K4's
integer window positions: the median windows start at (int)(crossing + 1 us),
so a crossing that moves by the FP32 error across an integer shifts a window
by one sample, the sync/porch levels by a noise sample, and that line's
refined start by up to ~0.1 sample, which reaches a few LSB on the steepest
colour-bar edges (measured on this prefix: luma 10 samples over 1 LSB, max 3;
chroma 74, max 5).  demod_05 is stored in FP32, so a sample within FP32
error of the sync threshold can flip its decision (7 to 12 over the 600-field
capture, no field-level consequence).

Mixed (K1 part A on the matrix cores with f16 or bf16 hi/lo taps, part B in
FP32, demod_05 stored in FP64, FP64 guard band): every sync-threshold
decision is the FP64 form's, because part B flags each 256 samples with a
demod_05 output within the guard width of the threshold (8 Hz f16, 32 Hz
bf16; the forms' measured error away from the recomputed blocks is at most
1.6 Hz f16 and 4.0 Hz bf16) and every FP64 part-B block holding such a
sample, or reading an envelope tile below the guard's ratio (0.1 f16, 0.5
bf16: dropouts, whose flattened sync edges would let K4 turn even a 1.5 Hz
error into a visible line-start shift), is recomputed in FP64.  The flips are therefore zero, not merely rare.  The remaining
output differences are continuous: demod_05's error moves K4's interpolated
crossings by ~3e-5 samples, which changes luma by at most 1 LSB and chroma
(a 3.58 MHz carrier of up to ~2e4 LSB) by at most 2 LSB where the error
meets a rounding boundary (1 sample in 1.4e8 over the 600-field capture).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MHZ = 28.0
MIXED = ("mixed-fp16", "mixed-bf16")
GUARD_DELTA = {"mixed-fp16": 8.0, "mixed-bf16": 32.0}


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def test_precision_config_validation(torch):
    import ctypes
    from cuvhs_amd import native as N
    h = N._vp()
    P = N.PRECISIONS
    cases = [(P["fp64"], 0, 0, True), (P["fp32"], 0, 0, True), (P["mixed-fp16"], 0, 0, True),
             (P["mixed-bf16"], 0, 0, True), (7, 0, 0, False),
             # the matrix-core part A takes u8 samples without a per-read DC
             (P["mixed-fp16"], 1, 0, False), (P["mixed-bf16"], 2, 0, False), (P["mixed-fp16"], 0, 1, False)]
    for prec, fmt, dc, ok in cases:
        cfg = N.Config(0, MHZ, 8, prec, None, 0, fmt, dc, 0)
        rc = N.lib().cuvhs_create(ctypes.byref(cfg), ctypes.byref(h))
        assert (rc == 0) == ok, (prec, fmt, dc, rc)
        if rc == 0:
            N.lib().cuvhs_destroy(h)


def test_fp32_form_vs_fp64_form(torch, cap1_prefix):
    from cuvhs_amd import precision
    d = torch.from_numpy(cap1_prefix).to("cuda:0")
    rep = precision.compare(d, MHZ, 64, "fp32")
    m = rep["mismatches"]
    print(rep)
    assert m["fields"] == 0 and m["file_loc"] == 0 and m["is_first"] == 0 and m["phase_id"] == 0
    assert m["dropout_fields"] == 0
    # sample-level sync threshold flips: rare (a sample within FP32 error of
    # the threshold); 7 to 12 over the whole 600-field capture when measured
    assert m["sync_threshold_flips"] <= 20
    k1 = rep["k1_error"]
    assert k1["demod_05"]["max"] < 64.0 and k1["demod_05"]["rms"] < 2.0
    assert k1["demod"]["max"] < 256.0 and k1["demod"]["rms"] < 2.0
    n_out = rep["fields"] * 263 * 910
    for k in ("luma", "chroma"):
        e = rep["output_error_lsb"][k]
        assert e["max"] <= 8 and e["rms"] < 0.2, (k, e)
        assert e["samples_over_1lsb"] <= 1e-5 * n_out, (k, e)


@pytest.mark.parametrize("form", MIXED)
def test_mixed_form_vs_fp64_form(torch, cap1_prefix, form):
    """Zero sync-threshold flips and identical field-level decisions: the
    guard band makes every threshold decision the FP64 form's."""
    from cuvhs_amd import precision
    d = torch.from_numpy(cap1_prefix).to("cuda:0")
    rep = precision.compare(d, MHZ, 64, form)
    print(rep)
    m = rep["mismatches"]
    assert m == {"fields": 0, "sync_threshold_flips": 0, "file_loc": 0, "is_first": 0, "phase_id": 0,
                 "dropout_fields": 0}, m
    k1 = rep["k1_error"]
    # outside the recomputed blocks the error stays well inside the guard
    # width (measured max 1.6 Hz f16, 4.0 Hz bf16)
    assert k1["demod_05"]["max"] < GUARD_DELTA[form] / 2, k1
    assert k1["demod_05"]["rms"] < 1.0 and k1["demod"]["rms"] < 1.0, k1
    n_out = rep["fields"] * 263 * 910
    lu, ch = rep["output_error_lsb"]["luma"], rep["output_error_lsb"]["chroma"]
    assert lu["max"] <= 1, lu
    assert ch["max"] <= 2 and ch["samples_over_1lsb"] <= 1e-6 * n_out, ch
    assert lu["rms"] < 0.2 and ch["rms"] < 0.2, (lu, ch)


@pytest.mark.parametrize("form", MIXED)
def test_mixed_form_guard_is_selective(torch, cap1_prefix, form):
    """The guard recomputes some blocks (the prefix has sync edges near the
    threshold and dropouts) but only a small share of them."""
    from cuvhs_amd import native as N
    dec = N.Decoder(MHZ, 64, precision=form)
    out = dec.decode_chunked(torch.from_numpy(cap1_prefix).to("cuda:0"))
    spfp = dec.fmt.spfp
    dec.close()
    ga = sum(c["guard_blocks"][0] for c in out["chunks"])
    gb = sum(c["guard_blocks"][1] for c in out["chunks"])
    fk1 = sum(c["fields_k1"] for c in out["chunks"] if c["fields"])
    # FP64 design at 28 MHz: 2925 and 2271 outputs per part-A / part-B block
    na, nb = fk1 * -(-spfp // 2925), fk1 * -(-spfp // 2271)
    print(form, "guard blocks", ga, gb, "of", na, nb)
    assert 0 < ga < 0.25 * na and 0 < gb < 0.15 * nb


@pytest.mark.parametrize("form", ("fp32",) + MIXED)
def test_reduced_form_vs_oracle(torch, cap1_prefix, ref1_prefix, form):
    """Field-level decisions of the reduced forms equal the oracle's."""
    from cuvhs_amd import native as N
    dec = N.Decoder(MHZ, 64, precision=form)
    out = dec.decode_capture(torch.from_numpy(cap1_prefix).to("cuda:0"))
    ref = ref1_prefix
    n = out["fields"]
    assert n == ref["fields"]
    for k in ("file_loc", "is_first", "phase_id", "do_count"):
        assert np.array_equal(out[k].cpu().numpy().astype(np.int64), np.asarray(ref[k]).astype(np.int64)), k
    # the FP64 form is within 1 LSB of the oracle; the mixed forms within
    # 1 (luma) and 2 (chroma) LSB of the FP64 form
    lim = 8 if form == "fp32" else 3
    for k in ("luma", "chroma"):
        d = np.abs(out[k].cpu().numpy().astype(np.int64) - ref[k].astype(np.int64))
        assert d.max() <= lim and (d > 1).sum() <= 1e-5 * d.size, (k, int(d.max()), int((d > 1).sum()))
    dec.close()
