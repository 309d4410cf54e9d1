"""GPU parity: every HIP stage on the same inputs as the oracle, then whole
decodes.  Integer decisions must be identical; u16 outputs within 1 LSB."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, s16

pytestmark = pytest.mark.gpu

MHZ = 28.0
SPF = 468140
SPFP = 503740
THR = 3542857.1428571427


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


@pytest.mark.parametrize("B,f32", [(1024, 0), (4096, 0), (8192, 0), (16384, 0), (1024, 1), (4096, 1), (8192, 1),
                                   (16384, 1)])
def test_conv_blocks_vs_numpy(torch, N, B, f32):
    """The convolution blocks the decoder runs (K1's 4096/8192-point blocks,
    the pre-band-pass's 16384-point FP64 and FP32 blocks, the chroma line's
    1024-point FP64 and FP32 filters), each on random complex data and a random spectrum given in
    natural bin order (permuted by the library as for its filters), against
    B * ifft(fft(x) * g)."""
    rng = np.random.default_rng(B + f32)
    x = rng.standard_normal(B) + 1j * rng.standard_normal(B)
    g = rng.standard_normal(B) + 1j * rng.standard_normal(B)
    dt, rt = (np.complex64, np.float32) if f32 else (np.complex128, np.float64)
    dx = dev(torch, x.astype(dt).view(rt))
    dg = dev(torch, g.astype(dt).view(rt))
    out = torch.empty_like(dx)
    assert N.lib().cuvhs_selftest_block(B, f32, N.ptr(dx), N.ptr(dg), N.ptr(out)) == 0
    y = out.cpu().numpy().view(dt)
    ref = np.fft.ifft(np.fft.fft(x) * g) * B
    tol = 2e-6 if f32 else 1e-13
    assert np.abs(y - ref).max() / np.abs(ref).max() < tol
    # unsupported shapes are refused, not run
    assert N.lib().cuvhs_selftest_block(2048, f32, N.ptr(dx), N.ptr(dg), N.ptr(out)) == N.E_ARG


def test_atan2_f64_vs_host(torch, N):
    """K1's FP64 arctangent (atan_f64.h) against numpy's: within 4 ulp of
    the result everywhere, signed zeros and axes exact; inputs spread over
    nine decades, near the tan(pi/8) reduction threshold and the diagonal."""
    rng = np.random.default_rng(7)
    n = 1 << 20
    y = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 6, n)
    x = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 6, n)
    k = n // 4
    x[:k] = np.abs(x[:k])
    y[:k] = x[:k] * 0.41421356237309503 * (1 + rng.uniform(-1e-12, 1e-12, k))
    x[k:2 * k] = y[k:2 * k] * (1 + rng.uniform(-1e-9, 1e-9, k))
    sy = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 0.0, 0.0, 1.0, 1.0, -1.0, 3.0])
    sx = np.array([0.0, 0.0, -0.0, -0.0, 0.0, 0.0, 1.0, -1.0, 1.0, -1.0, -1.0, -3.0])
    y = np.concatenate([y, sy])
    x = np.concatenate([x, sx])
    out = torch.empty(y.size, dtype=torch.float64, device="cuda:0")
    assert N.lib().cuvhs_selftest_atan2(N.ptr(dev(torch, y)), N.ptr(dev(torch, x)), N.ptr(out), y.size) == 0
    got = out.cpu().numpy()
    want = np.arctan2(y, x)
    ulp = np.abs(got - want) / np.spacing(np.abs(want))
    assert ulp[:n].max() <= 4.0
    assert np.array_equal(got[n:], want[n:]) and np.array_equal(np.signbit(got[n:]), np.signbit(want[n:]))


def test_fm_demod_vs_oracle(torch, N, oracle, cap0):
    nf = 4
    dec = N.Decoder(MHZ, 64)
    chunk = np.zeros(nf * SPFP, np.uint8)
    chunk[:cap0.size] = cap0
    dc = dev(torch, chunk)
    dm = torch.zeros(nf * SPFP, dtype=torch.float64, device="cuda:0")
    d5 = torch.zeros_like(dm)
    env = torch.zeros_like(dm)
    assert N.lib().cuvhs_fm_demod(dec.h, N.ptr(dc), cap0.size, nf, N.ptr(dm), N.ptr(d5), N.ptr(env)) == 0
    odm, od5, oenv = oracle.fm_demod(s16(cap0, nf * SPFP), MHZ, nf)
    # compare where the capture has data (past EOF the reference takes the
    # angle of round-off noise of an all-zero window)
    v = cap0.size - 4096
    g = {k: t.cpu().numpy()[:v] for k, t in (("dm", dm), ("d5", d5), ("env", env))}
    assert np.abs(g["env"] - oenv[:v]).max() / oenv[:v].max() < 1e-11
    assert np.abs(g["dm"] - odm[:v]).max() < 1e-3        # Hz
    assert np.abs(g["d5"] - od5[:v]).max() < 1e-3
    assert np.array_equal(g["d5"] <= THR, od5[:v] <= THR)
    dec.close()


def test_vsync_and_pulses_on_oracle_signal(torch, N, oracle, cap0, ref0):
    dec = N.Decoder(MHZ, 64)
    nf = 4
    _, od5, _ = oracle.fm_demod(s16(cap0, nf * SPFP), MHZ, nf)
    pad = np.zeros(64 * SPFP + SPFP)
    pad[:od5.size] = od5
    t5 = dev(torch, pad)
    cap = 80
    idx = torch.zeros(cap, dtype=torch.int32, device="cuda:0")
    cnt = ctypes.c_int()
    assert N.lib().cuvhs_vsync_discover(dec.h, N.ptr(t5), nf * SPFP, N.ptr(idx), cap, ctypes.byref(cnt)) == 0
    oidx = np.zeros(cap, np.int32)
    ocnt = oracle.lib().or_vsync_discover(MHZ, od5.ctypes.data, nf * SPFP, oidx.ctypes.data, cap)
    assert cnt.value == ocnt
    assert np.array_equal(idx.cpu().numpy()[:ocnt], oidx[:ocnt])
    c0 = ref0["chunks"][0]
    nfc = c0["nf"]
    offs = dev(torch, c0["offsets"].astype(np.uint64))
    st = torch.zeros(nfc * 800, dtype=torch.int32, device="cuda:0")
    ln = torch.zeros_like(st)
    ct = torch.zeros(nfc, dtype=torch.int32, device="cuda:0")
    assert N.lib().cuvhs_sync_pulses(dec.h, N.ptr(t5), nf * SPFP, N.ptr(offs), nfc, N.ptr(st), N.ptr(ln), N.ptr(ct)) == 0
    assert np.array_equal(ct.cpu().numpy(), c0["k2_count"])
    stn, lnn = st.cpu().numpy().reshape(nfc, 800), ln.cpu().numpy().reshape(nfc, 800)
    for f in range(nfc):
        k = c0["k2_count"][f]
        assert np.array_equal(stn[f, :k], c0["k2_starts"][f, :k])
        assert np.array_equal(lnn[f, :k], c0["k2_lens"][f, :k])
    dec.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_pulse_scans_brute_force_signals(torch, N, oracle, seed):
    """Synthetic runs: at sample 0, open at the end, every width class."""
    from test_oracle_stages import synth_pulses
    n = 3 * SPFP
    x = synth_pulses(seed + 10, n, 900)
    dec = N.Decoder(MHZ, 4)
    pad = np.zeros(4 * SPFP + SPFP)
    pad[:n] = x
    tx = dev(torch, pad)
    idx = torch.zeros(2000, dtype=torch.int32, device="cuda:0")
    cnt = ctypes.c_int()
    assert N.lib().cuvhs_vsync_discover(dec.h, N.ptr(tx), n, N.ptr(idx), 2000, ctypes.byref(cnt)) == 0
    oidx = np.zeros(2000, np.int32)
    ocnt = oracle.lib().or_vsync_discover(MHZ, x.ctypes.data, n, oidx.ctypes.data, 2000)
    assert cnt.value == ocnt and np.array_equal(idx.cpu().numpy()[:ocnt], oidx[:ocnt])
    offs_np = np.array([0, 99991, SPFP + 3, 2 * SPFP - 777], np.uint64)
    st = torch.zeros(4 * 800, dtype=torch.int32, device="cuda:0")
    ln = torch.zeros_like(st)
    ct = torch.zeros(4, dtype=torch.int32, device="cuda:0")
    assert N.lib().cuvhs_sync_pulses(dec.h, N.ptr(tx), n, N.ptr(dev(torch, offs_np)), 4, N.ptr(st), N.ptr(ln),
                                     N.ptr(ct)) == 0
    ost, oln, oct_ = np.zeros(3200, np.int32), np.zeros(3200, np.int32), np.zeros(4, np.int32)
    xo = np.zeros(4 * SPFP + SPFP)
    xo[:n] = x
    oracle.lib().or_sync_pulses(MHZ, xo.ctypes.data, offs_np.ctypes.data, 4, ost.ctypes.data, oln.ctypes.data,
                                oct_.ctypes.data)
    assert np.array_equal(ct.cpu().numpy(), oct_)
    g_st, g_ln = st.cpu().numpy(), ln.cpu().numpy()
    for f in range(4):
        k = oct_[f]
        assert np.array_equal(g_st[f * 800:f * 800 + k], ost[f * 800:f * 800 + k])
        assert np.array_equal(g_ln[f * 800:f * 800 + k], oln[f * 800:f * 800 + k])
    dec.close()


def k2b_on_chunk(torch, N, dec, c, state_from=None):
    nfc = c["nf"]
    offs = dev(torch, c["offsets"].astype(np.uint64))
    ps = dev(torch, c["k2_starts"].astype(np.int32).reshape(-1))
    pl = dev(torch, c["k2_lens"].astype(np.int32).reshape(-1))
    pc = dev(torch, c["k2_count"].astype(np.int32))
    pt = torch.zeros_like(ps)
    ll = torch.zeros(nfc * 525, dtype=torch.float64, device="cuda:0")
    isf = torch.zeros(nfc, dtype=torch.int32, device="cuda:0")
    assert N.lib().cuvhs_linelocs(dec.h, c["raw_offset"], N.ptr(offs), nfc, N.ptr(ps), N.ptr(pl), N.ptr(pc),
                                  N.ptr(pt), N.ptr(ll), N.ptr(isf)) == 0
    return (ll.cpu().numpy().reshape(nfc, 525), isf.cpu().numpy(), pc.cpu().numpy(),
            ps.cpu().numpy().reshape(nfc, 800), pt.cpu().numpy().reshape(nfc, 800))


def test_k2b_linelocs_on_oracle_pulses(torch, N, ref0):
    dec = N.Decoder(MHZ, 64)
    for c in ref0["chunks"]:
        ll, isf, cnt, starts, types = k2b_on_chunk(torch, N, dec, c)
        assert np.array_equal(ll, c["ll_coarse"])
        assert np.array_equal(isf, c["is_first_k2b"])
        assert np.array_equal(cnt, c["vp_count"])
        for f in range(c["nf"]):
            k = cnt[f]
            assert np.array_equal(starts[f, :k], c["vp_starts"][f, :k])
            assert np.array_equal(types[f, :k], c["vp_types"][f, :k])
    dec.close()


def test_hsync_refine_and_resample_on_oracle_inputs(torch, N, oracle, cap0, ref0):
    dec = N.Decoder(MHZ, 64)
    nf = 4
    odm, od5, _ = oracle.fm_demod(s16(cap0, nf * SPFP), MHZ, nf)
    c = ref0["chunks"][0]
    nfc = c["nf"]
    ll = dev(torch, c["ll_coarse"].reshape(-1).copy())
    assert N.lib().cuvhs_hsync_refine(dec.h, N.ptr(dev(torch, od5)), nfc * SPFP, N.ptr(ll), nfc) == 0
    assert np.array_equal(ll.cpu().numpy().reshape(nfc, 525), c["ll_refined"])
    luma = torch.zeros(nfc * 263 * 910, dtype=torch.uint16, device="cuda:0")
    llr = dev(torch, c["ll_refined"].reshape(-1).copy())
    assert N.lib().cuvhs_tbc_resample(dec.h, N.ptr(dev(torch, odm)), nfc * SPFP, N.ptr(llr), nfc, N.ptr(luma)) == 0
    exp = np.zeros(nfc * 263 * 910, np.uint16)
    oracle.lib().or_tbc_resample(MHZ, odm.ctypes.data, nfc * SPFP, c["ll_refined"].ctypes.data, nfc, exp.ctypes.data)
    assert np.array_equal(luma.cpu().numpy(), exp)
    dec.close()


def test_hsync_refine_closed_form(torch, N):
    from test_oracle_stages import trapezoid_field, HW
    x, ll0, b0s, P, S = trapezoid_field()
    dec = N.Decoder(MHZ, 1)
    ll = dev(torch, ll0)
    assert N.lib().cuvhs_hsync_refine(dec.h, N.ptr(dev(torch, x)), SPFP, N.ptr(ll), 1) == 0
    g = ll.cpu().numpy()
    for l in range(10, 525):
        if b0s[l] > 0:
            assert abs(g[l] - (b0s[l] + 4 * (THR - S) / (P - S) - HW + 2.25 * 0.7)) < 1e-9
    dec.close()


def compare_decode(out, ref, n_expected=None):
    n = out["fields"]
    assert n == ref["fields"]
    if n_expected is not None:
        assert n == n_expected
    for k in ("file_loc", "is_first", "phase_id", "do_count"):
        assert np.array_equal(out[k].cpu().numpy().astype(np.int64), np.asarray(ref[k]).astype(np.int64)), k
    dc = ref["do_count"]
    for k in ("do_line", "do_startx", "do_endx"):
        g = out[k].cpu().numpy()
        for i in range(n):
            assert np.array_equal(g[i, :dc[i]], ref[k][i, :dc[i]]), (k, i)
    worst = {}
    for k in ("luma", "chroma"):
        d = np.abs(out[k].cpu().numpy().astype(np.int64) - ref[k].astype(np.int64))
        worst[k] = int(d.max())
        assert worst[k] <= 1, (k, worst[k], int((d > 1).sum()))
    return worst


def test_decode_capture0(torch, N, cap0, ref0):
    dec = N.Decoder(MHZ, 64)
    out = dec.decode_capture(dev(torch, cap0))
    compare_decode(out, ref0, 4)
    dec.close()


def test_decode_capture1_prefix(torch, N, cap1_prefix, ref1_prefix):
    """140 fields of the 600-field workload: three chunks (64, 64, 12), SNR
    25-40 dB, dropouts and wow/flutter; carried K2b/chroma/parity state."""
    assert [c["nf"] for c in ref1_prefix["chunks"]] == [64, 64, 12]
    assert ref1_prefix["do_count"].sum() > 0
    dec = N.Decoder(MHZ, 64)
    out = dec.decode_capture(dev(torch, cap1_prefix))
    compare_decode(out, ref1_prefix, 140)
    # same context again after reset: identical
    dec.reset()
    out2 = dec.decode_capture(dev(torch, cap1_prefix))
    for k in ("luma", "chroma", "file_loc"):
        assert torch.equal(out[k], out2[k])
    dec.close()


def test_chroma_and_dropout_stages(torch, N, oracle, cap1_prefix, ref1_prefix):
    """K6 and K7 alone on the oracle's line locations of the first chunk."""
    c = ref1_prefix["chunks"][0]
    nfc = c["nf"]
    dec = N.Decoder(MHZ, 64)
    chunk = cap1_prefix[:64 * SPFP]
    out = torch.zeros(nfc * 263 * 910, dtype=torch.uint16, device="cuda:0")
    phase = np.zeros(nfc, np.int32)
    ll = dev(torch, c["ll_refined"].reshape(-1).copy())
    rc = N.lib().cuvhs_chroma(dec.h, N.ptr(dev(torch, chunk)), chunk.size, N.ptr(ll), nfc, N.ptr(out),
                              phase.ctypes.data)
    assert rc == 0
    assert np.array_equal(phase, ref1_prefix["phase_id"][:nfc])
    # oracle chroma before concealment = final chroma except on concealed rows
    _, _, env = oracle.fm_demod(s16(chunk, 64 * SPFP), MHZ, 64)
    luma = torch.zeros_like(out)
    lines = torch.zeros(nfc * 512, dtype=torch.int32, device="cuda:0")
    sx, ex, cnt = torch.zeros_like(lines), torch.zeros_like(lines), torch.zeros(nfc, dtype=torch.int32, device="cuda:0")
    envp = np.zeros(64 * SPFP + SPFP)
    envp[:env.size] = env
    rc = N.lib().cuvhs_dropouts(dec.h, N.ptr(dev(torch, envp)), N.ptr(ll), nfc, N.ptr(luma), N.ptr(out),
                                N.ptr(lines), N.ptr(sx), N.ptr(ex), N.ptr(cnt))
    assert rc == 0
    assert np.array_equal(cnt.cpu().numpy(), ref1_prefix["do_count"][:nfc])
    g = out.cpu().numpy().reshape(nfc, 263, 910).astype(np.int64)
    d = np.abs(g - ref1_prefix["chroma"][:nfc].astype(np.int64))
    assert d.max() <= 1
    dec.close()


def test_cli_matches_oracle_files(tmp_path, oracle, cap0):
    exe = os.path.join(ROOT, "cuvhs_amd", "lib", "cuvhs")
    src = str(tmp_path / "cap.u8")
    cap0.tofile(src)
    r = subprocess.run([exe, src, str(tmp_path / "g")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # refuses to overwrite without --overwrite (tbc_writer.cpp:16-31)
    r2 = subprocess.run([exe, src, str(tmp_path / "g")], capture_output=True, text=True, timeout=120)
    assert r2.returncode == 1 and "use --overwrite" in r2.stderr
    assert oracle.lib().or_decode_file(src.encode(), str(tmp_path / "o").encode(), MHZ, 64) == 4
    with open(tmp_path / "g.tbc.json") as a, open(tmp_path / "o.tbc.json") as b:
        assert a.read() == b.read()
    for s in (".tbc", "_chroma.tbc"):
        a = np.fromfile(str(tmp_path / ("g" + s)), np.uint16).astype(np.int64)
        b = np.fromfile(str(tmp_path / ("o" + s)), np.uint16).astype(np.int64)
        assert a.size == b.size == 4 * 263 * 910 and np.abs(a - b).max() <= 1
