"""K3 line location (the reference's CUVHS_K2B_DISABLE path) and the
K3/K4/K5 diagnostic counters on the GPU against the oracle."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from test_oracle_k3 import ideal_field

pytestmark = pytest.mark.gpu

MHZ = 28.0


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


def gpu_k3(torch, N, dec, st, ln, ct):
    nf = ct.size
    d_st, d_ln, d_ct = dev(torch, st.astype(np.int32)), dev(torch, ln.astype(np.int32)), dev(torch, ct.astype(np.int32))
    types = torch.zeros((nf, 800), dtype=torch.int32, device="cuda:0")
    ll = torch.zeros((nf, 525), dtype=torch.float64, device="cuda:0")
    isf = torch.zeros(nf, dtype=torch.int32, device="cuda:0")
    assert N.lib().cuvhs_k3_linelocs(dec.h, N.ptr(d_st), N.ptr(d_ln), N.ptr(d_ct), nf, N.ptr(types), N.ptr(ll),
                                     N.ptr(isf)) == 0
    return types.cpu().numpy(), ll.cpu().numpy(), isf.cpu().numpy()


def test_k3_on_ideal_fields(torch, N, oracle):
    dec = N.Decoder(MHZ, 64)
    for first in (True, False):
        st, ln, ct, _ = ideal_field(first)
        g = gpu_k3(torch, N, dec, st, ln, ct)
        o = oracle.k3_linelocs(st, ln, ct, MHZ)
        n = ct[0]
        assert np.array_equal(g[0][0, :n], o[0][0, :n])
        assert np.array_equal(g[1], o[1]) and np.array_equal(g[2], o[2])
    dec.close()


def test_k3_on_oracle_pulses(torch, N, oracle, ref1_prefix):
    """K3 on the K2 pulse lists of every chunk of the 140-field prefix."""
    dec = N.Decoder(MHZ, 64)
    for c in ref1_prefix["chunks"]:
        nf = c["nf"]
        st, ln, ct = c["k2_starts"][:nf], c["k2_lens"][:nf], c["k2_count"][:nf]
        g = gpu_k3(torch, N, dec, st, ln, ct)
        o = oracle.k3_linelocs(st, ln, ct, MHZ)
        for f in range(nf):
            assert np.array_equal(g[0][f, :ct[f]], o[0][f, :ct[f]])
        assert np.array_equal(g[1], o[1])
        assert np.array_equal(g[2], o[2])
    dec.close()


def compare(out, ref):
    n = out["fields"]
    assert n == ref["fields"] and n > 0
    for k in ("file_loc", "is_first", "phase_id", "do_count"):
        assert np.array_equal(out[k].cpu().numpy().astype(np.int64), np.asarray(ref[k]).astype(np.int64)), k
    dc = ref["do_count"]
    for k in ("do_line", "do_startx", "do_endx"):
        g = out[k].cpu().numpy()
        for i in range(n):
            assert np.array_equal(g[i, :dc[i]], ref[k][i, :dc[i]]), (k, i)
    for k in ("luma", "chroma"):
        d = np.abs(out[k].cpu().numpy().astype(np.int64) - ref[k].astype(np.int64))
        assert int(d.max()) <= 1, (k, int(d.max()))


@pytest.mark.parametrize("locator", ["k3", "k2b"])
def test_decode_with_diagnostics_vs_oracle(torch, N, oracle, cap1_prefix, locator):
    """The 140-field prefix decoded with either line locator and the debug
    counters on: outputs as the oracle's, counters per chunk equal to the
    oracle's.  One exception: the K4 "refined sync-like" counter reads
    demod_05 at lines up to 524, which for the capture's last field lie past
    the end of the capture, where K1 of the zero-filled tail is
    ill-conditioned in any implementation; there it may differ by a few."""
    ref = oracle.decode(cap1_prefix, MHZ, 64, trace=True, line_locator=1 if locator == "k3" else 0,
                        diagnostics=True)
    dec = N.Decoder(MHZ, 64, line_locator=locator)
    dec.diagnostics(True)
    out = dec.decode_chunked(dev(torch, cap1_prefix))
    compare(out, ref)
    got = [c["diag"] for c in out["chunks"] if c["fields"]]
    assert len(got) == len(ref["diag"])
    for i, (g, r) in enumerate(zip(got, ref["diag"])):
        r = [int(x) for x in r]
        if i + 1 < len(got):
            assert g == r, i
        else:
            assert g[:5] == r[:5] and g[6:] == r[6:]
            assert abs(g[5] - r[5]) <= 8
    totals, _ = dec.diagnostics_read()
    assert list(totals.values()) == [sum(c[k] for c in got) for k in range(9)]
    dec.close()


def test_cli_k3_debug_logs(N, oracle, tmp_path):
    """CUVHS_K2B_DISABLE selects K3 in the CLI, and the CUVHS_DEBUG_K*_LOG
    lines carry the oracle's counters.  Six fields in chunks of four: the
    first chunk's window ends before the end of the capture, so none of its
    pulses come from the zero-filled tail (where K1, and so the K2 pulses K3
    snaps to, are ill-conditioned in any implementation)."""
    cap = N.generate(1, n_samples=6 * 468140)
    src = str(tmp_path / "cap.u8")
    cap.tofile(src)
    env = dict(os.environ, CUVHS_K2B_DISABLE="1", CUVHS_DEBUG_K3_LOG="1", CUVHS_DEBUG_K4_LOG="1",
               CUVHS_DEBUG_K5_LOG="1")
    exe = os.path.join(N.LIB_DIR, "cuvhs")
    r = subprocess.run([exe, "-f", "28", "--batch", "4", src, str(tmp_path / "o")], env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ref = oracle.decode(cap, MHZ, 4, trace=True, line_locator=1, diagnostics=True)
    assert ref["chunks"][0]["nf"] == 4 and len(ref["chunks"]) >= 2
    d = ref["diag"][0]
    assert (f"[K3 spacing] chunk raw_offset=0: bad_spacing={d[0]} isolated_spacing={d[1]} large_jump={d[2]}"
            in r.stderr)
    assert f"[K4 debug] chunk raw_offset=0: large_delta={d[3]} isolated_jump={d[4]} refined_sync_like={d[5]}" in r.stderr
    assert (f"[K5 debug] chunk raw_offset=0: mode=refined oob_pixels={d[6]} bad_geom_lines={d[7]} "
            f"sync_like_pixels={d[8]}" in r.stderr)
    luma = np.fromfile(str(tmp_path / "o.tbc"), np.uint16).astype(np.int64).reshape(-1, 263, 910)
    assert np.abs(luma[:4] - ref["luma"][:4].astype(np.int64)).max() <= 1
