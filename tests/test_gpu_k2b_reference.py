"""GPU K2b line location (cuvhs_linelocs) against the reference's own host K2b.

The vectors in tests/golden/ref_k2b_*.npz hold, for seven degraded NTSC
captures (28 / 35.8 / 40 MHz, SNR down to 18 dB, eight dropouts per field, a
field cut out of the tape) and four PAL captures (the reference's PAL
branches, pipeline.cu:707-717, 811), the K2 pulse lists of each 64-field chunk and what the
reference's host K2b (pipeline.cu:133-148, 410-893, compiled as it lies; see
tests/golden/make_ref_k2b_fixtures.py) returns for them, with the cadence
state carried across chunks.  The HIP K2b gets the same pulses chunk after
chunk on one context (whose cadence state carries the same way) and must
return the same validated pulses, types, line locations and is_first, bit
for bit.
"""
import glob
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

sys.path.insert(0, GOLDEN)
import make_ref_k2b_fixtures as FX  # noqa: E402

pytestmark = pytest.mark.gpu

K2B_FILES = sorted(glob.glob(os.path.join(GOLDEN, "ref_k2b_*.npz")))


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.mark.parametrize("path", K2B_FILES, ids=lambda p: os.path.basename(p)[8:-4])
def test_gpu_k2b_equals_reference(torch, path):
    from cuvhs_amd import native as N
    z = np.load(path)
    mhz = float(z["mhz"])
    system = "pal" if "system" in z.files and int(z["system"]) == 1 else "ntsc"
    lpf = N.GEOMETRY[system][0]
    chunks = FX.unpack_k2b(z)
    print(f"case {os.path.basename(path)}: {len(chunks)} chunks at {mhz} MHz, {system}", flush=True)
    dec = N.Decoder(mhz, int(z["batch"]), system=system)
    try:
        for ci, c in enumerate(chunks):
            nf = c["nf"]
            dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
            offs = dev(c["offsets"].astype(np.uint64))
            ps = dev(c["k2_starts"].reshape(-1))
            pl = dev(c["k2_lens"].reshape(-1))
            pc = dev(c["k2_count"])
            pt = torch.zeros_like(ps)
            ll = torch.zeros(nf * lpf, dtype=torch.float64, device="cuda:0")
            isf = torch.zeros(nf, dtype=torch.int32, device="cuda:0")
            rc = N.lib().cuvhs_linelocs(dec.h, c["raw_offset"], N.ptr(offs), nf, N.ptr(ps), N.ptr(pl), N.ptr(pc),
                                        N.ptr(pt), N.ptr(ll), N.ptr(isf))
            assert rc == 0, rc
            e = c["expected"]
            assert np.array_equal(pc.cpu().numpy(), e["count"]), ci
            got = {"starts": ps, "lens": pl, "types": pt}
            cnt = e["count"]
            for k, t in got.items():
                a = t.cpu().numpy().reshape(nf, 800)
                for f in range(nf):
                    assert np.array_equal(a[f, :cnt[f]], e[k][f, :cnt[f]]), (ci, f, k)
            assert np.array_equal(ll.cpu().numpy().reshape(nf, lpf), e["linelocs"]), ci
            assert np.array_equal(isf.cpu().numpy(), e["is_first"]), ci
    finally:
        dec.close()
