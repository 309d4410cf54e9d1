"""cuvhs_dropouts_noalloc (K7 without a context and without allocating)
against cuvhs_dropouts (the context's K7 on caller buffers).

Inputs: 20 fields of the 600-field 28 MHz benchmark capture (two dropouts per
field on average; fields before the K7 window quirk cuts in get entries),
the envelope from the library's own K1 (cuvhs_fm_demod) and the oracle's
refined line locations, random u16 planes to conceal into.  Two calls in a
row, the second with fewer fields, must each match the context entry point
bit for bit: entries, counts and the concealed planes.  Beyond the scratch
limit of 128 fields the call is refused with CUVHS_E_CAPACITY and a reason.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MHZ = 28.0
SPF = 468140
SPFP = 503740
NF = 20


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def inputs(torch, oracle):
    from cuvhs_amd import native as N
    cap = N.generate(1, n_samples=(NF + 2) * SPF)
    ref = oracle.decode(cap, MHZ, 64, trace=True)
    c = ref["chunks"][0]
    assert c["nf"] >= NF
    ll = np.ascontiguousarray(c["ll_refined"][:NF])
    dec = N.Decoder(MHZ, 64)
    d_cap = torch.from_numpy(cap).to("cuda:0")
    n_valid = min(cap.size, NF * SPFP)
    bufs = [torch.empty(NF * SPFP, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    N.check(N.lib().cuvhs_fm_demod(dec.h, N.ptr(d_cap), n_valid, NF, *[N.ptr(b) for b in bufs]), dec.h, "fm_demod")
    env = bufs[2]
    return dec, env, torch.from_numpy(ll.reshape(-1)).to("cuda:0")


def _run(torch, N, fn, nf, planes):
    luma = planes[0][:nf].clone()
    chroma = planes[1][:nf].clone()
    out = {k: torch.full((nf * 512,), -7, dtype=torch.int32, device="cuda:0") for k in ("lines", "starts", "ends")}
    out["count"] = torch.full((nf,), -7, dtype=torch.int32, device="cuda:0")
    rc = fn(luma, chroma, out)
    torch.cuda.synchronize()
    assert rc == 0, (rc, N.lib().cuvhs_last_error(None))
    cnt = out["count"].cpu().numpy()
    res = {"count": cnt, "luma": luma.cpu().numpy(), "chroma": chroma.cpu().numpy()}
    for k in ("lines", "starts", "ends"):
        a = out[k].cpu().numpy().reshape(nf, 512)
        res[k] = [a[f, :cnt[f]].copy() for f in range(nf)]
    return res


def test_noalloc_dropouts_equal_context_entry_point(torch, inputs):
    from cuvhs_amd import native as N
    dec, env, ll = inputs
    rng = np.random.default_rng(5)
    planes = [torch.from_numpy(rng.integers(0, 65536, (NF, 263, 910), dtype=np.uint16)).to("cuda:0") for _ in range(2)]
    total_entries = 0
    for nf in (NF, NF - 7):   # two calls in a row, the second smaller
        def ctx_fn(luma, chroma, o):
            return N.lib().cuvhs_dropouts(dec.h, N.ptr(env), N.ptr(ll), nf, N.ptr(luma), N.ptr(chroma),
                                          N.ptr(o["lines"]), N.ptr(o["starts"]), N.ptr(o["ends"]), N.ptr(o["count"]))

        def free_fn(luma, chroma, o):
            return N.lib().cuvhs_dropouts_noalloc(MHZ, N.ptr(env), N.ptr(ll), nf, N.ptr(luma), N.ptr(chroma),
                                                  N.ptr(o["lines"]), N.ptr(o["starts"]), N.ptr(o["ends"]),
                                                  N.ptr(o["count"]), None)

        a = _run(torch, N, ctx_fn, nf, planes)
        b = _run(torch, N, free_fn, nf, planes)
        assert np.array_equal(a["count"], b["count"]), nf
        for k in ("lines", "starts", "ends"):
            for f in range(nf):
                assert np.array_equal(a[k][f], b[k][f]), (nf, k, f)
        assert np.array_equal(a["luma"], b["luma"]) and np.array_equal(a["chroma"], b["chroma"]), nf
        total_entries += int(a["count"].sum())
    # the inputs exercise detection and concealment
    assert total_entries >= 10, total_entries


def test_noalloc_dropouts_refuses_beyond_its_scratch(torch, inputs):
    from cuvhs_amd import native as N
    _, env, ll = inputs
    z = N.ptr(env)
    rc = N.lib().cuvhs_dropouts_noalloc(MHZ, z, N.ptr(ll), 129, z, z, z, z, z, z, None)
    assert rc == N.E_CAPACITY
    assert b"128 fields" in N.lib().cuvhs_last_error(None)
    assert N.lib().cuvhs_dropouts_noalloc(MHZ, z, N.ptr(ll), 0, z, z, z, z, z, z, None) == N.E_ARG
