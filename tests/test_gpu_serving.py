"""Contexts are independent: decodes from several host threads at once, on
one GPU, give exactly the outputs of the same decodes run one after another
(the serving loop relies on it)."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_concurrent_contexts_match_sequential():
    import torch
    from cuvhs_amd import native as N
    jobs = [(28.0, 40, 11), (40.0, 30, 12), (35.8, 24, 13)]
    caps = []
    for mhz, nf, seed in jobs:
        spf = N.format_for(mhz).samples_per_field
        caps.append(torch.from_numpy(N.generate(1, sample_rate_mhz=mhz, n_samples=nf * spf, seed=seed)).cuda())
    decs = [N.Decoder(mhz, 16) for mhz, _, _ in jobs]
    seq = [decs[i].decode_capture(caps[i]) for i in range(len(jobs))]
    seq = [{k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in o.items()} for o in seq]
    par = [None] * len(jobs)

    def work(i):
        decs[i].reset()
        o = decs[i].decode_capture(caps[i])
        par[i] = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in o.items()}

    for _ in range(2):
        th = [threading.Thread(target=work, args=(i,)) for i in range(len(jobs))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for i in range(len(jobs)):
            assert par[i]["fields"] == seq[i]["fields"] > 0
            for k in ("luma", "chroma", "file_loc", "is_first", "phase_id", "do_count"):
                assert np.array_equal(par[i][k], seq[i][k]), (i, k)
    for d in decs:
        d.close()
