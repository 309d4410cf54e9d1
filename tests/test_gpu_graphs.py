"""Graph mode of cuvhs_decode_capture (cuvhs_config.graphs): cuvhs_create
captures each stage's kernels into a HIP graph once per buffer slot, and the
decode replays them for every chunk, the chunk's values read on the device
from the slot's plan.
It must give exactly the outputs of the launch-by-launch schedule, on every
form and input path, on replays, on a shorter capture after a longer one
(the plan clears what the longer chunk left past the shorter one), after
an output-capacity error, and with contexts decoding in several threads."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def same(a, b):
    """Identical decodes: every output array, dropout lists up to their counts."""
    assert a["fields"] == b["fields"] and a["status"] == b["status"]
    for k in ("luma", "chroma", "file_loc", "is_first", "phase_id", "do_count"):
        assert np.array_equal(a[k].cpu().numpy(), b[k].cpu().numpy()), k
    dc = a["do_count"].cpu().numpy()
    for k in ("do_line", "do_startx", "do_endx"):
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        for i in range(a["fields"]):
            assert np.array_equal(x[i, :dc[i]], y[i, :dc[i]]), (k, i)


def u16_capture(N, fields):
    fmt = N.format_for(28.0)
    cap = N.generate(1, n_samples=fields * fmt.samples_per_field)
    s16 = np.clip(np.round((cap.astype(np.float64) - 128.0) * 200.0 - 2500.0), -32768, 32767).astype(np.int32)
    return (s16 + 32768).astype(np.uint16)


CASES = {
    # name: (sample rate, batch, decoder options, capture maker)
    "fp64_28mhz": (28.0, 64, {}, lambda N: N.generate(1, n_samples=140 * 468140)),
    "mixed_fp16_28mhz": (28.0, 64, {"precision": "mixed-fp16"}, lambda N: N.generate(1, n_samples=140 * 468140)),
    "fp32_40mhz_headswitch": (40.0, 64, {"precision": "fp32"},
                              lambda N: N.generate(2, n_samples=130 * N.format_for(40.0).samples_per_field)),
    "k3_28mhz_batch16": (28.0, 16, {"line_locator": "k3"}, lambda N: N.generate(1, n_samples=70 * 468140)),
    "u16_dc_correct_batch16": (28.0, 16, {"input_format": "u16", "dc_correct": True}, lambda N: u16_capture(N, 70)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_graph_replay_equals_launches(torch, N, case):
    mhz, batch, opts, make = CASES[case]
    cap = dev(torch, make(N))
    launches = N.Decoder(mhz, batch, graphs=False, **opts)
    want = launches.decode_capture(cap)
    launches.close()
    graphs = N.Decoder(mhz, batch, graphs=True, **opts)
    first = graphs.decode_capture(cap)       # the graphs cuvhs_create captured
    graphs.reset()
    second = graphs.decode_capture(cap)      # replayed again
    graphs.close()
    assert want["fields"] >= 2 * batch
    same(first, want)
    same(second, want)


def test_graph_shorter_capture_after_longer(torch, N):
    """A 100-field capture (chunks of 64 and 36 fields) and then its first
    70 fields (64 + 6) on one graph-mode context: the second capture's last
    chunk is shorter than what the slot held, so its plan clears the K1
    outputs past it, as the launch-by-launch schedule does."""
    cap = N.generate(1, n_samples=100 * 468140)
    short = cap[: 70 * 468140]
    g = N.Decoder(28.0, 64, graphs=True)
    g.decode_capture(dev(torch, cap))
    g.reset()
    got = g.decode_capture(dev(torch, short))
    g.close()
    ref = N.Decoder(28.0, 64, graphs=False)
    want = ref.decode_capture(dev(torch, short))
    ref.close()
    same(got, want)


def test_graph_capacity_error_then_decode(torch, N):
    """An output buffer one field short stops the graph-mode decode with
    CUVHS_E_CAPACITY before the chunk that does not fit; the context then
    decodes the capture in full, identical to the launch-by-launch decode."""
    cap = dev(torch, N.generate(1, n_samples=80 * 468140))
    ref = N.Decoder(28.0, 64, graphs=False)
    want = ref.decode_capture(cap)
    ref.close()
    g = N.Decoder(28.0, 64, graphs=True)
    with pytest.raises(N.CuvhsError):
        g.decode_capture(cap, max_fields=want["fields"] - 1)
    g.reset()
    same(g.decode_capture(cap), want)
    g.close()


def test_graph_contexts_decode_in_threads(torch, N):
    """Four graph-mode contexts (one per thread, created, and so captured,
    before any thread decodes) decoding captures of different lengths and
    rates at once give the launch-by-launch outputs; each context decodes
    two captures in turn."""
    import threading
    caps = {28.0: N.generate(1, n_samples=150 * 468140),
            40.0: N.generate(2, n_samples=90 * N.format_for(40.0).samples_per_field)}
    jobs = [(28.0, 150), (40.0, 90), (28.0, 70), (40.0, 40)]
    d = {m: dev(torch, a) for m, a in caps.items()}
    spf = {m: N.format_for(m).samples_per_field for m in caps}
    want = {}
    for m, nf in set(jobs):
        ref = N.Decoder(m, 64, precision="mixed-fp16")
        want[(m, nf)] = ref.decode_capture(d[m][: nf * spf[m]])
        ref.close()
    ctxs = [N.Decoder(m, 64, precision="mixed-fp16", graphs=True) for m, _ in jobs]
    got, errors = {}, []

    def work(i):
        try:
            m, nf = jobs[i]
            for n in (nf, jobs[(i + 2) % 4][1]):
                ctxs[i].reset()
                got[(i, n)] = (m, n, ctxs[i].decode_capture(d[m][: n * spf[m]]))
        except Exception as e:  # reported below
            errors.append(repr(e))

    th = [threading.Thread(target=work, args=(i,)) for i in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for c in ctxs:
        c.close()
    assert not errors, errors
    assert len(got) == 8
    for m, n, out in got.values():
        same(out, want[(m, n)])
