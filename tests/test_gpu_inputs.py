"""Input formats, --dc-correct and streaming input on the GPU.

s16 / u16 captures are made from the synthetic u8 captures with extra low
bits and a DC offset, so the per-read DC removal and the 16-bit
normalisations (raw_reader.cpp:71-100) change the decoder's input.  Streams
(stdin, FIFO) must decode exactly as the same bytes in a file."""
import os
import subprocess
import threading

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


def to_s16(cap, offset, seed):
    """u8 capture -> s16 with 8 extra noise bits and a DC offset."""
    rng = np.random.default_rng(seed)
    v = (cap.astype(np.int32) - 128) * 256 + rng.integers(-128, 128, cap.size) + offset
    return np.clip(v, -32768, 32767).astype(np.int16)


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


def test_s16_exact_u8_equivalent(torch, N, cap0):
    """s16 samples equal to the u8 normalisation decode bit-identically to
    the u8 capture."""
    s16 = ((cap0.astype(np.int32) - 128) * 256).astype(np.int16)
    d8 = N.Decoder(28.0, 64)
    d16 = N.Decoder(28.0, 64, input_format="s16")
    a = d8.decode_capture(dev(torch, cap0))
    b = d16.decode_capture(dev(torch, s16))
    assert a["fields"] == b["fields"] == 4
    for k in ("luma", "chroma", "file_loc", "is_first", "phase_id", "do_count"):
        assert torch.equal(a[k], b[k]), k
    d8.close()
    d16.close()


def test_s16_dc_correct_vs_oracle(torch, N, oracle, cap0):
    """4 fields of s16 with a +3000 DC offset, decoded with --dc-correct."""
    s16 = to_s16(cap0, 3000, 1)
    ref = oracle.decode(s16, 28.0, 64, trace=True, dc_correct=True)
    dec = N.Decoder(28.0, 64, input_format="s16", dc_correct=True)
    compare(dec.decode_capture(dev(torch, s16)), ref)
    dec.close()


def test_u16_dc_correct_chunks_vs_oracle(torch, N, oracle):
    """40 fields of u16 (offset -2500) in chunks of 16: each chunk's read has
    its own mean."""
    fmt = N.format_for(28.0)
    cap = N.generate(1, n_samples=40 * fmt.samples_per_field)
    u16 = (to_s16(cap, -2500, 2).astype(np.int32) + 32768).astype(np.uint16)
    ref = oracle.decode(u16, 28.0, 16, trace=True, dc_correct=True)
    assert len(ref["chunks"]) >= 3
    dec = N.Decoder(28.0, 16, input_format="u16", dc_correct=True)
    compare(dec.decode_capture(dev(torch, u16)), ref)
    dec.close()


def _cli(N, *args, stdin=None):
    exe = os.path.join(N.LIB_DIR, "cuvhs")
    return subprocess.run([exe, *args], stdin=stdin, capture_output=True, text=True, timeout=300)


def _same_outputs(a, b):
    for s in (".tbc", "_chroma.tbc", ".tbc.json"):
        with open(a + s, "rb") as fa, open(b + s, "rb") as fb:
            assert fa.read() == fb.read(), s


def test_cli_stdin_and_fifo_match_file(N, tmp_path):
    """70 fields, --batch 16 (five chunks): decoding from stdin and from a
    FIFO gives the same three files as decoding the file."""
    fmt = N.format_for(28.0)
    cap = N.generate(1, n_samples=70 * fmt.samples_per_field)
    src = str(tmp_path / "cap.u8")
    cap.tofile(src)
    r = _cli(N, "-f", "28", "--batch", "16", src, str(tmp_path / "file"))
    assert r.returncode == 0, r.stderr
    with open(src, "rb") as f:
        r = _cli(N, "-f", "28", "--batch", "16", "-", str(tmp_path / "stdin"), stdin=f)
    assert r.returncode == 0, r.stderr
    assert "streaming" in r.stderr
    _same_outputs(str(tmp_path / "file"), str(tmp_path / "stdin"))
    fifo = str(tmp_path / "cap.fifo")
    os.mkfifo(fifo)

    def feed():
        with open(fifo, "wb") as w:
            for i in range(0, cap.size, 1 << 20):   # in pieces, as a live capture arrives
                w.write(cap[i:i + (1 << 20)].tobytes())

    t = threading.Thread(target=feed, daemon=True)
    t.start()
    r = _cli(N, "-f", "28", "--batch", "16", "--format", "u8", fifo, str(tmp_path / "fifo"))
    t.join(timeout=60)
    assert r.returncode == 0, r.stderr
    _same_outputs(str(tmp_path / "file"), str(tmp_path / "fifo"))


def test_cli_s16_dc_correct_matches_oracle(N, oracle, cap0, tmp_path):
    """The CLI on an .s16 file with --dc-correct: JSON identical to the
    oracle's, TBC samples within 1 LSB."""
    s16 = to_s16(cap0, 3000, 1)
    src = str(tmp_path / "cap.s16")
    s16.tofile(src)
    r = _cli(N, "-f", "28", "--dc-correct", src, str(tmp_path / "g"))
    assert r.returncode == 0, r.stderr
    assert oracle.lib().or_decode_file_fmt(src.encode(), str(tmp_path / "o").encode(), 28.0, 64, 1, 1) == 4
    with open(tmp_path / "g.tbc.json") as a, open(tmp_path / "o.tbc.json") as b:
        assert a.read() == b.read()
    for s in (".tbc", "_chroma.tbc"):
        a = np.fromfile(str(tmp_path / ("g" + s)), np.uint16).astype(np.int64)
        b = np.fromfile(str(tmp_path / ("o" + s)), np.uint16).astype(np.int64)
        assert a.size == b.size and np.abs(a - b).max() <= 1
