"""The C ABI library loads and exports every symbol include/cuvhs.h declares;
argument validation that needs no GPU."""
import ctypes
import subprocess

import pytest

from cuvhs_amd import native


def test_header_symbols_exported():
    L = native.lib()
    syms = native.header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", native.os.path.join(native.LIB_DIR, "libcuvhs.so")],
                         capture_output=True, text=True).stdout
    for s in syms:
        assert f" T {s}\n" in out, s


def test_create_rejects_unsupported_configs():
    h = ctypes.c_void_p()
    for cfg in (native.Config(2, 28.0, 64, 0, None, 0),     # no such system
                native.Config(0, 50.0, 64, 0, None, 0),     # above 40 MHz
                native.Config(0, 28.0, 64, 7, None, 0)):    # unknown precision
        assert native.lib().cuvhs_create(ctypes.byref(cfg), ctypes.byref(h)) == native.E_UNSUPPORTED
    assert native.lib().cuvhs_create(ctypes.byref(native.Config(0, 28.0, 0, 0, None, 0)), ctypes.byref(h)) == native.E_ARG
    assert native.lib().cuvhs_create(None, ctypes.byref(h)) == native.E_ARG
    # block sizes outside the supported ones (K1 4096 / 8192, pre-band-pass
    # 8192 / 16384) are refused before any device work
    for blocks in ((2048, 0, 0), (0, 16384, 0), (0, 0, 4096), (4097, 4096, 16384)):
        cfg = native.Config(0, 28.0, 64, 0, None, 0, 0, 0, 0, *blocks)
        assert native.lib().cuvhs_create(ctypes.byref(cfg), ctypes.byref(h)) == native.E_ARG, blocks


def test_null_context_calls_fail_cleanly():
    L = native.lib()
    info = native.ChunkInfo()
    assert L.cuvhs_decode_chunk(None, None, 0, 0, 1, ctypes.byref(info)) == native.E_ARG
    assert L.cuvhs_destroy(None) == native.E_ARG
    assert L.cuvhs_reset(None) == native.E_ARG
    assert L.cuvhs_last_error(None) == b"no context"


@pytest.mark.parametrize("mhz,spl", [(28.0, 1780), (35.8, 2275), (40.0, 2542)])
def test_format_known_answers(mhz, spl):
    f = native.format_for(mhz)
    assert f.samples_per_line == spl
    assert f.samples_per_field == spl * 263
    assert f.spfp == spl * 283
    assert abs(f.pulse_threshold_hz - 3542857.142857143) < 1e-6


@pytest.mark.parametrize("mhz,spl", [(28.0, 1792), (35.8, 2291), (40.0, 2560)])
def test_pal_format_known_answers(mhz, spl):
    """PAL (video_format.cpp:48-87): 15625 Hz lines, 313 per field, the
    -20 IRE threshold at 4.1 MHz - 20 x 7000 Hz, 1135 x 313 output fields;
    the NTSC entry point keeps NTSC."""
    f = native.format_for(mhz, "pal")
    assert f.samples_per_line == spl
    assert f.samples_per_field == spl * 313
    assert f.spfp == spl * 333
    assert (f.system, f.lines_per_frame, f.field_lines, f.line_len) == (1, 625, 313, 1135)
    assert abs(f.hz_ire - 7000.0) < 1e-9 and abs(f.ire0 - 4.1e6) < 1e-6
    assert abs(f.pulse_threshold_hz - 3.96e6) < 1e-6
    assert f.output_zero == 256.0 and abs(f.output_scale - (54016.0 - 256.0) / (100.0 + 300.0 / 7.0)) < 1e-9
    assert (f.burst_start, f.burst_end) == (99, 139)
    assert f.fsc == 4433618.75 and f.output_rate == 4.0 * 4433618.75
    n = native.format_for(mhz)
    assert (n.system, n.lines_per_frame, n.field_lines, n.line_len) == (0, 525, 263, 910)
    bad = native.Format()
    assert native.lib().cuvhs_format_for_system(2, mhz, ctypes.byref(bad)) == native.E_ARG
