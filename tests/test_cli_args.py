"""`cuvhs` argument handling that exits before any GPU work (main.cpp:45-95)."""
import os
import subprocess

import pytest

from cuvhs_amd import native

EXE = os.path.join(native.LIB_DIR, "cuvhs")


def run(*args):
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=60)


def test_help_exits_zero():
    r = run("--help")
    assert r.returncode == 0 and "Usage" in r.stderr and "--system <NTSC|PAL>" in r.stderr


@pytest.mark.parametrize("args,msg", [
    ((), "input_file and output_base are required"),
    (("--bogus", "a.u8", "o"), "Unknown option: --bogus"),
    (("--system", "SECAM", "a.u8", "o"), "Unknown system"),
    (("--format", "s24", "a.u8", "o"), "Unknown format: s24"),
    (("--batch", "0", "a.u8", "o"), "--batch must be >= 1"),
    (("--precision", "fp16", "a.u8", "o"), "Unknown precision: fp16"),
    (("a.u8", "o", "extra"), "Too many arguments"),
])
def test_bad_arguments_exit_one(args, msg):
    r = run(*args)
    assert r.returncode == 1
    assert msg in r.stderr
