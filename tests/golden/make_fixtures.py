"""Regenerates the golden fixtures of tests/golden/.

* fft_vectors.npz        numpy.fft forward/inverse of seeded inputs (sizes with
                         factors 2, 3, 5, 7, as the reference plans)
* filter_refs.npz        scipy.signal magnitude responses of the reference's
                         Butterworth designs and the Hamming sync FIR at
                         sampled bins, 28 and 40 MHz
* ref_format.json        VideoFormat constants printed by the reference's own
                         video_format.cpp (oracle/_ref/ref_host format), NTSC
                         and PAL ("pal_28" ...)
* ref_convert_u8.txt     RawReader::read_at conversion of all 256 u8 values
* ref_convert_formats.npz  RawReader::read_at conversion of 4096 random u8,
                         s16 and u16 samples, with and without --dc-correct
                         by the reference's own raw_reader.cpp
* ref_writer_3fields.tbc.json  .tbc.json written by the reference's own
                         tbc_writer.cpp for fixed metadata (ref_host json);
                         ref_writer_3fields_pal.tbc.json the same for PAL

The reference-derived files need oracle/_ref/ref_host (make -C oracle ref,
only where the reference sources are mounted).
"""
import json
import os
import subprocess

import numpy as np
from scipy import signal

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_host")

FFT_SIZES = [2, 3, 4, 5, 6, 7, 8, 12, 30, 49, 64, 100, 210, 343, 1024]


def fft_vectors():
    rng = np.random.default_rng(1234)
    d = {}
    for n in FFT_SIZES:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        d[f"x{n}"] = x
        d[f"fwd{n}"] = np.fft.fft(x)
        d[f"inv{n}"] = np.fft.ifft(x) * n
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **d)


def k1_fb(mhz):
    fs = mhz * 1e6
    spl = int(round(fs / 15734.264))
    spfp = spl * 263 + 20 * spl
    n = spfp + 2048
    while True:
        m = n
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        if m == 1:
            break
        n += 1
    return n, n // 2 + 1, spfp


def filter_refs():
    d = {}
    for mhz in (28.0, 40.0):
        fs = mhz * 1e6
        N, fb, spfp = k1_fb(mhz)
        k = np.unique(np.linspace(1, fb - 2, 397).astype(int))
        w = np.pi * k / (fb - 1)
        bp = signal.butter(8, [1.3e6, 7.5e6], btype="bandpass", fs=fs, output="sos")
        lp = signal.butter(8, 8e6, btype="lowpass", fs=fs, output="sos")
        hp = signal.butter(8, 1e6, btype="highpass", fs=fs, output="sos")
        mag = np.ones_like(w)
        for sos in (bp, lp, hp):
            mag = mag * np.abs(signal.sosfreqz(sos, worN=w)[1])
        d[f"rf_bins_{int(mhz)}"] = k
        d[f"rf_mag_{int(mhz)}"] = mag * 2.0
        fir = signal.firwin(65, 5e5, window="hamming", fs=fs)
        H = signal.freqz(fir, worN=w)[1] * np.exp(1j * w * 32)
        d[f"sync_{int(mhz)}"] = H
        # chroma pre-band-pass on its own 7-smooth grid
        n2 = spfp
        while True:
            m = n2
            for p in (2, 3, 5, 7):
                while m % p == 0:
                    m //= p
            if m == 1:
                break
            n2 += 1
        fb2 = n2 // 2 + 1
        k2 = np.unique(np.linspace(1, fb2 - 2, 397).astype(int))
        w2 = np.pi * k2 / (fb2 - 1)
        cb = signal.butter(4, [6e4, 1.2e6], btype="bandpass", fs=fs, output="sos")
        d[f"pre_bins_{int(mhz)}"] = k2
        d[f"pre_mag2_{int(mhz)}"] = np.abs(signal.sosfreqz(cb, worN=w2)[1]) ** 2 / n2
    np.savez_compressed(os.path.join(HERE, "filter_refs.npz"), **d)


def reference_files():
    if not os.path.exists(REF):
        print("oracle/_ref/ref_host missing: reference-derived fixtures not regenerated")
        return
    fmts = {}
    for system in ("ntsc", "pal"):
        for mhz in ("28", "35.8", "40"):
            args = [REF, "format", mhz] + (["pal"] if system == "pal" else [])
            out = subprocess.run(args, capture_output=True, text=True, check=True).stdout
            fmts[("pal_" if system == "pal" else "") + mhz] = {k: float(v) for k, v in (l.split("=") for l in out.split())}
    with open(os.path.join(HERE, "ref_format.json"), "w") as f:
        json.dump(fmts, f, indent=1, sort_keys=True)
    tmp = "/tmp/_all_u8.u8"
    np.arange(256, dtype=np.uint8).tofile(tmp)
    out = subprocess.run([REF, "convert", tmp], capture_output=True, text=True, check=True).stdout
    with open(os.path.join(HERE, "ref_convert_u8.txt"), "w") as f:
        f.write(out)
    for system, name in (("ntsc", "ref_writer_3fields.tbc.json"), ("pal", "ref_writer_3fields_pal.tbc.json")):
        base = "/tmp/_ref_writer"
        args = [REF, "json", base] + (["pal"] if system == "pal" else [])
        subprocess.run(args, capture_output=True, text=True, check=True)
        with open(base + ".tbc.json") as f:
            text = f.read()
        with open(os.path.join(HERE, name), "w") as f:
            f.write(text)


def convert_formats():
    """RawReader::read_at conversion of u8 / s16 / u16 samples, with and
    without --dc-correct, by the reference's compiled raw_reader.cpp."""
    if not os.path.exists(REF):
        print("oracle/_ref/ref_host missing: conversion fixtures not regenerated")
        return
    rng = np.random.default_rng(16)
    data = {"u8": rng.integers(0, 256, 4096).astype(np.uint8),
            "s16": rng.integers(-32768, 32768, 4096).astype(np.int16),
            "u16": rng.integers(0, 65536, 4096).astype(np.uint16)}
    out = {}
    for name, arr in data.items():
        tmp = "/tmp/_conv." + name
        arr.tofile(tmp)
        out[name + "_in"] = arr
        for dc in ("", "dc"):
            args = [REF, "convert", tmp, name] + ([dc] if dc else [])
            txt = subprocess.run(args, capture_output=True, text=True, check=True).stdout
            out[name + ("_dc" if dc else "") + "_out"] = np.array([float(x) for x in txt.split()])
    np.savez(os.path.join(HERE, "ref_convert_formats.npz"), **out)


if __name__ == "__main__":
    fft_vectors()
    filter_refs()
    reference_files()
    convert_formats()
    print("fixtures written to", HERE)
