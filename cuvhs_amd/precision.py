"""Error and decision report of a reduced-precision form against the FP64
form on the same capture.

The FP64 form reproduces the reference's arithmetic (cuFFT FP64 in
fm_demod.cu:396-486) and is the one checked against the oracle.  A reduced
form changes only K1 (and the chroma pre-band-pass, FP32 in every reduced
form): CUVHS_PRECISION_FP32 runs the K1 transforms, phase and envelope in
FP32; the mixed forms (CUVHS_PRECISION_MIXED_F16 / _BF16) run K1 part A on
the matrix cores and part B in FP32, with an FP64 guard band that recomputes
every block whose sync-threshold decisions the reduced arithmetic could
change.  This module measures what that does downstream:

* K1 outputs chunk by chunk (demod, demod_05, envelope): max and RMS
  absolute error over the samples the capture covers, up to TAIL_GUARD
  samples before its end (past the end the chunk reads zeros, and
  demodulating the step into them is ill-conditioned in any precision,
  FP64 included), and the number of those samples whose sync-threshold decision
  on demod_05 flips (x <= pulse_threshold_hz, the comparison behind VSYNC
  discovery and the K2 pulse runs, vsync_discover.cu:17-20,
  sync_pulses.cu:33-41);
* decoded outputs: field count, fileLoc / isFirstField / fieldPhaseID and
  dropout-entry mismatches, and max / RMS error of the u16 luma and chroma.
"""
import ctypes

import numpy as np

from . import native as N

TAIL_GUARD = 16384


def _chunk_k1(dec, d_capture, raw_offset, fields_k1):
    """K1 outputs of one chunk (cuvhs_fm_demod) and the number of leading
    samples that lie in the capture (the rest is the zero-filled tail)."""
    import torch
    n = d_capture.numel() - raw_offset
    spfp = dec.fmt.spfp
    nv = min(n, dec.batch * spfp)
    shape = (fields_k1 * spfp,)
    out = [torch.empty(shape, dtype=torch.float64, device=d_capture.device) for _ in range(3)]
    # the library writes them on the decoder's non-blocking stream: torch's
    # work on their memory (earlier tensors it held) must be done first
    torch.cuda.current_stream(d_capture.device).synchronize()
    base = d_capture.data_ptr() + raw_offset * d_capture.element_size()
    N.check(N.lib().cuvhs_fm_demod(dec.h, ctypes.c_void_p(base), nv, fields_k1, *[N.ptr(t) for t in out]), dec.h,
            "fm_demod")
    return out, min(nv, fields_k1 * spfp)


def _err(a, b):
    import torch
    d = (a.to(torch.float64) - b.to(torch.float64)).abs()
    if d.numel() == 0:
        return {"max": 0.0, "rms": 0.0}
    return {"max": float(d.max()), "rms": float(d.square().mean().sqrt())}


def compare(d_capture, sample_rate_mhz=28.0, batch_fields=64, form="fp32", ref=None, system="ntsc"):
    """Decode d_capture (device u8 tensor) in `form` and in FP64; return the
    report dict.  ref: an FP64 decode_chunked result to reuse."""
    import torch
    dec64 = N.Decoder(sample_rate_mhz, batch_fields, precision="fp64", system=system)
    dec = N.Decoder(sample_rate_mhz, batch_fields, precision=form, system=system)
    try:
        r = ref if ref is not None else dec64.decode_chunked(d_capture)
        o = dec.decode_chunked(d_capture)
        pthr = float(dec64.fmt.pulse_threshold_hz)
        # K1 on the FP64 form's chunk schedule
        k1 = {"demod": [], "demod_05": [], "env": []}
        flips = 0
        samples = 0
        for c in r["chunks"]:
            if not c["fields"]:
                continue
            a, nv = _chunk_k1(dec64, d_capture, c["raw_offset"], c["fields_k1"])
            b, _ = _chunk_k1(dec, d_capture, c["raw_offset"], c["fields_k1"])
            if nv < a[0].numel():
                nv = max(nv - TAIL_GUARD, 0)
                a = [t[:nv] for t in a]
                b = [t[:nv] for t in b]
            for key, x, y in zip(k1, a, b):
                k1[key].append(_err(y, x))
            flips += int(((a[1] <= pthr) != (b[1] <= pthr)).sum())
            samples += a[1].numel()
            del a, b
        k1_rep = {k: {"max": max(e["max"] for e in v),
                      "rms": float(np.sqrt(np.mean([e["rms"] ** 2 for e in v])))} for k, v in k1.items() if v}
        nf = min(r["fields"], o["fields"])
        mism = {"fields": abs(r["fields"] - o["fields"]), "sync_threshold_flips": flips}
        for key in ("file_loc", "is_first", "phase_id"):
            mism[key] = int((r[key][:nf].to(torch.int64) != o[key][:nf].to(torch.int64)).sum())
        do_bad = 0
        rc, oc = r["do_count"][:nf].cpu().numpy(), o["do_count"][:nf].cpu().numpy()
        lists = {k: (r[k][:nf].cpu().numpy(), o[k][:nf].cpu().numpy()) for k in ("do_line", "do_startx", "do_endx")}
        for i in range(nf):
            if rc[i] != oc[i] or any(not np.array_equal(x[i, :rc[i]], y[i, :rc[i]]) for x, y in lists.values()):
                do_bad += 1
        mism["dropout_fields"] = do_bad
        out_err = {k: _err(o[k][:nf].to(torch.int32), r[k][:nf].to(torch.int32)) for k in ("luma", "chroma")}
        for k in ("luma", "chroma"):
            d = (o[k][:nf].to(torch.int32) - r[k][:nf].to(torch.int32)).abs()
            out_err[k]["samples_over_1lsb"] = int((d > 1).sum())
            # where they are: (field, line) of the first few, and their fields
            at = torch.nonzero(d > 1)[:, :2].cpu().numpy()
            out_err[k]["over_1lsb_fields"] = sorted({int(f) for f in at[:, 0]})[:32]
            out_err[k]["over_1lsb_first"] = [[int(f), int(l)] for f, l in at[:8]]
        return {"form": form, "fields": o["fields"], "k1_samples": samples, "k1_error": k1_rep,
                "output_error_lsb": out_err, "mismatches": mism}
    finally:
        dec.close()
        dec64.close()
