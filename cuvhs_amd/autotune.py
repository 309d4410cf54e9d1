"""Autotuner of the overlap-save block sizes (runs on the GPU).

    python -m cuvhs_amd.autotune [--fields 192] [--repeats 3]
                                 [--header cuvhs_amd/csrc/tuning_gfx950.h]
                                 [--json profiles/autotune_gfx950.json]

The decoder's convolution stages run as overlap-save blocks whose size is a
free choice: K1 part A and part B (4096 or 8192 points) and the chroma
pre-band-pass (8192 or 16384).  A larger block wastes less of each transform
on the filter's support but holds more LDS and registers, so fewer
workgroups share a CU; which wins depends on the filter (sample rate,
precision form) and on the device.  For each sample rate of the benchmark
(28, 35.8, 40 MHz) and each precision form, this script decodes a synthetic
capture of `--fields` fields (the benchmark's signal model at that rate)
with the stages serialised, once per block size each stage's filter fits
(the other stages fixed), takes the fastest of `--repeats` decodes per
size, and keeps the fastest size per stage.  It writes the table that
cuvhs_create reads by default (a C header, rebuilt into the library) and the
measurements as JSON.
"""
import argparse
import datetime
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

RATES = (28.0, 35.8, 40.0)
FORMS = ("fp64", "fp32", "mixed-fp16", "mixed-bf16")
# stage -> (profile stage name, candidate sizes, position in Decoder(blocks=...))
STAGES = {"k1a": ("k1_analytic", (4096, 8192), 0),
          "k1b": ("k1_video", (4096, 8192), 1),
          "pre": ("k6_prebpf", (8192, 16384), 2)}
BASE = (4096, 4096, 16384)   # sizes every filter fits, for the stages not being timed
PRECISION_ENUM = {"fp64": 0, "fp32": 1, "mixed-fp16": 2, "mixed-bf16": 3}


def capture(mhz, fields):
    """The benchmark's signal model at mhz: the 600-field workload's
    generator settings (config 1), or the 40 MHz head-switch workload's
    (config 2) at 40 MHz."""
    from cuvhs_amd import native as N
    spf = N.format_for(mhz).samples_per_field
    if mhz == 40.0:
        return N.generate(2, n_samples=fields * spf)
    return N.generate(1, n_samples=fields * spf, sample_rate_mhz=mhz)


def time_stage(torch, d_cap, mhz, form, blocks, stage, repeats):
    """Fastest serialised time (ms per decode) of one stage with the given
    block sizes, and the block shapes; None if a size does not fit."""
    from cuvhs_amd import native as N
    try:
        dec = N.Decoder(mhz, 64, precision=form, blocks=blocks)
    except N.CuvhsError:
        return None
    try:
        ka, va, kb, vb = dec.k1_blocks()
        kp, vp, _ = dec.prebpf_block()
        dec.reset()
        dec.decode_capture(d_cap)      # kernel loading, first-touch
        best = None
        for _ in range(repeats):
            dec.profile(True, serial=True)
            dec.reset()
            dec.decode_capture(d_cap)
            t = dec.profile_read().get(STAGES[stage][0], 0.0)
            dec.profile(False)
            best = t if best is None else min(best, t)
        torch.cuda.synchronize()
        return {"ms": round(best, 4), "k1a": [ka, va], "k1b": [kb, vb], "pre": [kp, vp]}
    finally:
        dec.close()


def tune(fields, repeats):
    import torch
    res = []
    for mhz in RATES:
        cap = capture(mhz, fields)
        d_cap = torch.from_numpy(cap).to("cuda:0")
        for form in FORMS:
            row = {"mhz": mhz, "form": form, "fields": fields, "stages": {}}
            for stage, (_, sizes, pos) in STAGES.items():
                if stage == "k1a" and form.startswith("mixed"):
                    continue   # part A runs on the matrix cores in the mixed forms
                trials = {}
                for b in sizes:
                    blocks = list(BASE)
                    blocks[pos] = b
                    r = time_stage(torch, d_cap, mhz, form, tuple(blocks), stage, repeats)
                    trials[str(b)] = r
                    print(f"{mhz:g} MHz {form} {stage} {b}: {r['ms'] if r else 'does not fit'}", flush=True)
                fit = {int(k): v["ms"] for k, v in trials.items() if v}
                row["stages"][stage] = {"trials": trials, "best": min(fit, key=fit.get) if fit else 0}
            res.append(row)
        del d_cap
    return res


def header(res, device):
    lines = []
    for row in res:
        st = row["stages"]
        k1a = st["k1a"]["best"] if "k1a" in st else 0
        lines.append(f"    {{{row['mhz']:.1f}, {PRECISION_ENUM[row['form']]}, {k1a}, {st['k1b']['best']}, "
                     f"{st['pre']['best']}}},   // {row['form']}")
    stamp = datetime.date.today().isoformat()
    return f"""// Block sizes of the overlap-save convolutions measured fastest on gfx950
// (MI355X) by cuvhs_amd/autotune.py, which writes this file: for each sample
// rate and precision form, every block size each stage's filter fits was
// timed in a decode with the stages serialised, and the fastest kept.
// cuvhs_create takes a row's sizes when its rate is within 0.05 MHz of the
// configuration's and the precision matches; other configurations use the
// cost model of design.cpp.  k1a is unused by the mixed forms.
// Measured {stamp} on {device}; timings: profiles/autotune_gfx950.json.
#pragma once

namespace cuvhs {{

struct TunedBlocks {{
    double mhz;
    int precision;   // CUVHS_PRECISION_*
    int k1a, k1b, pre;
}};

constexpr TunedBlocks kTunedBlocks[] = {{
{chr(10).join(lines)}
}};

}}  // namespace cuvhs
"""


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fields", type=int, default=192)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--header", default=os.path.join(HERE, "csrc", "tuning_gfx950.h"))
    ap.add_argument("--json", default=os.path.join(ROOT, "profiles", "autotune_gfx950.json"))
    args = ap.parse_args()
    import torch
    device = torch.cuda.get_device_name(0)
    res = tune(args.fields, args.repeats)
    with open(args.header, "w") as f:
        f.write(header(res, device))
    with open(args.json, "w") as f:
        json.dump({"device": device, "fields": args.fields, "repeats": args.repeats, "base_blocks": BASE,
                   "unit": "ms per decode, stage serialised", "results": res}, f, indent=1)
    print(f"wrote {args.header} and {args.json}", flush=True)


if __name__ == "__main__":
    main()
