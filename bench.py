"""Decode throughput of the MI355X VHS decoder on the 600-field 28 MHz workload.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

One step decodes the whole synthetic capture (600 NTSC fields of u8 RF at
28 MHz: SNR 25-40 dB per field, Poisson dropouts, wow/flutter; chunks of
F = 64 fields) with the capture already resident in HBM and outputs left in
HBM.  Prints one JSON line on stdout.

The headline value is the FP64 form (the reference's arithmetic, checked
against the oracle).  The line also carries the FP32 form (K1 transforms in
FP32) timed the same way, with its error and decision report against the
FP64 form (cuvhs_amd.precision); --no-forms skips it.
"""
import argparse
import json
import math
import os
import sys
import time

# Hardware queues for this process (read when HIP initialises, before torch
# touches the GPU).  The decoder runs four streams besides torch's; with
# HIP's default of four queues two of them share one, and a context then
# keeps its forked tail work on the head stream (cuvhs_create reads the
# same variable).  Eight gives every stream its own queue.
os.environ["GPU_MAX_HW_QUEUES"] = "8"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "NTSC fields/s at 28 MHz (MFMA mixed vs plain FP32, 1 GPU) + achieved HBM GB/s"
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64, vector and matrix

# Algorithmic bytes per unit of each stage (the stage's compulsory reads and
# writes of its own inputs/outputs, counted once).
STAGE_BYTES = {
    # per capture sample: u8 in, analytic envelope (f64) + instantaneous frequency (f64) out
    "k1_analytic": ("sample", 1 + 8 + 8),
    # per capture sample: instantaneous frequency in, demod + demod_05 out (f64)
    "k1_video": ("sample", 8 + 8 + 8),
    # per capture sample: u8 in, band-passed f32 out
    "k6_prebpf": ("sample", 1 + 4),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# the workloads a bench line can be taken on
WORKLOADS = {
    "600f_28mhz": {"config": 1, "mhz": 28.0, "fields": None,
                   "desc": "600 synthetic NTSC fields of u8 RF at 28 MHz (SNR 25-40 dB, dropouts, wow/flutter)"},
    "640f_40mhz_headswitch": {"config": 2, "mhz": 40.0, "fields": 640,
                              "desc": "first 640 fields of the 6000-field 40 MHz u8 workload (same signal model, "
                                      "0-4 us head-switch offsets in groups of 1-8 fields)"},
}


def cpu_baseline(cap, fmt, batch, mhz):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    sample = cap[: 2 * batch * fmt.spfp]  # two chunks
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t = time.perf_counter()
    ref = oracle_py.decode(sample, mhz, batch, trace=False)
    dt = time.perf_counter() - t
    return {"value": ref["fields"] / dt, "unit": "fields/s", "cores": threads, "kind": "port",
            "sample": f"first {ref['fields']} fields of the same {mhz:g} MHz capture ({sample.size} u8 samples), "
                      f"oracle liboracle.so (FP64, own FFT), OpenMP over fields, {dt:.1f} s"}


def serving_main(args):
    """Serving form: 16 requests at 4 req/s (Poisson), sample rates uniform in
    {28, 35.8, 40} MHz, lengths log-uniform in 30..600 fields, sliced from one
    600-field synthetic capture per rate resident in HBM; 4 workers with one
    context per rate each, 64 GB output budget."""
    import torch
    from cuvhs_amd import serving
    t = time.perf_counter()
    pool_fields = 600
    pool = serving.make_pool(pool_fields)
    log(f"generated the capture pool in {time.perf_counter() - t:.1f} s")
    d_pool = {m: torch.from_numpy(a).to("cuda:0") for m, a in pool.items()}
    torch.cuda.synchronize()
    reqs = serving.make_requests(16, 4.0, 30, pool_fields, pool_fields)
    res = {}
    burst = [dict(r, arrival=0.0) for r in reqs]
    for form in ("fp64", "fp32"):
        # one untimed pass warms the contexts' kernels, then the timed runs:
        # Poisson arrivals, and all 16 at once (the saturated aggregate rate)
        serving.run(d_pool, reqs[:2], precision=form)
        res[form] = serving.summarize(serving.run(d_pool, reqs, precision=form))
        res[form]["all_at_once"] = serving.summarize(serving.run(d_pool, burst, precision=form))
        log(f"{form}: {res[form]}")
    line = {
        "metric": "serving: per-field latency p50/p99 and aggregate NTSC fields/s, 16 concurrent captures",
        "value": res["fp64"]["aggregate_fields_per_s"],
        "unit": "fields/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "data": "synthetic",
        "config": {"workload": "16 requests, Poisson 4 req/s, fs in {28, 35.8, 40} MHz, 30..600 fields log-uniform",
                   "workers": 4, "batch_fields": 64, "hbm_budget_gb": 64},
        "forms": res,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the FP32 form and its error report")
    ap.add_argument("--workload", default="600f_28mhz", choices=sorted(WORKLOADS),
                    help="the capture a bench line is taken on (default: the 600-field 28 MHz workload)")
    ap.add_argument("--serving", action="store_true",
                    help="instead: the serving loop (16 concurrent captures at 28/35.8/40 MHz, Poisson arrivals), "
                         "one JSON line with latency percentiles and aggregate fields/s per precision form")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 4 plain device copies of MIB MiB (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the decode path shards by "
              "capture segment, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from cuvhs_amd import native

    if args.serving:
        serving_main(args)
        return

    if args.copy_probe > 0:
        n = args.copy_probe << 20
        src = torch.randint(0, 255, (n,), dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 4 copies of {n} bytes (read {n} B + write {n} B each)")
        return

    t = time.perf_counter()
    wl = WORKLOADS[args.workload]
    mhz = wl["mhz"]
    if wl["fields"]:
        cap = native.generate(wl["config"], n_samples=wl["fields"] * native.format_for(mhz).samples_per_field)
    else:
        cap = native.generate(wl["config"])
    log(f"generated {cap.size} samples in {time.perf_counter() - t:.1f} s")
    dev = torch.device("cuda:0")
    d_cap = torch.from_numpy(cap).to(dev)
    dec = native.Decoder(mhz, args.batch, device=0)
    fmt = dec.fmt
    fields = 0
    for _ in range(args.warmup):
        dec.reset()
        fields = dec.decode_capture(d_cap)["fields"]
    log(f"warmup done: {fields} fields")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dec.reset()
        out = dec.decode_capture(d_cap)
        fields = out["fields"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # one more decode, untimed, with stage events on the same overlapped
    # schedule (the events cost host time, so the timed region has none)
    dec.profile(True)
    dec.reset()
    dec.decode_capture(d_cap)
    prof = dec.profile_read()
    k1_samples, chunks, _ = dec.profile_work()
    # and one with the stages serialised: each stage's own time
    dec.profile(True, serial=True)
    dec.reset()
    dec.decode_capture(d_cap)
    iso = dec.profile_read()
    dec.profile(False)
    ms_per_step = dt * 1000.0 / args.steps
    value = fields * args.steps / dt
    bytes_per_field = cap.size / fields + 2 * 2 * 910 * 263  # u8 in + luma/chroma u16 out
    log(f"{fields} fields/step, {ms_per_step:.2f} ms/step, {value:.1f} fields/s")
    stage_ms = dict(prof)
    log("per-step stage ms (overlapped): " + ", ".join(f"{k}={v:.3f}" for k, v in stage_ms.items()))
    log("per-step stage ms (serialised): " + ", ".join(f"{k}={v:.3f}" for k, v in iso.items()))

    # dominant kernel (largest serialised time) among those with a byte count
    dom = max((k for k in STAGE_BYTES if k in iso), key=lambda k: iso[k])
    unit, per_unit = STAGE_BYTES[dom]
    launches = chunks                      # one launch per chunk of the profiled decode
    alg_bytes_per_launch = per_unit * k1_samples / launches
    avg_launch_s = stage_ms[dom] / 1000.0 / launches
    achieved = alg_bytes_per_launch / avg_launch_s / 1e9
    iso_launch_s = iso[dom] / 1000.0 / launches
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(pmc_path) and args.workload == "600f_28mhz":  # counters taken on that workload
        with open(pmc_path) as f:
            traffic = json.load(f).get("bytes_per_launch", {}).get(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "alg_bytes_per_launch": int(alg_bytes_per_launch), "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                "achieved_isolated": round(alg_bytes_per_launch / iso_launch_s / 1e9, 1),
                "avg_launch_ms_isolated": round(iso_launch_s * 1e3, 4)}

    # the same kernel against the FP64 arithmetic roof: its convolution blocks
    # as algorithmic flops, two B-point complex FFTs (5 B log2 B each) and the
    # B-point spectrum product (6 B) per block of V outputs, over the blocks
    # of one launch (the epilogue's arctangent, square root and difference are
    # not counted); peak 78.6 TFLOP/s FP64
    roofline_fp64 = None
    if dom in ("k1_analytic", "k1_video"):
        ba, va, bb, vb = dec.k1_blocks()
        B, V = (ba, va) if dom == "k1_analytic" else (bb, vb)
        blocks = (k1_samples / fmt.spfp) * -(-fmt.spfp // V) / launches
        flops = blocks * (2 * 5 * B * math.log2(B) + 6 * B)
        roofline_fp64 = {"bound": "fp64", "kernel": dom, "block": B, "valid_per_block": V,
                         "alg_flops_per_launch": int(flops), "unit": "TFLOP/s", "peak": PEAK_FP64_TFLOPS,
                         "achieved": round(flops / avg_launch_s / 1e12, 2),
                         "frac": round(flops / avg_launch_s / 1e12 / PEAK_FP64_TFLOPS, 4),
                         "achieved_isolated": round(flops / iso_launch_s / 1e12, 2),
                         "frac_isolated": round(flops / iso_launch_s / 1e12 / PEAK_FP64_TFLOPS, 4)}

    forms = None
    if not args.no_forms:
        from cuvhs_amd import precision
        d32 = native.Decoder(mhz, args.batch, device=0, precision="fp32")
        for _ in range(args.warmup):
            d32.reset()
            d32.decode_capture(d_cap)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            d32.reset()
            f32_fields = d32.decode_capture(d_cap)["fields"]
        torch.cuda.synchronize()
        dt32 = time.perf_counter() - t0
        d32.profile(True, serial=True)
        d32.reset()
        d32.decode_capture(d_cap)
        iso32 = d32.profile_read()
        d32.close()
        log(f"fp32 form: {f32_fields * args.steps / dt32:.1f} fields/s; comparing with fp64 ...")
        report = precision.compare(d_cap, mhz, args.batch, "fp32")
        forms = {"fp32": {"value": round(f32_fields * args.steps / dt32, 2), "unit": "fields/s",
                          "ms_per_step": round(dt32 * 1000.0 / args.steps, 3), "dtype": "f32 (K1), f64 (rest)",
                          "stage_ms_per_step_isolated": {k: round(v, 3) for k, v in iso32.items()},
                          "vs_fp64": report}}

    cpu = None
    if not args.no_cpu_baseline:
        log("running CPU baseline ...")
        cpu = cpu_baseline(cap, fmt, args.batch, mhz)
        log(f"cpu baseline {cpu['value']:.2f} fields/s on {cpu['cores']} threads")

    line = {
        "metric": METRIC if mhz == 28.0 else METRIC.replace("28 MHz", f"{mhz:g} MHz"),
        "value": round(value, 2),
        "unit": "fields/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["desc"],
                   "fields_per_step": fields, "batch_fields": args.batch, "sample_rate_mhz": mhz,
                   "precision": "fp64 (chroma pre-band-pass fp32)"},
        "hbm_gbs_algorithmic": round(value * bytes_per_field / 1e9, 2),
        "stage_ms_per_step": {k: round(v, 3) for k, v in stage_ms.items()},
        "stage_ms_per_step_isolated": {k: round(v, 3) for k, v in iso.items()},
        "roofline": roofline,
        "roofline_fp64": roofline_fp64,
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
