"""Decode throughput of the MI355X VHS decoder on the 600-field 28 MHz workload.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

One step decodes the whole synthetic capture (600 NTSC fields of u8 RF at
28 MHz: SNR 25-40 dB per field, Poisson dropouts, wow/flutter; chunks of
F = 64 fields) with the capture already resident in HBM and outputs left in
HBM.  Prints one JSON line on stdout.

The headline value is the MFMA mixed form (--form, default mixed-fp16: K1
part A on the matrix cores with f16 hi/lo taps, part B in FP32, an FP64
guard band that keeps every sync-threshold decision the FP64 form's).  The
line also carries the other forms timed the same way -- fp64 (the
reference's arithmetic, the one checked against the oracle), fp32 (K1
transforms in FP32) and mixed-bf16 -- each reduced form with its error and
decision report against the FP64 form (cuvhs_amd.precision), the mixed
form's MFMA utilisation and its speed over the plain FP32 form;
--no-forms skips them.  --workload picks another capture for a secondary
line (the 40 MHz head-switch workload, or the same 600-field signal model
on PAL's line structure and levels).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "NTSC fields/s at 28 MHz (MFMA mixed vs plain FP32, 1 GPU) + achieved HBM GB/s"
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6   # MI355X FP64, vector and matrix
PEAK_F16_TFLOPS = 2500.0  # MI355X dense f16 / bf16 matrix
FORMS = ("fp64", "fp32", "mixed-fp16", "mixed-bf16")
DTYPES = {"fp64": "f64", "fp32": "f32 (K1), f64 (rest)",
          "mixed-fp16": "f16 hi/lo MFMA (K1 part A), f32 (K1 part B), f64 (guard band, rest)",
          "mixed-bf16": "bf16 hi/lo MFMA (K1 part A), f32 (K1 part B), f64 (guard band, rest)"}

LINE_RATE = 15734.264      # NTSC line rate (video_format.cpp)
PAL_LINE_RATE = 15625.0
TBC_BYTES = 2 * 2 * 910 * 263  # luma + chroma u16 per field
PAL_TBC_BYTES = 2 * 2 * 1135 * 313

# The stages a roofline line can be taken on (their stage names in the
# decoder's profile); the dominant one is the largest serialised time.
ROOFLINE_STAGES = ("k1_analytic", "k1_video", "k6_prebpf", "k6_chroma")


# Compulsory HBM bytes of one field at the benchmark's sample rates: its
# unique u8 input (one field period of samples) plus the two u16 TBC planes,
# as the benchmark's measurement plan tabulates them.
ALG_BYTES_PER_FIELD = {28.0: 1424455, 35.8: 1553968, 40.0: 1624655}


def alg_bytes_per_field(mhz, system="ntsc"):
    """Compulsory bytes of one field (the table above; for other rates one
    field period of u8 samples, fs * 262.5 / line rate, plus the planes;
    PAL: fs * 312.5 / 15625 plus its 1135 x 313 planes)."""
    if system == "pal":
        return round(mhz * 1e6 * 312.5 / PAL_LINE_RATE) + PAL_TBC_BYTES
    return ALG_BYTES_PER_FIELD.get(round(mhz, 1), round(mhz * 1e6 * 262.5 / LINE_RATE) + TBC_BYTES)


def fft_flops(n):
    """5 N log2 N for a complex transform of N points."""
    return 5.0 * n * math.log2(n)


def k1_fft_flops_per_segment(stage, n):
    """The reference's K1 transforms of one segment in the FFT formulation
    (a real transform counted as half a complex one): part A one D2Z and one
    Z2Z, part B one D2Z and two Z2D (fm_demod.cu:396-486)."""
    return (0.5 + 1.0) * fft_flops(n) if stage == "k1_analytic" else (0.5 + 0.5 + 0.5) * fft_flops(n)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# the workloads a bench line can be taken on
WORKLOADS = {
    "600f_28mhz": {"config": 1, "mhz": 28.0, "fields": None,
                   "desc": "600 synthetic NTSC fields of u8 RF at 28 MHz (SNR 25-40 dB, dropouts, wow/flutter)"},
    "640f_40mhz_headswitch": {"config": 2, "mhz": 40.0, "fields": 640,
                              "desc": "first 640 fields of the 6000-field 40 MHz u8 workload (same signal model, "
                                      "0-4 us head-switch offsets in groups of 1-8 fields)"},
    "6000f_40mhz": {"config": 2, "mhz": 40.0, "fields": 6000,
                    "desc": "6000 synthetic NTSC fields of u8 RF at 40 MHz (4.0 GB; SNR 25-40 dB, dropouts, "
                            "wow/flutter, 0-4 us head-switch offsets in groups of 1-8 fields)"},
    "600f_28mhz_pal": {"config": 1, "mhz": 28.0, "fields": 600, "system": "pal",
                       "desc": "600 synthetic PAL fields of u8 RF at 28 MHz (the 600-field signal model on PAL's "
                               "312.5-line fields and levels; 1135 x 313 output fields)"},
}


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    return oracle_py


def cpu_decode(cap, mhz, batch, threads, system="ntsc"):
    """One oracle decode of cap with the given OpenMP thread count: fields/s
    and per-stage ms per field."""
    O = _oracle()
    used = O.set_threads(threads)
    t = time.perf_counter()
    ref = O.decode(cap, mhz, batch, trace=False, system=system)
    dt = time.perf_counter() - t
    n = max(ref["fields"], 1)
    return {"fields": ref["fields"], "seconds": round(dt, 2), "fields_per_s": round(ref["fields"] / dt, 3),
            "threads": used, "stage_ms_per_field": {k: round(v * 1e3 / n, 2) for k, v in ref["stage_seconds"].items()}}


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_baseline(cap, fmt, batch, mhz, system="ntsc"):
    """The oracle (the reference's stage logic restated in C++, FP64) on the
    host cores: the first two chunks of the workload with every host thread,
    and the 4-field capture with one thread, each with per-stage ms/field."""
    from cuvhs_amd import native
    threads = host_threads()
    sample = cap[: 2 * batch * fmt.spfp]
    multi = cpu_decode(sample, mhz, batch, threads, system)
    one = cpu_decode(native.generate(0), 28.0, batch, 1)
    _oracle().set_threads(threads)
    return {"value": multi["fields_per_s"], "unit": "fields/s", "cores": multi["threads"], "kind": "port",
            "sample": f"first {multi['fields']} fields of the same {mhz:g} MHz {system.upper()} capture "
                      f"({sample.size} u8 samples), "
                      f"oracle liboracle.so (FP64, own FFT), OpenMP over fields, {multi['seconds']} s",
            "stage_ms_per_field": multi["stage_ms_per_field"],
            "one_thread": {"workload": "4-field 28 MHz capture", **one}}


def cpu_baseline_full(args):
    """The full CPU baseline set: the 4-field and the 600-field 28 MHz
    captures in full and the first 640 fields of the 40 MHz head-switch
    capture, each with every host thread, and the 4-field capture with one
    thread; one JSON line."""
    from cuvhs_amd import native
    threads = host_threads()
    runs = {}
    c0 = native.generate(0)
    runs["4f_28mhz"] = cpu_decode(c0, 28.0, args.batch, threads)
    runs["4f_28mhz_1thread"] = cpu_decode(c0, 28.0, args.batch, 1)
    log(f"4 fields: {runs['4f_28mhz']}")
    c1 = native.generate(1)
    runs["600f_28mhz"] = cpu_decode(c1, 28.0, args.batch, threads)
    log(f"600 fields: {runs['600f_28mhz']}")
    del c1
    c2 = native.generate(2, n_samples=640 * native.format_for(40.0).samples_per_field)
    runs["640f_40mhz_headswitch"] = cpu_decode(c2, 40.0, args.batch, threads)
    log(f"640 fields at 40 MHz: {runs['640f_40mhz_headswitch']}")
    line = {"metric": "CPU baseline: oracle (reference stage logic, FP64) NTSC fields/s on the host cores",
            "unit": "fields/s", "threads": threads, "kind": "port", "batch_fields": args.batch, "runs": runs}
    print(json.dumps(line), flush=True)


def cli_end_to_end(cap, mhz, batch, form, system="ntsc"):
    """The `cuvhs` CLI on the capture written to a file: disk read, host to
    device copies, decode, device to host copies and .tbc/.tbc.json writes."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "cuvhs_amd", "lib", "cuvhs")
    with tempfile.TemporaryDirectory() as tmp:
        inp = os.path.join(tmp, "in.u8")
        cap.tofile(inp)
        out = os.path.join(tmp, "out")
        runs = []
        for _ in range(2):   # the first run pays for kernel loading
            t = time.perf_counter()
            r = subprocess.run([exe, "--system", system.upper(), "-f", f"{mhz:g}", "--batch", str(batch),
                                "--precision", form, "--overwrite", inp, out],
                               stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
            runs.append(time.perf_counter() - t)
            if r.returncode != 0:
                return {"error": r.stderr[-300:]}
        nf = json.load(open(out + ".tbc.json"))["videoParameters"]["numberOfSequentialFields"]
    return {"fields": nf, "precision": form, "seconds": round(runs[-1], 3), "fields_per_s": round(nf / runs[-1], 1),
            "includes": "process start, file read, PCIe copies, decode, file writes"}


def serving_main(args):
    """Serving form: 16 requests at 4 req/s (Poisson), sample rates uniform in
    {28, 35.8, 40} MHz, lengths log-uniform in 30..3000 fields, sliced from
    one 3000-field synthetic capture per rate resident in HBM (5.2 GB); 4
    workers with one context per rate each, 64 GB output budget."""
    import torch
    from cuvhs_amd import serving
    t = time.perf_counter()
    pool_fields = 3000
    pool = serving.make_pool(pool_fields)
    log(f"generated the capture pool in {time.perf_counter() - t:.1f} s")
    d_pool = {m: torch.from_numpy(a).to("cuda:0") for m, a in pool.items()}
    torch.cuda.synchronize()
    reqs = serving.make_requests(16, 4.0, 30, pool_fields, pool_fields)
    res = {}
    burst = [dict(r, arrival=0.0) for r in reqs]
    g = bool(args.graphs)
    for form in ("mixed-fp16", "fp32", "fp64"):
        # one untimed pass warms the contexts' kernels, then the timed runs:
        # Poisson arrivals, and all 16 at once (the saturated aggregate rate)
        serving.run(d_pool, reqs[:2], precision=form, graphs=g)
        res[form] = serving.summarize(serving.run(d_pool, reqs, precision=form, graphs=g))
        res[form]["all_at_once"] = serving.summarize(serving.run(d_pool, burst, precision=form, graphs=g))
        log(f"{form}: {res[form]}")
    line = {
        "metric": "serving: per-field latency p50/p99 and aggregate NTSC fields/s, 16 concurrent captures",
        "value": res["mixed-fp16"]["aggregate_fields_per_s"],
        "mixed_vs_fp32": round(res["mixed-fp16"]["aggregate_fields_per_s"] / res["fp32"]["aggregate_fields_per_s"], 3),
        "unit": "fields/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "data": "synthetic",
        "config": {"workload": "16 requests, Poisson 4 req/s, fs in {28, 35.8, 40} MHz, 30..3000 fields log-uniform",
                   "workers": 4, "batch_fields": 64, "hbm_budget_gb": 64, "stage_graphs": g},
        "forms": res,
    }
    print(json.dumps(line), flush=True)


def guard_blocks(dec, d_cap):
    """FP64 guard-band blocks (part A, part B) one decode recomputes."""
    ch = dec.decode_chunked(d_cap)["chunks"]
    dec.reset()
    return [sum(c["guard_blocks"][0] for c in ch), sum(c["guard_blocks"][1] for c in ch)]


def time_form(native, torch, d_cap, mhz, args, form, system="ntsc"):
    """One precision form on the same capture and schedule as the headline:
    warmup, K timed decodes, one serialised decode for stage times."""
    d = native.Decoder(mhz, args.batch, device=0, precision=form, graphs=bool(args.graphs), system=system)
    n = 0
    for _ in range(args.warmup):
        d.reset()
        d.decode_capture(d_cap)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        d.reset()
        n = d.decode_capture(d_cap)["fields"]
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    d.profile(True, serial=True)
    d.reset()
    d.decode_capture(d_cap)
    iso = d.profile_read()
    d.profile(False)
    r = {"value": round(n * args.steps / dt, 2), "unit": "fields/s", "ms_per_step": round(dt * 1000.0 / args.steps, 3),
         "dtype": DTYPES[form], "stage_ms_per_step_isolated": {k: round(v, 3) for k, v in iso.items()},
         "guard_blocks_per_step": guard_blocks(d, d_cap) if form.startswith("mixed") else None,
         "matrix_flops_per_segment": d.k1_matrix_flops()}
    d.close()
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the other precision forms and their reports")
    ap.add_argument("--form", default="mixed-fp16", choices=FORMS, help="precision form of the headline value")
    ap.add_argument("--workload", default="600f_28mhz", choices=sorted(WORKLOADS),
                    help="the capture a bench line is taken on (default: the 600-field 28 MHz workload)")
    ap.add_argument("--serving", action="store_true",
                    help="instead: the serving loop (16 concurrent captures at 28/35.8/40 MHz, Poisson arrivals), "
                         "one JSON line with latency percentiles and aggregate fields/s per precision form")
    ap.add_argument("--cpu-baseline-full", action="store_true",
                    help="instead: the CPU baseline on the 4-field and 600-field 28 MHz captures and 640 fields of "
                         "the 40 MHz capture with every host thread, and the 4-field capture with one thread")
    ap.add_argument("--no-cli", action="store_true", help="skip the CLI end-to-end figure")
    ap.add_argument("--graphs", type=int, default=0, choices=(0, 1),
                    help="1: the decode replays captured HIP graphs of its stages (cuvhs_config.graphs); "
                         "0: it launches the kernels one by one")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 4 plain device copies of MIB MiB (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the decode path shards by "
              "capture segment, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.cpu_baseline_full:
        cpu_baseline_full(args)
        return
    import numpy as np  # noqa: F401
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
    system = wl.get("system", "ntsc")
    if wl["fields"]:
        cap = native.generate(wl["config"], system=system,
                              n_samples=wl["fields"] * native.format_for(mhz, system).samples_per_field)
    else:
        cap = native.generate(wl["config"])
    log(f"generated {cap.size} samples in {time.perf_counter() - t:.1f} s")
    dev = torch.device("cuda:0")
    d_cap = torch.from_numpy(cap).to(dev)
    dec = native.Decoder(mhz, args.batch, device=0, precision=args.form, graphs=bool(args.graphs), system=system)
    fmt = dec.fmt
    matrix_flops_per_seg = dec.k1_matrix_flops()
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
    guard = guard_blocks(dec, d_cap)
    dec.close()
    ms_per_step = dt * 1000.0 / args.steps
    value = fields * args.steps / dt
    bytes_per_field = alg_bytes_per_field(mhz, system)
    log(f"{fields} fields/step, {ms_per_step:.2f} ms/step, {value:.1f} fields/s")
    stage_ms = dict(prof)
    log("per-step stage ms (overlapped): " + ", ".join(f"{k}={v:.3f}" for k, v in stage_ms.items()))
    log("per-step stage ms (serialised): " + ", ".join(f"{k}={v:.3f}" for k, v in iso.items()))

    # The dominant kernel (largest serialised stage time among the
    # transform stages) against the HBM roof, in the units of the metric:
    # the compulsory bytes of the fields one launch decodes (unique u8 input
    # plus the luma and chroma planes, alg_bytes_per_field) over its average
    # launch time, from HIP events on the stream it runs on.
    dom = max((k for k in ROOFLINE_STAGES if k in iso), key=lambda k: iso[k])
    launches = chunks                      # one launch per chunk of the profiled decode
    bpf = alg_bytes_per_field(mhz, system)
    fields_per_launch = fields / launches
    alg_bytes_per_launch = bpf * fields_per_launch
    avg_launch_s = stage_ms[dom] / 1000.0 / launches
    iso_launch_s = iso[dom] / 1000.0 / launches
    achieved = alg_bytes_per_launch / avg_launch_s / 1e9
    traffic = traffic_ratio = None
    pmc_path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(pmc_path) and args.workload == "600f_28mhz":  # counters taken on that workload
        with open(pmc_path) as f:
            pmc = json.load(f)
        if pmc.get("form") == args.form:   # and on one precision form
            traffic = pmc.get("bytes_per_launch", {}).get(dom)
            if pmc.get("bytes_per_decode"):
                traffic_ratio = round(pmc["bytes_per_decode"] / (bpf * fields), 2)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "alg_bytes_per_field": bpf, "fields_per_launch": round(fields_per_launch, 2),
                "alg_bytes_per_launch": int(alg_bytes_per_launch), "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                "achieved_isolated": round(alg_bytes_per_launch / iso_launch_s / 1e9, 1),
                "avg_launch_ms_isolated": round(iso_launch_s * 1e3, 4),
                "decode_traffic_ratio": traffic_ratio}

    # the same kernel against the FP64 arithmetic roof, in the reference's
    # FFT formulation (5 N log2 N per complex transform of a segment's
    # N-point window, half for a real one): part A one D2Z and one Z2Z, part
    # B one D2Z and two Z2D per segment (fm_demod.cu:396-486)
    roofline_fp64 = None
    if dom in ("k1_analytic", "k1_video") and args.form == "fp64":
        from cuvhs_amd import native as _n
        segs = k1_samples / fmt.spfp / launches
        n_fft = _n.k1_fft_size(mhz, system)
        flops = segs * k1_fft_flops_per_segment(dom, n_fft)
        roofline_fp64 = {"bound": "fp64", "kernel": dom, "fft_points": n_fft, "segments_per_launch": round(segs, 2),
                         "alg_flops_per_launch": int(flops), "unit": "TFLOP/s", "peak": PEAK_FP64_TFLOPS,
                         "achieved": round(flops / avg_launch_s / 1e12, 2),
                         "frac": round(flops / avg_launch_s / 1e12 / PEAK_FP64_TFLOPS, 4),
                         "achieved_isolated": round(flops / iso_launch_s / 1e12, 2),
                         "frac_isolated": round(flops / iso_launch_s / 1e12 / PEAK_FP64_TFLOPS, 4)}

    # the matrix cores' share of K1 part A in the mixed forms: executed MFMA
    # flops (every hi/lo product, padding steps included) over the part's
    # serialised stage time, against the dense f16 / bf16 peak
    def mfma_use(per_seg, iso_ms):
        if not per_seg or "k1_analytic" not in iso_ms:
            return None
        fl = per_seg * k1_samples / fmt.spfp
        tf = fl / (iso_ms["k1_analytic"] / 1000.0) / 1e12
        return {"kernel": "k1_part_a_mfma", "flops_per_segment": per_seg, "flops_per_step": fl,
                "achieved_tflops": round(tf, 1), "peak_tflops": PEAK_F16_TFLOPS, "frac": round(tf / PEAK_F16_TFLOPS, 4),
                "k1_analytic_ms_isolated": round(iso_ms["k1_analytic"], 3)}

    from cuvhs_amd import precision
    head_report = None
    if args.form != "fp64" and not args.no_forms:
        log(f"{args.form}: comparing with fp64 ...")
        head_report = precision.compare(d_cap, mhz, args.batch, args.form, system=system)
    forms = None
    if not args.no_forms:
        forms = {}
        for form in FORMS:
            if form == args.form:
                continue
            r = time_form(native, torch, d_cap, mhz, args, form, system)
            log(f"{form} form: {r['value']:.1f} fields/s")
            r["mfma"] = mfma_use(r.pop("matrix_flops_per_segment"), r["stage_ms_per_step_isolated"])
            if form != "fp64":
                log(f"{form}: comparing with fp64 ...")
                r["vs_fp64"] = precision.compare(d_cap, mhz, args.batch, form, system=system)
            forms[form] = r

    cli = None
    if not args.no_cli and wl["fields"] != 6000:   # the CLI figure would write a 4 GB file first
        cli = cli_end_to_end(cap, mhz, args.batch, args.form, system)
        log(f"cli end to end: {cli}")

    cpu = None
    if not args.no_cpu_baseline:
        log("running CPU baseline ...")
        cpu = cpu_baseline(cap, fmt, args.batch, mhz, system)
        log(f"cpu baseline {cpu['value']:.2f} fields/s on {cpu['cores']} threads")

    line = {
        "metric": (METRIC if mhz == 28.0 else METRIC.replace("28 MHz", f"{mhz:g} MHz")).replace(
            "NTSC", system.upper()),
        "value": round(value, 2),
        "unit": "fields/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.form],
        "data": "synthetic",
        "config": {"workload": wl["desc"],
                   "fields_per_step": fields, "batch_fields": args.batch, "sample_rate_mhz": mhz,
                   "precision": args.form + ("" if args.form == "fp64" else " (chroma pre-band-pass fp32)"),
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                   "stage_graphs": bool(args.graphs)},
        "hbm_gbs_algorithmic": round(value * bytes_per_field / 1e9, 2),
        "stage_ms_per_step": {k: round(v, 3) for k, v in stage_ms.items()},
        "stage_ms_per_step_isolated": {k: round(v, 3) for k, v in iso.items()},
        "roofline": roofline,
        "roofline_fp64": roofline_fp64,
        "mfma": mfma_use(matrix_flops_per_seg, iso),
        "guard_blocks_per_step": guard,
        "vs_fp64": head_report,
        "mixed_vs_fp32": (round(value / forms["fp32"]["value"], 3)
                          if forms and "fp32" in forms and args.form.startswith("mixed") else None),
        "cpu_baseline": cpu,
        "cli_end_to_end": cli,
        "forms": forms,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
