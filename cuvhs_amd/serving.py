"""Serving loop: many captures decoded concurrently on one GPU.

Requests arrive as a Poisson process; each names a device-resident capture
(sample rate from {28, 35.8, 40} MHz, length log-uniform in fields).  A pool
of worker threads owns one decode context per sample rate each (a context is
single-threaded; contexts are independent, so the workers' decodes overlap on
the GPU), takes requests in arrival order and decodes each with
cuvhs_decode_capture into device output buffers.  Admission is by HBM
budget: the workers' contexts are created up front and a request starts only
when its output buffers fit in what the budget leaves.

Reported per form (FP64 / FP32 K1): request latency (arrival to decode
complete), per-field latency (each field counted with its request's
latency), the decode-only time per field, and aggregate fields/s over the
whole run (first arrival to last completion).
"""
import math
import queue
import threading
import time

import numpy as np

from . import native as N

RATES = (28.0, 35.8, 40.0)
FIELD_LINES, LINE_LEN, MAX_DO = N.FIELD_LINES, N.LINE_LEN, N.MAX_DROPOUTS


def output_bytes(nfields):
    # luma + chroma u16, file_loc u64, 3 ints, 3 dropout lists
    return nfields * (2 * FIELD_LINES * LINE_LEN * 2 + 8 + 3 * 4 + 3 * MAX_DO * 4)


def make_pool(fields_per_rate, seed=4):
    """One synthetic capture per sample rate (the benchmark's signal model),
    returned as host arrays; requests are slices of these."""
    pool = {}
    for i, mhz in enumerate(RATES):
        spf = N.format_for(mhz).samples_per_field
        pool[mhz] = N.generate(1, sample_rate_mhz=mhz, n_samples=fields_per_rate * spf, seed=seed * 10 + i)
    return pool


def make_requests(n, rate_hz, min_fields, max_fields, pool_fields, seed=4):
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate_hz)
        mhz = RATES[int(rng.integers(0, len(RATES)))]
        nf = int(round(math.exp(rng.uniform(math.log(min_fields), math.log(max_fields)))))
        nf = max(min_fields, min(nf, pool_fields - 2))
        start = int(rng.integers(0, pool_fields - nf - 1))
        reqs.append({"id": i, "arrival": t, "mhz": mhz, "fields": nf, "start_field": start})
    return reqs


def run(d_pool, reqs, precision="fp64", workers=4, batch=64, hbm_budget_gb=64.0, device=0, keep_outputs=False,
        graphs=False):
    """Serve reqs (arrival times in seconds from now) against the device pool
    {mhz: uint8 cuda tensor}; returns the per-request records (with the
    decode's output tensors under "out" when keep_outputs)."""
    ctx = [{m: N.Decoder(m, batch, device=device, precision=precision, graphs=graphs) for m in RATES}
           for _ in range(workers)]
    budget = hbm_budget_gb * 1e9
    lock = threading.Condition()
    in_use = [0.0]
    q = queue.Queue()
    done = []
    errors = []

    def worker(w):
        while True:
            r = q.get()
            if r is None:
                return
            dec = ctx[w][r["mhz"]]
            spf = dec.fmt.samples_per_field
            # output room: the decoder's bound for the capture's length (a
            # capture yields more fields than field periods when the tape runs
            # fast, e.g. 2693 from 2689 at 40 MHz with head-switch offsets)
            cap_fields = dec.default_capacity(r["fields"] * spf)
            need = output_bytes(cap_fields)
            with lock:
                while in_use[0] + need > budget:
                    lock.wait()
                in_use[0] += need
            try:
                a = r["start_field"] * spf
                cap = d_pool[r["mhz"]][a: a + r["fields"] * spf]
                r["start"] = time.perf_counter() - t0
                dec.reset()
                # returns once the decoder's own streams are idle
                out = dec.decode_capture(cap, max_fields=cap_fields)
                r["end"] = time.perf_counter() - t0
                r["decoded"] = out["fields"]
                if keep_outputs:
                    r["out"] = out
                done.append(r)
            except Exception as e:  # reported after the run
                errors.append((r["id"], repr(e)))
            finally:
                with lock:
                    in_use[0] -= need
                    lock.notify_all()

    threads = [threading.Thread(target=worker, args=(w,), daemon=True) for w in range(workers)]
    t0 = time.perf_counter()
    for th in threads:
        th.start()
    for r in reqs:
        wait = r["arrival"] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        q.put(dict(r))
    for _ in threads:
        q.put(None)
    for th in threads:
        th.join()
    for c in ctx:
        for d in c.values():
            d.close()
    if errors:
        raise N.CuvhsError(f"serving errors: {errors[:3]}")
    return done


def summarize(done):
    lat = np.array([r["end"] - r["arrival"] for r in done])
    fields = np.array([r["decoded"] for r in done])
    per_field = np.repeat(lat, fields)
    decode_per_field = np.array([(r["end"] - r["start"]) / max(1, r["decoded"]) for r in done])
    span = max(r["end"] for r in done) - min(r["arrival"] for r in done)
    return {
        "requests": len(done),
        "fields": int(fields.sum()),
        "aggregate_fields_per_s": round(float(fields.sum()) / span, 1),
        "span_s": round(span, 3),
        "request_latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
                               "p99": round(float(np.percentile(lat, 99)) * 1e3, 2)},
        "field_latency_ms": {"p50": round(float(np.percentile(per_field, 50)) * 1e3, 2),
                             "p99": round(float(np.percentile(per_field, 99)) * 1e3, 2)},
        "decode_ms_per_field": {"p50": round(float(np.percentile(decode_per_field, 50)) * 1e3, 4),
                                "p99": round(float(np.percentile(decode_per_field, 99)) * 1e3, 4)},
    }
