"""Golden vectors of the reference's own K2b line location and chroma phase
helpers, for the parity tests.

The reference keeps its host K2b (pipeline.cu:133-148, 410-893) and its NTSC
phase table and burst metrics (chroma_decode.cu:329-450) as plain C++; the
Makefile target `refk2b` in oracle/ compiles them as they lie into
oracle/_ref/libref_k2b.so (that needs the reference mounted, so only in the
build container).  This script feeds that library the K2 pulse lists the
oracle finds on degraded synthetic captures, chunk by chunk with the cadence
state carried across chunks as Pipeline does (pipeline.cu:1717-1778), and
stores inputs and outputs as data:

  ref_k2b_<case>.npz   K2 pulses per field and chunk (inputs), and the
                       reference's validated pulses, types, good flags, line
                       locations, is_first and anchor values (outputs)
  ref_chroma_phase.npz the reference's burst metric (both tracks) and burst
                       phase of the field-0 track trial of small captures,
                       computed on the oracle's band-passed lines, and the
                       lookup_ntsc_phase table for every key

Run from the repository root:  python tests/golden/make_ref_k2b_fixtures.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

# name -> (sample rate MHz, generator config, overrides, fields, field cut out
# of the tape or None[, system])
K2B_CASES = {
    "cfg1_601f": (28.0, 1, {}, None, None),
    "cfg2_640f_40mhz": (40.0, 2, {}, 640, None),
    "snr25_128f": (28.0, 1, dict(snr_lo_db=25.0, snr_hi_db=25.0, seed=5), 128, None),
    "snr18_128f": (28.0, 1, dict(snr_lo_db=18.0, snr_hi_db=20.0, seed=6), 128, None),
    "dropouts8_96f": (28.0, 1, dict(dropouts_per_field=8.0, seed=7), 96, None),
    "cut_field_70f": (28.0, 1, dict(seed=9), 71, 6),
    "cfg1_130f_35_8mhz": (35.8, 1, dict(sample_rate_mhz=35.8, seed=8), 130, None),
    # PAL (the reference's PAL branches of pipeline.cu:707-717, 811)
    "pal_cfg1_200f": (28.0, 1, {}, 200, None, "pal"),
    "pal_cfg2_192f_40mhz": (40.0, 2, {}, 192, None, "pal"),
    "pal_snr18_96f": (28.0, 1, dict(snr_lo_db=18.0, snr_hi_db=20.0, seed=16), 96, None, "pal"),
    "pal_cut_field_70f": (28.0, 1, dict(seed=19), 71, 5, "pal"),
}
BATCH = 64

# chroma trial captures: (MHz, config, overrides, fields, start field[, system])
CHROMA_CASES = [(28.0, 1, {}, 3, s) for s in range(4)] + [
    (28.0, 0, {}, 4, 0),
    (40.0, 2, {}, 3, 0),
    (40.0, 2, {}, 4, 1),
    (35.8, 1, dict(sample_rate_mhz=35.8, seed=8), 3, 0),
    (28.0, 1, dict(snr_lo_db=18.0, snr_hi_db=20.0, seed=6), 3, 2),
] + [(28.0, 1, {}, 3, s, "pal") for s in range(4)] + [(40.0, 2, {}, 3, 1, "pal")]


def case_system(case, n):
    """The system of a case tuple with n fixed elements ("ntsc" by default)."""
    return case[n] if len(case) > n else "ntsc"


def capture(mhz, config, over, fields, cut=None, start=0, system="ntsc"):
    from cuvhs_amd import native
    spf = native.format_for(mhz, system).samples_per_field
    kw = dict(over)
    if fields is not None:
        kw["n_samples"] = (fields + start) * spf
    g = native.generate(config, system=system, **kw)
    if cut is not None:
        g = np.concatenate([g[:cut * spf], g[(cut + 1) * spf:]])
    return np.ascontiguousarray(g[start * spf:])


# chunks kept in the committed vectors (the whole capture is compared when the
# vectors are made; the two long captures keep their first three chunks)
KEEP_CHUNKS = {"cfg1_601f": 3, "cfg2_640f_40mhz": 3}


def _delta(a):
    a = np.asarray(a, np.int64)
    return np.diff(a, prepend=0).astype(np.int32) if a.size else a.astype(np.int32)


def _undelta(d):
    return np.cumsum(np.asarray(d, np.int64)).astype(np.int32)


def pack_k2b(name, mhz, chunks, ref_fn, keep=None, system="ntsc"):
    """Inputs of the first `keep` chunks and the outputs of ref_fn (the
    reference's K2b chunk runner) on them, plus the cadence state entering
    each chunk.  Every chunk is also run through the oracle's K2b; returns
    the record and the number of chunks where the two differ in any output."""
    import oracle_py as O
    st = O.K2bState.initial()
    so = O.K2bState.initial()
    rec = {k: [] for k in ("raw_offset", "nf", "state_in", "offsets", "k2_count", "k2_starts", "k2_lens",
                           "vp_count", "vp_starts", "vp_lens", "vp_types", "vp_good", "linelocs", "is_first",
                           "anchor")}
    differ = 0
    for ci, c in enumerate(chunks):
        nf = c["nf"]
        state_in = [st.prev_first_field, st.prev_progressive_field, st.prev_first_hsync_readloc,
                    st.prev_first_hsync_loc, st.prev_first_hsync_diff]
        out = O.k2b_chunk_with(ref_fn, mhz, st, c["raw_offset"], c["offsets"], c["k2_count"], c["k2_starts"],
                               c["k2_lens"], system)
        mine = O.k2b_chunk(mhz, so, c["raw_offset"], c["offsets"], c["k2_count"], c["k2_starts"], c["k2_lens"],
                           system)
        if any(not np.array_equal(out[k], mine[k]) for k in out):
            differ += 1
        if keep is not None and ci >= keep:
            continue
        rec["raw_offset"].append(c["raw_offset"])
        rec["nf"].append(nf)
        rec["state_in"].append(np.array(state_in, np.float64))
        rec["offsets"].append(np.asarray(c["offsets"], np.uint64))
        rec["k2_count"].append(c["k2_count"].astype(np.int32))
        for f in range(nf):
            n = int(c["k2_count"][f])
            rec["k2_starts"].append(_delta(c["k2_starts"][f][:n]))
            rec["k2_lens"].append(c["k2_lens"][f][:n].astype(np.int32))
        rec["vp_count"].append(out["count"])
        for f in range(nf):
            n = int(out["count"][f])
            rec["vp_starts"].append(_delta(out["starts"][f][:n]))
            rec["vp_lens"].append(out["lens"][f][:n])
            rec["vp_types"].append(out["types"][f][:n].astype(np.int8))
            rec["vp_good"].append(out["good"][f][:n])
        rec["linelocs"].append(out["linelocs"])
        rec["is_first"].append(out["is_first"])
        rec["anchor"].append(out["anchor"])
    cat = {k: np.concatenate(v) if k not in ("raw_offset", "nf", "state_in") else np.asarray(v)
           for k, v in rec.items()}
    cat["raw_offset"] = cat["raw_offset"].astype(np.int64)
    cat["nf"] = cat["nf"].astype(np.int64)
    cat["mhz"] = np.float64(mhz)
    cat["system"] = np.int64(O.SYSTEMS[system])
    cat["batch"] = np.int64(BATCH)
    cat["chunks_total"] = np.int64(len(chunks))
    cat["fields_total"] = np.int64(sum(c["nf"] for c in chunks))
    cat["chunks_differing"] = np.int64(differ)
    return cat, differ


def unpack_k2b(z):
    """Per-chunk inputs, entering cadence state and expected outputs of a
    ref_k2b_<case>.npz."""
    from oracle_py import MAX_PULSES
    chunks = []
    fpos = spos = vpos = 0
    for ci, (ro, nf) in enumerate(zip(z["raw_offset"], z["nf"])):
        nf = int(nf)
        cnt = z["k2_count"][fpos:fpos + nf]
        vcnt = z["vp_count"][fpos:fpos + nf]
        starts = np.zeros((nf, MAX_PULSES), np.int32)
        lens = np.zeros((nf, MAX_PULSES), np.int32)
        exp = {k: np.zeros((nf, MAX_PULSES), np.int32) for k in ("starts", "lens", "types")}
        exp["good"] = np.zeros((nf, MAX_PULSES), np.uint8)
        for f in range(nf):
            n = int(cnt[f])
            starts[f, :n] = _undelta(z["k2_starts"][spos:spos + n])
            lens[f, :n] = z["k2_lens"][spos:spos + n]
            spos += n
            m = int(vcnt[f])
            exp["starts"][f, :m] = _undelta(z["vp_starts"][vpos:vpos + m])
            exp["lens"][f, :m] = z["vp_lens"][vpos:vpos + m]
            exp["types"][f, :m] = z["vp_types"][vpos:vpos + m]
            exp["good"][f, :m] = z["vp_good"][vpos:vpos + m]
            vpos += m
        exp["count"] = vcnt.astype(np.int32)
        exp["linelocs"] = z["linelocs"][fpos:fpos + nf]
        exp["is_first"] = z["is_first"][fpos:fpos + nf]
        exp["anchor"] = z["anchor"][fpos:fpos + nf]
        si = z["state_in"][ci]
        chunks.append(dict(raw_offset=int(ro), nf=nf, offsets=z["offsets"][fpos:fpos + nf],
                           k2_count=cnt.astype(np.int32), k2_starts=starts, k2_lens=lens, expected=exp,
                           state_in=(int(si[0]), int(si[1]), int(si[2]), float(si[3]), float(si[4]))))
        fpos += nf
    return chunks


def chroma_trial(mhz, config, over, fields, start, system="ntsc"):
    """The oracle's field-0 track trial lines of a small capture."""
    import oracle_py as O
    cap = capture(mhz, config, over, fields, start=start, system=system)
    r = O.decode(cap, mhz, BATCH, trace=True, system=system)
    return r


def main():
    import oracle_py as O
    R = O.ref_k2b_lib()
    if R is None:
        sys.exit("oracle/_ref/libref_k2b.so is missing: run `make -C oracle refk2b` with the reference mounted")
    only = sys.argv[1:]
    for name, case in K2B_CASES.items():
        if only and name not in only:
            continue
        mhz, config, over, fields, cut = case[:5]
        system = case_system(case, 5)
        t = time.perf_counter()
        cap = capture(mhz, config, over, fields, cut, system=system)
        r = O.decode(cap, mhz, BATCH, trace=True, system=system)
        rec, differ = pack_k2b(name, mhz, r["chunks"], R.ref_k2b_chunk_sys, KEEP_CHUNKS.get(name), system)
        path = os.path.join(HERE, f"ref_k2b_{name}.npz")
        np.savez_compressed(path, **rec)
        print(f"{name}: {r['fields']} fields, {len(r['chunks'])} chunks ({len(rec['nf'])} kept), "
              f"oracle vs reference: {differ} chunks differ, {os.path.getsize(path)} B, "
              f"{time.perf_counter() - t:.0f} s", flush=True)
    if only and "chroma" not in only:
        return
    rows = []
    for case in CHROMA_CASES:
        mhz, config, over, fields, start = case[:5]
        system = case_system(case, 5)
        r = chroma_trial(mhz, config, over, fields, start, system)
        m0, m1, ph = O.ref_chroma_trial(R, mhz, r["track_trial"]["lines"], system)
        rows.append([mhz, config, fields, start, m0, m1, ph, O.SYSTEMS[system]])
        print(f"chroma {system} {mhz} cfg{config} start {start}: metrics {m0:.4f} {m1:.4f} phase {ph:.3f}",
              flush=True)
    table = np.zeros((2, 4, 5, 2), np.int32)   # [is_first][quadrant][delta + 1] -> (phase_id, offset)
    import ctypes
    for isf in range(2):
        for q in range(4):
            for d in range(-1, 4):
                off = ctypes.c_int(0)
                pid = R.ref_lookup_ntsc_phase(isf, q, d, ctypes.byref(off))
                table[isf, q, d + 1] = (pid, off.value)
    cases = np.array(rows)
    overs = np.array([repr(c[2]) for c in CHROMA_CASES])
    np.savez_compressed(os.path.join(HERE, "ref_chroma_phase.npz"), cases=cases, overrides=overs, lookup=table)
    print("chroma phase fixture written")


if __name__ == "__main__":
    main()
