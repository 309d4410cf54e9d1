"""ctypes access to the CPU FP64 oracle (liboracle.so).  Test infrastructure:
only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use it."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAX_PULSES = 800
MAX_DO = 512
LPF = 525
SYSTEMS = {"ntsc": 0, "pal": 1}
# (line locations per frame, output lines per field, samples per output line,
# chroma line transform points)
GEOMETRY = {"ntsc": (525, 263, 910, 1024), "pal": (625, 313, 1135, 2048)}

_vp = ctypes.c_void_p
_L = None


def build(verbose=False):
    import subprocess
    r = subprocess.run(["make", "-C", HERE, "liboracle.so"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    if verbose:
        print(r.stdout)


def lib():
    global _L
    if _L is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.or_create.restype = _vp
        L.or_create.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.or_create_sys.restype = _vp
        L.or_create_sys.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.or_format_init_sys.argtypes = [ctypes.c_int, ctypes.c_double, _vp]
        L.or_k3_linelocs_sys.argtypes = [ctypes.c_int, ctypes.c_double, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]
        L.or_k2b_chunk_sys.argtypes = [ctypes.c_int, ctypes.c_double, _vp, ctypes.c_uint64, _vp, ctypes.c_int] + [_vp] * 8
        L.or_destroy.argtypes = [_vp]
        L.or_decode.argtypes = [_vp, _vp, ctypes.c_size_t]
        L.or_num_fields.argtypes = [_vp]
        L.or_num_chunks.argtypes = [_vp]
        L.or_trace_candidates.argtypes = [_vp, ctypes.c_int]
        L.or_chroma_trial.argtypes = [_vp, _vp, _vp]
        L.or_luma.restype = _vp
        L.or_luma.argtypes = [_vp]
        L.or_chroma.restype = _vp
        L.or_chroma.argtypes = [_vp]
        L.or_field_meta.argtypes = [_vp] * 5
        L.or_field_dropouts.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp]
        L.or_write_outputs.argtypes = [_vp, ctypes.c_char_p]
        L.or_trace_chunk.argtypes = [_vp, ctypes.c_int] + [_vp] * 14
        L.or_fm_demod.argtypes = [ctypes.c_double, _vp, ctypes.c_int, _vp, _vp, _vp]
        L.or_vsync_discover.argtypes = [ctypes.c_double, _vp, ctypes.c_int, _vp, ctypes.c_int]
        L.or_sync_pulses.argtypes = [ctypes.c_double, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]
        L.or_hsync_refine.argtypes = [ctypes.c_double, _vp, ctypes.c_int, _vp, ctypes.c_int]
        L.or_tbc_resample.argtypes = [ctypes.c_double, _vp, ctypes.c_int, _vp, ctypes.c_int, _vp]
        L.or_filter_table.argtypes = [ctypes.c_double, ctypes.c_int, _vp]
        L.or_fft.argtypes = [ctypes.c_int, _vp, _vp, ctypes.c_int]
        L.or_format_init.argtypes = [ctypes.c_double, _vp]
        L.or_decode_file.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_int]
        L.or_decode_file_fmt.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int]
        L.or_set_input.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.or_convert.argtypes = [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_size_t, _vp]
        L.or_set_line_locator.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.or_chunk_diagnostics.argtypes = [_vp, ctypes.c_int, _vp]
        L.or_diag_counts.argtypes = [ctypes.c_double, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]
        L.or_k3_linelocs.argtypes = [ctypes.c_double, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]
        L.or_chroma_trial_lines.argtypes = [_vp, ctypes.c_int, _vp]
        L.or_stage_seconds.argtypes = [_vp, _vp]
        L.or_set_threads.argtypes = [ctypes.c_int]
        L.or_k2b_chunk.argtypes = [ctypes.c_double, _vp, ctypes.c_uint64, _vp, ctypes.c_int] + [_vp] * 8
        L.or_lookup_ntsc_phase.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_burst_metric.argtypes = [ctypes.c_double, _vp]
        L.or_burst_metric.restype = ctypes.c_double
        L.or_burst_phase.argtypes = [ctypes.c_double, _vp]
        L.or_burst_phase.restype = ctypes.c_double
        L.or_burst_metric_sys.argtypes = [ctypes.c_int, ctypes.c_double, _vp]
        L.or_burst_metric_sys.restype = ctypes.c_double
        L.or_burst_phase_sys.argtypes = [ctypes.c_int, ctypes.c_double, _vp]
        L.or_burst_phase_sys.restype = ctypes.c_double
        _L = L
    return _L


STAGES = ("convert", "k1_fm_demod", "vsync_cluster", "k2_pulses", "k2b_linelocs", "k4_hsync_refine",
          "k5_tbc_resample", "k6_chroma", "k7_dropouts")


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops; returns the count in effect."""
    return lib().or_set_threads(int(n))


FORMATS = {np.dtype(np.uint8): 0, np.dtype(np.int16): 1, np.dtype(np.uint16): 2}


def decode(data, mhz, batch, trace=True, dc_correct=False, line_locator=0, diagnostics=False, system="ntsc"):
    """Decode a u8 capture; returns a dict of numpy outputs (+ per-chunk traces)."""
    L = lib()
    lpf, rows, width, nfft = GEOMETRY[system]
    d = L.or_create_sys(SYSTEMS[system], float(mhz), int(batch), 1 if trace else 0)
    try:
        data = np.ascontiguousarray(data)
        if data.dtype not in FORMATS:
            raise ValueError(f"capture dtype {data.dtype} is not u8, s16 or u16")
        L.or_set_input(d, FORMATS[data.dtype], 1 if dc_correct else 0)
        L.or_set_line_locator(d, int(line_locator), 1 if diagnostics else 0)
        n = L.or_decode(d, data.ctypes.data, data.size)
        out = {"fields": n}
        st = np.zeros(9)
        L.or_stage_seconds(d, st.ctypes.data)
        out["stage_seconds"] = dict(zip(STAGES, st.tolist()))
        tbc = n * rows * width
        for key, fn in (("luma", L.or_luma), ("chroma", L.or_chroma)):
            if n == 0:  # no field was produced: the plane buffers were never allocated
                out[key] = np.zeros((0, rows, width), np.uint16)
            else:
                ptr = ctypes.cast(fn(d), ctypes.POINTER(ctypes.c_uint16))
                out[key] = np.ctypeslib.as_array(ptr, (tbc,)).reshape(n, rows, width).copy()
        fl = np.zeros(n, np.uint64)
        isf = np.zeros(n, np.int32)
        ph = np.zeros(n, np.int32)
        dc = np.zeros(n, np.int32)
        L.or_field_meta(d, fl.ctypes.data, isf.ctypes.data, ph.ctypes.data, dc.ctypes.data)
        out.update(file_loc=fl, is_first=isf, phase_id=ph, do_count=dc)
        dl = np.zeros((n, MAX_DO), np.int32)
        ds = np.zeros((n, MAX_DO), np.int32)
        de = np.zeros((n, MAX_DO), np.int32)
        for i in range(n):
            L.or_field_dropouts(d, i, dl[i].ctypes.data, ds[i].ctypes.data, de[i].ctypes.data)
        out.update(do_line=dl, do_startx=ds, do_endx=de)
        chunks = []
        if trace:
            for c in range(L.or_num_chunks(d)):
                B = batch
                ro = np.zeros(1, np.uint64)
                fk = np.zeros(1, np.int32)
                no = np.zeros(1, np.uint64)
                offs = np.zeros(B, np.uint64)
                pc = np.zeros(B, np.int32)
                ps = np.zeros(B * MAX_PULSES, np.int32)
                pl = np.zeros(B * MAX_PULSES, np.int32)
                vc = np.zeros(B, np.int32)
                vs = np.zeros(B * MAX_PULSES, np.int32)
                vl = np.zeros(B * MAX_PULSES, np.int32)
                vt = np.zeros(B * MAX_PULSES, np.int32)
                llc = np.zeros(B * lpf)
                llr = np.zeros(B * lpf)
                isk = np.zeros(B, np.int32)
                nf = L.or_trace_chunk(d, c, *[a.ctypes.data for a in (ro, fk, no, offs, pc, ps, pl, vc, vs, vl, vt, llc, llr, isk)])
                chunks.append(dict(raw_offset=int(ro[0]), fields_k1=int(fk[0]), next_offset=int(no[0]), nf=nf,
                                   offsets=offs[:nf], k2_count=pc[:nf], k2_starts=ps.reshape(B, -1)[:nf],
                                   k2_lens=pl.reshape(B, -1)[:nf], vp_count=vc[:nf], vp_starts=vs.reshape(B, -1)[:nf],
                                   vp_lens=vl.reshape(B, -1)[:nf], vp_types=vt.reshape(B, -1)[:nf],
                                   ll_coarse=llc.reshape(B, -1)[:nf], ll_refined=llr.reshape(B, -1)[:nf],
                                   is_first_k2b=isk[:nf], candidates=L.or_trace_candidates(d, c)))
        out["chunks"] = chunks
        m = np.zeros(2)
        tk = np.zeros(1, np.int32)
        if L.or_chroma_trial(d, m.ctypes.data, tk.ctypes.data) == 0:
            out["track_trial"] = {"metrics": m, "track": int(tk[0])}
            if trace:
                lines = []
                for which in range(3):
                    buf = np.zeros(rows * nfft)
                    if L.or_chroma_trial_lines(d, which, buf.ctypes.data) == buf.size:
                        lines.append(buf.reshape(rows, nfft))
                if len(lines) == 3:
                    out["track_trial"]["lines"] = lines
        if diagnostics:
            diag = []
            for c in range(L.or_num_chunks(d) if trace else 10 ** 6):
                v = np.zeros(9, np.int32)
                if L.or_chunk_diagnostics(d, c, v.ctypes.data) != 0:
                    break
                diag.append(v)
            out["diag"] = np.array(diag).reshape(-1, 9)
        return out
    finally:
        L.or_destroy(d)


def fm_demod(raw_s16, mhz, nfields):
    L = lib()
    raw = np.ascontiguousarray(raw_s16, dtype=np.float64)
    n = raw.size
    dm = np.zeros(n)
    d5 = np.zeros(n)
    env = np.zeros(n)
    L.or_fm_demod(float(mhz), raw.ctypes.data, int(nfields), dm.ctypes.data, d5.ctypes.data, env.ctypes.data)
    return dm, d5, env


def convert(data, dc_correct=False):
    """RawReader::convert of a u8 / s16 / u16 array (numpy) to doubles."""
    data = np.ascontiguousarray(data)
    out = np.zeros(data.size)
    rc = lib().or_convert(FORMATS[data.dtype], 1 if dc_correct else 0, data.ctypes.data, data.size, out.ctypes.data)
    if rc != 0:
        raise ValueError("or_convert failed")
    return out


def k3_linelocs(starts, lens, count, mhz, system="ntsc"):
    """K3 line location on K2 pulse lists ([nf, 800] int arrays, count [nf]):
    (nominal types, linelocs [nf, 525 or 625], is_first [nf])."""
    starts = np.ascontiguousarray(starts, np.int32)
    lens = np.ascontiguousarray(lens, np.int32)
    count = np.ascontiguousarray(count, np.int32)
    nf = count.size
    types = np.zeros((nf, MAX_PULSES), np.int32)
    ll = np.zeros((nf, GEOMETRY[system][0]))
    isf = np.zeros(nf, np.int32)
    lib().or_k3_linelocs_sys(SYSTEMS[system], float(mhz), starts.ctypes.data, lens.ctypes.data, count.ctypes.data, nf,
                             types.ctypes.data, ll.ctypes.data, isf.ctypes.data)
    return types, ll, isf


def diag_counts(which, mhz, a, b=None, c=None, nf=1):
    """One of the reference's debug counter sets (3: K3 spacing on linelocs a;
    4: K4 with demod_05 a, linelocs before b and after c; 5: K5 with demod a
    and linelocs b) over nf fields."""
    a = np.ascontiguousarray(a, np.float64)
    b = None if b is None else np.ascontiguousarray(b, np.float64)
    c = None if c is None else np.ascontiguousarray(c, np.float64)
    out = np.zeros(3, np.int32)
    rc = lib().or_diag_counts(float(mhz), int(which), a.ctypes.data, None if b is None else b.ctypes.data,
                              None if c is None else c.ctypes.data, a.size, int(nf), out.ctypes.data)
    if rc != 0:
        raise ValueError("bad counter set")
    return tuple(int(x) for x in out)


class K2bState(ctypes.Structure):
    """Cross-field K2b cadence state (pipeline.h:92-97), initial values of
    Pipeline's members."""
    _fields_ = [("prev_first_field", ctypes.c_int), ("prev_progressive_field", ctypes.c_int),
                ("prev_first_hsync_readloc", ctypes.c_longlong), ("prev_first_hsync_loc", ctypes.c_double),
                ("prev_first_hsync_diff", ctypes.c_double)]

    @classmethod
    def initial(cls):
        return cls(-1, -1, -1, -1.0, -1.0)


def k2b_chunk_with(fn, mhz, state, raw_offset, offsets, count, starts, lens, system=None):
    """Run one chunk of K2b through fn (or_k2b_chunk, or the reference's
    ref_k2b_chunk with the same signature; with system "ntsc" / "pal" the
    _sys forms, which take the system first) on K2 pulses count[nf],
    starts/lens[nf][<=800]; returns the validated pulses, types, good flags,
    line locations, is_first and the anchor values per field."""
    nf = len(offsets)
    st = np.zeros((nf, MAX_PULSES), np.int32)
    ln = np.zeros((nf, MAX_PULSES), np.int32)
    st[:, :starts.shape[1]] = starts
    ln[:, :lens.shape[1]] = lens
    cnt = np.ascontiguousarray(count, np.int32).copy()
    offs = np.ascontiguousarray(offsets, np.uint64)
    types = np.zeros((nf, MAX_PULSES), np.int32)
    good = np.zeros((nf, MAX_PULSES), np.uint8)
    ll = np.zeros((nf, GEOMETRY[system or "ntsc"][0]))
    isf = np.zeros(nf, np.int32)
    anchor = np.zeros((nf, 9))
    pre = () if system is None else (ctypes.c_int(SYSTEMS[system]),)
    rc = fn(*pre, ctypes.c_double(float(mhz)), ctypes.byref(state), ctypes.c_uint64(int(raw_offset)), offs.ctypes.data,
            nf, st.ctypes.data, ln.ctypes.data, cnt.ctypes.data, types.ctypes.data, good.ctypes.data,
            ll.ctypes.data, isf.ctypes.data, anchor.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"k2b chunk failed ({rc})")
    return dict(count=cnt, starts=st, lens=ln, types=types, good=good, linelocs=ll, is_first=isf, anchor=anchor)


def k2b_chunk(mhz, state, raw_offset, offsets, count, starts, lens, system="ntsc"):
    """The oracle's K2b (pipeline.cu:1717-1778 restated) on one chunk."""
    return k2b_chunk_with(lib().or_k2b_chunk_sys, mhz, state, raw_offset, offsets, count, starts, lens, system)


def lookup_ntsc_phase(is_first, quadrant, delta):
    return lib().or_lookup_ntsc_phase(int(is_first), int(quadrant), int(delta))


def burst_metric(mhz, lines, system="ntsc"):
    lines = np.ascontiguousarray(lines, np.float64)
    return lib().or_burst_metric_sys(SYSTEMS[system], float(mhz), lines.ctypes.data)


def burst_phase(mhz, lines, system="ntsc"):
    lines = np.ascontiguousarray(lines, np.float64)
    return lib().or_burst_phase_sys(SYSTEMS[system], float(mhz), lines.ctypes.data)


_R = None


def ref_k2b_lib():
    """The reference's own host K2b and chroma helpers, compiled by the
    Makefile target `refk2b` into oracle/_ref/libref_k2b.so when the reference
    sources are mounted; None when that library is absent."""
    global _R
    if _R is None:
        path = os.path.join(HERE, "_ref", "libref_k2b.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_k2b_chunk.argtypes = [ctypes.c_double, _vp, ctypes.c_uint64, _vp, ctypes.c_int] + [_vp] * 8
        R.ref_k2b_chunk_sys.argtypes = [ctypes.c_int, ctypes.c_double, _vp, ctypes.c_uint64, _vp, ctypes.c_int] + [_vp] * 8
        R.ref_lookup_ntsc_phase.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]
        R.ref_burst_cancellation.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_double]
        R.ref_burst_cancellation.restype = ctypes.c_double
        R.ref_burst_phase.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double]
        R.ref_burst_phase.restype = ctypes.c_double
        R.ref_chroma_params.argtypes = [ctypes.c_double, _vp]
        R.ref_chroma_params_sys.argtypes = [ctypes.c_int, ctypes.c_double, _vp]
        _R = R
    return _R


def ref_chroma_trial(R, mhz, lines, system="ntsc"):
    """The reference's measure_burst_cancellation on the track-0 and track-1
    trial lines and measure_burst_phase on the detected-track lines
    (chroma_decode.cu:688-731), with its own format constants."""
    p = np.zeros(6)
    R.ref_chroma_params_sys(SYSTEMS[system], float(mhz), p.ctypes.data)
    fsc, rate, L, rows, bs, be = p
    nfft = 1
    while nfft < int(L):
        nfft *= 2
    args = (int(rows), nfft, int(L), int(bs), int(be), 1.0 / nfft)
    m = [R.ref_burst_cancellation(np.ascontiguousarray(lines[t]).ctypes.data, *args) for t in (0, 1)]
    ph = R.ref_burst_phase(np.ascontiguousarray(lines[2]).ctypes.data, *args, fsc, rate)
    return m[0], m[1], ph
