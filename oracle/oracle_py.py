"""ctypes access to the CPU FP64 oracle (liboracle.so).  Test infrastructure:
only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use it."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAX_PULSES = 800
MAX_DO = 512
LPF = 525

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
        _L = L
    return _L


FORMATS = {np.dtype(np.uint8): 0, np.dtype(np.int16): 1, np.dtype(np.uint16): 2}


def decode(data, mhz, batch, trace=True, dc_correct=False, line_locator=0, diagnostics=False):
    """Decode a u8 capture; returns a dict of numpy outputs (+ per-chunk traces)."""
    L = lib()
    d = L.or_create(float(mhz), int(batch), 1 if trace else 0)
    try:
        data = np.ascontiguousarray(data)
        if data.dtype not in FORMATS:
            raise ValueError(f"capture dtype {data.dtype} is not u8, s16 or u16")
        L.or_set_input(d, FORMATS[data.dtype], 1 if dc_correct else 0)
        L.or_set_line_locator(d, int(line_locator), 1 if diagnostics else 0)
        n = L.or_decode(d, data.ctypes.data, data.size)
        out = {"fields": n}
        tbc = n * 263 * 910
        for key, fn in (("luma", L.or_luma), ("chroma", L.or_chroma)):
            if n == 0:  # no field was produced: the plane buffers were never allocated
                out[key] = np.zeros((0, 263, 910), np.uint16)
            else:
                ptr = ctypes.cast(fn(d), ctypes.POINTER(ctypes.c_uint16))
                out[key] = np.ctypeslib.as_array(ptr, (tbc,)).reshape(n, 263, 910).copy()
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
                llc = np.zeros(B * LPF)
                llr = np.zeros(B * LPF)
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


def k3_linelocs(starts, lens, count, mhz):
    """K3 line location on K2 pulse lists ([nf, 800] int arrays, count [nf]):
    (nominal types, linelocs [nf, 525], is_first [nf])."""
    starts = np.ascontiguousarray(starts, np.int32)
    lens = np.ascontiguousarray(lens, np.int32)
    count = np.ascontiguousarray(count, np.int32)
    nf = count.size
    types = np.zeros((nf, MAX_PULSES), np.int32)
    ll = np.zeros((nf, 525))
    isf = np.zeros(nf, np.int32)
    lib().or_k3_linelocs(float(mhz), starts.ctypes.data, lens.ctypes.data, count.ctypes.data, nf, types.ctypes.data,
                         ll.ctypes.data, isf.ctypes.data)
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
