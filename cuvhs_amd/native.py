"""ctypes bindings of include/cuvhs.h and the synthetic capture generator.

The decoder runs only through libcuvhs.so (HIP, gfx950).  Loading fails with
an exception when the library has not been built; there is no fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")

MAX_PULSES = 800
MAX_DROPOUTS = 512
LINES_PER_FRAME = 525      # NTSC geometry; a decoder's own is Decoder.geometry
FIELD_LINES = 263
LINE_LEN = 910
SYSTEMS = {"ntsc": 0, "pal": 1}
# (line locations per frame, output lines per field, samples per output line)
GEOMETRY = {"ntsc": (525, 263, 910), "pal": (625, 313, 1135)}

E_ARG, E_HIP, E_ALLOC, E_UNSUPPORTED, E_CAPACITY = -1, -2, -3, -4, -5
INPUT_U8, INPUT_S16, INPUT_U16 = 0, 1, 2
INPUT_FORMATS = {"u8": INPUT_U8, "s16": INPUT_S16, "u16": INPUT_U16}
LINE_LOCATORS = {"k2b": 0, "k3": 1}
PRECISIONS = {"fp64": 0, "fp32": 1, "mixed-fp16": 2, "mixed-bf16": 3}
DIAG_NAMES = ("k3_bad_spacing", "k3_isolated_spacing", "k3_large_jump", "k4_large_delta", "k4_isolated_jump",
              "k4_refined_sync_like", "k5_oob_pixels", "k5_bad_geom_lines", "k5_sync_like_pixels")
W_VSYNC_OVERFLOW, W_NO_VSYNC, W_EOF = 1, 2, 4

_vp = ctypes.c_void_p
_i = ctypes.c_int
_u64 = ctypes.c_uint64
_d = ctypes.c_double


class CuvhsError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [("system", _i), ("sample_rate_mhz", _d), ("batch_fields", _i), ("precision", _i),
                ("stream", _vp), ("device", _i), ("input_format", _i), ("dc_correct", _i), ("line_locator", _i),
                ("block_k1a", _i), ("block_k1b", _i), ("block_pre", _i), ("graphs", _i)]


class Format(ctypes.Structure):
    _fields_ = [("sample_rate", _d), ("samples_per_line", _i), ("samples_per_field", _i), ("spfp", _i),
                ("pulse_threshold_hz", _d), ("ire0", _d), ("hz_ire", _d), ("output_scale", _d),
                ("output_zero", _d), ("output_rate", _d), ("black16b", _d), ("white16b", _d),
                ("burst_start", _i), ("burst_end", _i), ("system", _i), ("lines_per_frame", _i),
                ("field_lines", _i), ("line_len", _i), ("fsc", _d)]


class ChunkInfo(ctypes.Structure):
    _fields_ = [("fields", _i), ("fields_k1", _i), ("vsync_candidates", _i), ("status", _i),
                ("raw_offset", _u64), ("next_raw_offset", _u64), ("guard_blocks_a", _i), ("guard_blocks_b", _i)]


class ChunkOutputs(ctypes.Structure):
    _fields_ = [(n, _vp) for n in ("luma", "chroma", "file_loc", "is_first", "phase_id", "do_count", "do_line",
                                   "do_startx", "do_endx")]


class GenParams(ctypes.Structure):
    _fields_ = [("sample_rate_mhz", _d), ("n_samples", ctypes.c_size_t), ("seed", _u64), ("lead_lines", _d),
                ("snr_lo_db", _d), ("snr_hi_db", _d), ("chroma_db", _d), ("dropouts_per_field", _d),
                ("dropout_len_lo_us", _d), ("dropout_len_hi_us", _d), ("wow_max_us", _d), ("wow_f_lo_hz", _d),
                ("wow_f_hi_hz", _d), ("headswitch_max_us", _d), ("headswitch_group_max", _i), ("system", _i)]


_SIGS = {
    "cuvhs_create": (_i, [ctypes.POINTER(Config), ctypes.POINTER(_vp)]),
    "cuvhs_destroy": (_i, [_vp]),
    "cuvhs_get_format": (_i, [_vp, ctypes.POINTER(Format)]),
    "cuvhs_format_for": (_i, [_d, ctypes.POINTER(Format)]),
    "cuvhs_format_for_system": (_i, [_i, _d, ctypes.POINTER(Format)]),
    "cuvhs_reset": (_i, [_vp]),
    "cuvhs_decode_chunk": (_i, [_vp, _vp, _u64, _u64, _i, ctypes.POINTER(ChunkInfo)]),
    "cuvhs_chunk_outputs_get": (_i, [_vp, ctypes.POINTER(ChunkOutputs)]),
    "cuvhs_decode_capture": (_i, [_vp, _vp, _u64, _i] + [_vp] * 9 + [ctypes.POINTER(_i)]),
    "cuvhs_fm_demod": (_i, [_vp, _vp, _u64, _i, _vp, _vp, _vp]),
    "cuvhs_vsync_discover": (_i, [_vp, _vp, _i, _vp, _i, ctypes.POINTER(_i)]),
    "cuvhs_sync_pulses": (_i, [_vp, _vp, _i, _vp, _i, _vp, _vp, _vp]),
    "cuvhs_linelocs": (_i, [_vp, _u64, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cuvhs_hsync_refine": (_i, [_vp, _vp, _i, _vp, _i]),
    "cuvhs_k3_linelocs": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "cuvhs_diagnostics": (_i, [_vp, _i]),
    "cuvhs_diagnostics_read": (_i, [_vp, _vp, _vp]),
    "cuvhs_tbc_resample": (_i, [_vp, _vp, _i, _vp, _i, _vp]),
    "cuvhs_chroma": (_i, [_vp, _vp, _u64, _vp, _i, _vp, _vp]),
    "cuvhs_dropouts": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cuvhs_dropouts_noalloc": (_i, [_d, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cuvhs_selftest_block": (_i, [_i, _i, _vp, _vp, _vp]),
    "cuvhs_selftest_atan2": (_i, [_vp, _vp, _vp, ctypes.c_int64]),
    "cuvhs_last_error": (ctypes.c_char_p, [_vp]),
    "cuvhs_profile": (_i, [_vp, _i]),
    "cuvhs_profile_read": (_i, [_vp, _vp, _vp, _i]),
    "cuvhs_profile_work": (_i, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "cuvhs_k1_matrix_flops": (_i, [_vp, ctypes.POINTER(_d)]),
    "cuvhs_k1_blocks": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "cuvhs_prebpf_block": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
}

_lib = None
_gen = None


def header_symbols():
    """Function names declared in include/cuvhs.h."""
    import re
    path = os.path.join(os.path.dirname(HERE), "include", "cuvhs.h")
    with open(path) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(cuvhs_[a-z_0-9]+)\s*\(", text)))


def lib():
    """The decoder library (raises CuvhsError when it is not built)."""
    global _lib
    if _lib is None:
        path = os.path.join(LIB_DIR, "libcuvhs.so")
        if not os.path.exists(path):
            raise CuvhsError(f"{path} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def gen_lib():
    global _gen
    if _gen is None:
        path = os.path.join(LIB_DIR, "librfgen.so")
        if not os.path.exists(path):
            raise CuvhsError(f"{path} is missing: build it with __graft_entry__.build()")
        G = ctypes.CDLL(path)
        G.cuvhs_gen_defaults.argtypes = [ctypes.POINTER(GenParams), _i]
        G.cuvhs_gen_defaults.restype = None
        G.cuvhs_gen_defaults_system.argtypes = [ctypes.POINTER(GenParams), _i, _i]
        G.cuvhs_gen_defaults_system.restype = None
        G.cuvhs_gen.argtypes = [ctypes.POINTER(GenParams), _vp]
        G.cuvhs_gen.restype = _i
        _gen = G
    return _gen


def gen_params(config, system="ntsc", **overrides):
    """Generator settings of a benchmark configuration; system "pal" gives
    the same signal model on PAL's line structure and levels (n_samples
    then counts PAL fields)."""
    p = GenParams()
    gen_lib().cuvhs_gen_defaults_system(ctypes.byref(p), int(config), SYSTEMS[system])
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def generate(config=0, system="ntsc", **overrides):
    """Synthetic u8 capture of one of the benchmark configurations (numpy array)."""
    import numpy as np
    p = gen_params(config, system, **overrides)
    out = np.empty(p.n_samples, dtype=np.uint8)
    rc = gen_lib().cuvhs_gen(ctypes.byref(p), out.ctypes.data)
    if rc != 0:
        raise CuvhsError(f"cuvhs_gen failed ({rc})")
    return out


_hip = None


def _hip_d2d(dst, src, nbytes):
    """Synchronous device-to-device copy through the HIP runtime."""
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, _i]
        _hip.hipMemcpy.restype = _i
    rc = _hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, 3)  # hipMemcpyDeviceToDevice
    if rc != 0:
        raise CuvhsError(f"hipMemcpy failed ({rc})")


def next_7smooth(n):
    """Smallest 2^a 3^b 5^c 7^d >= n (next_fft_size, fm_demod.cu:22-33)."""
    while True:
        m = n
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def k1_fft_size(mhz, system="ntsc"):
    """The reference's K1 transform length: a padded field window plus the
    2048-sample overlap, rounded up to 7-smooth (fm_demod.cu:273-280)."""
    return next_7smooth(int(format_for(mhz, system).spfp) + 2048)


def format_for(mhz, system="ntsc"):
    f = Format()
    rc = lib().cuvhs_format_for_system(SYSTEMS[system], float(mhz), ctypes.byref(f))
    if rc != 0:
        raise CuvhsError(f"cuvhs_format_for_system({system}, {mhz}) -> {rc}")
    return f


def check(rc, ctx=None, what="call"):
    if rc < 0:
        msg = lib().cuvhs_last_error(ctx).decode() if ctx else ""
        raise CuvhsError(f"{what} failed with {rc} {msg}")
    return rc


def _addr(t):
    """Device address of a torch tensor, torch's stream already synchronised."""
    return ctypes.c_void_p(t.data_ptr())


def ptr(t):
    """Device address of a torch tensor (or None), for a library call.

    The library runs on its own non-blocking streams, which do not wait for
    torch's: the tensor's memory may still be written (a fill, a copy) or
    read (an earlier tensor that held it) by work queued on torch's stream,
    so that stream is synchronised first."""
    if t is None:
        return None
    if t.is_cuda:
        import torch
        torch.cuda.current_stream(t.device).synchronize()
    return ctypes.c_void_p(t.data_ptr())


class Decoder:
    """One decode context (Pipeline, pipeline.h:28-114) on a HIP device."""

    def __init__(self, sample_rate_mhz=28.0, batch_fields=64, device=0, input_format="u8", dc_correct=False,
                 line_locator="k2b", precision="fp64", blocks=None, graphs=False, system="ntsc"):
        """blocks: optional (k1a, k1b, pre) block sizes, 0 for the tuned
        or modelled choice (cuvhs_config.block_*).  graphs: decode_capture
        replays captured stage graphs (cuvhs_config.graphs).  system: "ntsc"
        or "pal" (cuvhs_config.system)."""
        self.input_format = INPUT_FORMATS[input_format]
        self.precision = precision
        self.system = system
        self.geometry = GEOMETRY[system]
        ba, bb, bp = blocks or (0, 0, 0)
        cfg = Config(SYSTEMS[system], float(sample_rate_mhz), int(batch_fields), PRECISIONS[precision], None, int(device),
                     self.input_format, 1 if dc_correct else 0, LINE_LOCATORS[line_locator], int(ba), int(bb),
                     int(bp), 1 if graphs else 0)
        h = _vp()
        rc = lib().cuvhs_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise CuvhsError(f"cuvhs_create failed with {rc}")
        self.h = h
        self.mhz = float(sample_rate_mhz)
        self.batch = int(batch_fields)
        self.device = int(device)
        self.fmt = Format()
        check(lib().cuvhs_get_format(self.h, ctypes.byref(self.fmt)), self.h, "get_format")

    def close(self):
        if self.h:
            lib().cuvhs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def profile(self, enable=True, serial=False):
        """Stage timing; serial=True runs the stages one after another so each
        total is that stage's own device time."""
        mode = (2 if serial else 1) if enable else 0
        check(lib().cuvhs_profile(self.h, mode), self.h, "profile")

    def profile_read(self):
        """{stage name: accumulated device ms} since profiling was enabled."""
        n = 32
        ms = (ctypes.c_double * n)()
        names = (ctypes.c_char_p * n)()
        k = lib().cuvhs_profile_read(self.h, ms, names, n)
        return {names[i].decode(): ms[i] for i in range(k)}

    def diagnostics(self, enable=True):
        """Collect the reference's K3/K4/K5 debug counters per chunk (clears totals)."""
        check(lib().cuvhs_diagnostics(self.h, 1 if enable else 0), self.h, "diagnostics")

    def diagnostics_read(self):
        """({name: total since enabled}, {name: last chunk})."""
        tot = (ctypes.c_int64 * 9)()
        last = (ctypes.c_int32 * 9)()
        check(lib().cuvhs_diagnostics_read(self.h, tot, last), self.h, "diagnostics_read")
        return dict(zip(DIAG_NAMES, list(tot))), dict(zip(DIAG_NAMES, list(last)))

    def profile_work(self):
        """(K1 samples demodulated, chunks, fields) since profiling was enabled."""
        smp, ch, fl = _u64(), _i(), _i()
        check(lib().cuvhs_profile_work(self.h, ctypes.byref(smp), ctypes.byref(ch), ctypes.byref(fl)), self.h,
              "profile_work")
        return smp.value, ch.value, fl.value

    def k1_matrix_flops(self):
        """Matrix-core flops of K1 part A per segment (mixed forms; else 0)."""
        v = _d()
        check(lib().cuvhs_k1_matrix_flops(self.h, ctypes.byref(v)), self.h, "k1_matrix_flops")
        return v.value

    def k1_blocks(self):
        """K1 block shapes: (block_a, valid_a, block_b, valid_b)."""
        v = [_i() for _ in range(4)]
        check(lib().cuvhs_k1_blocks(self.h, *[ctypes.byref(x) for x in v]), self.h, "k1_blocks")
        return tuple(x.value for x in v)

    def prebpf_block(self):
        """Chroma pre-band-pass block shape and where the context's block
        sizes came from: (block, valid, source), source 0 config, 1 tuned
        table, 2 cost model."""
        v = [_i() for _ in range(3)]
        check(lib().cuvhs_prebpf_block(self.h, *[ctypes.byref(x) for x in v]), self.h, "prebpf_block")
        return tuple(x.value for x in v)

    def decode_chunked(self, d_capture):
        """Decode a device-resident u8 capture chunk by chunk through
        cuvhs_decode_chunk (the loop of Pipeline::run, pipeline.cu:2626-2713,
        as the CLI runs it), gathering each chunk's outputs.  Returns the same
        dict as decode_capture plus the per-chunk info records."""
        import torch
        self._check_input(d_capture)
        n = d_capture.numel()
        esz = d_capture.element_size()
        spf = self.fmt.samples_per_field
        keys = ("luma", "chroma", "file_loc", "is_first", "phase_id", "do_count", "do_line", "do_startx", "do_endx")
        parts = {k: [] for k in keys}
        # the capture was written on torch's stream; the decoder's streams are its own
        torch.cuda.current_stream(d_capture.device).synchronize()
        infos = []
        off = 0
        base = d_capture.data_ptr()
        while off < n and (n - off) // spf > 0:
            this_chunk = int(min(self.batch, (n - off) // spf))
            info = ChunkInfo()
            rc = lib().cuvhs_decode_chunk(self.h, ctypes.c_void_p(base + off * esz), n - off, off, this_chunk,
                                          ctypes.byref(info))
            check(rc, self.h, "decode_chunk")
            infos.append({"fields": info.fields, "fields_k1": info.fields_k1, "raw_offset": info.raw_offset,
                          "next_raw_offset": info.next_raw_offset, "status": info.status,
                          "vsync_candidates": info.vsync_candidates,
                          "guard_blocks": (info.guard_blocks_a, info.guard_blocks_b),
                          "diag": list(self.diagnostics_read()[1].values()) if info.fields else None})
            if info.fields == 0:
                break
            o = ChunkOutputs()
            check(lib().cuvhs_chunk_outputs_get(self.h, ctypes.byref(o)), self.h, "chunk_outputs_get")
            k = info.fields
            _, fl, lw = self.geometry
            shapes = {"luma": ((k, fl, lw), torch.uint16), "chroma": ((k, fl, lw), torch.uint16),
                      "file_loc": ((k,), torch.uint64), "is_first": ((k,), torch.int32), "phase_id": ((k,), torch.int32),
                      "do_count": ((k,), torch.int32), "do_line": ((k, MAX_DROPOUTS), torch.int32),
                      "do_startx": ((k, MAX_DROPOUTS), torch.int32), "do_endx": ((k, MAX_DROPOUTS), torch.int32)}
            for key in keys:
                shape, dt = shapes[key]
                t = torch.empty(shape, dtype=dt, device=d_capture.device)
                _hip_d2d(t.data_ptr(), getattr(o, key), t.numel() * t.element_size())
                parts[key].append(t)
            off = info.next_raw_offset
        out = {k: (torch.cat(v) if v else None) for k, v in parts.items()}
        # the concatenation runs on torch's stream and the parts' memory goes
        # back to torch's allocator: the next buffer a library call writes on
        # the decoder's (non-blocking) streams can be that memory, so the
        # copies must have read it first
        del parts
        torch.cuda.current_stream(d_capture.device).synchronize()
        out["fields"] = sum(i["fields"] for i in infos)
        out["chunks"] = infos
        return out

    def _check_input(self, d_capture):
        import torch
        want = {INPUT_U8: torch.uint8, INPUT_S16: torch.int16, INPUT_U16: torch.uint16}[self.input_format]
        if d_capture.dtype != want:
            raise CuvhsError(f"capture dtype {d_capture.dtype} does not match the decoder's input format ({want})")

    def reset(self):
        check(lib().cuvhs_reset(self.h), self.h, "reset")

    def default_capacity(self, n):
        """Output fields a capture of n samples can yield: one per field
        period, plus one per chunk (a chunk whose VSYNC search fails advances
        by F field periods less 10 lines, pipeline.cu:1614-1651), plus 1%
        for a tape running fast, plus one."""
        spf = int(self.fmt.samples_per_field)
        step = max(1, self.batch * spf - 10 * int(self.fmt.samples_per_line))
        per_period = n // spf
        return int(per_period + -(-n // step) + per_period // 100 + 1)

    def decode_capture(self, d_capture, max_fields=None):
        """Decode a device-resident capture (torch uint8 / int16 / uint16
        tensor on the GPU, matching the decoder's input format).  Returns a
        dict of device tensors trimmed to the decoded field count."""
        import torch
        self._check_input(d_capture)
        n = d_capture.numel()
        if max_fields is None:
            max_fields = self.default_capacity(n)
        dev = d_capture.device
        _, fl, lw = self.geometry
        luma = torch.empty((max_fields, fl, lw), dtype=torch.uint16, device=dev)
        chroma = torch.empty_like(luma)
        file_loc = torch.empty(max_fields, dtype=torch.uint64, device=dev)
        is_first = torch.empty(max_fields, dtype=torch.int32, device=dev)
        phase = torch.empty(max_fields, dtype=torch.int32, device=dev)
        doc = torch.empty(max_fields, dtype=torch.int32, device=dev)
        dol = torch.empty((max_fields, MAX_DROPOUTS), dtype=torch.int32, device=dev)
        dos = torch.empty_like(dol)
        doe = torch.empty_like(dol)
        nf = _i()
        # the capture was written on torch's stream; the decoder's streams are
        # its own (a device-wide sync would also wait for other decoders)
        torch.cuda.current_stream(dev).synchronize()
        rp = _addr   # torch's stream was synchronised just above
        rc = lib().cuvhs_decode_capture(self.h, rp(d_capture), n, max_fields, rp(luma), rp(chroma), rp(file_loc),
                                        rp(is_first), rp(phase), rp(doc), rp(dol), rp(dos), rp(doe),
                                        ctypes.byref(nf))
        check(rc, self.h, "decode_capture")
        k = nf.value
        return {"fields": k, "status": rc, "luma": luma[:k], "chroma": chroma[:k], "file_loc": file_loc[:k],
                "is_first": is_first[:k], "phase_id": phase[:k], "do_count": doc[:k], "do_line": dol[:k],
                "do_startx": dos[:k], "do_endx": doe[:k]}
