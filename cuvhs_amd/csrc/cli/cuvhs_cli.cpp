// `cuvhs` command line: the reference's drop-in surface (src/main.cpp:24-165,
// Pipeline::run pipeline.cu:2626-2713) on top of the C ABI in include/cuvhs.h.
//
//   cuvhs [--system NTSC|PAL] [-f MHz] [--format u8|s16|u16] [--dc-correct]
//         [--gpu id] [--batch F] [--precision fp64|fp32|mixed-fp16|mixed-bf16] [--overwrite]
//         <input | -> <output_base>
//
// Writes output_base.tbc, output_base_chroma.tbc and output_base.tbc.json.
// Exit status 0 on success, 1 on any failure.  --batch sets the chunk size F
// (default 64); the reference derives F from free VRAM and its results
// depend on it, so it is explicit here.  Input "-" (or a FIFO) is decoded as
// a stream with the same results as the same bytes in a file.  --precision
// fp32 runs the FM demodulation transforms in FP32 (default fp64, the
// reference's arithmetic).
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <strings.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/cuvhs.h"
#include "tbc_io.h"

namespace {

struct Args {
    std::string input, output;
    bool pal = false;
    double mhz = 28.0;
    std::string format;
    bool dc_correct = false;
    int gpu = 0;
    int batch = 64;
    int precision = CUVHS_PRECISION_FP64;
    bool overwrite = false;
};

void usage(const char* prog) {
    fprintf(stderr,
            "cuvhs (MI355X) - VHS RF decoder\n\n"
            "Usage: %s [options] <input_file> <output_base>\n\n"
            "Arguments:\n"
            "  input_file       Raw RF capture (.u8/.raw, .s16, .u16), or - for stdin\n"
            "  output_base      Output prefix (creates .tbc, _chroma.tbc, .tbc.json)\n\n"
            "Options:\n"
            "  --system <NTSC|PAL>   Video system (default: NTSC)\n"
            "  -f <MHz>              Sample rate in MHz (default: 28)\n"
            "  --format <u8|s16|u16> Input sample format (default: from extension)\n"
            "  --dc-correct          Subtract per-read DC offset after normalization\n"
            "  --gpu <id>            GPU device ID (default: 0, -1 = most memory)\n"
            "  --batch <fields>      Fields per chunk (default: 64)\n"
            "  --precision <fp64|fp32|mixed-fp16|mixed-bf16>\n"
            "                        decode arithmetic (default: fp64, the reference's)\n"
            "  --overwrite           Overwrite existing output files\n"
            "  -h, --help            Show this help\n",
            prog);
}

bool parse(int argc, char** argv, Args& a) {
    int pos = 0;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-h") || !strcmp(argv[i], "--help")) {
            usage(argv[0]);
            exit(0);
        } else if (!strcmp(argv[i], "--system") && i + 1 < argc) {
            i++;
            if (!strcasecmp(argv[i], "NTSC")) a.pal = false;
            else if (!strcasecmp(argv[i], "PAL")) a.pal = true;
            else { fprintf(stderr, "Unknown system: %s\n", argv[i]); return false; }
        } else if (!strcmp(argv[i], "-f") && i + 1 < argc) {
            a.mhz = atof(argv[++i]);
        } else if (!strcmp(argv[i], "--format") && i + 1 < argc) {
            a.format = argv[++i];
            if (a.format != "u8" && a.format != "s16" && a.format != "u16") {
                fprintf(stderr, "Unknown format: %s\n", a.format.c_str());
                return false;
            }
        } else if (!strcmp(argv[i], "--dc-correct")) {
            a.dc_correct = true;
        } else if (!strcmp(argv[i], "--gpu") && i + 1 < argc) {
            a.gpu = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--batch") && i + 1 < argc) {
            a.batch = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--precision") && i + 1 < argc) {
            i++;
            if (!strcmp(argv[i], "fp64")) a.precision = CUVHS_PRECISION_FP64;
            else if (!strcmp(argv[i], "fp32")) a.precision = CUVHS_PRECISION_FP32;
            else if (!strcmp(argv[i], "mixed-fp16")) a.precision = CUVHS_PRECISION_MIXED_F16;
            else if (!strcmp(argv[i], "mixed-bf16")) a.precision = CUVHS_PRECISION_MIXED_BF16;
            else { fprintf(stderr, "Unknown precision: %s\n", argv[i]); return false; }
        } else if (!strcmp(argv[i], "--overwrite")) {
            a.overwrite = true;
        } else if (argv[i][0] == '-' && strlen(argv[i]) > 1) {
            fprintf(stderr, "Unknown option: %s\n", argv[i]);
            return false;
        } else {
            if (pos == 0) a.input = argv[i];
            else if (pos == 1) a.output = argv[i];
            else { fprintf(stderr, "Too many arguments\n"); return false; }
            pos++;
        }
    }
    if (pos < 2) {
        fprintf(stderr, "Error: input_file and output_base are required\n\n");
        usage(argv[0]);
        return false;
    }
    return true;
}

std::string detect_format(const std::string& path) {
    auto ext = path.substr(path.find_last_of('.') + 1);
    if (ext == "s16") return "s16";
    if (ext == "u16") return "u16";
    return "u8";  // .u8, .raw and anything else (main.cpp:89-95)
}

// env_flag_enabled (pipeline.cu:37-40)
bool env_flag(const char* name) {
    const char* e = getenv(name);
    return e && e[0] && strcmp(e, "0") != 0;
}

#define CHECK_HIP(x)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "HIP error %s at %s\n", hipGetErrorString(e_), #x);           \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

// Capture input.  A regular file is read at chunk offsets with pread
// (RawReader::read_at, raw_reader.cpp:128-152).  A pipe, FIFO or stdin ("-")
// is read sequentially into a buffer that holds the samples from the current
// chunk start on (the reference refuses read_at on streams,
// raw_reader.cpp:129-132; here a stream decodes exactly as the same bytes in a
// file would): a chunk needs batch * spfp samples past its start, or the
// end of the stream, and chunk starts only move forward.
class InputSource {
  public:
    bool open(const std::string& path, int bps) {
        bps_ = bps;
        if (path == "-") {
            fd_ = STDIN_FILENO;
            own_ = false;
            stream_ = true;
            return true;
        }
        fd_ = ::open(path.c_str(), O_RDONLY);
        if (fd_ < 0) {
            perror(path.c_str());
            return false;
        }
        own_ = true;
        struct stat st;
        if (fstat(fd_, &st) < 0) {
            perror("fstat");
            return false;
        }
        stream_ = S_ISFIFO(st.st_mode) || S_ISCHR(st.st_mode);
        if (!stream_) {
            total_ = (uint64_t)st.st_size / (uint64_t)bps_;
            total_known_ = true;
        }
        return true;
    }
    ~InputSource() {
        if (own_ && fd_ >= 0) ::close(fd_);
    }
    bool is_stream() const { return stream_; }
    bool total_known() const { return total_known_; }
    uint64_t total() const { return total_; }

    // Make samples [offset, offset + want) available (fewer at the end of the
    // input) and copy them to dst; returns the count, or -1 on a read error
    // or an offset behind the stream buffer.
    long long read(uint64_t offset, size_t want, uint8_t* dst) {
        if (!stream_) {
            if (offset >= total_) return 0;
            size_t n = (size_t)std::min<uint64_t>(want, total_ - offset);
            ssize_t got = pread(fd_, dst, n * bps_, (off_t)(offset * bps_));
            return got < 0 ? -1 : got / bps_;
        }
        if (offset < base_) {
            fprintf(stderr, "stream input: chunk start %llu precedes buffered data %llu\n",
                    (unsigned long long)offset, (unsigned long long)base_);
            return -1;
        }
        // drop what precedes this chunk, then read until it is covered
        const size_t drop = (size_t)std::min<uint64_t>((offset - base_) * bps_, buf_.size());
        buf_.erase(buf_.begin(), buf_.begin() + drop);
        base_ += drop / bps_;
        if (base_ < offset) {  // the chunk starts past the data read so far: skip input
            if (!skip((offset - base_) * bps_)) return -1;
            base_ = offset;
        }
        const size_t need = want * bps_;
        while (!eof_ && buf_.size() < need) {
            size_t old = buf_.size();
            buf_.resize(need);
            ssize_t got = read_full(buf_.data() + old, need - old);
            if (got < 0) return -1;
            buf_.resize(old + (size_t)got);
            if ((size_t)got < need - old) eof_ = true;
        }
        if (eof_ && !total_known_) {
            total_ = base_ + buf_.size() / bps_;
            total_known_ = true;
        }
        size_t n = std::min(want, buf_.size() / bps_);
        memcpy(dst, buf_.data(), n * bps_);
        return (long long)n;
    }

  private:
    // read(2) until count bytes or end of input, retrying on EINTR
    // (RawReader::read_full, raw_reader.cpp:105-124)
    ssize_t read_full(uint8_t* p, size_t count) {
        size_t total = 0;
        while (total < count) {
            ssize_t n = ::read(fd_, p + total, count - total);
            if (n > 0) {
                total += (size_t)n;
            } else if (n == 0) {
                break;
            } else {
                if (errno == EINTR) continue;
                if (total > 0) break;
                return -1;
            }
        }
        return (ssize_t)total;
    }
    bool skip(uint64_t bytes) {
        std::vector<uint8_t> tmp(1 << 20);
        while (bytes > 0 && !eof_) {
            size_t n = (size_t)std::min<uint64_t>(bytes, tmp.size());
            ssize_t got = read_full(tmp.data(), n);
            if (got < 0) return false;
            if ((size_t)got < n) eof_ = true;
            bytes -= (uint64_t)got;
        }
        return true;
    }

    int fd_ = -1;
    bool own_ = false, stream_ = false, eof_ = false, total_known_ = false;
    int bps_ = 1;
    uint64_t total_ = 0;
    uint64_t base_ = 0;  // sample offset of buf_[0] (stream mode)
    std::vector<uint8_t> buf_;
};

}  // namespace

int main(int argc, char** argv) {
    Args a;
    if (!parse(argc, argv, a)) return 1;
    if (a.format.empty()) a.format = detect_format(a.input);
    if (a.batch < 1) { fprintf(stderr, "--batch must be >= 1\n"); return 1; }
    const int in_fmt = a.format == "s16" ? CUVHS_INPUT_S16 : (a.format == "u16" ? CUVHS_INPUT_U16 : CUVHS_INPUT_U8);
    const int bps = in_fmt == CUVHS_INPUT_U8 ? 1 : 2;

    cuvhs_config cfg{};
    cfg.system = a.pal ? CUVHS_SYSTEM_PAL : CUVHS_SYSTEM_NTSC;
    cfg.sample_rate_mhz = a.mhz;
    cfg.batch_fields = a.batch;
    cfg.precision = a.precision;
    cfg.stream = nullptr;
    cfg.device = a.gpu;
    cfg.input_format = in_fmt;
    cfg.dc_correct = a.dc_correct ? 1 : 0;
    // the reference switches to K3 line location when CUVHS_K2B_DISABLE is
    // set at all (pipeline.cu:1694)
    cfg.line_locator = getenv("CUVHS_K2B_DISABLE") ? CUVHS_LINELOC_K3 : CUVHS_LINELOC_K2B;
    const bool log_k3 = env_flag("CUVHS_DEBUG_K3_LOG"), log_k4 = env_flag("CUVHS_DEBUG_K4_LOG"),
               log_k5 = env_flag("CUVHS_DEBUG_K5_LOG");
    cuvhs_ctx* ctx = nullptr;
    int rc = cuvhs_create(&cfg, &ctx);
    if (rc != 0) {
        fprintf(stderr, "Failed to initialize GPU decoder (%d)\n", rc);
        return 1;
    }
    cuvhs_format fmt;
    cuvhs_get_format(ctx, &fmt);
    fprintf(stderr, "Format: %s @ %.1f MHz capture rate\n  Samples/line: %d  Samples/field: %d\n",
            a.pal ? "PAL" : "NTSC", a.mhz, fmt.samples_per_line, fmt.samples_per_field);

    if ((log_k3 || log_k4 || log_k5) && cuvhs_diagnostics(ctx, 1) != 0) {
        fprintf(stderr, "Failed to enable diagnostics\n");
        cuvhs_destroy(ctx);
        return 1;
    }
    InputSource src;
    if (!src.open(a.input, bps)) {
        fprintf(stderr, "Failed to open input: %s\n", a.input.c_str());
        cuvhs_destroy(ctx);
        return 1;
    }
    const char* dc_note = a.dc_correct ? ", dc-corrected" : "";
    if (src.is_stream())
        fprintf(stderr, "Input: %s (%.1f MHz, %s, streaming%s)\n", a.input == "-" ? "stdin" : a.input.c_str(), a.mhz,
                a.format.c_str(), dc_note);
    else
        fprintf(stderr, "Input: %s (%.1f MHz, %s, %.2f GB%s)\n", a.input.c_str(), a.mhz, a.format.c_str(),
                src.total() * (double)bps / (1024.0 * 1024.0 * 1024.0), dc_note);

    cuvhs_cli::TbcWriter writer;
    writer.set_format(fmt);
    if (!writer.open(a.output, a.overwrite)) {
        fprintf(stderr, "Failed to open output: %s\n", a.output.c_str());
        cuvhs_destroy(ctx);
        return 1;
    }

    const size_t chunk_cap = (size_t)a.batch * (size_t)fmt.spfp;
    const size_t tbc = (size_t)fmt.field_lines * fmt.line_len;
    uint8_t* h_chunk = nullptr;
    uint8_t* d_chunk = nullptr;
    CHECK_HIP(hipHostMalloc((void**)&h_chunk, chunk_cap * bps, hipHostMallocDefault));
    CHECK_HIP(hipMalloc((void**)&d_chunk, chunk_cap * bps));
    std::vector<uint16_t> luma(a.batch * tbc), chroma(a.batch * tbc);
    std::vector<uint64_t> file_loc(a.batch);
    std::vector<int32_t> is_first(a.batch), phase(a.batch), doc(a.batch);
    std::vector<int32_t> dol((size_t)a.batch * CUVHS_MAX_DROPOUTS), dos(dol.size()), doe(dol.size());

    const int total_est = src.total_known() ? (int)(src.total() / (uint64_t)fmt.samples_per_field) : 0;
    if (total_est > 0) fprintf(stderr, "Starting decode: ~%d fields in chunks of %d\n\n\n", total_est, a.batch);
    else fprintf(stderr, "Starting decode: streaming, chunks of %d fields\n\n\n", a.batch);
    auto t0 = std::chrono::steady_clock::now();
    uint64_t raw_offset = 0;
    int done = 0;
    bool ok = true;
    const uint64_t spf = (uint64_t)fmt.samples_per_field;
    while (true) {
        // Pipeline::run chunk sizing (pipeline.cu:2651-2658).  A stream that
        // has not ended holds at least batch * spfp samples past raw_offset,
        // so its remaining field count is at least batch, as the file's is.
        long long got = src.read(raw_offset, chunk_cap, h_chunk);
        if (got < 0) { fprintf(stderr, "Input read failed\n"); ok = false; break; }
        if (src.total_known() && raw_offset >= src.total()) break;
        uint64_t remaining = src.total_known() ? (src.total() - raw_offset) / spf : (uint64_t)a.batch;
        if (remaining == 0) break;
        int this_chunk = (int)std::min<uint64_t>((uint64_t)a.batch, remaining);
        size_t want = (size_t)this_chunk * (size_t)fmt.spfp;
        size_t n_read = std::min((size_t)got, want);
        if (n_read == 0) break;
        CHECK_HIP(hipMemcpy(d_chunk, h_chunk, n_read * bps, hipMemcpyHostToDevice));
        cuvhs_chunk_info info;
        rc = cuvhs_decode_chunk(ctx, d_chunk, n_read, raw_offset, this_chunk, &info);
        if (rc < 0) {
            fprintf(stderr, "Decode failed (%d): %s\n", rc, cuvhs_last_error(ctx));
            ok = false;
            break;
        }
        if (info.status & CUVHS_W_VSYNC_OVERFLOW)
            fprintf(stderr, "Warning: VSYNC candidate buffer overflow (%d > %d), truncating chunk results\n",
                    info.vsync_candidates, info.fields_k1 * 20);
        if (info.status & CUVHS_W_NO_VSYNC)
            fprintf(stderr, "Warning: no VSYNC candidates found in chunk at raw offset %llu, using nominal field spacing fallback\n",
                    (unsigned long long)raw_offset);
        if (info.fields == 0) break;
        if (log_k3 || log_k4 || log_k5) {
            // the reference's per-chunk debug lines (pipeline.cu:1874, 2036, 2074)
            int32_t dg[9];
            cuvhs_diagnostics_read(ctx, nullptr, dg);
            if (log_k3)
                fprintf(stderr, "  [K3 spacing] chunk raw_offset=%zu: bad_spacing=%d isolated_spacing=%d large_jump=%d\n",
                        (size_t)raw_offset, dg[0], dg[1], dg[2]);
            if (log_k4)
                fprintf(stderr, "  [K4 debug] chunk raw_offset=%zu: large_delta=%d isolated_jump=%d refined_sync_like=%d\n",
                        (size_t)raw_offset, dg[3], dg[4], dg[5]);
            if (log_k5)
                fprintf(stderr, "  [K5 debug] chunk raw_offset=%zu: mode=%s oob_pixels=%d bad_geom_lines=%d sync_like_pixels=%d\n",
                        (size_t)raw_offset, "refined", dg[6], dg[7], dg[8]);
        }
        const int n = info.fields;
        cuvhs_chunk_outputs o;
        cuvhs_chunk_outputs_get(ctx, &o);
        CHECK_HIP(hipMemcpy(luma.data(), o.luma, n * tbc * 2, hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(chroma.data(), o.chroma, n * tbc * 2, hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(file_loc.data(), o.file_loc, n * 8, hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(is_first.data(), o.is_first, n * 4, hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(phase.data(), o.phase_id, n * 4, hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(doc.data(), o.do_count, n * 4, hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(dol.data(), o.do_line, (size_t)n * CUVHS_MAX_DROPOUTS * 4, hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(dos.data(), o.do_startx, (size_t)n * CUVHS_MAX_DROPOUTS * 4, hipMemcpyDeviceToHost));
        CHECK_HIP(hipMemcpy(doe.data(), o.do_endx, (size_t)n * CUVHS_MAX_DROPOUTS * 4, hipMemcpyDeviceToHost));
        if (!writer.write_fields(luma.data(), chroma.data(), n)) { fprintf(stderr, "TBC write failed\n"); ok = false; break; }
        for (int i = 0; i < n; i++) {
            cuvhs_cli::FieldMeta m;
            m.is_first = is_first[i] != 0;
            m.phase_id = phase[i];
            m.file_loc = file_loc[i];
            const size_t b = (size_t)i * CUVHS_MAX_DROPOUTS;
            for (int k = 0; k < doc[i]; k++) {
                m.do_line.push_back(dol[b + k]);
                m.do_startx.push_back(dos[b + k]);
                m.do_endx.push_back(doe[b + k]);
            }
            writer.add_field(std::move(m));
        }
        writer.write_json();
        done += n;
        raw_offset = info.next_raw_offset;
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (total_est > 0)
            fprintf(stderr, "\033[2A\033[K  %d/%d fields (%.1f%%) | %.1f FPS\n\033[K\n", done, total_est,
                    100.0 * done / total_est, (done / 2.0) / std::max(el, 1e-9));
        else
            fprintf(stderr, "\033[2A\033[K  %d fields | %.1f FPS\n\033[K\n", done, (done / 2.0) / std::max(el, 1e-9));
    }
    double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fprintf(stderr, "\033[2A\033[KDecode complete: %d fields in %.1f seconds (%.1f FPS)\n\033[K\n", done, el,
            (done / 2.0) / std::max(el, 1e-9));
    (void)hipFree(d_chunk);
    (void)hipHostFree(h_chunk);
    cuvhs_destroy(ctx);
    if (!writer.finalize() || !ok) {
        fprintf(stderr, "Pipeline failed\n");
        return 1;
    }
    fprintf(stderr, "Done.\n");
    return 0;
}
