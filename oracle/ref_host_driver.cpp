// Driver around the reference's own host-only sources (video_format.cpp,
// raw_reader.cpp, tbc_writer.cpp), built by `make ref` into _ref/.  It prints
// the reference's VideoFormat constants, converts u8 samples through
// RawReader::read_at, and writes a .tbc.json through TBCWriter, so tests can
// compare the oracle's and the decoder's restatements with the reference's
// compiled code.  Test infrastructure only.
//
// usage: ref_host format <mhz> [pal]
//        ref_host convert <file> [u8|s16|u16] [dc]   (one converted value per line)
//        ref_host json <out_base> [pal]        (fixed metadata, see below)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "format/video_format.h"
#include "io/raw_reader.h"
#include "io/tbc_writer.h"

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const bool pal = argc > 3 && !strcmp(argv[3], "pal");
    const VideoSystem sys = pal ? VideoSystem::PAL : VideoSystem::NTSC;
    if (!strcmp(argv[1], "format")) {
        VideoFormat f(sys, atof(argv[2]));
        printf("sample_rate=%.17g\nsamples_per_line=%d\nsamples_per_field=%d\n", f.sample_rate, f.samples_per_line,
               f.samples_per_field);
        printf("hsync_width=%.17g\neq_pulse_width=%.17g\nvsync_width=%.17g\n", f.hsync_width, f.eq_pulse_width,
               f.vsync_width);
        printf("ire0=%.17g\nhz_ire=%.17g\nvsync_ire=%.17g\nsync_tip_hz=%.17g\npulse_threshold_hz=%.17g\n", f.ire0,
               f.hz_ire, f.vsync_ire, f.sync_tip_hz, f.pulse_threshold_hz);
        printf("output_zero=%.17g\noutput_scale=%.17g\noutput_rate=%.17g\nfsc=%.17g\nchroma_under=%.17g\n",
               f.output_zero, f.output_scale, f.output_rate, f.fsc, f.chroma_under);
        printf("burst_abs_ref=%.17g\n", f.burst_abs_ref);
        printf("lines_per_frame=%d\noutput_line_len=%d\noutput_field_lines=%d\nnum_eq_pulses=%d\n",
               f.lines_per_frame, f.output_line_len, f.output_field_lines, f.num_eq_pulses);
        printf("field_lines_first=%d\nfield_lines_second=%d\nactive_line_start=%d\n", f.field_lines_first,
               f.field_lines_second, f.active_line_start);
        printf("burst_start_us=%.17g\nburst_end_us=%.17g\n", f.burst_start_us, f.burst_end_us);
        return 0;
    }
    if (!strcmp(argv[1], "convert")) {
        // ref_host convert <file> [u8|s16|u16] [dc]
        InputFormat fmt = InputFormat::U8;
        if (argc > 3 && !strcmp(argv[3], "s16")) fmt = InputFormat::S16;
        if (argc > 3 && !strcmp(argv[3], "u16")) fmt = InputFormat::U16;
        RawReader r;
        InputConditioning cond;
        cond.dc_correct = argc > 4 && !strcmp(argv[4], "dc");
        r.set_conditioning(cond);
        if (!r.open(argv[2], fmt)) return 1;
        size_t n = r.total_samples();
        std::vector<double> v(n);
        size_t got = r.read_at(v.data(), 0, n);
        for (size_t i = 0; i < got; i++) printf("%.17g\n", v[i]);
        return 0;
    }
    if (!strcmp(argv[1], "json")) {
        // three fields: alternating parity from false, phase IDs 1,4,0 (0 = unset),
        // dropouts only on the second field
        TBCWriter w;
        VideoFormat f(sys, 28.0);
        if (!w.open(argv[2], f, true)) return 1;
        std::vector<uint16_t> blank((size_t)f.output_line_len * f.output_field_lines, 1024);
        const int phase[3] = {1, 4, 0};
        const size_t loc[3] = {32032, 499165, 966298};
        for (int i = 0; i < 3; i++) {
            w.set_file_loc(loc[i]);
            w.write_luma_field(blank.data());
            w.write_chroma_field(blank.data());
            w.set_first_field(i % 2 == 1);
            if (phase[i] > 0) w.set_field_phase_id(phase[i]);
            if (i == 1) {
                w.add_dropout(12, 100, 910);
                w.add_dropout(13, 0, 57);
            }
            w.finish_field();
        }
        return w.finalize() ? 0 : 1;
    }
    return 2;
}
