// ld-tools TBC/JSON writer; key order, indentation and number formats follow
// the reference's TBCWriter::write_json (tbc_writer.cpp:89-175).
#include "tbc_io.h"

#include <sys/stat.h>

#include <cstdio>

namespace cuvhs_cli {

namespace {
bool exists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}
}  // namespace

TbcWriter::~TbcWriter() {
    if (luma_) fclose(luma_);
    if (chroma_) fclose(chroma_);
}

bool TbcWriter::open(const std::string& base, bool overwrite) {
    std::string l = base + ".tbc", c = base + "_chroma.tbc";
    json_ = base + ".tbc.json";
    if (!overwrite) {
        for (const std::string* p : {&l, &c, &json_}) {
            if (exists(*p)) {
                fprintf(stderr, "Output exists: %s (use --overwrite)\n", p->c_str());
                return false;
            }
        }
    }
    luma_ = fopen(l.c_str(), "wb");
    if (!luma_) { perror(l.c_str()); return false; }
    chroma_ = fopen(c.c_str(), "wb");
    if (!chroma_) { perror(c.c_str()); return false; }
    return true;
}

bool TbcWriter::write_fields(const uint16_t* luma, const uint16_t* chroma, int n) {
    cuvhs_format f = fmt_;
    if (!has_fmt_) cuvhs_format_for(28.0, &f);
    size_t cnt = (size_t)n * f.line_len * f.field_lines;
    return fwrite(luma, sizeof(uint16_t), cnt, luma_) == cnt && fwrite(chroma, sizeof(uint16_t), cnt, chroma_) == cnt;
}

bool TbcWriter::write_json() const {
    std::string tmp = json_ + ".tmp";
    FILE* fp = fopen(tmp.c_str(), "w");
    if (!fp) { perror(tmp.c_str()); return false; }
    // the levels, rates and geometry come from the decoder's format (the
    // sample rate does not enter them); the active-video columns are the
    // reference's per-system constants (tbc_writer.cpp:106-107)
    cuvhs_format f = fmt_;
    if (!has_fmt_) cuvhs_format_for(28.0, &f);
    const bool pal = f.system == CUVHS_SYSTEM_PAL;
    const int n = (int)meta_.size();
    fprintf(fp, "{\n  \"videoParameters\": {\n");
    fprintf(fp, "    \"system\": \"%s\",\n", pal ? "PAL" : "NTSC");
    fprintf(fp, "    \"isSubcarrierLocked\": false,\n");
    fprintf(fp, "    \"isSourcePal\": %s,\n", pal ? "true" : "false");
    fprintf(fp, "    \"numberOfSequentialFields\": %d,\n", n);
    fprintf(fp, "    \"black16bIre\": %.1f,\n", f.black16b);
    fprintf(fp, "    \"white16bIre\": %.1f,\n", f.white16b);
    fprintf(fp, "    \"sampleRate\": %.0f,\n", f.output_rate);
    fprintf(fp, "    \"fieldWidth\": %d,\n", f.line_len);
    fprintf(fp, "    \"fieldHeight\": %d,\n", f.field_lines);
    fprintf(fp, "    \"colourBurstStart\": %d,\n", f.burst_start);
    fprintf(fp, "    \"colourBurstEnd\": %d,\n", f.burst_end);
    fprintf(fp, "    \"activeVideoStart\": %d,\n", pal ? 185 : 134);
    fprintf(fp, "    \"activeVideoEnd\": %d,\n", pal ? 1107 : 894);
    fprintf(fp, "    \"tapeFormat\": \"VHS\",\n");
    fprintf(fp, "    \"isMapped\": false\n");
    fprintf(fp, "  },\n  \"fields\": [\n");
    for (int i = 0; i < n; i++) {
        const FieldMeta& m = meta_[(size_t)i];
        fprintf(fp, "    {\n");
        fprintf(fp, "      \"isFirstField\": %s,\n", m.is_first ? "true" : "false");
        fprintf(fp, "      \"seqNo\": %d,\n", i + 1);
        fprintf(fp, "      \"fileLoc\": %llu,\n", (unsigned long long)m.file_loc);
        if (m.phase_id > 0) fprintf(fp, "      \"fieldPhaseID\": %d,\n", m.phase_id);
        fprintf(fp, "      \"dropOuts\": {\n        \"fieldLine\": [");
        for (size_t j = 0; j < m.do_line.size(); j++) fprintf(fp, j ? ", %d" : "%d", m.do_line[j]);
        fprintf(fp, "],\n        \"startx\": [");
        for (size_t j = 0; j < m.do_startx.size(); j++) fprintf(fp, j ? ", %d" : "%d", m.do_startx[j]);
        fprintf(fp, "],\n        \"endx\": [");
        for (size_t j = 0; j < m.do_endx.size(); j++) fprintf(fp, j ? ", %d" : "%d", m.do_endx[j]);
        fprintf(fp, "]\n      }\n");
        fprintf(fp, "    }%s\n", (i < n - 1) ? "," : "");
    }
    fprintf(fp, "  ]\n}\n");
    fclose(fp);
    if (rename(tmp.c_str(), json_.c_str()) != 0) { perror("rename .tbc.json"); return false; }
    return true;
}

bool TbcWriter::finalize() {
    if (luma_) fflush(luma_);
    if (chroma_) fflush(chroma_);
    if (!write_json()) return false;
    fprintf(stderr, "\nWrote %s (%d fields)\n", json_.c_str(), (int)meta_.size());
    return true;
}

}  // namespace cuvhs_cli
