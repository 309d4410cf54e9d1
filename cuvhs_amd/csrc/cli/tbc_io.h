// ld-tools output files: raw u16 luma/chroma TBC streams and the .tbc.json
// metadata (reference: src/io/tbc_writer.{h,cpp}).
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../../include/cuvhs.h"

namespace cuvhs_cli {

struct FieldMeta {
    bool is_first = false;
    int phase_id = 0;
    uint64_t file_loc = 0;
    std::vector<int> do_line, do_startx, do_endx;
};

class TbcWriter {
  public:
    ~TbcWriter();
    // The decoder's format (system, field geometry, levels); NTSC until set.
    void set_format(const cuvhs_format& f) { fmt_ = f; has_fmt_ = true; }
    // Refuses existing outputs unless overwrite (tbc_writer.cpp:16-31).
    bool open(const std::string& base, bool overwrite);
    bool write_fields(const uint16_t* luma, const uint16_t* chroma, int nfields);
    void add_field(FieldMeta m) { meta_.push_back(std::move(m)); }
    // Rewritten after every chunk through a temporary file and rename.
    bool write_json() const;
    bool finalize();
    int fields() const { return (int)meta_.size(); }

  private:
    cuvhs_format fmt_{};
    bool has_fmt_ = false;
    FILE* luma_ = nullptr;
    FILE* chroma_ = nullptr;
    std::string json_;
    std::vector<FieldMeta> meta_;
};

}  // namespace cuvhs_cli
