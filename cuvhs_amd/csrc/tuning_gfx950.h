// Block sizes of the overlap-save convolutions measured fastest on gfx950
// (MI355X) by cuvhs_amd/autotune.py, which writes this file: for each sample
// rate and precision form, every block size each stage's filter fits was
// timed in a decode with the stages serialised, and the fastest kept.
// cuvhs_create takes a row's sizes when its rate is within 0.05 MHz of the
// configuration's and the precision matches; other configurations use the
// cost model of design.cpp.  k1a is unused by the mixed forms.
// Measured 2026-12-04 on AMD Radeon Graphics; timings: profiles/autotune_gfx950.json.
#pragma once

namespace cuvhs {

struct TunedBlocks {
    double mhz;
    int precision;   // CUVHS_PRECISION_*
    int k1a, k1b, pre;
};

constexpr TunedBlocks kTunedBlocks[] = {
    {28.0, 0, 4096, 4096, 16384},   // fp64
    {28.0, 1, 4096, 4096, 16384},   // fp32
    {28.0, 2, 0, 4096, 16384},   // mixed-fp16
    {28.0, 3, 0, 4096, 16384},   // mixed-bf16
    {35.8, 0, 4096, 8192, 16384},   // fp64
    {35.8, 1, 4096, 8192, 16384},   // fp32
    {35.8, 2, 0, 8192, 16384},   // mixed-fp16
    {35.8, 3, 0, 8192, 16384},   // mixed-bf16
    {40.0, 0, 4096, 8192, 16384},   // fp64
    {40.0, 1, 4096, 8192, 16384},   // fp32
    {40.0, 2, 0, 8192, 16384},   // mixed-fp16
    {40.0, 3, 0, 8192, 16384},   // mixed-bf16
};

}  // namespace cuvhs
