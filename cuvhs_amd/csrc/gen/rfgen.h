// Synthetic VHS RF capture generator, NTSC or PAL (host C ABI).
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cuvhs_gen_params {
    double sample_rate_mhz;     /* capture rate */
    size_t n_samples;           /* capture length in u8 samples */
    uint64_t seed;
    double lead_lines;          /* lines of signal before the first vblank */
    double snr_lo_db, snr_hi_db;/* per-field SNR drawn uniformly (luma carrier power) */
    double chroma_db;           /* colour-under level re luma carrier for 40 IRE chroma */
    double dropouts_per_field;  /* Poisson mean; 0 disables */
    double dropout_len_lo_us, dropout_len_hi_us;
    double wow_max_us;          /* wow/flutter amplitude ~ U(0, max); 0 disables */
    double wow_f_lo_hz, wow_f_hi_hz;
    double headswitch_max_us;   /* head-switch step ~ U(0, max) per field group; 0 disables */
    int headswitch_group_max;   /* group length ~ U{1..max} fields */
    int system;                 /* 0 NTSC, 1 PAL */
} cuvhs_gen_params;

/* config: 0 = 4 fields @28 MHz, 35 dB, seed 0, exactly 4*spf samples;
 *         1 = 600 fields @28 MHz, SNR U(25,40), dropouts, wow, seed 1;
 *         2 = 6000 fields @40 MHz, config-1 model + head-switch steps, seed 2 */
void cuvhs_gen_defaults(cuvhs_gen_params* p, int config);   /* NTSC */
/* the same settings on a system's line structure and levels (0 NTSC, 1 PAL;
 * n_samples counts that system's fields) */
void cuvhs_gen_defaults_system(cuvhs_gen_params* p, int config, int system);
int cuvhs_gen(const cuvhs_gen_params* p, uint8_t* out);

#ifdef __cplusplus
}
#endif
