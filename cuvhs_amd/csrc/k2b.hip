// K2b line location on the GPU.
//
// Reference: the host path pipeline.cu:410-893 (functions) and its driver
// :1694-1791.  The reference runs it serially on the CPU after copying every
// pulse list to the host and copying line locations back.  Here:
//   k_k2b_field   one workgroup per field: pulse refinement with the vblank
//                 state machine, mean line length, vblank pulse scan and the
//                 first-HSYNC anchor for both field parities and both
//                 "no previous anchor" cases -- everything that does not
//                 depend on the previous field.  Data stays in LDS.
//   k_k2b_chain   one thread: the cross-field cadence (pipeline.h:92-97),
//                 O(1) per field except the rare estimate-only anchor.
//   k_k2b_linelocs one workgroup per field: sort pulse starts (bitonic) and
//                 snap the 525 line locations (pipeline.cu:857-893).
// The statement order of every floating-point expression follows the
// reference so results round identically.
#include <hip/hip_runtime.h>

#include "common.h"
#include "launch.h"
#include "wave.h"

namespace cuvhs {

struct VP {            // validated pulse, structure of arrays in LDS/global
    int* start;
    int* len;
    unsigned char* tg;  // type | good << 4
};

struct Timings {
    double hsync_min, hsync_max, eq_min, eq_max, vsync_min, vsync_max;
};

struct AnchorVariant {
    int count;              // valid_location_count (0 = none)
    double loc;             // first_hsync_loc after rounding and hsync offset
    double loc_rounded;     // first_hsync_loc after rounding, before the hsync offset
};

// Per-field summary written by k_k2b_field.
struct K2bField {
    int nvp;                // validated pulses (possibly > 800)
    int npulses_k2;         // K2 count (capped at 800)
    int overflow;
    int id_zero;            // interlaced_detected == 0
    int round_consensus1;   // round(ic / max(1, id)) == 1
    double ic, id;
    double meanlinelen;
    int first_k2_start;
    AnchorVariant var[2][2];  // [first_field][prev_loc_nonpositive]
    int vblank_pulses[8];
};

struct K2bAnchorOut {
    int valid;
    int first_field;
    double first_hsync_loc;
    double first_hsync_line;
    double meanlinelen;
};

__device__ __forceinline__ bool inr(double v, double lo, double hi) { return v >= lo && v <= hi; }

__device__ __forceinline__ bool qualitycheck(int prev_type, int type, int start_delta, int in_line_len) {
    double mn, mx;
    if (prev_type > 0 && type > 0) { mn = 0.4; mx = 0.6; }
    else if (prev_type == 0 && type == 0) { mn = 0.9; mx = 1.1; }
    else { mn = 0.4; mx = 1.1; }
    double span = (double)start_delta / (double)in_line_len;
    return span >= mn && span <= mx;
}

// qualitycheck with the span start_delta / in_line_len already divided out
__device__ __forceinline__ bool quality_span(int prev_type, int type, double span) {
    double mn, mx;
    if (prev_type > 0 && type > 0) { mn = 0.4; mx = 0.6; }
    else if (prev_type == 0 && type == 0) { mn = 0.9; mx = 1.1; }
    else { mn = 0.4; mx = 1.1; }
    return span >= mn && span <= mx;
}

// Per-K2-pulse data of the refinement, computed across the wavefront before
// the sequential passes: width class bits (kClsHs/Eq/Vs) and, from the
// spacing to the previous K2 pulse in lines, (double)(rs[i] - rs[i-1]) /
// line length (the span the quality check takes when pulse i-1 is the last
// validated one), quality_span's result for the three type pairings: both
// types > 0 ([0.4, 0.6]), both HSYNC ([0.9, 1.1]), mixed ([0.4, 1.1]).  The
// sequential passes then branch on bits, not on FP64 compares.
constexpr unsigned char kClsHs = 1, kClsEq = 2, kClsVs = 8;
constexpr unsigned char kQPp = 16, kQHh = 32, kQMx = 64;
__device__ __forceinline__ bool quality_bits(int prev_type, int type, int c) {
    if (prev_type > 0 && type > 0) return (c & kQPp) != 0;
    if (prev_type == 0 && type == 0) return (c & kQHh) != 0;
    return (c & kQMx) != 0;
}
struct PulseInfo {
    const unsigned char* cls;   // class and quality bits
};

__device__ __forceinline__ double round_line_loc(double line_number) {
    return round(0.5 * round(line_number / 0.5) * 10.0) / 10.0;
}

// host_run_vblank_state_machine (pipeline.cu:519-609) appending to vp[n..)
// Returns done; *n_out is the new length (entries past the cap are dropped
// and flagged).  Run by the whole wavefront with uniform values: the pulse
// data are broadcast reads of LDS, lane 0 stores the appended pulses, and
// the last validated pulse's type and start come in (prev_type, prev_start,
// n > 0) and go out (*last_type, *last_start) in registers, so nothing the
// machine stores is read back.  prev_src: the K2 index the last validated
// pulse came from, or -1.
__device__ bool vblank_state_machine(const int* rs, const int* rl, PulseInfo pi, int npc, int start_idx,
                                     int num_pulses, int in_line_len, VP vp, int n, int prev_src, int prev_type,
                                     int prev_start, int* n_out, int* ovf, int* last_tg, int* last_start) {
    const int lane = threadIdx.x & 63;
    const bool writer = lane == 0;
    bool done = false;
    double num_pulses_half = num_pulses / 2.0;
    int vsync_start = -1;
    double state_end = 0.0;
    double state_length = -1.0;
    int state = prev_type;
    int tg_out = -1;
    // the machine reads at most 26 pulses from start_idx: lane k holds
    // pulse start_idx + k, read back by readlane in the loop
    const int kk = start_idx + lane;
    const int w_rs = kk < npc ? rs[kk] : 0, w_rl = kk < npc ? rl[kk] : 0, w_cls = kk < npc ? (int)pi.cls[kk] : 0;
    // state_end (a double) against integer starts: ps > state_end iff
    // ps > floor(state_end), ps < state_end iff ps < ceil(state_end)
    int end_fl = 0, end_cl = 0;
    for (int i = start_idx; i < npc && i < start_idx + 26; i++) {
        const int jj = i - start_idx;
        const int ps = __builtin_amdgcn_readlane(w_rs, jj);
        const int c = __builtin_amdgcn_readlane(w_cls, jj);
        const bool hs = (c & kClsHs) != 0, eq = (c & kClsEq) != 0, vsn = (c & kClsVs) != 0;
        int type = -1;
        bool have = false;
        if (state == -1) {
            if (hs) { type = kPulseHsync; have = true; }
        } else if (state == kPulseHsync) {
            if (hs) { type = kPulseHsync; have = true; }
            else if (eq) { type = kPulseEq1; state_length = num_pulses_half; have = true; }
            else if (vsn) { vsync_start = n - 1; type = kPulseVsync; have = true; }
        } else if (state == kPulseEq1) {
            if (eq) { type = kPulseEq1; have = true; }
            else if (vsn) { vsync_start = n - 1; type = kPulseVsync; state_length = num_pulses_half; have = true; }
            else if (hs) { type = kPulseHsync; have = true; }
        } else if (state == kPulseVsync) {
            if (eq) { type = kPulseEq2; state_length = num_pulses_half; have = true; }
            else if (vsn) { type = kPulseVsync; have = true; }
            else if (ps > end_fl && hs) { type = kPulseHsync; have = true; }
        } else if (state == kPulseEq2) {
            if (eq) { type = kPulseEq2; have = true; }
            else if (hs) { type = kPulseHsync; have = true; done = true; }
        }
        if (have && type != state) {
            if (ps < end_cl) {
                have = false;
            } else if (state_length > 0.0) {
                state_end = (double)ps + ((state_length - 0.1) * (double)in_line_len);
                state_length = -1.0;
                end_fl = (int)floor(state_end);
                end_cl = (int)ceil(state_end);
            }
        }
        if (have) {
            bool good = false;
            if (n > 0)
                good = prev_src == i - 1 ? quality_bits(state, type, c)
                                         : qualitycheck(state, type, ps - prev_start, in_line_len);
            if (n < kVpCap) {
                const int tg = type | (good ? 16 : 0);
                const int len = __builtin_amdgcn_readlane(w_rl, jj);
                if (writer) {
                    vp.start[n] = ps;
                    vp.len[n] = len;
                    vp.tg[n] = (unsigned char)tg;
                }
                n++;
                state = type;
                prev_start = ps;
                prev_src = i;
                tg_out = tg;
            } else {
                *ovf = 1;
            }
        }
        if (done && vsync_start >= 0) break;
    }
    *n_out = n;
    *last_tg = tg_out;
    *last_start = prev_start;
    return done;
}

// Line spacings around the vblanks of a first and of a second field: entry
// and exit of the field's own vblank, then of the next one
// (pipeline.cu:707-717; NTSC and PAL differ).
__device__ __forceinline__ void field_order_lengths(bool pal, double ffl[4], double sfl[4]) {
    const double n_f[4] = {1, 0.5, 0.5, 1}, n_s[4] = {0.5, 1, 1, 0.5};
    const double p_f[4] = {0.5, 0.5, 1, 1}, p_s[4] = {1, 1, 0.5, 0.5};
    for (int i = 0; i < 4; i++) {
        ffl[i] = pal ? p_f[i] : n_f[i];
        sfl[i] = pal ? p_s[i] : n_s[i];
    }
}

// first-HSYNC anchor given the field parity (pipeline.cu:747-845, without the
// estimate-only branch which needs the previous field).  Called by all 64
// lanes of the wave with the same arguments; every lane returns the result.
// The HSYNC offset average runs over the validated pulses in order: each
// lane computes the terms of its pulses, and the sum is taken in pulse order
// from the lanes' values (the reference's sequential rounding).
__device__ AnchorVariant anchor_variant(VP vp, int nvp, const int* vbp, double mll, bool first_field,
                                        bool prev_nonpos, int num_eq, int fl0, int fl1, bool pal) {
    AnchorVariant r;
    r.count = 0;
    r.loc = 0.0;
    r.loc_rounded = 0.0;
    const double TOL = 0.5;
    double ffl[4], sfl[4];
    field_order_lengths(pal, ffl, sfl);
    double* cfl = first_field ? ffl : sfl;
    double vsl = num_eq / 2.0;
    double cur_lines = first_field ? fl0 : fl1;
    double vl[8];
    vl[0] = 0.0 + cfl[0];
    vl[1] = vl[0] + vsl;
    vl[2] = vl[1] + vsl;
    vl[3] = vl[2] + vsl - 0.5;
    double hsl = vl[3] + cfl[1];
    vl[4] = cur_lines + cfl[2];
    vl[5] = vl[4] + vsl;
    vl[6] = vl[5] + vsl;
    vl[7] = vl[6] + vsl - 0.5;
    auto calc = [&](int a, int b, double& loc, int& cnt, double& off) {
        int p1 = vbp[a], p2 = vbp[b];
        if (p1 == -1 || p2 == -1 || mll == 0.0) return;
        double actual = ((double)p1 - (double)p2) / mll;
        double expected = vl[a] - vl[b];
        if (actual < expected + TOL && actual > expected - TOL) {
            off += actual - expected;
            loc += (double)p2 + mll * (hsl - vl[b]);
            cnt += 1;
        }
    };
    double fv_loc = 0.0, fv_off = 0.0, lv_loc = 0.0, lv_off = 0.0;
    int fv_n = 0, lv_n = 0;
    for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++) calc(a, b, fv_loc, fv_n, fv_off);
    for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++) calc(4 + a, 4 + b, lv_loc, lv_n, lv_off);
    double loc = 0.0, off = 0.0;
    int cnt = 0;
    double fe = fv_n ? fv_loc / fv_n : 0.0;
    double le = lv_n ? lv_loc / lv_n : 0.0;
    if (fv_n && lv_n && fe < le + TOL * mll && fe > le - TOL * mll) {
        loc = fv_loc + lv_loc;
        cnt = fv_n + lv_n;
        off = fv_off + lv_off;
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) calc(a, 4 + b, loc, cnt, off);
    } else if (fv_n == 6 || (prev_nonpos && fv_n > lv_n)) {
        loc = fv_loc; cnt = fv_n; off = fv_off;
    } else if (lv_n == 6 || (prev_nonpos && lv_n > fv_n)) {
        loc = lv_loc; cnt = lv_n; off = lv_off;
    }
    if (cnt <= 0) return r;
    off /= cnt;
    loc = round((loc + off) / cnt);
    r.loc_rounded = loc;
    double hoff = 0.0;
    int hc = 0;
    const int lane = threadIdx.x & 63;
    for (int base = 0; base < nvp; base += 64) {
        const int i = base + lane;
        bool brk = false, take = false;
        double term = 0.0;
        if (i < nvp) {
            const int tg = vp.tg[i];
            if ((tg & 15) == kPulseHsync && (tg & 16)) {
                const double ps = (double)vp.start[i];
                const double lineloc = (ps - loc) / mll + hsl;
                const int rl = (int)llround(lineloc);
                brk = rl > cur_lines;
                take = !brk && rl >= hsl;
                term = loc + mll * ((double)rl - hsl) - ps;
            }
        }
        // pulses after the first break are not scanned.  The sum runs in
        // pulse order over the lanes' terms, the pulses not taken adding
        // +0.0 (which leaves every sum unchanged but the sign of a zero one,
        // and loc - (+-0) / hc is the same either way)
        const unsigned long long mb = __ballot(brk);
        unsigned long long mt = __ballot(take);
        if (mb) mt &= (mb & (~mb + 1ull)) - 1ull;
        if (mt) {
            const double tv = ((mt >> lane) & 1ull) ? term : 0.0;
#pragma unroll
            for (int j = 0; j < 64; j++) hoff += lane_value(tv, j);
            hc += __builtin_popcountll(mt);
        }
        if (mb) break;
    }
    if (hc > 0) {
        hoff /= hc;
        loc -= hoff;
    }
    r.count = cnt;
    r.loc = loc;
    return r;
}

// One workgroup (64 threads) per field.
__global__ __launch_bounds__(64) void k_k2b_field(const ChunkCtl* ctl, int nf_override, const int* k2_starts,
                                                  const int* k2_lens, const int* k2_count, double tol_half_us,
                                                  double hsync_width, double eq_width, double vsync_width,
                                                  int spl, int num_eq, int fl0, int fl1, int pal, int* g_vp_start,
                                                  int* g_vp_len, unsigned char* g_vp_tg, K2bField* summ) {
    __shared__ int rs[kMaxPulses], rl[kMaxPulses];
    __shared__ int vs[kVpCap], vlen[kVpCap];
    __shared__ unsigned char vtg[kVpCap];
    __shared__ int s_nvp, s_ovf;
    __shared__ int s_vbp[8];
    __shared__ double s_mll;
    const int f = blockIdx.x;
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    if (f >= nf) return;
    const int npc = min(max(k2_count[f], 0), kMaxPulses);
    for (int i = threadIdx.x; i < npc; i += 64) {
        rs[i] = k2_starts[(size_t)f * kMaxPulses + i];
        rl[i] = k2_lens[(size_t)f * kMaxPulses + i];
    }
    __syncthreads();
    VP vp{vs, vlen, vtg};
    K2bField& S = summ[f];
    // host_get_line_timings (pipeline.cu:450-477): the lengths are integers,
    // so their sum is exact in any order and the wavefront can share it
    const double tol = tol_half_us;
    const double gmin = hsync_width - 3.5 * tol, gmax = hsync_width + 4.0 * tol;
    double psum = 0.0;
    int pcnt = 0;
    for (int i = threadIdx.x; i < npc; i += 64) {
        double len = (double)rl[i];
        if (len >= gmin && len <= gmax) { psum += len; pcnt++; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        psum += __shfl_xor(psum, o, 64);
        pcnt += __shfl_xor(pcnt, o, 64);
    }
    Timings lt;
    {
        const double sum = psum;
        const int cnt = pcnt;
        double est = cnt > 0 ? (sum / cnt) : hsync_width;
        double hoff = est - hsync_width;
        double nom = hsync_width + hoff;
        lt.hsync_min = nom - tol;
        lt.hsync_max = nom + tol;
        lt.eq_min = eq_width + hoff - tol;
        lt.eq_max = eq_width + hoff + tol;
        lt.vsync_min = (vsync_width + hoff) * 0.5;
        lt.vsync_max = vsync_width + hoff + 2.0 * tol;
    }
    // per-pulse width classes and quality bits (PulseInfo), across the wavefront
    __shared__ unsigned char s_cls[kMaxPulses];
    for (int i = threadIdx.x; i < npc; i += 64) {
        const double len = (double)rl[i];
        unsigned char c = 0;
        if (inr(len, lt.hsync_min, lt.hsync_max)) c |= kClsHs;
        if (inr(len, lt.eq_min, lt.eq_max)) c |= kClsEq;
        if (inr(len, lt.vsync_min, lt.vsync_max)) c |= kClsVs;
        const double span = i > 0 ? (double)(rs[i] - rs[i - 1]) / (double)spl : 0.0;
        if (quality_span(1, 1, span)) c |= kQPp;
        if (quality_span(0, 0, span)) c |= kQHh;
        if (quality_span(0, 1, span)) c |= kQMx;
        s_cls[i] = c;
    }
    __syncthreads();
    const PulseInfo pi{s_cls};
    {
        // host_refine_pulses_exactish (pipeline.cu:611-648), run by the whole
        // wavefront with uniform values: pulse i's class, start, length and
        // span come from lane i mod 64 of registers loaded once per 64
        // pulses (readlane, no LDS latency in the loop); lane 0 stores the
        // validated pulses; the last one's type and start stay in registers.
        // prev_src: the K2 index the last validated pulse came from, when the
        // main loop appended it (-1 after the state machine's appends)
        const int lane = threadIdx.x & 63;
        int ovf = 0;
        int n = 0;
        int prev_tg = 0, prev_start = 0, prev_src = -1;
        int w_cls = 0, w_rs = 0, w_rl = 0;
        for (int i = 0; i < npc; i++) {
            const int j = i & 63;
            if (j == 0) {
                const int k = i + lane;
                w_cls = k < npc ? (int)s_cls[k] : 0;
                w_rs = k < npc ? rs[k] : 0;
                w_rl = k < npc ? rl[k] : 0;
            }
            const int c = __builtin_amdgcn_readlane(w_cls, j);
            if (c & kClsHs) {
                const int cs = __builtin_amdgcn_readlane(w_rs, j);
                bool good = false;
                if (n > 0)
                    good = prev_src == i - 1 ? quality_bits(prev_tg & 15, kPulseHsync, c)
                                             : qualitycheck(prev_tg & 15, kPulseHsync, cs - prev_start, spl);
                if (n < kVpCap) {
                    const int tg = kPulseHsync | (good ? 16 : 0);
                    if (lane == 0) {
                        vs[n] = cs;
                        vlen[n] = __builtin_amdgcn_readlane(w_rl, j);
                        vtg[n] = (unsigned char)tg;
                    }
                    n++;
                    prev_tg = tg;
                    prev_start = cs;
                    prev_src = i;
                } else {
                    ovf = 1;
                }
            } else if (i > 2 && (c & kClsEq) && n > 0 && (prev_tg & 15) == kPulseHsync) {
                int n2 = n, ltg = -1, lstart = 0;
                bool done = vblank_state_machine(rs, rl, pi, npc, i - 2, num_eq, spl, vp, n, prev_src, prev_tg & 15,
                                                 prev_start, &n2, &ovf, &ltg, &lstart);
                if (done && n2 > n) {  // keep the appended sequence
                    n = n2;
                    prev_tg = ltg;
                    prev_start = lstart;
                    prev_src = -1;
                }
            }
            // every branch advances by one pulse (pipeline.cu:637-642)
        }
        if (lane == 0) {
            s_nvp = n;
            s_ovf = ovf;
        }
    }
    __syncthreads();
    const int n = s_nvp;
    const int lane = threadIdx.x;
    // host_compute_meanlinelen_from_valid_pulses (pipeline.cu:479-517): the
    // reference keeps the first longest run of consecutive HSYNCs (its
    // "length" counts the pulses after the first) and averages the spacings
    // within +-5 % of nominal over its pulses 1 .. len-1.  Here: the HSYNC
    // flags as a bitmask in LDS, each run start finds its end in the mask,
    // and a wavefront max picks the longest (earliest on ties); the spacings
    // are integers, so their sum is exact in any order.
    __shared__ unsigned long long s_hmask[kVpCap / 64];
    const int nw = (n + 63) / 64;
    for (int w = 0; w < nw; w++) {
        const int i = w * 64 + lane;
        const unsigned long long m = __ballot(i < n && (vtg[i] & 15) == kPulseHsync);
        if (lane == 0) s_hmask[w] = m;
    }
    __syncthreads();
    unsigned long long best = 0;  // (len + 1) << 32 | (0x7fffffff - start); 0 = no run
    for (int w = 0; w < nw; w++) {
        const int i = w * 64 + lane;
        const unsigned long long m = s_hmask[w];
        const bool h = ((m >> lane) & 1ull) != 0;
        const bool prev_h = lane > 0 ? ((m >> (lane - 1)) & 1ull) != 0 : (w > 0 && (s_hmask[w - 1] >> 63) != 0);
        if (i < n && h && !prev_h) {
            // end of this run: the first clear bit after i (or n)
            int e = n;
            unsigned long long rest = ~m & (lane < 63 ? (~0ull << (lane + 1)) : 0ull);
            if (rest) e = w * 64 + __builtin_ctzll(rest);
            else
                for (int w2 = w + 1; w2 < nw; w2++)
                    if (~s_hmask[w2]) { e = w2 * 64 + __builtin_ctzll(~s_hmask[w2]); break; }
            if (e > n) e = n;
            const unsigned long long key =
                ((unsigned long long)(e - 1 - i + 1) << 32) | (unsigned long long)(0x7fffffff - i);
            if (key > best) best = key;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(best, o, 64);
        if (other > best) best = other;
    }
    const int lr_start = best ? 0x7fffffff - (int)(best & 0xffffffffull) : -1;
    const int lr_len = best ? (int)(best >> 32) - 1 : -1;
    const double nominal = (double)spl;
    double lsum = 0.0;
    int lcnt = 0;
    if (lr_start >= 1 && lr_len >= 1) {
        for (int i = lr_start + 1 + lane; i < lr_start + lr_len; i += 64) {
            double ll = (double)(vs[i] - vs[i - 1]);
            double ratio = ll / nominal;
            if (ratio >= 0.95 && ratio <= 1.05) { lsum += ll; lcnt++; }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        lsum += __shfl_xor(lsum, o, 64);
        lcnt += __shfl_xor(lcnt, o, 64);
    }
    const double mll = lcnt == 0 ? nominal : lsum / (double)lcnt;
    // vblank pulse scan (pipeline.cu:675-727).  Pulse i (i >= 1, good) is
    // compared with pulse i-1; the reference switches to the second group at
    // the first good pulse past vs[0] + fl0 * mll and lets the last match of
    // each kind in each group win.  The kinds are mutually exclusive, so each
    // (kind, group) is the last matching index, found by ballots.
    __shared__ int s_jstar;
    __shared__ int s_vb[8];
    __shared__ double s_fo[4];
    if (lane == 0) {
        s_jstar = n;
        for (int k = 0; k < 8; k++) s_vb[k] = -1;
        for (int k = 0; k < 4; k++) s_fo[k] = -1.0;
    }
    __syncthreads();
    const double gthr = vs[0] + fl0 * mll;
    for (int w = 0; w < nw; w++) {
        const int i = w * 64 + lane;
        const bool c = i >= 1 && i < n && (vtg[i] & 16) && vs[i] > gthr;
        const unsigned long long m = __ballot(c);
        if (m) {
            if (lane == 0) s_jstar = w * 64 + __builtin_ctzll(m);
            break;
        }
    }
    __syncthreads();
    const int jstar = s_jstar;
    for (int w = 0; w < nw; w++) {
        const int i = w * 64 + lane;
        int kind = -1;
        if (i >= 1 && i < n && (vtg[i] & 16)) {
            const int lt_ = vtg[i - 1] & 15, ty = vtg[i] & 15;
            if (lt_ == kPulseHsync && ty > kPulseHsync) kind = 0;
            else if (lt_ == kPulseEq1 && ty == kPulseVsync) kind = 1;
            else if (lt_ == kPulseVsync && ty == kPulseEq2) kind = 2;
            else if (lt_ > kPulseHsync && ty == kPulseHsync) kind = 3;
        }
        const int grp = i >= jstar ? 1 : 0;
        for (int k = 0; k < 4; k++)
            for (int g = 0; g < 2; g++) {
                const unsigned long long m = __ballot(kind == k && grp == g);
                if (m && lane == 63 - __builtin_clzll(m)) {
                    s_vb[k + 4 * g] = k == 3 ? vs[i - 1] : vs[i];
                    if (k == 0) s_fo[0 + 2 * g] = round_line_loc((vs[i] - vs[i - 1]) / mll);
                    if (k == 3) s_fo[1 + 2 * g] = round_line_loc((vs[i] - vs[i - 1]) / mll);
                }
            }
    }
    __syncthreads();
    if (lane == 0) {
        double fol[4];
        int vbp[8];
        for (int k = 0; k < 4; k++) fol[k] = s_fo[k];
        for (int k = 0; k < 8; k++) vbp[k] = s_vb[k];
        double ffl[4], sfl[4];
        field_order_lengths(pal != 0, ffl, sfl);
        double ic = 0.0, id = 0.0;
        for (int i = 0; i < 4; i++) {
            if (fol[i] == ffl[i]) { ic += 1.0; id += 1.0; }
            if (fol[i] == sfl[i]) id += 1.0;
        }
        S.nvp = n;
        S.npulses_k2 = npc;
        S.overflow = s_ovf;
        S.ic = ic;
        S.id = id;
        S.id_zero = id == 0.0;
        S.round_consensus1 = round(ic / fmax(1.0, id)) == 1.0;
        S.meanlinelen = mll;
        S.first_k2_start = npc > 0 ? rs[0] : 0;
        for (int k = 0; k < 8; k++) { S.vblank_pulses[k] = vbp[k]; s_vbp[k] = vbp[k]; }
        s_mll = mll;
    }
    __syncthreads();
    for (int var = 0; var < 4; var++) {
        const bool ff = (var & 1) != 0;
        const bool pn = (var & 2) != 0;
        const AnchorVariant av = anchor_variant(vp, n, s_vbp, s_mll, ff, pn, num_eq, fl0, fl1, pal != 0);
        if (threadIdx.x == 0) S.var[ff ? 1 : 0][pn ? 1 : 0] = av;
    }
    for (int i = threadIdx.x; i < n; i += 64) {
        g_vp_start[(size_t)f * kVpCap + i] = vs[i];
        g_vp_len[(size_t)f * kVpCap + i] = vlen[i];
        g_vp_tg[(size_t)f * kVpCap + i] = vtg[i];
    }
}

// Cross-field cadence (pipeline.cu:1717-1772, :729-745, :811-827), one thread.
__global__ void k_k2b_chain(ChunkCtl* ctl, DecodeState* dst, int nf_override, const unsigned long long* offsets,
                            const K2bField* summ, const int* g_vp_start, const unsigned char* g_vp_tg,
                            K2bAnchorOut* anchors, int* is_first, int num_eq, int fl0, int fl1, int pal) {
    // the field summaries and offsets are staged in LDS by the whole
    // workgroup, so the sequential loop below waits on LDS, not on L2
    constexpr int kStage = 64;
    static_assert(sizeof(K2bField) % sizeof(int) == 0, "summaries are copied as words");
    __shared__ K2bField s_summ[kStage];
    __shared__ unsigned long long s_off[kStage];
    if (blockIdx.x != 0) return;
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    const int ns = nf < kStage ? nf : kStage;
    {
        const int* src = reinterpret_cast<const int*>(summ);
        int* dstw = reinterpret_cast<int*>(s_summ);
        const int words = ns * (int)(sizeof(K2bField) / sizeof(int));
        for (int i = threadIdx.x; i < words; i += blockDim.x) dstw[i] = src[i];
        for (int i = threadIdx.x; i < ns; i += blockDim.x) s_off[i] = offsets[i];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    K2bState st = dst->k2b;
    const unsigned long long raw_offset = ctl->raw_offset;
    int ovf = 0;
    for (int f = 0; f < nf; f++) {
        const K2bField& S = f < kStage ? s_summ[f] : summ[f];
        const unsigned long long off_f = f < kStage ? s_off[f] : offsets[f];
        ovf |= S.overflow;
        K2bAnchorOut A;
        A.valid = 0;
        A.first_field = -1;
        A.first_hsync_loc = 0.0;
        A.first_hsync_line = 0.0;
        double mll = S.meanlinelen;
        A.meanlinelen = mll;
        double prev_offset_lines = 0.0;
        if (st.prev_first_hsync_readloc != -1 && mll > 0.0)
            prev_offset_lines = ((double)st.prev_first_hsync_readloc - (double)(raw_offset + off_f)) / mll;
        bool valid = false;
        bool first_field = false;
        double loc = 0.0, hsl = 0.0, new_diff = st.prev_first_hsync_diff;
        if (S.nvp > 0 && mll > 0.0) {
            if (st.prev_first_field == -1) first_field = S.id_zero || S.round_consensus1;
            else first_field = !st.prev_first_field;
            int conf = 75;
            if (S.id > 0.0) {
                if (st.prev_first_hsync_loc < 0.0) conf = min(conf, 50);
                double weighting = S.id / 4.0;
                int ffc = (int)llround((S.ic / S.id) * weighting * 100.0);
                int sfc = (int)llround(((S.id - S.ic) / S.id) * weighting * 100.0);
                if (ffc >= conf && ffc > sfc) first_field = true;
                else if (sfc >= conf && sfc > ffc) first_field = false;
            }
            double ffl[4], sfl[4];
            field_order_lengths(pal != 0, ffl, sfl);
            double* cfl = first_field ? ffl : sfl;
            double vsl = num_eq / 2.0;
            hsl = 0.0 + cfl[0] + vsl + vsl + vsl - 0.5 + cfl[1];
            {   // same association as the reference's chained vblank_lines
                double v0 = 0.0 + cfl[0];
                double v1 = v0 + vsl;
                double v2 = v1 + vsl;
                double v3 = v2 + vsl - 0.5;
                hsl = v3 + cfl[1];
            }
            double cur_lines = first_field ? fl0 : fl1;
            // the estimate steps one field back from the previous anchor:
            // the previous field's lines in NTSC, this field's in PAL
            // (pipeline.cu:811)
            double prev_lines = pal ? cur_lines : (first_field ? fl1 : fl0);
            bool prev_nonpos = st.prev_first_hsync_loc <= 0.0;
            const AnchorVariant& V = S.var[first_field ? 1 : 0][prev_nonpos ? 1 : 0];
            int est_loc = (int)llround((prev_offset_lines + prev_lines + st.prev_first_hsync_loc / mll) * mll);
            if (V.count > 0) {
                valid = true;
                loc = V.loc;
                new_diff = (V.loc_rounded - est_loc) / mll;
            } else if (st.prev_first_hsync_loc > 0.0) {
                // estimate-only anchor (pipeline.cu:814-844)
                double ewo = (st.prev_first_hsync_diff <= 0.5 && st.prev_first_hsync_diff >= -0.5)
                                 ? est_loc + mll * st.prev_first_hsync_diff
                                 : (double)est_loc;
                if (ewo <= 0.0) ewo = (double)g_vp_start[(size_t)f * kVpCap + 0];
                double l = 0.0 + ewo;
                l = round((l + 0.0 / 1) / 1);
                double hoff = 0.0;
                int hc = 0;
                for (int i = 0; i < S.nvp && i < kVpCap; i++) {
                    int tg = g_vp_tg[(size_t)f * kVpCap + i];
                    if ((tg & 15) != kPulseHsync || !(tg & 16)) continue;
                    double ps = (double)g_vp_start[(size_t)f * kVpCap + i];
                    double lineloc = (ps - l) / mll + hsl;
                    int rl = (int)llround(lineloc);
                    if (rl > cur_lines) break;
                    if (rl >= hsl) {
                        hoff += l + mll * ((double)rl - hsl) - ps;
                        hc++;
                    }
                }
                if (hc > 0) {
                    hoff /= hc;
                    l -= hoff;
                }
                valid = true;
                loc = l;
                new_diff = st.prev_first_hsync_diff;
            }
        }
        if (valid) {
            A.valid = 1;
            A.first_field = first_field ? 1 : 0;
            A.first_hsync_loc = loc;
            A.first_hsync_line = hsl;
            is_first[f] = A.first_field;
            st.prev_first_field = A.first_field;
            st.prev_progressive_field = 0;
            st.prev_first_hsync_readloc = (long long)(raw_offset + off_f);
            st.prev_first_hsync_loc = loc;
            st.prev_first_hsync_diff = new_diff;
        } else {
            is_first[f] = (st.prev_first_field >= 0) ? (1 - st.prev_first_field) : 1;
        }
        anchors[f] = A;
    }
    dst->k2b = st;
    ctl->vp_overflow = ovf;
}

// Line locations + pulse outputs, one workgroup (256 threads) per field;
// LPF line locations per frame (525 NTSC, 625 PAL).
template <int LPF>
__global__ __launch_bounds__(256) void k_k2b_linelocs(const ChunkCtl* ctl, int nf_override, const K2bField* summ,
                                                      const K2bAnchorOut* anchors, const int* g_vp_start,
                                                      const int* g_vp_len, const unsigned char* g_vp_tg,
                                                      int* out_starts, int* out_lens, int* out_count,
                                                      int* out_types, double* linelocs) {
    constexpr int lpf = LPF;
    __shared__ int key[kVpCap];
    const int f = blockIdx.x;
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    if (f >= nf) return;
    const K2bField& S = summ[f];
    const K2bAnchorOut A = anchors[f];
    const int n = min(S.nvp, kVpCap);
    // pulse outputs (pipeline.cu:1739-1750)
    const int after = min(n, kMaxPulses);
    for (int i = threadIdx.x; i < kMaxPulses; i += 256) {
        size_t o = (size_t)f * kMaxPulses + i;
        if (i < after) {
            size_t g = (size_t)f * kVpCap + i;
            out_starts[o] = g_vp_start[g];
            out_lens[o] = g_vp_len[g];
            out_types[o] = g_vp_tg[g] & 15;
        } else {
            out_starts[o] = 0;
            out_lens[o] = 0;
            out_types[o] = kPulseHsync;
        }
    }
    if (threadIdx.x == 0) out_count[f] = after;
    double* ll = linelocs + (size_t)f * lpf;
    const double mll = S.meanlinelen;
    if (!A.valid) {
        double ref = (double)S.first_k2_start;
        if (S.npulses_k2 <= 0) ref = 0.0;
        for (int line = threadIdx.x; line < lpf; line += 256) ll[line] = ref + mll * (double)line;
        return;
    }
    // bitonic sort of the starts (pad with INT_MAX), skipped when the list
    // is already ascending (the usual case: pulses are validated in order)
    __shared__ int s_unsorted;
    int P = 1;
    while (P < n) P <<= 1;
    if (threadIdx.x == 0) s_unsorted = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < P; i += 256) {
        key[i] = i < n ? g_vp_start[(size_t)f * kVpCap + i] : 0x7fffffff;
        if (i > 0 && i < n && g_vp_start[(size_t)f * kVpCap + i - 1] > key[i]) s_unsorted = 1;
    }
    __syncthreads();
    for (int k = 2; k <= P && s_unsorted; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += 256) {
                int ixj = i ^ j;
                if (ixj > i) {
                    int a = key[i], b = key[ixj];
                    bool up = (i & k) == 0;
                    if ((a > b) == up) { key[i] = b; key[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
    // host_valid_pulses_to_linelocs_exactish (pipeline.cu:857-893), restated.
    // The reference walks forward from a cursor cur while |key - l| does not
    // increase, keeping the last index walked with distance <= mll/1.5, and
    // never looks at index n-1.  With the keys ascending the distance is
    // unimodal in the index, so the walk from cur ends at g = the last index
    // of the nearest run over the whole list when g >= cur, and otherwise at
    // the end of the run of keys equal to key[cur]; the kept index is that
    // end clamped to n-2, if within tolerance.  Wave 0 finds g for every line
    // by binary search (lane j: lines j + 64 k), then carries the cursor
    // through the frame's lines reading those results from the lanes.
    __shared__ int s_pick[LPF];
    const double max_allowed = mll / 1.5;
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        constexpr int K = (LPF + 63) / 64;
        int g[K];
        int ok[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int line = lane + 64 * k;
            g[k] = -1;
            ok[k] = 0;
            if (line < lpf && n >= 2) {
                const double l = A.first_hsync_loc + mll * ((double)line - A.first_hsync_line);
                int lo = 0, hi = n;  // first index with key >= l
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if ((double)key[mid] < l) lo = mid + 1;
                    else hi = mid;
                }
                int p;
                if (lo == 0) p = 0;
                else if (lo == n) p = n - 1;
                else p = fabs((double)key[lo] - l) <= fabs((double)key[lo - 1] - l) ? lo : lo - 1;
                while (p + 1 < n && key[p + 1] == key[p]) p++;  // last of the nearest run
                g[k] = p;
                const int sl = p < n - 2 ? p : n - 2;
                ok[k] = fabs((double)key[sl] - l) <= max_allowed;
            }
        }
        int cur = 0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            for (int j = 0; j < 64; j++) {
                const int line = 64 * k + j;
                if (line >= lpf) break;
                const int gl = __builtin_amdgcn_readlane(g[k], j);
                int pick = -1;
                if (cur <= n - 2) {
                    if (gl >= cur) {
                        if (__builtin_amdgcn_readlane(ok[k], j)) pick = gl < n - 2 ? gl : n - 2;
                    } else {
                        int e = cur;
                        while (e + 1 < n && key[e + 1] == key[e]) e++;
                        const int sl = e < n - 2 ? e : n - 2;
                        const double l = A.first_hsync_loc + mll * ((double)line - A.first_hsync_line);
                        if (fabs((double)key[sl] - l) <= max_allowed) pick = sl;
                    }
                }
                if (pick >= 0) cur = pick + 1;
                if (lane == 0) s_pick[line] = pick;
            }
        }
    }
    __syncthreads();
    for (int line = threadIdx.x; line < lpf; line += 256) {
        const int pk = s_pick[line];
        ll[line] = pk >= 0 ? (double)key[pk] : A.first_hsync_loc + mll * ((double)line - A.first_hsync_line);
    }
}

// ---- host launchers ----
hipError_t run_k2b_fields(const ChunkCtl* ctl_dev, int nf_override, int max_fields, const Fmt& fmt,
                          const int* k2_starts, const int* k2_lens, const int* k2_count, int* vp_start, int* vp_len,
                          unsigned char* vp_tg, K2bField* summ, hipStream_t st) {
    double tol = fmt.sample_rate * 0.5e-6;
    hipLaunchKernelGGL(k_k2b_field, dim3(max_fields), dim3(64), 0, st, ctl_dev, nf_override, k2_starts, k2_lens,
                       k2_count, tol, fmt.hsync_width, fmt.eq_pulse_width, fmt.vsync_width, fmt.samples_per_line,
                       fmt.num_eq_pulses, fmt.field_lines_first, fmt.field_lines_second, fmt.system == kSystemPal,
                       vp_start, vp_len, vp_tg, summ);
    return hipGetLastError();
}

hipError_t run_k2b_cadence(const ChunkCtl* ctl_dev, ChunkCtl* ctl_mut, DecodeState* dst, int nf_override,
                           int max_fields, const Fmt& fmt, const unsigned long long* offsets, const int* vp_start,
                           const int* vp_len, const unsigned char* vp_tg, const K2bField* summ, K2bAnchorOut* anchors,
                           int* out_starts, int* out_lens, int* out_count, int* out_types, double* linelocs,
                           int* is_first, hipStream_t st) {
    hipLaunchKernelGGL(k_k2b_chain, dim3(1), dim3(64), 0, st, ctl_mut, dst, nf_override, offsets, summ, vp_start, vp_tg,
                       anchors, is_first, fmt.num_eq_pulses, fmt.field_lines_first, fmt.field_lines_second,
                       fmt.system == kSystemPal);
    if (fmt.lines_per_frame == 625)
        hipLaunchKernelGGL(k_k2b_linelocs<625>, dim3(max_fields), dim3(256), 0, st, ctl_dev, nf_override, summ,
                           anchors, vp_start, vp_len, vp_tg, out_starts, out_lens, out_count, out_types, linelocs);
    else if (fmt.lines_per_frame == 525)
        hipLaunchKernelGGL(k_k2b_linelocs<525>, dim3(max_fields), dim3(256), 0, st, ctl_dev, nf_override, summ,
                           anchors, vp_start, vp_len, vp_tg, out_starts, out_lens, out_count, out_types, linelocs);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t run_k2b(const ChunkCtl* ctl_dev, ChunkCtl* ctl_mut, DecodeState* dst, int nf_override, int max_fields, const Fmt& fmt,
                   const unsigned long long* offsets, const int* k2_starts, const int* k2_lens, const int* k2_count,
                   int* vp_start, int* vp_len, unsigned char* vp_tg, K2bField* summ, K2bAnchorOut* anchors,
                   int* out_starts, int* out_lens, int* out_count, int* out_types, double* linelocs, int* is_first,
                   hipStream_t st) {
    hipError_t e = run_k2b_fields(ctl_dev, nf_override, max_fields, fmt, k2_starts, k2_lens, k2_count, vp_start,
                                  vp_len, vp_tg, summ, st);
    if (e != hipSuccess) return e;
    return run_k2b_cadence(ctl_dev, ctl_mut, dst, nf_override, max_fields, fmt, offsets, vp_start, vp_len, vp_tg, summ,
                           anchors, out_starts, out_lens, out_count, out_types, linelocs, is_first, st);
}

size_t k2b_field_summary_bytes() { return sizeof(K2bField); }
size_t k2b_anchor_bytes() { return sizeof(K2bAnchorOut); }

}  // namespace cuvhs
