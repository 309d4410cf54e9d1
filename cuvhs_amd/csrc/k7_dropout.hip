// K7 dropout detection, TBC mapping and concealment, plus the per-field
// output metadata (file location, field parity).
//
// Reference: dropout_detect.cu:17-339 (one thread per field scanning the
// whole window serially; 32-lane shuffles for the mean) and
// pipeline.cu:2575-2610 (parity alternation, fileLoc).
//
// Here one 256-thread workgroup handles a field: block means are computed in
// parallel, the hysteresis state machine (start at mean <= 0.18*env, end at
// mean >= 1.25 * that) is resolved with a scan of "last decisive class"
// across threads, and only the short event list is processed serially.
// Concealment keeps the reference's row copy (row l-1, or l+1 for row 0,
// into row l over [startx, endx)) and applies entries in order; each thread
// owns whole columns, so a row written by one entry and read by a later one
// is handled by the same thread in program order.
#include <hip/hip_runtime.h>

#include "common.h"
#include "launch.h"

namespace cuvhs {

constexpr int kDoThreads = 256;
constexpr int kEnvBlock = 16;            // dropout_detect.cu:11
constexpr int kMaxClassBlocks = 53504;   // >= spfp / 16 for captures up to 40 MHz (PAL: 53,280)
constexpr int kMaxEvents = 2048;

__device__ __forceinline__ int find_tbc_line(const double* ll, int als, int n_lines, double rf) {
    if (rf < ll[als] || rf >= ll[als + n_lines]) return -1;
    int lo = 0, hi = n_lines - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (ll[als + mid] <= rf) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Workgroups per field for the envelope sum and the block classes.  Only the
// fields whose scan window is non-empty do work (a dozen or so per chunk, the
// window quirk of dropout_detect.cu:106-114), so each is split finely enough
// to spread over the whole GPU.
constexpr int kSumSlices = 128;
constexpr int kClassSlices = 128;

struct DoWindow {
    int start, end;
};

// scan window of dropout_detect.cu:106-114 (field_start uses the padded stride)
__device__ __forceinline__ DoWindow do_window(const double* linelocs, int f, int spfp, int als, int lpf,
                                              int field_lines) {
    const double* fll = linelocs + (size_t)f * lpf;
    const int field_start = f * spfp;
    DoWindow w;
    w.start = max((int)fll[als], field_start);
    w.end = min((int)fll[als + field_lines], field_start + spfp);
    return w;
}

// Partial sums of the field's envelope over [f*spfp, (f+1)*spfp), only for
// fields with a non-empty scan window (the mean is unused otherwise).
// S: storage type of the envelope (double, or float in the reduced-precision
// forms); sums are FP64.
template <typename S>
__global__ __launch_bounds__(kDoThreads) void k_do_sum(const S* __restrict__ env, const double* __restrict__ linelocs,
                                                        const ChunkCtl* ctl, int nf_override, int spfp, int als,
                                                        int lpf, int field_lines, double* partial) {
    __shared__ double red[kDoThreads];
    const int f = blockIdx.x, sl = blockIdx.y, t = threadIdx.x;
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    if (f >= nf) return;
    DoWindow w = do_window(linelocs, f, spfp, als, lpf, field_lines);
    if (w.end <= w.start) return;
    const int L = (spfp + kSumSlices - 1) / kSumSlices;
    const int a = sl * L, b = min(spfp, a + L);
    const S* fe = env + (size_t)f * spfp;
    double s = 0.0;
    for (int i = a + t; i < b; i += kDoThreads) s += (double)fe[i];
    red[t] = s;
    __syncthreads();
    for (int o = kDoThreads / 2; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) partial[f * kSumSlices + sl] = red[0];
}

__device__ __forceinline__ double field_mean(const double* partial, int f, int spfp) {
    double s = 0.0;
    for (int k = 0; k < kSumSlices; k++) s += partial[f * kSumSlices + k];
    return s / (double)spfp;
}

// 16-sample block means vs the hysteresis thresholds: 1 = <= down, 2 = >= up.
template <typename S>
__global__ __launch_bounds__(kDoThreads) void k_do_class(const S* __restrict__ env, const double* __restrict__ linelocs,
                                                          const ChunkCtl* ctl, int nf_override, int spfp, int als,
                                                          int lpf, int field_lines, const double* partial,
                                                          unsigned char* cls) {
    const int f = blockIdx.x, sl = blockIdx.y;
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    if (f >= nf) return;
    DoWindow w = do_window(linelocs, f, spfp, als, lpf, field_lines);
    if (w.end <= w.start) return;
    __shared__ double s_mean;
    if (threadIdx.x == 0) s_mean = field_mean(partial, f, spfp);
    __syncthreads();
    const double down = s_mean * 0.18;
    const double up = down * 1.25;
    const int nblk = min((w.end - w.start + kEnvBlock - 1) / kEnvBlock, kMaxClassBlocks);
    const int per = (nblk + kClassSlices - 1) / kClassSlices;
    const int b0 = sl * per, b1 = min(nblk, b0 + per);
    for (int b = b0 + threadIdx.x; b < b1; b += kDoThreads) {
        int bs = w.start + b * kEnvBlock;
        int be = min(bs + kEnvBlock, w.end);
        double bsum = 0.0;
        if (be - bs == kEnvBlock) {
            // full block: all loads issued together, summed in sample order
            double e[kEnvBlock];
#pragma unroll
            for (int i = 0; i < kEnvBlock; i++) e[i] = (double)env[bs + i];
#pragma unroll
            for (int i = 0; i < kEnvBlock; i++) bsum += e[i];
        } else {
            for (int i = bs; i < be; i++) bsum += (double)env[i];
        }
        double bm = bsum / (double)(be - bs);
        cls[(size_t)f * kMaxClassBlocks + b] = bm <= down ? 1 : (bm >= up ? 2 : 0);
    }
}

__global__ __launch_bounds__(kDoThreads) void k_dropouts(const unsigned char* __restrict__ gcls,
                                                          const double* __restrict__ linelocs,
                                                          const ChunkCtl* ctl, int nf_override, int spfp, int als,
                                                          int lpf, int field_lines, int line_len, uint16_t* luma,
                                                          uint16_t* chroma, int* do_lines, int* do_starts,
                                                          int* do_ends, int* do_count, const ChunkPlan* plan) {
    __shared__ unsigned char cls[kMaxClassBlocks];
    if (plan) {   // a replayed graph: the chunk's place in the output arrays
        luma = plan->luma;
        chroma = plan->chroma;
        do_lines = plan->do_line;
        do_starts = plan->do_startx;
        do_ends = plan->do_endx;
        do_count = plan->do_count;
    }
    __shared__ int ev[kMaxEvents];
    __shared__ int s_last[kDoThreads];
    __shared__ int s_cnt[kDoThreads];
    __shared__ int s_ent[3 * kMaxDropouts];
    __shared__ int s_nent;
    __shared__ int s_nev;
    const int f = blockIdx.x;
    const int nf = nf_override > 0 ? nf_override : ctl->nf;
    if (f >= nf) return;
    const int t = threadIdx.x;
    const double* fll = linelocs + (size_t)f * lpf;
    const int field_start = f * spfp;
    const int scan_start = max((int)fll[als], field_start);
    const int scan_end = min((int)fll[als + field_lines], field_start + spfp);
    if (t == 0) s_nent = 0;
    if (scan_end > scan_start) {
        const int nblk = (scan_end - scan_start + kEnvBlock - 1) / kEnvBlock;
        const int nb = min(nblk, kMaxClassBlocks);
        // classes from k_do_class: 0 neutral, 1 down, 2 up
        for (int b = t; b < nb; b += kDoThreads) cls[b] = gcls[(size_t)f * kMaxClassBlocks + b];
        __syncthreads();
        const int per = (nb + kDoThreads - 1) / kDoThreads;
        const int b0 = t * per, b1 = min(nb, b0 + per);
        int last = 0;
        for (int b = b0; b < b1; b++)
            if (cls[b]) last = cls[b];
        s_last[t] = last;
        __syncthreads();
        if (t == 0) {
            int state = 2;  // not in a dropout
            for (int k = 0; k < kDoThreads; k++) {
                int l = s_last[k];
                s_last[k] = state;  // incoming state of thread k
                if (l) state = l;
            }
        }
        __syncthreads();
        int state = s_last[t];
        int cnt = 0;
        for (int b = b0; b < b1; b++) {
            int c = cls[b];
            if ((c == 1 && state == 2) || (c == 2 && state == 1)) cnt++;
            if (c) state = c;
        }
        s_cnt[t] = cnt;
        __syncthreads();
        if (t == 0) {
            int acc = 0;
            for (int k = 0; k < kDoThreads; k++) { int c = s_cnt[k]; s_cnt[k] = acc; acc += c; }
            s_nev = acc;
        }
        __syncthreads();
        const int nev = s_nev;
        if (nev <= kMaxEvents) {
            int pos = s_cnt[t];
            state = s_last[t];
            for (int b = b0; b < b1; b++) {
                int c = cls[b];
                if ((c == 1 && state == 2) || (c == 2 && state == 1)) ev[pos++] = b;
                if (c) state = c;
            }
        }
        __syncthreads();
        if (t == 0) {
            // serial replay of dropout_detect.cu:121-182 on the event list
            int rs[kMaxRfDropouts], re[kMaxRfDropouts];
            int n = 0;
            int rf_start = -1;
            auto close_at = [&](int end_pos) {
                if (n > 0 && rf_start - re[n - 1] <= 30) re[n - 1] = end_pos;
                else if (n < kMaxRfDropouts) { rs[n] = rf_start; re[n] = end_pos; n++; }
                rf_start = -1;
            };
            if (nev <= kMaxEvents) {
                for (int k = 0; k < nev; k++) {
                    int bpos = scan_start + ev[k] * kEnvBlock;
                    if (rf_start < 0) rf_start = bpos;  // start event
                    else close_at(bpos);                // end event
                }
            } else {
                int st = 2;
                for (int b = 0; b < nb; b++) {
                    int c = cls[b];
                    int bpos = scan_start + b * kEnvBlock;
                    if (c == 1 && st == 2) rf_start = bpos;
                    else if (c == 2 && st == 1) close_at(bpos);
                    if (c) st = c;
                }
            }
            if (rf_start >= 0) close_at(scan_end);
            int m = 0;
            for (int i = 0; i < n; i++)
                if (re[i] - rs[i] >= 10) { rs[m] = rs[i]; re[m] = re[i]; m++; }
            int tc = 0;
            for (int i = 0; i < m && tc < kMaxDropouts; i++) {
                double srf = (double)rs[i], erf = (double)re[i];
                int sl = find_tbc_line(fll, als, field_lines, srf);
                if (sl < 0) continue;
                int el = find_tbc_line(fll, als, field_lines, erf - 1.0);
                if (el < 0) el = field_lines - 1;
                for (int l = sl; l <= el && tc < kMaxDropouts; l++) {
                    double lls = fll[als + l], lle = fll[als + l + 1];
                    double len = lle - lls;
                    if (len <= 0.0) continue;
                    int sx, ex;
                    if (l == sl) { sx = (int)((srf - lls) / len * (double)line_len); sx = max(0, sx); }
                    else sx = 0;
                    if (l == el) { ex = (int)((erf - lls) / len * (double)line_len + 0.999); ex = min(line_len, ex); }
                    else ex = line_len;
                    if (ex > sx) {
                        s_ent[3 * tc] = l;
                        s_ent[3 * tc + 1] = sx;
                        s_ent[3 * tc + 2] = ex;
                        tc++;
                    }
                }
            }
            s_nent = tc;
        }
    }
    __syncthreads();
    const int ne = s_nent;
    for (int i = t; i < ne; i += kDoThreads) {
        size_t o = (size_t)f * kMaxDropouts + i;
        do_lines[o] = s_ent[3 * i];
        do_starts[o] = s_ent[3 * i + 1];
        do_ends[o] = s_ent[3 * i + 2];
    }
    if (t == 0) do_count[f] = ne;
    // concealment, entries in order: column col belongs to thread
    // col % kDoThreads in every entry, so a row written by one entry and read
    // by a later one (row l concealed, then row l + 1 copied from it) is
    // handled by one thread in program order.  Starting each entry at sx + t
    // instead would hand a column to different threads in entries with
    // different sx, a race with no barrier between the entries.
    uint16_t* lu = luma + (size_t)f * field_lines * line_len;
    uint16_t* ch = chroma + (size_t)f * field_lines * line_len;
    for (int d = 0; d < ne; d++) {
        int line = s_ent[3 * d], sx = s_ent[3 * d + 1], ex = s_ent[3 * d + 2];
        int src = (line > 0) ? line - 1 : line + 1;
        if (src >= field_lines) continue;
        for (int col = t; col < ex; col += kDoThreads) {
            if (col < sx) continue;
            lu[line * line_len + col] = lu[src * line_len + col];
            ch[line * line_len + col] = ch[src * line_len + col];
        }
    }
}

// Output metadata: fileLoc and the alternating field parity
// (pipeline.cu:2577-2593), phase IDs passed through.
__global__ void k_field_meta(const ChunkCtl* ctl, DecodeState* dst, const unsigned long long* offsets,
                             const int* is_first_k2b, unsigned long long* file_loc, int* is_first_out,
                             const ChunkPlan* plan) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (plan) {
        file_loc = plan->file_loc;
        is_first_out = plan->is_first;
    }
    const int nf = ctl->nf;
    int last = dst->last_parity;
    for (int i = 0; i < nf; i++) {
        file_loc[i] = ctl->raw_offset + offsets[i];
        int h = is_first_k2b[i];
        int first;
        if (h >= 0 && (h == 1) != (last != 0)) first = (h == 1);
        else first = !last;
        last = first;
        is_first_out[i] = first;
    }
    dst->last_parity = last;
}

size_t dropout_class_bytes(int max_fields) { return (size_t)max_fields * kMaxClassBlocks; }
size_t dropout_partial_count(int max_fields) { return (size_t)max_fields * kSumSlices; }

hipError_t run_dropout_classes(const void* env, int f32, const double* ll, const ChunkCtl* ctl, int nf_override,
                               int max_fields, const Fmt& fmt, double* partial, unsigned char* cls, hipStream_t st) {
    if ((fmt.spfp + kEnvBlock - 1) / kEnvBlock > kMaxClassBlocks) return hipErrorInvalidValue;
    const dim3 gs(max_fields, kSumSlices), gc(max_fields, kClassSlices);
    if (f32) {
        const float* e = static_cast<const float*>(env);
        hipLaunchKernelGGL(k_do_sum<float>, gs, dim3(kDoThreads), 0, st, e, ll, ctl, nf_override, fmt.spfp,
                           fmt.active_line_start, fmt.lines_per_frame, fmt.field_lines, partial);
        hipLaunchKernelGGL(k_do_class<float>, gc, dim3(kDoThreads), 0, st, e, ll, ctl, nf_override, fmt.spfp,
                           fmt.active_line_start, fmt.lines_per_frame, fmt.field_lines, partial, cls);
    } else {
        const double* e = static_cast<const double*>(env);
        hipLaunchKernelGGL(k_do_sum<double>, gs, dim3(kDoThreads), 0, st, e, ll, ctl, nf_override, fmt.spfp,
                           fmt.active_line_start, fmt.lines_per_frame, fmt.field_lines, partial);
        hipLaunchKernelGGL(k_do_class<double>, gc, dim3(kDoThreads), 0, st, e, ll, ctl, nf_override, fmt.spfp,
                           fmt.active_line_start, fmt.lines_per_frame, fmt.field_lines, partial, cls);
    }
    return hipGetLastError();
}

hipError_t run_dropout_scan(const unsigned char* cls, const double* ll, const ChunkCtl* ctl, int nf_override,
                            int max_fields, const Fmt& fmt, uint16_t* luma, uint16_t* chroma, int* lines, int* starts,
                            int* ends, int* count, hipStream_t st, const ChunkPlan* plan) {
    hipLaunchKernelGGL(k_dropouts, dim3(max_fields), dim3(kDoThreads), 0, st, cls, ll, ctl, nf_override, fmt.spfp,
                       fmt.active_line_start, fmt.lines_per_frame, fmt.field_lines, fmt.line_len, luma, chroma, lines,
                       starts, ends, count, plan);
    return hipGetLastError();
}

hipError_t run_dropouts(const double* env, const double* ll, const ChunkCtl* ctl, int nf_override, int max_fields,
                        const Fmt& fmt, double* partial, unsigned char* cls, uint16_t* luma, uint16_t* chroma,
                        int* lines, int* starts, int* ends, int* count, hipStream_t st) {
    hipError_t e = run_dropout_classes(env, 0, ll, ctl, nf_override, max_fields, fmt, partial, cls, st);
    if (e != hipSuccess) return e;
    return run_dropout_scan(cls, ll, ctl, nf_override, max_fields, fmt, luma, chroma, lines, starts, ends, count, st);
}

hipError_t run_field_meta(const ChunkCtl* ctl, DecodeState* dst, const unsigned long long* offsets,
                          const int* is_first_k2b, unsigned long long* file_loc, int* is_first_out, hipStream_t st,
                          const ChunkPlan* plan) {
    hipLaunchKernelGGL(k_field_meta, dim3(1), dim3(64), 0, st, ctl, dst, offsets, is_first_k2b, file_loc,
                       is_first_out, plan);
    return hipGetLastError();
}

}  // namespace cuvhs
