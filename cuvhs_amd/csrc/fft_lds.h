// Convolution blocks in LDS for gfx950 (wave64): y = IFFT(FFT(x) G) over
// one block, with the 16 complex values a thread holds in registers.
//
// * conv4096: 4096 = 16 x 256 on 256 threads (4 waves): a radix-16 stage
//   across the workgroup on the values each thread loaded itself, one
//   cross-wave exchange, then per-wave 256-point transforms (16 x 16) with
//   wave-local transposes; the spectrum is applied in that transposed bin
//   order, and the inverse retraces the steps.
// * conv_rows<R>: 16 R points on R threads (R = 512 or 1024), the same
//   scheme with 16 rows of R points transformed as 16 x 16 x R/256 (a wave
//   owns 64 / (R / 16) rows, so only the first and last exchange cross
//   waves).
// * filter1024_wave: 1024 = 16 x 64 on one wavefront, a per-bin function
//   applied in the middle (the chroma line band-pass).
//
// Exchanges go through LDS either as whole complex elements or as real and
// then imaginary parts (half the LDS for two more barriers); the LDS layout
// is padded by one element per 16 (element p lives at p + p/16), which keeps
// the strided exchange patterns off shared banks.  Twiddles come from a table
// of W_B^m = exp(-2 pi i m / B) in global memory: four loads per run (w1, w2,
// w4, w8), the other powers by complex products.
#pragma once
#include <hip/hip_runtime.h>

#include "wave.h"

namespace cuvhs {

// padded LDS index, and the LDS elements a B-point buffer needs
__device__ __forceinline__ int pidx(int p) { return p + (p >> 4); }
constexpr int lds_elems(int B) { return B + B / 16; }

template <typename T> struct C2;
template <> struct C2<double> { using type = double2; };
template <> struct C2<float> { using type = float2; };

// Complex product, contracted to FMAs: the transforms are compared with the
// oracle by tolerance, not bit for bit, and FMAs cut the twiddle cost.
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
#pragma clang fp contract(fast)
    V r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// multiply by sign*i  (sign = -1: -i, +1: +i)
template <int S, typename V> __device__ __forceinline__ V mul_si(V a) {
    V r;
    if (S < 0) { r.x = a.y; r.y = -a.x; } else { r.x = -a.y; r.y = a.x; }
    return r;
}
template <typename V> __device__ __forceinline__ V cconj(V a) { V r; r.x = a.x; r.y = -a.y; return r; }

template <int S, typename V> __device__ __forceinline__ void dft2(V& a, V& b) {
    V t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int S, typename V> __device__ __forceinline__ void dft4(V& a0, V& a1, V& a2, V& a3) {
    V t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_si<S>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

// In-place DFT16 of v[0..15] (natural order in and out).
template <int S, typename T> __device__ __forceinline__ void dft16(typename C2<T>::type* v) {
    using V = typename C2<T>::type;
    // cos/sin(2 pi k / 16), k = 1, 2, 3
    const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173;
    const T c2 = (T)0.70710678118654752440;
    // step 1: DFT4 over n1 for each n2 (elements n2, n2+4, n2+8, n2+12)
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) dft4<S>(v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]);
    // now v[n2 + 4*k1] = a[n2][k1]; step 2: twiddle by W16^(n2*k1)
    auto tw = [&](V a, int e) -> V {
        // W16^e with sign S, e in [0, 9]
        T c, s;
        switch (e) {
            case 0: return a;
            case 1: c = c1; s = s1; break;
            case 2: c = c2; s = c2; break;
            case 3: c = s1; s = c1; break;
            case 4: return mul_si<S>(a);
            case 6: c = -c2; s = c2; break;
            case 9: c = -c1; s = -s1; break;
            default: c = 1; s = 0; break;
        }
        V w;
        w.x = c;
        w.y = (S < 0) ? -s : s;
        return cmul(a, w);
    };
#pragma unroll
    for (int n2 = 1; n2 < 4; n2++)
#pragma unroll
        for (int k1 = 1; k1 < 4; k1++) v[n2 + 4 * k1] = tw(v[n2 + 4 * k1], n2 * k1);
    // step 3: DFT4 over n2 for each k1 -> X[k1 + 4*k2]
    V out[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        V a0 = v[0 + 4 * k1], a1 = v[1 + 4 * k1], a2 = v[2 + 4 * k1], a3 = v[3 + 4 * k1];
        dft4<S>(a0, a1, a2, a3);
        out[k1 + 0] = a0;
        out[k1 + 4] = a1;
        out[k1 + 8] = a2;
        out[k1 + 12] = a3;
    }
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = out[i];
}

// ---- 4096-point convolution block with wave-local sub-transforms ----
//
// y = IFFT(FFT(x) * G) for B = 4096 = 16 x 256 on 256 threads (4 waves):
//   X[k1 + 16 k2] = sum_n2 W256^(n2 k2) W4096^(n2 k1) sum_n1 x[256 n1 + n2] W16^(n1 k1)
// Stage 1: thread t = n2 does the DFT16 over n1 on values it already holds
// and the W4096^(t k1) twiddles, then stores row k1, column t.  One block
// barrier.  Wave w then owns rows k1 = 4w..4w+3 (1024 contiguous elements)
// and does their four 256-point transforms as 16 x 16 (lane (j, q): row
// 4w + j, column class q) with a transpose through its own rows, needing only
// wavefront-scope ordering.  The spectrum product happens there, in the
// transposed bin order k = k1 + 16 c + 256 d (G is stored permuted so the
// loads are coalesced: Gp[1024 w + 64 d + lane], see conv4096_gperm), and the
// inverse retraces the same steps.  Two block barriers per block instead of
// two per Stockham stage.


// WB^(e0 * {1, 2, 4, 8}) from the table of WB^m, m < B (B a power of two).
template <typename T, int B = 4096>
__device__ __forceinline__ void twiddle_base(typename C2<T>::type* wb, const typename C2<T>::type* tw, int e0) {
#pragma unroll
    for (int i = 0; i < 4; i++) wb[i] = tw[(e0 << i) & (B - 1)];
}

// v[k] *= WB^(S * e0 * k), k = 1..15 (S = -1 forward, +1 inverse), from
// the four base powers of twiddle_base and eleven products.
template <int S, typename T>
__device__ __forceinline__ void twiddle_run(typename C2<T>::type* v, const typename C2<T>::type* wb) {
    using V = typename C2<T>::type;
    V w[16];
    w[1] = wb[0];
    w[2] = wb[1];
    w[4] = wb[2];
    w[8] = wb[3];
    if (S > 0) {
        w[1] = cconj(w[1]);
        w[2] = cconj(w[2]);
        w[4] = cconj(w[4]);
        w[8] = cconj(w[8]);
    }
    // each power is applied as it is formed, so at most eight are live
    w[3] = cmul(w[1], w[2]);
    w[5] = cmul(w[4], w[1]);
    w[6] = cmul(w[4], w[2]);
    w[7] = cmul(w[4], w[3]);
#pragma unroll
    for (int r = 1; r < 8; r++) v[r] = cmul(v[r], w[r]);
    v[8] = cmul(v[8], w[8]);
#pragma unroll
    for (int r = 9; r < 16; r++) v[r] = cmul(v[r], cmul(w[8], w[r - 8]));
}

// Host-side bin permutation of a natural-order spectrum for conv4096.
inline int conv4096_gperm(int k) {
    int k1 = k & 15, c = (k >> 4) & 15, d = k >> 8;
    return (k1 >> 2) * 1024 + d * 64 + (k1 & 3) * 16 + c;
}

// LDS exchange of the 16 values a lane holds: v[i] is stored at pidx(ws(i))
// and v[i] reloaded from pidx(rs(i)).  Block: the exchange crosses waves
// (workgroup barriers), else it stays within the wave.  Whole: complex
// elements through a buffer of lds_elems(n) V; otherwise the real parts and
// then the imaginary parts through a buffer of lds_elems(n) T, half the LDS
// for two more barriers.  The stores may overwrite only locations whose
// earlier loads were made by the same lane.
template <bool Block> __device__ __forceinline__ void lds_sync() {
    if constexpr (Block) __syncthreads();
    else wave_sync();
}
// Padded: ws / rs return padded indices already (for patterns whose padding
// the caller can write as a base plus constants, so the compiler keeps one
// base register and immediate offsets instead of sixteen addresses).
template <bool Block, bool Whole, bool Padded = false, typename V, typename WS, typename RS>
__device__ __forceinline__ void lds_exchange(void* lds, V* v, WS ws, RS rs) {
    auto pw = [&](int i) { return Padded ? ws(i) : pidx(ws(i)); };
    auto pr = [&](int i) { return Padded ? rs(i) : pidx(rs(i)); };
    if constexpr (Whole) {
        V* b = static_cast<V*>(lds);
#pragma unroll
        for (int i = 0; i < 16; i++) b[pw(i)] = v[i];
        lds_sync<Block>();
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = b[pr(i)];
    } else {
        using T = decltype(v[0].x);
        T* b = static_cast<T*>(lds);
#pragma unroll
        for (int i = 0; i < 16; i++) b[pw(i)] = v[i].x;
        lds_sync<Block>();
#pragma unroll
        for (int i = 0; i < 16; i++) v[i].x = b[pr(i)];
        lds_sync<Block>();
#pragma unroll
        for (int i = 0; i < 16; i++) b[pw(i)] = v[i].y;
        lds_sync<Block>();
#pragma unroll
        for (int i = 0; i < 16; i++) v[i].y = b[pr(i)];
    }
}

// In: v[n1] = x[256 n1 + t].  Out: v[n1] = y[256 n1 + t].  lds: lds_elems(4096)
// elements of V (Whole) or of T; the caller synchronises before reusing it.
template <typename T, bool Whole = true>
__device__ __forceinline__ void conv4096(void* lds, typename C2<T>::type* v, const typename C2<T>::type* tw,
                                         const typename C2<T>::type* Gp) {
    using V = typename C2<T>::type;
    const int t = threadIdx.x;
    const int w = t >> 6, l = t & 63, q = l & 15;
    const int base = ((w << 2) + (l >> 4)) << 8;
    // padded LDS indices as a lane base plus compile-time constants (base is
    // a multiple of 256): (r << 8) + t -> 272 r + pidx(t);
    // base + 16 a + q -> pb + 17 a; base + 16 q + b -> pq + b
    const int pt = pidx(t);
    const int pb = base + (base >> 4) + q;
    const int pq = base + (base >> 4) + 17 * q;
    // FP64: the inverse stages reload their twiddles and the spectrum is
    // loaded four bins at a time (indices opaque to the compiler, memory
    // clobbers), so forward-stage values are not held across the transform
    auto fresh = [](int e) {
        if constexpr (sizeof(T) == 8) asm volatile("" : "+v"(e));
        return e;
    };
    V wb[4];
    dft16<-1, T>(v);
    twiddle_base<T>(wb, tw, t);
    twiddle_run<-1, T>(v, wb);
    // rows k1 to the waves; 256-point transforms: n2 = 16 a + b, k2 = c + 16 d
    lds_exchange<true, Whole, true>(lds, v, [&](int r) { return 272 * r + pt; },
                                    [&](int a) { return pb + 17 * a; });          // lane q = b
    dft16<-1, T>(v);                                                            // a -> c
    twiddle_base<T>(wb, tw, 16 * q);
    twiddle_run<-1, T>(v, wb);                                                  // W256^(b c)
    lds_exchange<false, Whole, true>(lds, v, [&](int c) { return pb + 17 * c; },
                                     [&](int b) { return pq + b; });              // lane q = c
    dft16<-1, T>(v);                                                            // b -> d
#pragma unroll
    for (int d0 = 0; d0 < 16; d0 += 4) {
        V g[4];
#pragma unroll
        for (int d = 0; d < 4; d++) g[d] = Gp[(w << 10) + ((d0 + d) << 6) + l];
#pragma unroll
        for (int d = 0; d < 4; d++) v[d0 + d] = cmul(v[d0 + d], g[d]);
        if constexpr (sizeof(T) == 8) asm volatile("" ::: "memory");
    }
    dft16<+1, T>(v);                                                            // d -> b
    twiddle_base<T>(wb, tw, fresh(16 * q));
    twiddle_run<+1, T>(v, wb);
    lds_exchange<false, Whole, true>(lds, v, [&](int b) { return pq + b; },
                                     [&](int c) { return pb + 17 * c; });         // lane q = b
    dft16<+1, T>(v);                                                            // c -> a
    const int pt2 = fresh(pt);
    lds_exchange<true, Whole, true>(lds, v, [&](int a) { return pb + 17 * a; },
                                    [&](int r) { return 272 * r + pt2; });
    twiddle_base<T>(wb, tw, fresh(t));
    twiddle_run<+1, T>(v, wb);
    dft16<+1, T>(v);
}

// ---- 1024-point filter block on one wavefront ----
//
// y = IFFT(M(FFT(x))) for 1024 = 16 x 64 on 64 lanes, M applied bin by bin:
//   X[k1 + 16 k2] = sum_n2 W64^(n2 k2) W1024^(n2 k1) sum_n1 x[64 n1 + n2] W16^(n1 k1)
// Lane l first does the DFT16 over n1 of the inputs it holds (x[64 n1 + l])
// and the W1024^(l k1) twiddles.  Row k1 (64 points) then belongs to lanes
// 4 k1 .. 4 k1 + 3 (lane l = 4 j + q: row j, column class q) and is
// transformed as n2 = 4 a + b, k2 = c + 16 d: a DFT16 over a (b = q), the
// W64^(b c) twiddles, a transpose through the row so that lane (j, q) holds
// c = 4 q + i, b = 0..3 at v[4 i + b], and DFT4s over b.  v[4 i + d] is then
// bin k = j + 16 (4 q + i) + 256 d, where M is applied, and the inverse
// retraces the steps.  Every exchange stays inside the wavefront (no
// workgroup barrier).  In: v[n1] = x[64 n1 + l].  Out: v[n1] = 1024 y[64 n1 + l].
// lds: lds_elems(1024) elements of V (Whole) or of T, owned by this wavefront.
template <typename T, bool Whole, typename M>
__device__ __forceinline__ void filter1024_wave(void* lds, typename C2<T>::type* v, const typename C2<T>::type* tw,
                                                M apply) {
    using V = typename C2<T>::type;
    const int l = (int)threadIdx.x & 63;
    const int j = l >> 2, q = l & 3;
    V wb[4];
    dft16<-1, T>(v);                                                     // n1 -> k1
    twiddle_base<T, 1024>(wb, tw, l);
    twiddle_run<-1, T>(v, wb);                                           // W1024^(l k1)
    lds_exchange<false, Whole>(lds, v, [&](int k1) { return 64 * k1 + l; },
                               [&](int a) { return 64 * j + 4 * a + q; });    // row j, b = q
    dft16<-1, T>(v);                                                     // a -> c
    twiddle_base<T, 1024>(wb, tw, 16 * q);
    twiddle_run<-1, T>(v, wb);                                           // W64^(b c)
    lds_exchange<false, Whole>(lds, v, [&](int c) { return 64 * j + 4 * c + q; },
                               [&](int m) { return 64 * j + 16 * q + m; });   // c = 4 q + i, b
#pragma unroll
    for (int i = 0; i < 4; i++) dft4<-1>(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);  // b -> d
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int d = 0; d < 4; d++) v[4 * i + d] = apply(v[4 * i + d], j + 16 * (4 * q + i) + 256 * d);
#pragma unroll
    for (int i = 0; i < 4; i++) dft4<+1>(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);  // d -> b
    lds_exchange<false, Whole>(lds, v, [&](int m) { return 64 * j + 16 * q + m; },
                               [&](int c) { return 64 * j + 4 * c + q; });
    twiddle_base<T, 1024>(wb, tw, 16 * q);
    twiddle_run<+1, T>(v, wb);
    dft16<+1, T>(v);                                                     // c -> a
    lds_exchange<false, Whole>(lds, v, [&](int a) { return 64 * j + 4 * a + q; },
                               [&](int k1) { return 64 * k1 + l; });
    twiddle_base<T, 1024>(wb, tw, l);
    twiddle_run<+1, T>(v, wb);
    dft16<+1, T>(v);                                                     // k1 -> n1
}

// ---- 16 R-point convolution blocks, same scheme, R = 512 or 1024 ----
//
// B = 16 R on R threads.  Stage 1 as in conv4096 (thread t = n2 < R, rows k1
// of R), one block barrier, then each row is transformed by L = R/16 lanes of
// one wave (a wave owns 64/L rows) as 16 x 16 x G, G = R/256, with two
// transposes through its own rows:
//   n2 = L a + G b + g,  k2 = c + 16 e1 + 256 e2  (e1 = (16/G) e1h + e1l)
//   pass A, lane m = G b + g: DFT16 over a -> c, twiddle WR^(m c)
//   pass B, lane G c + g:     DFT16 over b -> e1, twiddle WL^(g e1)
//   pass C, lane G c + e1h:   DFT_G over g -> e2 (held as v[G e1l + g])
// The spectrum is applied in pass C's order (Gp[R k1 + L j + lane in row],
// j = G e1l + e2, see conv_rows_gperm) and the inverse retraces the passes.

template <int R>
inline int conv_rows_gperm(int k) {
    constexpr int G = R / 256, L = R / 16, Q = 16 / G;
    int k1 = k & 15, k2 = k >> 4;
    int c = k2 & 15, e1 = (k2 >> 4) & 15, e2 = k2 >> 8;
    int lane = G * c + e1 / Q, j = G * (e1 % Q) + e2;
    return k1 * R + j * L + lane;
}

template <int G, int S, typename V>
__device__ __forceinline__ void dft_small(V* v) {
    if constexpr (G == 4) dft4<S>(v[0], v[1], v[2], v[3]);
    else dft2<S>(v[0], v[1]);
}

// In: v[n1] = x[R n1 + t].  Out: v[n1] = y[R n1 + t].  lds: lds_elems(16 R)
// elements of V (Whole) or of T (real and imaginary parts exchanged in turn,
// half the LDS for more barriers).  tw: the table of W(16 R)^m.
template <typename T, int R, bool Whole = true>
__device__ __forceinline__ void conv_rows(void* lds, typename C2<T>::type* v, const typename C2<T>::type* tw,
                                          const typename C2<T>::type* Gp) {
    using V = typename C2<T>::type;
    constexpr int NB = 16 * R, G = R / 256, L = R / 16;
    static_assert(R == 512 || R == 1024, "row length");
    const int t = threadIdx.x;
    const int rw = t / L;                   // row owned in the passes (L lanes per row)
    const int l = t % L;                    // lane within the row
    const int row = rw * R;
    const int hi = l / G, lo = l % G;       // (c, g) in pass B, (c, e1h) in pass C
    // In the FP64 1024-row block (128 registers at 1024 threads) the inverse
    // stages reload their twiddles: indices made opaque to the compiler keep
    // it from holding the forward stages' values across the whole transform.
    auto fresh = [](int e) {
        if constexpr (sizeof(T) == 8 && R == 1024) asm volatile("" : "+v"(e));
        return e;
    };
    V wb[4];
    dft16<-1, T>(v);
    twiddle_base<T, NB>(wb, tw, t);
    twiddle_run<-1, T>(v, wb);
    // Padded LDS indices, written as a lane base plus a compile-time
    // constant (pidx(A + B) = pidx(A) + pidx(B) for A a multiple of 16, and
    // row, L and 16 lo are): the compiler then keeps one base register per
    // pattern and immediate offsets.
    constexpr int RP = R + R / 16, LP = L + L / 16;
    const int pt = pidx(t);                              // r R + t  ->  r RP + pt
    const int prow = row + row / 16;
    const int pa = prow + pidx(l);                       // pass A: row + L a + l
    const int pb = prow + LP * hi + lo;                  // pass B: row + L hi + G b + lo
    const int pc = prow + LP * hi + 17 * lo;             // pass C: row + L hi + 16 lo + j
    // rows k1 to their waves; pass A: lane m = l holds n2 = L a + m
    lds_exchange<true, Whole, true>(lds, v, [&](int r) { return r * RP + pt; }, [&](int a) { return pa + LP * a; });
    dft16<-1, T>(v);                                               // a -> c
    twiddle_base<T, NB>(wb, tw, 16 * l);
    twiddle_run<-1, T>(v, wb);                                     // WR^(m c)
    // pass B: lane (c = hi, g = lo) holds m = G b + g
    lds_exchange<false, Whole, true>(lds, v, [&](int c) { return pa + LP * c; },
                                     [&](int b) { return pb + pidx(G * b); });
    dft16<-1, T>(v);                                               // b -> e1
    twiddle_base<T, NB>(wb, tw, 256 * lo);
    twiddle_run<-1, T>(v, wb);                                     // WL^(g e1)
    // pass C: lane (c = hi, e1h = lo) holds (e1l, g) at v[G e1l + g]
    lds_exchange<false, Whole, true>(lds, v, [&](int e1) { return pb + pidx(G * e1); },
                                     [&](int j) { return pc + j; });
#pragma unroll
    for (int e = 0; e < 16; e += G) dft_small<G, -1>(v + e);
    // the spectrum in groups of four loads (the compiler would otherwise hold
    // all sixteen beside the data: 64 more registers in FP64)
#pragma unroll
    for (int j0 = 0; j0 < 16; j0 += 4) {
        V g[4];
#pragma unroll
        for (int j = 0; j < 4; j++) g[j] = Gp[row + (j0 + j) * L + l];
#pragma unroll
        for (int j = 0; j < 4; j++) v[j0 + j] = cmul(v[j0 + j], g[j]);
        if constexpr (sizeof(T) == 8) asm volatile("" ::: "memory");
    }
#pragma unroll
    for (int e = 0; e < 16; e += G) dft_small<G, +1>(v + e);
    // inverse pass B: lane (c, g)
    lds_exchange<false, Whole, true>(lds, v, [&](int j) { return pc + j; },
                                     [&](int e1) { return pb + pidx(G * e1); });
    twiddle_base<T, NB>(wb, tw, fresh(256 * lo));
    twiddle_run<+1, T>(v, wb);
    dft16<+1, T>(v);                                               // e1 -> b
    // inverse pass A: lane m
    lds_exchange<false, Whole, true>(lds, v, [&](int b) { return pb + pidx(G * b); },
                                     [&](int c) { return pa + LP * c; });
    twiddle_base<T, NB>(wb, tw, fresh(16 * l));
    twiddle_run<+1, T>(v, wb);
    dft16<+1, T>(v);                                               // c -> a
    const int pt2 = fresh(pt);
    lds_exchange<true, Whole, true>(lds, v, [&](int a) { return pa + LP * a; }, [&](int r) { return r * RP + pt2; });
    twiddle_base<T, NB>(wb, tw, fresh(t));
    twiddle_run<+1, T>(v, wb);
    dft16<+1, T>(v);
}

}  // namespace cuvhs
