"""The bit algebra of the sync-threshold events (k1_filters.hip: k1_part_b's
bitmask ORs; sync.hip: k_thr_bits, k_ev_count, k_ev_write), emulated with
numpy and compared with a direct scan for run boundaries of x <= thr
(vsync_discover.cu:17-20: a start is a sample at or below the threshold
whose predecessor is not, sample 0 included; an end is the first sample
above it after a run).  The block ranges are the K1 video blocks of the
28 MHz and 40 MHz formats, so ballots straddle bitmask words at every
offset.  The GPU parity tests pin the kernels themselves."""
import numpy as np
import pytest

M32 = (1 << 32) - 1
B_MAX = 8192


def direct_events(b):
    ev = []
    for i in range(len(b)):
        p = b[i - 1] if i > 0 else False
        if b[i] and not p:
            ev.append(2 * i)
        elif not b[i] and p:
            ev.append(2 * i + 1)
    return np.array(ev, np.int64)


def k1b_bits(b, spfp, V, nt=256, lp=1000):
    """k1_part_b: every block (b, f) ballots its 64-lane rows, ORs the up to
    three words each straddles into its LDS words (word 0 = the word of the
    block's first sample), then stores the words inside its range and ORs
    the partly covered first / last word into the bitmask.  Blocks run in
    reverse order and every stored word must be one no other block touches."""
    T = len(b)
    nf = T // spfp
    words = np.zeros(((T + 63) // 64) * 2, np.uint64)
    touched = np.zeros(len(words), np.int64)
    whole = []
    for f in reversed(range(nf)):
        for blk in reversed(range((spfp + V - 1) // V)):
            s_lo = f * spfp + blk * V
            s_end = f * spfp + min((blk + 1) * V, spfp)
            w_lo, w_hi = s_lo >> 5, (s_end - 1) >> 5
            off = s_lo & 31
            local = [0] * (B_MAX // 32 + 4)
            for r in range(16):
                for w in range(nt // 64):
                    o = r * nt + 64 * w + np.arange(64) - lp
                    q_rel = blk * V + o            # q - kOverlap
                    ok = (o >= 0) & (o < V) & (q_rel < spfp)
                    s = f * spfp + q_rel
                    bit = ok & np.where(ok, b[np.clip(s, 0, T - 1)], False)
                    m = sum(1 << int(j) for j in np.nonzero(bit)[0])
                    if m == 0:
                        continue
                    rel = off + int(o[0])
                    sh = rel & 31
                    parts = [(m << sh) & M32, ((m << sh) & ((1 << 64) - 1)) >> 32,
                             (m >> (64 - sh)) & M32 if sh else 0]
                    for k, pv in enumerate(parts):
                        if pv:
                            assert 0 <= (rel >> 5) + k < len(local)
                            local[(rel >> 5) + k] |= pv
            nw = w_hi - w_lo + 1
            for i in range(nw):
                partial = (i == 0 and off != 0) or (i == nw - 1 and (s_end & 31) != 0)
                if partial:
                    words[w_lo + i] |= local[i]
                else:
                    words[w_lo + i] = local[i]
                    whole.append(w_lo + i)
                touched[w_lo + i] += 1
    assert all(touched[w] == 1 for w in whole), "a whole-word store hits a word another block writes"
    return words


def thr_bits(b):
    """k_thr_bits: 64-aligned ballots, two whole words each."""
    T = len(b)
    words = np.zeros(((T + 63) // 64) * 2, np.uint64)
    for i in range(0, T, 64):
        m = 0
        for j in range(64):
            if i + j < T and b[i + j]:
                m |= 1 << j
        words[i >> 5] = m & M32
        words[(i >> 5) + 1] = m >> 32
    return words


def events_from_bits(words, T, wg=256):
    """k_ev_count + k_ev_scan + k_ev_write."""
    nwg = (T + 8191) // 8192
    masks = []
    for j in range(nwg * wg):
        s = 32 * j
        fs = fe = 0
        if s < T:
            bw = int(words[j])
            p = ((bw << 1) & M32) | ((int(words[j - 1]) >> 31) if j > 0 else 0)
            fs = bw & ~p & M32
            fe = ~bw & p & M32
            if T - s < 32:
                valid = (1 << (T - s)) - 1
                fs &= valid
                fe &= valid
        masks.append((fs, fe))
    ev = []
    for j, (fs, fe) in enumerate(masks):  # threads in order = exclusive scan order
        a = fs | fe
        while a:
            k = (a & -a).bit_length() - 1
            ev.append(2 * (32 * j + k) + ((fe >> k) & 1))
            a &= a - 1
    return np.array(ev, np.int64)


# (segment length, valid outputs per block, threads per block, causal taps):
# the 28 MHz 4096-point and 40 MHz 8192-point video blocks on shortened
# segments, and a small case
@pytest.mark.parametrize("spfp,V,nt,lp", [(58517, 2271, 256, 1082), (83596, 5586, 512, 1544), (1000, 700, 256, 1000)])
def test_bitmask_events_match_direct_scan(spfp, V, nt, lp):
    rng = np.random.default_rng(spfp)
    nf = 2
    T = nf * spfp
    # runs of random lengths above / below the threshold, plus isolated flips
    x = np.repeat(rng.integers(0, 2, 4000).astype(bool), rng.integers(1, 90, 4000))[:T]
    if len(x) < T:
        x = np.concatenate([x, np.zeros(T - len(x), bool)])
    x[rng.integers(0, T, 50)] ^= True
    x[0] = True  # a run open at sample 0 starts there
    want = direct_events(x)
    wk = k1b_bits(x, spfp, V, nt, lp)
    assert np.array_equal(wk, thr_bits(x)), "K1 part B ORs and the array form disagree"
    got = events_from_bits(wk, T)
    assert np.array_equal(got, want)
    # a total that ends mid-word: no event at or past it
    T2 = T - 13
    got2 = events_from_bits(thr_bits(x[:T2]), T2)
    assert np.array_equal(got2, direct_events(x[:T2]))
