"""VSYNC clustering as a chain of upper-bound searches (sync.hip:
k_vsync_cluster) against the reference's sequential loop over the sorted
candidates (pipeline.cu:1598-1626): accept a candidate when it lies more than
spf/2 past the last accepted safe offset (position less 5 lines, clamped at
0), cap the offsets at the chunk's field count.  Random sorted candidate sets
with bursts (several VSYNC-width runs per field, noise runs), positions
inside the first 5 lines, and caps below and above the accepted count."""
import numpy as np
import pytest


def sequential(cands, spf, spl, num_fields):
    offs = []
    for pos in cands:
        if not offs or (pos - offs[-1]) % (1 << 64) > spf // 2:
            offs.append(pos - 5 * spl if pos > 5 * spl else 0)
    return offs[:num_fields]


def chained(cands, spf, spl, num_fields):
    safe = lambda p: p - 5 * spl if p > 5 * spl else 0
    n = len(cands)
    nxt = [0] * n
    for i in range(n):
        thr = safe(cands[i]) + spf // 2
        lo, hi = i + 1, n
        while lo < hi:
            mid = (lo + hi) >> 1
            if cands[mid] > thr:
                hi = mid
            else:
                lo = mid + 1
        nxt[i] = lo
    offs, i = [], 0
    while i < n and len(offs) < num_fields:
        offs.append(safe(cands[i]))
        i = nxt[i]
    return offs


@pytest.mark.parametrize("seed", range(40))
def test_chain_matches_sequential(seed):
    rng = np.random.default_rng(seed)
    spf = int(rng.integers(200_000, 700_000))
    spl = spf // 263
    nf = int(rng.integers(1, 70))
    pos = []
    start = int(rng.integers(0, 6 * spl))
    for f in range(nf):
        base = start + f * spf + int(rng.integers(-spl, spl))
        pos += [base + k * (spl // 2) for k in range(int(rng.integers(0, 8)))]
    pos += list(rng.integers(0, nf * spf + spf, int(rng.integers(0, 40))))
    cands = sorted(set(int(p) for p in pos if p >= 0))
    for cap in (1, max(1, nf // 2), nf, nf + 5):
        assert chained(cands, spf, spl, cap) == sequential(cands, spf, spl, cap)
