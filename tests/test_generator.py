"""Synthetic capture generator: lengths, determinism, thread independence."""
import hashlib
import os
import subprocess
import sys

import numpy as np

from conftest import ROOT
from cuvhs_amd import native


def test_config_lengths():
    assert native.gen_params(0).n_samples == 4 * 468140       # exactly 4 * spf, 28 MHz
    assert native.gen_params(1).n_samples == 600 * 468140
    p2 = native.gen_params(2)
    assert p2.n_samples == 6000 * 2542 * 263 and p2.sample_rate_mhz == 40.0


def test_deterministic_and_in_range(cap0):
    again = native.generate(0)
    assert np.array_equal(again, cap0)
    assert 0 < cap0.min() and cap0.max() < 255 and abs(cap0.mean() - 128) < 0.5
    other = native.generate(0, seed=99)
    assert not np.array_equal(other, cap0)


def test_thread_count_independent():
    code = ("import sys,hashlib; sys.path.insert(0,%r); from cuvhs_amd import native; "
            "print(hashlib.sha256(native.generate(1, n_samples=3000000).tobytes()).hexdigest())") % ROOT
    outs = []
    for t in ("1", "6"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.append(subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                                   check=True).stdout.strip())
    assert outs[0] == outs[1]


def test_dropouts_dip_the_envelope():
    # one long dropout forced by a high Poisson rate: the RF amplitude drops
    cap = native.generate(1, n_samples=468140 * 2, snr_lo_db=60.0, snr_hi_db=60.0, dropouts_per_field=20.0)
    x = cap.astype(np.float64) - 128.0
    blk = np.abs(x[: x.size // 256 * 256]).reshape(-1, 256).mean(1)
    assert blk.min() < 0.2 * np.median(blk)
