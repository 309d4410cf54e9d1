"""Oracle filter design (fm_demod.cu:273-374, chroma_decode.cu:568-587)
against scipy.signal and closed-form invariants."""
import os

import numpy as np
import pytest

from conftest import GOLDEN


def table(oracle, mhz, which):
    n = oracle.lib().or_filter_table(float(mhz), which, None)
    out = np.zeros(n * (2 if which in (1, 2) else 1))
    oracle.lib().or_filter_table(float(mhz), which, out.ctypes.data)
    if which in (1, 2):
        return out[0::2] + 1j * out[1::2]
    return out


@pytest.mark.parametrize("mhz", [28, 40])
def test_rf_response_matches_scipy_butterworth(oracle, mhz):
    g = np.load(os.path.join(GOLDEN, "filter_refs.npz"))
    rf = table(oracle, mhz, 0)
    k = g[f"rf_bins_{mhz}"]
    ref = g[f"rf_mag_{mhz}"]
    assert np.allclose(rf[k], ref, rtol=1e-8, atol=1e-12)
    assert rf[0] == 0.0  # band-pass is zero at DC


@pytest.mark.parametrize("mhz", [28, 40])
def test_sync_fir_is_hamming_firwin(oracle, mhz):
    g = np.load(os.path.join(GOLDEN, "filter_refs.npz"))
    fv = table(oracle, mhz, 1)
    f05 = table(oracle, mhz, 2)
    k = g[f"rf_bins_{mhz}"]
    with np.errstate(invalid="ignore", divide="ignore"):
        ratio = f05[k] / fv[k]
    ok = np.abs(fv[k]) > 1e-12 * np.abs(fv).max()
    assert np.allclose(ratio[ok], g[f"sync_{mhz}"][ok], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("mhz", [28, 40])
def test_video_response_invariants(oracle, mhz):
    """De-emphasis = inverted RBJ high shelf (273,755.82 Hz, 13.9794 dB):
    unity at DC and 1/A = 10^(-13.9794/40) at the shelf frequency; the 3 MHz
    super-Gaussian low-pass is flat there and halves the gain at 1.5 MHz
    relative to the de-emphasis alone; everything carries the 1/N scale."""
    fv = table(oracle, mhz, 1)
    fb = fv.size
    n = oracle.lib().or_filter_table(float(mhz), 0, None) * 2 - 2
    fs = mhz * 1e6
    assert abs(fv[0].real * n - 1.0) < 1e-12 and fv[0].imag == 0.0
    k0 = int(round(273755.82 / (fs / 2) * (fb - 1)))
    assert abs(abs(fv[k0]) * n - 10 ** (-13.9794 / 40)) < 2e-4
    mag = np.abs(fv) * n
    k3 = int(3.0e6 / (fs / 2) * (fb - 1))
    assert np.all(np.diff(mag[: k3]) <= 1e-15)  # monotone low-pass below 3 MHz
    # swapping the shelf's numerator and denominator would give A, not 1/A
    assert abs(fv[k0]) * n < 1.0


@pytest.mark.parametrize("mhz", [28, 40])
def test_chroma_prebpf_is_squared_butterworth(oracle, mhz):
    g = np.load(os.path.join(GOLDEN, "filter_refs.npz"))
    h = table(oracle, mhz, 3)
    k = g[f"pre_bins_{mhz}"]
    assert np.allclose(h[k], g[f"pre_mag2_{mhz}"], rtol=1e-8, atol=1e-20)
