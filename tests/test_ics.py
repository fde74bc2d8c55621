"""The synthetic initial states: the separable band-limited generator used for
grids too large for a full-size FFT gives the FFT generator's field, and both
fields have the advertised spectrum and amplitude."""
import numpy as np
import pytest

import pcmx_ics as I


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 24, 20), (20, 12, 28)])
def test_separable_matches_fft(shape):
    nx, ny, nz = shape
    ref = I.bandlimited(3, shape, 1, 4, 0.05, 11)
    out = np.zeros((3, nz, ny, nx))
    I.bandlimited_into(out, 1, 4, 0.05, 11, zslab=5)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-14)


def test_writes_into_a_ghosted_view():
    n, g = 24, 3
    f = np.full((4, n + 2 * g, n + 2 * g, n + 2 * g), 7.0)
    I.bandlimited_into(f[1:4, g:-g, g:-g, g:-g], 1, 4, 0.02, 5)
    np.testing.assert_allclose(f[1:4, g:-g, g:-g, g:-g], I.bandlimited(3, n, 1, 4, 0.02, 5), atol=1e-14)
    assert np.all(f[0] == 7.0)                       # other variables untouched
    assert np.all(f[1:4, :g] == 7.0) and np.all(f[1:4, :, :, -g:] == 7.0)   # ghosts untouched


def test_shell_spectrum_and_rms():
    n = 32
    out = np.zeros((3, n, n, n))
    I.bandlimited_into(out, 1, 4, 0.05, 3)
    rms = np.sqrt(np.mean(np.sum(out * out, axis=0)))
    assert abs(rms - 0.05) < 1e-15
    spec = np.abs(np.fft.rfftn(out, axes=(1, 2, 3))) ** 2
    kx = np.arange(n // 2 + 1)[None, None, :]
    k = np.fft.fftfreq(n, 1.0 / n)
    kk = np.sqrt(kx ** 2 + k[None, :, None] ** 2 + k[:, None, None] ** 2)
    inside = (kk >= 1) & (kk <= 4)
    assert spec[:, ~inside].max() < 1e-20 * spec[:, inside].max()
    assert np.all(spec[:, inside].sum(axis=1) > 0)


def test_rejects_grid_smaller_than_shell():
    with pytest.raises(ValueError):
        I.bandlimited_into(np.zeros((1, 8, 8, 8)), 1, 4, 0.1, 0)
