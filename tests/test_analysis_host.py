"""Spectral analysis, host side (no GPU): the reference's helper functions
(analysis.py:8-42) on closed forms, and spectrum's slicing, windowing and
normalisation with the GPU transform replaced by numpy's DFT, against the
reference's own expression (numpy fft / fft2 with norm="ortho")."""
import numpy as np
import pytest

from phonomena_amd import fdtd
from phonomena_amd.simulation import analysis


def test_nonlinspace_skips_the_first_spacing():
    assert np.array_equal(analysis.nonlinspace(np.array([5.0, 1.0, 2.0, 4.0])), [0.0, 1.0, 3.0, 7.0])
    with pytest.raises(TypeError):
        analysis.nonlinspace(np.zeros((2, 2)))


def test_trim_trailing_zeros():
    a = np.array([1.0, 50.0, 100.0, 0.5, 0.2, 0.0])
    out, idx = analysis.trim_trailing_zeros(a.copy(), threshold=1)
    assert np.array_equal(out, [1.0, 50.0, 100.0])
    assert list(idx) == [0, 1, 2]
    b = np.array([[1.0, 2.0, 0.001], [3.0, 4.0, 5.0]])
    out, idx = analysis.trim_trailing_zeros(b.copy(), threshold=1)
    assert out.shape == (2, 2) and list(idx) == [0, 1]


def _numpy_dft(x, nf=None, window=None, scale=1.0):
    x = np.asarray(x)
    n = x.shape[-1]
    w = np.ones(n) if window is None else window
    return (np.fft.fft(x * w) * scale)[..., : (n if nf is None else nf)]


@pytest.mark.parametrize("n", [1, 2, 7, 64, 101])
def test_spectrum_matches_reference_expression(monkeypatch, n):
    monkeypatch.setattr(fdtd, "dft", _numpy_dft)
    rng = np.random.default_rng(n)
    u = rng.standard_normal((9, 4, 3, n))
    ds = analysis._Dataset({"uy": u, "ux": u}, {"x": np.linspace(0, 1, 9), "dt": 0.125,
                                                "fdx": np.full((9, 4, 3), 0.5)})
    nf = n // 2
    win = np.hanning(n)
    x, f, d1 = analysis.spectrum(ds, "uy", 2, 1, x_index=4)
    assert np.array_equal(f, np.fft.fftfreq(n, d=0.125)[:nf])
    assert np.allclose(d1, np.abs(np.fft.fft(u[4, 1, 2, :] * win, norm="ortho"))[:nf], rtol=1e-12, atol=1e-14)
    x, f, d2 = analysis.spectrum(ds, "uy", 2, 1)
    ref = np.abs(np.fft.fft2(u[:, 1, 2, :] * win, norm="ortho"))[:, :nf]
    assert d2.shape == ref.shape
    assert np.allclose(d2, ref, rtol=1e-12, atol=1e-14)
    x, _, _ = analysis.spectrum(ds, "ux", 2, 1, x_index=0)
    assert np.array_equal(x, analysis.nonlinspace(np.full(9, 0.5)))


def test_probe_dataset_layout():
    frames = np.arange(2 * 3 * 4, dtype=float).reshape(2, 3, 4)
    ds = analysis.probe_dataset(frames, x=[0, 1, 2], y=[0, 1, 2, 3], frame_dt=0.5)
    assert ds["uz"].shape == (3, 4, 1, 2)
    assert ds["uz"][1, 2, 0, 1] == frames[1, 1, 2]
    assert ds.attrs["dt"] == 0.5


def test_hdf_path_needs_h5py():
    try:
        import h5py  # noqa: F401
        pytest.skip("h5py present")
    except ImportError:
        pass
    with pytest.raises(ImportError, match="h5py"):
        analysis.spectrum("missing.hdf", "uz", 0, 0)
