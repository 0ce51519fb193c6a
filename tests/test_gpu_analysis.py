"""Spectral analysis on the GPU (pf_dft) against numpy's FFT, which is what
the reference's analysis.spectrum calls (analysis.py:44-96).  pf_dft's FP64
FFT (radix-2, Bluestein's chirp form for lengths that are not powers of two)
and numpy's FFT round differently, so the bar is 1e-11 of the largest
magnitude of each transform.

Lengths by path: a power of two up to 8192 is a plain FFT in one block;
any other length whose chirp convolution (M >= 2n - 1) fits 8192 values
runs the chirp form in one block (3 ... 4096); longer traces, and powers of
two above 8192, run the three-launch four-step chirp form."""
import numpy as np
import pytest

from helpers import uniform_crystal, spec_from
from phonomena_amd import fdtd
from phonomena_amd.simulation import analysis

pytestmark = pytest.mark.gpu
TOL = 1e-11


def _close(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    assert a.shape == b.shape
    assert float(np.max(np.abs(a - b))) <= TOL * scale


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 100, 1000, 1001, 4095, 4096, 8192, 4097, 10007, 16384, 100003])
def test_dft_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((5, n))
    _close(fdtd.dft(x), np.fft.fft(x))
    z = x + 1j * rng.standard_normal((5, n))
    w = np.hanning(n)
    nf = max(1, n // 2)
    _close(fdtd.dft(z, nf=nf, window=w, scale=0.25), (np.fft.fft(z * w) * 0.25)[:, :nf])


def test_dft_of_a_sinusoid_peaks_at_its_bin():
    n, k = 960, 37
    t = np.arange(n)
    y = np.abs(fdtd.dft(np.cos(2 * np.pi * k * t / n)))
    assert int(np.argmax(y[: n // 2])) == k
    assert abs(y[k] - n / 2) < 1e-9 * n


def test_long_probe_traces():
    # many long traces in one call (four-step path, several scratch rows)
    rng = np.random.default_rng(3)
    n = 50000
    x = rng.standard_normal((48, n))
    w = np.hanning(n)
    _close(fdtd.dft(x, nf=n // 2, window=w, scale=1 / np.sqrt(n)),
           (np.fft.fft(x * w, norm="ortho"))[:, : n // 2])


@pytest.mark.parametrize("n", [2 ** 23, 2 ** 23 - 1])
def test_longest_transforms(n):
    # the largest chirp convolution (2^24 values): one column and one row
    # per four-step block, 256 MB of scratch per trace
    rng = np.random.default_rng(n % 1000)
    x = rng.standard_normal((2, n))
    _close(fdtd.dft(x, nf=4096, scale=0.5), np.fft.fft(x)[:, :4096] * 0.5)


def test_dft_is_linear_and_shifts_to_a_phase():
    # X of x(t - s) (circular) is X_k e^{-2 pi i k s / n}: checks the chirp
    # form's index arithmetic independently of numpy, at a four-step length
    rng = np.random.default_rng(4)
    n, sft = 30011, 977
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = rng.standard_normal(n)
    X, Y, S = fdtd.dft(x), fdtd.dft(y), fdtd.dft(np.roll(x, sft))
    k = np.arange(n)
    _close(S, X * np.exp(-2j * np.pi * ((k * sft) % n) / n))
    _close(fdtd.dft(2.5 * x - y), 2.5 * X - Y)
    # Parseval
    assert abs(np.sum(np.abs(X) ** 2) / n - np.sum(np.abs(x) ** 2)) < 1e-10 * np.sum(np.abs(x) ** 2)


def test_many_traces_in_one_call():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((70000, 16))  # more rows than one launch takes
    _close(fdtd.dft(x, nf=8), np.fft.fft(x)[:, :8])


@pytest.mark.parametrize("x_index", [None, 3])
def test_spectrum_matches_reference_expression(x_index):
    rng = np.random.default_rng(7)
    n = 300
    u = rng.standard_normal((12, 5, 4, n))
    ds = analysis._Dataset({"uz": u}, {"x": np.linspace(0, 11, 12), "dt": 1e-3})
    x, f, d = analysis.spectrum(ds, "uz", 2, 1, x_index=x_index)
    win = np.hanning(n)
    if x_index is None:
        ref = np.abs(np.fft.fft2(u[:, 1, 2, :] * win, norm="ortho"))[:, : n // 2]
    else:
        ref = np.abs(np.fft.fft(u[x_index, 1, 2, :] * win, norm="ortho"))[: n // 2]
    _close(d, ref)
    assert np.array_equal(f, np.fft.fftfreq(n, d=1e-3)[: n // 2])


def test_probe_traces_of_a_run():
    # a sin-driven crystal, probes every 2 steps: the spectrum of a probe
    # trace equals numpy's, and the source frequency is its strongest line
    steps, every = 800, 2
    x, y, z, ids, rows, dt, wave = uniform_crystal((24, 20, 12), 1.0, steps=steps, wave_kind="sin",
                                                   f=None, seed=2)
    spec = spec_from(x, y, z, ids, rows, dt, wave, probe=(0, 2, 12, 0, 2, 10), every=every,
                     cap=steps // every)
    ctx = fdtd.Context(spec, precision="fp64")
    ctx.step(steps)
    frames = ctx.probes()
    ctx.close()
    ds = analysis.probe_dataset(frames, x[0:24:2], y[0:20:2], dt * every)
    xs, f, d = analysis.spectrum(ds, "uz", 0, 3, x_index=2)
    n = frames.shape[0]
    ref = np.abs(np.fft.fft(frames[:, 2, 3] * np.hanning(n), norm="ortho"))[: n // 2]
    _close(d, ref)
    # 20 source periods in 400 frames: the driven line is bin 20
    assert abs(int(np.argmax(d)) - 20) <= 2
    _, _, d2 = analysis.spectrum(ds, "uz", 0, 3)
    _close(d2, np.abs(np.fft.fft2(frames[:, :, 3].T * np.hanning(n), norm="ortho"))[:, : n // 2])
