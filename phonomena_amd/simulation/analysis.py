"""Spectral analysis of recorded fields (phonomena/simulation/analysis.py).

Same names and meaning as the reference module: ``nonlinspace``,
``trim_trailing_zeros`` and ``spectrum(file, u_id, z_index, y_index,
x_index=None)``, which returns (x, f, |DFT|) of a Hann-windowed time trace
(x_index given) or of an (x, t) slice (x_index None), ortho-normalised,
first N//2 frequencies.  The transforms run on the GPU (``fdtd.dft``, a
batched FP64 DFT kernel); there is no CPU fallback.

``file`` is an HDF path in the reference's Writer schema (needs h5py, which
this image lacks) or any mapping with the same layout: the dataset
``file[u_id]`` of shape (nx, ny, nz, T) and ``file.attrs`` (or
``file["attrs"]``) holding ``x``, ``fdx`` and ``dt``.  ``probe_dataset``
wraps the surface u_z probe frames of a run in that layout.
"""
import numpy as np

from .. import fdtd


def nonlinspace(spacing):
    """Positions from spacings (analysis.py:8-17): X[0] = 0 and
    X[i] = X[i-1] + spacing[i] (the first spacing is not used)."""
    spacing = np.asarray(spacing)
    if len(spacing.shape) > 1:
        raise TypeError('Only supports 1D arrays')
    out = np.zeros(spacing.size)
    for i in range(1, spacing.size):
        out[i] = out[i - 1] + spacing[i]
    return out


def trim_trailing_zeros(arr, threshold=1):
    """Drop trailing values below threshold % of the maximum
    (analysis.py:19-42).  Values below the threshold are set to 0 in ``arr``
    itself, as in the reference.  Returns (trimmed, kept index range)."""
    og = np.copy(arr)
    arr[arr < (threshold * np.max(arr) / 100)] = 0
    if len(arr.shape) == 1:
        out = np.trim_zeros(arr, 'b')
        if out.size == 0:
            out = og
        idx = range(0, out.size)
    elif len(arr.shape) == 2:
        flat = np.prod(arr, axis=0)
        tmp = np.trim_zeros(flat, 'b')
        if tmp.size == 0:
            tmp = flat
        idx = range(0, tmp.size)
        out = arr[:, idx]
    else:
        raise TypeError("Only accepts 1D arrays")
    return out, idx


class _Dataset(dict):
    """Mapping with an ``attrs`` member, the shape of an h5py.File."""

    def __init__(self, data, attrs):
        super().__init__(data)
        self.attrs = dict(attrs)


def probe_dataset(frames, x, y, frame_dt, u_id="uz"):
    """The probe frames of a run, (T, ni, nj) as ``Context.probes()`` or
    ``Solver.probes`` return them, as a dataset ``spectrum`` reads: u_id of
    shape (ni, nj, 1, T), attrs x / y (the probe nodes' coordinates) and dt
    (the time between frames: solver dt x record_every)."""
    f = np.asarray(frames, dtype=np.float64)
    u = np.ascontiguousarray(np.transpose(f, (1, 2, 0))[:, :, None, :])
    return _Dataset({u_id: u}, {"x": np.asarray(x, dtype=np.float64), "y": np.asarray(y, dtype=np.float64),
                                "dt": float(frame_dt)})


def _open(file):
    if isinstance(file, str):
        try:
            import h5py
        except ImportError as e:
            raise ImportError("reading an HDF file needs h5py, which is not installed; pass the datasets as a "
                              "mapping (see probe_dataset)") from e
        return h5py.File(file, 'r'), True
    if hasattr(file, "get"):
        return file, False
    raise TypeError("file must be an HDF path or a mapping of datasets")


def _attrs(hdf):
    a = getattr(hdf, "attrs", None)
    return a if a is not None else hdf["attrs"]


def spectrum(file, u_id, z_index, y_index, x_index=None):
    """One- or two-dimensional spectrum of a recorded field
    (analysis.py:44-96).

    x_index given: |DFT_t(u[x, y, z, :] * hann)| (ortho), first N//2
    frequencies.  x_index None: |DFT_x DFT_t(u[:, y, z, :] * hann)| (ortho
    in both axes, numpy's fft2), first N//2 frequencies of every x row.
    Returns (x, f, dft)."""
    hdf, close = _open(file)
    try:
        u = hdf.get(u_id)
        assert u is not None
        attrs = _attrs(hdf)
        if u_id == 'ux':
            x = nonlinspace(np.asarray(attrs["fdx"])[:, 0, 0])
        else:
            x = np.array(attrs["x"])
        dt = attrs["dt"]
        n = u.shape[3]
        nf = n // 2
        f = np.fft.fftfreq(n, d=dt)[:nf]
        window = np.hanning(n)
        if nf == 0:
            shape = (u.shape[0], 0) if x_index is None else (0,)
            return x, f, np.zeros(shape)
        if x_index is None:
            a = np.asarray(u[:, y_index, z_index, :], dtype=np.float64)
            t = fdtd.dft(a, nf=nf, window=window, scale=1.0 / np.sqrt(n))             # (nx, nf)
            s = fdtd.dft(np.ascontiguousarray(t.T), scale=1.0 / np.sqrt(a.shape[0]))  # (nf, nx)
            dft = np.abs(s.T)
        else:
            a = np.asarray(u[x_index, y_index, z_index, :], dtype=np.float64)
            dft = np.abs(fdtd.dft(a, nf=nf, window=window, scale=1.0 / np.sqrt(n)))
    finally:
        if close:
            hdf.close()
    return x, f, dft
