"""ctypes wrapper for ``oracle/brute.c`` (the scalar restatement).

TEST INFRASTRUCTURE ONLY; see ``oracle/reference_math.py``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def _lib(omp=False):
    """The serial library, or the OpenMP build (same bits, all cores)."""
    if omp not in _LIBS:
        name = "liboracle_brute_omp.so" if omp else "liboracle_brute.so"
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError("oracle/%s missing: run `make -C oracle`" % name)
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        lib.oracle_run.restype = ctypes.c_int
        lib.oracle_run.argtypes = ([ctypes.c_int] * 3 + [P] * 6 + [P, P, ctypes.c_double, P]
                                   + [P] * 3 + [P, ctypes.c_int, ctypes.c_int] + [P] * 9
                                   + [P, ctypes.c_int, P])
        _LIBS[omp] = lib
    return _LIBS[omp]


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def material_table(c_list, p_list):
    """(n,13) rows C00 C01 C02 C10 C11 C12 C20 C21 C22 C33 C44 C55 rho."""
    rows = []
    for c, p in zip(c_list, p_list):
        c = np.asarray(c, dtype=np.float64)
        rows.append([c[0, 0], c[0, 1], c[0, 2], c[1, 0], c[1, 1], c[1, 2],
                     c[2, 0], c[2, 1], c[2, 2], c[3, 3], c[4, 4], c[5, 5], float(p)])
    return np.ascontiguousarray(rows, dtype=np.float64)


def run(x, y, z, ids, table, dt, mur12, wave, steps, state=None, t0=0,
        probe=None, every=0, sponge=None, si=1.0, omp=False):
    """Run the scalar oracle; returns (state dict, frames or None).

    ``state`` maps ux, uy, uz, ux_new, ..., uz_old to float64 arrays of the
    reference shapes and is updated in place (fresh zeros if None).
    """
    nx, ny, nz = x.size, y.size, z.size
    xs, ys, zs = x * si, y * si, z * si
    fdx = np.ascontiguousarray(xs[1:] - xs[:-1])
    fdy = np.ascontiguousarray(ys[1:] - ys[:-1])
    fdz = np.ascontiguousarray(zs[1:] - zs[:-1])
    sdx = np.ascontiguousarray(np.mean([fdx[1:], fdx[:-1]], axis=0))
    sdy = np.ascontiguousarray(np.mean([fdy[1:], fdy[:-1]], axis=0))
    sdz = np.ascontiguousarray(np.mean([fdz[1:], fdz[:-1]], axis=0))
    shp = {"x": (nx - 1, ny, nz), "y": (nx, ny - 1, nz), "z": (nx, ny, nz - 1)}
    if state is None:
        state = {}
        for c in "xyz":
            for suf in ("", "_new", "_old"):
                state["u" + c + suf] = np.zeros(shp[c])
    ids = np.ascontiguousarray(ids, dtype=np.uint8)
    table = np.ascontiguousarray(table, dtype=np.float64)
    mur12 = np.ascontiguousarray(mur12, dtype=np.float64)
    wave = np.ascontiguousarray(wave, dtype=np.float64)
    frames = None
    pr = None
    if probe is not None:
        pr = np.ascontiguousarray(probe, dtype=np.int32)
        nfr = sum(1 for tt in range(t0, t0 + steps) if every and (tt + 1) % every == 0)
        frames = np.zeros((max(nfr, 1), probe[2], probe[5]))
    dxs = dys = dzs = None
    if sponge is not None:
        dxs, dys, dzs = (np.ascontiguousarray(a, dtype=np.float64) for a in sponge)
    D = dt ** 2
    n = _lib(omp).oracle_run(nx, ny, nz, _p(fdx), _p(fdy), _p(fdz), _p(sdx), _p(sdy), _p(sdz),
                          _p(ids), _p(table), D, _p(mur12), _p(dxs), _p(dys), _p(dzs),
                          _p(wave), t0, steps,
                          *[_p(state["u" + c + suf]) for suf in ("", "_new", "_old") for c in "xyz"],
                          _p(pr), int(every), _p(frames))
    if n < 0:
        raise MemoryError("oracle_run failed")
    if frames is not None:
        frames = frames[:n]
    return state, frames
