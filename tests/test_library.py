"""The C ABI library without a GPU: it loads, exports every symbol the header
declares, and rejects invalid descriptors before touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

from phonomena_amd import fdtd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    out = []
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            out += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pf_\w+)\s*\(", txt, re.M)
    return sorted(set(out))


def test_library_is_built_and_loads():
    assert os.path.exists(fdtd.LIB_PATH), "run __graft_entry__.build() first"
    lib = fdtd.load()
    assert lib.pf_abi_version() == 2


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(fdtd.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s


def _desc(**over):
    nx, ny, nz = over.pop("n", (6, 6, 6))
    arrs = dict(
        fd=[np.ones(nx - 1), np.ones(ny - 1), np.ones(nz - 1)],
        sd=[np.ones(nx - 2), np.ones(ny - 2), np.ones(nz - 2)],
        mat=np.array([[1.0] * 12 + [1.0]]), ids=np.zeros((nx, ny, nz), np.uint8),
        mur=np.zeros(12), wave=np.zeros(10))
    arrs.update(over.pop("arrs", {}))
    d = fdtd.PfDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.precision, d.kernel, d.n_materials = 0, 0, arrs["mat"].shape[0]
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    d.fdx, d.fdy, d.fdz = (p(a) for a in arrs["fd"])
    d.sdx, d.sdy, d.sdz = (p(a) for a in arrs["sd"])
    d.materials, d.ids, d.mur, d.wave = p(arrs["mat"]), p(arrs["ids"]), p(arrs["mur"]), p(arrs["wave"])
    d.wave_len = arrs["wave"].size
    d.dt, d.dt2 = 1.0, 1.0
    for k, v in over.items():
        setattr(d, k, v)
    return d, arrs


@pytest.mark.parametrize("change,msg", [
    (dict(n=(3, 6, 6)), "node counts"),
    (dict(precision=7), "precision"),
    (dict(kernel=9), "kernel"),
    (dict(n_materials=0), "n_materials"),
    (dict(dt=0.0), "dt"),
    (dict(arrs={"ids": np.full((6, 6, 6), 3, np.uint8)}), "out of range"),
    (dict(arrs={"fd": [np.r_[1.0, -1.0, 1, 1, 1], np.ones(5), np.ones(5)]}), "fdx"),
    (dict(arrs={"mat": np.array([[1.0] * 12 + [0.0]])}), "rho"),
])
def test_create_rejects_bad_descriptors(change, msg):
    lib = fdtd.load()
    d, keep = _desc(**change)
    h = ctypes.c_void_p()
    rc = lib.pf_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == -1
    assert msg in lib.pf_last_error().decode()
    assert not h.value


def test_bad_probe_grid_rejected():
    lib = fdtd.load()
    d, keep = _desc()
    for n, v in enumerate((0, 2, 4, 0, 1, 3)):  # i0 + si*(ni-1) = 6 > nx-1
        d.probe[n] = v
    d.record_every, d.probe_capacity = 1, 4
    h = ctypes.c_void_p()
    assert lib.pf_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert "probe" in lib.pf_last_error().decode()


def test_null_handles_are_errors_not_crashes():
    lib = fdtd.load()
    assert lib.pf_step(None, 1) == -1
    assert lib.pf_sync(None) == -1
    assert lib.pf_destroy(None) == 0
    e = ctypes.c_double()
    assert lib.pf_energy(None, ctypes.byref(e)) == -1


def test_context_wrapper_raises_fdtd_error():
    x = np.arange(6.0)
    from phonomena_amd.fdtd import MeshSpec
    spec = MeshSpec((6, 6, 6), [np.ones(5)] * 3, [np.ones(4)] * 3, np.array([[1.0] * 13]),
                    np.full((6, 6, 6), 2, np.uint8), 1.0, 1.0, np.zeros(12), np.zeros(4))
    with pytest.raises(fdtd.FdtdError, match="PF_ERR_ARG"):
        fdtd.Context(spec)
    del x


def test_dft_arguments_are_checked_before_any_device_work():
    # pf_dft's longest transform is 2^23 values (the chirp convolution of
    # length 2^24); argument errors return before the device is touched
    with pytest.raises(fdtd.FdtdError, match="longest transform"):
        fdtd.dft(np.zeros(2 ** 23 + 1), nf=1)
    lib = fdtd.load()
    one = np.zeros(1)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    # lengths far beyond the limit fail the same way (no shift or size overflow)
    assert lib.pf_dft(vp(one), None, 1, 2 ** 62 + 1, 1, None, 1.0, vp(one), vp(one)) == -1
    assert "longest transform" in lib.pf_last_error().decode()
    lib = fdtd.load()
    out = np.zeros(4)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert lib.pf_dft(vp(out), None, 1, 4, 5, None, 1.0, vp(out), vp(out)) == -1
    assert "nf" in lib.pf_last_error().decode()
