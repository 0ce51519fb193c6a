"""ctypes binding of libphonofdtd (the C ABI in include/phonofdtd.h).

This is the only way the Python side reaches the GPU: there is no CPU
fallback.  If the library is missing or fails to load, every call raises
``FdtdUnavailable``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libphonofdtd.so")

PF_OK = 0
ERRORS = {-1: "PF_ERR_ARG", -2: "PF_ERR_HIP", -3: "PF_ERR_STATE", -4: "PF_ERR_NOMEM", -5: "PF_ERR_RANGE"}
PRECISION = {"fp64": 0, "fp32": 1, "fp32c": 2}
NOALLOC_FIELD_MAX = 4 * 1024 * 1024  # PF_NOALLOC_FIELD_MAX
KERNEL = {"auto": 0, "naive": 1, "tiled": 2, "strip": 3, "tiled2": 4}
FIELDS = {"ux": 0, "uy": 1, "uz": 2, "ux_old": 3, "uy_old": 4, "uz_old": 5}


class FdtdError(RuntimeError):
    """A libphonofdtd call returned an error code."""

    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, code), msg))
        self.code = code


class FdtdUnavailable(FdtdError):
    """The HIP library is not built or cannot be loaded."""

    def __init__(self, msg):
        RuntimeError.__init__(self, msg)
        self.code = None


class PfDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("kernel", ctypes.c_int32), ("n_materials", ctypes.c_int32),
        ("fdx", ctypes.c_void_p), ("fdy", ctypes.c_void_p), ("fdz", ctypes.c_void_p),
        ("sdx", ctypes.c_void_p), ("sdy", ctypes.c_void_p), ("sdz", ctypes.c_void_p),
        ("materials", ctypes.c_void_p), ("ids", ctypes.c_void_p),
        ("dt", ctypes.c_double), ("dt2", ctypes.c_double),
        ("mur", ctypes.c_void_p),
        ("sponge_x", ctypes.c_void_p), ("sponge_y", ctypes.c_void_p), ("sponge_z", ctypes.c_void_p),
        ("wave", ctypes.c_void_p), ("wave_len", ctypes.c_int64),
        ("probe", ctypes.c_int32 * 6), ("record_every", ctypes.c_int32), ("probe_capacity", ctypes.c_int32),
        ("tile_j", ctypes.c_int32), ("tile_x", ctypes.c_int32), ("tile_z", ctypes.c_int32),
        ("node_x", ctypes.c_void_p), ("node_y", ctypes.c_void_p),
        ("n_targets", ctypes.c_int32), ("targets", ctypes.c_void_p),
        ("surface", ctypes.c_int32),
    ]


_lib = None


def load(path=None):
    """Load (once) and return the library; raise FdtdUnavailable if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FdtdUnavailable("libphonofdtd.so is not built (expected at %s); run "
                              "`python -c 'import __graft_entry__ as g; g.build()'`" % p)
    try:
        lib = ctypes.CDLL(p)
    except OSError as e:
        raise FdtdUnavailable("cannot load %s: %s" % (p, e))
    V, I32, I64, D = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "pf_abi_version": ([], ctypes.c_int),
        "pf_last_error": ([], ctypes.c_char_p),
        "pf_device_info": ([ctypes.c_char_p, I32, P(I32), P(D)], ctypes.c_int),
        "pf_create": ([P(PfDesc), P(V)], ctypes.c_int),
        "pf_destroy": ([V], ctypes.c_int),
        "pf_step": ([V, I64], ctypes.c_int),
        "pf_sync": ([V], ctypes.c_int),
        "pf_steps_done": ([V, P(I64)], ctypes.c_int),
        "pf_kernel_info": ([V, P(I32), P(I32), P(I32)], ctypes.c_int),
        "pf_get_field": ([V, I32, V, I64], ctypes.c_int),
        "pf_get_ids": ([V, V, I64], ctypes.c_int),
        "pf_get_field_noalloc": ([V, I32, V, I64], ctypes.c_int),
        "pf_snapshot_open": ([V, I32], ctypes.c_int),
        "pf_snapshot": ([V, I32], ctypes.c_int),
        "pf_snapshot_wait": ([V, I32, P(P(D)), P(I64)], ctypes.c_int),
        "pf_snapshot_release": ([V, I32], ctypes.c_int),
        "pf_set_field": ([V, I32, V, I64], ctypes.c_int),
        "pf_get_probes": ([V, V, I64, P(I64)], ctypes.c_int),
        "pf_energy": ([V, P(D)], ctypes.c_int),
        "pf_profile": ([V, I64, P(D), P(D)], ctypes.c_int),
        "pf_batch_create": ([I32, P(V)], ctypes.c_int),
        "pf_batch_add": ([V, V], ctypes.c_int),
        "pf_batch_remove": ([V, V], ctypes.c_int),
        "pf_batch_step": ([V, I64], ctypes.c_int),
        "pf_batch_sync": ([V], ctypes.c_int),
        "pf_batch_set_tiles": ([V, I32, I32, I32, I32], ctypes.c_int),
        "pf_batch_autotune": ([V, I32, P(I32), P(I32), P(I32), P(D)], ctypes.c_int),
        "pf_batch_destroy": ([V], ctypes.c_int),
        "pf_stream_probe": ([I32, I32, I64, I32, P(D), P(D)], ctypes.c_int),
        "pf_dft": ([V, V, I64, I64, I64, V, D, V, V], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.pf_abi_version() != 2:
        raise FdtdUnavailable("libphonofdtd ABI %d, expected 2" % lib.pf_abi_version())
    if path is None:
        _lib = lib
    return lib


def check(rc):
    if rc != PF_OK:
        msg = load().pf_last_error()
        raise FdtdError(rc, msg.decode() if msg else "")
    return rc


def device_info():
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int32()
    gib = ctypes.c_double()
    check(load().pf_device_info(name, 256, ctypes.byref(cu), ctypes.byref(gib)))
    return {"name": name.value.decode(), "cu": cu.value, "hbm_gib": gib.value}


def stream_probe(reads, writes, n, reps=20):
    """Practical HBM bandwidth of a read/write mix: (GB/s, ms per pass)."""
    g, t = ctypes.c_double(), ctypes.c_double()
    check(load().pf_stream_probe(int(reads), int(writes), int(n), int(reps), ctypes.byref(g), ctypes.byref(t)))
    return g.value, t.value


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class MeshSpec:
    """Everything a context needs, as host arrays (reference layouts).

    nodes: (nx, ny, nz); fd/sd: per-axis spacing arrays; materials: (n, 13)
    rows C00 C01 C02 C10 C11 C12 C20 C21 C22 C33 C44 C55 rho; ids: uint8
    (nx, ny, nz); dt and dt2 (= dt**2 as the host computes it); mur: 12
    face coefficients; wave: FP64 source table (its length is the step
    budget); sponge: None or three per-axis damping arrays; probe:
    (i0, si, ni, j0, sj, nj) or None.  With ids None, raster = (x, y,
    targets) gives the node coordinates along x and y and the (n, 4)
    inclusion list (tx, ty, r, z nodes covered) the device rasterizes the
    material map from (problem.raster_targets).  surface: 0 for the
    reference's free surface (bit for bit), 1 for the one with the u_z
    update's precedence fixed (pf_desc.surface).
    """

    def __init__(self, nodes, fd, sd, materials, ids, dt, dt2, mur, wave,
                 sponge=None, probe=None, record_every=1, probe_capacity=1, raster=None, surface=0):
        self.nodes = tuple(int(n) for n in nodes)
        self.fd = [_c(a, np.float64) for a in fd]
        self.sd = [_c(a, np.float64) for a in sd]
        self.materials = _c(materials, np.float64).reshape(-1, 13)
        self.ids = None if ids is None else _c(ids, np.uint8)
        self.raster = None
        if self.ids is None:
            if raster is None:
                raise ValueError("ids or an inclusion list to rasterize is required")
            x, y, tg = raster
            self.raster = (_c(x, np.float64), _c(y, np.float64), _c(tg, np.float64).reshape(-1, 4))
        self.dt = float(dt)
        self.dt2 = float(dt2)
        self.mur = _c(mur, np.float64).reshape(12)
        self.wave = _c(wave, np.float64)
        self.sponge = None if sponge is None else [_c(a, np.float64) for a in sponge]
        self.probe = None if probe is None else tuple(int(v) for v in probe)
        self.record_every = int(record_every)
        self.probe_capacity = int(probe_capacity)
        if surface not in (0, 1):
            raise ValueError("surface must be 0 (the reference's) or 1 (precedence fixed)")
        self.surface = int(surface)
        nx, ny, nz = self.nodes
        if self.raster is not None:
            if self.raster[0].shape != (nx,) or self.raster[1].shape != (ny,):
                raise ValueError("node coordinates do not match nodes %s" % (self.nodes,))
        elif self.ids.shape != (nx, ny, nz):
            raise ValueError("ids shape %s != nodes %s" % (self.ids.shape, self.nodes))

    @property
    def cells(self):
        nx, ny, nz = self.nodes
        return nx * ny * nz


# tile shape a context takes without measuring (rows along y, lanes along z)
STATIC_TILES = {"fp64": (8, 32), "fp32": (16, 64), "fp32c": (16, 64)}


class Context:
    """One mesh resident on the GPU (a pf_ctx).

    With tile_j = tile_z = 0 the one-step sweep's tile shape is measured at
    creation (pf_create times each built shape on the new, all-zero fields);
    tune=False takes a fixed shape instead, for contexts that only step
    inside a batch (which has its own tiles)."""

    def __init__(self, spec, precision="fp64", kernel="auto", tile_j=0, tile_x=0, tile_z=0, tune=True):
        if not tune and not tile_j and not tile_z and kernel in ("auto", "tiled", "tiled2"):
            # the fixed shape of the tiled sweeps; the strip and naive
            # sweeps take their own defaults
            tile_j, tile_z = STATIC_TILES[precision]
        self.lib = load()
        self.spec = spec
        self.precision = precision
        nx, ny, nz = spec.nodes
        d = PfDesc()
        d.nx, d.ny, d.nz = nx, ny, nz
        d.precision = PRECISION[precision]
        d.kernel = KERNEL[kernel]
        d.n_materials = spec.materials.shape[0]
        keep = []

        def ptr(a):
            keep.append(a)
            return a.ctypes.data_as(ctypes.c_void_p)

        d.fdx, d.fdy, d.fdz = (ptr(a) for a in spec.fd)
        d.sdx, d.sdy, d.sdz = (ptr(a) for a in spec.sd)
        d.materials = ptr(spec.materials)
        if spec.ids is not None:
            d.ids = ptr(spec.ids)
        else:
            x, y, tg = spec.raster
            d.node_x, d.node_y = ptr(x), ptr(y)
            d.n_targets = tg.shape[0]
            d.targets = ptr(tg)
        d.dt, d.dt2 = spec.dt, spec.dt2
        d.mur = ptr(spec.mur)
        if spec.sponge is not None:
            d.sponge_x, d.sponge_y, d.sponge_z = (ptr(a) for a in spec.sponge)
        d.wave = ptr(spec.wave)
        d.wave_len = spec.wave.size
        if spec.probe is not None:
            for n, v in enumerate(spec.probe):
                d.probe[n] = v
            d.record_every = spec.record_every
            d.probe_capacity = spec.probe_capacity
        d.tile_j, d.tile_x, d.tile_z = int(tile_j), int(tile_x), int(tile_z)
        d.surface = getattr(spec, "surface", 0)
        h = ctypes.c_void_p()
        check(self.lib.pf_create(ctypes.byref(d), ctypes.byref(h)))
        self.handle = h
        self._shapes = {"ux": (nx - 1, ny, nz), "uy": (nx, ny - 1, nz), "uz": (nx, ny, nz - 1)}

    # -- life cycle
    def close(self):
        if getattr(self, "handle", None):
            check(self.lib.pf_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping
    def step(self, n):
        check(self.lib.pf_step(self.handle, int(n)))

    def sync(self):
        check(self.lib.pf_sync(self.handle))

    @property
    def steps_done(self):
        v = ctypes.c_int64()
        check(self.lib.pf_steps_done(self.handle, ctypes.byref(v)))
        return v.value

    def kernel_info(self):
        """(kernel name, tile rows along y, tile lanes along z) of the sweep
        this context runs, 'auto' resolved."""
        k, j, z = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(self.lib.pf_kernel_info(self.handle, ctypes.byref(k), ctypes.byref(j), ctypes.byref(z)))
        name = {v: n for n, v in KERNEL.items()}[k.value]
        return name, j.value, z.value

    # -- fields
    # -- asynchronous snapshots (pf_snapshot_*)
    def snapshot_open(self, slots):
        check(self.lib.pf_snapshot_open(self.handle, int(slots)))

    def snapshot(self, slot):
        """Queue a snapshot of the current level into ``slot``; returns at once."""
        check(self.lib.pf_snapshot(self.handle, int(slot)))

    def snapshot_wait(self, slot):
        """Wait for ``slot``'s copy; returns ({'ux','uy','uz'}: views of the
        pinned frame, valid until snapshot_release), steps done when taken)."""
        ptr = ctypes.POINTER(ctypes.c_double)()
        steps = ctypes.c_int64()
        check(self.lib.pf_snapshot_wait(self.handle, int(slot), ctypes.byref(ptr), ctypes.byref(steps)))
        out, off = {}, 0
        for c in ("ux", "uy", "uz"):
            shape = self._shapes[c]
            n = int(np.prod(shape))
            out[c] = np.ctypeslib.as_array(ptr, shape=(off + n,))[off:].reshape(shape)
            off += n
        return out, steps.value

    def snapshot_release(self, slot):
        check(self.lib.pf_snapshot_release(self.handle, int(slot)))

    def ids(self):
        """The device's material index map, (nx, ny, nz) uint8."""
        out = np.empty(self.spec.nodes, dtype=np.uint8)
        check(self.lib.pf_get_ids(self.handle, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def get(self, which):
        shape = self._shapes[which.replace("_old", "")]
        out = np.empty(shape, dtype=np.float64)
        check(self.lib.pf_get_field(self.handle, FIELDS[which], out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def get_noalloc(self, which):
        """As get(), staged in the library's static per-device array
        (pf_get_field_noalloc: no device allocation; fields of at most
        NOALLOC_FIELD_MAX elements)."""
        shape = self._shapes[which.replace("_old", "")]
        out = np.empty(shape, dtype=np.float64)
        check(self.lib.pf_get_field_noalloc(self.handle, FIELDS[which], out.ctypes.data_as(ctypes.c_void_p),
                                            out.size))
        return out

    def set(self, which, arr):
        shape = self._shapes[which.replace("_old", "")]
        a = _c(arr, np.float64)
        if a.shape != shape:
            raise ValueError("%s must have shape %s" % (which, shape))
        check(self.lib.pf_set_field(self.handle, FIELDS[which], a.ctypes.data_as(ctypes.c_void_p), a.size))

    def probes(self):
        p = self.spec.probe
        if p is None:
            return np.zeros((0, 0, 0))
        cap = self.spec.probe_capacity
        out = np.empty((cap, p[2], p[5]), dtype=np.float64)
        n = ctypes.c_int64()
        check(self.lib.pf_get_probes(self.handle, out.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(n)))
        return out[:n.value]

    def energy(self):
        v = ctypes.c_double()
        check(self.lib.pf_energy(self.handle, ctypes.byref(v)))
        return v.value

    def profile(self, n):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(self.lib.pf_profile(self.handle, int(n), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


def dft(x, nf=None, window=None, scale=1.0):
    """Batched 1-D DFT along the last axis on the GPU (pf_dft), first nf
    frequencies: scale * sum_t x[..., t] window[t] exp(-2 pi i k t / n)."""
    lib = load()
    x = np.asarray(x)
    n = x.shape[-1]
    nf = n if nf is None else int(nf)
    lead = x.shape[:-1]
    batch = int(np.prod(lead)) if lead else 1
    re = np.ascontiguousarray(x.real, dtype=np.float64).reshape(batch, n)
    im = np.ascontiguousarray(x.imag, dtype=np.float64).reshape(batch, n) if np.iscomplexobj(x) else None
    w = None if window is None else np.ascontiguousarray(window, dtype=np.float64)
    if w is not None and w.shape != (n,):
        raise ValueError("window must have %d values" % n)
    ore = np.empty((batch, nf))
    oim = np.empty((batch, nf))
    vp = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    for b0 in range(0, batch, 65535):  # one launch covers at most 65535 rows
        b1 = min(batch, b0 + 65535)
        check(lib.pf_dft(vp(re[b0:b1]), None if im is None else vp(im[b0:b1]), b1 - b0, n, nf, vp(w),
                         float(scale), vp(ore[b0:b1]), vp(oim[b0:b1])))
    return (ore + 1j * oim).reshape(lead + (nf,))


class Batch:
    """Many contexts of one precision stepped together (a pf_batch).

    Members keep their own step budgets; each member's fields and probes are
    bit-identical to stepping it alone.  While a context is a member, step it
    through the batch only.
    """

    def __init__(self, precision="fp32"):
        self.lib = load()
        self.precision = precision
        h = ctypes.c_void_p()
        check(self.lib.pf_batch_create(PRECISION[precision], ctypes.byref(h)))
        self.handle = h
        self.members = []

    def add(self, ctx):
        check(self.lib.pf_batch_add(self.handle, ctx.handle))
        self.members.append(ctx)

    def remove(self, ctx):
        check(self.lib.pf_batch_remove(self.handle, ctx.handle))
        self.members.remove(ctx)

    def step(self, n):
        check(self.lib.pf_batch_step(self.handle, int(n)))

    def sync(self):
        check(self.lib.pf_batch_sync(self.handle))

    def set_tiles(self, tile_j=0, tile_z=0, tile_x=0, graph=True):
        check(self.lib.pf_batch_set_tiles(self.handle, int(tile_j), int(tile_z), int(tile_x), 1 if graph else 0))

    def autotune(self, trial_steps=4):
        """Time every candidate tile shape on real steps; keep the fastest."""
        j, z, x, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        check(self.lib.pf_batch_autotune(self.handle, int(trial_steps), ctypes.byref(j), ctypes.byref(z),
                                         ctypes.byref(x), ctypes.byref(ms)))
        return {"tile_j": j.value, "tile_z": z.value, "tile_x": x.value, "ms_per_step": ms.value}

    def close(self):
        if getattr(self, "handle", None):
            check(self.lib.pf_batch_destroy(self.handle))
            self.handle = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
