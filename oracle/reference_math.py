"""CPU oracle: FP64 NumPy restatement of phonomena's FDTD time step.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this module, and only as
the checker or as the timed CPU baseline; the solver library never calls it.

What it restates (reference file:line, all under ``phonomena/simulation``):

* ``stress``            -> ``base_solver.py:323-372``  (``update_T``)
* ``stress_surface``    -> ``base_solver.py:402-433``  (``apply_T_tfbc``)
* ``displacement``      -> ``base_solver.py:435-463``  (``update_u``)
* ``displacement_surface`` -> ``base_solver.py:488-517`` (``apply_u_tfbc``)
* ``mur``               -> ``base_solver.py:519-554``  (``apply_u_abc``)
* ``rotate``            -> ``base_solver.py:556-571``  (``time_step``)
* ``step`` / ``run``    -> ``base_solver.py:245-260``  (loop body of ``run``)
* ``source``            -> ``base_solver.py:294-312``  (``update_sin`` / ``update_ricker``)
* ``spacings``          -> ``grid.py:112-127``
* ``dense_material``    -> ``material.py:55-63``      (``setConstants``)
* ``time_step_size``    -> ``material.py:80-93``      (``setTimeStep``)

Evaluation order is kept expression by expression (NumPy evaluates ``a*b/c``
as ``(a*b)/c`` and sums left to right), so on the same inputs this module
reproduces the reference bit for bit; the golden trajectories in
``tests/golden`` pin that.  The material arrays use the reference's own dense
layout (``C[nx,ny,nz,6,6]``, ``P[nx,ny,nz]``), so timing this module also
reproduces the reference's CPU cost model.

Extensions with no reference counterpart (pinned by their own tests, not by
the reference): the ``gauss`` source, the damping ``sponge`` layer, the
``energy`` audit and the fixed free surface (``Problem(surface_fixed=True)``:
the surface u_z update divides the T3 difference by fdz; tests/test_surface.py
pins it by its equality with the reference at unit spacing and its
invariance under grid scaling).  They are defined here once and the GPU path
is checked against these definitions.
"""
import numpy as np


# ---------------------------------------------------------------- set-up ----

def spacings(x, y, z, si=1.0):
    """Full (fd, n-1) and staggered (sd, n-2) spacings shaped for broadcasting.

    grid.py:112-127: fd = diff of SI coordinates; sd = np.mean of adjacent fd.
    """
    xs, ys, zs = x * si, y * si, z * si
    fdx = (xs[1:] - xs[:-1]).reshape((-1, 1, 1))
    fdy = (ys[1:] - ys[:-1]).reshape((1, -1, 1))
    fdz = (zs[1:] - zs[:-1]).reshape((1, 1, -1))
    sdx = np.mean([fdx[1:, :, :], fdx[:-1, :, :]], axis=0)
    sdy = np.mean([fdy[:, 1:, :], fdy[:, :-1, :]], axis=0)
    sdz = np.mean([fdz[:, :, 1:], fdz[:, :, :-1]], axis=0)
    return dict(fdx=fdx, fdy=fdy, fdz=fdz, sdx=sdx, sdy=sdy, sdz=sdz)


def dense_material(ids, c_table, p_table):
    """Per-node 6x6 stiffness and density from a material-index map.

    material.py:55-63 fills every node with the primary material and then the
    inclusion nodes with the secondary one; ``ids`` holds that choice per node
    (0 = primary, 1 = secondary, more entries allowed).
    """
    c_table = np.asarray(c_table, dtype=np.float64)
    p_table = np.asarray(p_table, dtype=np.float64)
    C = c_table[ids]
    P = p_table[ids]
    return C, P


def time_step_size(materials, fd, c_max, si=1.0):
    """CFL time step, material.py:80-93 (incl. the second SI factor)."""
    dts = []
    for c, p in materials:
        vl = np.sqrt(c[0][0] / p)
        vt = np.sqrt(c[3][3] / p)
        vmax = max((vl, vt))
        dxmin = min((np.amin(fd[0]), np.amin(fd[1]), np.amin(fd[2]))) * si
        dts.append(c_max * dxmin / vmax)
    return min(dts)


def source(kind, tt, dt, f, source_delay=0.0, width=None):
    """Wall source sample w(tt).

    ``sin`` and ``ricker`` follow base_solver.py:294-312.  ``gauss`` is an
    extension: a Gaussian-modulated sinusoid
    sin(2 pi f (t - t0)) * exp(-((t - t0)/s)^2 / 2) with s = 1/f and
    t0 = 4 s unless given.
    """
    if kind == "sin":
        return np.sin(2 * np.pi * f * tt * dt)
    if kind == "ricker":
        arg = (np.pi * f * (dt * tt - source_delay)) ** 2
        return (1 - 2 * arg) * np.exp(-arg)
    if kind == "gauss":
        s = (1.0 / f) if width is None else width
        t0 = 4.0 * s if not source_delay else source_delay
        t = tt * dt - t0
        return np.sin(2 * np.pi * f * t) * np.exp(-0.5 * (t / s) ** 2)
    raise ValueError(kind)


def source_table(kind, steps, dt, args):
    """w(tt) for tt = 0..steps-1 in FP64, one scalar call per step."""
    a = dict(args)
    f = a.pop("f")
    return np.array([source(kind, tt, dt, f, **a) for tt in range(steps)], dtype=np.float64)


def mur_coefficients(C, P, dt, sp):
    """Face coefficients of the first-order Mur condition, base_solver.py:525-537.

    Only node (0,0,0) sets the wave speeds.  Returned as the 12 scalars the
    four faces use: {face: (c_ux, c_uy, c_uz)}.
    """
    c11 = C[0, 0, 0, 0, 0]
    c44 = C[0, 0, 0, 3, 3]
    vl = np.sqrt(c11 / P[0, 0, 0])
    vt = np.sqrt(c44 / P[0, 0, 0])
    ctx = ((vt * dt - sp["fdx"]) / (vt * dt + sp["fdx"]))[:, 0, 0]
    clx = ((vl * dt - sp["sdx"]) / (vl * dt + sp["sdx"]))[:, 0, 0]
    cty = ((vt * dt - sp["fdy"]) / (vt * dt + sp["fdy"]))[0, :, 0]
    cly = ((vl * dt - sp["sdy"]) / (vl * dt + sp["sdy"]))[0, :, 0]
    ctz = ((vt * dt - sp["fdz"]) / (vt * dt + sp["fdz"]))[0, 0, :]
    clz = ((vl * dt - sp["sdz"]) / (vl * dt + sp["sdz"]))[0, 0, :]
    return {
        "xmax": (clx[-1], ctx[-1], ctx[-1]),
        "y0": (cty[0], cly[0], cty[0]),
        "ymax": (cty[-1], cly[-1], cty[-1]),
        "zmax": (ctz[-1], ctz[-1], clz[-1]),
    }


def sponge_profile(n, width, smax, lo, hi):
    """Damping term s = sigma*dt/2 per node index along one axis (extension).

    s(d) = smax * ((width - d)/width)^2 for a node d < width cells inside a
    damped face; ``lo``/``hi`` select the faces at index 0 and n-1.
    """
    s = np.zeros(n, dtype=np.float64)
    if width <= 0:
        return s
    for i in range(n):
        v = 0.0
        if lo and i < width:
            r = (width - i) / width
            v += smax * r * r
        if hi and (n - 1 - i) < width:
            r = (width - (n - 1 - i)) / width
            v += smax * r * r
        s[i] = v
    return s


# ----------------------------------------------------------------- state ----

class State:
    """Field arrays of grid.py:83-110 (copy-rotation semantics kept)."""

    def __init__(self, nx, ny, nz):
        self.shape = (nx, ny, nz)
        shp = {"x": (nx - 1, ny, nz), "y": (nx, ny - 1, nz), "z": (nx, ny, nz - 1)}
        for c in "xyz":
            for suf in ("", "_new", "_old", "_temp"):
                setattr(self, "u" + c + suf, np.zeros(shp[c]))
        self.T1 = np.zeros((nx, ny, nz))
        self.T2 = np.zeros((nx, ny, nz))
        self.T3 = np.zeros((nx, ny, nz))
        self.T4 = np.zeros((nx, ny - 1, nz - 1))
        self.T5 = np.zeros((nx - 1, ny, nz - 1))
        self.T6 = np.zeros((nx - 1, ny - 1, nz))


# ------------------------------------------------------------ the step ----

def _row_stress(C, a, dux, duy, duz, sx, sy, sz):
    # C[a,0]*dux/sx + C[a,1]*duy/sy + C[a,2]*duz/sz, left to right.
    return C[..., a, 0] * dux / sx + C[..., a, 1] * duy / sy + C[..., a, 2] * duz / sz


def stress(s, C, sp):
    """Interior stresses, base_solver.py:329-372."""
    Ci = C[1:-1, 1:-1, 1:-1]
    dux = s.ux[1:, 1:-1, 1:-1] - s.ux[:-1, 1:-1, 1:-1]
    duy = s.uy[1:-1, 1:, 1:-1] - s.uy[1:-1, :-1, 1:-1]
    duz = s.uz[1:-1, 1:-1, 1:] - s.uz[1:-1, 1:-1, :-1]
    for a, T in enumerate((s.T1, s.T2, s.T3)):
        T[1:-1, 1:-1, 1:-1] = _row_stress(Ci, a, dux, duy, duz, sp["sdx"], sp["sdy"], sp["sdz"])
    s.T4[1:-1, :, :] = C[1:-1, 1:, 1:, 3, 3] * (
        (s.uy[1:-1, :, 1:] - s.uy[1:-1, :, :-1]) / sp["fdz"]
        + (s.uz[1:-1, 1:, :] - s.uz[1:-1, :-1, :]) / sp["fdy"])
    s.T5[:, 1:-1, :] = C[1:, 1:-1, 1:, 4, 4] * (
        (s.ux[:, 1:-1, 1:] - s.ux[:, 1:-1, :-1]) / sp["fdz"]
        + (s.uz[1:, 1:-1, :] - s.uz[:-1, 1:-1, :]) / sp["fdx"])
    s.T6[:, :, 1:-1] = C[1:, 1:, 1:-1, 5, 5] * (
        (s.ux[:, 1:, 1:-1] - s.ux[:, :-1, 1:-1]) / sp["fdy"]
        + (s.uy[1:, :, 1:-1] - s.uy[:-1, :, 1:-1]) / sp["fdx"])


def stress_surface(s, C, sp):
    """Traction-free plane k=0, base_solver.py:408-433.

    Only the first spacing of each axis is used, and the shear terms divide
    by the spacings in the order the reference writes them.
    """
    sx0 = sp["sdx"][0, :, :]
    sy0 = sp["sdy"][:, 0, :]
    sz0 = sp["sdz"][:, :, 0]
    fx0 = sp["fdx"][0, :, :]
    fy0 = sp["fdy"][:, 0, :]
    fz0 = sp["fdz"][:, :, 0]
    Cs = C[1:-1, 1:-1, 0]
    dux = s.ux[1:, 1:-1, 0] - s.ux[:-1, 1:-1, 0]
    duy = s.uy[1:-1, 1:, 0] - s.uy[1:-1, :-1, 0]
    duz = s.uz[1:-1, 1:-1, 0] - 0
    s.T1[1:-1, 1:-1, 0] = _row_stress(Cs, 0, dux, duy, duz, sx0, sy0, sz0)
    s.T2[1:-1, 1:-1, 0] = _row_stress(Cs, 1, dux, duy, duz, sx0, sy0, sz0)
    s.T3[1:-1, 1:-1, 0] = 0
    s.T4[1:-1, :, 0] = C[1:-1, 1:, 0, 3, 3] * (
        (s.uy[1:-1, :, 1] - s.uy[1:-1, :, 0]) / fy0
        + (s.uz[1:-1, 1:, 0] - s.uz[1:-1, :-1, 0]) / fz0)
    s.T5[:, 1:-1, 0] = C[1:, 1:-1, 0, 4, 4] * (
        (s.ux[:, 1:-1, 1] - s.ux[:, 1:-1, 0]) / fx0
        + (s.uz[1:, 1:-1, 0] - s.uz[:-1, 1:-1, 0]) / fz0)
    s.T6[:, :, 0] = C[1:, 1:, 0, 5, 5] * (
        (s.ux[:, 1:, 0] - s.ux[:, :-1, 0]) / fx0
        + (s.uy[1:, :, 0] - s.uy[:-1, :, 0]) / fz0)


def _leap(u, u_old, coef, div, damp):
    # 2*u - u_old + coef*div (base_solver.py:441-463); with a sponge the
    # extension form ((2u - (1-s) u_old) + coef*div) / (1+s).
    if damp is None:
        return 2 * u - u_old + coef * div
    return (2 * u - (1 - damp) * u_old + coef * div) / (1 + damp)


def _sl(damp, comp, sl):
    return None if damp is None else damp[comp][sl]


def displacement(s, P, dt, sp, damp=None):
    """Interior displacements, base_solver.py:441-463.

    ``damp`` (extension) maps component -> per-cell sponge term shaped like
    that component's array; None keeps the reference arithmetic exactly.
    """
    D = dt ** 2
    sl = (slice(None), slice(1, -1), slice(1, -1))
    s.ux_new[sl] = _leap(s.ux[sl], s.ux_old[sl], D / P[1:, 1:-1, 1:-1],
                         (s.T1[1:, 1:-1, 1:-1] - s.T1[:-1, 1:-1, 1:-1]) / sp["fdx"]
                         + (s.T6[:, 1:, 1:-1] - s.T6[:, :-1, 1:-1]) / sp["sdy"]
                         + (s.T5[:, 1:-1, 1:] - s.T5[:, 1:-1, :-1]) / sp["sdz"],
                         _sl(damp, "x", sl))
    sl = (slice(1, -1), slice(None), slice(1, -1))
    s.uy_new[sl] = _leap(s.uy[sl], s.uy_old[sl], D / P[1:-1, 1:, 1:-1],
                         (s.T6[1:, :, 1:-1] - s.T6[:-1, :, 1:-1]) / sp["sdx"]
                         + (s.T2[1:-1, 1:, 1:-1] - s.T2[1:-1, :-1, 1:-1]) / sp["fdy"]
                         + (s.T4[1:-1, :, 1:] - s.T4[1:-1, :, :-1]) / sp["sdz"],
                         _sl(damp, "y", sl))
    sl = (slice(1, -1), slice(1, -1), slice(None))
    s.uz_new[sl] = _leap(s.uz[sl], s.uz_old[sl], D / P[1:-1, 1:-1, 1:],
                         (s.T5[1:, 1:-1, :] - s.T5[:-1, 1:-1, :]) / sp["sdx"]
                         + (s.T4[1:-1, 1:, :] - s.T4[1:-1, :-1, :]) / sp["sdy"]
                         + (s.T3[1:-1, 1:-1, 1:] - s.T3[1:-1, 1:-1, :-1]) / sp["fdz"],
                         _sl(damp, "z", sl))


def displacement_surface(s, P, dt, sp, damp=None, fixed=False):
    """Displacements on the free surface k=0, base_solver.py:495-517.

    The u_z line keeps the reference's operator precedence: T3[...,1] enters
    unscaled and only T3[...,0] (always 0) is divided by fdz.  ``fixed``
    (extension, not the reference) divides the difference instead,
    (T3[...,1] - T3[...,0]) / fdz, as the interior u_z update does
    (base_solver.py:461-463); with fdz = 1 both give the same bits.
    """
    D = dt ** 2
    sx0 = sp["sdx"][0, :, :]
    sy0 = sp["sdy"][:, 0, :]
    sz0 = sp["sdz"][:, :, 0]
    fx0 = sp["fdx"][0, :, :]
    fy0 = sp["fdy"][:, 0, :]
    fz0 = sp["fdz"][:, :, 0]
    s.ux_new[:, 1:-1, 0] = _leap(
        s.ux[:, 1:-1, 0], s.ux_old[:, 1:-1, 0], D / P[1:, 1:-1, 0],
        (s.T1[1:, 1:-1, 0] - s.T1[:-1, 1:-1, 0]) / fx0
        + (s.T6[:, 1:, 0] - s.T6[:, :-1, 0]) / sy0
        + (s.T5[:, 1:-1, 0] - 0) / sz0,
        _sl(damp, "x", (slice(None), slice(1, -1), 0)))
    s.uy_new[1:-1, :, 0] = _leap(
        s.uy[1:-1, :, 0], s.uy_old[1:-1, :, 0], D / P[1:-1, 1:, 0],
        (s.T6[1:, :, 0] - s.T6[:-1, :, 0]) / sx0
        + (s.T2[1:-1, 1:, 0] - s.T2[1:-1, :-1, 0]) / fy0
        + (s.T4[1:-1, :, 0] - 0) / sz0,
        _sl(damp, "y", (slice(1, -1), slice(None), 0)))
    s.uz_new[1:-1, 1:-1, 0] = _leap(
        s.uz[1:-1, 1:-1, 0], s.uz_old[1:-1, 1:-1, 0], D / P[1:-1, 1:-1, 0],
        (s.T5[1:, 1:-1, 0] - s.T5[:-1, 1:-1, 0]) / sx0
        + (s.T4[1:-1, 1:, 0] - s.T4[1:-1, :-1, 0]) / sy0
        + ((s.T3[1:-1, 1:-1, 1] - s.T3[1:-1, 1:-1, 0]) / fz0 if fixed
           else s.T3[1:-1, 1:-1, 1] - s.T3[1:-1, 1:-1, 0] / fz0),
        _sl(damp, "z", (slice(1, -1), slice(1, -1), 0)))


def mur(s, mc):
    """First-order Mur faces in the reference's order, base_solver.py:539-554.

    Each face reads the plane one cell inside (as already updated this step)
    and overwrites the edge cells left by earlier faces.
    """
    cx, cy, cz = mc["xmax"]
    s.ux_new[-1, :, :] = s.ux[-2, :, :] + cx * (s.ux_new[-2, :, :] - s.ux[-1, :, :])
    s.uy_new[-1, :, :] = s.uy[-2, :, :] + cy * (s.uy_new[-2, :, :] - s.uy[-1, :, :])
    s.uz_new[-1, :, :] = s.uz[-2, :, :] + cz * (s.uz_new[-2, :, :] - s.uz[-1, :, :])
    cx, cy, cz = mc["y0"]
    s.ux_new[:, 0, :] = s.ux[:, 1, :] + cx * (s.ux_new[:, 1, :] - s.ux[:, 0, :])
    s.uy_new[:, 0, :] = s.uy[:, 1, :] + cy * (s.uy_new[:, 1, :] - s.uy[:, 0, :])
    s.uz_new[:, 0, :] = s.uz[:, 1, :] + cz * (s.uz_new[:, 1, :] - s.uz[:, 0, :])
    cx, cy, cz = mc["ymax"]
    s.ux_new[:, -1, :] = s.ux[:, -2, :] + cx * (s.ux_new[:, -2, :] - s.ux[:, -1, :])
    s.uy_new[:, -1, :] = s.uy[:, -2, :] + cy * (s.uy_new[:, -2, :] - s.uy[:, -1, :])
    s.uz_new[:, -1, :] = s.uz[:, -2, :] + cz * (s.uz_new[:, -2, :] - s.uz[:, -1, :])
    cx, cy, cz = mc["zmax"]
    s.ux_new[:, :, -1] = s.ux[:, :, -2] + cx * (s.ux_new[:, :, -2] - s.ux[:, :, -1])
    s.uy_new[:, :, -1] = s.uy[:, :, -2] + cy * (s.uy_new[:, :, -2] - s.uy[:, :, -1])
    s.uz_new[:, :, -1] = s.uz[:, :, -2] + cz * (s.uz_new[:, :, -2] - s.uz[:, :, -1])


def rotate(s):
    """Copy rotation, base_solver.py:561-571: old <- u, u <- new (new kept)."""
    for c in "xyz":
        getattr(s, "u" + c + "_temp")[...] = getattr(s, "u" + c)
        getattr(s, "u" + c)[...] = getattr(s, "u" + c + "_new")
        getattr(s, "u" + c + "_old")[...] = getattr(s, "u" + c + "_temp")


def sponge_fields(shape, prof):
    """Per-cell damping term for each component array (extension)."""
    nx, ny, nz = shape
    sx, sy, sz = prof
    full = sx[:, None, None] + sy[None, :, None] + sz[None, None, :]
    return {"x": full[:nx - 1, :, :], "y": full[:, :ny - 1, :], "z": full[:, :, :nz - 1]}


class Problem:
    """Everything one run needs, in the reference's layouts."""

    def __init__(self, x, y, z, C, P, dt, wave, mur_c=None, sponge=None, si=1.0, surface_fixed=False):
        self.x, self.y, self.z = x, y, z
        self.surface_fixed = surface_fixed
        self.sp = spacings(x, y, z, si)
        self.C, self.P, self.dt = C, P, dt
        self.wave = np.asarray(wave, dtype=np.float64)
        self.mur_c = mur_c if mur_c is not None else mur_coefficients(C, P, dt, self.sp)
        self.shape = (x.size, y.size, z.size)
        self.damp = None
        if sponge is not None:
            d = sponge_fields(self.shape, sponge)
            self.damp = d


def run(pb, steps, probe=None, record_every=0, state=None):
    """Step ``steps`` times; optionally record surface-u_z probe frames.

    ``probe`` = (i0, si, ni, j0, sj, nj) picks uz[i0 + si*a, j0 + sj*b, 0]
    after the rotation of every step with (tt+1) % record_every == 0.
    """
    s = state if state is not None else State(*pb.shape)
    frames = []
    for tt in range(steps):
        step(s, pb, tt)
        if probe is not None and record_every and (tt + 1) % record_every == 0:
            frames.append(probe_frame(s, probe))
    return s, (np.array(frames) if probe is not None else None)


def step(s, pb, tt):
    """One time step, base_solver.py:251-256 (sponge: extension)."""
    s.uz[0, :, 0] = pb.wave[tt]
    stress(s, pb.C, pb.sp)
    stress_surface(s, pb.C, pb.sp)
    displacement(s, pb.P, pb.dt, pb.sp, pb.damp)
    displacement_surface(s, pb.P, pb.dt, pb.sp, pb.damp, pb.surface_fixed)
    mur(s, pb.mur_c)
    rotate(s)


def probe_frame(s, probe):
    i0, si_, ni, j0, sj, nj = probe
    return s.uz[i0:i0 + si_ * ni:si_, j0:j0 + sj * nj:sj, 0].copy()


# ---------------------------------------------------------------- energy ----

def energy(s, pb, c_comp=None):
    """Energy audit (extension), summed in FP64.

    E = dV/2 * [ sum_q sum_cells rho_q * ((u_q - u_q_old)/dt)^2
               + sum over stress sites of T(u) : strain(u_old) ]
    with u = the newest level (after rotation) and u_old the previous one.
    The mixed-level strain term makes E the quantity the leapfrog scheme
    conserves (E^{n+1/2}); it is exactly constant without source and
    boundary flux.  rho_q is the density at the node the update of u_q
    samples (ux: i+1, uy: j+1, uz: k+1); stresses are those of
    ``stress``/``stress_surface``; each strain is the bracket its stress
    multiplies.  ``c_comp`` optionally adds per-component terms to the
    increment.  dV is the product of the mean full spacings.
    """
    sp, P, C, dt = pb.sp, pb.P, pb.C, pb.dt
    t = State(*pb.shape)
    t.ux[...] = s.ux
    t.uy[...] = s.uy
    t.uz[...] = s.uz
    stress(t, C, sp)
    stress_surface(t, C, sp)
    o = State(*pb.shape)
    o.ux[...] = s.ux_old
    o.uy[...] = s.uy_old
    o.uz[...] = s.uz_old
    vx = s.ux - s.ux_old
    vy = s.uy - s.uy_old
    vz = s.uz - s.uz_old
    if c_comp is not None:
        vx = vx + c_comp[0]
        vy = vy + c_comp[1]
        vz = vz + c_comp[2]
    # The x = 0 wall of uy and uz is clamped/driven, not a dynamic degree of
    # freedom, and is left out of the kinetic term.
    ke = (np.sum(P[1:, :, :] * (vx / dt) ** 2) + np.sum(P[1:, 1:, :] * (vy[1:] / dt) ** 2)
          + np.sum(P[1:, :, 1:] * (vz[1:] / dt) ** 2))
    e = strains(o, sp)
    se = (np.sum(t.T1 * e["e11"]) + np.sum(t.T2 * e["e22"]) + np.sum(t.T3 * e["e33"])
          + np.sum(t.T4 * e["g23"]) + np.sum(t.T5 * e["g13"]) + np.sum(t.T6 * e["g12"]))
    dV = np.mean(sp["fdx"]) * np.mean(sp["fdy"]) * np.mean(sp["fdz"])
    return 0.5 * dV * (ke + se)


def strains(t, sp):
    """The strain brackets that ``stress``/``stress_surface`` multiply by C.

    Zero wherever the matching stress is never written.
    """
    nx, ny, nz = t.shape
    e11 = np.zeros((nx, ny, nz))
    e22 = np.zeros((nx, ny, nz))
    e33 = np.zeros((nx, ny, nz))
    e11[1:-1, 1:-1, 1:-1] = (t.ux[1:, 1:-1, 1:-1] - t.ux[:-1, 1:-1, 1:-1]) / sp["sdx"]
    e22[1:-1, 1:-1, 1:-1] = (t.uy[1:-1, 1:, 1:-1] - t.uy[1:-1, :-1, 1:-1]) / sp["sdy"]
    e33[1:-1, 1:-1, 1:-1] = (t.uz[1:-1, 1:-1, 1:] - t.uz[1:-1, 1:-1, :-1]) / sp["sdz"]
    e11[1:-1, 1:-1, 0] = (t.ux[1:, 1:-1, 0] - t.ux[:-1, 1:-1, 0]) / sp["sdx"][0, :, :]
    e22[1:-1, 1:-1, 0] = (t.uy[1:-1, 1:, 0] - t.uy[1:-1, :-1, 0]) / sp["sdy"][:, 0, :]
    e33[1:-1, 1:-1, 0] = (t.uz[1:-1, 1:-1, 0] - 0) / sp["sdz"][:, :, 0]
    g23 = np.zeros(t.T4.shape)
    g13 = np.zeros(t.T5.shape)
    g12 = np.zeros(t.T6.shape)
    g23[1:-1, :, :] = ((t.uy[1:-1, :, 1:] - t.uy[1:-1, :, :-1]) / sp["fdz"]
                       + (t.uz[1:-1, 1:, :] - t.uz[1:-1, :-1, :]) / sp["fdy"])
    g13[:, 1:-1, :] = ((t.ux[:, 1:-1, 1:] - t.ux[:, 1:-1, :-1]) / sp["fdz"]
                       + (t.uz[1:, 1:-1, :] - t.uz[:-1, 1:-1, :]) / sp["fdx"])
    g12[:, :, 1:-1] = ((t.ux[:, 1:, 1:-1] - t.ux[:, :-1, 1:-1]) / sp["fdy"]
                       + (t.uy[1:, :, 1:-1] - t.uy[:-1, :, 1:-1]) / sp["fdx"])
    g23[1:-1, :, 0] = ((t.uy[1:-1, :, 1] - t.uy[1:-1, :, 0]) / sp["fdy"][:, 0, :]
                       + (t.uz[1:-1, 1:, 0] - t.uz[1:-1, :-1, 0]) / sp["fdz"][:, :, 0])
    g13[:, 1:-1, 0] = ((t.ux[:, 1:-1, 1] - t.ux[:, 1:-1, 0]) / sp["fdx"][0, :, :]
                       + (t.uz[1:, 1:-1, 0] - t.uz[:-1, 1:-1, 0]) / sp["fdz"][:, :, 0])
    g12[:, :, 0] = ((t.ux[:, 1:, 0] - t.ux[:, :-1, 0]) / sp["fdx"][0, :, :]
                    + (t.uy[1:, :, 0] - t.uy[:-1, :, 0]) / sp["fdz"][:, :, 0])
    return dict(e11=e11, e22=e22, e33=e33, g23=g23, g13=g13, g12=g12)
