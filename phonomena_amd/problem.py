"""Host-side assembly of a device problem from a Grid, a Material and a cfg.

Everything here is set-up arithmetic the reference does on the host before
or inside its step loop, restated so the device receives the same FP64
numbers: spacings (grid.py:112-127), dt**2 (base_solver.py:443), the Mur
face coefficients (base_solver.py:525-537) and the source table
(base_solver.py:251, 294-312).  Extensions: the sponge profile and the
probe grid.
"""
import numpy as np

from .fdtd import MeshSpec

# Faces that carry the Mur condition in the reference (and the sponge here):
# x = max, y = 0, y = max, z = max.  No absorber at x = 0 (source wall) or
# z = 0 (free surface).
SPONGE_FACES = {"x": (False, True), "y": (True, True), "z": (False, True)}
DEFAULT_SPONGE_STRENGTH = 0.15


def mur_coefficients(row, dt, fd, sd):
    """12 Mur coefficients from the material at node (0,0,0).

    ``row`` is that node's 13-entry material row; ``fd``/``sd`` the 1-D
    spacing arrays per axis.  Order: faces x=max, y=0, y=max, z=max, each
    (ux, uy, uz); longitudinal speed for the face-normal component,
    transverse for the others (base_solver.py:525-554).
    """
    c11, c44, p = row[0], row[9], row[12]
    vl = np.sqrt(c11 / p)
    vt = np.sqrt(c44 / p)
    fdx, fdy, fdz = fd
    sdx, sdy, sdz = sd
    ctx = (vt * dt - fdx) / (vt * dt + fdx)
    clx = (vl * dt - sdx) / (vl * dt + sdx)
    cty = (vt * dt - fdy) / (vt * dt + fdy)
    cly = (vl * dt - sdy) / (vl * dt + sdy)
    ctz = (vt * dt - fdz) / (vt * dt + fdz)
    clz = (vl * dt - sdz) / (vl * dt + sdz)
    return np.array([clx[-1], ctx[-1], ctx[-1],
                     cty[0], cly[0], cty[0],
                     cty[-1], cly[-1], cty[-1],
                     ctz[-1], ctz[-1], clz[-1]], dtype=np.float64)


def sponge_axis(n, width, strength, lo, hi):
    """Damping term s = sigma*dt/2 per node along one axis (extension).

    s(d) = strength * ((width - d) / width)^2 for a node d < width cells from
    a damped face; the step becomes
    u_new = ((2u - (1-s) u_old) + coef*div) / (1 + s).
    """
    s = np.zeros(n, dtype=np.float64)
    if width <= 0:
        return s
    for i in range(n):
        v = 0.0
        if lo and i < width:
            r = (width - i) / width
            v += strength * r * r
        if hi and (n - 1 - i) < width:
            r = (width - (n - 1 - i)) / width
            v += strength * r * r
        s[i] = v
    return s


def sponge_tables(nodes, sponge):
    """Per-axis damping tables from cfg {'width': cells, 'strength': s}."""
    if not sponge:
        return None
    width = int(sponge.get("width", 0))
    strength = float(sponge.get("strength", DEFAULT_SPONGE_STRENGTH))
    return [sponge_axis(n, width, strength, *SPONGE_FACES[a]) for n, a in zip(nodes, "xyz")]


def axis_spacings(g):
    """1-D full and staggered spacings of a Grid (grid.py:117-127)."""
    fd = [np.ascontiguousarray(a.ravel()) for a in (g.fdx, g.fdy, g.fdz)]
    sd = [np.ascontiguousarray(a.ravel()) for a in (g.sdx, g.sdy, g.sdz)]
    return fd, sd


def probe_grid(spec, nodes):
    """Normalise a probe spec to (i0, si, ni, j0, sj, nj) on the u_z surface.

    ``spec`` is None, a 6-sequence, or {'stride': s} / {'count': n} for an
    evenly strided grid covering the surface.
    """
    if spec is None:
        return None
    nx, ny, _ = nodes
    if isinstance(spec, dict):
        if "count" in spec:
            n = int(spec["count"])
            si, sj = max(1, nx // n), max(1, ny // n)
            ni, nj = min(n, (nx - 1) // si + 1), min(n, (ny - 1) // sj + 1)
        else:
            s = int(spec.get("stride", 1))
            si = sj = s
            ni, nj = (nx - 1) // s + 1, (ny - 1) // s + 1
        return (0, si, ni, 0, sj, nj)
    p = tuple(int(v) for v in spec)
    if len(p) != 6:
        raise ValueError("probe must be (i0, si, ni, j0, sj, nj)")
    return p


def raster_targets(g):
    """(x, y, targets) for the device rasterizer: node coordinates along x
    and y, and one row (tx, ty, r, K) per inclusion of the grid, K = the
    number of leading z nodes the inclusion covers.

    K follows the z test of Grid.inclusionIndices (grid.py:158-175): the
    first target compares the z coordinates with its depth, and every later
    one compares the index array the previous test returned (the reference
    rebinds ``z``).  Each selection is a leading run of nodes, so a count
    describes it."""
    z = g.z
    rows = []
    for t in g.targets:
        idx = np.array(np.where(z <= t['z'])).flatten()
        if not np.array_equal(idx, np.arange(idx.size)):
            raise ValueError("z selection of an inclusion is not a leading run of nodes")
        rows.append((float(t['x']), float(t['y']), float(t['r']), float(idx.size)))
        z = idx
    tg = np.array(rows, dtype=np.float64).reshape(-1, 4)
    return np.ascontiguousarray(g.x, np.float64), np.ascontiguousarray(g.y, np.float64), tg


def raster_id(x, y, k, targets):
    """Material index of node (x, y, k) under the rasterizer's rule (host)."""
    for tx, ty, r, kk in targets:
        X = x - tx
        Y = y - ty
        if np.sqrt(np.add(np.square(X), np.square(Y))) < r and k < kk:
            return 1
    return 0


MAX_MATERIALS = 256
# entries of the 6x6 stiffness the stencils read (base_solver.py:329-372):
# the normal 3x3 block, then the shear diagonal C33 C44 C55
_ROW_ENTRIES = [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2), (2, 0), (2, 1), (2, 2), (3, 3), (4, 4), (5, 5)]


def dense_rows(C, P):
    """Per-node 13-entry rows (C00 .. C22, C33, C44, C55, rho) of the
    reference's dense arrays C (nx, ny, nz, 6, 6) and P (nx, ny, nz), as
    Material.update leaves them (material.py:48-63)."""
    C = np.asarray(C, dtype=np.float64)
    P = np.asarray(P, dtype=np.float64)
    if C.ndim != 5 or C.shape[3:] != (6, 6) or P.shape != C.shape[:3]:
        raise ValueError("material C must be (nx, ny, nz, 6, 6) and P (nx, ny, nz); got %s and %s"
                         % (C.shape, P.shape))
    cols = [C[..., a, b] for a, b in _ROW_ENTRIES] + [P]
    return np.stack(cols, axis=-1)


def material_table(m, nodes):
    """(rows, ids) of a Material: ``rows`` (n, 13) float64 and ``ids``
    (nx, ny, nz) uint8 with ids[i, j, k] the row of node (i, j, k).

    This package's Material carries them (``rows()``, ``ids``).  The
    reference's Material (phonomena/simulation/material.py:22-93) holds the
    dense per-node arrays ``C`` (nx, ny, nz, 6, 6) and ``P`` instead; then
    the rows are the distinct 13-entry rows of its nodes (every node holds
    the primary or the secondary material after setConstants,
    material.py:55-63, but any per-node assignment of at most 256 distinct
    materials is taken as it stands), in the order primary, secondary, then
    the others by first appearance.  The values are copied bit for bit.
    """
    ids = getattr(m, "ids", None)
    if callable(getattr(m, "rows", None)) and ids is not None and tuple(np.shape(ids)) == tuple(nodes):
        return m.rows(), ids
    C, P = getattr(m, "C", None), getattr(m, "P", None)
    if C is None or P is None:
        raise ValueError("material has neither rows()/ids nor the dense C/P arrays of material.py:48-63")
    per_node = dense_rows(C, P)
    if per_node.shape[:3] != tuple(nodes):
        raise ValueError("material arrays are %s, the grid has %s nodes (call Material.update after "
                         "Grid.update)" % (per_node.shape[:3], tuple(nodes)))
    flat = per_node.reshape(-1, 13)
    # distinct rows by their bytes (exact equality, no tolerance)
    keys = np.ascontiguousarray(flat).view(np.dtype((np.void, 13 * 8))).ravel()
    uniq, first, inverse = np.unique(keys, return_index=True, return_inverse=True)
    if uniq.size > MAX_MATERIALS:
        raise ValueError("material map has %d distinct materials (at most %d)" % (uniq.size, MAX_MATERIALS))
    prefer = []
    for name in ("primary", "secondary"):
        mat = getattr(m, name, None)
        if mat is not None and "c" in mat and "p" in mat:
            c = np.asarray(mat["c"], dtype=np.float64)
            row = np.array([c[a, b] for a, b in _ROW_ENTRIES] + [float(mat["p"])], dtype=np.float64)
            hit = np.nonzero(keys[first] == row.view(np.dtype((np.void, 13 * 8)))[0])[0]
            if hit.size:
                prefer.append(int(hit[0]))
    order = []
    for u in prefer + list(np.argsort(first, kind="stable")):
        if u not in order:
            order.append(u)
    remap = np.empty(uniq.size, dtype=np.uint8)
    remap[np.array(order, dtype=np.int64)] = np.arange(len(order), dtype=np.uint8)
    rows = np.ascontiguousarray(flat[first[np.array(order, dtype=np.int64)]])
    ids = remap[inverse.ravel()].reshape(nodes)
    # the stencils only ever see rows 0 and 1 of a two-material mesh as
    # primary and secondary; a mesh of the primary alone still gets a
    # secondary row so the inclusion list may name it
    if rows.shape[0] == 1:
        sec = getattr(m, "secondary", None)
        if sec is not None and "c" in sec and "p" in sec:
            c = np.asarray(sec["c"], dtype=np.float64)
            rows = np.vstack([rows, [c[a, b] for a, b in _ROW_ENTRIES] + [float(sec["p"])]])
    return rows, np.ascontiguousarray(ids)


def mesh_spec(g, m, wave, cfg, device_ids=False):
    """MeshSpec of a solver's grid/material/source under its cfg.  With
    device_ids the material map is rasterized on the device from the grid's
    inclusion list (m.ids is not needed).

    ``m`` is this package's Material or the reference's (dense ``C``/``P``,
    material.py:22-93): see material_table.  ``cfg["free_surface"]``:
    "reference" (default; the reference's free surface, bit for bit) or
    "fixed" (its u_z precedence corrected, for physical spacings; no
    counterpart in the reference)."""
    nodes = (g.x.size, g.y.size, g.z.size)
    fd, sd = axis_spacings(g)
    dt = m.dt
    dt2 = dt ** 2
    raster = None
    if device_ids:
        rows = m.rows() if callable(getattr(m, "rows", None)) else material_table(m, nodes)[0]
        raster = raster_targets(g)
        id0 = raster_id(raster[0][0], raster[1][0], 0, raster[2])
        ids = None
    else:
        rows, ids = material_table(m, nodes)
        id0 = int(ids[0, 0, 0])
    mur = mur_coefficients(rows[id0], dt, fd, sd)
    probe = probe_grid(cfg.get("probe"), nodes)
    every = int(cfg.get("record_every", 10))
    steps = len(wave)
    cap = int(cfg.get("probe_capacity", 0)) or max(1, steps // max(every, 1))
    return MeshSpec(nodes, fd, sd, rows, ids, dt, dt2, mur, wave,
                    sponge=sponge_tables(nodes, cfg.get("sponge")),
                    probe=probe, record_every=every, probe_capacity=cap, raster=raster,
                    surface=free_surface(cfg))


FREE_SURFACES = {"reference": 0, "fixed": 1}


def free_surface(cfg):
    """pf_desc.surface for a solver cfg's "free_surface" entry."""
    v = cfg.get("free_surface", "reference")
    if v not in FREE_SURFACES:
        raise ValueError("free_surface must be one of %s (got %r)" % (sorted(FREE_SURFACES), v))
    return FREE_SURFACES[v]
