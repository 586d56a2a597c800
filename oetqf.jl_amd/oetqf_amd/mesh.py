"""Mesh generation (src/BEM/mesh.jl), host side.

`gen_mesh("RectOkada", ...)` follows mesh.jl:39-56 exactly.  gmsh is not
available, so `gen_gmsh_mesh("BEMHex8Mesh", ...)` builds the same
structured box (uniform transfinite x/y, nz extruded layers with heights
normalize(cumsum(rfzh), Inf), mesh.jl:95-133) itself and writes it to a
small .npz file, which `gen_mesh("BEMHex8Mesh", file)` reads back as
mesh.jl:149-186 reads a .msh (barycentres, sizes, q = (cx, cy - dy/2,
cz + dz/2)).  Cell order: x fastest, then y, then layer.
"""
import math
import os

import numpy as np


def sind(x):
    """sin of an angle in degrees, exact at multiples of 90 (Julia sind)."""
    rx = math.fmod(x, 360.0)
    arx = abs(rx)
    if rx == 0.0:
        return math.copysign(0.0, x)
    if arx == 180.0:
        return math.copysign(0.0, rx)
    if arx < 45.0:
        s = math.sin(math.radians(arx))
    elif arx <= 135.0:
        s = math.cos(math.radians(90.0 - arx))
    elif arx < 225.0:
        s = math.sin(math.radians(180.0 - arx))
    elif arx <= 315.0:
        s = -math.cos(math.radians(270.0 - arx))
    else:
        s = -math.sin(math.radians(360.0 - arx))
    return -s if rx < 0 else s


def cosd(x):
    """cos of an angle in degrees, exact at multiples of 90 (Julia cosd)."""
    arx = abs(math.fmod(x, 360.0))
    if arx < 45.0:
        return math.cos(math.radians(arx))
    if arx <= 135.0:
        return math.sin(math.radians(90.0 - arx))
    if arx < 225.0:
        return -math.cos(math.radians(180.0 - arx))
    if arx <= 315.0:
        return math.sin(math.radians(arx - 270.0))
    return math.cos(math.radians(360.0 - arx))


def _range_len(start, stop, step):
    q = (stop - start) / step
    r = round(q)
    if abs(q - r) <= 1e-10 * max(1.0, abs(r)):
        q = r
    return 0 if q < 0 else int(math.floor(q)) + 1


class RectOkadaMesh:
    """Fields of RectOkadaMesh (mesh.jl:5-21)."""

    def __init__(self, x, dx, nx, ax, xi, dxi, nxi, axi, y, z, dep, dip):
        x = np.ascontiguousarray(x, dtype=np.float64)
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        ax = np.ascontiguousarray(np.asarray(ax, dtype=np.float64).reshape(-1, 2))
        axi = np.ascontiguousarray(np.asarray(axi, dtype=np.float64).reshape(-1, 2))
        y = np.ascontiguousarray(y, dtype=np.float64)
        z = np.ascontiguousarray(z, dtype=np.float64)
        # mesh.jl:19-20
        assert len(x) == len(ax) == nx, "length(x) == length(ax) == nx"
        assert len(xi) == len(axi) == nxi == len(y) == len(z), \
            "length(ξ) == length(aξ) == nξ == length(y) == length(z)"
        self.x, self.dx, self.nx, self.ax = x, float(dx), int(nx), ax
        self.xi, self.dxi, self.nxi, self.axi = xi, dxi, int(nxi), axi
        self.y, self.z, self.dep, self.dip = y, z, float(dep), float(dip)

    # Julia-style field aliases
    @property
    def Δx(self):
        return self.dx

    @property
    def nξ(self):
        return self.nxi

    @property
    def ξ(self):
        return self.xi

    @property
    def aξ(self):
        return self.axi


class BEMHex8Mesh:
    """Fields of BEMHex8Mesh (mesh.jl:58-72)."""

    def __init__(self, cx, cy, cz, qx, qy, qz, dx, dy, dz, theta=0.0):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (cx, cy, cz, qx, qy, qz, dx, dy, dz)]
        assert len({a.shape for a in arrs}) == 1, "all hex8 mesh arrays must have the same size"
        (self.cx, self.cy, self.cz, self.qx, self.qy, self.qz, self.dx, self.dy, self.dz) = arrs
        self.theta = float(theta)

    @property
    def ne(self):
        return len(self.cx)


def _gen_rect_okada(x, xi, dx, dxi, dip):
    # _equidist_mesh_downdip (mesh.jl:45-50)
    nxi = _range_len(0.0, -xi + dxi, -dxi)
    xis = np.array([(0.0 - j * dxi) - dxi / 2 for j in range(nxi)])
    axi = np.stack([xis - dxi / 2, xis + dxi / 2], axis=1)
    y = xis * cosd(dip)
    z = xis * sind(dip)
    # _equidist_mesh_strike (mesh.jl:52-56)
    nx = _range_len(-x / 2 + dx / 2, x / 2 - dx / 2, dx)
    xs = np.array([(-x / 2 + dx / 2) + i * dx for i in range(nx)])
    ax = np.stack([xs - dx / 2, xs + dx / 2], axis=1)
    return RectOkadaMesh(xs, dx, nx, ax, xis, dxi, nxi, axi, y, z, 0.0, dip)


def gen_mesh(kind, *args, **kwargs):
    """gen_mesh(Val(:RectOkada), x, ξ, Δx, Δξ, dip) (mesh.jl:39-43) or
    gen_mesh(Val(:BEMHex8Mesh), file; rotation, transpose) (mesh.jl:149-186)."""
    kind = str(kind).lstrip(":")
    if kind == "RectOkada":
        return _gen_rect_okada(*[float(a) for a in args])
    if kind == "BEMHex8Mesh":
        return _read_hex8(*args, **kwargs)
    raise ValueError("unknown mesh kind %r" % kind)


def gen_mesh_graded(x, xi, dx, nxi, dip, ratio=1.0):
    """Rectangular Okada mesh with equidistant strike cells and nxi down-dip
    cells whose widths grow geometrically by `ratio` with depth (finer
    toward the surface).  Produces the same RectOkadaMesh fields (the
    Green's functions use only aξ, y, z, Δx and nx, GF.jl:38,49-53)."""
    w = ratio ** np.arange(nxi)
    edges = np.concatenate([[0.0], -np.cumsum(w / w.sum() * xi)])
    edges[-1] = -xi
    xis = 0.5 * (edges[:-1] + edges[1:])
    axi = np.stack([edges[1:], edges[:-1]], axis=1)
    m = _gen_rect_okada(x, xi, dx, xi / nxi, dip)
    return RectOkadaMesh(m.x, dx, m.nx, m.ax, xis, xi / nxi, nxi, axi, xis * cosd(dip),
                         xis * sind(dip), 0.0, dip)


def hex8_box(llx, lly, llz, dx, dy, dz, nx, ny, nz, rfzh=None):
    """Structured box of hex8 cells (the geometry of gen_gmsh_mesh with
    rfx = rfy = 1) returned directly as a BEMHex8Mesh."""
    rf = np.ones(nz) if rfzh is None else np.asarray(rfzh, dtype=np.float64)
    assert len(rf) == nz, "length(rfzh) must equal nz"
    cum = np.cumsum(rf)
    zf = np.concatenate([[llz], llz + dz * (cum / cum[-1])])
    xf = llx + dx * np.arange(nx + 1) / nx
    yf = lly + dy * np.arange(ny + 1) / ny
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    cx = 0.5 * (xf[I] + xf[I + 1])
    cy = 0.5 * (yf[J] + yf[J + 1])
    cz = 0.5 * (zf[K] + zf[K + 1])
    ddx = np.abs(xf[I + 1] - xf[I])
    ddy = np.abs(yf[J + 1] - yf[J])
    ddz = np.abs(zf[K + 1] - zf[K])
    return BEMHex8Mesh(cx, cy, cz, cx, cy - ddy / 2, cz + ddz / 2, ddx, ddy, ddz, 0.0)


def gen_gmsh_mesh(kind, llx, lly, llz, dx, dy, dz, nx, ny, nz, rfx=1.0, rfy=1.0, rfzh=None,
                  rfxType="Bump", rfyType="Bump", output="temp.npz"):
    """gen_gmsh_mesh(Val(:BEMHex8Mesh), ...) (mesh.jl:95-133): writes the box
    to `output` (npz instead of gmsh's .msh)."""
    if str(kind).lstrip(":") != "BEMHex8Mesh":
        raise ValueError("unknown mesh kind %r" % kind)
    if rfx != 1.0 or rfy != 1.0:
        raise NotImplementedError("gmsh Bump refinement (rfx, rfy != 1) needs gmsh")
    m = hex8_box(llx, lly, llz, dx, dy, dz, nx, ny, nz, rfzh)
    with open(output, "wb") as f:
        np.savez(f, cx=m.cx, cy=m.cy, cz=m.cz, dx=m.dx, dy=m.dy, dz=m.dz)
    return output


def _read_hex8(file, rotation=0.0, transpose=False):
    if not os.path.isfile(file):
        raise FileNotFoundError("Mesh file %s does not exist." % file)
    d = np.load(file, allow_pickle=False)
    cx, cy, cz, ddx, ddy, ddz = (d[k] for k in ("cx", "cy", "cz", "dx", "dy", "dz"))
    if transpose:  # mesh.jl:171-173 swaps the in-plane edge directions
        ddx, ddy = ddy, ddx
    return BEMHex8Mesh(cx, cy, cz, cx, cy - ddy / 2, cz + ddz / 2, ddx, ddy, ddz, rotation)
