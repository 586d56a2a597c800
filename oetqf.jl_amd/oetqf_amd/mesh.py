"""Mesh generation (src/BEM/mesh.jl), host side.

`gen_mesh("RectOkada", ...)` follows mesh.jl:39-56 exactly.  gmsh is not
available, so `gen_gmsh_mesh("BEMHex8Mesh", ...)` builds the same box
(transfinite x/y edges with "Bump" or "Progression" node distributions,
nz extruded layers with heights normalize(cumsum(rfzh), Inf), mesh.jl:95-133)
itself and writes it as a gmsh MSH 4.1 ASCII file of 8-node hexahedra
(element type 5).  `gen_mesh("BEMHex8Mesh", file)` reads MSH 4.1 or 2.2
ASCII files (gmsh's or these) as mesh.jl:149-186 does: barycentre = mean
of the 8 nodes, Δx = |p1 - p4| and Δy = |p1 - p2| in the horizontal plane
(nodes 2 and 4 swapped with transpose=true), Δz = |z1 - z5|, q = (cx,
cy - Δy/2, cz + Δz/2).  Cell order of the files written here: x fastest,
then y, then layer.  gmsh's own cell numbering and the exact node
distribution of its "Bump" rule are not reproduced (gmsh is absent):
`bump_nodes` below is this package's profile with the same meaning of the
coefficient (coef < 1 refines toward both ends, coef > 1 toward the
middle).
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


def hex8_box(llx, lly, llz, dx, dy, dz, nx, ny, nz, rfzh=None, rfx=1.0, rfy=1.0, rfxType="Bump",
             rfyType="Bump"):
    """Structured box of hex8 cells (the geometry of gen_gmsh_mesh) returned
    directly as a BEMHex8Mesh."""
    xf, yf, zf = _box_nodes(llx, lly, llz, dx, dy, dz, nx, ny, nz, rfx, rfy, rfzh, rfxType, rfyType)
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    cx = 0.5 * (xf[I] + xf[I + 1])
    cy = 0.5 * (yf[J] + yf[J + 1])
    cz = 0.5 * (zf[K] + zf[K + 1])
    ddx = np.abs(xf[I + 1] - xf[I])
    ddy = np.abs(yf[J + 1] - yf[J])
    ddz = np.abs(zf[K + 1] - zf[K])
    return BEMHex8Mesh(cx, cy, cz, cx, cy - ddy / 2, cz + ddz / 2, ddx, ddy, ddz, 0.0)


def bump_nodes(n, coef):
    """n + 1 node parameters in [0, 1] of a "Bump" distribution: element
    sizes follow h(t) ∝ 1 + (coef - 1)(2t - 1)^2, so the elements at both
    ends are `coef` times the size of the middle one (coef < 1: refined
    toward the ends).  coef = 1 is uniform."""
    if coef <= 0:
        raise ValueError("Bump coefficient must be > 0")
    if coef == 1.0:
        return np.arange(n + 1) / n
    # cumulative density of 1/h (trapezoids on a fine grid), inverted at k/n
    # by linear interpolation
    t = np.linspace(0.0, 1.0, 200001)
    dens = 1.0 / (1.0 + (coef - 1.0) * (2.0 * t - 1.0) ** 2)
    F = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(t))])
    F /= F[-1]
    x = np.interp(np.arange(n + 1) / n, F, t)
    x[0], x[-1] = 0.0, 1.0
    return x


def progression_nodes(n, coef):
    """n + 1 node parameters of a geometric "Progression": each element is
    `coef` times the previous one."""
    if coef <= 0:
        raise ValueError("Progression coefficient must be > 0")
    w = coef ** np.arange(n)
    x = np.concatenate([[0.0], np.cumsum(w / w.sum())])
    x[-1] = 1.0
    return x


def _edge_nodes(n, coef, kind):
    if kind == "Bump":
        return bump_nodes(n, coef)
    if kind == "Progression":
        return progression_nodes(n, coef)
    raise ValueError("unknown transfinite distribution %r (Bump or Progression)" % kind)


def _box_nodes(llx, lly, llz, dx, dy, dz, nx, ny, nz, rfx, rfy, rfzh, rfxType, rfyType):
    # uniform edges as llx + dx * i / nx (the oracle's structured box)
    xs = llx + (dx * np.arange(nx + 1) / nx if rfx == 1.0 else dx * _edge_nodes(nx, rfx, rfxType))
    ys = lly + (dy * np.arange(ny + 1) / ny if rfy == 1.0 else dy * _edge_nodes(ny, rfy, rfyType))
    rf = np.ones(nz) if rfzh is None else np.asarray(rfzh, dtype=np.float64)
    assert len(rf) == nz, "length(rfzh) must equal nz"
    cum = np.cumsum(rf)
    zs = np.concatenate([[llz], llz + dz * (cum / cum[-1])])  # normalize(cumsum(rfzh), Inf)
    return xs, ys, zs


def gen_gmsh_mesh(kind, llx, lly, llz, dx, dy, dz, nx, ny, nz, rfx=1.0, rfy=1.0, rfzh=None,
                  rfxType="Bump", rfyType="Bump", output="temp.msh"):
    """gen_gmsh_mesh(Val(:BEMHex8Mesh), ...) (mesh.jl:95-133): the box as a
    gmsh MSH 4.1 ASCII file of hexahedra.  rfx/rfy refine the x/y edges
    ("Bump": bump_nodes, "Progression": progression_nodes); rfzh are the
    layer heights (normalised)."""
    if str(kind).lstrip(":") != "BEMHex8Mesh":
        raise ValueError("unknown mesh kind %r" % kind)
    xs, ys, zs = _box_nodes(llx, lly, llz, dx, dy, dz, nx, ny, nz, rfx, rfy, rfzh, rfxType, rfyType)
    nxn, nyn = nx + 1, ny + 1
    tag = lambda i, j, k: 1 + i + nxn * (j + nyn * k)
    nn = nxn * nyn * (nz + 1)
    ne = nx * ny * nz
    with open(output, "w") as f:
        f.write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
        f.write("$Entities\n0 0 0 1\n1 %.17g %.17g %.17g %.17g %.17g %.17g 0 0\n$EndEntities\n" % (
            xs.min(), ys.min(), zs.min(), xs.max(), ys.max(), zs.max()))
        f.write("$Nodes\n1 %d 1 %d\n3 1 0 %d\n" % (nn, nn, nn))
        f.write("".join("%d\n" % (t + 1) for t in range(nn)))
        for k in range(nz + 1):
            for j in range(nyn):
                for i in range(nxn):
                    f.write("%.17g %.17g %.17g\n" % (xs[i], ys[j], zs[k]))
        f.write("$EndNodes\n$Elements\n1 %d 1 %d\n3 1 5 %d\n" % (ne, ne, ne))
        e = 0
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    e += 1
                    # node 2 one step in y, node 4 one step in x from node 1;
                    # nodes 5-8 the next layer (mesh.jl:167-180 reads them so)
                    b = [tag(i, j, k), tag(i, j + 1, k), tag(i + 1, j + 1, k), tag(i + 1, j, k)]
                    t = [tag(i, j, k + 1), tag(i, j + 1, k + 1), tag(i + 1, j + 1, k + 1),
                         tag(i + 1, j, k + 1)]
                    f.write("%d %s\n" % (e, " ".join(str(x) for x in b + t)))
        f.write("$EndElements\n")
    return output


def _msh_hexahedra(file):
    """(node coordinates by tag, list of 8-node connectivities) of the type-5
    elements of a gmsh MSH 4.1 or 2.2 ASCII file."""
    with open(file) as f:
        tok = f.read().split("\n")
    i, nodes, hexes, version = 0, {}, [], None
    while i < len(tok):
        ln = tok[i].strip()
        if ln == "$MeshFormat":
            version = tok[i + 1].split()[0]
            if int(tok[i + 1].split()[1]) != 0:
                raise ValueError("binary MSH files are not supported (write ASCII)")
            i += 2
        elif ln == "$Nodes":
            if version.startswith("4"):
                nb = int(tok[i + 1].split()[0])
                i += 2
                for _ in range(nb):
                    _, _, par, cnt = (int(v) for v in tok[i].split())
                    tags = [int(tok[i + 1 + q]) for q in range(cnt)]
                    for q, t in enumerate(tags):
                        xyz = tok[i + 1 + cnt + q].split()
                        nodes[t] = (float(xyz[0]), float(xyz[1]), float(xyz[2]))
                    i += 1 + 2 * cnt
            else:
                cnt = int(tok[i + 1])
                for q in range(cnt):
                    v = tok[i + 2 + q].split()
                    nodes[int(v[0])] = (float(v[1]), float(v[2]), float(v[3]))
                i += 2 + cnt
        elif ln == "$Elements":
            if version.startswith("4"):
                nb = int(tok[i + 1].split()[0])
                i += 2
                for _ in range(nb):
                    _, _, etype, cnt = (int(v) for v in tok[i].split())
                    if etype == 5:
                        for q in range(cnt):
                            hexes.append([int(v) for v in tok[i + 1 + q].split()[1:9]])
                    i += 1 + cnt
            else:
                cnt = int(tok[i + 1])
                for q in range(cnt):
                    v = [int(x) for x in tok[i + 2 + q].split()]
                    if v[1] == 5:
                        hexes.append(v[3 + v[2]:3 + v[2] + 8])
                i += 2 + cnt
        else:
            i += 1
    if version is None:
        raise ValueError("%s is not a gmsh MSH file" % file)
    return nodes, hexes


def _read_hex8(file, rotation=0.0, transpose=False):
    """gen_mesh(Val(:BEMHex8Mesh), file; rotation, transpose) (mesh.jl:149-186)."""
    if not os.path.isfile(file):
        raise FileNotFoundError("Mesh file %s does not exist." % file)
    if file.endswith(".npz"):
        d = np.load(file, allow_pickle=False)
        cx, cy, cz, ddx, ddy, ddz = (d[k] for k in ("cx", "cy", "cz", "dx", "dy", "dz"))
        if transpose:
            ddx, ddy = ddy, ddx
        return BEMHex8Mesh(cx, cy, cz, cx, cy - ddy / 2, cz + ddz / 2, ddx, ddy, ddz, rotation)
    nodes, hexes = _msh_hexahedra(file)
    P = np.array([[nodes[t] for t in h] for h in hexes], dtype=np.float64)  # (ne, 8, 3)
    c = P.mean(axis=1)
    p1, p2, p4, p5 = P[:, 0], P[:, 1], P[:, 3], P[:, 4]
    if transpose:  # mesh.jl:171-173
        p2, p4 = p4, p2
    ddx = np.hypot(p1[:, 0] - p4[:, 0], p1[:, 1] - p4[:, 1])
    ddy = np.hypot(p1[:, 0] - p2[:, 0], p1[:, 1] - p2[:, 1])
    ddz = np.abs(p1[:, 2] - p5[:, 2])
    return BEMHex8Mesh(c[:, 0], c[:, 1], c[:, 2], c[:, 0], c[:, 1] - ddy / 2, c[:, 2] + ddz / 2,
                       ddx, ddy, ddz, rotation)
