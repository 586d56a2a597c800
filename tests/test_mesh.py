"""Mesh generation: the reference's known-answer tests (test/BEM/tests.jl:4-37)
for the host mesh builders and the oracle's C builders, and their agreement."""
import os

import numpy as np
import pytest


def test_rect_okada_known_answers(oq, oracle):
    # test/BEM/tests.jl:5-13
    m = oq.gen_mesh("RectOkada", 100.0, 50.0, 2.0, 2.0, 33.0)
    assert m.nx == 50 and m.nxi == 25
    assert m.axi[-1][0] == m.xi[-1] - m.dxi / 2
    assert m.y[-1] == m.xi[-1] * oq.cosd(m.dip)
    assert m.z[-1] == m.xi[-1] * oq.sind(m.dip)
    assert m.x[-1] - m.x[0] == m.dx * (m.nx - 1)
    o = oracle.mesh_rect_okada(100.0, 50.0, 2.0, 2.0, 33.0)
    assert o["nx"] == 50 and o["nxi"] == 25
    for k in ("x", "ax", "xi", "axi", "y", "z"):
        assert np.array_equal(np.asarray(getattr(m, k)), o[k]), k


@pytest.mark.parametrize("args", [(40e3, 20e3, 1250.0, 1250.0, 90.0), (80e3, 8e3, 10e3, 2e3, 90.0),
                                  (10.0, 10.0, 2.0, 2.0, 90.0), (100.0, 100.0, 10.0, 20.0, 90.0)])
def test_rect_okada_sizes(oq, args):
    m = oq.gen_mesh("RectOkada", *args)
    assert m.nx == round(args[0] / args[2]) and m.nxi == round(args[1] / args[3])
    assert m.z.max() < 0 and m.x[0] == -args[0] / 2 + args[2] / 2
    assert np.all(m.ax[:, 1] - m.ax[:, 0] == args[2])


def test_hex8_mesh_invariants(oq, oracle, rng, tmp_path):
    # test/BEM/tests.jl:15-36
    llx, lly, llz = rng.random(3)
    dx, dy, dz = rng.random(3) * 5
    nx, ny, nz = (int(v) for v in rng.integers(2, 11, 3))
    f = str(tmp_path / "m.msh")
    oq.gen_gmsh_mesh("BEMHex8Mesh", llx, lly, llz, dx, dy, -dz, nx, ny, nz, output=f)
    me = oq.gen_mesh("BEMHex8Mesh", f)
    for x, tx in zip([me.cx, me.cy, me.cz], [nx, ny, nz]):
        assert len(np.unique(np.round(x, 6))) == tx
    assert np.allclose(me.cx, me.qx)
    assert np.allclose(me.cy - me.dy / 2, me.qy)
    assert np.allclose(me.cz + me.dz / 2, me.qz)
    assert np.isclose(np.max(me.cx + me.dx / 2), llx + dx)
    assert np.isclose(np.min(me.cx - me.dx / 2), llx)
    assert np.isclose(np.max(me.cy + me.dy / 2), lly + dy)
    assert np.isclose(np.min(me.cy - me.dy / 2), lly)
    assert np.isclose(np.max(me.cz + me.dz / 2), llz)
    assert np.isclose(np.min(me.cz - me.dz / 2), llz - dz)
    o = oracle.mesh_hex8(llx, lly, llz, dx, dy, -dz, nx, ny, nz)
    for k in ("cx", "cy", "cz", "qx", "qy", "qz", "dx", "dy", "dz"):
        assert np.allclose(getattr(me, k), o[k], rtol=1e-15, atol=1e-15), k


def test_hex8_layer_heights(oq):
    """rfzh: cumulative layer heights normalised by their maximum
    (normalize(cumsum(rfzh), Inf), mesh.jl:127)."""
    m = oq.hex8_box(-40e3, -2.5e3, -8e3, 80e3, 5e3, -22e3, 4, 3, 3, rfzh=np.cumprod(np.ones(3) * 1.5))
    tops = np.unique(np.round(m.cz + m.dz / 2, 6))[::-1]
    w = np.cumprod(np.ones(3) * 1.5)
    cum = np.cumsum(w) / np.sum(w)
    assert np.allclose(np.sort(np.unique(np.round(m.dz, 6))), np.sort(22e3 * w / w.sum()))
    assert np.isclose(tops[0], -8e3) and len(tops) == 3
    assert np.isclose(m.cz.min() - m.dz[np.argmin(m.cz)] / 2, -8e3 - 22e3 * cum[-1])


def test_graded_fault_mesh(oq):
    m = oq.gen_mesh_graded(80e3, 20e3, 625.0, 32, 90.0, ratio=1.05)
    assert m.nx == 128 and m.nxi == 32
    w = m.axi[:, 1] - m.axi[:, 0]
    assert np.all(np.diff(w) > 0) and np.isclose(w.sum(), 20e3)
    assert m.axi[0][1] == 0.0 and m.axi[-1][0] == -20e3
    assert np.allclose(m.z, 0.5 * (m.axi[:, 0] + m.axi[:, 1]))


def test_missing_mesh_file(oq, tmp_path):
    with pytest.raises(FileNotFoundError):
        oq.gen_mesh("BEMHex8Mesh", str(tmp_path / "nope.npz"))


def test_gmsh_file_round_trip(oq, tmp_path):
    """gen_gmsh_mesh writes an MSH 4.1 file of hexahedra; gen_mesh reads it
    as mesh.jl:149-186 does and gives the structured box cell for cell."""
    f = str(tmp_path / "box.msh")
    oq.gen_gmsh_mesh("BEMHex8Mesh", -40e3, -20e3, -10e3, 80e3, 40e3, -30e3, 4, 3, 2, output=f)
    assert open(f).readline().strip() == "$MeshFormat"
    me = oq.gen_mesh("BEMHex8Mesh", f)
    box = oq.hex8_box(-40e3, -20e3, -10e3, 80e3, 40e3, -30e3, 4, 3, 2)
    for k in ("cx", "cy", "cz", "qx", "qy", "qz", "dx", "dy", "dz"):
        assert np.allclose(getattr(me, k), getattr(box, k), rtol=1e-14, atol=1e-9), k
    # transpose swaps the in-plane edge roles of nodes 2 and 4 (mesh.jl:171-173)
    mt = oq.gen_mesh("BEMHex8Mesh", f, transpose=True)
    assert np.allclose(mt.dx, box.dy) and np.allclose(mt.dy, box.dx)
    assert np.allclose(mt.qy, mt.cy - box.dx / 2)


def test_gmsh_msh22_reader(oq, tmp_path):
    """An MSH 2.2 ASCII file (gmsh's older format) of one unit hexahedron with
    nodes listed as gmsh lists them after mesh.jl's extrusion (node 2 along
    y, node 4 along x from node 1, nodes 5-8 one layer down)."""
    f = tmp_path / "one.msh"
    pts = [(0, 0, 0), (0, 2, 0), (3, 2, 0), (3, 0, 0), (0, 0, -5), (0, 2, -5), (3, 2, -5), (3, 0, -5)]
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", "8"]
    lines += ["%d %g %g %g" % (i + 1, *p) for i, p in enumerate(pts)]
    lines += ["$EndNodes", "$Elements", "2", "1 15 2 0 1 1", "2 5 2 0 1 1 2 3 4 5 6 7 8", "$EndElements"]
    f.write_text("\n".join(lines) + "\n")
    m = oq.gen_mesh("BEMHex8Mesh", str(f))
    assert m.ne == 1
    assert (m.cx[0], m.cy[0], m.cz[0]) == (1.5, 1.0, -2.5)
    assert (m.dx[0], m.dy[0], m.dz[0]) == (3.0, 2.0, 5.0)
    assert (m.qx[0], m.qy[0], m.qz[0]) == (1.5, 0.0, 0.0)


@pytest.mark.parametrize("coef", [0.3, 1.0, 2.5])
def test_bump_refinement(oq, tmp_path, coef):
    """Bump: symmetric edge nodes; the end elements are coef times the middle
    one (refined toward both ends for coef < 1); the invariants of
    test/BEM/tests.jl:15-36 hold on the refined box read back from its file."""
    from oetqf_amd.mesh import bump_nodes
    t = bump_nodes(20, coef)
    h = np.diff(t)
    assert t[0] == 0.0 and t[-1] == 1.0 and np.all(h > 0)
    assert np.allclose(h, h[::-1], rtol=1e-6)
    assert np.all(np.diff(h[:10]) * (1.0 - coef) >= 0)  # monotone from the ends to the middle
    hf = np.diff(bump_nodes(400, coef))                   # the size ratio, in the fine limit
    assert abs(hf[0] / hf[200] - coef) <= 0.02 * coef
    f = str(tmp_path / "b.msh")
    llx, lly, llz, dx, dy, dz = 1.0, 2.0, -3.0, 40.0, 30.0, -20.0
    oq.gen_gmsh_mesh("BEMHex8Mesh", llx, lly, llz, dx, dy, dz, 9, 7, 3, rfx=coef, rfy=coef,
                     rfzh=[1.0, 2.0, 3.0], output=f)
    me = oq.gen_mesh("BEMHex8Mesh", f)
    for x, tx in zip([me.cx, me.cy, me.cz], [9, 7, 3]):
        assert len(np.unique(np.round(x, 6))) == tx
    assert np.isclose(np.max(me.cx + me.dx / 2), llx + dx) and np.isclose(np.min(me.cx - me.dx / 2), llx)
    assert np.isclose(np.max(me.cy + me.dy / 2), lly + dy) and np.isclose(np.min(me.cy - me.dy / 2), lly)
    assert np.isclose(np.max(me.cz + me.dz / 2), llz) and np.isclose(np.min(me.cz - me.dz / 2), llz + dz)
    assert np.isclose(np.sum(me.dx * me.dy * me.dz), dx * dy * -dz)
    box = oq.hex8_box(llx, lly, llz, dx, dy, dz, 9, 7, 3, rfzh=[1.0, 2.0, 3.0], rfx=coef, rfy=coef)
    assert np.allclose(box.cx, me.cx) and np.allclose(box.dy, me.dy)


def test_progression_refinement(oq):
    from oetqf_amd.mesh import progression_nodes
    h = np.diff(progression_nodes(8, 1.3))
    assert np.allclose(h[1:] / h[:-1], 1.3) and np.isclose(h.sum(), 1.0)
    with pytest.raises(ValueError):
        oq.hex8_box(0, 0, 0, 1, 1, -1, 2, 2, 1, rfx=2.0, rfxType="Beta")
