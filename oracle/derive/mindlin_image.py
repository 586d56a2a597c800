"""Derivation script (TEST INFRASTRUCTURE / code generation).

Generates the C body of the smooth "image" part of the half-space
eigenstrain kernel used by the hex8 strain-volume stress:

    D[i][l][p] = sum_{(j,k) in pair p} d^2 G^I_ij(x, y) / dx_l dy_k

where G^I = G^Mindlin - G^Kelvin is the part of Mindlin's (1936) point-force
Green's function for a traction-free half-space that is regular inside the
medium, written in the Okada frame (z up, medium z <= 0).  p runs over the
symmetric pairs (xx, xy, xz, yy, yz, zz); off-diagonal pairs sum both
orders so that contraction with the 6 unique moment components is exact.

Mindlin's solution is checked here before use: Navier equilibrium,
traction-free surface and reciprocity, symbolically evaluated at a point.

Run:  python oracle/derive/mindlin_image.py OUT.inc  [--check] [--device]
      python oracle/derive/mindlin_image.py --traction FT OUT.inc [--halves]
      python oracle/derive/mindlin_image.py --stress OUT.inc [--halves]
The output is written verbatim to oracle/hex8_image.inc (--device raises
every negative power of R and of zr + ys - R from the one reciprocal of
each that the common subexpressions already hold).  --stress and --traction
write the device's combinations of the kernel (the diagonal and symmetric
sums a stress takes, as oetqf.jl_amd/csrc/hex8_image_stress.inc; the pair
sums a fault's shear traction takes, hex8_image_trac0.inc with --halves and
hex8_image_trac1.inc), simplified as a whole.
"""
import sys

import sympy as sp

X, Y, zr, ys, nu = sp.symbols("X Y zr ys nu", real=True)


def mindlin(x, y, z, c, nu, kelvin=True):
    """Mindlin's Green's function, depth-positive frame (z, c >= 0), unit
    force, factor 16 pi mu (1-nu) removed.  Returns G[i][j] (u_i from F_j)."""
    R1 = sp.sqrt(x**2 + y**2 + (z - c) ** 2)
    R2 = sp.sqrt(x**2 + y**2 + (z + c) ** 2)
    A = 4 * (1 - nu) * (1 - 2 * nu)
    k = 1 if kelvin else 0
    h = [x, y]
    G = [[0] * 3 for _ in range(3)]
    for a in range(2):
        for b in range(2):
            d = 1 if a == b else 0
            G[a][b] = (k * ((3 - 4 * nu) * d / R1 + h[a] * h[b] / R1**3)
                       + d / R2 + (3 - 4 * nu) * h[a] * h[b] / R2**3
                       + 2 * c * z / R2**3 * (d - 3 * h[a] * h[b] / R2**2)
                       + A / (R2 + z + c) * (d - h[a] * h[b] / (R2 * (R2 + z + c))))
        G[2][a] = h[a] * (k * (z - c) / R1**3 + (3 - 4 * nu) * (z - c) / R2**3
                          - 6 * c * z * (z + c) / R2**5 + A / (R2 * (R2 + z + c)))
        G[a][2] = h[a] * (k * (z - c) / R1**3 + (3 - 4 * nu) * (z - c) / R2**3
                          + 6 * c * z * (z + c) / R2**5 - A / (R2 * (R2 + z + c)))
    G[2][2] = (k * ((3 - 4 * nu) / R1 + (z - c) ** 2 / R1**3)
               + (8 * (1 - nu) ** 2 - (3 - 4 * nu)) / R2
               + ((3 - 4 * nu) * (z + c) ** 2 - 2 * c * z) / R2**3
               + 6 * c * z * (z + c) ** 2 / R2**5)
    return G


def check():
    x, y, z, c = sp.symbols("x y z c", real=True)
    mu = sp.Integer(1)
    lam = 2 * mu * nu / (1 - 2 * nu)
    G = mindlin(x, y, z, c, nu)
    V = [x, y, z]
    vals = {x: 0.37, y: -0.81, z: 1.3, c: 0.9, nu: 0.27}
    vals0 = dict(vals)
    vals0[z] = 0
    worst = 0.0
    for j in range(3):
        u = [G[i][j] for i in range(3)]
        div = sum(sp.diff(u[i], V[i]) for i in range(3))
        for i in range(3):
            nav = mu * sum(sp.diff(u[i], V[k], 2) for k in range(3)) + (lam + mu) * sp.diff(div, V[i])
            worst = max(worst, abs(float(nav.subs(vals))))
        for i in range(3):
            s = lam * div * (1 if i == 2 else 0) + mu * (sp.diff(u[i], z) + sp.diff(u[2], V[i]))
            worst = max(worst, abs(float(s.subs(vals0))))
    rec = G[2][0] - G[0][2].subs({x: -x, z: c, c: z}, simultaneous=True)
    worst = max(worst, abs(float(rec.subs(vals))))
    print("Mindlin check (equilibrium, free surface, reciprocity): max residual", worst)
    assert worst < 1e-12


def image_kernel():
    """D[i][l][p] in the z-up frame for receiver-minus-source horizontal
    offsets (X, Y), receiver z = zr <= 0 and source z = ys <= 0."""
    S = [1, 1, -1]
    Gm = mindlin(X, Y, -zr, -ys, nu, kelvin=False)
    G = [[S[i] * S[j] * Gm[i][j] for j in range(3)] for i in range(3)]
    dx = [lambda f: sp.diff(f, X), lambda f: sp.diff(f, Y), lambda f: sp.diff(f, zr)]
    dy = [lambda f: -sp.diff(f, X), lambda f: -sp.diff(f, Y), lambda f: sp.diff(f, ys)]
    pairs = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    out = []
    for i in range(3):
        for l in range(3):
            for (j, k) in pairs:
                e = dx[l](dy[k](G[i][j]))
                if j != k:
                    e = e + dx[l](dy[j](G[i][k]))
                out.append(e)
    return out


def reciprocal_powers(lines):
    """Rewrites OQ_IPOW(b, -k) (and a division by OQ_IPOW(b, k)) as a positive
    power of a reciprocal of b that
    the generated code defines as 1.0/b (or of its negation, for a base
    defined as the negation of such a b), and 1.0/b' for that negated base
    likewise."""
    import re
    defs = {}
    for ln in lines:
        m = re.match(r"const double (x\d+) = (.*);$", ln)
        if m:
            defs[m.group(1)] = m.group(2)
    recip = {"R": "x0"}
    assert defs["x0"] == "1.0/R"
    for name, rhs in defs.items():
        m = re.fullmatch(r"1\.0/(x\d+)", rhs)
        if m and m.group(1) not in recip:
            recip[m.group(1)] = name
    for name, rhs in defs.items():
        m = re.fullmatch(r"-(x\d+)", rhs)
        if m and m.group(1) in recip:  # -1/b: defined where b's reciprocal is
            recip[name] = "(-%s)" % recip[m.group(1)]
    out = []
    for ln in lines:
        m = re.match(r"const double (x\d+) = 1\.0/(x\d+);$", ln)
        if m and recip.get(m.group(2), m.group(1)) != m.group(1):
            ln = "const double %s = %s;" % (m.group(1), recip[m.group(2)])
        ln = re.sub(r"/OQ_IPOW\((\w+), (\d+)\)",
                    lambda mm: "*OQ_IPOW(%s, %s)" % (recip[mm.group(1)], mm.group(2)), ln)
        ln = re.sub(r"OQ_IPOW\((\w+), -(\d+)\)",
                    lambda mm: "OQ_IPOW(%s, %s)" % (recip[mm.group(1)], mm.group(2)), ln)
        out.append(ln)
    return out


def emit(path, device=False):
    import re
    R = sp.Symbol("R", positive=True)
    r2 = X**2 + Y**2 + (-ys - zr) ** 2
    exprs = [sp.powsimp(e.subs(r2, R**2), force=True) for e in image_kernel()]
    reps, red = sp.cse(exprs, optimizations="basic")
    lines = ["/* generated by oracle/derive/mindlin_image.py -- do not edit.",
             " * Inputs: X, Y (receiver - source, horizontal), zr, ys (receiver and",
             " * source z, both <= 0), nu, R = |(X, Y, zr + ys)|.  Output D[54]. */"]

    def cc(e):
        t = sp.ccode(e)
        t = re.sub(r"pow\(([A-Za-z0-9_]+), (-?\d+)\)", r"OQ_IPOW(\1, \2)", t)
        assert "pow(" not in t and "sqrt(" not in t, t
        return t
    for sym, e in reps:
        lines.append("const double %s = %s;" % (sym, cc(e)))
    for n, e in enumerate(red):
        lines.append("D[%d] = %s;" % (n, cc(e)))
    if device:
        lines = reciprocal_powers(lines)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", path, len(reps), "temporaries")


# The pairs of D entries that the shear traction on a fault takes (the
# traction projection of GF.jl:89-96 applied to the image stress): strike
# slip -sigma_xy sin(dip) + sigma_xz cos(dip) needs D[0][1] + D[1][0] and
# D[0][2] + D[2][0]; dip slip (sigma_zz - sigma_yy)/2 sin(2 dip) + sigma_yz
# cos(2 dip) needs D[2][2] - D[1][1] and D[1][2] + D[2][1] (per eigenstrain
# pair p; the coefficients multiply them in the kernel).
TRACTION_PAIRS = {0: (((0, 1), (1, 0), 1), ((0, 2), (2, 0), 1)),
                  1: (((2, 2), (1, 1), -1), ((1, 2), (2, 1), 1))}


def traction_pairs(ft):
    R = sp.Symbol("R", positive=True)
    r2 = X**2 + Y**2 + (-ys - zr) ** 2
    D = [sp.powsimp(e.subs(r2, R**2), force=True) for e in image_kernel()]
    at = lambda il, p: D[il[0] * 18 + il[1] * 6 + p]
    out = []
    for a, b, sgn in TRACTION_PAIRS[ft]:
        for p in range(6):
            out.append(at(a, p) + sgn * at(b, p))
    return out, [sp.simplify(e) for e in out]


# The combinations of D entries that a full stress takes (the image part of
# grad u enters sigma = lambda tr(eps) I + 2 mu eps only through the
# diagonal and the symmetric sums): per moment pair q, D[0][0], D[1][1],
# D[2][2], D[0][1] + D[1][0], D[0][2] + D[2][0], D[1][2] + D[2][1].
STRESS_COMBOS = (((0, 0),), ((1, 1),), ((2, 2),), ((0, 1), (1, 0)), ((0, 2), (2, 0)), ((1, 2), (2, 1)))


def stress_combos():
    R = sp.Symbol("R", positive=True)
    r2 = X**2 + Y**2 + (-ys - zr) ** 2
    D = [sp.powsimp(e.subs(r2, R**2), force=True) for e in image_kernel()]
    out = []
    for q in range(6):
        for combo in STRESS_COMBOS:
            out.append(sum(D[i * 18 + l * 6 + q] for i, l in combo))
    return out, [sp.simplify(e) for e in out]


def emit_traction(path, ft, halves=False):
    """P[12] = the two pair sums of TRACTION_PAIRS[ft] for the six p,
    simplified before the common subexpressions are taken (half the
    operations of the 24 entries they come from); checked against the
    unsimplified sums at random points."""
    import random
    import re
    R = sp.Symbol("R", positive=True)
    raw, exprs = traction_pairs(ft) if ft >= 0 else stress_combos()
    rnd = random.Random(7)
    worst = 0.0
    for _ in range(20):
        v = {X: rnd.uniform(-3, 3), Y: rnd.uniform(-3, 3), zr: -rnd.uniform(0.1, 3),
             ys: -rnd.uniform(0.1, 3), nu: rnd.uniform(0.1, 0.45)}
        v[R] = float(sp.sqrt(v[X] ** 2 + v[Y] ** 2 + (v[zr] + v[ys]) ** 2))
        for a, b in zip(raw, exprs):
            fa, fb = float(a.subs(v)), float(b.subs(v))
            worst = max(worst, abs(fa - fb) / max(abs(fa), 1e-300))
    print("%s: simplified vs raw, max relative difference %.1e" % (
        "traction pairs ft=%d" % ft if ft >= 0 else "stress combinations", worst))
    assert worst < 1e-12
    if ft >= 0:
        lines = ["/* generated by oracle/derive/mindlin_image.py --traction %d OUT%s -- do not edit."
                 % (ft, " --halves" if halves else ""),
                 " * Inputs: X, Y (receiver - source, horizontal), zr, ys (receiver and",
                 " * source z, both <= 0), nu, R = |(X, Y, zr + ys)|.  Output P[12]: the",
                 " * pair sums of the image kernel the shear traction takes, P[p] and",
                 " * P[6 + p] for eigenstrain pair p. */"]
    else:
        lines = ["/* generated by oracle/derive/mindlin_image.py --stress OUT%s -- do not edit."
                 % (" --halves" if halves else ""),
                 " * Inputs: X, Y (receiver - source, horizontal), zr, ys (receiver and",
                 " * source z, both <= 0), nu, R = |(X, Y, zr + ys)|.  Output P[36]:",
                 " * P[6 q + c], moment pair q, c = D[0][0], D[1][1], D[2][2], D[0][1] +",
                 " * D[1][0], D[0][2] + D[2][0], D[1][2] + D[2][1] of the image kernel. */"]

    def cc(e):
        t = sp.ccode(e)
        t = re.sub(r"pow\(([A-Za-z0-9_]+), (-?\d+)\)", r"OQ_IPOW(\1, \2)", t)
        assert "pow(" not in t and "sqrt(" not in t, t
        return t
    # halves: the two sums of each pair in scopes of their own (fewer values
    # live at once, some subexpressions computed twice)
    h = len(exprs) // 2
    groups = [(0, exprs[:h]), (h, exprs[h:])] if halves else [(0, exprs)]
    ntemp = 0
    for base, grp in groups:
        reps, red = sp.cse(grp, optimizations="basic")
        ntemp += len(reps)
        if halves:
            lines.append("{")
        for sym, e in reps:
            lines.append("const double %s = %s;" % (sym, cc(e)))
        for n, e in enumerate(red):
            lines.append("P[%d] = %s;" % (base + n, cc(e)))
        if halves:
            lines.append("}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", path, ntemp, "temporaries")


if __name__ == "__main__":
    if "--check" in sys.argv:
        check()
    if "--stress" in sys.argv:
        emit_traction(sys.argv[sys.argv.index("--stress") + 1], -1, halves="--halves" in sys.argv)
        sys.exit(0)
    if "--traction" in sys.argv:
        ft = int(sys.argv[sys.argv.index("--traction") + 1])
        emit_traction(sys.argv[sys.argv.index("--traction") + 2], ft, halves="--halves" in sys.argv)
        sys.exit(0)
    for a in sys.argv[1:]:
        if not a.startswith("--"):
            emit(a, device="--device" in sys.argv)
