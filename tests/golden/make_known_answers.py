"""Writes known_answers.json: hand-derived known-answer vectors for the index
map (fdutil.c:1363-1411, 1474-1539) and the spherical one-way boundary rule
(spher_utils.c:583-631).  Nothing here runs the reference or the oracle.

Index map cases put points at dyadic fractions of a cell, so every float
operation of the map is exact and the expected weights follow from the
bilinear (trilinear) formula with those fractions:
    w00 = (1-f1)(1-f2), w01 = f1(1-f2), w10 = (1-f1)f2, w11 = f1 f2
with f1 the z fraction and f2 the x fraction; the 3-D weights carry (1-f3)
or f3 for y in the upper bit.  A point outside [o, o + (n-1)d) on any axis
maps to node 0 with the whole weight on the second (z+1) node.

Rule cases: 0 none, 1 theta-low, 2 theta-high, 3 r-low, 4 r-high, 5 phi-low,
6 phi-high; the answer is the last rule the serial sweep applies at the node,
worked out by hand from the else-if choice of iop (first of r, theta, phi
within NOP = 4 of a face) and the six range tests.
"""
import json
import os
from fractions import Fraction as F

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")


def w2(f1, f2):
    return [float((1 - f1) * (1 - f2)), float(f1 * (1 - f2)), float((1 - f1) * f2), float(f1 * f2)]


def w3(f1, f2, f3):
    out = []
    for q in range(8):
        a = f1 if q & 1 else 1 - f1
        b = f2 if q & 2 else 1 - f2
        c = f3 if q & 4 else 1 - f3
        out.append(float(a * b * c))
    return out


OUTSIDE2 = [0.0, 1.0, 0.0, 0.0]
OUTSIDE3 = [0.0, 1.0] + [0.0] * 6


def lint2d_cases():
    # grid: nzpad, nxpad, ozpad, oxpad, dz, dx
    g = [20, 30, -10.0, -15.0, 2.5, 5.0]
    oz, ox, dz, dx = F(-10), F(-15), F(5, 2), F(5)

    def at(cz, cx):  # position in cells from the padded origin
        return [float(ox + cx * dx), float(oz + cz * dz)]

    cases = [
        ("on node", at(6, 5), [6, 5], w2(F(0), F(0))),
        ("mid cell", at(F(13, 2), F(11, 2)), [6, 5], w2(F(1, 2), F(1, 2))),
        ("quarter z, three quarters x", at(F(25, 4), F(23, 4)), [6, 5], w2(F(1, 4), F(3, 4))),
        ("padded origin", at(0, 0), [0, 0], w2(F(0), F(0))),
        ("inside the padding", at(1, F(1, 2)), [1, 0], w2(F(0), F(1, 2))),
        ("last z node is outside", at(19, 3), [0, 0], OUTSIDE2),
        ("last x node is outside", at(3, 29), [0, 0], OUTSIDE2),
        ("just below the last z node", at(F(75, 4), 3), [18, 3], w2(F(3, 4), F(0))),
        ("beyond x", [1000.0, 5.0], [0, 0], OUTSIDE2),
        ("above the origin", [0.0, -10.5], [0, 0], OUTSIDE2),
    ]
    return [dict(name=n, xz=p, grid=g, j=j, w=w) for n, p, j, w in cases]


def lint3d_cases():
    # grid: nzpad, nxpad, nypad, ozpad, oxpad, oypad, dz, dx, dy
    g = [16, 18, 20, -20.0, -20.0, -40.0, 5.0, 5.0, 10.0]
    oz, ox, oy, dz, dx, dy = F(-20), F(-20), F(-40), F(5), F(5), F(10)

    def at(cz, cx, cy):
        return [float(ox + cx * dx), float(oy + cy * dy), float(oz + cz * dz)]

    cases = [
        ("on node", at(7, 8, 9), [7, 8, 9], w3(F(0), F(0), F(0))),
        ("cell centre", at(F(15, 2), F(17, 2), F(19, 2)), [7, 8, 9], w3(F(1, 2), F(1, 2), F(1, 2))),
        ("f = (1/4, 1/2, 3/4)", at(F(29, 4), F(17, 2), F(39, 4)), [7, 8, 9],
         w3(F(1, 4), F(1, 2), F(3, 4))),
        ("padded origin", at(0, 0, 0), [0, 0, 0], w3(F(0), F(0), F(0))),
        ("last y node is outside", at(3, 3, 19), [0, 0, 0], OUTSIDE3),
        ("negative outside y", at(3, 3, F(-1, 2)), [0, 0, 0], OUTSIDE3),
        ("beyond z", at(16, 3, 3), [0, 0, 0], OUTSIDE3),
    ]
    return [dict(name=n, xyz=p, grid=g, j=j, w=w) for n, p, j, w in cases]


def rule_cases():
    grid = [20, 18, 16]      # padded n_r, n_theta, n_phi
    c = [10, 9, 8]           # an interior node
    cases = [
        # r faces (theta, phi interior: iop comes from r)
        ((0, 9, 8), 3), ((1, 9, 8), 3), ((2, 9, 8), 3), ((3, 9, 8), 0), ((4, 9, 8), 3),
        ((5, 9, 8), 0), ((14, 9, 8), 0), ((15, 9, 8), 4), ((16, 9, 8), 4), ((17, 9, 8), 0),
        ((18, 9, 8), 4), ((19, 9, 8), 4),
        # theta faces
        ((10, 0, 8), 1), ((10, 3, 8), 0), ((10, 4, 8), 1), ((10, 12, 8), 0), ((10, 13, 8), 2),
        ((10, 14, 8), 2), ((10, 17, 8), 2),
        # phi faces
        ((10, 9, 0), 5), ((10, 9, 3), 0), ((10, 9, 10), 0), ((10, 9, 11), 6), ((10, 9, 12), 6),
        ((10, 9, 15), 6),
        # edges and corners: later rules overwrite earlier ones
        ((0, 0, 8), 3), ((0, 0, 0), 5), ((10, 0, 15), 6), ((3, 0, 8), 1), ((3, 2, 8), 0),
        ((2, 2, 2), 5),
        (tuple(c), 0),
    ]
    return [dict(node=list(n), grid=grid, rule=r) for n, r in cases]


def main():
    data = {"lint2d": lint2d_cases(), "lint3d": lint3d_cases(), "spher3d_rule": rule_cases()}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
