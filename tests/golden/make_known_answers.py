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

Reflectivity cases (expl=y injection, spher_utils.c:153-173 in 2-D,
250-310 in 3-D): the reference takes |v0-vk|/(v0+vk) of the FIRST
neighbour k, in its fixed search order, whose velocity differs from the
source node's.  Case k sets every neighbour at search position >= k to its
own velocity and leaves the earlier ones at v0, so the answer names position
k.  The 3-D list holds (phi+1, r-1) twice (positions 16 and 18): the second
copy can never be the first differing neighbour, so case 18 answers with
position 19.  The orders below are transcribed from the reference's
else-if chains, one entry per branch, as (dr, dtheta[, dphi]) with
x = r, z = theta, y = phi of its index expressions.

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


# spher_utils.c:155-170, branch by branch: z+1, x+1, z+1 x+1, x-1, z-1 x-1,
# z+1 x-1, z-1 x+1 (x = r, z = theta)
REFL2D_ORDER = [(0, 1), (1, 0), (1, 1), (-1, 0), (-1, -1), (-1, 1), (1, -1)]
# spher_utils.c:250-304: cardinal (252-262), diagonals (264-286), corners
# (288-304); (dr, dtheta, dphi) with x = r, z = theta, y = phi
REFL3D_ORDER = [
    (0, 1, 0), (0, 0, 1), (1, 0, 0), (0, -1, 0), (0, 0, -1), (-1, 0, 0),
    (1, 1, 0), (-1, -1, 0), (-1, 1, 0), (1, -1, 0),
    (0, 1, 1), (0, -1, -1), (0, -1, 1), (0, 1, -1),
    (1, 0, 1), (-1, 0, -1), (-1, 0, 1), (1, 0, -1),
    (-1, 0, 1),
    (1, 1, 1), (-1, 1, 1), (1, -1, 1), (-1, -1, 1),
    (1, 1, -1), (-1, 1, -1), (1, -1, -1), (-1, -1, -1),
]


def _f32_refl(v0, vk):
    """fabs((v0 - vk) / (v0 + vk)) in float32, as the reference evaluates it."""
    import numpy as np
    a, b = np.float32(v0), np.float32(vk)
    return float(np.float32(abs(np.float32(np.float32(a - b) / np.float32(a + b)))))


def refl_cases(order, grid, node):
    """One case per search position k: neighbours at positions >= k get
    velocity V0 + (pos + 1) * 2^20 (distinct, exact in float32), the rest V0."""
    v0 = 1.68e8
    cases = []
    for k in range(len(order)):
        early = set(tuple(o) for o in order[:k])   # these nodes keep v0
        nbr = {}
        for pos in range(k, len(order)):
            off = tuple(order[pos])
            if off not in early:
                nbr.setdefault(off, v0 + (pos + 1) * 2.0 ** 20)
        first = next(p for p in range(k, len(order)) if tuple(order[p]) in nbr)
        vk = nbr[tuple(order[first])]
        cases.append(dict(case=k, grid=grid, node=node, v0=v0,
                          set=[list(o) + [v] for o, v in nbr.items()],
                          position=first, r=_f32_refl(v0, vk)))
    # homogeneous neighbourhood: nothing differs, r = 0
    cases.append(dict(case=len(order), grid=grid, node=node, v0=v0, set=[], position=-1, r=0.0))
    return cases


def main():
    data = {"lint2d": lint2d_cases(), "lint3d": lint3d_cases(), "spher3d_rule": rule_cases(),
            # padded grids (n_r, n_theta[, n_phi]) and the source node (r, theta[, phi])
            "refl2d": refl_cases(REFL2D_ORDER, [12, 10], [5, 4]),
            "refl3d": refl_cases(REFL3D_ORDER, [10, 9, 8], [4, 4, 3])}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
