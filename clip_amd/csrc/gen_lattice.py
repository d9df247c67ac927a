#!/usr/bin/env python3
"""Generate lattice_tables.inc: D2Q9 / D3Q19 velocity sets, weights and the
moment transforms of the weighted-MRT operator.

Velocity order and weights follow CLIP's WMRT::WMRTvelSet
(include/solver/WMRT.cuh:189-206); the order is part of the checkpoint format
(populations are stored by direction index).

The moment matrix M is built from its defining polynomials of the lattice
velocity e = (ex, ey[, ez]) — the rows CLIP's convertD2Q9Weighted /
convertD3Q19Weighted apply (WMRT.cuh:54-126) — and M^-1 (reconvert*,
WMRT.cuh:77-184) is obtained here by exact rational inversion, so neither
table is typed in by hand.  Stress moments (relaxed with s = 1/(tau+1/2)):
D2Q9 rows {7, 8}, D3Q19 rows {4..8} (NsAllen.cu:428-429, 448-452).

Run:  python gen_lattice.py > lattice_tables.inc
"""
from fractions import Fraction as Fr

D2Q9_E = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]
D2Q9_W = [Fr(4, 9)] + [Fr(1, 9)] * 4 + [Fr(1, 36)] * 4

D3Q19_E = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
           (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1),
           (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)]
D3Q19_W = [Fr(1, 3)] + [Fr(1, 18)] * 6 + [Fr(1, 36)] * 12

# In-directions of the minus side of each axis, in the order of the
# reference's boundary maps: WMRT::wallBCMap (WMRT.cuh:214-237) and
# WMRT::slipWallBCMap (WMRT.cuh:243-264) list the same directions in the same
# order for X-, Y- and Z-.  Entry m of a plus side's list is the opposite of
# the minus side's entry m in the wall map and its mirror image along the
# axis in the slip map.
D2Q9_IN_MINUS = [[1, 5, 8], [2, 5, 6]]
D3Q19_IN_MINUS = [[1, 9, 7, 13, 11], [3, 10, 7, 17, 15], [5, 14, 11, 18, 15]]


def d2q9_rows():
    def e2(x, y):
        return x * x + y * y
    return [
        lambda x, y: 1,
        lambda x, y: 3 * e2(x, y) - 4,
        lambda x, y: Fr(9, 2) * e2(x, y) ** 2 - Fr(21, 2) * e2(x, y) + 4,
        lambda x, y: x,
        lambda x, y: (3 * e2(x, y) - 5) * x,
        lambda x, y: y,
        lambda x, y: (3 * e2(x, y) - 5) * y,
        lambda x, y: x * x - y * y,
        lambda x, y: x * y,
    ]


def d3q19_rows():
    def e2(x, y, z):
        return x * x + y * y + z * z
    return [
        lambda x, y, z: 1,
        lambda x, y, z: x,
        lambda x, y, z: y,
        lambda x, y, z: z,
        lambda x, y, z: x * y,
        lambda x, y, z: y * z,
        lambda x, y, z: x * z,
        lambda x, y, z: 3 * x * x - e2(x, y, z),
        lambda x, y, z: y * y - z * z,
        lambda x, y, z: e2(x, y, z) - 1,
        lambda x, y, z: (3 * e2(x, y, z) - 5) * x,
        lambda x, y, z: (3 * e2(x, y, z) - 5) * y,
        lambda x, y, z: (3 * e2(x, y, z) - 5) * z,
        lambda x, y, z: x * (y * y - z * z),
        lambda x, y, z: y * (z * z - x * x),
        lambda x, y, z: z * (x * x - y * y),
        lambda x, y, z: 3 * e2(x, y, z) ** 2 - 6 * e2(x, y, z) + 1,
        lambda x, y, z: (2 * e2(x, y, z) - 3) * (3 * x * x - e2(x, y, z)),
        lambda x, y, z: (2 * e2(x, y, z) - 3) * (y * y - z * z),
    ]


def build(rows, E):
    return [[Fr(r(*e)) for e in E] for r in rows]


def inverse(M):
    n = len(M)
    A = [row[:] + [Fr(int(i == j)) for j in range(n)] for i, row in enumerate(M)]
    for c in range(n):
        p = next(r for r in range(c, n) if A[r][c] != 0)
        A[c], A[p] = A[p], A[c]
        pv = A[c][c]
        A[c] = [v / pv for v in A[c]]
        for r in range(n):
            if r != c and A[r][c] != 0:
                f = A[r][c]
                A[r] = [a - f * b for a, b in zip(A[r], A[c])]
    return [row[n:] for row in A]


def lcm(a, b):
    from math import gcd
    return a * b // gcd(a, b)


def mirror(E, q, a):
    return E.index(tuple(-c if d == a else c for d, c in enumerate(E[q])))


def side_lists(E, in_minus):
    """Per side (X-, X+, Y-, Y+[, Z-, Z+]): the wall map's and the slip map's
    direction lists."""
    opp = [E.index(tuple(-c for c in e)) for e in E]
    wall, slip = [], []
    for a, lst in enumerate(in_minus):
        wall += [list(lst), [opp[q] for q in lst]]
        slip += [list(lst), [mirror(E, q, a) for q in lst]]
    return wall, slip


def emit(name, E, W, rows, stress, in_minus):
    Q = len(E)
    D = len(E[0])
    M = build(rows, E)
    Mi = inverse(M)
    out = []
    out.append(f"// ---- {name} (generated by gen_lattice.py; do not edit) ----")
    comps = "xyz"
    for d in range(3):
        vals = [e[d] if d < D else 0 for e in E]
        out.append(f"static constexpr int {name}_E{comps[d]}[{Q}] = {{{', '.join(map(str, vals))}}};")
    out.append(f"static constexpr double {name}_W[{Q}] = {{{', '.join(f'{w.numerator}.0 / {w.denominator}.0' for w in W)}}};")
    out.append(f"static constexpr int {name}_M[{Q}][{Q}] = {{")
    for r in M:
        assert all(v.denominator == 1 for v in r)
        out.append("    {" + ", ".join(str(int(v)) for v in r) + "},")
    out.append("};")
    # M^-1 rows as integer numerators over a per-row common denominator.
    dens = []
    nums = []
    for r in Mi:
        den = 1
        for v in r:
            den = lcm(den, v.denominator)
        dens.append(den)
        nums.append([int(v * den) for v in r])
    out.append(f"static constexpr int {name}_MINV_NUM[{Q}][{Q}] = {{")
    for r in nums:
        out.append("    {" + ", ".join(map(str, r)) + "},")
    out.append("};")
    out.append(f"static constexpr int {name}_MINV_DEN[{Q}] = {{{', '.join(map(str, dens))}}};")
    out.append(f"static constexpr int {name}_STRESS[{len(stress)}] = {{{', '.join(map(str, stress))}}};")
    # opposite direction
    opp = [E.index(tuple(-c for c in e)) for e in E]
    out.append(f"static constexpr int {name}_OPP[{Q}] = {{{', '.join(map(str, opp))}}};")
    # direction mirrored along each axis (identity along axes beyond D)
    mir = [[mirror(E, q, a) if a < D else q for q in range(Q)] for a in range(3)]
    out.append(f"static constexpr int {name}_MIRROR[3][{Q}] = {{")
    for r in mir:
        out.append("    {" + ", ".join(map(str, r)) + "},")
    out.append("};")
    # position of each direction in each side's wall / slip list (-1: absent)
    wall, slip = side_lists(E, in_minus)
    for tag, lists in (("WALL_POS", wall), ("SLIP_POS", slip)):
        out.append(f"static constexpr int {name}_{tag}[6][{Q}] = {{")
        for sd in range(6):
            lst = lists[sd] if sd < len(lists) else []
            out.append("    {" + ", ".join(str(lst.index(q) if q in lst else -1) for q in range(Q)) + "},")
        out.append("};")
    return "\n".join(out)


if __name__ == "__main__":
    print("// lattice_tables.inc — generated by clip_amd/csrc/gen_lattice.py")
    print(emit("D2Q9", D2Q9_E, D2Q9_W, d2q9_rows(), [7, 8], D2Q9_IN_MINUS))
    print(emit("D3Q19", D3Q19_E, D3Q19_W, d3q19_rows(), [4, 5, 6, 7, 8], D3Q19_IN_MINUS))
