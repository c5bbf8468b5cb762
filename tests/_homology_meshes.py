"""Meshes with known (co)homology for the exact-topology tests: the
reference's own fixtures (homology.rs:165-210: the cubes, the two-sphere, the
square annulus; harmonic.rs:189-204: the flat tori) and a few more (a solid
torus slab, two disjoint disks). Each builder returns (dim, cells, geom):
geom is vertex coordinates, or a callable giving the squared edge lengths of
the complex's edges (skeleton(1) pairs) for the periodic and closed meshes."""
import numpy as np

from oracle import oracle as o


def kuhn_ball(dim, n):
    return o.kuhn(dim, n)


def _compact(cells, coords):
    used, inv = np.unique(cells, return_inverse=True)
    return inv.reshape(cells.shape).astype(np.int64), coords[used]


def cube(dim):
    cells, coords = o.kuhn(dim, 2)
    return dim, cells, coords


def annulus():
    """CartesianTopology::cube(2, 3) without box 4 (its cells 8 and 9)."""
    cells, coords = o.kuhn(2, 3)
    return 2, np.delete(cells, [8, 9], axis=0), coords


def tunnel():
    """The z = 0 layer of the 3x3x3 Kuhn grid without its centre box: a solid
    torus (b = 1, 1, 0, 0)."""
    cells, coords = o.kuhn(3, 3)
    boxes = [b for b in range(9) if b != 4]  # box index x + 3y + 9z, 6 cells each
    keep = np.concatenate([np.arange(6 * b, 6 * b + 6) for b in boxes])
    c, x = _compact(cells[keep], coords)
    return 3, c, x


def two_disks():
    cells, coords = o.kuhn(2, 1)
    nv = len(coords)
    return 2, np.vstack([cells, cells + nv]), np.vstack([coords, coords + [3.0, 0.0]])


def two_sphere():
    """The boundary of a tetrahedron (homology.rs:170-177), unit edges."""
    cells = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]], np.int64)
    return 2, cells, lambda cx: np.ones(cx.nsimplices(1))


def _periodic(dim, n):
    cells, _ = o.kuhn(dim, n)
    nv = (n + 1) ** dim
    idx = np.arange(nv)
    digits = np.stack([(idx // (n + 1) ** a) % (n + 1) for a in range(dim)], 1) % n
    m = sum(digits[:, a] * n ** a for a in range(dim))
    tcells = np.sort(m[cells], axis=1)

    def lengths(cx):
        e = cx.skeleton(1)
        pos = np.stack([(e // n ** a) % n for a in range(dim)], -1)  # [edges, 2, dim]
        diff = pos[:, 1] - pos[:, 0]
        diff = (diff + 1) % n - 1  # the Kuhn edge steps are -1, 0 or +1 per axis
        return (diff.astype(float) ** 2).sum(1) / n ** 2
    return dim, tcells, lengths


def torus():
    """FlatQuotient::unit_torus(2, 3)."""
    return _periodic(2, 3)


def circle():
    """FlatQuotient::unit_torus(1, 3)."""
    return _periodic(1, 3)


def mobius(nx=4):
    """A Moebius strip: nx unit boxes in a row, the last one glued to the first
    with a flip, two triangles a box. Non-orientable: b = (1, 1, 0), relative
    (0, 0, 0) over Q (H_1(M, dM) is Z/2)."""
    def v(i, j):
        return (1 - j) if i == nx else 2 * i + j
    cells, sq = [], {}
    for i in range(nx):
        a, b, c, d = v(i, 0), v(i + 1, 0), v(i, 1), v(i + 1, 1)
        cells += [sorted([a, b, d]), sorted([a, c, d])]
        for e, l2 in (((a, b), 1.0), ((c, d), 1.0), ((a, c), 1.0), ((b, d), 1.0), ((a, d), 2.0)):
            sq[tuple(sorted(e))] = l2
    return 2, np.array(cells, np.int64), lambda cx: np.array([sq[tuple(e)] for e in cx.skeleton(1).tolist()])


def squared_lengths(cx, geom):
    if callable(geom):
        return np.ascontiguousarray(geom(cx), dtype=np.float64)
    return cx.lengths_from_coords(np.ascontiguousarray(geom, dtype=np.float64))


FIXTURES = {
    "cube1": lambda: cube(1), "cube2": lambda: cube(2), "cube3": lambda: cube(3),
    "two_sphere": two_sphere, "annulus": annulus, "circle": circle, "torus": torus,
    "tunnel": tunnel, "two_disks": two_disks, "mobius": mobius,
}
