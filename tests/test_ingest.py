"""Request ingestion: gmsh MSH 4.1 and CBOR (host code, no GPU).

- gmsh: the reference's own cases (regge/src/io/gmsh.rs:178-241): the two
  counterclockwise triangles of the unit square read as two cells, four
  vertices and the file's node words [0, 1, 2], [1, 3, 2], and every
  not-a-mesh input is an error, never a crash; plus renumbering of unused
  nodes, lower-dimensional elements as faces, skipped non-simplex elements and
  sections, and a 3D Kuhn mesh written and read back.
- CBOR: the decoder against RFC 8949 Appendix A's known answers (integers,
  floats of every width, nested and indefinite containers, tags); the writer's
  float widths against the same table; round trips of cells and lengths. The
  schema (serde's struct encoding of Skeleton's simplices and nalgebra's
  DVector storage through ciborium 0.2.2) is restated from the serde data
  model: parity unpinned by a reference-written file, none ships with the
  reference.
"""
import struct

import numpy as np
import pytest

import formoniq_amd as fa
from oracle import oracle as o

SQUARE = b"""$MeshFormat
4.1 0 8
$EndMeshFormat
$Nodes
1 4 1 4
2 1 0 4
1
2
3
4
0 0 0
1 0 0
0 1 0
1 1 0
$EndNodes
$Elements
1 2 1 2
2 1 2 2
1 1 2 3
2 2 4 3
$EndElements
"""


def test_gmsh_square_keeps_the_files_node_words():
    dim, cells, coords = fa.gmsh2coord_cells(SQUARE)
    assert dim == 2
    np.testing.assert_array_equal(cells, [[0, 1, 2], [1, 3, 2]])
    np.testing.assert_array_equal(coords, [[0, 0], [1, 0], [0, 1], [1, 1]])


@pytest.mark.parametrize("data", [
    b"not a mesh at all",
    b"",
    b"$MeshFormat\n4.1 0 8\n$EndMeshFormat\n",
    b"$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 4 1 4\n",
    b"$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 1 1 1\n2 1 0 1\n1\n0 0 x\n$EndNodes\n",
    b"$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 999999999 1 999999999\n",
    SQUARE.replace(b"2 2 4 3", b"2 2 9 3"),   # a node the file lacks
    SQUARE.replace(b"1 1 2 3", b"1 0 2 3"),   # node tag 0
])
def test_gmsh_not_a_mesh_is_an_error(data):
    with pytest.raises(fa.FeecError) as e:
        fa.gmsh2coord_cells(data)
    assert e.value.status == 1


def test_gmsh_binary_and_old_versions_are_unsupported():
    for head in (b"4.1 1 8", b"2.2 0 8"):
        with pytest.raises(fa.FeecError) as e:
            fa.gmsh2coord_cells(SQUARE.replace(b"4.1 0 8", head))
        assert e.value.status == 8


def _msh(nodes, blocks, extra=b""):
    """An MSH 4.1 ASCII file: nodes [n, 3] in one block, element blocks
    [(dim, type, [[tag nodes...], ...])]."""
    out = [b"$MeshFormat\n4.1 0 8\n$EndMeshFormat\n", extra]
    n = len(nodes)
    out.append(b"$Nodes\n1 %d 1 %d\n3 1 0 %d\n" % (n, n, n))
    out += [b"%d\n" % (i + 1) for i in range(n)]
    out += [(" ".join(repr(float(x)) for x in p) + "\n").encode() for p in nodes]
    out.append(b"$EndNodes\n$Elements\n")
    total = sum(len(b[2]) for b in blocks)
    out.append(b"%d %d 1 %d\n" % (len(blocks), total, total))
    tag = 1
    for dim, typ, els in blocks:
        out.append(b"%d 1 %d %d\n" % (dim, typ, len(els)))
        for e in els:
            out.append((" ".join(str(x) for x in [tag] + [v + 1 for v in e]) + "\n").encode())
            tag += 1
    out.append(b"$EndElements\n")
    return b"".join(out)


def test_gmsh_cells_are_the_top_dimension_and_unused_nodes_go():
    nodes = [[0, 0, 0], [9, 9, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]]  # node 1 is unused
    data = _msh(nodes, [
        (1, 1, [[0, 2], [2, 4]]),            # lines: faces, not cells
        (2, 3, [[0, 2, 4, 3]]),              # a quadrangle: skipped
        (2, 2, [[0, 2, 3], [2, 4, 3]]),
        (0, 15, [[4]]),                      # a point
    ], extra=b"$PhysicalNames\n1\n2 1 \"domain\"\n$EndPhysicalNames\n")
    dim, cells, coords = fa.gmsh2coord_cells(data)
    assert dim == 2
    # nodes 0, 2, 3, 4 are used: renumbered 0, 1, 2, 3 by tag order
    np.testing.assert_array_equal(cells, [[0, 1, 2], [1, 3, 2]])
    np.testing.assert_array_equal(coords, [[0, 0], [1, 0], [0, 1], [1, 1]])
    # z is dropped only when it is 0 on every node the file lists, used or not
    _, _, c3 = fa.gmsh2coord_cells(_msh([[0, 0, 0], [9, 9, 9]] + nodes[2:],
                                        [(2, 2, [[0, 2, 3], [2, 4, 3]])]))
    assert c3.shape == (4, 3)
    only_quads = _msh(nodes[:4], [(2, 3, [[0, 1, 2, 3]])])
    with pytest.raises(fa.FeecError):
        fa.gmsh2coord_cells(only_quads)


def test_gmsh_3d_kuhn_mesh_round_trip():
    cells, coords = o.kuhn(3, 3)
    coords = coords + 0.01 * np.sin(np.arange(coords.size)).reshape(coords.shape)
    rng = np.random.default_rng(0)
    words = np.array([rng.permutation(c) for c in cells])  # any node order per element
    dim, rc, rx = fa.gmsh2coord_cells(_msh(coords, [(3, 4, words.tolist())]))
    assert dim == 3 and rx.shape == coords.shape
    np.testing.assert_array_equal(rc, words)
    np.testing.assert_array_equal(rx, coords)


# ---- CBOR ------------------------------------------------------------------------------
# RFC 8949 Appendix A: (encoding, value)
RFC_NUMBERS = [
    ("00", 0), ("01", 1), ("0a", 10), ("17", 23), ("1818", 24), ("1819", 25), ("1864", 100),
    ("1903e8", 1000), ("1a000f4240", 1000000), ("1b000000e8d4a51000", 1000000000000),
    ("20", -1), ("29", -10), ("3863", -100), ("3903e7", -1000),
    ("f90000", 0.0), ("f98000", -0.0), ("f93c00", 1.0), ("fb3ff199999999999a", 1.1),
    ("f93e00", 1.5), ("f97bff", 65504.0), ("fa47c35000", 100000.0),
    ("fa7f7fffff", 3.4028234663852886e+38), ("fb7e37e43c8800759c", 1.0e+300),
    ("f90001", 5.960464477539063e-8), ("f90400", 0.00006103515625), ("f9c400", -4.0),
    ("fbc010666666666666", -4.1), ("f97c00", float("inf")), ("f9fc00", float("-inf")),
    ("fa7f800000", float("inf")), ("fb7ff0000000000000", float("inf")),
    ("c11a514b67b0", 1363896240), ("c1fb41d452d9ec200000", 1363896240.5),
]
VECTOR = bytes.fromhex("a1") + b"\x66vector"


@pytest.mark.parametrize("enc,value", RFC_NUMBERS)
def test_cbor_decodes_rfc8949_numbers(enc, value):
    got = fa.load_cbor_lengths(VECTOR + bytes.fromhex("81" + enc))
    assert got.shape == (1,)
    assert got[0] == value and np.signbit(got[0]) == np.signbit(value)


def test_cbor_decodes_containers():
    # [1, [2, 3], [4, 5]] with indefinite outer and inner arrays is not a vector of numbers,
    # but its first element array form is read as a skeleton: [[1,2,3],[4,5,6]]
    d, cells = fa.load_cbor_cells(bytes.fromhex("829f010203ff83040506"))
    assert d == 2
    np.testing.assert_array_equal(cells, [[1, 2, 3], [4, 5, 6]])
    # a map with an indefinite text key ("vertices" in two chunks)
    key = bytes.fromhex("7f") + b"\x64vert" + b"\x64ices" + bytes.fromhex("ff")
    d, cells = fa.load_cbor_cells(bytes.fromhex("81a1") + key + bytes.fromhex("820709"))
    assert d == 1
    np.testing.assert_array_equal(cells, [[7, 9]])
    # the storage form of the vector, and nrows checked against it
    storage = bytes.fromhex("a1") + b"\x66vector" + bytes.fromhex("a3") + b"\x64data" + \
        bytes.fromhex("82f93c00f94000") + b"\x65nrows" + bytes.fromhex("02") + b"\x65ncols" + \
        bytes.fromhex("f6")
    np.testing.assert_array_equal(fa.load_cbor_lengths(storage), [1.0, 2.0])
    with pytest.raises(fa.FeecError):
        fa.load_cbor_lengths(storage.replace(b"\x65nrows\x02", b"\x65nrows\x03"))


@pytest.mark.parametrize("data", [
    b"", bytes.fromhex("9f01"), bytes.fromhex("1b0000"), bytes.fromhex("83010203ff"),
    bytes.fromhex("9b7fffffffffffffff"), bytes.fromhex("ff"), bytes.fromhex("a1"),
    bytes.fromhex("81a1") + b"\x68vertices" + bytes.fromhex("8120"),  # a negative vertex
    bytes.fromhex("82a1") + b"\x68vertices" + bytes.fromhex("8101") +
    bytes.fromhex("a1") + b"\x68vertices" + bytes.fromhex("820102"),  # mixed dimensions
    b"\x9f" * 100,                                                       # nested too deeply
])
def test_cbor_malformed_is_an_error(data):
    with pytest.raises(fa.FeecError):
        fa.load_cbor_cells(data)


def test_cbor_writer_uses_the_shortest_exact_float():
    for enc, value in RFC_NUMBERS:
        if not isinstance(value, float) or enc.startswith("c1") or enc in ("fa7f800000",
                                                                          "fb7ff0000000000000"):
            continue  # integers, tags, and the wider encodings of infinity
        b = fa.save_cbor_lengths([value])
        head = VECTOR + bytes.fromhex("a3") + b"\x64data" + bytes.fromhex("81")
        assert b.startswith(head)
        assert b[len(head):len(head) + len(enc) // 2] == bytes.fromhex(enc), (enc, value)


def test_cbor_skeleton_layout_and_round_trips():
    b = fa.save_cbor_cells(np.array([[0, 1, 2], [1, 2, 24]]))
    assert b == bytes.fromhex("82a1") + b"\x68vertices" + bytes.fromhex("83000102") + \
        bytes.fromhex("a1") + b"\x68vertices" + bytes.fromhex("8301021818")
    cells, _ = o.kuhn(3, 4)
    d, back = fa.load_cbor_cells(fa.save_cbor_cells(cells))
    assert d == 3
    np.testing.assert_array_equal(back, cells)
    rng = np.random.default_rng(1)
    lengths = np.concatenate([rng.random(500), [0.0, -0.0, 1.0, 0.5, 2.0 ** -20, 1e300, 3.25]])
    back = fa.load_cbor_lengths(fa.save_cbor_lengths(lengths))
    np.testing.assert_array_equal(back, lengths)
    assert np.array_equal(np.signbit(back), np.signbit(lengths))


@pytest.mark.gpu
def test_gmsh_mesh_assembles_as_the_oracle():
    """A 3D mesh read from MSH text (elements in shuffled node order) builds the
    library's complex and mass matrices bit-identical to the oracle's on the
    same cells and coordinates."""
    from _common import assert_csr_parity
    cells, coords = o.kuhn(3, 4)
    rng = np.random.default_rng(3)
    inner = np.all((coords > 0) & (coords < 1), axis=1)
    coords[inner] += 0.05 / 4 * (2 * rng.random((inner.sum(), 3)) - 1)
    words = np.array([rng.permutation(c) for c in cells])
    dim, rc, rx = fa.gmsh2coord_cells(_msh(coords, [(3, 4, words.tolist())]))
    gcx = fa.Complex(dim, rc, nverts=len(rx))
    geo = fa.MeshLengthsSq.from_coords(gcx, rx)
    ocx = o.Complex(dim, np.sort(rc, axis=1))
    oL = ocx.lengths_from_coords(rx)
    wc = fa.WhitneyComplex(geo)
    for k in range(4):
        assert_csr_parity(wc.mass(k), o.assemble(ocx, oL, o.MASS, k), 0.0, f"M_{k}")
    # and a geometry stored as CBOR reads back to the same lengths
    np.testing.assert_array_equal(fa.load_cbor_lengths(fa.save_cbor_lengths(geo.lengths())), oL)
