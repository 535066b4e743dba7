"""File formats around the substep: the binary SDF file (Geometry/SDF.cpp:
75-151, WriteToFile / LoadFromFile) and Wavefront OBJ meshes
(Geometry/Assets.cpp:309-451, LoadMeshFromObj). Host code only; the bytes are
checked against files built here by hand, the OBJ corner rules against a
small file whose answer is worked out below, and the reference's own sphere
model (when the reference tree is present) against Euler's formula."""
import os

import numpy as np
import pytest

REF_SPHERE = "/root/reference/Models/sphere.obj"


def sdf_bytes(cells, bmin, bmax, values64):
    head = np.asarray(cells, dtype=np.int32).tobytes()
    box = np.asarray([bmin[0], bmax[0], bmin[1], bmax[1], bmin[2], bmax[2]], dtype=np.float64)
    return head + box.tobytes() + np.asarray(values64, dtype=np.float64).tobytes()


def test_sdf_write_layout(sc, tmp_path):
    rng = np.random.default_rng(5)
    cells = (3, 4, 5)
    vals = rng.standard_normal((6, 5, 4)).astype(np.float32)        # (nz+1, ny+1, nx+1)
    bmin, bmax = (-1.0, -0.5, -2.0), (1.0, 1.5, 0.25)
    s = sc.SDF(vals, cells, bmin, bmax)
    p = tmp_path / "a.sdf"
    s.write(p)
    want = sdf_bytes(cells, np.float32(bmin), np.float32(bmax), vals.ravel().astype(np.float64))
    assert p.read_bytes() == want


def test_sdf_load_narrows_and_round_trips(sc, tmp_path):
    rng = np.random.default_rng(6)
    cells = (2, 3, 4)
    v64 = rng.standard_normal(3 * 4 * 5) * (1 + 1e-9)                  # not representable in float
    bmin, bmax = (0.1, -0.3, 0.7), (1.1, 0.9, 2.2)
    p = tmp_path / "b.sdf"
    p.write_bytes(sdf_bytes(cells, bmin, bmax, v64))
    s = sc.SDF.from_file(p)
    assert s.n == cells
    assert s.bmin == tuple(float(np.float32(x)) for x in bmin)
    assert s.bmax == tuple(float(np.float32(x)) for x in bmax)
    got = s.values()
    assert got.shape == (5, 4, 3)
    assert np.array_equal(got.ravel(), v64.astype(np.float32))
    q = tmp_path / "c.sdf"
    s.write(q)
    assert q.read_bytes() == sdf_bytes(cells, np.float32(bmin), np.float32(bmax),
                                       v64.astype(np.float32).astype(np.float64))


def test_sdf_file_errors(sc, tmp_path):
    with pytest.raises(sc.SolidClothError, match="cannot open"):
        sc.SDF.from_file(tmp_path / "missing.sdf")
    p = tmp_path / "short.sdf"
    p.write_bytes(sdf_bytes((2, 2, 2), (0, 0, 0), (1, 1, 1), np.zeros(26)))    # 27 nodes needed
    with pytest.raises(sc.SolidClothError, match="truncated"):
        sc.SDF.from_file(p)
    p.write_bytes(sdf_bytes((0, 2, 2), (0, 0, 0), (1, 1, 1), np.zeros(9)))
    with pytest.raises(sc.SolidClothError, match="out of range"):
        sc.SDF.from_file(p)


def test_sdf_grid_beyond_int_range_is_refused(sc, tmp_path):
    """1400^3 cells are 1401^3 = 2.75e9 nodes, past the 32-bit node index of
    the device lookup: the header alone must be refused as out of range
    (not read as a truncated file of 22 GB of values)."""
    p = tmp_path / "huge.sdf"
    p.write_bytes(sdf_bytes((1400, 1400, 1400), (0, 0, 0), (1, 1, 1), np.zeros(0)))
    with pytest.raises(sc.SolidClothError, match="out of range"):
        sc.SDF.from_file(p)
    # 1289^3 cells (1290^3 = 2,146,689,000 < 2^31 - 1 nodes) passes the range
    # check and then fails on the missing values; 1290^3 cells would not
    p.write_bytes(sdf_bytes((1289, 1289, 1289), (0, 0, 0), (1, 1, 1), np.zeros(0)))
    with pytest.raises(sc.SolidClothError, match="truncated"):
        sc.SDF.from_file(p)
    p.write_bytes(sdf_bytes((1290, 1290, 1290), (0, 0, 0), (1, 1, 1), np.zeros(0)))
    with pytest.raises(sc.SolidClothError, match="out of range"):
        sc.SDF.from_file(p)


def test_sdf_grid_copy_respects_capacity(sc):
    """scGetSDFGrid writes at most `capacity` values and returns the node count."""
    import ctypes as C

    vals = np.arange(27, dtype=np.float32).reshape(3, 3, 3)
    s = sc.SDF(vals, (2, 2, 2), (0, 0, 0), (1, 1, 1))
    out = np.full(40, -7.0, dtype=np.float32)
    n = sc.lib().scGetSDFGrid(s.h, None, None, out.ctypes.data_as(C.c_void_p), 10)
    assert n == 27
    assert np.array_equal(out[:10], vals.ravel()[:10]) and np.all(out[10:] == -7.0)
    n = sc.lib().scGetSDFGrid(s.h, None, None, out.ctypes.data_as(C.c_void_p), 40)
    assert n == 27 and np.array_equal(out[:27], vals.ravel()) and np.all(out[27:] == -7.0)


QUAD_OBJ = """# two triangles and a quad face
o quad
v 0 0 0
v 0.1 0 0
v 0.1 0.2 0
v 0 0.2 0.3
vt 0 0
vt 1 0
vt 1 1
vt 0 1 0.5
vn 0 0 1
f 1/1/1 2/2/1 3/3/1
f 1/1/1 3/3/1 4/4/1
f 4/4 2/3 3/3 1/1
"""


def test_obj_corners(sc, tmp_path):
    p = tmp_path / "quad.obj"
    p.write_text(QUAD_OBJ)
    P = np.array([[0, 0, 0], [0.1, 0, 0], [0.1, 0.2, 0], [0, 0.2, 0.3]], dtype=np.float32)
    # corners (position, normal, uv), 0-based, normal 0 where absent:
    # face 1: (0,0,0) (1,0,1) (2,0,2) -> 0 1 2; face 2: (0,0,0) (2,0,2) (3,0,3) -> 0 2 3;
    # face 3 keeps its first three corners: (3,0,3) (1,0,2) (2,0,2) -> 3 4 2.
    # The last vt line has three numbers: u is the first, v the last.
    m = sc.Mesh.from_obj(p)
    assert np.array_equal(m.indices, [0, 1, 2, 0, 2, 3, 3, 4, 2])
    assert np.array_equal(m.vertices, P[[0, 1, 2, 3, 1]])
    T = np.array([[0, 0], [1, 0], [1, 1], [0, 0.5]], dtype=np.float32)
    assert np.array_equal(m.uvs, T[[0, 1, 2, 3, 2]])
    n = sc.Mesh.from_obj(p, naive=True)
    assert np.array_equal(n.indices, [0, 1, 2, 0, 2, 3, 3, 1, 2])
    assert np.array_equal(n.vertices, P)
    assert n.uvs is None


def test_obj_errors(sc, tmp_path):
    with pytest.raises(sc.SolidClothError, match="cannot open"):
        sc.Mesh.from_obj(tmp_path / "missing.obj")
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n")
    with pytest.raises(sc.SolidClothError, match="missing vertex"):
        sc.Mesh.from_obj(p, naive=True)
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 -1\n")
    with pytest.raises(sc.SolidClothError, match="unsupported index"):
        sc.Mesh.from_obj(p)


@pytest.mark.skipif(not os.path.exists(REF_SPHERE), reason="reference models not present")
def test_obj_reference_sphere(sc):
    n = sc.Mesh.from_obj(REF_SPHERE, naive=True)
    V = sum(1 for line in open(REF_SPHERE) if line.startswith("v "))
    F = sum(1 for line in open(REF_SPHERE) if line.startswith("f "))
    assert len(n.vertices) == V and len(n.indices) == 3 * F
    tri = n.indices.reshape(-1, 3).astype(np.int64)
    edges = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    uniq, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2), "a closed sphere shares every edge between two triangles"
    assert V - len(uniq) + F == 2
    m = sc.Mesh.from_obj(REF_SPHERE)
    assert len(m.indices) == 3 * F and m.uvs is not None and len(m.uvs) == len(m.vertices)
    # welding the non-naive vertices by position gives back the sphere's positions
    assert len(np.unique(m.vertices, axis=0)) == len(np.unique(n.vertices, axis=0))
    assert np.array_equal(m.vertices[m.indices.reshape(-1, 3)], n.vertices[n.indices.reshape(-1, 3)])
