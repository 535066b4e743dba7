"""SDF::CreateFromMesh on the GPU (scCreateSDFFromMesh, over the mesh's
triangle LBVH) against the compiled reference's values
(tests/golden/ref_sdf_from_mesh.npz: the reference's own tree and query at
every node) and against the oracle's brute-force restatement on a larger
grid, bit for bit."""
import os
import time

import numpy as np
import pytest

from solidcloth_amd import scenes as S

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "ref_sdf_from_mesh.npz"))


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(int(G["cases"])))
def test_sdf_from_mesh_matches_the_reference(sc, case):
    g = lambda k: G[f"c{case}_{k}"]
    box = g("box")
    print(f"sdf from mesh golden case {case}", flush=True)
    sdf = sc.SDF.from_mesh(sc.Mesh(g("vertices"), g("indices")), tuple(int(c) for c in g("cells")), box[:3], box[3:])
    got = sdf.values()
    assert np.array_equal(got.view(np.uint32), g("values").view(np.uint32))


@pytest.mark.gpu
def test_sdf_from_mesh_matches_the_oracle_on_a_larger_grid(sc, oracle_mod):
    v, t = S.icosphere(3, 0.4, (0.013, -0.021, 0.007))
    v = np.asarray(v, np.float32)
    t = np.asarray(t, np.uint32)
    cells, lo, hi = (48, 40, 44), (-0.55, -0.6, -0.5), (0.6, 0.5, 0.55)
    print("sdf from mesh 48x40x44 over 1280 triangles", flush=True)
    t0 = time.perf_counter()
    sdf = sc.SDF.from_mesh(sc.Mesh(v, t), cells, lo, hi)
    got = sdf.values()
    gpu_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    want = oracle_mod.sdf_from_mesh(v, t, cells, lo, hi)
    print(f"GPU {gpu_s * 1e3:.1f} ms (with the upload and LBVH build), oracle {time.perf_counter() - t0:.2f} s",
          flush=True)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the unsigned distance of an icosphere of radius 0.4: about | |p - c| - 0.4 |
    # (a flat facet lies up to 0.4 (1 - cos of half its angle) inside the sphere)
    nx, ny, nz = cells
    x = lo[0] + np.arange(nx + 1) * (hi[0] - lo[0]) / nx
    y = lo[1] + np.arange(ny + 1) * (hi[1] - lo[1]) / ny
    z = lo[2] + np.arange(nz + 1) * (hi[2] - lo[2]) / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    r = np.sqrt((X - 0.013) ** 2 + (Y + 0.021) ** 2 + (Z - 0.007) ** 2)
    assert np.abs(got - np.abs(r - 0.4)).max() < 0.02


@pytest.mark.gpu
def test_sdf_from_mesh_on_one_triangle(sc):
    """One triangle (0,0,0), (1,0,0), (0,0,1) on the unit cell: a node at one of
    its vertices has every triangle at distance 0, which the reference's query
    skips, so it keeps FLT_MAX; (1, 0, 1) is 1/sqrt(2) from the hypotenuse and
    (1, 1, 1) sqrt(3/2) from its midpoint; nodes above a vertex are 1 away."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    t = np.array([0, 1, 2], np.uint32)
    vals = sc.SDF.from_mesh(sc.Mesh(v, t), (1, 1, 1), (0, 0, 0), (1, 1, 1)).values()  # [z, y, x]
    big = np.float32(3.4028235e38)
    assert vals[0, 0, 0] == big and vals[0, 0, 1] == big and vals[1, 0, 0] == big
    assert vals[1, 0, 1] == pytest.approx(np.sqrt(0.5), abs=1e-7)
    assert vals[1, 1, 1] == pytest.approx(np.sqrt(1.5), abs=1e-7)
    assert vals[0, 1, 0] == 1.0 and vals[0, 1, 1] == 1.0 and vals[1, 1, 0] == 1.0
