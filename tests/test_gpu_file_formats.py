"""Obstacles read from files on the GPU substep: a drape over a sphere SDF
that went through the binary SDF file, and over an icosphere mesh that went
through a Wavefront OBJ file (naive loading keeps the vertex order), each
bit-exact against the oracle built from the original arrays."""
import numpy as np
import pytest

from solidcloth_amd import scenes as S
from parity_util import pair, same_positions

pytestmark = pytest.mark.gpu


def test_sdf_from_file_drape_bit_exact(sc, oracle_mod, tmp_path, monkeypatch):
    real = sc.SDF

    def via_file(values, n, bmin, bmax):
        p = tmp_path / "sphere.sdf"
        real(values, n, bmin, bmax).write(p)
        return real.from_file(p)

    monkeypatch.setattr(sc, "SDF", via_file)
    scene = S.DRAPE256.with_(n=40)
    print("sdf via file, drape n=40", flush=True)
    world, patch, m, o = pair(sc, oracle_mod, scene)
    for f in range(16):
        patch.Step(scene.dt)
        o.step(scene.dt)
        same_positions(m, o, f"sdf-file frame {f}")


def test_mesh_from_obj_drape_bit_exact(sc, oracle_mod, tmp_path, monkeypatch):
    real = sc.Mesh

    def via_obj(v, t):
        p = tmp_path / "ico.obj"
        with open(p, "w") as fh:
            for x in np.asarray(v, dtype=np.float32).reshape(-1, 3):
                fh.write("v %.9g %.9g %.9g\n" % tuple(float(c) for c in x))
            for tri in np.asarray(t).reshape(-1, 3):
                fh.write("f %d %d %d\n" % tuple(int(k) + 1 for k in tri))
        m = real.from_obj(p, naive=True)
        assert np.array_equal(m.vertices, np.asarray(v, dtype=np.float32).reshape(-1, 3))
        return m

    monkeypatch.setattr(sc, "Mesh", via_obj)
    scene = S.MESHSMALL.with_(n=40, mesh_subdiv=3, offset=(0.0, 0.265, 0.0), frames=6)
    print("mesh via obj, n=40 subdiv 3", flush=True)
    world, patch, m, o = pair(sc, oracle_mod, scene)
    for f in range(6):
        patch.Step(scene.dt)
        o.step(scene.dt)
        same_positions(m, o, f"obj-mesh frame {f}")
