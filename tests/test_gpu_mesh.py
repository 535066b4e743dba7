"""GPU parity of the triangle-mesh obstacle path (LBVH nearest-feature
detection, vertex / edge / triangle contact projection) against the CPU
oracle, through the C ABI.

The oracle searches features exhaustively (ties to the lowest index); the
GPU walks an LBVH with conservative pruning, so contact sets must be equal.
Edge and triangle contacts couple several cloth vertices; the GPU solves
them by colour groups of cloth edges / triangles, and the oracle is given the
same processing order, so positions are expected bit for bit (bar 1e-5
relative to the cloth extent, bit-exactness asserted on top).
"""
import numpy as np
import pytest

from solidcloth_amd import scenes as S

from parity_util import pair as _pair
from parity_util import same_contacts as _same_contacts
from parity_util import same_positions as _same_positions

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1])
def test_meshsmall_vertex_edge_triangle_contacts_bit_exact(sc, oracle_mod, mode):
    scene = S.MESHSMALL
    print(f"meshsmall mode={mode}", flush=True)
    world, patch, m, o = _pair(sc, oracle_mod, scene, mode)
    seen = [0, 0, 0]
    for f in range(scene.frames):
        patch.Step(scene.dt)
        o.step(scene.dt)
        counts = _same_contacts(m, o)
        seen = [a + b for a, b in zip(seen, counts)]
        _same_positions(m, o, f"meshsmall frame {f}")
    assert min(seen) > 0, seen  # every contact kind occurred


def test_meshsmall_reference_early_exit(sc, oracle_mod):
    """Early-exit mode: the exit test includes the edge and triangle contact
    errors of each iteration."""
    scene = S.MESHSMALL.with_(fixed_iterations=False, iterations=15)
    world, patch, m, o = _pair(sc, oracle_mod, scene)
    for f in range(scene.frames):
        patch.Step(scene.dt)
        o.step(scene.dt)
        assert m.stats()["iterations"] == o.stats()["iterations"], f
    _same_positions(m, o, "meshsmall early exit")


def test_mesh_vertex_contacts_fused_into_slab_sweep(sc, oracle_mod):
    """Vertex contacts only: they are applied by the FEM slab kernels of the
    next iteration, re-projected onto their obstacle triangles there."""
    scene = S.MESHSMALL.with_(flags=S.CF_VERTICES | S.CF_WALLS)
    world, patch, m, o = _pair(sc, oracle_mod, scene)
    for f in range(scene.frames):
        patch.Step(scene.dt)
        o.step(scene.dt)
    _same_contacts(m, o)
    _same_positions(m, o, "mesh vertex contacts")
    assert len(m.contacts()[0]) > 0


def test_mesh_mid_size_parity(sc, oracle_mod):
    """40x40 cloth on a subdivision-3 icosphere (642 vertices, 1280
    triangles), perturbed start, all contact kinds."""
    scene = S.MESHSMALL.with_(n=40, mesh_subdiv=3, offset=(0.0, 0.265, 0.0), frames=6)
    world, patch, m, o = _pair(sc, oracle_mod, scene)
    for f in range(scene.frames):
        patch.Step(scene.dt)
        o.step(scene.dt)
    counts = _same_contacts(m, o)
    assert min(counts) > 0, counts
    _same_positions(m, o, "mesh mid size")
