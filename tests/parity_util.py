"""Helpers shared by the GPU parity tests: pair a GPU patch with the oracle
in the GPU's processing order and compare positions and contact sets."""
import numpy as np

from solidcloth_amd import scenes as S

REL_TOL = 1e-5


def pair(sc, oracle_mod, scene, mode=1):
    world, patch = S.build_patch(scene, sc)
    m = patch.GetModel()
    m.SetSweep(mode)
    o = oracle_mod.from_scene(scene, tri_order=m.constraint_order("tri"), bend_order=m.constraint_order("bend"))
    o.set_contact_order("edge", m.constraint_order("edge_contact"))
    o.set_contact_order("tri", m.constraint_order("tri_contact"))
    return world, patch, m, o


def same_positions(m, o, what):
    """1e-5 of the cloth extent is the bar; bit-exactness is asserted on top."""
    gp, op = m.positions(), o.get("pos")
    assert np.isfinite(gp).all(), what
    ext = max(1e-6, float(np.abs(op).max()))
    err = float(np.abs(gp.astype(np.float64) - op).max())
    assert err <= REL_TOL * ext, f"{what}: max abs diff {err:.3e}"
    assert err == 0.0, f"{what}: expected bit-exact, max abs diff {err:.3e}"


def _rows(ids, data):
    return sorted((tuple(int(x) for x in i), tuple(np.float32(d).tolist())) for i, d in zip(ids, data))


def same_contacts(m, o):
    """Vertex, edge and triangle contact sets of the last substep are equal.
    Returns (vertex, edge, triangle) contact counts."""
    pts, nrm = m.contacts()
    feats = m.contact_features()
    ids, data = o.contacts()
    ofeat = o.contact_features()[:, 1]
    g = sorted(tuple(np.r_[p, n].astype(np.float32).tolist()) + (int(f),) for p, n, f in zip(pts, nrm, feats))
    r = sorted(tuple(d[:6].astype(np.float32).tolist()) + (int(f),) for d, f in zip(data, ofeat))
    assert g == r
    # edge: (cloth edge, obstacle edge) + point, normal, w1, w2, lambda, lambdaF
    gi, gd = m.edge_contacts()
    oi, od = o.edge_contacts()
    assert _rows(gi[:, :2], gd) == _rows(oi[:, 2:4], od)
    # triangle: (cloth triangle, obstacle vertex) + point, normal, w1..w3, lambda, lambdaF
    ti, td = m.tri_contacts()
    oi, od = o.tri_contacts()
    assert _rows(ti[:, :2], td) == _rows(oi, od)
    return len(pts), len(gi), len(ti)
