"""Synthetic scenes for the cloth substep workloads.

The reference's frame is y-up with gravity (0, -9.8, 0) x unit
(Physics/Common.h:18, Physics/ClothPBD.cpp:263); the workloads use unit = 1.
A "z = 0" cloth of the workload descriptions is the reference's horizontal
grid (Physics/ClothPatch.cpp:181), which lies in the y = offset.y plane.

  drape32   32x32 grid, side 1 m, two top corners pinned, FEM (Young 1e4,
            Poisson 0.3) + dihedral bending 0.01, dt 1/60, 4 substeps,
            10 iterations, no collisions, 100 frames.
  drape256  256x256 grid, side 2 m, dropped from 0.5 m onto a sphere of
            radius 0.3 given as a 64^3-cell grid SDF (analytic distance
            sampled in FP32) plus a ground plane; 20 iterations, 4 substeps,
            friction 0.3, thickness 2 mm, 200 frames.
"""
from dataclasses import dataclass, field, replace

import numpy as np

CF_VERTICES, CF_TRIANGLES, CF_SELF, CF_WALLS, CF_EDGES = 1, 2, 4, 8, 16


@dataclass(frozen=True)
class Scene:
    name: str
    n: int
    side: float
    offset: tuple = (0.0, 0.0, 0.0)
    horizontal: bool = True
    attached: bool = False
    unit: float = 1.0
    young: float = 1e4
    poisson: float = 0.3
    bend: float = 0.01
    dt: float = 1.0 / 60.0
    substeps: int = 4
    iterations: int = 10
    fixed_iterations: bool = True
    frames: int = 100
    flags: int = 0
    thickness: float = 0.3
    tolerance: float = 0.2
    friction: float = 0.4
    # sphere SDF obstacle (None = no SDF)
    sphere_center: tuple = None
    sphere_radius: float = 0.0
    sdf_cells: int = 0
    sdf_half: float = 0.0
    # walls box (None = no walls)
    walls_min: tuple = None
    walls_max: tuple = None
    extra: dict = field(default_factory=dict)

    @property
    def inc(self):
        return np.float32(self.side / (self.n - 1))

    @property
    def num_vertices(self):
        return self.n * self.n

    @property
    def num_triangles(self):
        return 2 * (self.n - 1) * (self.n - 1)

    def with_(self, **kw):
        return replace(self, **kw)


DRAPE32 = Scene(name="drape32", n=32, side=1.0, attached=True, iterations=10, frames=100)

DRAPE256 = Scene(
    name="drape256", n=256, side=2.0, offset=(0.0, 0.5, 0.0), iterations=20, frames=200,
    flags=CF_VERTICES | CF_WALLS, thickness=0.002, tolerance=0.005, friction=0.3,
    sphere_center=(0.0, 0.0, 0.0), sphere_radius=0.3, sdf_cells=64, sdf_half=0.4,
    walls_min=(-10.0, -0.3, -10.0), walls_max=(10.0, 10.0, 10.0),
)


def sdf_box(scene):
    c = np.asarray(scene.sphere_center, dtype=np.float32)
    h = np.float32(scene.sdf_half)
    return tuple(float(v) for v in c - h), tuple(float(v) for v in c + h)


def sphere_sdf_values(scene):
    """Node samples of the signed distance to the sphere, FP32, x fastest.

    Node positions follow SDF::GetPosAt (Geometry/SDF.h:64-70): index times
    the cell step (extent times 1/cells, Geometry/SDF.cpp:37-39) plus the box
    minimum.
    """
    n = scene.sdf_cells
    bmin, bmax = sdf_box(scene)
    bmin = np.asarray(bmin, dtype=np.float32)
    bmax = np.asarray(bmax, dtype=np.float32)
    step = (bmax - bmin) * (np.float32(1.0) / np.float32(n))
    idx = np.arange(n + 1, dtype=np.float32)
    # axis order z (slowest), y, x (fastest)
    Z, Y, X = np.meshgrid(idx, idx, idx, indexing="ij")
    px = X * step[0] + bmin[0]
    py = Y * step[1] + bmin[1]
    pz = Z * step[2] + bmin[2]
    c = np.asarray(scene.sphere_center, dtype=np.float32)
    dx, dy, dz = px - c[0], py - c[1], pz - c[2]
    d = np.sqrt(dx * dx + dy * dy + dz * dz, dtype=np.float32) - np.float32(scene.sphere_radius)
    return np.ascontiguousarray(d.astype(np.float32).ravel())


def build_patch(scene, sc=None):
    """Create the world and patch of a scene in the GPU library."""
    if sc is None:
        import solidcloth_amd as sc
    world = sc.CollisionWorld()
    if scene.sdf_cells:
        bmin, bmax = sdf_box(scene)
        sdf = sc.SDF(sphere_sdf_values(scene), (scene.sdf_cells,) * 3, bmin, bmax)
        world.AddSDF(sdf)
    if scene.walls_min is not None:
        world.AddWalls(scene.walls_min, scene.walls_max)
    patch = sc.ClothPatch(world)
    m = patch.GetModel()
    m.SetUnit(scene.unit)
    m.SetFEM(True)
    m.SetShearStiffness(scene.young)
    m.SetPoisson(scene.poisson)
    m.SetBendStiffness(scene.bend)
    m.SetNumIterations(scene.iterations)
    m.SetFixedIterations(scene.fixed_iterations)
    m.SetCollisionFlags(scene.flags)
    m.SetThickness(scene.thickness)
    m.SetTolerance(scene.tolerance)
    m.SetFriction(scene.friction)
    patch.SetNumSteps(scene.substeps)
    patch.Init(scene.n, scene.n, scene.inc, scene.offset, scene.horizontal, scene.attached)
    return world, patch
