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
  mesh512   512x512 grid, side 2 m, height perturbed by three seeded
            sinusoids of amplitude 0.05 m (|k| <= 8 pi), dropped onto an
            icosphere triangle mesh (subdivision 6: 40,962 vertices, 81,920
            triangles) of radius 0.4 plus a ground plane; vertex, edge and
            triangle contacts against the mesh, continuous self collision
            (thresholds scaled to the grid spacing), thickness 2 mm, 10
            iterations, 4 substeps, 300 frames.
  fold24    24x24 grid folded over itself around a half cylinder, falling on
            the ground: the self-collision path at oracle size.
  meshsmall 24x24 grid over an icosphere of subdivision 2, the mesh
            contact path at a size the exhaustive oracle runs in seconds.
  batch512  512 independent cloths (batch_scenes): grid side uniform
            integer in [16, 128], Young log-uniform [1e3, 1e5], Poisson
            uniform [0.1, 0.45], bending log-uniform [1e-3, 1e-1], 1-4
            random pinned boundary vertices, each over its own sphere of
            radius uniform [0.2, 0.5] given as a 32^3-cell grid SDF;
            side 1 m at 0.6 m, 10 iterations, 4 substeps, 100 frames.
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
    # the sphere as an analytic collidable (scCreateSphere) instead of an SDF
    analytic_sphere: bool = False
    # walls box (None = no walls)
    walls_min: tuple = None
    walls_max: tuple = None
    # icosphere triangle-mesh obstacle (subdivision < 0 = none)
    mesh_subdiv: int = -1
    # capsule collidable: (centre, radius, half height, rotation (s, x, y, z))
    capsule: tuple = None
    mesh_center: tuple = (0.0, 0.0, 0.0)
    mesh_radius: float = 0.0
    # height perturbation: three sinusoids of this amplitude (0 = flat)
    perturb_amplitude: float = 0.0
    perturb_seed: int = 0
    # fold: the half x > 0 laid over the half x < 0 at this height gap (0 = none)
    fold_gap: float = 0.0
    # continuous self-collision thresholds (volume, bisection, vertex-triangle
    # distance); None = the reference's constants
    ccd: tuple = None
    # vertices pinned after Init (inverse mass 0); None = Init's own pins
    pins: tuple = None
    # velocity damping per substep (ClothModel::SetDragCoeff; 1 = none)
    drag: float = 1.0
    # self edge-edge displacement weights (scSetSelfEdgeWeights): 0 = the reference's
    self_edge_weights: int = 0
    # stretch model: FEM triangles (True) or links (False, ClothPBD.cpp:9-50);
    # dihedral bends (True) or bend links; link stretch stiffness (None = the
    # library default)
    fem: bool = True
    dihedral: bool = True
    stretch: float = None
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


def ccd_for_spacing(d):
    """The reference's absolute self-collision thresholds (volume 0.001,
    bisection 0.01, distance 0.1) were set for its demo cloth, whose grid
    spacing is 80/30 units (Demo/ClothDemo.cpp:141); scaled here to grid
    spacing d (volumes by the cube of the ratio)."""
    r = d / (80.0 / 30.0)
    return (0.001 * r ** 3, 0.01 * r ** 3, 0.1 * r)


# 512x512 perturbed cloth dropped onto the icosphere mesh (subdivision 6,
# radius 0.4), resting on a ground 5 cm below the sphere. Two choices the
# workload fixes where the reference gives no value: velocities are damped
# by 0.98 per substep (ClothModel::SetDragCoeff; the reference's default is
# no damping), so that the hanging cloth lands on the ground at about 2 m/s
# instead of 5; and self edge-edge contacts move each edge end by its own
# barycentric weight (scSetSelfEdgeWeights 1). As the reference writes the
# projection (ClothCollisionHandler.cpp:531-534) the weights of the two ends
# are exchanged, and the dense self contact of this drop then collapses
# patches of the cloth into knots whose contacts grow without bound (the
# oracle and the GPU agree on that collapse bit for bit).
MESH512 = Scene(
    name="mesh512", n=512, side=2.0, offset=(0.0, 0.5, 0.0), iterations=10, frames=300,
    flags=CF_VERTICES | CF_EDGES | CF_TRIANGLES | CF_WALLS | CF_SELF, thickness=0.002, tolerance=0.005,
    friction=0.3, walls_min=(-10.0, -0.45, -10.0), walls_max=(10.0, 10.0, 10.0),
    mesh_subdiv=6, mesh_radius=0.4, perturb_amplitude=0.05, perturb_seed=3,
    ccd=ccd_for_spacing(2.0 / 511), drag=0.98, self_edge_weights=1,
)

MESHSMALL = Scene(
    name="meshsmall", n=24, side=1.0, offset=(0.0, 0.3, 0.0), iterations=10, frames=30,
    flags=CF_VERTICES | CF_EDGES | CF_TRIANGLES | CF_WALLS, thickness=0.01, tolerance=0.01, friction=0.3,
    walls_min=(-10.0, -0.3, -10.0), walls_max=(10.0, 10.0, 10.0),
    mesh_subdiv=2, mesh_radius=0.25, perturb_amplitude=0.01, perturb_seed=1,
)


def batch_scenes(count=512, seed=4, n_range=(16, 128), frames=100):
    """The batched workload's cloths, drawn from one seeded generator."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        young = float(10.0 ** rng.uniform(3.0, 5.0))
        poisson = float(rng.uniform(0.1, 0.45))
        bend = float(10.0 ** rng.uniform(-3.0, -1.0))
        radius = float(rng.uniform(0.2, 0.5))
        border = [j * n + i for j in range(n) for i in range(n) if i in (0, n - 1) or j in (0, n - 1)]
        want = int(rng.integers(1, 5))
        chosen = []
        # no two pins in one triangle: the reference's FEM projection divides
        # by zero when a triangle has a single free vertex in the right place
        # (ClothPBD.cpp:52-108), so pins keep a grid distance of at least 2
        while len(chosen) < want:
            v = int(border[int(rng.integers(len(border)))])
            if all(max(abs(v // n - u // n), abs(v % n - u % n)) >= 2 for u in chosen):
                chosen.append(v)
        pins = tuple(sorted(chosen))
        out.append(Scene(
            name=f"batch{k}", n=n, side=1.0, offset=(0.0, 0.6, 0.0), iterations=10, frames=frames,
            young=young, poisson=poisson, bend=bend, flags=CF_VERTICES, thickness=0.002, tolerance=0.005,
            friction=0.3, sphere_center=(0.0, 0.0, 0.0), sphere_radius=radius, sdf_cells=32,
            sdf_half=radius + 0.1, pins=pins))
    return out


def batch_counts(scenes):
    """Total vertices, triangles and bends of a list of square-grid scenes."""
    v = sum(s.n * s.n for s in scenes)
    t = sum(2 * (s.n - 1) ** 2 for s in scenes)
    b = sum(3 * (s.n - 1) ** 2 - 2 * (s.n - 1) for s in scenes)
    return v, t, b


def icosphere(subdiv, radius=1.0, center=(0.0, 0.0, 0.0)):
    """Icosahedron subdivided `subdiv` times (every triangle into four, new
    vertices projected onto the sphere). Computed in float64 and rounded to
    FP32 once, so the arrays are identical on every host. Returns vertices
    (V, 3) float32 and triangle indices (T, 3) uint32, counter-clockwise seen
    from outside: V = 10 * 4^subdiv + 2, T = 20 * 4^subdiv."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.asarray(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.asarray(verts) * radius + np.asarray(center, dtype=np.float64)
    return v.astype(np.float32), np.asarray(faces, dtype=np.uint32)


def perturbed_positions(scene, pos):
    """Grid positions with the scene's height perturbation added to y:
    sum over three seeded sinusoids A sin(kx x + kz z + phi), |k| <= 8 pi,
    evaluated in float64 from the FP32 grid and rounded to FP32."""
    pos = np.asarray(pos, dtype=np.float32).copy()
    if scene.perturb_amplitude == 0.0:
        return pos
    rng = np.random.default_rng(scene.perturb_seed)
    ang = rng.uniform(0.0, 2.0 * np.pi, 3)
    mag = rng.uniform(0.0, 8.0 * np.pi, 3)
    phi = rng.uniform(0.0, 2.0 * np.pi, 3)
    x = pos[:, 0].astype(np.float64)
    z = pos[:, 2].astype(np.float64)
    h = np.zeros_like(x)
    for j in range(3):
        h += scene.perturb_amplitude * np.sin(mag[j] * np.cos(ang[j]) * x + mag[j] * np.sin(ang[j]) * z + phi[j])
    pos[:, 1] = (pos[:, 1].astype(np.float64) + h).astype(np.float32)
    return pos


FOLD24 = Scene(
    name="fold24", n=24, side=0.5, offset=(0.0, 0.0, 0.0), iterations=10, frames=12,
    flags=CF_VERTICES | CF_WALLS | CF_SELF, thickness=0.004, tolerance=0.004, friction=0.3,
    walls_min=(-10.0, -0.004, -10.0), walls_max=(10.0, 10.0, 10.0),
    fold_gap=0.02, ccd=ccd_for_spacing(0.5 / 23),
)


def folded_positions(scene, pos):
    """Grid positions with the half x > 0 folded over the half x < 0: it
    wraps around a half cylinder of diameter `fold_gap` along the line x = 0
    (arc length preserved) and continues flat at height `fold_gap`."""
    pos = np.asarray(pos, dtype=np.float32).copy()
    if scene.fold_gap == 0.0:
        return pos
    r = scene.fold_gap / 2.0
    s = pos[:, 0].astype(np.float64)
    y = pos[:, 1].astype(np.float64)
    right = s > 0
    on_arc = right & (s < np.pi * r)
    beyond = right & ~on_arc
    x = s.copy()
    x[on_arc] = r * np.sin(s[on_arc] / r)
    y[on_arc] += r * (1.0 - np.cos(s[on_arc] / r))
    x[beyond] = -(s[beyond] - np.pi * r)
    y[beyond] += 2.0 * r
    pos[:, 0] = x.astype(np.float32)
    pos[:, 1] = y.astype(np.float32)
    return pos


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
    if scene.analytic_sphere:
        world.AddSphere(scene.sphere_center, scene.sphere_radius)
    elif scene.sdf_cells:
        bmin, bmax = sdf_box(scene)
        sdf = sc.SDF(sphere_sdf_values(scene), (scene.sdf_cells,) * 3, bmin, bmax)
        world.AddSDF(sdf)
    if scene.mesh_subdiv >= 0:
        v, t = icosphere(scene.mesh_subdiv, scene.mesh_radius, scene.mesh_center)
        world.AddMesh(sc.Mesh(v, t))
    if scene.capsule is not None:
        world.AddCapsule(*scene.capsule)
    if scene.walls_min is not None:
        world.AddWalls(scene.walls_min, scene.walls_max)
    patch = sc.ClothPatch(world)
    m = patch.GetModel()
    m.SetUnit(scene.unit)
    m.SetFEM(scene.fem)
    m.SetDihedral(scene.dihedral)
    if scene.stretch is not None:
        m.SetStretchStiffness(scene.stretch)
    m.SetShearStiffness(scene.young)
    m.SetPoisson(scene.poisson)
    m.SetBendStiffness(scene.bend)
    m.SetNumIterations(scene.iterations)
    m.SetFixedIterations(scene.fixed_iterations)
    m.SetCollisionFlags(scene.flags)
    m.SetThickness(scene.thickness)
    m.SetTolerance(scene.tolerance)
    m.SetFriction(scene.friction)
    m.SetDragCoeff(scene.drag)
    m.SetSelfEdgeWeights(scene.self_edge_weights)
    patch.SetNumSteps(scene.substeps)
    patch.Init(scene.n, scene.n, scene.inc, scene.offset, scene.horizontal, scene.attached)
    if scene.perturb_amplitude != 0.0:
        m.set_positions(perturbed_positions(scene, m.positions()))
    if scene.fold_gap != 0.0:
        m.set_positions(folded_positions(scene, m.positions()))
    if scene.ccd is not None:
        m.SetSelfCCD(*scene.ccd)
    if scene.pins is not None:
        im = np.ones(scene.num_vertices, dtype=np.float32)
        im[list(scene.pins)] = 0.0
        m.set_inv_masses(im)
    return world, patch
