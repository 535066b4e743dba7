"""solidcloth_amd — Python mirror of SolidCloth's cloth API over the MI355X
C-ABI library libsolidcloth.so (include/solidcloth.h).

The classes keep the reference's names and method names (Physics/ClothPatch.h,
Physics/ClothModel.h, Physics/CollisionWorld.h, Demo/ClothAPI.h) so code and
tests read like the reference's. Every call goes through the C ABI into the
HIP implementation; there is no CPU fallback — loading fails loudly when the
library has not been built.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsolidcloth.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "solidcloth.h")

CF_VERTICES, CF_TRIANGLES, CF_SELF, CF_WALLS, CF_EDGES = 1, 2, 4, 8, 16
CT_INVALID, CT_WALLS, CT_SLAB, CT_SPHERE, CT_MESH, CT_BOX, CT_CAPSULE, CT_SDF = range(8)


class scVector3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class scVector2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class scAabb3(C.Structure):
    _fields_ = [("min", scVector3), ("max", scVector3)]


class scQuaternion(C.Structure):
    _fields_ = [("s", C.c_float), ("v", scVector3)]


class _Handle(C.Structure):
    _fields_ = [("ptr", C.c_void_p)]


scClothPatch = scCollisionWorld = scMesh = scModel = scSDF = _Handle

_lib = None


class SolidClothError(RuntimeError):
    pass


def v3(p):
    return scVector3(float(p[0]), float(p[1]), float(p[2]))


def _sig(L):
    H, vp, i, f, b = _Handle, C.c_void_p, C.c_int, C.c_float, C.c_bool
    P3 = C.POINTER(scVector3)
    sigs = {
        "scCreateCollisionWorld": (H, []), "scFreeCollisionWorld": (None, [H]),
        "scClearCollidables": (None, [H]), "scGetNumCollidables": (i, [H]),
        "scAddCollidable": (None, [H, vp]), "scGetCollidable": (vp, [H, i]),
        "scCreateWalls": (vp, [scAabb3]), "scCreateSphere": (vp, [scVector3, f]),
        "scCreateCapsule": (vp, [scVector3, f, f, scQuaternion]), "scCreateSDF": (vp, [H]),
        "scCreateMesh": (vp, [H]), "scGetCollidableType": (i, [vp]), "scGetMeshTree": (vp, [vp]),
        "scGetWallsInfo": (None, [vp, C.POINTER(scAabb3)]),
        "scGetSphereInfo": (None, [vp, C.POINTER(f), P3]),
        "scGetCapsuleInfo": (None, [vp, C.POINTER(f), C.POINTER(f), P3, C.POINTER(scQuaternion)]),
        "scCreateClothPatch": (H, [H]), "scFreeClothPatch": (None, [H]), "scGetClothModel": (H, [H]),
        "scGetClothMesh": (H, [H]), "scInitCloth": (None, [H, i, i, f, scVector3, b, b]),
        "scStepCloth": (None, [H, f]), "scUpdateClothMesh": (None, [H]),
        "scComputeStrainMap": (None, [H, P3, i]),
        "scGetNumParticles": (i, [H]), "scGetParticlePos": (None, [H, i, P3]),
        "scGetParticleInvMass": (f, [H, i]), "scGetNumTris": (i, [H]),
        "scGetTriangle": (None, [H, i, C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
        "scGetTriangleInfo": (None, [H, i, C.POINTER(f), C.POINTER(scVector2), C.POINTER(scVector2),
                                     C.POINTER(scVector2)]),
        "scGetNumQuads": (i, [H]), "scGetQuad": (None, [H, i] + [C.POINTER(i)] * 4),
        "scGetNumLinks": (i, [H]), "scGetLink": (None, [H, i, C.POINTER(i), C.POINTER(i)]),
        "scGetNumContacts": (i, [H]), "scGetContact": (None, [H, i, P3, P3]),
        "scGetNumTriContacts": (i, [H]), "scGetTriContact": (None, [H, i, P3, P3]),
        "scGetCollisionFlags": (i, [H]), "scSetCollisionFlags": (None, [H, i]),
        "scGetThickness": (f, [H]), "scSetThickness": (None, [H, f]),
        "scGetTolerance": (f, [H]), "scSetTolerance": (None, [H, f]),
        "scGetFriction": (f, [H]), "scSetFriction": (None, [H, f]),
        "scGetNumIterations": (i, [H]), "scSetNumIterations": (None, [H, i]),
        "scGetUseFEM": (b, [H]), "scSetUseFEM": (None, [H, b]), "scGetClothTree": (vp, [H]),
        "scInitClothFromMesh": (None, [H, H, scVector3, b]), "scSetMeshUVs": (i, [H, vp, i]),
        "scSetMeshColors": (i, [H, vp, i]), "scSetSolver": (None, [H, i]), "scGetSolver": (i, [H]),
        "scAddMouseSpring": (None, [H, i, i, i, scVector3, scVector3]),
        "scUpdateMouseSpring": (None, [H, scVector3]), "scRemoveMouseSpring": (None, [H]),
        "scGetMeshVertexCount": (i, [H]), "scGetMeshIndexCount": (i, [H]),
        "scGetMeshNormalBuffer": (P3, [H]), "scGetMeshVertexBuffer": (P3, [H]),
        "scGetMeshIndexBuffer": (C.POINTER(C.c_uint), [H]),
        "scSetUnit": (None, [H, f]), "scGetUnit": (f, [H]), "scSetShearStiffness": (None, [H, f]),
        "scSetStretchStiffness": (None, [H, f]), "scSetBendStiffness": (None, [H, f]),
        "scSetPoisson": (None, [H, f]), "scSetDihedral": (None, [H, b]), "scSetDragCoeff": (None, [H, f]),
        "scGetShearStiffness": (f, [H]), "scGetStretchStiffness": (f, [H]), "scGetBendStiffness": (f, [H]),
        "scGetPoisson": (f, [H]), "scGetDihedral": (b, [H]), "scGetDragCoeff": (f, [H]),
        "scGetSelfParams": (None, [H, C.POINTER(f), C.POINTER(f)]),
        "scSetNumSteps": (None, [H, i]), "scGetNumSteps": (i, [H]),
        "scSetFixedIterations": (None, [H, b]), "scSetInitIncline": (None, [H, f]),
        "scCreateGridSDF": (H, [vp, i, i, i, scAabb3]),
        "scCreateSDFFromMesh": (H, [H, i, i, i, scAabb3]), "scFreeSDF": (None, [H]),
        "scLoadSDFFile": (H, [C.c_char_p]), "scWriteSDFFile": (i, [H, C.c_char_p]),
        "scGetSDFGrid": (i, [H, vp, vp, vp, i]),
        "scLoadMeshFromObj": (H, [C.c_char_p, b]), "scGetMeshUVBuffer": (C.POINTER(scVector2), [H]),
        "scCreateMeshFromArrays": (H, [vp, i, vp, i]), "scFreeMesh": (None, [H]),
        "scGetParticlePositions": (i, [H, vp, i]), "scSetParticlePositions": (i, [H, vp, i]),
        "scGetParticleVelocities": (i, [H, vp, i]), "scSetParticleVelocities": (i, [H, vp, i]),
        "scGetNumBends": (i, [H]), "scGetBends": (i, [H, vp, i]), "scGetLinks": (i, [H, vp, vp, i]),
        "scGetConstraintOrder": (i, [H, i, vp, i]), "scGetColouring": (i, [H, i, vp, i]),
        "scGetSchedule": (i, [H, i, vp, vp, vp, i]), "scSetSweep": (None, [H, i, i]),
        "scGetSolverStats": (None, [H, vp]), "scEnableKernelTiming": (None, [H, b]),
        "scGetKernelTiming": (i, [H, i, C.POINTER(i), C.POINTER(C.c_double)]),
        "scSetUseGraph": (None, [H, b]), "scBandwidthProbe": (C.c_double, [C.c_longlong, i]),
        "scGetContactFeatures": (i, [H, vp, i]), "scGetEdgeContactData": (i, [H, vp, vp, i]),
        "scGetTriContactData": (i, [H, vp, vp, i]),
        "scLbvhSelfCheck": (i, [i, C.c_uint, i, vp]),
        "scGetSelfContacts": (i, [H, i, vp, vp, vp, i]), "scSetSelfParams": (None, [H, f, f]),
        "scSetSelfCCD": (None, [H, f, f, f]), "scSetSelfCapacity": (None, [H, i]),
        "scSetSelfEdgeWeights": (None, [H, i]), "scGetTileSchedule": (i, [H, vp, vp, vp, vp, vp, vp]), "scGetSelfCandidates": (C.c_longlong, [H, i, vp, C.c_longlong]),
        "scGetSelfStats": (i, [H, vp, i]),
        "scSetParticleInvMasses": (i, [H, vp, i]),
        "scCreateWarmStart": (H, [C.c_uint]), "scFreeWarmStart": (None, [H]),
        "scWarmStartGetParams": (i, [H, vp, i]), "scWarmStartSetParams": (i, [H, vp, i]),
        "scWarmStartSetLearningRate": (None, [H, f]), "scWarmStartGetSteps": (i, [H]),
        "scWarmStartForward": (i, [H, vp, i, vp]), "scWarmStartForwardNoAlloc": (i, [H, vp, i, vp, vp]), "scWarmStartGradient": (f, [H, vp, vp, vp, i, vp]),
        "scWarmStartTrain": (f, [H, vp, vp, vp, i, i]), "scSetWarmStart": (None, [H, H, i, i]),
        "scSetWarmStartFallback": (None, [H, f]), "scGetResidual": (C.c_double, [H]),
        "scGetResidualTerms": (None, [H, vp]),
        "scGetWarmStartStats": (None, [H, vp]), "scGetWarmStartData": (i, [H, i, vp, i]),
        "scSaveState": (None, [H]), "scRestoreState": (None, [H]), "scGetLastStepTime": (C.c_double, [H]),
        "scCreateClothBatch": (H, []), "scFreeClothBatch": (None, [H]), "scBatchAddPatch": (None, [H, H]),
        "scBatchGetNumPatches": (i, [H]), "scStepClothBatch": (None, [H, f]),
        "scBatchSetLaunch": (None, [H, i, i]),
        "scBatchGetLaunch": (None, [H, C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
        "scBatchGetTuning": (i, [H, vp, vp, i]), "scBatchGetStepTime": (C.c_double, [H, C.POINTER(i)]),
        "scBatchSetUseGraph": (None, [H, b]), "scBatchRecordWorkgroupTimes": (None, [H, b]),
        "scBatchGetWorkgroupTimes": (i, [H, vp, vp, vp, i]),
        "scGetLastError": (C.c_char_p, []),
        "scGetErrorCount": (i, []), "scGetBuildInfo": (C.c_char_p, []),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return sigs


def lib():
    """Load libsolidcloth.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python __graft_entry__.py` "
                              "or `make -C solidcloth_amd`")
        L = C.CDLL(LIB_PATH)
        _sig(L)
        _lib = L
    return _lib


def exported_names():
    """Names bound by this module (every function of include/solidcloth.h)."""
    return sorted(_sig(lib()).keys())


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _checked(name, *args):
    """Call a C-ABI function; raise SolidClothError if it recorded an error."""
    L = lib()
    before = L.scGetErrorCount()
    r = getattr(L, name)(*args)
    if L.scGetErrorCount() != before:
        raise SolidClothError(L.scGetLastError().decode())
    return r


# ------------------------------------------------------------------ objects

class SDF:
    """Node-sampled grid SDF (Geometry/SDF.h); values (nx+1)(ny+1)(nz+1), x fastest."""

    def __init__(self, values, n, bmin, bmax):
        v = np.ascontiguousarray(values, dtype=np.float32)
        nx, ny, nz = n
        if v.size != (nx + 1) * (ny + 1) * (nz + 1):
            raise ValueError("SDF value count does not match the grid")
        box = scAabb3(v3(bmin), v3(bmax))
        self.h = _checked("scCreateGridSDF", _ptr(v), nx, ny, nz, box)
        self.n, self.bmin, self.bmax = tuple(n), tuple(bmin), tuple(bmax)

    @classmethod
    def from_mesh(cls, mesh, n, bmin, bmax):
        """SDF::CreateFromMesh on the GPU: the unsigned distance of every node
        of an n = (nx, ny, nz)-cell grid over [bmin, bmax] to the mesh."""
        self = cls.__new__(cls)
        nx, ny, nz = n
        self.h = _checked("scCreateSDFFromMesh", mesh.h, nx, ny, nz, scAabb3(v3(bmin), v3(bmax)))
        self.n, self.bmin, self.bmax = tuple(n), tuple(bmin), tuple(bmax)
        return self

    @classmethod
    def from_file(cls, path):
        """SDF read from the reference's binary file (SDF::LoadFromFile)."""
        self = cls.__new__(cls)
        self.h = _checked("scLoadSDFFile", os.fsencode(path))
        cells, box = (C.c_int * 3)(), scAabb3()
        _checked("scGetSDFGrid", self.h, C.cast(cells, C.c_void_p), C.cast(C.pointer(box), C.c_void_p), None, 0)
        self.n = tuple(cells)
        self.bmin = (box.min.x, box.min.y, box.min.z)
        self.bmax = (box.max.x, box.max.y, box.max.z)
        return self

    def write(self, path):
        """Writes the SDF in the reference's binary format (SDF::WriteToFile)."""
        _checked("scWriteSDFFile", self.h, os.fsencode(path))

    def values(self):
        """Node values, shape (nz+1, ny+1, nx+1)."""
        nx, ny, nz = self.n
        out = np.empty((nz + 1, ny + 1, nx + 1), dtype=np.float32)
        _checked("scGetSDFGrid", self.h, None, None, _ptr(out), out.size)
        return out


class Mesh:
    """Indexed triangle mesh (Geometry/Mesh.h) built from vertex and index
    arrays; normals and unique edges are computed on creation."""

    def __init__(self, vertices, indices):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, dtype=np.uint32).ravel()
        self.h = _checked("scCreateMeshFromArrays", _ptr(v), len(v), _ptr(i), i.size)
        self.vertices, self.indices = v, i

    @classmethod
    def from_obj(cls, path, naive=False):
        """Mesh read from a Wavefront OBJ file (Assets LoadMeshFromObj); its
        uvs, when the file has them and naive is False, are in self.uvs."""
        self = cls.__new__(cls)
        self.h = _checked("scLoadMeshFromObj", os.fsencode(path), bool(naive))
        L = lib()
        nv, ni = L.scGetMeshVertexCount(self.h), L.scGetMeshIndexCount(self.h)
        vb, ib, ub = L.scGetMeshVertexBuffer(self.h), L.scGetMeshIndexBuffer(self.h), L.scGetMeshUVBuffer(self.h)
        self.vertices = (np.ctypeslib.as_array(C.cast(vb, C.POINTER(C.c_float)), shape=(nv, 3)).copy()
                         if nv else np.zeros((0, 3), np.float32))
        self.indices = np.ctypeslib.as_array(ib, shape=(ni,)).copy() if ni else np.zeros(0, np.uint32)
        self.uvs = (np.ctypeslib.as_array(C.cast(ub, C.POINTER(C.c_float)), shape=(nv, 2)).copy()
                    if ub else None)
        return self

    def set_uvs(self, uvs):
        """Material coordinates per vertex (Mesh::uvs), used by a mesh-based patch."""
        u = np.ascontiguousarray(uvs, dtype=np.float32).reshape(-1, 2)
        _checked("scSetMeshUVs", self.h, _ptr(u), len(u))

    def set_colors(self, colors):
        """Colours per vertex (Mesh::colors); a mesh-based patch with attached
        pins the vertices whose red is non-zero."""
        c = np.ascontiguousarray(colors, dtype=np.float32).reshape(-1, 3)
        _checked("scSetMeshColors", self.h, _ptr(c), len(c))

    def __del__(self):
        try:
            if self.h.ptr:
                lib().scFreeMesh(self.h)
        except Exception:
            pass


class CollisionWorld:
    """Physics/CollisionWorld.h:10-29 over scCreateCollisionWorld."""

    def __init__(self):
        self.h = _checked("scCreateCollisionWorld")
        self._keep = []

    def AddCollidable(self, handle):
        _checked("scAddCollidable", self.h, handle)

    def ClearCollidables(self):
        _checked("scClearCollidables", self.h)

    def GetNumCollidables(self):
        return _checked("scGetNumCollidables", self.h)

    def AddWalls(self, bmin, bmax):
        h = _checked("scCreateWalls", scAabb3(v3(bmin), v3(bmax)))
        self.AddCollidable(h)
        return h

    def AddSphere(self, center, radius):
        h = _checked("scCreateSphere", v3(center), float(radius))
        self.AddCollidable(h)
        return h

    def AddCapsule(self, center, radius, half_height, rot=(1.0, 0.0, 0.0, 0.0)):
        """Capsule collidable (scCreateCapsule): axis along the rotated y axis,
        rot = (s, x, y, z)."""
        q = scQuaternion(float(rot[0]), v3(rot[1:4]))
        h = _checked("scCreateCapsule", v3(center), float(radius), float(half_height), q)
        self.AddCollidable(h)
        return h

    def AddSDF(self, sdf):
        self._keep.append(sdf)
        h = _checked("scCreateSDF", sdf.h)
        self.AddCollidable(h)
        return h

    def AddMesh(self, mesh):
        """A static triangle-mesh obstacle (CollisionMesh, Physics/Common.h:113-138)."""
        self._keep.append(mesh)
        h = _checked("scCreateMesh", mesh.h)
        self.AddCollidable(h)
        return h


class ClothModel:
    """Physics/ClothModel.h:42-251 — parameters and particle state."""

    def __init__(self, h, owner):
        self.h = h
        self._owner = owner  # the patch owns the model; keep it alive

    def _set(self, name, v):
        _checked(name, self.h, v)

    SetThickness = lambda self, v: self._set("scSetThickness", float(v))
    SetTolerance = lambda self, v: self._set("scSetTolerance", float(v))
    SetFriction = lambda self, v: self._set("scSetFriction", float(v))
    SetNumIterations = lambda self, v: self._set("scSetNumIterations", int(v))
    SetFEM = lambda self, v: self._set("scSetUseFEM", bool(v))
    SetCollisionFlags = lambda self, v: self._set("scSetCollisionFlags", int(v))
    SetUnit = lambda self, v: self._set("scSetUnit", float(v))
    SetShearStiffness = lambda self, v: self._set("scSetShearStiffness", float(v))
    SetStretchStiffness = lambda self, v: self._set("scSetStretchStiffness", float(v))
    SetBendStiffness = lambda self, v: self._set("scSetBendStiffness", float(v))
    SetPoisson = lambda self, v: self._set("scSetPoisson", float(v))
    SetDihedral = lambda self, v: self._set("scSetDihedral", bool(v))
    SetDragCoeff = lambda self, v: self._set("scSetDragCoeff", float(v))
    SetFixedIterations = lambda self, v: self._set("scSetFixedIterations", bool(v))
    SetUseGraph = lambda self, v: self._set("scSetUseGraph", bool(v))

    def GetThickness(self):
        return _checked("scGetThickness", self.h)

    GetShearStiffness = lambda self: _checked("scGetShearStiffness", self.h)
    GetStretchStiffness = lambda self: _checked("scGetStretchStiffness", self.h)
    GetBendStiffness = lambda self: _checked("scGetBendStiffness", self.h)
    GetPoisson = lambda self: _checked("scGetPoisson", self.h)
    GetDihedral = lambda self: _checked("scGetDihedral", self.h)
    GetDragCoeff = lambda self: _checked("scGetDragCoeff", self.h)

    def GetSelfParams(self):
        shrink, threshold = C.c_float(), C.c_float()
        _checked("scGetSelfParams", self.h, C.byref(shrink), C.byref(threshold))
        return shrink.value, threshold.value

    def GetNumIterations(self):
        return _checked("scGetNumIterations", self.h)

    def GetNumParticles(self):
        return _checked("scGetNumParticles", self.h)

    def GetNumTris(self):
        return _checked("scGetNumTris", self.h)

    def GetNumLinks(self):
        return _checked("scGetNumLinks", self.h)

    def GetNumQuads(self):
        return _checked("scGetNumQuads", self.h)

    def GetNumBends(self):
        return _checked("scGetNumBends", self.h)

    def GetParticlePos(self, i):
        p = scVector3()
        _checked("scGetParticlePos", self.h, i, C.byref(p))
        return (p.x, p.y, p.z)

    def GetParticleInvMass(self, i):
        return _checked("scGetParticleInvMass", self.h, i)

    def positions(self):
        n = self.GetNumParticles()
        out = np.zeros((n, 3), dtype=np.float32)
        _checked("scGetParticlePositions", self.h, _ptr(out), n)
        return out

    def set_positions(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        _checked("scSetParticlePositions", self.h, _ptr(a), a.shape[0])

    def velocities(self):
        n = self.GetNumParticles()
        out = np.zeros((n, 3), dtype=np.float32)
        _checked("scGetParticleVelocities", self.h, _ptr(out), n)
        return out

    def set_velocities(self, v):
        a = np.ascontiguousarray(v, dtype=np.float32)
        _checked("scSetParticleVelocities", self.h, _ptr(a), a.shape[0])

    def inv_masses(self):
        return np.array([self.GetParticleInvMass(i) for i in range(self.GetNumParticles())], dtype=np.float32)

    # learned warm start
    WARM_MODES = {"off": 0, "train": 1, "solve": 2, "track": 3}

    def SetWarmStart(self, ws, mode="solve", iterations=4):
        m = self.WARM_MODES[mode] if isinstance(mode, str) else int(mode)
        self._warm = ws  # keep the network alive while attached
        _checked("scSetWarmStart", self.h, ws.h if ws is not None else _Handle(None), m, int(iterations))

    def SetWarmStartFallback(self, threshold=-1.0):
        _checked("scSetWarmStartFallback", self.h, float(threshold))

    def residual(self):
        return _checked("scGetResidual", self.h)

    def residual_terms(self):
        """RMS of (warp/weft, shear, bend error)."""
        out = np.zeros(3)
        _checked("scGetResidualTerms", self.h, _ptr(out))
        return out

    def warm_stats(self):
        out = np.zeros(7)
        _checked("scGetWarmStartStats", self.h, _ptr(out))
        return dict(residual=out[0], threshold=out[1], fallbacks=int(out[2]), plain_residual=out[3],
                    steps=int(out[4]), last_plain=bool(out[5]), warm_residual=out[6])

    def warm_data(self, which):
        """'features' (n,12), 'xhat' (n,3), 'corr' (n,3), 'targets' (n,4), 'scale' (ell, free count)."""
        k = {"features": 0, "xhat": 1, "corr": 2, "targets": 3, "scale": 4}[which]
        if k == 4:
            out = np.zeros(2, dtype=np.float32)
            _checked("scGetWarmStartData", self.h, 4, _ptr(out), 2)
            return float(out[0]), int(out[1])
        n = self.GetNumParticles()
        w = {0: 12, 1: 3, 2: 3, 3: 4}[k]
        out = np.zeros((n, w), dtype=np.float32)
        _checked("scGetWarmStartData", self.h, k, _ptr(out), n)
        return out

    def last_step_ms(self):
        return _checked("scGetLastStepTime", self.h)

    def save_state(self):
        _checked("scSaveState", self.h)

    def restore_state(self):
        _checked("scRestoreState", self.h)

    def set_inv_masses(self, m):
        a = np.ascontiguousarray(m, dtype=np.float32)
        _checked("scSetParticleInvMasses", self.h, _ptr(a), a.shape[0])

    def triangles(self):
        n = self.GetNumTris()
        out = np.zeros((n, 3), dtype=np.int32)
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        for k in range(n):
            _checked("scGetTriangle", self.h, k, C.byref(a), C.byref(b), C.byref(c))
            out[k] = (a.value, b.value, c.value)
        return out

    def bends(self):
        n = self.GetNumBends()
        out = np.zeros((n, 4), dtype=np.int32)
        _checked("scGetBends", self.h, _ptr(out), n)
        return out

    _KINDS = {0: 0, "tri": 0, 1: 1, "bend": 1, 2: 2, "edge_contact": 2, 3: 3, "tri_contact": 3, 4: 4, "link": 4}

    def constraint_order(self, kind):
        """kind 'tri' / 'bend' / 'edge_contact' / 'tri_contact' / 'link' (or 0-4)."""
        k = self._KINDS[kind]
        n = _checked("scGetConstraintOrder", self.h, k, None, 0)
        out = np.zeros(n, dtype=np.int32)
        _checked("scGetConstraintOrder", self.h, k, _ptr(out), n)
        return out

    def colour_offsets(self, kind):
        """Start offsets of the launch phases in constraint_order(kind)."""
        k = self._KINDS[kind]
        nc = _checked("scGetColouring", self.h, k, None, 0)
        out = np.zeros(nc + 1, dtype=np.int32)
        _checked("scGetColouring", self.h, k, _ptr(out), nc)
        return out

    def schedule(self, kind):
        """Per position of constraint_order(kind): (phase, slab, colour)."""
        k = self._KINDS[kind]
        n = _checked("scGetSchedule", self.h, k, None, None, None, 0)
        ph, sl, co = (np.zeros(n, dtype=np.int32) for _ in range(3))
        _checked("scGetSchedule", self.h, k, _ptr(ph), _ptr(sl), _ptr(co), n)
        return ph, sl, co

    def SetSweep(self, mode, slab_size=0):
        """mode 0: one launch per colour; mode 1: slab phases staged in LDS
        (strips of slab_size constraints); mode 2: strip interiors of about
        slab_size constraints, then the seams between them; mode 3: FEM
        triangles on tiles with halos, bends on slab phases; mode 4 (the
        default): mode 3's schedules with every launch's data requested
        before its first colour; in modes 3 and 4 slab_size sets the
        interiors of the bend and link slabs. slab_size 0: the library's
        choice."""
        _checked("scSetSweep", self.h, int(mode), int(slab_size))

    def GetNumContacts(self):
        return _checked("scGetNumContacts", self.h)

    def contacts(self):
        n = _checked("scGetNumContacts", self.h)
        pts = np.zeros((n, 3), dtype=np.float32)
        nrm = np.zeros((n, 3), dtype=np.float32)
        p, q = scVector3(), scVector3()
        for k in range(n):
            _checked("scGetContact", self.h, k, C.byref(p), C.byref(q))
            pts[k] = (p.x, p.y, p.z)
            nrm[k] = (q.x, q.y, q.z)
        return pts, nrm

    def contact_features(self):
        """Obstacle triangle of each vertex contact (-1 if not a mesh contact)."""
        n = _checked("scGetContactFeatures", self.h, None, 0)
        out = np.zeros(n, dtype=np.int32)
        _checked("scGetContactFeatures", self.h, _ptr(out), n)
        return out

    def _feature_contacts(self, name, width):
        n = _checked(name, self.h, None, None, 0)
        ids = np.zeros((n, 3), dtype=np.int32)
        data = np.zeros((n, width), dtype=np.float32)
        _checked(name, self.h, _ptr(ids), _ptr(data), n)
        return ids, data

    def edge_contacts(self):
        """ids (cloth edge, obstacle edge, collidable), data (point, normal, w1,
        w2, lambda, lambdaF)."""
        return self._feature_contacts("scGetEdgeContactData", 10)

    def tri_contacts(self):
        """ids (cloth triangle, obstacle vertex, collidable), data (point,
        normal, w1, w2, w3, lambda, lambdaF)."""
        return self._feature_contacts("scGetTriContactData", 11)

    def stats(self):
        out = np.zeros(6, dtype=np.float32)
        _checked("scGetSolverStats", self.h, _ptr(out))
        return dict(iterations=int(out[0]), bendError=float(out[1]), contactError=float(out[2]),
                    selfVTError=float(out[3]), selfEEError=float(out[4]), linkError=float(out[5]))

    def links(self):
        """(L, 2) vertex pairs and (L, 2) float32 (rest length, stiffness), index order."""
        n = _checked("scGetLinks", self.h, None, None, 0)
        ids = np.zeros((n, 2), dtype=np.int32)
        data = np.zeros((n, 2), dtype=np.float32)
        _checked("scGetLinks", self.h, _ptr(ids), _ptr(data), n)
        return ids, data

    def self_contacts(self, kind):
        """kind 'vt' / 'ee' -> keys (n,), ids (n, 4), data (n, 8): w1 w2 w3,
        normal, side, lambda; in the order they are solved."""
        k = 0 if kind == "vt" else 1
        n = _checked("scGetSelfContacts", self.h, k, None, None, None, 0)
        keys = np.zeros(n, dtype=np.uint64)
        ids = np.zeros((n, 4), dtype=np.int32)
        data = np.zeros((n, 8), dtype=np.float32)
        _checked("scGetSelfContacts", self.h, k, _ptr(keys), _ptr(ids), _ptr(data), n)
        return keys, ids, data

    _SOLVERS = {"jacobi": 0, "gauss_seidel": 1, "jacobi_cr": 2, "implicit": 3, "explicit": 4}

    def SetSolver(self, solver):
        """Physics/ClothModel.h:138: "gauss_seidel" (default), "jacobi", or a
        conjugate-residual type ("jacobi_cr", "implicit", "explicit"), or the
        SolverType value."""
        _checked("scSetSolver", self.h, self._SOLVERS.get(solver, solver) if isinstance(solver, str) else int(solver))

    def GetSolver(self):
        return _checked("scGetSolver", self.h)

    def AddMouseSpring(self, i1, i2, i3, coords, point):
        """Physics/ClothModel.h:150: pulls the barycentric point (coords) of
        particles i1, i2, i3 towards `point` in every following substep."""
        _checked("scAddMouseSpring", self.h, int(i1), int(i2), int(i3), v3(coords), v3(point))

    def UpdateMouseSpring(self, point):
        _checked("scUpdateMouseSpring", self.h, v3(point))

    def RemoveMouseSpring(self):
        _checked("scRemoveMouseSpring", self.h)

    def SetSelfParams(self, shrink, threshold):
        _checked("scSetSelfParams", self.h, float(shrink), float(threshold))

    def SetSelfCCD(self, volume, bisect, vt_distance):
        _checked("scSetSelfCCD", self.h, float(volume), float(bisect), float(vt_distance))

    def self_candidates(self, kind):
        """The last substep's broad-phase candidate pairs, 'vt' (triangle,
        vertex) or 'ee' (edge1, edge2), as a sorted (n, 2) int32 array."""
        k = 0 if kind == "vt" else 1
        n = _checked("scGetSelfCandidates", self.h, k, None, 0)
        out = np.zeros((n, 2), dtype=np.int32)
        if n:
            _checked("scGetSelfCandidates", self.h, k, _ptr(out), n)
        return out

    def tile_schedule(self):
        """The triangles' tile schedule (sweep mode 3) as a dict: vert_begin,
        verts, owned, col_offsets (tiles x colours + 1), instances, colours;
        None when the triangles are not on tiles."""
        counts = np.zeros(4, dtype=np.int32)
        nt = _checked("scGetTileSchedule", self.h, _ptr(counts), None, None, None, None, None)
        if nt == 0:
            return None
        ntiles, nverts, nc, ninst = (int(x) for x in counts)
        vb = np.zeros(ntiles + 1, dtype=np.int32)
        vs = np.zeros(nverts, dtype=np.int32)
        ow = np.zeros(nverts, dtype=np.int32)
        co = np.zeros(ntiles * nc + 1, dtype=np.int32)
        ins = np.zeros(ninst, dtype=np.int32)
        _checked("scGetTileSchedule", self.h, None, _ptr(vb), _ptr(vs), _ptr(ow), _ptr(co), _ptr(ins))
        return dict(vert_begin=vb, verts=vs, owned=ow, col_offsets=co, instances=ins, colours=nc)

    def SetSelfEdgeWeights(self, mode):
        """0: the reference's edge-edge displacement weights; 1: each end's own."""
        _checked("scSetSelfEdgeWeights", self.h, int(mode))

    def SetSelfCapacity(self, capacity):
        _checked("scSetSelfCapacity", self.h, int(capacity))

    def self_stats(self):
        """Contact capacity in use, steps rerun after an overflow, most
        vertex-triangle / edge-edge contacts in one substep of the last step,
        candidate-pair pool capacity (both kinds), colour groups and distinct
        contact vertices of the last substep."""
        out = np.zeros(9, dtype=np.int32)
        _checked("scGetSelfStats", self.h, _ptr(out), 9)
        return dict(capacity=int(out[0]), reruns=int(out[1]), peak_vt=int(out[2]), peak_ee=int(out[3]),
                    pool=int(out[4]), colours_vt=int(out[5]), colours_ee=int(out[6]),
                    vertices_vt=int(out[7]), vertices_ee=int(out[8]))

    def enable_kernel_timing(self, on=True):
        _checked("scEnableKernelTiming", self.h, bool(on))

    def kernel_timing(self, which):
        n, ms = C.c_int(), C.c_double()
        _checked("scGetKernelTiming", self.h, int(which), C.byref(n), C.byref(ms))
        return n.value, ms.value


class ClothPatch:
    """Physics/ClothPatch.h:24-72 over the scClothPatch handle."""

    def __init__(self, world):
        self.world = world
        self.h = _checked("scCreateClothPatch", world.h)
        self._model = ClothModel(_checked("scGetClothModel", self.h), self)

    def __del__(self):
        try:
            lib().scFreeClothPatch(self.h)
        except Exception:
            pass

    def InitFromMesh(self, mesh, offset=(0, 0, 0), attached=False):
        """Physics/ClothPatch.cpp:43-133: a patch from a Mesh (welded vertices,
        the mesh's uvs and colours)."""
        self._mesh = mesh
        _checked("scInitClothFromMesh", self.h, mesh.h, v3(offset), bool(attached))

    def Init(self, divX, divY, inc, offset=(0, 0, 0), horizontal=True, attached=False):
        _checked("scInitCloth", self.h, int(divX), int(divY), float(inc), v3(offset), bool(horizontal),
                 bool(attached))

    def SetInitIncline(self, incline):
        _checked("scSetInitIncline", self.h, float(incline))

    def Step(self, dt):
        _checked("scStepCloth", self.h, float(dt))

    def SetNumSteps(self, n):
        _checked("scSetNumSteps", self.h, int(n))

    def GetNumSteps(self):
        return _checked("scGetNumSteps", self.h)

    def GetModel(self):
        return self._model

    def UpdateMesh(self):
        _checked("scUpdateClothMesh", self.h)

    def mesh_arrays(self):
        """Copies of the render mesh (scGetClothMesh): vertices (V, 3),
        normals (V, 3) and triangle indices (T, 3); the normals are those of
        the last UpdateMesh."""
        L = lib()
        mesh = L.scGetClothMesh(self.h)
        nv = L.scGetMeshVertexCount(mesh)
        ni = L.scGetMeshIndexCount(mesh)
        vb, nb, ib = L.scGetMeshVertexBuffer(mesh), L.scGetMeshNormalBuffer(mesh), L.scGetMeshIndexBuffer(mesh)
        if nv <= 0 or not vb:
            raise SolidClothError("the patch has no render mesh")
        verts = np.ctypeslib.as_array(C.cast(vb, C.POINTER(C.c_float)), shape=(nv, 3)).copy()
        normals = (np.ctypeslib.as_array(C.cast(nb, C.POINTER(C.c_float)), shape=(nv, 3)).copy() if nb
                   else np.zeros((nv, 3), dtype=np.float32))
        idx = np.ctypeslib.as_array(ib, shape=(ni,)).copy().reshape(-1, 3)
        return verts, normals, idx


class ClothBatch:
    """Many patches stepped together, one launch per substep (scClothBatch).

    Keeps references to its patches so they outlive the batch."""

    def __init__(self, patches=()):
        self.h = _checked("scCreateClothBatch")
        self.patches = []
        for p in patches:
            self.Add(p)

    def __del__(self):
        try:
            lib().scFreeClothBatch(self.h)
        except Exception:
            pass

    def Add(self, patch):
        _checked("scBatchAddPatch", self.h, patch.h)
        self.patches.append(patch)

    def GetNumPatches(self):
        return _checked("scBatchGetNumPatches", self.h)

    def Step(self, dt):
        _checked("scStepClothBatch", self.h, float(dt))

    def SetLaunch(self, block_size=0, stage=-1):
        _checked("scBatchSetLaunch", self.h, int(block_size), int(stage))

    def GetLaunch(self):
        b, s, t = C.c_int(), C.c_int(), C.c_int()
        _checked("scBatchGetLaunch", self.h, C.byref(b), C.byref(s), C.byref(t))
        return b.value, s.value, bool(t.value)

    def SetUseGraph(self, enable):
        _checked("scBatchSetUseGraph", self.h, bool(enable))

    def RecordWorkgroupTimes(self, enable):
        _checked("scBatchRecordWorkgroupTimes", self.h, bool(enable))

    def workgroup_times(self):
        """Per patch, for the last recorded substep launch: start and end in
        microseconds from its earliest start, and the unit the cloth's
        workgroup ran on as (XCC, shader engine, array, compute unit); None
        before a recorded step."""
        n = _checked("scBatchGetWorkgroupTimes", self.h, None, None, None, 0)
        if n == 0:
            return None
        st, en = np.zeros(n), np.zeros(n)
        raw = np.zeros(n, dtype=np.int64)
        _checked("scBatchGetWorkgroupTimes", self.h, _ptr(st), _ptr(en), _ptr(raw), n)
        hw = raw & 0xffffffff
        unit = np.stack([raw >> 32, (hw >> 13) & 7, (hw >> 12) & 1, (hw >> 8) & 15], axis=1)
        return dict(start_us=st, end_us=en, unit=unit)

    def last_step_time(self):
        """(device ms of the last step's substep launches, launch count)."""
        n = C.c_int()
        ms = _checked("scBatchGetStepTime", self.h, C.byref(n))
        return ms, n.value

    def tuning(self):
        n = _checked("scBatchGetTuning", self.h, None, None, 0)
        cfg = np.zeros((n, 2), dtype=np.int32)
        ms = np.zeros(n, dtype=np.float64)
        _checked("scBatchGetTuning", self.h, _ptr(cfg), _ptr(ms), n)
        return [(int(c[0]), int(c[1]), float(t)) for c, t in zip(cfg, ms)]


class WarmStart:
    """The learned warm-start network (scWarmStart): 12-64-64-3 MLP."""

    NUM_PARAMS = 5187

    def __init__(self, seed=0):
        self.h = _checked("scCreateWarmStart", int(seed))

    def __del__(self):
        try:
            lib().scFreeWarmStart(self.h)
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.NUM_PARAMS, dtype=np.float32)
        _checked("scWarmStartGetParams", self.h, _ptr(out), out.size)
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32).ravel()
        _checked("scWarmStartSetParams", self.h, _ptr(a), a.size)

    def SetLearningRate(self, lr):
        _checked("scWarmStartSetLearningRate", self.h, float(lr))

    def steps(self):
        return _checked("scWarmStartGetSteps", self.h)

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros((X.shape[0], 3), dtype=np.float32)
        _checked("scWarmStartForward", self.h, _ptr(X), X.shape[0], _ptr(out))
        return out

    def forward_noalloc(self, features_ptr, n, out_ptr, stream=None):
        """scWarmStartForwardNoAlloc: device pointers in and out (n x 12 and
        n x 3 FP32), staged in the library's static scratch; enqueued on
        `stream` (a hipStream_t as an int, None = the null stream)."""
        return _checked("scWarmStartForwardNoAlloc", self.h, C.c_void_p(int(features_ptr)), int(n),
                        C.c_void_p(int(out_ptr)), C.c_void_p(int(stream) if stream else None))

    def _tw(self, X, T, w):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        return X, T, w, (None if w is None else _ptr(w))

    def gradient(self, X, T, w=None):
        X, T, w, wp = self._tw(X, T, w)
        g = np.zeros(self.NUM_PARAMS, dtype=np.float32)
        loss = _checked("scWarmStartGradient", self.h, _ptr(X), _ptr(T), wp, X.shape[0], _ptr(g))
        return loss, g

    def train(self, X, T, w=None, steps=1):
        X, T, w, wp = self._tw(X, T, w)
        return _checked("scWarmStartTrain", self.h, _ptr(X), _ptr(T), wp, X.shape[0], int(steps))
