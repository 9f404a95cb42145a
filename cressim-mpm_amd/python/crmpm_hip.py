"""ctypes binding of the MI355X MPM solver C-ABI (include/crmpm_hip.h).

PyTorch only provides device memory and the stream; every solver operation
runs in libcrmpm_hip.so.  Loading fails loudly when the library is missing:
there is no CPU fallback.

`HipSolver` plays the role of the reference engine for one scene: it owns
the engine-side device buffers (particles, shape/geometry/SDF tables) in the
reference's layouts, binds them to a solver handle and drives substeps,
applying kinematic targets between frames like SceneImpl::_advanceTaskFn
(src/engine/scene_impl.cpp:308-330).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libcrmpm_hip.so")

import crmpm_scenes as S  # noqa: E402

MLS, PB, STANDARD = 1, 3, 5
KIND_FROM_SCENE = {S.MLS: MLS, S.PB: PB, S.STD: STANDARD}

_vp = ctypes.c_void_p
_f3 = ctypes.c_float * 3


class ShapeData(ctypes.Structure):
    _fields_ = [(n, _vp) for n in ("type", "geometryIdx", "position", "rotation", "invScale", "linearVelocity",
                                    "angularVelocity", "comInvMass", "inertiaInv0", "inertiaInv1", "force", "torque",
                                    "params0")] + [("size", ctypes.c_int)]


class GeometryData(ctypes.Structure):
    _fields_ = [("type", _vp), ("params0", _vp), ("size", ctypes.c_int)]


class GeometrySdfData(ctypes.Structure):
    _fields_ = [("dimension", _vp), ("lowerBoundCellSize", _vp), ("gradientSignedDistance", _vp)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("numMaxParticles", ctypes.c_int), ("particleOffset", ctypes.c_int), ("cellSize", ctypes.c_float),
                ("boundMin", _f3), ("boundMax", _f3)]


_f4 = ctypes.c_float * 4

EXPORTS = {
    "crHipSvd3": [_vp, ctypes.c_int, _vp, _vp, _vp, _vp],
    "crHipParticleLattice": [ctypes.c_int, _f4, ctypes.POINTER(GeometrySdfData), _f3, _f3, ctypes.c_float, _f3, _f4, _f3,
                             ctypes.c_float, _vp, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint), _vp],
    "crHipParticleLatticeNoAlloc": [ctypes.c_int, _f4, ctypes.POINTER(GeometrySdfData), _f3, _f3, ctypes.c_float, _f3,
                                    _f4, _f3, ctypes.c_float, _vp, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint), _vp],
    "crHipMeshSdfDesc": [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)],
    "crHipMeshSdf": [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float), _vp, ctypes.c_ulonglong, _vp],
    "crHipMeshSlicerSdf": [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                           ctypes.POINTER(ctypes.c_float),
                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float), _vp, ctypes.c_ulonglong,
                           _vp],
    "crHipMeshSdfWithRule": [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float), _vp, ctypes.c_ulonglong,
                             _vp],
    "crHipMeshSlicerSdfWithRule": [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                   ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_float), _vp, ctypes.c_ulonglong, _vp],
    "crHipEngineCreate": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)],
    "crHipEngineDestroy": [_vp],
    "crHipEngineCreateScene": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, _f3, _f3, ctypes.POINTER(ctypes.c_int)],
    "crHipEngineSetSceneParams": [_vp, ctypes.c_int, ctypes.c_float, _f3, ctypes.c_int, ctypes.c_int],
    "crHipEngineGetSceneHostData": [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                    ctypes.POINTER(_vp), ctypes.POINTER(_vp)],
    "crHipEngineMarkSceneDirty": [_vp, ctypes.c_int, ctypes.c_uint],
    "crHipEngineSetSceneNumParticles": [_vp, ctypes.c_int, ctypes.c_uint],
    "crHipEngineAddComputeInitialData": [_vp, ctypes.c_int, ctypes.c_uint, _vp],
    "crHipEngineSetActiveMaskRange": [_vp, ctypes.c_int, ctypes.c_uint, ctypes.c_uint, ctypes.c_ubyte],
    "crHipEngineSetSceneShapes": [_vp, ctypes.c_int, ctypes.c_int, _vp],
    "crHipEngineGetTables": [_vp, ctypes.c_int, ctypes.POINTER(ShapeData), ctypes.POINTER(GeometryData),
                             ctypes.POINTER(GeometrySdfData)],
    "crHipEngineMarkTablesDirty": [_vp, ctypes.c_uint],
    "crHipEngineAdvanceAll": [_vp, ctypes.c_float],
    "crHipEngineFetchResultsAll": [_vp],
    "crHipEngineSetGraphMode": [_vp, ctypes.c_int],
    "crHipEngineAutotune": [_vp, ctypes.c_int],
    "crHipEngineGetSceneSubsteps": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    "crHipEngineGetSceneStateArrays": [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                       ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "crHipSolverCreate": [ctypes.c_int, ctypes.c_int, ctypes.c_float, _f3, _f3, ctypes.POINTER(_vp)],
    "crHipSolverDestroy": [_vp],
    "crHipSolverGetGridDims": [_vp, ctypes.POINTER(ctypes.c_int)],
    "crHipSolverSetStream": [_vp, _vp],
    "crHipSolverSetGravity": [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float],
    "crHipSolverSetIntegrationStepSize": [_vp, ctypes.c_float],
    "crHipSolverSetSolverIterations": [_vp, ctypes.c_int],
    "crHipSolverSetShapeContactModel": [_vp, ctypes.c_int],
    "crHipSolverBindParticleData": [_vp, _vp, _vp, ctypes.c_int],
    "crHipSolverBindParticleMaterials": [_vp, _vp, _vp, ctypes.c_int],
    "crHipSolverBindActiveMask": [_vp, _vp],
    "crHipSolverSetNumActiveParticles": [_vp, ctypes.c_uint],
    "crHipSolverSetNumShapes": [_vp, ctypes.c_int],
    "crHipSolverBindShapeIds": [_vp, _vp],
    "crHipSolverBindShapeData": [_vp, ctypes.POINTER(ShapeData)],
    "crHipSolverBindGeometryData": [_vp, ctypes.POINTER(GeometryData)],
    "crHipSolverBindGeometrySdfData": [_vp, ctypes.POINTER(GeometrySdfData)],
    "crHipSolverInitialize": [_vp],
    "crHipSolverComputeInitialData": [_vp, ctypes.c_uint, _vp],
    "crHipSolverStep": [_vp, ctypes.POINTER(ctypes.c_float)],
    "crHipSolverFetchResults": [_vp],
    "crHipSolverGetStateArrays": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                  ctypes.POINTER(ctypes.c_int)],
    "crHipSolverCopyGrid": [_vp, _vp],
    "crHipSolverBinParticles": [_vp, _vp, _vp],
    "crHipSolverProfileBegin": [_vp, ctypes.c_int],
    "crHipSolverProfileEnd": [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int), ctypes.c_int],
    "crHipSolverProfileMask": [_vp, ctypes.c_uint],
    "crHipSolverAutotune": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "crHipBatchCreate": [ctypes.c_int, ctypes.c_int, ctypes.POINTER(SceneDesc), ctypes.POINTER(_vp)],
    "crHipBatchDestroy": [_vp],
    "crHipBatchSetStream": [_vp, _vp],
    "crHipBatchSetSolverIterations": [_vp, ctypes.c_int],
    "crHipBatchBindParticleData": [_vp, _vp, _vp, ctypes.c_int],
    "crHipBatchBindParticleMaterials": [_vp, _vp, _vp, ctypes.c_int],
    "crHipBatchBindActiveMask": [_vp, _vp],
    "crHipBatchInitialize": [_vp],
    "crHipBatchGetGridDims": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
    "crHipBatchSetSceneGravity": [_vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float],
    "crHipBatchSetSceneStepSize": [_vp, ctypes.c_int, ctypes.c_float],
    "crHipBatchSetSceneContactModel": [_vp, ctypes.c_int, ctypes.c_int],
    "crHipBatchSetSceneNumActiveParticles": [_vp, ctypes.c_int, ctypes.c_uint],
    "crHipBatchSetSceneResting": [_vp, ctypes.c_int, ctypes.c_int],
    "crHipBatchBindSceneShapes": [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.POINTER(ShapeData),
                                  ctypes.POINTER(GeometryData), ctypes.POINTER(GeometrySdfData)],
    "crHipBatchComputeInitialData": [_vp, ctypes.c_int, ctypes.c_uint, _vp],
    "crHipBatchStep": [_vp],
    "crHipBatchSetGraphMode": [_vp, ctypes.c_int],
    "crHipBatchAutotune": [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "crHipBatchSetTileSplit": [_vp, ctypes.c_int, ctypes.c_int],
    "crHipBatchSetLaunchSizes": [_vp, ctypes.c_int, ctypes.c_int],
    "crHipBatchSetGatherSplit": [_vp, ctypes.c_int, ctypes.c_int],
    "crHipBatchGetGatherSplit": [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "crHipSolverGetGatherSplit": [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "crHipBatchSetScatterWaves": [_vp, ctypes.c_int],
    "crHipBatchGetScatterWaves": [_vp, ctypes.POINTER(ctypes.c_int)],
    "crHipSolverSetScatterWaves": [_vp, ctypes.c_int],
    "crHipSolverGetScatterWaves": [_vp, ctypes.POINTER(ctypes.c_int)],
    "crHipSolverSetLaunchConfig": [_vp] + [ctypes.c_int] * 6,
    "crHipSolverSetGraphMode": [_vp, ctypes.c_int],
    "crHipBatchFetchResults": [_vp],
    "crHipBatchGetStateArrays": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                 ctypes.POINTER(ctypes.c_int)],
    "crHipBatchCopyGrid": [_vp, ctypes.c_int, _vp],
    "crHipBatchProfileBegin": [_vp, ctypes.c_int],
    "crHipBatchProfileMask": [_vp, ctypes.c_uint],
    "crHipBatchProfileEnd": [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int), ctypes.c_int],
    "crHipAdjointCreate": [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)],
    "crHipAdjointForward": [_vp],
    "crHipAdjointBackward": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "crHipAdjointSetTapeBudget": [_vp, ctypes.c_ulonglong],
    "crHipAdjointGetSegmentTapeBytes": [_vp, ctypes.POINTER(ctypes.c_ulonglong)],
    "crHipAdjointDestroy": [_vp],
}

PROF_CATEGORIES = ("pb_fused", "pb_p2g", "pb_g2p_integrate", "update", "mls_p2g", "mls_g2p", "std_p2g", "std_g2p",
                   "binning", "pb_scatter")

_LIB = None


def load_library(path=None):
    """Load libcrmpm_hip.so (the in-tree build, or the file CRMPM_HIP_LIB
    names) and declare every C-ABI entry point."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = path or os.environ.get("CRMPM_HIP_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"libcrmpm_hip.so not found at {path}; run __graft_entry__.build()")
    # torch first: its HIP runtime (soname libamdhip64.so.7) then satisfies the
    # library's dependency, so one runtime serves both in this process.
    import torch  # noqa: F401

    lib = ctypes.CDLL(path)
    for name, args in EXPORTS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _LIB = lib
    return lib


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


def particle_lattice(obj, device="cuda:0", capacity=None, no_alloc=False):
    """ParticleObjectImpl::initialize on the device (crHipParticleLattice, or
    crHipParticleLatticeNoAlloc with no_alloc).  obj: a dict from
    crmpm_scenes.lattice_object.  Returns (count, (m, 4) float32 tensor of
    position|mass) with m = min(count, capacity)."""
    import torch

    lib = load_library()
    fn = lib.crHipParticleLatticeNoAlloc if no_alloc else lib.crHipParticleLattice
    dev = torch.device(device)
    stream = torch.cuda.current_stream(dev).cuda_stream
    sdf_t = [torch.as_tensor(np.ascontiguousarray(obj[k], dt), device=dev)
             for k, dt in (("sdf_dims", np.int32), ("sdf_lbcs", np.float32), ("sdf_data", np.float32))]
    sdf = GeometrySdfData(sdf_t[0].data_ptr(), sdf_t[1].data_ptr(), sdf_t[2].data_ptr())
    args = (int(obj["type"]), _f4(*obj["params0"]), ctypes.byref(sdf), _f3(*obj["bounds_min"]),
            _f3(*obj["bounds_max"]), float(obj["spacing"]), _f3(*obj["position"]), _f4(*obj["rotation"]),
            _f3(*obj["inv_scale"]), float(obj["mass"]))
    n = ctypes.c_uint(0)
    check(fn(*args, None, 0, ctypes.byref(n), _vp(stream)), "particle lattice count")
    m = n.value if capacity is None else min(n.value, capacity)
    out = torch.empty((max(m, 1), 4), dtype=torch.float32, device=dev)
    check(fn(*args, _vp(out.data_ptr()), m, ctypes.byref(n), _vp(stream)), "particle lattice")
    return n.value, out[:m]


def svd3(A, device="cuda:0"):
    """The device 3x3 SVD (crHipSvd3) of row-major matrices A (n, 3, 3):
    returns U, Vt (n, 3, 3) and S (n, 3) as numpy float32 arrays with
    U diag(S) Vt = A^T, as the reference's svd() factors the transpose."""
    import torch

    lib = load_library()
    A = np.asarray(A, np.float32)
    n = len(A)
    dev = torch.device(device)
    a = torch.as_tensor(np.ascontiguousarray(A.transpose(0, 2, 1).reshape(n, 9)), device=dev)
    U = torch.empty((max(n, 1), 9), dtype=torch.float32, device=dev)
    Vt = torch.empty_like(U)
    S = torch.empty((max(n, 1), 3), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    check(lib.crHipSvd3(_vp(a.data_ptr()), n, _vp(U.data_ptr()), _vp(Vt.data_ptr()), _vp(S.data_ptr()), _vp(stream)),
          "svd3")
    torch.cuda.synchronize(dev)
    to = lambda t: t[:n].cpu().numpy().reshape(n, 3, 3).transpose(0, 2, 1)  # noqa: E731
    return to(U), to(Vt), S[:n].cpu().numpy()


def load_obj(path):
    """Triangle mesh from a Wavefront OBJ file, as the reference's ObjLoader
    reads one (tools/obj_loader.h:62-95): `v` lines give vertices (x y z),
    `f` lines the first three vertex references of a face, each the 1-based
    index before any '/'; every other line is skipped.  Returns float32
    points (n, 3) and uint32 triangles (m, 3), ready for mesh_sdf."""
    pts, tris = [], []
    with open(path) as f:
        for line in f:
            w = line.split()
            if not w:
                continue
            if w[0] == "v":
                pts.append([float(x) for x in w[1:4]])
            elif w[0] == "f":
                tris.append([int(t.split("/")[0]) - 1 for t in w[1:4]])
    return np.asarray(pts, np.float32).reshape(-1, 3), np.asarray(tris, np.uint32).reshape(-1, 3)


MESH_RULES = {"reference": 0, "exact": 1}


def mesh_sdf(points, triangles, cell_size, fatten, device="cuda:0", rule="reference"):
    """SdfBuilder's mesh -> SDF table on the device (crHipMeshSdf, or
    crHipMeshSdfWithRule for rule "exact": signs by the exact winding
    number).  points (n, 3) float32 and triangles (m, 3) uint32 on the host.
    Returns dims, lower bound (numpy) and a (dx*dy*dz, 4) float32 device
    tensor of rows (gradient, signed distance)."""
    import torch

    lib = load_library()
    dev = torch.device(device)
    stream = torch.cuda.current_stream(dev).cuda_stream
    pts = np.ascontiguousarray(points, np.float32)
    tris = np.ascontiguousarray(triangles, np.uint32)
    dims = (ctypes.c_int * 3)()
    lower = (ctypes.c_float * 3)()
    args = (_vp(pts.ctypes.data), len(pts), _vp(tris.ctypes.data), len(tris), float(cell_size), float(fatten), dims,
            lower)
    check(lib.crHipMeshSdfDesc(*args), "mesh sdf desc")
    n = dims[0] * dims[1] * dims[2]
    out = torch.empty((n, 4), dtype=torch.float32, device=dev)
    check(lib.crHipMeshSdfWithRule(*args[:6], MESH_RULES[rule], *args[6:], _vp(out.data_ptr()), n, _vp(stream)),
          "mesh sdf")
    return np.array(dims[:], np.int32), np.array(lower[:], np.float32), out


def mesh_slicer_sdf(points, triangles, cell_size, fatten, spine, device="cuda:0", rule="reference"):
    """The mesh-slicer table on the device (crHipMeshSlicerSdf, or
    crHipMeshSlicerSdfWithRule for rule "exact": rows dropped by the final
    closest point only): dims, lower bound and a (dx*dy*dz + 1, 4) float32
    device tensor, the spine area in row 0."""
    import torch

    lib = load_library()
    dev = torch.device(device)
    stream = torch.cuda.current_stream(dev).cuda_stream
    pts = np.ascontiguousarray(points, np.float32)
    tris = np.ascontiguousarray(triangles, np.uint32)
    dims = (ctypes.c_int * 3)()
    lower = (ctypes.c_float * 3)()
    check(lib.crHipMeshSdfDesc(_vp(pts.ctypes.data), len(pts), _vp(tris.ctypes.data), len(tris), float(cell_size),
                               float(fatten), dims, lower), "mesh sdf desc")
    n = dims[0] * dims[1] * dims[2] + 1
    out = torch.empty((n, 4), dtype=torch.float32, device=dev)
    check(lib.crHipMeshSlicerSdfWithRule(_vp(pts.ctypes.data), len(pts), _vp(tris.ctypes.data), len(tris),
                                         float(cell_size), float(fatten), _f4(*spine), MESH_RULES[rule], dims, lower,
                                         _vp(out.data_ptr()), n, _vp(stream)), "mesh slicer sdf")
    return np.array(dims[:], np.int32), np.array(lower[:], np.float32), out


def _read_state(torch, device, stream, F, C, V0, stride, a, e):
    """F, C as (n, 9) with element 3 * column + row, and V0 (n,), of particles
    a..e-1, float64.  The solver keeps F and C as 9 floats per particle in
    that order and V0 as one float per particle."""
    out = {}
    for name, ptr in (("F", F), ("C", C)):
        t = torch.empty((stride, 9), dtype=torch.float32, device=device)
        hip_memcpy_d2d(t.data_ptr(), ptr, t.numel() * 4, stream)
        out[name] = t[a:e].cpu().numpy().astype(np.float64)
    t = torch.empty(stride, dtype=torch.float32, device=device)
    hip_memcpy_d2d(t.data_ptr(), V0, stride * 4, stream)
    out["V0"] = t[a:e].cpu().numpy().astype(np.float64)
    return out


class HipSolver:
    """One scene on the GPU: engine-side buffers + a solver handle."""

    def __init__(self, scene, device="cuda:0", stream=None):
        import torch

        self.torch = torch
        self.lib = load_library()
        self.scene = scene
        self.device = torch.device(device)
        self.n = n = len(scene["xm"])
        lib = self.lib
        h = _vp()
        check(lib.crHipSolverCreate(KIND_FROM_SCENE[scene["kind"]], n, scene["dx"], _f3(*scene["lo"]),
                                    _f3(*scene["hi"]), ctypes.byref(h)), "create")
        self.h = h
        dims = (ctypes.c_int * 3)()
        check(lib.crHipSolverGetGridDims(h, dims), "dims")
        self.dims = tuple(dims)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        check(lib.crHipSolverSetStream(h, _vp(self.stream.cuda_stream)), "stream")
        check(lib.crHipSolverSetGravity(h, *scene["gravity"]), "gravity")
        check(lib.crHipSolverSetIntegrationStepSize(h, scene["dt"]), "dt")
        check(lib.crHipSolverSetSolverIterations(h, scene["iterations"]), "iterations")
        check(lib.crHipSolverSetShapeContactModel(h, scene.get("contact_model", 0)), "contact model")
        dev = self.device
        T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).to(dev)
        self.xm = T(scene["xm"], np.float32)
        self.v = T(scene["v"], np.float32)
        self.params = T(scene["params"], np.float32)
        self.type = T(scene["type"], np.int32)
        self.mask = T(scene["mask"], np.uint8)
        check(lib.crHipSolverBindParticleData(h, self.xm.data_ptr(), self.v.data_ptr(), n), "bind particles")
        check(lib.crHipSolverBindParticleMaterials(h, self.params.data_ptr(), self.type.data_ptr(), n), "bind materials")
        check(lib.crHipSolverBindActiveMask(h, self.mask.data_ptr()), "bind mask")
        self._upload_shapes(first=True)
        check(lib.crHipSolverInitialize(h), "initialize")
        check(lib.crHipSolverSetNumActiveParticles(h, n), "num active")
        F, C, V0, stride = _vp(), _vp(), _vp(), ctypes.c_int()
        check(lib.crHipSolverGetStateArrays(h, ctypes.byref(F), ctypes.byref(C), ctypes.byref(V0), ctypes.byref(stride)),
              "state arrays")
        self._state_ptrs = (F.value, C.value, V0.value, stride.value)

    def _upload_shapes(self, first=False):
        torch = self.torch
        sc = self.scene
        t = S.shape_tables(sc)
        dev = self.device
        T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).to(dev)
        nf4 = len(S.SHAPE_F4)
        if first or not hasattr(self, "sh"):
            # the float4 shape fields live in one device block (one view per
            # field) so a frame's kinematic update is a single pinned copy
            cap = len(t["type"])
            self.sh_block = torch.from_numpy(np.ascontiguousarray(np.stack([t[k] for k in S.SHAPE_F4]))).to(dev)
            self.sh_stage = torch.empty((nf4, cap, 4), dtype=torch.float32).pin_memory()
            self.sh = {k: self.sh_block[i] for i, k in enumerate(S.SHAPE_F4)}
            self.sh["type"] = T(t["type"], np.int32)
            self.sh["geom"] = T(t["geom"], np.int32)
            self.sh["ids"] = T(np.maximum(t["ids"], 0) if len(t["ids"]) else np.zeros(1, np.int32), np.int32)
            cap = len(t["type"])
            self.force = torch.zeros((cap, 4), dtype=torch.float32, device=dev)
            self.torque = torch.zeros((cap, 4), dtype=torch.float32, device=dev)
            self.gtype = T(t["gtype"], np.int32)
            self.gparams = T(t["gparams"], np.float32)
            self._gparams_host = np.array(t["gparams"], np.float32)
            self.sdf_dims = T(sc["sdf_dims"], np.int32)
            self.sdf_lbcs = T(sc["sdf_lbcs"], np.float32)
            self.sdf_data = T(sc["sdf_data"], np.float32)
        else:
            # pinned staging keeps the per-frame shape update asynchronous on
            # the stream; wait for the previous frame's copy before refilling it
            if getattr(self, "sh_copied", None) is not None:
                self.sh_copied.synchronize()
            self.sh_stage.numpy()[:] = np.stack([t[k] for k in S.SHAPE_F4])
            with torch.cuda.stream(self.stream):
                self.sh_block.copy_(self.sh_stage, non_blocking=True)
                if self.gparams.shape == t["gparams"].shape and not np.array_equal(self._gparams_host, t["gparams"]):
                    # Geometry::setParams between frames
                    self._gparams_host = t["gparams"].copy()
                    self.gparams.copy_(torch.from_numpy(self._gparams_host))
                self.sh_copied = torch.cuda.Event()
                self.sh_copied.record(self.stream)
        p = lambda x: x.data_ptr()
        # field order: type, geometryIdx, position, rotation, invScale, linearVelocity, angularVelocity,
        # comInvMass, inertiaInv0, inertiaInv1, force, torque, params0
        sd = ShapeData(p(self.sh["type"]), p(self.sh["geom"]), p(self.sh["pos"]), p(self.sh["rot"]),
                       p(self.sh["invScale"]), p(self.sh["linVel"]), p(self.sh["angVel"]), p(self.sh["comInvMass"]),
                       p(self.sh["inertia0"]), p(self.sh["inertia1"]), p(self.force), p(self.torque),
                       p(self.sh["params0"]), len(t["type"]))
        gd = GeometryData(p(self.gtype), p(self.gparams), len(t["gtype"]))
        sdf = GeometrySdfData(p(self.sdf_dims), p(self.sdf_lbcs), p(self.sdf_data))
        lib, h = self.lib, self.h
        check(lib.crHipSolverBindShapeData(h, ctypes.byref(sd)), "bind shapes")
        check(lib.crHipSolverBindGeometryData(h, ctypes.byref(gd)), "bind geometry")
        check(lib.crHipSolverBindGeometrySdfData(h, ctypes.byref(sdf)), "bind sdf")
        check(lib.crHipSolverBindShapeIds(h, p(self.sh["ids"])), "bind shape ids")
        check(lib.crHipSolverSetNumShapes(h, len(t["ids"])), "num shapes")

    def compute_initial_data(self, idx=None):
        torch = self.torch
        if idx is None:
            idx = np.arange(self.n, dtype=np.uint32)
        self._idx = torch.from_numpy(np.ascontiguousarray(idx, np.uint32).view(np.int32)).to(self.device)
        check(self.lib.crHipSolverComputeInitialData(self.h, len(idx), self._idx.data_ptr()), "initial data")

    def set_stream(self, stream):
        """MpmSolverGpu::setCudaStream: later work goes to `stream` (a torch
        stream); the caller orders it after the work already on the old one."""
        self.stream = stream
        check(self.lib.crHipSolverSetStream(self.h, _vp(stream.cuda_stream)), "stream")

    def step(self, n=1):
        dt = ctypes.c_float()
        for _ in range(n):
            check(self.lib.crHipSolverStep(self.h, ctypes.byref(dt)), "step")

    def run(self, steps=None):
        sc = self.scene
        steps = sc.get("steps", 1) if steps is None else steps
        if not sc["frames"]:
            self.step(steps)
            return
        done = 0
        for fr in sc["frames"]:
            if done >= steps:
                break
            S.apply_frame(sc, fr)
            self._upload_shapes()
            k = min(fr["substeps"], steps - done)
            self.step(k)
            done += k

    def sync(self):
        check(self.lib.crHipSolverFetchResults(self.h), "fetch results")

    def state(self):
        torch = self.torch
        self.sync()
        F, C, V0, stride = self._state_ptrs
        n = self.n
        out = dict(xm=self.xm.cpu().numpy().astype(np.float64), v=self.v[:, :3].cpu().numpy().astype(np.float64))
        out.update(_read_state(torch, self.device, self.stream.cuda_stream, F, C, V0, stride, 0, n))
        return out

    def grid(self):
        torch = self.torch
        total = self.dims[0] * self.dims[1] * self.dims[2]
        g = torch.empty((total, 4), dtype=torch.float32, device=self.device)
        check(self.lib.crHipSolverCopyGrid(self.h, g.data_ptr()), "copy grid")
        self.sync()
        return g.cpu().numpy().astype(np.float64)

    def binning(self):
        """Sorted block keys and the particle order of the current state."""
        torch = self.torch
        keys = torch.empty(self.n, dtype=torch.int32, device=self.device)
        order = torch.empty(self.n, dtype=torch.int32, device=self.device)
        check(self.lib.crHipSolverBinParticles(self.h, keys.data_ptr(), order.data_ptr()), "bin particles")
        self.sync()
        return keys.cpu().numpy().view(np.uint32), order.cpu().numpy().view(np.uint32)

    def profile_begin(self, max_launches):
        check(self.lib.crHipSolverProfileBegin(self.h, max_launches), "profile begin")

    def autotune(self, steps_per_trial=2):
        """Tile split and launch sizes for this scene (state unchanged);
        returns (tile waves, update waves, sub-tile size, sub-tiles per pblock)."""
        t, u, ss, sm = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(self.lib.crHipSolverAutotune(self.h, steps_per_trial, ctypes.byref(t), ctypes.byref(u), ctypes.byref(ss),
                                           ctypes.byref(sm)), "autotune")
        return t.value, u.value, ss.value, sm.value

    def set_graph_mode(self, on):
        """Replay each substep as a hipGraph (captured on first use)."""
        check(self.lib.crHipSolverSetGraphMode(self.h, 1 if on else 0), "graph mode")

    def set_launch(self, cfg):
        """Set a launch configuration (tile waves, update waves, sub-tile size,
        sub-tiles per pblock, gather sub-tile size, gather sub-tiles[, scatter
        wavefronts per tile, default 1]) without timing trials; returns it as
        autotune() + gather_split() + (scatter_waves(),) would."""
        cfg = [int(c) for c in cfg]
        waves = cfg[6] if len(cfg) > 6 else 1
        check(self.lib.crHipSolverSetLaunchConfig(self.h, *cfg[:6]), "launch config")
        self.set_scatter_waves(waves)
        return tuple(cfg[:6]) + (waves,)

    def set_scatter_waves(self, waves):
        """Wavefronts per tile of the PB and MLS scatters and the adjoints'
        velocity scatters (1, 2 or 4, or 0: four for tiles of more than 64
        particles, one for the others)."""
        check(self.lib.crHipSolverSetScatterWaves(self.h, int(waves)), "scatter waves")

    def scatter_waves(self):
        w = ctypes.c_int()
        check(self.lib.crHipSolverGetScatterWaves(self.h, ctypes.byref(w)), "scatter waves")
        return w.value

    def gather_split(self):
        """(sub-tile size, sub-tiles per pblock) of the G2P kernels."""
        a, b = ctypes.c_int(), ctypes.c_int()
        check(self.lib.crHipSolverGetGatherSplit(self.h, ctypes.byref(a), ctypes.byref(b)), "gather split")
        return a.value, b.value

    def profile_mask(self, categories):
        """Record events only for launches of the named categories (empty: pause)."""
        mask = 0
        for c in categories:
            mask |= 1 << PROF_CATEGORIES.index(c)
        check(self.lib.crHipSolverProfileMask(self.h, mask), "profile mask")

    def profile_end(self):
        n = len(PROF_CATEGORIES)
        ms = (ctypes.c_float * n)()
        cnt = (ctypes.c_int * n)()
        check(self.lib.crHipSolverProfileEnd(self.h, ms, cnt, n), "profile end")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(PROF_CATEGORIES)}

    def shape_force(self):
        self.sync()
        return self.force.cpu().numpy().astype(np.float64), self.torque.cpu().numpy().astype(np.float64)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.crHipSolverDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SCENE_POSITION_MASS, SCENE_VELOCITY, SCENE_PARAMS0, SCENE_TYPE = 1, 2, 4, 8
SCENE_NUM_PARTICLES, SCENE_COMPUTE_INITIAL_DATA, SCENE_ACTIVE_MASK = 16, 32, 64
ENGINE_SHAPES, ENGINE_GEOMETRIES, ENGINE_SDF, ENGINE_SDF_ROWS = 1, 2, 4, 8


def _host_view(ptr, shape, ctype):
    n = int(np.prod(shape))
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(n,)).reshape(shape)


class HipEngine:
    """The engine runtime (crHipEngine*): the factory's particle block and
    shape tables as pinned host arrays (numpy views), scenes as slices,
    dirty flags, advance_all / fetch_results_all."""

    def __init__(self, particle_capacity, shape_capacity=1, geometry_capacity=1, sdf_tables=1, sdf_rows=1):
        self.lib = load_library()
        h = _vp()
        check(self.lib.crHipEngineCreate(particle_capacity, shape_capacity, geometry_capacity, sdf_tables, sdf_rows,
                                         ctypes.byref(h)), "engine create")
        self.h = h
        self.caps = (shape_capacity, geometry_capacity, sdf_tables, sdf_rows)
        self.sizes = []

    def create_scene(self, kind, num_max, cell, lo, hi, dt=0.002, gravity=(0.0, -9.81, 0.0), contact_model=0,
                     iterations=10):
        sid = ctypes.c_int()
        check(self.lib.crHipEngineCreateScene(self.h, kind, num_max, cell, _f3(*lo), _f3(*hi), ctypes.byref(sid)),
              "create scene")
        check(self.lib.crHipEngineSetSceneParams(self.h, sid.value, dt, _f3(*gravity), contact_model, iterations),
              "scene params")
        self.sizes.append(num_max)
        return sid.value

    def host(self, sid):
        """numpy views of the scene's slice of the pinned host arrays."""
        p = [_vp() for _ in range(5)]
        check(self.lib.crHipEngineGetSceneHostData(self.h, sid, *[ctypes.byref(x) for x in p]), "host data")
        n = self.sizes[sid]
        return dict(xm=_host_view(p[0].value, (n, 4), ctypes.c_float), v=_host_view(p[1].value, (n, 4), ctypes.c_float),
                    params=_host_view(p[2].value, (n, 4), ctypes.c_float),
                    type=_host_view(p[3].value, (n,), ctypes.c_int),
                    mask=_host_view(p[4].value, (n,), ctypes.c_ubyte))

    def tables(self):
        """numpy views of the pinned host tables (shape fields by name, geometry and SDF arrays)."""
        sd, gd, sdf = ShapeData(), GeometryData(), GeometrySdfData()
        check(self.lib.crHipEngineGetTables(self.h, 0, ctypes.byref(sd), ctypes.byref(gd), ctypes.byref(sdf)), "tables")
        sc, gc, tc, rc = self.caps
        f4 = lambda ptr, n: _host_view(ptr, (n, 4), ctypes.c_float)  # noqa: E731
        out = dict(type=_host_view(sd.type, (sc,), ctypes.c_int), geom=_host_view(sd.geometryIdx, (sc,), ctypes.c_int))
        for k, field in zip(S.SHAPE_F4, ("position", "rotation", "invScale", "linearVelocity", "angularVelocity",
                                         "comInvMass", "inertiaInv0", "inertiaInv1", "params0")):
            out[k] = f4(getattr(sd, field), sc)
        out["force"], out["torque"] = f4(sd.force, sc), f4(sd.torque, sc)
        out["gtype"] = _host_view(gd.type, (gc,), ctypes.c_int)
        out["gparams"] = f4(gd.params0, gc)
        out["sdf_dims"] = _host_view(sdf.dimension, (tc, 4), ctypes.c_int)
        out["sdf_lbcs"] = f4(sdf.lowerBoundCellSize, tc)
        out["sdf_data"] = f4(sdf.gradientSignedDistance, rc)
        return out

    def mark(self, sid, flags):
        check(self.lib.crHipEngineMarkSceneDirty(self.h, sid, flags), "mark dirty")

    def mark_tables(self, flags):
        check(self.lib.crHipEngineMarkTablesDirty(self.h, flags), "mark tables")

    def set_num_particles(self, sid, n):
        check(self.lib.crHipEngineSetSceneNumParticles(self.h, sid, n), "num particles")

    def add_initial_data(self, sid, idx):
        self._idx = np.ascontiguousarray(idx, np.uint32)
        check(self.lib.crHipEngineAddComputeInitialData(self.h, sid, len(self._idx), _vp(self._idx.ctypes.data)),
              "initial data")

    def set_shapes(self, sid, ids):
        a = np.ascontiguousarray(ids, np.int32)
        check(self.lib.crHipEngineSetSceneShapes(self.h, sid, len(a), _vp(a.ctypes.data) if len(a) else None),
              "scene shapes")

    def advance_all(self, dt):
        check(self.lib.crHipEngineAdvanceAll(self.h, dt), "advance all")

    def fetch_results_all(self):
        check(self.lib.crHipEngineFetchResultsAll(self.h), "fetch results all")

    def set_graph_mode(self, on):
        check(self.lib.crHipEngineSetGraphMode(self.h, 1 if on else 0), "graph mode")

    def autotune(self, steps_per_trial=2):
        check(self.lib.crHipEngineAutotune(self.h, steps_per_trial), "autotune")

    def device_tables(self):
        """Device pointers of the engine's tables (CrHipShapeData,
        CrHipGeometryData, CrHipGeometrySdfData)."""
        sd, gd, sdf = ShapeData(), GeometryData(), GeometrySdfData()
        check(self.lib.crHipEngineGetTables(self.h, 1, ctypes.byref(sd), ctypes.byref(gd), ctypes.byref(sdf)),
              "device tables")
        return sd, gd, sdf

    def substeps(self, sid):
        n = ctypes.c_int()
        check(self.lib.crHipEngineGetSceneSubsteps(self.h, sid, ctypes.byref(n)), "substeps")
        return n.value

    def state(self, sid):
        """F, C (n, 9) and V0 of the scene, float64 (after an advance)."""
        import torch

        F, C, V0, stride, first = _vp(), _vp(), _vp(), ctypes.c_int(), ctypes.c_int()
        check(self.lib.crHipEngineGetSceneStateArrays(self.h, sid, ctypes.byref(F), ctypes.byref(C), ctypes.byref(V0),
                                                      ctypes.byref(stride), ctypes.byref(first)), "state arrays")
        dev = torch.device("cuda:0")
        stream = torch.cuda.current_stream(dev)
        a = first.value
        out = _read_state(torch, dev, stream.cuda_stream, F.value, C.value, V0.value, stride.value, a,
                          a + self.sizes[sid])
        stream.synchronize()
        return out

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.crHipEngineDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# geometry types whose params0 (x, y) hold an SDF row offset and table id as int bits
_TABLE_GEOMETRIES = (S.TRIMESH, S.TRIMESH_SLICER, S.LINES)


def engine_from_scenes(scenes, held_back=None):
    """An engine holding scene dicts (as HipSolver takes them): particles
    written to the host block; shapes, geometries and SDF tables appended to
    the factory tables, geometry rows and SDF offsets and ids shifted; every
    particle but held_back[i] of scene i allocated and queued for
    computeInitialData.  Returns (engine, [(scene id, first shape row, shape
    count, first SDF row, first SDF table)], host tables)."""
    held_back = held_back or {}
    n_shapes = sum(len(sc["shape_list"]) for sc in scenes)
    n_geoms = sum(len(sc["geoms"]) for sc in scenes)
    n_tables = sum(len(sc["sdf_dims"]) for sc in scenes)
    n_rows = sum(len(sc["sdf_data"]) for sc in scenes)
    eng = HipEngine(sum(len(sc["xm"]) for sc in scenes), max(1, n_shapes), max(1, n_geoms), n_tables, n_rows)
    tabs = eng.tables()
    ids, s0, g0, r0, t0 = [], 0, 0, 0, 0
    for i, sc in enumerate(scenes):
        n = len(sc["xm"])
        sid = eng.create_scene(KIND_FROM_SCENE[sc["kind"]], n, sc["dx"], sc["lo"], sc["hi"], dt=sc["dt"],
                               gravity=sc["gravity"], contact_model=sc.get("contact_model", 0),
                               iterations=sc["iterations"])
        hd = eng.host(sid)
        for k in ("xm", "v", "params", "type", "mask"):
            hd[k][:] = sc[k]
        t = S.shape_tables(sc)
        k = len(sc["shape_list"])
        for j in range(k):
            tabs["type"][s0 + j] = t["type"][j]
            tabs["geom"][s0 + j] = t["geom"][j] + g0
            for f in S.SHAPE_F4:
                tabs[f][s0 + j] = t[f][j]
        for j, (gt, gp) in enumerate(sc["geoms"]):
            tabs["gtype"][g0 + j] = gt
            gp = np.array(gp, np.float32)
            if gt in _TABLE_GEOMETRIES:
                bits = gp.view(np.int32)
                bits[0] += r0
                bits[1] += t0
            tabs["gparams"][g0 + j] = gp
        nt, nr = len(sc["sdf_dims"]), len(sc["sdf_data"])
        tabs["sdf_dims"][t0:t0 + nt] = sc["sdf_dims"]
        tabs["sdf_lbcs"][t0:t0 + nt] = sc["sdf_lbcs"]
        tabs["sdf_data"][r0:r0 + nr] = sc["sdf_data"]
        eng.set_shapes(sid, np.arange(s0, s0 + k))
        m = n - held_back.get(i, 0)
        eng.set_num_particles(sid, m)
        eng.add_initial_data(sid, np.arange(m))
        eng.mark(sid, SCENE_POSITION_MASS | SCENE_VELOCITY | SCENE_PARAMS0 | SCENE_TYPE | SCENE_ACTIVE_MASK)
        ids.append((sid, s0, k, r0, t0))
        s0 += k
        g0 += len(sc["geoms"])
        r0 += nr
        t0 += nt
    eng.mark_tables(ENGINE_SHAPES | ENGINE_GEOMETRIES | ENGINE_SDF | ENGINE_SDF_ROWS)
    return eng, ids, tabs


class DeviceShapeTables:
    """A scene's ShapeData / GeometryData / GeometrySdfData tables in device memory."""

    def __init__(self, scene, device):
        import torch

        t = S.shape_tables(scene)
        T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dt)).to(device)
        self.sh = {k: T(t[k], np.float32) for k in S.SHAPE_F4}
        self.sh["type"] = T(t["type"], np.int32)
        self.sh["geom"] = T(t["geom"], np.int32)
        self.ids = T(np.maximum(t["ids"], 0) if len(t["ids"]) else np.zeros(1, np.int32), np.int32)
        cap = len(t["type"])
        self.force = torch.zeros((cap, 4), dtype=torch.float32, device=device)
        self.torque = torch.zeros((cap, 4), dtype=torch.float32, device=device)
        self.gtype = T(t["gtype"], np.int32)
        self.gparams = T(t["gparams"], np.float32)
        self.sdf_dims = T(scene["sdf_dims"], np.int32)
        self.sdf_lbcs = T(scene["sdf_lbcs"], np.float32)
        self.sdf_data = T(scene["sdf_data"], np.float32)
        self.num_shapes = len(t["ids"])
        p = lambda x: x.data_ptr()
        self.sd = ShapeData(p(self.sh["type"]), p(self.sh["geom"]), p(self.sh["pos"]), p(self.sh["rot"]),
                            p(self.sh["invScale"]), p(self.sh["linVel"]), p(self.sh["angVel"]),
                            p(self.sh["comInvMass"]), p(self.sh["inertia0"]), p(self.sh["inertia1"]), p(self.force),
                            p(self.torque), p(self.sh["params0"]), cap)
        self.gd = GeometryData(p(self.gtype), p(self.gparams), len(t["gtype"]))
        self.sdf = GeometrySdfData(p(self.sdf_dims), p(self.sdf_lbcs), p(self.sdf_data))


class HipBatch:
    """Many scenes of one solver kind advanced together (crHipBatch*): one
    launch per phase over all scenes.  Scene particles are slices of one set
    of device arrays, like the engine's per-scene ranges of the factory's
    particle block.  Scenes with kinematic frames are not supported here."""

    def __init__(self, scenes, device="cuda:0", stream=None):
        import torch

        self.torch = torch
        self.lib = lib = load_library()
        self.scenes = scenes
        self.device = dev = torch.device(device)
        kinds = {sc["kind"] for sc in scenes}
        iters = {sc["iterations"] for sc in scenes}
        if len(kinds) != 1 or len(iters) != 1:
            raise ValueError("a batch holds scenes of one solver kind and one iteration count")
        if any(sc["frames"] for sc in scenes):
            raise ValueError("HipBatch does not apply kinematic frames")
        self.counts = [len(sc["xm"]) for sc in scenes]
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        ns = len(scenes)
        descs = (SceneDesc * ns)()
        for i, sc in enumerate(scenes):
            descs[i] = SceneDesc(self.counts[i], int(self.offsets[i]), sc["dx"], _f3(*sc["lo"]), _f3(*sc["hi"]))
        h = _vp()
        check(lib.crHipBatchCreate(KIND_FROM_SCENE[scenes[0]["kind"]], ns, descs, ctypes.byref(h)), "batch create")
        self.h = h
        self.dims = []
        for i in range(ns):
            d = (ctypes.c_int * 3)()
            check(lib.crHipBatchGetGridDims(h, i, d), "dims")
            self.dims.append(tuple(d))
        self.stream = stream if stream is not None else torch.cuda.current_stream(dev)
        check(lib.crHipBatchSetStream(h, _vp(self.stream.cuda_stream)), "stream")
        check(lib.crHipBatchSetSolverIterations(h, scenes[0]["iterations"]), "iterations")
        cat = lambda key, dt: torch.from_numpy(np.ascontiguousarray(np.concatenate([sc[key] for sc in scenes]), dt)).to(dev)
        self.xm = cat("xm", np.float32)
        self.v = cat("v", np.float32)
        self.params = cat("params", np.float32)
        self.type = cat("type", np.int32)
        self.mask = cat("mask", np.uint8)
        n = int(self.offsets[-1])
        self.n = n
        check(lib.crHipBatchBindParticleData(h, self.xm.data_ptr(), self.v.data_ptr(), n), "bind particles")
        check(lib.crHipBatchBindParticleMaterials(h, self.params.data_ptr(), self.type.data_ptr(), n), "bind materials")
        check(lib.crHipBatchBindActiveMask(h, self.mask.data_ptr()), "bind mask")
        self.tables = []
        for i, sc in enumerate(scenes):
            check(lib.crHipBatchSetSceneGravity(h, i, *sc["gravity"]), "gravity")
            check(lib.crHipBatchSetSceneStepSize(h, i, sc["dt"]), "dt")
            check(lib.crHipBatchSetSceneContactModel(h, i, sc.get("contact_model", 0)), "contact model")
            t = DeviceShapeTables(sc, dev)
            self.tables.append(t)
            check(lib.crHipBatchBindSceneShapes(h, i, t.num_shapes, t.ids.data_ptr(), ctypes.byref(t.sd),
                                                ctypes.byref(t.gd), ctypes.byref(t.sdf)), "bind shapes")
        check(lib.crHipBatchInitialize(h), "initialize")
        for i in range(ns):
            check(lib.crHipBatchSetSceneNumActiveParticles(h, i, self.counts[i]), "num active")
        F, C, V0, stride = _vp(), _vp(), _vp(), ctypes.c_int()
        check(lib.crHipBatchGetStateArrays(h, ctypes.byref(F), ctypes.byref(C), ctypes.byref(V0), ctypes.byref(stride)),
              "state arrays")
        self._state_ptrs = (F.value, C.value, V0.value, stride.value)

    def compute_initial_data(self):
        torch = self.torch
        self._idx = []
        for i, c in enumerate(self.counts):
            idx = torch.arange(c, dtype=torch.int32, device=self.device)
            self._idx.append(idx)
            check(self.lib.crHipBatchComputeInitialData(self.h, i, c, idx.data_ptr()), "initial data")

    def step(self, n=1):
        for _ in range(n):
            check(self.lib.crHipBatchStep(self.h), "batch step")

    def set_graph_mode(self, on):
        check(self.lib.crHipBatchSetGraphMode(self.h, 1 if on else 0), "graph mode")

    def autotune(self, steps_per_trial=3):
        """Returns (tile waves, update waves, sub-tile size, sub-tiles per pblock)."""
        t, u, ss, sm = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(self.lib.crHipBatchAutotune(self.h, steps_per_trial, ctypes.byref(t), ctypes.byref(u), ctypes.byref(ss),
                                          ctypes.byref(sm)), "autotune")
        return t.value, u.value, ss.value, sm.value

    def set_tile_split(self, sub_size, sub_max):
        check(self.lib.crHipBatchSetTileSplit(self.h, sub_size, sub_max), "tile split")

    def set_gather_split(self, sub_size, sub_max):
        check(self.lib.crHipBatchSetGatherSplit(self.h, sub_size, sub_max), "gather split")

    def gather_split(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        check(self.lib.crHipBatchGetGatherSplit(self.h, ctypes.byref(a), ctypes.byref(b)), "gather split")
        return a.value, b.value

    def set_launch_sizes(self, tile_waves, update_waves):
        check(self.lib.crHipBatchSetLaunchSizes(self.h, tile_waves, update_waves), "launch sizes")

    def set_scatter_waves(self, waves):
        check(self.lib.crHipBatchSetScatterWaves(self.h, int(waves)), "scatter waves")

    def scatter_waves(self):
        w = ctypes.c_int()
        check(self.lib.crHipBatchGetScatterWaves(self.h, ctypes.byref(w)), "scatter waves")
        return w.value

    def sync(self):
        check(self.lib.crHipBatchFetchResults(self.h), "fetch results")

    def state(self, i):
        """State of scene i (as HipSolver.state)."""
        torch = self.torch
        self.sync()
        F, C, V0, stride = self._state_ptrs
        a, e = int(self.offsets[i]), int(self.offsets[i + 1])
        out = dict(xm=self.xm[a:e].cpu().numpy().astype(np.float64), v=self.v[a:e, :3].cpu().numpy().astype(np.float64))
        out.update(_read_state(torch, self.device, self.stream.cuda_stream, F, C, V0, stride, a, e))
        return out

    def grid(self, i):
        torch = self.torch
        d = self.dims[i]
        g = torch.empty((d[0] * d[1] * d[2], 4), dtype=torch.float32, device=self.device)
        check(self.lib.crHipBatchCopyGrid(self.h, i, g.data_ptr()), "copy grid")
        self.sync()
        return g.cpu().numpy().astype(np.float64)

    def profile_begin(self, max_launches):
        check(self.lib.crHipBatchProfileBegin(self.h, max_launches), "profile begin")

    def profile_end(self):
        n = len(PROF_CATEGORIES)
        ms = (ctypes.c_float * n)()
        cnt = (ctypes.c_int * n)()
        check(self.lib.crHipBatchProfileEnd(self.h, ms, cnt, n), "profile end")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(PROF_CATEGORIES)}

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.crHipBatchDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DifferentiableMls:
    """Differentiable rollout of one MLS scene without rigid shapes: `steps`
    substeps from the scene's initial positions, with per-particle Lame
    parameters and initial velocities as inputs and final positions as the
    output.  Gradients come from the hand-written adjoint kernels
    (crHipAdjoint*), checkpointing every `checkpoint_every` substeps.
    V0 comes from computeInitialData on the initial positions once; every
    rollout restarts from those positions with F = I and C = 0."""

    KIND = S.MLS

    def __init__(self, scene, steps, checkpoint_every=10, device="cuda:0", tune=False, graph=True):
        import torch

        if scene["kind"] != self.KIND or scene["shape_list"]:
            raise ValueError(f"{type(self).__name__} covers scenes of solver kind {self.KIND} without rigid shapes")
        self.torch = torch
        self.solver = HipSolver(scene, device=device)
        self.solver.compute_initial_data()
        # tune: True runs the autotuner; a launch configuration (as
        # HipSolver.set_launch takes) is set as given
        if isinstance(tune, (tuple, list)):
            self.tuned = self.solver.set_launch(tune)
        elif tune:
            self.tuned = (self.solver.autotune(steps_per_trial=2) + self.solver.gather_split() +
                          (self.solver.scatter_waves(),))
        # the forward pass (and the backward's re-runs that keep no tape)
        # replay each substep as a hipGraph
        self.solver.set_graph_mode(graph)
        self.steps = steps
        self.n = self.solver.n
        h = _vp()
        check(self.solver.lib.crHipAdjointCreate(self.solver.h, steps, checkpoint_every, ctypes.byref(h)), "adjoint create")
        self.h = h
        self.x0 = self.solver.xm.clone()
        F, C, V0, stride = self.solver._state_ptrs
        self.stride = stride
        dev = self.solver.device
        eye = torch.zeros((stride, 9), dtype=torch.float32, device=dev)
        eye[:, 0] = eye[:, 4] = eye[:, 8] = 1.0
        self._eye = eye
        self._zero = torch.zeros((stride, 9), dtype=torch.float32, device=dev)

    def _reset(self, lam, mu, v0):
        s = self.solver
        torch = self.torch
        s.xm.copy_(self.x0)
        s.v[:, :3] = v0.detach().to(torch.float32)
        s.v[:, 3] = 0.0
        s.params[:, 0] = lam.detach().to(torch.float32)
        s.params[:, 1] = mu.detach().to(torch.float32)
        F, C, _, _ = s._state_ptrs
        sh = s.stream.cuda_stream
        hip_memcpy_d2d(F, self._eye.data_ptr(), self._eye.numel() * 4, sh)
        hip_memcpy_d2d(C, self._zero.data_ptr(), self._zero.numel() * 4, sh)

    def forward(self, lam, mu, v0):
        """Final positions (N, 3), float32; no graph (use rollout() for autograd).
        backward() differentiates the latest forward: each forward starts a
        new generation, which a rollout's autograd node checks."""
        self._reset(lam, mu, v0)
        check(self.solver.lib.crHipAdjointForward(self.h), "adjoint forward")
        self.generation = getattr(self, "generation", 0) + 1
        return self.solver.xm[:, :3].clone()

    def backward(self, grad_x, grad_v=None):
        """dL/dlambda, dL/dmu (float64) and dL/dv0 (float32) for dL/dx of the last forward."""
        torch = self.torch
        dev = self.solver.device
        n = self.n
        gx = torch.zeros((n, 4), dtype=torch.float32, device=dev)
        gx[:, :3] = grad_x.to(torch.float32)
        gv = None
        if grad_v is not None:
            gv = torch.zeros((n, 4), dtype=torch.float32, device=dev)
            gv[:, :3] = grad_v.to(torch.float32)
        dlam = torch.empty(n, dtype=torch.float64, device=dev)
        dmu = torch.empty(n, dtype=torch.float64, device=dev)
        dv0 = torch.empty((n, 4), dtype=torch.float32, device=dev)
        check(self.solver.lib.crHipAdjointBackward(self.h, gx.data_ptr(), gv.data_ptr() if gv is not None else None,
                                                   dlam.data_ptr(), dmu.data_ptr(), None, dv0.data_ptr()),
              "adjoint backward")
        return dlam, dmu, dv0[:, :3]

    def rollout(self, lam, mu, v0):
        """Final positions as a differentiable function of (lambda, mu, v0)."""
        return _MlsRollout.apply(lam, mu, v0, self)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.solver.lib.crHipAdjointDestroy(self.h)
            self.h = None
        self.solver.close()


def _rollout_function():
    import torch

    class MlsRollout(torch.autograd.Function):
        """torch.autograd.Function over the GPU forward and adjoint kernels."""

        @staticmethod
        def forward(ctx, lam, mu, v0, sim):
            ctx.sim = sim
            ctx.dtypes = (lam.dtype, mu.dtype, v0.dtype)
            x = sim.forward(lam, mu, v0)
            ctx.generation = sim.generation
            return x

        @staticmethod
        def backward(ctx, grad_x):
            # the adjoint recomputes from the checkpoints and parameters of the
            # sim's latest forward: a later forward would give this graph its gradients
            if ctx.sim.generation != ctx.generation:
                raise RuntimeError("DifferentiableMls: another forward ran on this sim after this rollout; "
                                   "backward of an earlier rollout is not available (one rollout per backward)")
            dlam, dmu, dv0 = ctx.sim.backward(grad_x)
            tl, tm, tv = ctx.dtypes
            return dlam.to(tl), dmu.to(tm), dv0.to(tv), None

    return MlsRollout


def _set_tape_budget(self, nbytes):
    """Device memory the backward pass may take for checkpoint segments'
    tapes (default 8 GiB).  PB: two segments' tapes let the re-run of one
    segment overlap the reverse of the next; with one, each substep is re-run
    once in the backward pass; below it, twice.  MLS: a second set of
    per-segment buffers within the budget lets the re-runs overlap the
    reverses.  The gradients are bitwise the same either way."""
    check(self.solver.lib.crHipAdjointSetTapeBudget(self.h, int(nbytes)), "adjoint tape budget")


DifferentiableMls.set_tape_budget = _set_tape_budget


class _LazyRollout:
    fn = None

    @classmethod
    def apply(cls, *args):
        if cls.fn is None:
            cls.fn = _rollout_function()
        return cls.fn.apply(*args)


_MlsRollout = _LazyRollout


class DifferentiablePb(DifferentiableMls):
    """Differentiable rollout of one PB-MPM scene without rigid shapes: the
    per-particle inputs are the constraint's relaxation alpha (params0.x)
    and rotation ratio (params0.y) and the initial velocity; forward(alpha,
    ratio, v0) gives the final positions and backward(dL/dx) returns
    dL/dalpha, dL/dratio (float64) and dL/dv0.  The adjoint runs the SVD's
    polar-rotation derivative (crHipAdjoint* on a PB solver); the solver
    iterations are the scene's."""

    KIND = S.PB

    def segment_tape_bytes(self):
        """Bytes of one checkpoint segment's tape."""
        v = ctypes.c_ulonglong()
        check(self.solver.lib.crHipAdjointGetSegmentTapeBytes(self.h, ctypes.byref(v)), "segment tape bytes")
        return v.value


_HIPRT = None


def hip_runtime():
    """The HIP runtime library that torch itself loaded."""
    global _HIPRT
    if _HIPRT is None:
        import glob

        import torch

        cands = sorted(glob.glob(os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so*")))
        cands += ["libamdhip64.so"]
        for c in cands:
            try:
                _HIPRT = ctypes.CDLL(c)
                break
            except OSError:
                continue
        if _HIPRT is None:
            raise RuntimeError("HIP runtime library not found")
        _HIPRT.hipMemcpyAsync.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp]
        _HIPRT.hipMemcpyAsync.restype = ctypes.c_int
        _HIPRT.hipStreamSynchronize.argtypes = [_vp]
        _HIPRT.hipStreamSynchronize.restype = ctypes.c_int
    return _HIPRT


def hip_memcpy_d2d(dst, src, nbytes, stream_handle):
    rt = hip_runtime()
    check(rt.hipMemcpyAsync(_vp(dst), _vp(src), nbytes, 3, _vp(stream_handle)), "hipMemcpyAsync")  # 3: device to device
    check(rt.hipStreamSynchronize(_vp(stream_handle)), "hipStreamSynchronize")
