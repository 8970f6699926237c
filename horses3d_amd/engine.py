"""Python host binding of the hx3d C ABI (include/hx3d.h, horses3d_amd/libhx3d.so).

The engine is the drop-in for the reference's ComputeTimeDerivative / TakeRK3Step
(Solver/src/NavierStokesSolver/SpatialDiscretization.f90:227-320,
Solver/src/libs/timeintegrator/ExplicitMethods.f90:667-788).  A *case* is a dict of numpy
arrays in the reference's HexMesh element/face storage contract (see include/hx3d.h); the
case builders in horses3d_amd.mesh produce one.

There is no CPU fallback: if the HIP library is missing or no GPU is visible, every call
raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhx3d.so")

RIEMANN = {"roe": 0, "lax-friedrichs": 1, "lxf": 1}
INVISCID = {"standard": 0, "split-form": 1}
AVERAGING = {"standard": 0, "ducros": 1, "kennedy-gruber": 2, "pirozzoli": 3, "morinishi": 4}
LES = {None: 0, "none": 0, "smagorinsky": 1, "wale": 2}
LES_DEFAULT_C = {0: 0.0, 1: 0.2, 2: 0.325}   # LESModels.f90:248, :352
NODES = {"gauss-lobatto": 0, "gauss": 1}      # "discretization nodes"
MONITORS = {"kinetic energy": 0, "kinetic energy rate": 1, "enstrophy": 2, "mean velocity": 3,
            "internal energy": 4, "entropy": 5, "math entropy": 6}


class CaseDesc(C.Structure):
    _fields_ = [("K", C.c_int), ("N", C.c_void_p), ("faceID", C.c_void_p), ("faceSide", C.c_void_p),
                ("F", C.c_int), ("fElem", C.c_void_p), ("fESide", C.c_void_p), ("fRot", C.c_void_p),
                ("Ja", C.c_void_p), ("J", C.c_void_p),
                ("fNormal", C.c_void_p), ("fT1", C.c_void_p), ("fT2", C.c_void_p),
                ("fJac", C.c_void_p),
                ("mach", C.c_double), ("reynolds", C.c_double), ("prandtl", C.c_double),
                ("euler", C.c_int), ("riemann", C.c_int), ("inviscid", C.c_int),
                ("averaging", C.c_int), ("lambda_stab", C.c_double), ("les", C.c_int),
                ("les_c", C.c_double), ("nodes", C.c_int)]


class ControlRun(C.Structure):
    _fields_ = [("cfl", C.c_double), ("dcfl", C.c_double), ("dt", C.c_double), ("tolerance", C.c_double),
                ("steps", C.c_int), ("output_interval", C.c_int), ("restart", C.c_int)]


class HX3DError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class HX3DDiverged(HX3DError):
    """HX3D_ENAN: a NaN appeared in the state (the reference's checkForNan stop)."""


HX3D_ENAN = 5


def _error(msg, code):
    return (HX3DDiverged if code == HX3D_ENAN else HX3DError)(msg, code)


_lib = None

EXPORTS = ["hx3d_engine_create", "hx3d_engine_create_multi", "hx3d_multi_set_active", "hx3d_multi_status",
           "hx3d_engine_destroy", "hx3d_last_error", "hx3d_num_nodes",
           "hx3d_set_state", "hx3d_get_state", "hx3d_get_state_nodes", "hx3d_set_state_device", "hx3d_get_state_device",
           "hx3d_compute_qdot", "hx3d_rk3_step", "hx3d_set_ctd_after_steps", "hx3d_get_qdot", "hx3d_max_dt", "hx3d_advance", "hx3d_advance_dt", "hx3d_residuals",
           "hx3d_residuals_noalloc",
           "hx3d_kinetic_energy", "hx3d_check_nan", "hx3d_synchronize", "hx3d_query", "hx3d_get_time", "hx3d_engine_device_bytes",
           "hx3d_device_memory", "hx3d_set_timing", "hx3d_set_step_graph",
           "hx3d_autotune", "hx3d_set_tiling", "hx3d_get_tiling",
           "hx3d_kernel_times", "hx3d_mesh_read_specmesh", "hx3d_mesh_read_specmesh_nodes", "hx3d_mesh_error",
           "hx3d_mesh_num_nodes",
           "hx3d_mesh_num_face_nodes", "hx3d_mesh_coords", "hx3d_mesh_fill_desc", "hx3d_mesh_destroy",
           "hx3d_hsol_error", "hx3d_hsol_reference_values", "hx3d_hsol_write", "hx3d_hsol_write_nodes",
           "hx3d_hsol_read_header",
           "hx3d_hsol_read", "hx3d_save_solution", "hx3d_load_solution", "hx3d_volume_monitor",
           "hx3d_control_read", "hx3d_control_error", "hx3d_control_value", "hx3d_control_boundary",
           "hx3d_control_case", "hx3d_control_destroy"]


def load_library(path=LIB_PATH):
    """Load the HIP engine library (does not touch the GPU)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise HX3DError("hx3d HIP library not built: %s (run __graft_entry__.build())" % path)
    L = C.CDLL(path)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    L.hx3d_engine_create.argtypes = [C.POINTER(CaseDesc), C.POINTER(C.c_void_p)]
    L.hx3d_engine_destroy.argtypes = [P]
    L.hx3d_engine_create_multi.argtypes = [P, I, C.POINTER(C.c_void_p)]
    L.hx3d_multi_set_active.argtypes = [P, I, I]
    L.hx3d_multi_status.argtypes = [P, P, P]
    L.hx3d_last_error.argtypes = [P]
    L.hx3d_last_error.restype = C.c_char_p
    L.hx3d_num_nodes.argtypes = [P]
    L.hx3d_num_nodes.restype = C.c_long
    for fn in ("hx3d_set_state", "hx3d_get_state", "hx3d_set_state_device", "hx3d_get_state_device"):
        getattr(L, fn).argtypes = [P, P]
    L.hx3d_get_state_nodes.argtypes = [P, C.c_long, C.c_long, P]
    L.hx3d_compute_qdot.argtypes = [P, D, P]
    L.hx3d_rk3_step.argtypes = [P, D, D]
    L.hx3d_set_ctd_after_steps.argtypes = [P, I]
    L.hx3d_get_qdot.argtypes = [P, P]
    L.hx3d_max_dt.argtypes = [P, D, D, C.POINTER(D)]
    L.hx3d_advance.argtypes = [P, I, D, D, C.POINTER(D)]
    L.hx3d_advance_dt.argtypes = [P, I, D, C.POINTER(D)]
    L.hx3d_residuals.argtypes = [P, P]
    L.hx3d_residuals_noalloc.argtypes = [P, P]
    L.hx3d_kinetic_energy.argtypes = [P, C.POINTER(D)]
    L.hx3d_check_nan.argtypes = [P, C.POINTER(I)]
    L.hx3d_synchronize.argtypes = [P]
    L.hx3d_query.argtypes = [P, C.POINTER(I)]
    L.hx3d_get_time.argtypes = [P, C.POINTER(D)]
    L.hx3d_engine_device_bytes.argtypes = [P]
    L.hx3d_engine_device_bytes.restype = C.c_long
    L.hx3d_device_memory.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    L.hx3d_set_timing.argtypes = [P, I]
    L.hx3d_set_step_graph.argtypes = [P, I]
    L.hx3d_autotune.argtypes = [P, I, C.POINTER(I), P]
    L.hx3d_set_tiling.argtypes = [P, I]
    L.hx3d_get_tiling.argtypes = [P, C.POINTER(I), C.POINTER(I)]
    L.hx3d_kernel_times.argtypes = [P, P]
    L.hx3d_mesh_read_specmesh.argtypes = [C.c_char_p, I, P, C.c_char_p, C.POINTER(C.c_void_p)]
    L.hx3d_mesh_read_specmesh_nodes.argtypes = [C.c_char_p, I, P, C.c_char_p, I, C.POINTER(C.c_void_p)]
    L.hx3d_mesh_error.argtypes = [P]
    L.hx3d_mesh_error.restype = C.c_char_p
    L.hx3d_mesh_num_nodes.argtypes = [P]
    L.hx3d_mesh_num_nodes.restype = C.c_long
    L.hx3d_mesh_num_face_nodes.argtypes = [P]
    L.hx3d_mesh_num_face_nodes.restype = C.c_long
    L.hx3d_mesh_coords.argtypes = [P]
    L.hx3d_mesh_coords.restype = C.POINTER(C.c_double)
    L.hx3d_mesh_fill_desc.argtypes = [P, C.POINTER(CaseDesc)]
    L.hx3d_mesh_destroy.argtypes = [P]
    L.hx3d_hsol_error.argtypes = []
    L.hx3d_hsol_error.restype = C.c_char_p
    L.hx3d_hsol_reference_values.argtypes = [D, P]
    L.hx3d_hsol_write.argtypes = [C.c_char_p, C.c_char_p, I, P, I, D, P, P, P]
    L.hx3d_hsol_write_nodes.argtypes = [C.c_char_p, C.c_char_p, I, P, I, D, P, P, P, I]
    L.hx3d_hsol_read_header.argtypes = [C.c_char_p, P, P, P, P, P, P]
    L.hx3d_hsol_read.argtypes = [C.c_char_p, I, P, P, P]
    L.hx3d_save_solution.argtypes = [P, C.c_char_p, C.c_char_p, I, D]
    L.hx3d_load_solution.argtypes = [P, C.c_char_p, C.POINTER(I), C.POINTER(D)]
    L.hx3d_volume_monitor.argtypes = [P, I, C.POINTER(D)]
    L.hx3d_control_read.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    L.hx3d_control_error.argtypes = [P]
    L.hx3d_control_error.restype = C.c_char_p
    L.hx3d_control_value.argtypes = [P, C.c_char_p]
    L.hx3d_control_value.restype = C.c_char_p
    L.hx3d_control_boundary.argtypes = [P, C.c_char_p, C.c_char_p, C.c_char_p, I]
    L.hx3d_control_case.argtypes = [P, C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(CaseDesc), C.POINTER(ControlRun)]
    L.hx3d_control_destroy.argtypes = [P]
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def read_specmesh(path, N=None, orders=None, periodic=None, nodes="gauss-lobatto"):
    """A SpecMesh file as a case dict, through the C front end (hx3d_mesh_read_specmesh_nodes).

    N: one order for every element, or orders: (K, 3) per-element orders.  periodic: the
    periodic boundaries of the .control file as {name: coupled name} (or pairs).  nodes: the
    "discretization nodes" the geometry is built on ("gauss-lobatto" or "gauss").  Runs on the
    host only."""
    L = load_library()
    if orders is not None:
        ords = np.ascontiguousarray(orders, dtype=np.int32)
        op = _p(ords)
    else:
        ords, op = None, None
    if periodic is None:
        spec = None
    else:
        items = periodic.items() if isinstance(periodic, dict) else periodic
        spec = ";".join("%s=%s" % (a, b) for a, b in items).encode()
    h = C.c_void_p()
    rc = L.hx3d_mesh_read_specmesh_nodes(os.fsencode(path), int(N or 0), op, spec, NODES[nodes.lower()],
                                         C.byref(h))
    try:
        if rc != 0:
            msg = L.hx3d_mesh_error(h).decode() if h else "out of memory"
            raise HX3DError("hx3d_mesh_read_specmesh: %s (code %d)" % (msg, rc))
        d = CaseDesc()
        L.hx3d_mesh_fill_desc(h, C.byref(d))
        return _case_from_desc(L, h, d)
    finally:
        L.hx3d_mesh_destroy(h)


def _case_from_desc(L, h, d):
    """Copies of the mesh arrays a filled descriptor points to, as a case dict."""
    nodes, fnodes, K, F = L.hx3d_mesh_num_nodes(h), L.hx3d_mesh_num_face_nodes(h), d.K, d.F

    def arr(ptr, n, ty, shape):
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ty)), shape=(n,)).copy().reshape(shape)

    Nk = arr(d.N, 3 * K, C.c_int32, (K, 3))
    case = dict(K=K, F=F, N=Nk, faceID=arr(d.faceID, 6 * K, C.c_int32, (K, 6)),
                faceSide=arr(d.faceSide, 6 * K, C.c_int32, (K, 6)),
                fElem=arr(d.fElem, 2 * F, C.c_int32, (F, 2)), fESide=arr(d.fESide, 2 * F, C.c_int32, (F, 2)),
                fRot=arr(d.fRot, F, C.c_int32, (F,)), Ja=arr(d.Ja, 9 * nodes, C.c_double, (nodes, 9)),
                J=arr(d.J, nodes, C.c_double, (nodes,)), normal=arr(d.fNormal, 3 * fnodes, C.c_double, (fnodes, 3)),
                t1=arr(d.fT1, 3 * fnodes, C.c_double, (fnodes, 3)), t2=arr(d.fT2, 3 * fnodes, C.c_double, (fnodes, 3)),
                fjac=arr(d.fJac, fnodes, C.c_double, (fnodes,)),
                x=arr(L.hx3d_mesh_coords(h), 3 * nodes, C.c_double, (nodes, 3)))
    ne = np.prod(Nk + 1, axis=1)
    case["off"] = np.concatenate([[0], np.cumsum(ne)]).astype(np.int64)
    return case


_NAMES = lambda table: {v: k for k, v in table.items() if k not in (None, "lxf")}


class Control:
    """A HORSES3D .control file read by the C reader (hx3d_control_read): key values with the
    reference's parsing and defaults, and the boundary conditions of named zones.  Host only."""

    def __init__(self, path):
        L = load_library()
        h = C.c_void_p()
        rc = L.hx3d_control_read(os.fsencode(path), C.byref(h))
        self.h = h
        if rc != 0:
            msg = L.hx3d_control_error(h).decode() if h else "out of memory"
            self.close()
            raise HX3DError("hx3d_control_read: %s (code %d)" % (msg, rc))

    def close(self):
        if getattr(self, "h", None):
            load_library().hx3d_control_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __getitem__(self, key):
        v = load_library().hx3d_control_value(self.h, key.encode())
        if v is None:
            raise KeyError(key)
        return v.decode()

    def get(self, key, default=None):
        try:
            return self[key]
        except KeyError:
            return default

    def boundary(self, zone):
        """(type, coupled zone or "") of a mesh zone from the "#define boundary" blocks."""
        L = load_library()
        t, c = C.create_string_buffer(256), C.create_string_buffer(256)
        rc = L.hx3d_control_boundary(self.h, zone.encode(), t, c, 256)
        if rc != 0:
            raise HX3DError("hx3d_control_boundary: %s (code %d)" % (L.hx3d_control_error(self.h).decode(), rc))
        return t.value.decode(), c.value.decode()

    def case(self, root=None):
        """(case dict, engine keyword arguments, run settings) of the file (hx3d_control_case).
        Relative file names are taken from root, else from the working directory."""
        L = load_library()
        m, d, r = C.c_void_p(), CaseDesc(), ControlRun()
        rc = L.hx3d_control_case(self.h, os.fsencode(root) if root else None, C.byref(m), C.byref(d), C.byref(r))
        try:
            if rc != 0:
                raise HX3DError("hx3d_control_case: %s (code %d)" % (L.hx3d_control_error(self.h).decode(), rc))
            case = _case_from_desc(L, m, d)
        finally:
            if m:
                L.hx3d_mesh_destroy(m)
        physics = dict(mach=d.mach, reynolds=d.reynolds, prandtl=d.prandtl, euler=bool(d.euler),
                       riemann=_NAMES(RIEMANN)[d.riemann], inviscid=_NAMES(INVISCID)[d.inviscid],
                       averaging=_NAMES(AVERAGING)[d.averaging], lambda_stab=d.lambda_stab,
                       les=_NAMES(LES)[d.les], les_c=d.les_c, nodes=_NAMES(NODES)[d.nodes])
        run = dict(cfl=r.cfl, dcfl=r.dcfl, dt=r.dt, tolerance=r.tolerance, steps=r.steps,
                   output_interval=r.output_interval, restart=bool(r.restart))
        return case, physics, run


def device_memory():
    """(free, total) bytes of the current HIP device."""
    f, t = C.c_size_t(), C.c_size_t()
    if load_library().hx3d_device_memory(C.byref(f), C.byref(t)) != 0:
        raise HX3DError("hx3d_device_memory failed")
    return int(f.value), int(t.value)


def read_control(path, root=None):
    """(case dict, engine keyword arguments, run settings) of a .control file and the mesh and
    order files it names: Engine(case, **physics) sets up the case it describes."""
    c = Control(path)
    try:
        return c.case(root)
    finally:
        c.close()


def _desc(case, keep, mach, reynolds, prandtl=0.72, riemann="roe", inviscid="standard", averaging="standard",
          les=None, les_c=None, lambda_stab=1.0, euler=False, nodes="gauss-lobatto"):
    """A CaseDesc of a case dict; the arrays it points to are kept in `keep`."""
    keep.update(
        N=np.ascontiguousarray(case["N"], dtype=np.int32),
        faceID=np.ascontiguousarray(case["faceID"], dtype=np.int32),
        faceSide=np.ascontiguousarray(case["faceSide"], dtype=np.int32),
        fElem=np.ascontiguousarray(case["fElem"], dtype=np.int32),
        fESide=np.ascontiguousarray(case["fESide"], dtype=np.int32),
        fRot=np.ascontiguousarray(case["fRot"], dtype=np.int32),
        Ja=np.ascontiguousarray(case["Ja"], dtype=np.float64),
        J=np.ascontiguousarray(case["J"], dtype=np.float64))
    face_geo = all(x in case for x in ("normal", "t1", "t2", "fjac"))
    if face_geo:
        for x in ("normal", "t1", "t2", "fjac"):
            keep[x] = np.ascontiguousarray(case[x], dtype=np.float64)
    k = keep
    fg = (lambda x: _p(k[x])) if face_geo else (lambda x: None)
    les_id = LES[les]
    return CaseDesc(int(case["K"]), _p(k["N"]), _p(k["faceID"]), _p(k["faceSide"]), int(case["F"]),
                    _p(k["fElem"]), _p(k["fESide"]), _p(k["fRot"]), _p(k["Ja"]), _p(k["J"]),
                    fg("normal"), fg("t1"), fg("t2"), fg("fjac"), mach, reynolds, prandtl, 1 if euler else 0,
                    RIEMANN[riemann], INVISCID[inviscid], AVERAGING[averaging], lambda_stab, les_id,
                    LES_DEFAULT_C[les_id] if les_c is None else les_c, NODES[nodes.lower()])


class Engine:
    """Device-resident DGSEM case.  State arrays on the host are (nodes, 5) float64."""

    def __init__(self, case, mach, reynolds, prandtl=0.72, riemann="roe", inviscid="standard",
                 averaging="standard", les=None, les_c=None, lambda_stab=1.0, euler=False,
                 nodes="gauss-lobatto"):
        L = load_library()
        self._keep = {}
        d = _desc(case, self._keep, mach, reynolds, prandtl, riemann, inviscid, averaging, les, les_c,
                  lambda_stab, euler, nodes)
        h = C.c_void_p()
        rc = L.hx3d_engine_create(C.byref(d), C.byref(h))
        self._created(L, h, rc, "hx3d_engine_create")

    def _created(self, L, h, rc, what):
        self.h = h
        if rc != 0:
            msg = L.hx3d_last_error(h).decode() if h else "create failed"
            L.hx3d_engine_destroy(h)
            self.h = None
            raise _error("%s: %s (code %d)" % (what, msg, rc), rc)
        self.nodes = int(L.hx3d_num_nodes(self.h))
        self.t = 0.0


    def _chk(self, rc, what):
        if rc != 0:
            raise _error("%s: %s (code %d)" % (what, load_library().hx3d_last_error(self.h).decode(), rc), rc)

    def close(self):
        if getattr(self, "h", None):
            load_library().hx3d_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        assert Q.shape == (self.nodes, 5), Q.shape
        self._chk(load_library().hx3d_set_state(self.h, _p(Q)), "hx3d_set_state")

    def get_state(self):
        out = np.empty((self.nodes, 5))
        self._chk(load_library().hx3d_get_state(self.h, _p(out)), "hx3d_get_state")
        return out

    def qdot(self, t=0.0):
        out = np.empty((self.nodes, 5))
        self._chk(load_library().hx3d_compute_qdot(self.h, t, _p(out)), "hx3d_compute_qdot")
        return out

    def rk3_step(self, t, dt):
        self._chk(load_library().hx3d_rk3_step(self.h, t, dt), "hx3d_rk3_step")

    def set_ctd_after_steps(self, on):
        """CTD_AFTER_STEPS: recompute the time derivative at the state each step leaves (on), or
        keep the last stage's (off, the reference's default)."""
        self._chk(load_library().hx3d_set_ctd_after_steps(self.h, 1 if on else 0), "hx3d_set_ctd_after_steps")

    def get_qdot(self):
        """The current time derivative (after qdot() or a step), (nodes, 5)."""
        out = np.empty((self.nodes, 5))
        self._chk(load_library().hx3d_get_qdot(self.h, _p(out)), "hx3d_get_qdot")
        return out

    def max_dt(self, cfl, dcfl):
        dt = C.c_double()
        self._chk(load_library().hx3d_max_dt(self.h, cfl, dcfl, C.byref(dt)), "hx3d_max_dt")
        return dt.value

    def advance(self, nsteps, cfl, dcfl, t=None):
        tt = C.c_double(self.t if t is None else t)
        self._chk(load_library().hx3d_advance(self.h, nsteps, cfl, dcfl, C.byref(tt)), "hx3d_advance")
        self.t = tt.value
        return self.t

    def advance_dt(self, nsteps, dt, t=None):
        """nsteps RK3 steps of the fixed size dt; returns the time reached."""
        tt = C.c_double(self.t if t is None else t)
        self._chk(load_library().hx3d_advance_dt(self.h, nsteps, dt, C.byref(tt)), "hx3d_advance_dt")
        self.t = tt.value
        return self.t

    def advance_async(self, nsteps, cfl, dcfl):
        """Queues nsteps steps on the engine's stream and returns at once (poll with done())."""
        self._chk(load_library().hx3d_advance(self.h, nsteps, cfl, dcfl, None), "hx3d_advance")

    def done(self):
        d = C.c_int()
        self._chk(load_library().hx3d_query(self.h, C.byref(d)), "hx3d_query")
        return bool(d.value)

    def time(self):
        t = C.c_double()
        self._chk(load_library().hx3d_get_time(self.h, C.byref(t)), "hx3d_get_time")
        self.t = t.value
        return self.t

    def device_bytes(self):
        return int(load_library().hx3d_engine_device_bytes(self.h))

    def residuals(self):
        r = np.zeros(5)
        self._chk(load_library().hx3d_residuals(self.h, _p(r)), "hx3d_residuals")
        return r

    def kinetic_energy(self):
        ke = C.c_double()
        self._chk(load_library().hx3d_kinetic_energy(self.h, C.byref(ke)), "hx3d_kinetic_energy")
        return ke.value

    def volume_monitor(self, kind):
        """A volume monitor of the resident state (VolumeMonitor.f90): "kinetic energy",
        "kinetic energy rate", "enstrophy", "mean velocity", "internal energy", "entropy" or
        "math entropy"."""
        L = load_library()
        v = C.c_double()
        self._chk(L.hx3d_volume_monitor(self.h, MONITORS[kind.lower()], C.byref(v)), "hx3d_volume_monitor")
        return v.value

    def has_nan(self):
        f = C.c_int()
        self._chk(load_library().hx3d_check_nan(self.h, C.byref(f)), "hx3d_check_nan")
        return bool(f.value)

    def save_solution(self, path, iteration=0, time=None, header_name=None):
        """HexMesh_SaveSolution of the resident state (a .hsol solution file)."""
        L = load_library()
        self._chk(L.hx3d_save_solution(self.h, os.fsencode(path), None if header_name is None else header_name.encode(),
                                       int(iteration), self.t if time is None else float(time)), "hx3d_save_solution")

    def load_solution(self, path):
        """LoadSolutionForRestart: the state of a .hsol file; returns (iteration, time)."""
        L = load_library()
        it, tm = C.c_int(), C.c_double()
        self._chk(L.hx3d_load_solution(self.h, os.fsencode(path), C.byref(it), C.byref(tm)), "hx3d_load_solution")
        self.t = tm.value
        return it.value, tm.value

    def synchronize(self):
        self._chk(load_library().hx3d_synchronize(self.h), "hx3d_synchronize")

    def set_step_graph(self, on):
        """advance() replays a captured HIP graph per step (on) or issues the launches (default)."""
        self._chk(load_library().hx3d_set_step_graph(self.h, 1 if on else 0), "hx3d_set_step_graph")

    def autotune(self, reps=5):
        """Mixed orders: times each element tiling on the resident state and keeps the fastest;
        returns (policy, ms per evaluation of each policy), or (-1, None) for one order."""
        ch = C.c_int()
        ms = np.zeros(3)
        self._chk(load_library().hx3d_autotune(self.h, int(reps), C.byref(ch), _p(ms)), "hx3d_autotune")
        return ch.value, (ms if ch.value >= 0 else None)

    def set_tiling(self, policy, sizes=0):
        """Mixed orders: element tiling = packing policy (0 first-fit decreasing, 1 order-triple
        groups, 2 one element per tile) and tile sizes (0 256-thread tiles when every element
        fits, else 512; 1 512-thread tiles; 2 both, each element in the smaller that fits)."""
        self._chk(load_library().hx3d_set_tiling(self.h, int(policy) + 4 * int(sizes)), "hx3d_set_tiling")

    def tiling(self):
        pol, nt = C.c_int(), C.c_int()
        self._chk(load_library().hx3d_get_tiling(self.h, C.byref(pol), C.byref(nt)), "hx3d_get_tiling")
        return pol.value, nt.value

    def set_timing(self, on):
        self._chk(load_library().hx3d_set_timing(self.h, 1 if on else 0), "hx3d_set_timing")

    def kernel_times(self):
        out = np.zeros(4)
        self._chk(load_library().hx3d_kernel_times(self.h, _p(out)), "hx3d_kernel_times")
        return out


class MultiEngine(Engine):
    """Several independent cases of one order in one engine (hx3d_engine_create_multi): each
    case keeps its own physics constants (Mach, Reynolds and Prandtl numbers, LES constant), its
    own dt (MaxTimeStep of its own state, every step), time and divergence flag, and can be
    switched off when it has taken its steps; the discretisation is shared.  Every kernel launch
    carries all the active cases.  The state arrays are the cases' states one after the other
    (`slices`)."""

    def __init__(self, cases, physics):
        L = load_library()
        if isinstance(physics, dict):
            physics = [physics] * len(cases)
        self._keeps = [{} for _ in cases]
        descs = (CaseDesc * len(cases))(*[_desc(c, k, **ph) for c, k, ph in zip(cases, self._keeps, physics)])
        self._descs = descs
        h = C.c_void_p()
        rc = L.hx3d_engine_create_multi(descs, len(cases), C.byref(h))
        self._created(L, h, rc, "hx3d_engine_create_multi")
        self.ncases = len(cases)
        n = [int(c["off"][-1]) if "off" in c else len(c["J"]) for c in cases]
        o = np.concatenate([[0], np.cumsum(n)])
        self.slices = [slice(int(a), int(b)) for a, b in zip(o[:-1], o[1:])]

    def set_states(self, states):
        Q = np.empty((self.nodes, 5))
        for s, q in zip(self.slices, states):
            Q[s] = q
        self.set_state(Q)

    def get_states(self):
        Q = self.get_state()
        return [Q[s].copy() for s in self.slices]

    def get_case_state(self, c):
        """Case c's state alone (waits for the queued work)."""
        s = self.slices[c]
        out = np.empty((s.stop - s.start, 5))
        self._chk(load_library().hx3d_get_state_nodes(self.h, s.start, s.stop - s.start, _p(out)), "hx3d_get_state_nodes")
        return out

    def set_active(self, c, on):
        """Case c takes part in the following steps (on) or is left as it is (off)."""
        self._chk(load_library().hx3d_multi_set_active(self.h, int(c), 1 if on else 0), "hx3d_multi_set_active")

    def status(self):
        """(times, diverged flags) of the cases once the queued work is done."""
        t = np.zeros(self.ncases)
        f = np.zeros(self.ncases, dtype=np.int32)
        self._chk(load_library().hx3d_multi_status(self.h, _p(t), _p(f)), "hx3d_multi_status")
        return t, f.astype(bool)

    def advance(self, nsteps, cfl, dcfl, t=None):
        """Queues nsteps steps of every active case, each with its own CFL dt; returns at once."""
        assert t is None, "several cases: read their times with status()"
        self._chk(load_library().hx3d_advance(self.h, nsteps, cfl, dcfl, None), "hx3d_advance")
