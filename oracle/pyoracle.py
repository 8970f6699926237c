"""ctypes binding of the C oracle (oracle/libhx3d_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg, never by the product package.

A *case* is a dict of numpy arrays in the element/face storage contract of the reference's
HexMesh (see oracle/hx3d_oracle.h); ``from_reference_dump`` builds one from a dump written by
oracle/ref/ProblemFile.f90.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhx3d_oracle.so")

RIEMANN = {"roe": 0, "lax-friedrichs": 1, "lxf": 1}
INVISCID = {"standard": 0, "split-form": 1}
AVERAGING = {"standard": 0, "ducros": 1, "kennedy-gruber": 2, "pirozzoli": 3, "morinishi": 4}
LES = {None: 0, "none": 0, "smagorinsky": 1, "wale": 2}
NODES = {"gauss-lobatto": 0, "gauss": 1}


class Params(C.Structure):
    _fields_ = [("mach", C.c_double), ("reynolds", C.c_double), ("prandtl", C.c_double),
                ("riemann", C.c_int), ("inviscid", C.c_int), ("averaging", C.c_int),
                ("les", C.c_int), ("lambda_stab", C.c_double), ("les_c", C.c_double),
                ("euler", C.c_int), ("nodes", C.c_int)]


def make_params(mach, reynolds, prandtl=0.72, riemann="roe", inviscid="standard",
                averaging="standard", les=None, lambda_stab=1.0, les_c=None, euler=False,
                nodes="gauss-lobatto"):
    if les_c is None:
        les_c = {0: 0.0, 1: 0.2, 2: 0.325}[LES[les]]
    return Params(mach, reynolds, prandtl, RIEMANN[riemann], INVISCID[inviscid],
                  AVERAGING[averaging], LES[les], lambda_stab, les_c, 1 if euler else 0,
                  NODES[nodes])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: %s (run __graft_entry__.build())" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.ox_case_create.restype = P
        L.ox_case_create.argtypes = [C.c_int, P, P, P, P, P, P, C.c_int, P, P, P, P, P, P, P, P, P,
                                     C.POINTER(Params)]
        L.ox_case_destroy.argtypes = [P]
        L.ox_compute_qdot.argtypes = [P, P, C.c_double, P, P]
        L.ox_rk3_step.argtypes = [P, P, P, C.c_double, C.c_double]
        L.ox_max_dt.restype = C.c_double
        L.ox_max_dt.argtypes = [P, P, C.c_double, C.c_double]
        L.ox_kinetic_energy.restype = C.c_double
        L.ox_kinetic_energy.argtypes = [P, P]
        L.ox_nodal_gl.argtypes = [C.c_int] + [P] * 9
        L.ox_nodal_gauss.argtypes = [C.c_int] + [P] * 8
        L.ox_nodal_cgl.argtypes = [C.c_int, P, P, P]
        L.ox_tset.argtypes = [C.c_int, C.c_int, P]
        L.ox_tset_nodes.argtypes = [C.c_int, C.c_int, C.c_int, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def nodal_gl(N):
    n = N + 1
    x, w = np.zeros(n), np.zeros(n)
    D, hatD, sharpD = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    vL, vR, bL, bR = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    lib().ox_nodal_gl(N, *[_p(a) for a in (x, w, D, hatD, sharpD, vL, vR, bL, bR)])
    return dict(x=x, w=w, D=D, hatD=hatD, sharpD=sharpD, vL=vL, vR=vR, bL=bL, bR=bR)


def nodal_gauss(N):
    n = N + 1
    x, w = np.zeros(n), np.zeros(n)
    D, hatD = np.zeros((n, n)), np.zeros((n, n))
    vL, vR, bL, bR = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    lib().ox_nodal_gauss(N, *[_p(a) for a in (x, w, D, hatD, vL, vR, bL, bR)])
    return dict(x=x, w=w, D=D, hatD=hatD, vL=vL, vR=vR, bL=bL, bR=bR)


def nodal_cgl(N):
    n = N + 1
    x, D, T = np.zeros(n), np.zeros((n, n)), np.zeros((n, n))
    lib().ox_nodal_cgl(N, _p(x), _p(D), _p(T))
    return dict(xcgl=x, DCGL=D, Tcheb2gauss=T)


def tset(N, M, nodes="gauss-lobatto"):
    T = np.zeros((M + 1, N + 1))
    lib().ox_tset_nodes(NODES[nodes], N, M, _p(T))
    return T


def from_reference_dump(d):
    """Flatten a parsed reference dump (tests/golden/refcase.read_dump) into a case dict."""
    el, fa = d["elements"], d["faces"]
    K, F = len(el), len(fa)
    N = np.array([e["N"] for e in el], dtype=np.int32)
    ne = np.prod(N + 1, axis=1)
    off = np.concatenate([[0], np.cumsum(ne)]).astype(np.int64)
    flat = lambda a, nv: a.reshape(nv, -1, order="F").T
    Ja = np.concatenate([np.concatenate([flat(e["JaXi"], 3), flat(e["JaEta"], 3), flat(e["JaZeta"], 3)], axis=1)
                         for e in el])
    J = np.concatenate([e["J"].reshape(-1, order="F") for e in el])
    Q = np.concatenate([flat(e["Q"], 5) for e in el])
    QDot = np.concatenate([flat(e["QDot"], 5) for e in el])
    X = np.concatenate([flat(e["x"], 3) for e in el])
    faceID = np.array([e["faceIDs"] for e in el], dtype=np.int32) - 1
    faceSide = np.array([e["faceSide"] for e in el], dtype=np.int32)
    fElem = np.array([f["elementIDs"] for f in fa], dtype=np.int32) - 1
    fESide = np.array([f["elementSide"] for f in fa], dtype=np.int32)
    fRot = np.array([f["rotation"] for f in fa], dtype=np.int32)
    fNf = np.array([f["Nf"] for f in fa], dtype=np.int32)
    fn = np.prod(fNf + 1, axis=1)
    foff = np.concatenate([[0], np.cumsum(fn)]).astype(np.int64)
    normal = np.concatenate([flat(f["normal"], 3) for f in fa])
    t1 = np.concatenate([flat(f["t1"], 3) for f in fa])
    t2 = np.concatenate([flat(f["t2"], 3) for f in fa])
    fjac = np.concatenate([f["jac"].reshape(-1, order="F") for f in fa])
    fx = np.concatenate([flat(f["x"], 3) for f in fa])
    return dict(K=K, N=N, off=off, faceID=faceID, faceSide=faceSide, Ja=np.ascontiguousarray(Ja),
                J=J, F=F, fElem=fElem, fESide=fESide, fRot=fRot, fNf=fNf, foff=foff,
                normal=np.ascontiguousarray(normal), t1=np.ascontiguousarray(t1),
                t2=np.ascontiguousarray(t2), fjac=fjac, x=np.ascontiguousarray(X),
                fx=np.ascontiguousarray(fx), Q=np.ascontiguousarray(Q),
                QDot=np.ascontiguousarray(QDot), time=d["time"])


class Oracle:
    """One oracle case (mesh + geometry + physics); state arrays are (nodes, 5) float64."""

    def __init__(self, case, params):
        self.case = case
        self.params = params
        c = {k: np.ascontiguousarray(case[k]) for k in
             ("N", "off", "faceID", "faceSide", "Ja", "J", "fElem", "fESide", "fRot", "fNf", "foff",
              "normal", "t1", "t2", "fjac")}
        c["N"] = c["N"].astype(np.int32)
        for k in ("faceID", "faceSide", "fElem", "fESide", "fRot", "fNf"):
            c[k] = c[k].astype(np.int32)
        for k in ("off", "foff"):
            c[k] = c[k].astype(np.int64)
        self._keep = c
        self.nodes = int(c["off"][-1])
        self.h = lib().ox_case_create(int(case["K"]), _p(c["N"]), _p(c["off"]), _p(c["faceID"]),
                                      _p(c["faceSide"]), _p(c["Ja"]), _p(c["J"]), int(case["F"]),
                                      _p(c["fElem"]), _p(c["fESide"]), _p(c["fRot"]), _p(c["fNf"]),
                                      _p(c["foff"]), _p(c["normal"]), _p(c["t1"]), _p(c["t2"]),
                                      _p(c["fjac"]), C.byref(params))
        if not self.h:
            raise ValueError("oracle: unsupported case (split forms run on Gauss-Lobatto nodes)")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ox_case_destroy(self.h)
            self.h = None

    def qdot(self, Q, t=0.0, want_grad=False):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        out = np.zeros((self.nodes, 5))
        grad = np.zeros((self.nodes, 15)) if want_grad else None
        lib().ox_compute_qdot(self.h, _p(Q), t, _p(out), _p(grad) if want_grad else None)
        return (out, grad) if want_grad else out

    def rk3_step(self, Q, G, t, dt):
        assert Q.flags.c_contiguous and G.flags.c_contiguous
        lib().ox_rk3_step(self.h, _p(Q), _p(G), t, dt)

    def max_dt(self, Q, cfl, dcfl):
        return lib().ox_max_dt(self.h, _p(np.ascontiguousarray(Q)), cfl, dcfl)

    def kinetic_energy(self, Q):
        return lib().ox_kinetic_energy(self.h, _p(np.ascontiguousarray(Q)))


def normwise_error(a, b):
    """Per conserved variable: max_i |a-b| / max_i |b| (normwise relative error)."""
    a = np.asarray(a).reshape(-1, 5)
    b = np.asarray(b).reshape(-1, 5)
    return np.abs(a - b).max(axis=0) / np.maximum(np.abs(b).max(axis=0), 1e-300)
