"""Solution files (.hsol) against files the reference solver wrote.

tests/golden/hsol holds .hsol files from three reference runs (a curved uniform-order box, a
box with per-element orders, and a curved box on Gauss nodes), made by make_hsol.py, with the state the reference's ProblemFile
dump held at the same instant.  Reading them must give that state bit for bit, and writing the
same state with the same header must give the same bytes.  Host only (ctypes, no GPU)."""
import ctypes as C
import os

import numpy as np
import pytest

from horses3d_amd import engine as E

HS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hsol")
CASES = ["box2_p3_curved", "box2_mixed_p2to4", "box2_p3_curved_gauss"]
NODE_TYPE = {"box2_p3_curved_gauss": 1}   # the file's node type field: 1 Gauss, 2 Gauss-Lobatto


def lib():
    L = E.load_library()
    return L


def header(path):
    L = lib()
    t, nt, k, it = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    tm = C.c_double()
    refs = (C.c_double * 6)()
    rc = L.hx3d_hsol_read_header(os.fsencode(path), C.byref(t), C.byref(nt), C.byref(k), C.byref(it), C.byref(tm), refs)
    assert rc == 0, L.hx3d_hsol_error()
    return t.value, nt.value, k.value, it.value, tm.value, np.array(refs[:])


def read(path, N, grad=False):
    L = lib()
    N = np.ascontiguousarray(N, dtype=np.int32)
    nodes = int(np.prod(N + 1, axis=1).sum())
    Q = np.zeros((nodes, 5))
    rc = L.hx3d_hsol_read(os.fsencode(path), len(N), N.ctypes.data_as(C.c_void_p), Q.ctypes.data_as(C.c_void_p), None)
    return rc, Q


@pytest.mark.parametrize("name", CASES)
def test_read_reference_files(name):
    z = np.load(os.path.join(HS, name + ".npz"), allow_pickle=False)
    L = lib()
    ref = (C.c_double * 6)()
    assert L.hx3d_hsol_reference_values(C.c_double(float(z["mach"])), ref) == 0
    for tag, it in (("step0", 0), ("final", int(z["steps"]))):
        p = os.path.join(HS, "%s_%s.hsol" % (name, tag))
        t, nt, k, iters, tm, refs = header(p)
        assert (t, nt, k, iters) == (2, NODE_TYPE.get(name, 2), len(z["N"]), it)
        assert tm == float(z["t_" + tag])
        # the default reference state of the physics setup, bit for bit
        assert np.array_equal(refs, np.array(ref[:]))
        rc, Q = read(p, z["N"])
        assert rc == 0, L.hx3d_hsol_error()
        assert np.array_equal(Q, z["Q_" + tag])


@pytest.mark.parametrize("name", CASES)
def test_write_reproduces_reference_bytes(name, tmp_path):
    z = np.load(os.path.join(HS, name + ".npz"), allow_pickle=False)
    L = lib()
    src = os.path.join(HS, name + "_final.hsol")
    t, nt, k, it, tm, refs = header(src)
    N = np.ascontiguousarray(z["N"], dtype=np.int32)
    Q = np.ascontiguousarray(z["Q_final"])
    out = os.path.join(str(tmp_path), "w.hsol")
    R = (C.c_double * 6)(*refs)
    # the reference stores the name it was given: RESULTS/case.hsol
    nodes = E.NODES["gauss" if nt == 1 else "gauss-lobatto"]
    rc = L.hx3d_hsol_write_nodes(os.fsencode(out), b"RESULTS/case.hsol", len(N), N.ctypes.data_as(C.c_void_p), it,
                                 C.c_double(tm), R, Q.ctypes.data_as(C.c_void_p), None, nodes)
    assert rc == 0, L.hx3d_hsol_error()
    assert open(out, "rb").read() == open(src, "rb").read()


def test_gradient_files_and_errors(tmp_path):
    L = lib()
    z = np.load(os.path.join(HS, "box2_mixed_p2to4.npz"), allow_pickle=False)
    N = np.ascontiguousarray(z["N"], dtype=np.int32)
    Q = np.ascontiguousarray(z["Q_final"])
    nodes = len(Q)
    G = np.random.default_rng(1).standard_normal((3 * nodes, 5))
    p = os.path.join(str(tmp_path), "g.hsol")
    R = (C.c_double * 6)()
    L.hx3d_hsol_reference_values(C.c_double(0.1), R)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    assert L.hx3d_hsol_write(os.fsencode(p), None, len(N), vp(N), 7, C.c_double(0.5), R, vp(Q), vp(G)) == 0
    assert header(p)[:4] == (3, 2, len(N), 7)
    Q2, G2 = np.zeros_like(Q), np.zeros_like(G)
    assert L.hx3d_hsol_read(os.fsencode(p), len(N), vp(N), vp(Q2), vp(G2)) == 0
    assert np.array_equal(Q2, Q) and np.array_equal(G2, G)
    # a type 3 file read for the state alone skips the gradients
    assert read(p, N)[0] == 0 and np.array_equal(read(p, N)[1], Q)
    # element count and per-element orders are checked; truncation is reported
    assert read(p, N[:-1])[0] != 0 and b"elements" in L.hx3d_hsol_error()
    bad = N.copy()
    bad[3, 0] += 1
    rc, _ = read(p, bad)
    assert rc != 0 and b"orders" in L.hx3d_hsol_error()
    t = os.path.join(str(tmp_path), "t.hsol")
    open(t, "wb").write(open(p, "rb").read()[:1000])
    assert read(t, N)[0] != 0 and b"truncated" in L.hx3d_hsol_error()
    assert read(os.path.join(str(tmp_path), "none.hsol"), N)[0] != 0
