"""Case writer and runner for the reference oracle build (oracle/_ref/bin/horses3d.ns).

Used only to write the golden fixtures in this directory.  It writes a periodic box as a
SpecMesh file (format read by Solver/src/libs/mesh/Read_SpecMesh.f90:36-348), a
``.control`` file (keys from Solver/src/libs/foundation/Setup.f90:1-33 and
RiemannSolvers_NS.f90:1-49), optionally a polynomial-order file
(Solver/src/libs/io/ReadInputFile.f90:132-153), runs the reference, and parses the binary
dumps written by oracle/ref/ProblemFile.f90.
"""
import os
import subprocess
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_EXE = os.path.join(REPO, "oracle", "_ref", "bin", "horses3d.ns")
sys.path.insert(0, REPO)

from horses3d_amd.specmesh import (CORNER_XI, FACE_FIXED, FACE_NAMES_BOX, FACE_UV, L,  # noqa: E402,F401
                                  PERIODIC_PAIRS, box_map, write_orders, write_specmesh)


def write_control(path, mesh, order, steps, mach, re, riemann="roe", inviscid="standard",
                  averaging=None, cfl=0.4, dcfl=0.4, les=None, les_intensity=None,
                  order_file=None, pr=0.72, lambda_stab=None, ctd_after=True, extra=None,
                  monitors=(("KinEn", "Kinetic energy"),), nodes="Gauss-Lobatto"):
    kv = [("Flow equations", '"NS"'),
          ("mesh file name", '"%s"' % mesh),
          ("solution file name", '"RESULTS/case.hsol"'),
          ("restart", ".false."),
          ("restart file name", '"RESULTS/case.rst"'),
          ("Number of time steps", str(steps)),
          ("Output Interval", "1"),
          ("Convergence tolerance", "1.d-30"),
          ("cfl", repr(cfl)),
          ("dcfl", repr(dcfl)),
          ("mach number", repr(mach)),
          ("Reynolds number", repr(re)),
          ("Prandtl number", repr(pr)),
          ("riemann solver", '"%s"' % riemann),
          ("inviscid discretization", '"%s"' % inviscid),
          ("viscous discretization", '"BR1"'),
          ("discretization nodes", '"%s"' % nodes),
          ("compute time derivative after timestep", ".true." if ctd_after else ".false."),
          ("AOA theta", "0.0"),
          ("AOA phi", "0.0")]
    if order_file:
        kv.append(("polynomial order file", '"%s"' % order_file))
    else:
        kv.append(("Polynomial order", str(order)))
    if averaging:
        kv.append(("averaging", '"%s"' % averaging))
    if lambda_stab is not None:
        kv.append(("lambda stabilization", repr(lambda_stab)))
    if les:
        kv.append(("les model", '"%s"' % les))
        if les_intensity is not None:
            kv.append(("les model intensity", repr(les_intensity)))
    for k, v in (extra or []):
        kv.append((k, v))
    out = ["%-40s = %s" % (k, v) for k, v in kv]
    for a, b in PERIODIC_PAIRS:
        out += ["", "#define boundary %s" % a, "   type = periodic",
                "   coupled boundary = %s" % b, "#end"]
    for i, (name, var) in enumerate(monitors):
        out += ["", "#define volume monitor %d" % (i + 1), "   Name = %s" % name, "   Variable = %s" % var, "#end"]
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def run_reference(case_dir, control="case.control", env=None, threads=1, timeout=3600, exe=None):
    os.makedirs(os.path.join(case_dir, "RESULTS"), exist_ok=True)
    os.makedirs(os.path.join(case_dir, "MESH"), exist_ok=True)
    e = dict(os.environ)
    e.update({"OMP_NUM_THREADS": str(threads), "OMP_SCHEDULE": "static"})
    e.update(env or {})
    r = subprocess.run([exe or REF_EXE, control], cwd=case_dir, env=e, capture_output=True,
                       text=True, timeout=timeout)
    with open(os.path.join(case_dir, "ref_stdout.log"), "w") as fh:
        fh.write(r.stdout + r.stderr)
    if r.returncode != 0:
        raise RuntimeError("reference run failed (%d):\n%s" % (r.returncode, (r.stdout + r.stderr)[-4000:]))
    return r.stdout


def read_dump(path):
    """Parse a dump written by oracle/ref/ProblemFile.f90 (hx3d_dump)."""
    b = open(path, "rb").read()
    off = 0

    def ints(n):
        nonlocal off
        v = np.frombuffer(b, dtype="<i4", count=n, offset=off)
        off += 4 * n
        return v.copy()

    def dbl(shape):
        nonlocal off
        cnt = int(np.prod(shape))
        v = np.frombuffer(b, dtype="<f8", count=cnt, offset=off)
        off += 8 * cnt
        return v.reshape(shape, order="F").copy()

    ne, nf = ints(2)
    time = dbl((1,))[0]
    el = []
    for _ in range(ne):
        N = ints(3)
        fid = ints(6)
        fside = ints(6)
        s = tuple(int(x) + 1 for x in N)
        d = dict(N=N, faceIDs=fid, faceSide=fside)
        d["x"] = dbl((3,) + s)
        d["JaXi"] = dbl((3,) + s)
        d["JaEta"] = dbl((3,) + s)
        d["JaZeta"] = dbl((3,) + s)
        d["J"] = dbl(s)
        d["Q"] = dbl((5,) + s)
        d["QDot"] = dbl((5,) + s)
        el.append(d)
    faces = []
    for _ in range(nf):
        h = ints(14)
        Nf = h[0:2]
        s = (int(Nf[0]) + 1, int(Nf[1]) + 1)
        d = dict(Nf=Nf, elementIDs=h[2:4], elementSide=h[4:6], rotation=int(h[6]),
                 faceType=int(h[7]), projectionType=h[8:10], NfLeft=h[10:12], NfRight=h[12:14])
        d["x"] = dbl((3,) + s)
        d["normal"] = dbl((3,) + s)
        d["t1"] = dbl((3,) + s)
        d["t2"] = dbl((3,) + s)
        d["jac"] = dbl(s)
        faces.append(d)
    assert off == len(b), (off, len(b))
    return dict(time=time, elements=el, faces=faces)


def read_dt_log(path):
    rows = [l.split() for l in open(path) if l.strip()]
    return np.array([[float(x) for x in r] for r in rows])


def read_volume_monitor(case_dir, name="kinen"):
    """A volume monitor written by the reference (RESULTS/case.<name>.volume): iteration, time,
    value per row."""
    p = os.path.join(case_dir, "RESULTS", "case.%s.volume" % name.lower())
    rows = []
    for l in open(p):
        t = l.split()
        if len(t) >= 3:
            try:
                rows.append([float(t[0]), float(t[1]), float(t[2])])
            except ValueError:
                pass
    return np.array(rows)
