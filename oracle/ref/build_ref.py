#!/usr/bin/env python3
"""Build the HORSES3D compressible Navier-Stokes solver from its own Fortran sources.

This is test infrastructure: the resulting executable is the reference oracle used to
write golden vectors (tests/golden/) and, optionally, to time the reference OpenMP CPU
build.  Nothing on the product path links or runs it.

Recipe
------
* Sources are read in place from the reference tree (``--ref``, default
  ``/root/reference``); nothing is copied.  Outputs go only to ``oracle/_ref/``.
* The compile order of each library is the object list its own Makefile declares
  (``LIBOBJS_NS`` for the physics-specialised libraries, ``LIBOBJS`` otherwise;
  Solver/Makefile.in:227-231 gives the library order).  The Makefiles are read as data,
  they are not executed.
* Compiler: ROCm's ``amdflang`` (flang-22); gfortran/ifort are absent from the image.
  Flags follow the reference's RELEASE set (Solver/Makefile.in:42-43) minus the
  gfortran-only options: ``-cpp -O2 -fopenmp -DNAVIERSTOKES -D_has_Quad``.  This -O2
  build (``bin/horses3d.ns``) wrote the golden vectors.  ``--opt=-O3 --tag O3`` builds the
  timing variant ``bin/horses3d_O3.ns`` (objects under ``O3/``) that the CPU baseline runs.
* The per-case user plugin (the reference's ProblemFile API, libs/problemfile/
  ProblemFile.f90:22-237) is our own ``oracle/ref/ProblemFile.f90``; it sets the
  Taylor-Green initial condition and dumps state, time derivative and geometry.
* ``WallDistance.f90:203`` calls an external ``HessianEstimation_DFP`` that the
  reference never defines (only ``HessianEstimation_DFP_DFP`` exists, :243).  It is
  reachable only from wall-distance optimisation, which periodic boxes never run.  The
  link therefore uses ``--unresolved-symbols=ignore-in-object-files`` (the same lazy
  binding the reference's own shared-library build relies on); no stand-in code is
  written for it.
"""
import argparse
import os
import re
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(REPO, "oracle", "_ref")

# Library build order for the NS solver (Solver/Makefile.in:227-231), physics-specialised
# libraries built in their NS flavour.
LIBS = [
    ("ftobject", "LIBOBJS"),
    ("foundation", "LIBOBJS"),
    ("mpiutils", "LIBOBJS"),
    ("ml", "LIBOBJS"),
    ("spectral", "LIBOBJS"),
    ("io", "LIBOBJS"),
    ("physics/navierstokes", "LIBOBJS"),
    ("physics/common", "LIBOBJS_NS"),
    ("mesh", "LIBOBJS_NS"),
    ("particles", "LIBOBJS_NS"),
    ("sources", "LIBOBJS_NS"),
    ("monitors", "LIBOBJS_NS"),
    ("problemfile", None),
    ("discretization", "LIBOBJS_NS"),
    ("jacobian", "LIBOBJS_NS"),
    ("timeintegrator", "LIBOBJS_NS"),
]
# Link order (HORSES_LIBS_NS, Solver/Makefile.in:231)
LINK_ORDER = ["timeintegrator", "jacobian", "discretization", "monitors", "particles", "mesh",
              "physics/common", "physics/navierstokes", "io", "spectral", "mpiutils",
              "foundation", "ftobject", "sources", "ml"]

FC = "amdflang"


def makefile_objects(libdir, var):
    text = open(os.path.join(libdir, "Makefile")).read()
    text = text.replace("\\\n", " ")
    m = re.search(r"^%s\s*=(.*)$" % re.escape(var), text, re.M)
    if not m:
        raise RuntimeError("no %s in %s/Makefile" % (var, libdir))
    return [os.path.splitext(os.path.basename(t))[0] for t in m.group(1).split() if t.endswith(".o")]


def find_source(libdir, name):
    p = os.path.join(libdir, name + ".f90")
    if os.path.exists(p):
        return p
    for root, _, files in os.walk(libdir):
        if name + ".f90" in files:
            return os.path.join(root, name + ".f90")
    raise FileNotFoundError(name + ".f90 under " + libdir)


BUILD = OUT   # objects, modules and archives of the variant being built


def fflags(ref, opt):
    return ["-cpp", opt, "-fopenmp", "-DNAVIERSTOKES", "-D_has_Quad",
            "-I", os.path.join(ref, "Solver", "src", "libs", "foundation"),
            "-module-dir", os.path.join(BUILD, "mod"), "-I", os.path.join(BUILD, "mod")]


def compile_one(ref, src, obj, opt):
    cmd = [FC] + fflags(ref, opt) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return r.returncode, r.stdout + r.stderr


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--opt", default="-O2")
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--tag", default="", help="variant name: objects under oracle/_ref/<tag>/, "
                                             "executable oracle/_ref/bin/horses3d_<tag>.ns")
    args = ap.parse_args()
    global BUILD
    BUILD = os.path.join(OUT, args.tag) if args.tag else OUT
    libs_root = os.path.join(args.ref, "Solver", "src", "libs")
    for d in ("obj", "mod", "lib"):
        os.makedirs(os.path.join(BUILD, d), exist_ok=True)
    os.makedirs(os.path.join(OUT, "bin"), exist_ok=True)
    t0 = time.time()
    archives = {}
    for lib, var in LIBS:
        libdir = os.path.join(libs_root, lib)
        if var is None:
            # The default plugin file also declares module ProblemFileFunctions (the plugin
            # interfaces the solver libraries use).  Compile it for that module only; its
            # default routines are not linked (our own plugin supplies them).
            src = os.path.join(libdir, "ProblemFile.f90")
            obj = os.path.join(BUILD, "obj", "problemfile_interfaces.o")
            rc, log = compile_one(args.ref, src, obj, args.opt)
            if rc != 0:
                sys.stderr.write(log)
                raise SystemExit("compile failed: %s" % src)
            continue
        names = makefile_objects(libdir, var)
        objdir = os.path.join(BUILD, "obj", lib.replace("/", "_"))
        os.makedirs(objdir, exist_ok=True)
        objs, pending = [], []
        for n in names:
            src = find_source(libdir, n)
            obj = os.path.join(objdir, n + ".o")
            objs.append(obj)
            if (args.force or not os.path.exists(obj)
                    or os.path.getmtime(obj) < os.path.getmtime(src)):
                pending.append((src, obj))
        # The object lists are not always in module-dependency order (the Makefiles add
        # make.deps rules), so compile in passes: what fails on a module that is not built
        # yet is retried after the rest of the pass.
        while pending:
            with ThreadPoolExecutor(max_workers=args.jobs) as ex:
                results = list(ex.map(lambda so: compile_one(args.ref, so[0], so[1], args.opt), pending))
            failed = [(so, log) for so, (rc, log) in zip(pending, results) if rc != 0]
            if len(failed) == len(pending):
                for (src, _), log in failed:
                    sys.stderr.write(log)
                raise SystemExit("compile failed in %s" % lib)
            pending = [so for so, _ in failed]
        a = os.path.join(BUILD, "lib", "lib%s.a" % lib.replace("/", "_"))
        if os.path.exists(a):
            os.remove(a)
        subprocess.check_call(["ar", "crs", a] + objs)
        archives[lib] = a
        print("[build_ref] %-22s %3d files  (%.0fs)" % (lib, len(objs), time.time() - t0), flush=True)

    # Solver objects + our ProblemFile plugin
    solver = os.path.join(args.ref, "Solver", "src", "NavierStokesSolver")
    objdir = os.path.join(BUILD, "obj", "solver")
    os.makedirs(objdir, exist_ok=True)
    extra = []
    for src in (os.path.join(HERE, "ProblemFile.f90"),
                os.path.join(solver, "SpatialDiscretization.f90"),
                os.path.join(solver, "main.f90")):
        obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
        rc, log = compile_one(args.ref, src, obj, args.opt)
        if rc != 0:
            sys.stderr.write(log)
            raise SystemExit("compile failed: %s" % src)
        extra.append(obj)
    exe = os.path.join(OUT, "bin", "horses3d%s.ns" % ("_" + args.tag if args.tag else ""))
    link = ([FC, "-fopenmp", "-o", exe] + extra
            + ["-Wl,--start-group"] + [archives[l] for l in LINK_ORDER] + ["-Wl,--end-group"]
            + ["-Wl,--unresolved-symbols=ignore-in-object-files", "-ldl"])
    subprocess.check_call(link)
    print("[build_ref] linked %s (%.0fs)" % (exe, time.time() - t0), flush=True)


if __name__ == "__main__":
    main()
