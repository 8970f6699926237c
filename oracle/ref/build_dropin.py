#!/usr/bin/env python3
"""Build the drop-in driver: the reference solver whose explicit RK3 step runs on the hx3d engine.

Test infrastructure.  It links the reference build that oracle/ref/build_ref.py left in
oracle/_ref/ (its -O2 libraries, modules and solver objects: the same build that wrote the
golden vectors) with
  * horses3d_amd/fortran/hx3d_binding.f90 and hx3d_horses3d.f90 (the drop-in adapter),
  * tests/fortran/hx3d_dropin.f90 (the driver program, in place of the reference's main.f90),
  * horses3d_amd/libhx3d.so (the engine),
into oracle/_ref/bin/hx3d_dropin.ns.  Nothing is copied from the reference tree; its include
directory (Includes.h) is read in place.

    python oracle/ref/build_dropin.py [--ref /root/reference]
"""
import argparse
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(REPO, "oracle", "_ref")
sys.path.insert(0, HERE)
import build_ref as br  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    lib = os.path.join(REPO, "horses3d_amd", "libhx3d.so")
    mod = os.path.join(OUT, "mod")
    solver_objs = [os.path.join(OUT, "obj", "solver", n + ".o") for n in ("ProblemFile", "SpatialDiscretization")]
    archives = [os.path.join(OUT, "lib", "lib%s.a" % l.replace("/", "_")) for l in br.LINK_ORDER]
    for p in [lib] + solver_objs + archives:
        if not os.path.exists(p):
            raise SystemExit("missing %s (build the engine and oracle/ref/build_ref.py first)" % p)
    work = os.path.join(OUT, "dropin")
    os.makedirs(work, exist_ok=True)
    inc = os.path.join(args.ref, "Solver", "src", "libs", "foundation")
    flags = ["-cpp", "-O2", "-fopenmp", "-DNAVIERSTOKES", "-D_has_Quad", "-I", inc, "-I", mod, "-I", work,
             "-module-dir", work]
    objs = []
    for src in (os.path.join(REPO, "horses3d_amd", "fortran", "hx3d_binding.f90"),
                os.path.join(REPO, "horses3d_amd", "fortran", "hx3d_horses3d.f90"),
                os.path.join(REPO, "tests", "fortran", "hx3d_dropin.f90")):
        obj = os.path.join(work, os.path.basename(src)[:-4] + ".o")
        subprocess.check_call([br.FC] + flags + ["-c", src, "-o", obj])
        objs.append(obj)
    os.makedirs(os.path.join(OUT, "bin"), exist_ok=True)
    exe = os.path.join(OUT, "bin", "hx3d_dropin.ns")
    subprocess.check_call([br.FC, "-fopenmp", "-o", exe] + objs + solver_objs
                          + ["-Wl,--start-group"] + archives + ["-Wl,--end-group"]
                          + [lib, "-Wl,-rpath,$ORIGIN/../../../horses3d_amd",
                             "-Wl,--unresolved-symbols=ignore-in-object-files", "-ldl"])
    print("[build_dropin] linked %s" % exe, flush=True)


if __name__ == "__main__":
    main()
