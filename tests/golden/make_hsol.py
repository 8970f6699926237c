"""Solution-file fixtures: runs the reference build (oracle/_ref/bin/horses3d.ns) on three small
periodic boxes and keeps the .hsol files it writes (HexMesh_SaveSolution, HexMesh.f90:3376-3556)
together with the state, derivative and time its ProblemFile dump holds at the same instant.

    python tests/golden/make_hsol.py [case ...]    # writes tests/golden/hsol/*.hsol and *.npz
"""
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(REPO, "oracle"))
import refcase as rc  # noqa: E402
import pyoracle as po  # noqa: E402

OUT = os.path.join(HERE, "hsol")
CASES = {
    # 2^3 curved box, N = 3, Roe + BR1 standard form, 3 steps
    "box2_p3_curved": dict(n=2, N=3, amp=0.1, seed=3, steps=3, orders=None),
    # 2^3 straight box with per-element orders on {2..4} (element arrays of different sizes)
    "box2_mixed_p2to4": dict(n=2, N=4, amp=0.0, seed=0, steps=2, orders=(2, 4, 21)),
    # 2^3 curved box, N = 3 on Gauss nodes (the file's node type field is 1)
    "box2_p3_curved_gauss": dict(n=2, N=3, amp=0.1, seed=4, steps=2, orders=None, nodes="Gauss"),
}


def run(name, spec, work):
    d = os.path.join(work, name)
    os.makedirs(d)
    ph = np.zeros(9) if not spec["seed"] else np.random.default_rng(spec["seed"]).uniform(0.0, 2.0 * np.pi, 9)
    rc.write_specmesh(os.path.join(d, "box.mesh"), spec["n"], spec["amp"], ph, bface_order=max(4, spec["N"]))
    kw = {}
    if spec["orders"]:
        lo, hi, seed = spec["orders"]
        orders = np.random.default_rng(seed).integers(lo, hi + 1, (spec["n"] ** 3, 3))
        rc.write_orders(os.path.join(d, "orders.txt"), orders)
        kw["order_file"] = "orders.txt"
    if spec.get("nodes"):
        kw["nodes"] = spec["nodes"]
    rc.write_control(os.path.join(d, "case.control"), "box.mesh", spec["N"], spec["steps"], 0.1, 1600.0, **kw)
    rc.run_reference(d, env={"HX3D_DUMP_PREFIX": os.path.join(d, "dump"), "HX3D_DUMP_STEPS": "0"})
    for f, tag in (("case.hsol", "final"), ("case_0000000000.hsol", "step0")):
        shutil.copy(os.path.join(d, "RESULTS", f), os.path.join(OUT, "%s_%s.hsol" % (name, tag)))
    out = {}
    for tag, dump in (("final", "dump_final.bin"), ("step0", "dump_step0.bin")):
        g = po.from_reference_dump(rc.read_dump(os.path.join(d, dump)))
        out["Q_" + tag] = g["Q"]
        out["QDot_" + tag] = g["QDot"]
        out["t_" + tag] = np.float64(g["time"])
        out["N"] = g["N"]
    out["steps"] = np.int64(spec["steps"])
    out["mach"] = np.float64(0.1)
    out["meta"] = np.array(repr(dict(spec, mach=0.1, re=1600.0)))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)


def main(names):
    os.makedirs(OUT, exist_ok=True)
    work = tempfile.mkdtemp(prefix="hx3d_hsol_")
    try:
        for name in names or list(CASES):
            run(name, CASES[name], work)
            print(name, flush=True)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main(sys.argv[1:])
