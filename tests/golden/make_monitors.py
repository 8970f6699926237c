"""Volume-monitor fixtures: runs the reference build (oracle/_ref/bin/horses3d.ns) on two small
periodic boxes with the volume monitors of VolumeMonitor.f90 switched on, and keeps each
monitor's values together with the state and time its ProblemFile dump holds at the end.

    python tests/golden/make_monitors.py     # writes tests/golden/monitors/*.npz
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

OUT = os.path.join(HERE, "monitors")
MONITORS = (("kinen", "Kinetic energy"), ("kinrate", "Kinetic energy rate"), ("enstr", "Enstrophy"),
            ("meanvel", "Mean velocity"), ("inten", "Internal energy"), ("entropy", "Entropy"),
            ("mathent", "Math entropy"))
CASES = {
    # 3^3 curved box, N = 3, split-form Pirozzoli + LxF, 4 steps
    "box3_p3_curved_split": dict(n=3, N=3, amp=0.1, seed=6, steps=4, orders=None,
                                 ctl=dict(riemann="lax-friedrichs", inviscid="split-form", averaging="Pirozzoli")),
    # 2^3 straight box with per-element orders on {2..4}, Roe + standard form, 3 steps
    "box2_mixed_p2to4": dict(n=2, N=4, amp=0.0, seed=0, steps=3, orders=(2, 4, 21), ctl=dict()),
}


def run(name, spec, work):
    d = os.path.join(work, name)
    os.makedirs(d)
    ph = np.zeros(9) if not spec["seed"] else np.random.default_rng(spec["seed"]).uniform(0.0, 2.0 * np.pi, 9)
    rc.write_specmesh(os.path.join(d, "box.mesh"), spec["n"], spec["amp"], ph, bface_order=max(4, spec["N"]))
    kw = dict(spec["ctl"])
    if spec["orders"]:
        lo, hi, seed = spec["orders"]
        orders = np.random.default_rng(seed).integers(lo, hi + 1, (spec["n"] ** 3, 3))
        rc.write_orders(os.path.join(d, "orders.txt"), orders)
        kw["order_file"] = "orders.txt"
    rc.write_control(os.path.join(d, "case.control"), "box.mesh", spec["N"], spec["steps"], 0.1, 1600.0,
                     monitors=MONITORS, **kw)
    rc.run_reference(d, env={"HX3D_DUMP_PREFIX": os.path.join(d, "dump"), "HX3D_DUMP_STEPS": "0"})
    g = po.from_reference_dump(rc.read_dump(os.path.join(d, "dump_final.bin")))
    out = dict(Q_final=g["Q"], QDot_final=g["QDot"], t_final=np.float64(g["time"]), N=g["N"],
               steps=np.int64(spec["steps"]))
    for key, _ in MONITORS:
        out["mon_" + key] = rc.read_volume_monitor(d, key)
    out["meta"] = np.array(repr(dict(spec, mach=0.1, re=1600.0)))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **out)


def main():
    os.makedirs(OUT, exist_ok=True)
    work = tempfile.mkdtemp(prefix="hx3d_mon_")
    try:
        for name, spec in CASES.items():
            run(name, spec, work)
            print(name, flush=True)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main()
