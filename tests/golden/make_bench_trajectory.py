#!/usr/bin/env python3
"""Write tests/golden/bench/tgv16_n4_curved_split_pirozzoli_lxf_100steps.npz: the reference
solver (oracle/_ref/bin/horses3d.ns, built from the reference's sources) run on the benchmark
case itself -- 16^3 curvilinear periodic box (amplitude 0.1, phases of seed 2, boundary order 4),
N = 4 (512,000 nodes), Taylor-Green at Mach 0.1, Re 1600, split form Pirozzoli, Lax-Friedrichs,
BR1, CFL 0.4 -- for 100 RK3 steps, without CTD_AFTER_STEPS.

The full state is too large to keep, so the fixture holds what the full-size parity test compares:
the entries where the reference's initial state differs from mesh.taylor_green (a few last-bit
differences of the Fortran runtime's sines and cosines), the dt of every step, the kinetic-energy monitor of every step, the state and time derivative at
4096 seeded nodes plus every node of two elements at steps 0, 50 and 100, and per-variable
maxima and sums of squares of the state at those steps.

Run from the repository root:  python tests/golden/make_bench_trajectory.py [--threads 8]
"""
import argparse
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))
import refcase as rc  # noqa: E402
import pyoracle as po  # noqa: E402
from horses3d_amd import mesh  # noqa: E402

N_EL, ORDER, AMP, SEED, STEPS = 16, 4, 0.1, 2, 100
DUMPS = (0, 50, 100)
OUT = os.path.join(HERE, "bench", "tgv16_n4_curved_split_pirozzoli_lxf_100steps.npz")


def sample_nodes(nodes, ne):
    r = np.random.default_rng(1616)
    pick = r.choice(nodes, 4096, replace=False)
    el = np.concatenate([np.arange(ne), 2047 * ne + np.arange(ne)])   # elements 0 and 2047 whole
    return np.unique(np.concatenate([pick, el])).astype(np.int64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    work = tempfile.mkdtemp(prefix="hx3d_bench_")
    try:
        rc.write_specmesh(os.path.join(work, "box.mesh"), N_EL, AMP, mesh.box_phases(SEED), bface_order=4)
        rc.write_control(os.path.join(work, "case.control"), "box.mesh", ORDER, STEPS, 0.1, 1600.0,
                         riemann="lax-friedrichs", inviscid="split-form", averaging="Pirozzoli",
                         ctd_after=False)
        env = {"HX3D_DUMP_PREFIX": os.path.join(work, "dump"),
               "HX3D_DUMP_STEPS": ",".join(str(s) for s in DUMPS)}
        rc.run_reference(work, env=env, threads=args.threads)
        out = {}
        sel = None
        for s in DUMPS:
            name = "dump_step%d.bin" % s if s < STEPS else "dump_final.bin"
            d = po.from_reference_dump(rc.read_dump(os.path.join(work, name)))
            if sel is None:
                sel = sample_nodes(len(d["Q"]), (ORDER + 1) ** 3)
                out["nodes"] = np.int64(len(d["Q"]))
                out["sample"] = sel
                out["x_sample"] = d["x"][sel]
                # The reference's initial state against mesh.taylor_green of the same node
                # coordinates: the formula and its operation order agree, but the Fortran runtime's
                # sines and cosines round a few arguments the other way.  The differing entries
                # (flat index into the (nodes, 5) state, reference value) make the state exact.
                Qpy = mesh.taylor_green({"x": d["x"]}, 0.1).ravel()
                Qref = np.ascontiguousarray(d["Q"]).ravel()
                fix = np.flatnonzero(Qpy != Qref)
                out["ic_fix_index"] = fix.astype(np.int64)
                out["ic_fix_value"] = Qref[fix]
                print("initial state: %d of %d entries differ from mesh.taylor_green" % (len(fix), len(Qref)))
            out["Q_step%d" % s] = d["Q"][sel]
            out["QDot_step%d" % s] = d["QDot"][sel]
            out["t_step%d" % s] = np.float64(d["time"])
            out["Qmax_step%d" % s] = np.abs(d["Q"]).max(axis=0)
            out["Qsq_step%d" % s] = (d["Q"] * d["Q"]).sum(axis=0)
        out["dt_log"] = rc.read_dt_log(os.path.join(work, "dump_dt.txt"))
        out["ke_monitor"] = rc.read_volume_monitor(work)
        out["meta"] = np.array(repr(dict(n=N_EL, order=ORDER, amp=AMP, phase_seed=SEED, steps=STEPS, mach=0.1,
                                         re=1600.0, riemann="lax-friedrichs", inviscid="split-form",
                                         averaging="Pirozzoli", ctd_after=False, bface_order=4)))
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        np.savez_compressed(OUT, **out)
        print("%s %.1f KB" % (OUT, os.path.getsize(OUT) / 1024.0))
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main()
