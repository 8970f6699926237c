"""Drop-in fixtures: the reference solver (oracle/_ref/bin/horses3d.ns, its own TakeRK3Step on the
CPU) on three small cases, keeping the residual history (ComputeMaxResiduals after every step,
RESULTS/case.residuals) and the kinetic-energy monitor (RESULTS/case.kinen.volume) it writes.
tests/test_dropin_gpu.py runs the same set-up with the RK3 step taken by the hx3d engine through
the Fortran adapter (oracle/_ref/bin/hx3d_dropin.ns) and compares the two.

    python tests/golden/make_dropin.py [case ...]    # writes tests/golden/dropin/*.npz
"""
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import refcase as rc  # noqa: E402

OUT = os.path.join(HERE, "dropin")
# The keys the reference's CheckInputIntegrity would otherwise default (main.f90:238-318); the
# drop-in driver replaces main.f90 and with it that routine.
DEFAULTED = [("save gradients with solution", ".false."), ("save sensor with solution", ".false.")]
CASES = {
    # the plain 4^3 x N=3 Taylor-Green case: Roe, standard form, 100 steps, the reference's
    # default residual source (the last RK stage's QDot)
    "c1_tgv_n4_p3_roe_standard_100steps": dict(n=4, N=3, amp=0.0, seed=0, steps=100, ctd_after=False,
                                               ctl=dict(riemann="roe", inviscid="standard", averaging="Standard")),
    # 3^3 curved box, N = 4, split-form Pirozzoli + LxF, 20 steps, derivative after every step
    "c2_tgv_n3_p4_curved_split_pirozzoli_lxf_20steps": dict(
        n=3, N=4, amp=0.1, seed=6, steps=20, ctd_after=True,
        ctl=dict(riemann="lax-friedrichs", inviscid="split-form", averaging="Pirozzoli")),
    # 3^3 curved box, N = 3 on Gauss nodes (the reference's default node type), Roe, standard
    # form, 30 steps, the last RK stage's residuals
    "c3_tgv_n3_p3_curved_gauss_roe_30steps": dict(
        n=3, N=3, amp=0.1, seed=9, steps=30, ctd_after=False,
        ctl=dict(riemann="roe", inviscid="standard", averaging="Standard", nodes="Gauss")),
}


def write_case(d, spec):
    """Mesh and control file of a case into directory d (shared with the test)."""
    ph = np.zeros(9) if not spec["seed"] else np.random.default_rng(spec["seed"]).uniform(0.0, 2.0 * np.pi, 9)
    rc.write_specmesh(os.path.join(d, "box.mesh"), spec["n"], spec["amp"], ph, bface_order=max(4, spec["N"]))
    rc.write_control(os.path.join(d, "case.control"), "box.mesh", spec["N"], spec["steps"], 0.1, 1600.0,
                     ctd_after=spec["ctd_after"], extra=DEFAULTED, **spec["ctl"])
    os.makedirs(os.path.join(d, "RESULTS"), exist_ok=True)
    os.makedirs(os.path.join(d, "MESH"), exist_ok=True)


def read_residuals(d):
    """RESULTS/case.residuals: iteration, time, then the five residuals (the elapsed-time columns
    are dropped)."""
    rows = []
    for line in open(os.path.join(d, "RESULTS", "case.residuals")):
        t = line.split()
        if t and not t[0].startswith("#"):
            v = [float(x) for x in t]
            rows.append([v[0], v[1]] + v[4:9])
    return np.array(rows)


def main(names):
    os.makedirs(OUT, exist_ok=True)
    work = tempfile.mkdtemp(prefix="hx3d_dropin_")
    try:
        for name in names or list(CASES):
            spec = CASES[name]
            d = os.path.join(work, name)
            os.makedirs(d)
            write_case(d, spec)
            rc.run_reference(d)
            np.savez_compressed(os.path.join(OUT, name + ".npz"), residuals=read_residuals(d),
                                kinen=rc.read_volume_monitor(d, "kinen"), meta=np.array(repr(spec)))
            print(name, flush=True)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main(sys.argv[1:])
