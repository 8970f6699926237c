#!/usr/bin/env python3
"""Write the golden fixtures of tests/golden/ by running the reference solver built from
its own sources (oracle/_ref/bin/horses3d.ns, made by oracle/ref/build_ref.py).

Each fixture holds the reference's own mesh connectivity and geometry (the HexMesh
element/face storage) together with states Q and time derivatives QDot dumped by
oracle/ref/ProblemFile.f90 at chosen steps, the reference's dt sequence (MaxTimeStep,
printed before every step) and its kinetic-energy volume monitor.

Run from the repository root:  python tests/golden/make_golden.py [case ...]
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

# name: (mesh n, N, amp, phase seed, steps, dump steps, control kwargs, ic noise amp, noise seed)
CASES = {
    # the reference's CPU case: 4^3 straight, N=3, Roe, BR1, standard form, M=0.1, Re=1600, 100 steps
    "c1_tgv_n4_p3_roe_standard": dict(n=4, N=3, amp=0.0, seed=0, steps=100, dumps=[0, 1, 10],
                                      ctl=dict(mach=0.1, re=1600.0)),
    # the 16^3 x N=4 benchmark case in small: curvilinear map, split form Pirozzoli (KEP), LxF
    "c2s_tgv_n3_p4_curved_split_pirozzoli_lxf": dict(
        n=3, N=4, amp=0.1, seed=2, steps=5, dumps=[0, 2],
        ctl=dict(mach=0.1, re=1600.0, riemann="lax-friedrichs", inviscid="split-form",
                 averaging="Pirozzoli")),
    # the 48^3 x N=5 case in small: curvilinear (amplitude 0.15), split KEP, Roe, WALE, N=5
    "c5s_tgv_n2_p5_curved_split_kg_roe_wale": dict(
        n=2, N=5, amp=0.15, seed=5, steps=4, dumps=[0, 2],
        ctl=dict(mach=0.1, re=1600.0, riemann="roe", inviscid="split-form",
                 averaging="Kennedy-Gruber", les="Wale")),
    # a serving-stream case in small: Smagorinsky LES, Mach 0.5, Re 800, N=3
    "c4s_tgv_n3_p3_smagorinsky": dict(
        n=3, N=3, amp=0.0, seed=0, steps=4, dumps=[0, 2],
        ctl=dict(mach=0.5, re=800.0, les="Smagorinsky", les_intensity=0.2)),
    # further flux combinations on a perturbed state
    "x_noise_n3_p3_curved_split_ducros_lxf": dict(
        n=3, N=3, amp=0.1, seed=7, steps=3, dumps=[0], noise=0.05, noise_seed=71,
        ctl=dict(mach=0.3, re=500.0, riemann="lax-friedrichs", inviscid="split-form",
                 averaging="Ducros")),
    "x_noise_n2_p6_curved_split_standard_roe": dict(
        n=2, N=6, amp=0.1, seed=8, steps=2, dumps=[0], noise=0.05, noise_seed=81,
        ctl=dict(mach=0.2, re=1000.0, riemann="roe", inviscid="split-form",
                 averaging="Standard")),
    "x_noise_n2_p7_standard_lxf_kg": dict(
        n=2, N=7, amp=0.1, seed=9, steps=2, dumps=[0], noise=0.05, noise_seed=91,
        ctl=dict(mach=0.4, re=2000.0, riemann="lax-friedrichs", inviscid="standard",
                 averaging="Kennedy-Gruber")),
    # the p-anisotropic 20^3 case in small: p-anisotropic orders (Nx, Ny, Nz) uniform on {2..5} per element, mortar
    # faces between mismatched orders, velocity noise, Mach 0.2, Roe + BR1
    "c3s_tgv_n3_mixed_p2to5_roe": dict(
        n=3, N=5, orders=(2, 5, 33), amp=0.0, seed=0, steps=3, dumps=[0, 2], noise=1e-3, noise_seed=333,
        ctl=dict(mach=0.2, re=1600.0)),
    "x_noise_n2_mixed_p2to6_curved_split_pirozzoli_lxf": dict(
        n=2, N=6, orders=(2, 6, 44), amp=0.1, seed=4, steps=2, dumps=[0], noise=0.05, noise_seed=441,
        ctl=dict(mach=0.3, re=800.0, riemann="lax-friedrichs", inviscid="split-form",
                 averaging="Pirozzoli")),
    # Morinishi averaging: the split-form two-point flux and, in the standard form, the LxF face
    # average (AveragedStates)
    "x_noise_n2_p4_curved_split_morinishi_lxf": dict(
        n=2, N=4, amp=0.1, seed=12, steps=2, dumps=[0], noise=0.05, noise_seed=121,
        ctl=dict(mach=0.3, re=900.0, riemann="lax-friedrichs", inviscid="split-form",
                 averaging="Morinishi")),
    "x_noise_n2_p3_curved_standard_lxf_morinishi": dict(
        n=2, N=3, amp=0.1, seed=13, steps=2, dumps=[0], noise=0.05, noise_seed=131,
        ctl=dict(mach=0.25, re=1200.0, riemann="lax-friedrichs", inviscid="standard",
                 averaging="Morinishi")),
    "x_noise_n3_p2_standard_roe": dict(
        n=3, N=2, amp=0.1, seed=10, steps=2, dumps=[0], noise=0.05, noise_seed=101,
        ctl=dict(mach=0.8, re=400.0)),
    # Gauss nodes (the reference's default "discretization nodes"; standard form only): dense
    # prolongation to the faces and surface weights on every node
    "g1_tgv_n3_p3_gauss_roe_standard": dict(
        n=3, N=3, amp=0.0, seed=0, steps=4, dumps=[0, 2],
        ctl=dict(mach=0.1, re=1600.0, nodes="Gauss")),
    "g2_noise_n2_p4_curved_gauss_lxf_smagorinsky": dict(
        n=2, N=4, amp=0.1, seed=14, steps=2, dumps=[0], noise=0.05, noise_seed=141,
        ctl=dict(mach=0.3, re=900.0, riemann="lax-friedrichs", les="Smagorinsky", les_intensity=0.2,
                 nodes="Gauss")),
    "g3_noise_n3_mixed_p2to5_gauss_roe": dict(
        n=3, N=5, orders=(2, 5, 55), amp=0.1, seed=15, steps=2, dumps=[0], noise=0.02, noise_seed=151,
        ctl=dict(mach=0.2, re=1600.0, nodes="Gauss")),
}


# Runs without "compute time derivative after timestep" (the reference's default): the QDot the
# final dump holds is the last RK stage's, the derivative ComputeMaxResiduals reads after a step
# (TimeIntegrator.f90:852).  Written to stage3/ (their final QDot is not QDot(Q_final)).
STAGE3_CASES = {
    "c1_tgv_n4_p3_roe_standard_10steps": dict(n=4, N=3, amp=0.0, seed=0, steps=10, dumps=[0],
                                              ctl=dict(mach=0.1, re=1600.0, ctd_after=False)),
    "c2s_tgv_n3_p4_curved_split_pirozzoli_lxf_5steps": dict(
        n=3, N=4, amp=0.1, seed=2, steps=5, dumps=[0],
        ctl=dict(mach=0.1, re=1600.0, riemann="lax-friedrichs", inviscid="split-form",
                 averaging="Pirozzoli", ctd_after=False)),
}


def phases(seed):
    if seed == 0:
        return np.zeros(9)
    return np.random.default_rng(seed).uniform(0.0, 2.0 * np.pi, 9)


def run_case(name, spec, workdir, outdir=HERE):
    d = os.path.join(workdir, name)
    os.makedirs(d, exist_ok=True)
    rc.write_specmesh(os.path.join(d, "box.mesh"), spec["n"], spec["amp"], phases(spec["seed"]),
                      bface_order=max(4, spec["N"]))
    ctl = dict(spec["ctl"])
    mach, re = ctl.pop("mach"), ctl.pop("re")
    if spec.get("orders"):
        lo, hi, oseed = spec["orders"]
        orders = np.random.default_rng(oseed).integers(lo, hi + 1, (spec["n"] ** 3, 3))
        rc.write_orders(os.path.join(d, "orders.txt"), orders)
        ctl["order_file"] = "orders.txt"
    env = {}
    if spec.get("noise"):
        # pass 1 (no steps) to obtain the node coordinates the reference builds
        rc.write_control(os.path.join(d, "case.control"), "box.mesh", spec["N"], 0, mach, re, **ctl)
        rc.run_reference(d, env={"HX3D_DUMP_PREFIX": os.path.join(d, "geo")})
        g = po.from_reference_dump(rc.read_dump(os.path.join(d, "geo_final.bin")))
        rng = np.random.default_rng(spec["noise_seed"])
        Q = g["Q"].copy()
        Q[:, 0] *= 1.0 + spec["noise"] * rng.standard_normal(len(Q))
        Q[:, 1:4] += spec["noise"] * rng.standard_normal((len(Q), 3))
        Q[:, 4] *= 1.0 + 0.1 * spec["noise"] * rng.standard_normal(len(Q))
        # back to element-ordered Fortran layout Q(1:5,i,j,k) per element: node-major packed
        Q.astype("<f8").tofile(os.path.join(d, "ic.bin"))
        env["HX3D_IC_FILE"] = os.path.join(d, "ic.bin")
    rc.write_control(os.path.join(d, "case.control"), "box.mesh", spec["N"], spec["steps"], mach, re, **ctl)
    env.update({"HX3D_DUMP_PREFIX": os.path.join(d, "dump"),
                "HX3D_DUMP_STEPS": ",".join(str(s) for s in spec["dumps"])})
    rc.run_reference(d, env=env)
    dumps = {}
    for s in spec["dumps"]:
        dumps[s] = po.from_reference_dump(rc.read_dump(os.path.join(d, "dump_step%d.bin" % s)))
    final = po.from_reference_dump(rc.read_dump(os.path.join(d, "dump_final.bin")))
    dts = rc.read_dt_log(os.path.join(d, "dump_dt.txt"))
    ke = rc.read_volume_monitor(d)
    base = dumps[spec["dumps"][0]]
    out = {k: base[k] for k in ("N", "off", "faceID", "faceSide", "Ja", "J", "fElem", "fESide",
                                "fRot", "fNf", "foff", "normal", "t1", "t2", "fjac", "x")}
    out["K"] = np.int64(base["K"])
    out["F"] = np.int64(base["F"])
    for s, dd in dumps.items():
        out["Q_step%d" % s] = dd["Q"]
        out["QDot_step%d" % s] = dd["QDot"]
        out["t_step%d" % s] = np.float64(dd["time"])
    out["Q_final"] = final["Q"]
    out["QDot_final"] = final["QDot"]
    out["t_final"] = np.float64(final["time"])
    out["steps"] = np.int64(spec["steps"])
    out["dt_log"] = dts
    out["ke_monitor"] = ke
    meta = dict(spec["ctl"])
    meta.update(n=spec["n"], order=spec["N"], amp=spec["amp"], phase_seed=spec["seed"],
                noise=spec.get("noise", 0.0), orders=spec.get("orders"))
    out["meta"] = np.array(repr(meta))
    path = os.path.join(outdir, name + ".npz")
    np.savez_compressed(path, **out)
    return path


def main(argv):
    names = argv or list(CASES) + list(STAGE3_CASES)
    work = tempfile.mkdtemp(prefix="hx3d_golden_")
    try:
        for nm in names:
            if nm in STAGE3_CASES:
                os.makedirs(os.path.join(HERE, "stage3"), exist_ok=True)
                p = run_case(nm, STAGE3_CASES[nm], work, os.path.join(HERE, "stage3"))
            else:
                p = run_case(nm, CASES[nm], work)
            print("%-48s %8.1f KB" % (os.path.basename(p), os.path.getsize(p) / 1024.0), flush=True)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main(sys.argv[1:])
