#!/usr/bin/env python3
"""Benchmark: DOF-updates/s (FP64) of the plain GPU engine on one MI355X.

Workload (default): the single-GPU case the metric is quoted on -- periodic box
[-pi,pi]^3 of 16^3 curvilinear hexahedra (x' = x + 0.1 prod_k sin(x_k + phi_dk), phases
from seed 2), Gauss-Lobatto N=4 (512,000 nodes), Taylor-Green vortex at Mach 0.1, Re 1600,
split-form kinetic-energy-preserving volume flux (Pirozzoli averaging), Lax-Friedrichs face
flux, BR1 viscous fluxes, low-storage RK3 with dt from CFL 0.4 every step.  One "step" is one
full RK3 step (three time-derivative stages plus MaxTimeStep); one DOF-update is one node
advanced one step (the reference's convention, main.f90:404-409).

Prints ONE JSON line.  `roofline` is for the dominant kernel (the fused volume/surface/RK
kernel) from its HIP-event-timed average launch duration; `cpu_baseline` times the reference
OpenMP build (oracle/_ref, built from the reference's sources) on a bounded sample when it is
present, otherwise the C restatement.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6


def alg_bytes(N):
    """Compulsory FP64 traffic of the two-sweep DGSEM+BR1 stage (bytes):
    per node  - gradient sweep: Q(5) Ja(9) 1/J(1) in, grad(15) out  = 240 B
              - flux sweep (volume kernel): Q(5) grad(15) Ja(9) J(1) G(5) in, G(5) Q(5) out = 360 B
    per face node - gradient sweep: both traces of Q (10), n and |Ja| (4) = 112 B
                  - flux sweep (face kernel): both traces of Q and grad (40), n t1 |Ja| (7) = 376 B
                    (the second tangent is n x t1, formed in the kernel)
    with 3/(N+1) face nodes per node on a conforming hexahedral mesh."""
    fpn = 3.0 / (N + 1)
    return dict(grad=240.0 + 112.0 * fpn, face=376.0 * fpn, volume=360.0,
                stage=600.0 + 488.0 * fpn)


def host_cpu():
    """(model name, physical cores, logical CPUs) of this host from /proc/cpuinfo."""
    model, cores, logical = "unknown", set(), 0
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name":
                model = v
                logical += 1
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
                cores.add((phys, core))
    except OSError:
        pass
    return model, len(cores) or logical, logical


def cpu_threads():
    """The threads this process may run: its CPU affinity, capped by OMP_NUM_THREADS when the
    environment sets it (a GPU box grants each GPU a share of its cores that way)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline_reference(n, N, steps, threads):
    """Time the reference solver on the same box: the -O3 build oracle/_ref/bin/horses3d_O3.ns
    (the -O2 golden-vector build if that is absent), its own 'Solver efficiency' line (seconds
    per million DOF and iteration, main.f90:404-409)."""
    exe, flags = os.path.join(REPO, "oracle", "_ref", "bin", "horses3d_O3.ns"), "-O3"
    if not os.path.exists(exe):
        exe, flags = os.path.join(REPO, "oracle", "_ref", "bin", "horses3d.ns"), "-O2"
    if not os.path.exists(exe):
        return None
    sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
    import refcase as rc
    work = tempfile.mkdtemp(prefix="hx3d_cpu_")
    # straight-sided box: the per-step work of the reference does not depend on the mapping
    rc.write_specmesh(os.path.join(work, "box.mesh"), n)
    rc.write_control(os.path.join(work, "case.control"), "box.mesh", N, steps, 0.1, 1600.0,
                     riemann="lax-friedrichs", inviscid="split-form", averaging="Pirozzoli",
                     ctd_after=False)
    t0 = time.time()
    out = rc.run_reference(work, threads=threads, timeout=900, exe=exe)
    wall = time.time() - t0
    m = re.search(r"Solver efficiency:\s*([0-9.Ee+-]+)", out)
    if not m:
        return None
    eff = float(m.group(1))   # seconds / (1e6 DOF * iteration)
    model, phys, logical = host_cpu()
    return dict(value=1e6 / eff, unit="DOF-updates/s", cores=threads, kind="reference",
                cpu=model, host_physical_cores=phys, host_logical_cpus=logical,
                sample="reference OpenMP build (amdflang %s -fopenmp from the reference sources) on %s, "
                       "%d threads (OMP_NUM_THREADS=%d, OMP_SCHEDULE=static: this process's CPU share; the host "
                       "has %d physical cores), %d^3 straight periodic box, N=%d, split-form Pirozzoli + LxF + BR1, "
                       "%d RK3 steps; a straight box does the same per-step work as the GPU line's curvilinear one "
                       "(the map changes no size, no loop and no flux); its own 'Solver efficiency' line "
                       "(%.3e s/(MDOF*iter)), wall %.1f s incl. setup"
                       % (flags, model, threads, threads, phys, n, N, steps, eff, wall))


def cpu_baseline_port(case, kw, steps):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np
    import pyoracle as po
    from horses3d_amd import mesh
    o = po.Oracle(case, po.make_params(kw["mach"], kw["reynolds"], riemann=kw["riemann"],
                                       inviscid=kw["inviscid"], averaging=kw["averaging"]))
    Q = mesh.taylor_green(case, kw["mach"])
    G = np.zeros_like(Q)
    t0 = time.time()
    for _ in range(steps):
        dt = o.max_dt(Q, 0.4, 0.4)
        o.rk3_step(Q, G, 0.0, dt)
    el = time.time() - t0
    nd = mesh.ndof(case)
    return dict(value=nd * steps / el, unit="DOF-updates/s", cores=1, kind="port",
                sample="C restatement (oracle), %d nodes, %d RK3 steps, 1 thread" % (nd, steps))


def mixed_bench(args):
    """The p-anisotropic case: one JSON line with its rate and per-kernel-group times.  The
    element tiling of the persistent element kernels is autotuned first (hx3d_autotune), then
    the timed steps run with direct launches and with the step as a replayed HIP graph."""
    import torch
    from horses3d_amd import mesh
    from horses3d_amd.engine import Engine
    steps = 30 if args.steps == 200 else args.steps
    print("[bench] building the p-anisotropic 20^3 case", file=sys.stderr, flush=True)
    case = mesh.periodic_box_mixed(20, mesh.random_orders(8000, 2, 7, 3))
    Q0 = mesh.taylor_green(case, 0.2, noise=1e-3, seed=33)
    nd = len(Q0)
    eng = Engine(case, mach=0.2, reynolds=1600.0, riemann="roe")
    eng.set_state(Q0)
    choice, tune_ms = eng.autotune(reps=5)
    print("[bench] autotuned tiling %d (ms per evaluation %s)" % (choice, tune_ms), file=sys.stderr, flush=True)
    t_cw = time.perf_counter()
    while time.perf_counter() - t_cw < args.clock_warmup_s:   # from the initial state each time
        eng.set_state(Q0)
        eng.advance(5, 0.4, 0.4, t=0.0)
        eng.synchronize()
    rates = {}
    for graph in (False, True):
        eng.set_step_graph(graph)
        eng.set_state(Q0)
        eng.advance(args.warmup, 0.4, 0.4, t=0.0)
        eng.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.advance(steps, 0.4, 0.4)
        eng.synchronize()
        rates[graph] = time.perf_counter() - t0
    assert not eng.has_nan()
    eng.set_step_graph(False)
    eng.set_timing(True)
    eng.advance(steps, 0.4, 0.4)
    kt = eng.kernel_times()
    eng.close()
    el = min(rates.values())
    print(json.dumps({"metric": "DOF-updates/s (FP64)", "value": nd * steps / el, "unit": "DOF-updates/s",
                      "steps": steps, "ms_per_step": el / steps * 1e3, "dtype": "f64",
                      "ms_per_step_direct": rates[False] / steps * 1e3, "ms_per_step_graph": rates[True] / steps * 1e3,
                      "tiling": {"policy": choice, "ms_per_evaluation": list(tune_ms)},
                      "stage_ms": {"gradient (face unStar + lifting)": kt[0], "face flux": kt[1], "volume": kt[2]},
                      "config": {"workload": "p-anisotropic 20^3 periodic box, (Nx,Ny,Nz) uniform on {2..7} (seed 3), "
                                             "%d nodes, TGV + 1e-3 noise, Mach 0.2, Re 1600, Roe + BR1" % nd}}))


def gauss_bench(args):
    """Gauss nodes (the reference's default node type): the 16^3 x N=4 curvilinear box
    (amplitude 0.1) built by the SpecMesh front end on Gauss nodes, standard form, Roe + BR1, run
    on the face-trace kernels; the same case on Gauss-Lobatto nodes (the uniform-order kernels)
    beside it.  One JSON line."""
    import tempfile
    import torch
    from horses3d_amd import mesh, specmesh
    from horses3d_amd.engine import Engine, read_specmesh
    steps = 30 if args.steps == 200 else args.steps
    d = tempfile.mkdtemp(prefix="hx3d_gauss_")
    path = os.path.join(d, "box.mesh")
    specmesh.write_specmesh(path, 16, 0.1, mesh.box_phases(2), bface_order=4)
    out = {}
    for nodes in ("gauss", "gauss-lobatto"):
        print("[bench] 16^3 x N=4 on %s nodes" % nodes, file=sys.stderr, flush=True)
        case = read_specmesh(path, N=4, periodic=dict(specmesh.PERIODIC_PAIRS), nodes=nodes)
        Q0 = mesh.taylor_green(case, 0.1)
        nd = len(Q0)
        eng = Engine(case, mach=0.1, reynolds=1600.0, riemann="roe", nodes=nodes)
        del case
        eng.set_state(Q0)
        if nodes == "gauss":
            eng.autotune(reps=5)
        t_cw = time.perf_counter()
        while time.perf_counter() - t_cw < args.clock_warmup_s:
            eng.set_state(Q0)
            eng.advance(10, 0.4, 0.4, t=0.0)
            eng.synchronize()
        eng.set_state(Q0)
        eng.advance(args.warmup, 0.4, 0.4, t=0.0)
        eng.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.advance(steps, 0.4, 0.4)
        eng.synchronize()
        el = time.perf_counter() - t0
        assert not eng.has_nan()
        eng.set_timing(True)
        eng.advance(steps, 0.4, 0.4)
        kt = eng.kernel_times()
        eng.close()
        out[nodes] = dict(value=nd * steps / el, ms_per_step=el / steps * 1e3,
                          stage_ms={"gradient": kt[0], "face flux": kt[1], "volume": kt[2], "stage": kt[3]})
    os.remove(path)
    os.rmdir(d)
    g = out["gauss"]
    print(json.dumps({"metric": "DOF-updates/s (FP64)", "value": g["value"], "unit": "DOF-updates/s",
                      "steps": steps, "ms_per_step": g["ms_per_step"], "dtype": "f64", "stage_ms": g["stage_ms"],
                      "gauss_lobatto_same_case": out["gauss-lobatto"],
                      "config": {"workload": "16^3 curvilinear hexes (amplitude 0.1) x N=4 on Gauss nodes (%d nodes), "
                                             "standard form, Roe + BR1, TGV Mach 0.1 Re 1600, LS-RK3, CFL 0.4" % nd}}))


def batch_bench(args):
    """Many small cases: one engine over their union against one engine per case, same steps."""
    import numpy as np
    import torch
    from horses3d_amd import batch, mesh
    from horses3d_amd.engine import Engine
    rng = np.random.default_rng(7)
    ncase, steps, dt = 64, 100, 1e-3
    cases = [mesh.straight_box(4, 3) for _ in range(ncase)]
    Q0 = [mesh.taylor_green(c, 0.1, phase=tuple(rng.uniform(0.0, 2.0 * np.pi, 3))) for c in cases]
    kw = dict(mach=0.1, reynolds=1600.0, riemann="roe")
    nodes = sum(len(q) for q in Q0)
    print("[bench] one engine per case", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    alone = []
    for c, q in zip(cases, Q0):
        e = Engine(c, **kw)
        e.set_state(q)
        e.advance_dt(steps, dt, t=0.0)
        alone.append(e.get_state())
        e.close()
    t_alone = time.perf_counter() - t0
    print("[bench] one engine for all cases", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = batch.Batch(cases, **kw)
    b.set_states(Q0)
    b.advance_dt(steps, dt, t=0.0)
    got = b.get_states()
    b.close()
    t_batch = time.perf_counter() - t0
    same = all(np.array_equal(a, g) for a, g in zip(alone, got))
    print(json.dumps({"cases": ncase, "steps": steps, "nodes": nodes,
                      "batched_dof_updates_per_s": nodes * steps / t_batch,
                      "one_engine_per_case_dof_updates_per_s": nodes * steps / t_alone,
                      "ratio": t_alone / t_batch, "results_identical": same,
                      "config": {"workload": "%d independent 4^3 x N=3 periodic boxes, TGV with seeded phases, "
                                             "Roe + BR1, Mach 0.1, Re 1600, %d steps of dt %g; wall time incl. "
                                             "engine set-up, upload and read-back" % (ncase, steps, dt)}}))


def serve_bench(args):
    """The serving stream: a Poisson stream of --serve-cases periodic-box cases of n^3 straight
    hexes, n uniform on {8..--serve-nmax}, order N uniform on {3..6}, Mach uniform on [0.1, 0.8],
    Re log-uniform on [400, 5000], Taylor-Green state with seeded phase offsets, Roe + BR1 +
    Smagorinsky (Cs 0.2), step counts uniform on [100, 1000], dt from CFL 0.4 per case.  Each case
    is built when it is set up and dropped once on the device.  Served with shared launches
    (serve_batched: several-case engines grouped by order), then the same stream one case at a
    time on the plain engine; final states compared by digest."""
    import numpy as np
    from horses3d_amd import mesh, serve
    rng = np.random.default_rng(args.serve_seed)

    def req(n, N, mach, re, phase, steps, arrival):
        def build():
            case = mesh.straight_box(n, N)
            return case, mesh.taylor_green(case, mach, phase=phase)
        return serve.Request(None, dict(mach=mach, reynolds=re, riemann="roe", les="smagorinsky", les_c=0.2), None,
                             steps=steps, arrival=arrival, build=build, nodes=n ** 3 * (N + 1) ** 3, order=N)

    reqs, t = [], 0.0
    for k in range(args.serve_cases):
        n, N = int(rng.integers(8, args.serve_nmax + 1)), int(rng.integers(3, 7))
        mach = float(rng.uniform(0.1, 0.8))
        re = float(np.exp(rng.uniform(np.log(400.0), np.log(5000.0))))
        phase = tuple(float(x) for x in rng.uniform(0.0, 2.0 * np.pi, 3))
        t += float(rng.exponential(args.serve_gap))
        reqs.append(req(n, N, mach, re, phase, int(rng.integers(100, 1001)), t))
    work = sum(r.nodes * r.steps for r in reqs)
    print("[bench] serving %d cases, %.3g DOF-updates in all" % (len(reqs), work), file=sys.stderr, flush=True)
    res, st = serve.serve_batched(reqs, chunk=args.serve_chunk, max_batch=64, max_batches=args.serve_running,
                                  workers=args.serve_workers, progress=16, keep_states=False)
    print("[bench] served in %.1f s; one at a time" % st["wall_s"], file=sys.stderr, flush=True)
    alone, st1 = serve.one_at_a_time(reqs, keep_states=False, progress=16)
    same = all(r.t == a.t and r.Q == a.Q for r, a in zip(res, alone))
    print(json.dumps({"serve": st, "one_at_a_time": st1, "results_identical": same,
                      "throughput_ratio": st["dof_updates_per_s"] / st1["dof_updates_per_s"],
                      "one_at_a_time_stepping_only_dof_updates_per_s": st1["dof_updates"] / st1["stepping_s"],
                      "max_batches": args.serve_running, "host_workers": args.serve_workers,
                      "config": {"workload": "Poisson stream of %d periodic-box cases (n^3 straight hexes, n on 8..%d; "
                                             "N on 3..6; Mach 0.1-0.8; Re 400-5000 log-uniform; Roe+BR1+Smagorinsky "
                                             "Cs 0.2; 100-1000 steps, CFL 0.4 per case; mean arrival gap %g s, seed %d)"
                                             % (len(reqs), args.serve_nmax, args.serve_gap, args.serve_seed)}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=16, help="elements per box side")
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--amplitude", type=float, default=0.1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: this process's CPU share, see cpu_threads)")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--clock-warmup-s", type=float, default=0.5,
                    help="seconds of untimed steps that bring the GPU to its sustained clock first")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy kernel 5 times (the "
                         "known-size reference for calibrating FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--gauss", action="store_true",
                    help="Gauss nodes: the 16^3 x N=4 curvilinear box from the SpecMesh front end on Gauss nodes "
                         "(standard form, Roe + BR1) against the same case on Gauss-Lobatto nodes")
    ap.add_argument("--largest", action="store_true",
                    help="the largest single-GPU case instead: 48^3 curvilinear periodic hexes (amplitude 0.15) x N=5 (23.9 M "
                         "nodes), split-form Kennedy-Gruber + Roe + BR1 + WALE, 20 timed steps, no CPU baseline "
                         "(the curvilinear map changes no array size and no kernel)")
    ap.add_argument("--mixed", action="store_true",
                    help="the p-anisotropic case instead: 20^3 straight periodic hexes with orders (Nx, Ny, Nz) "
                         "uniform on {2..7} per element (seed 3, 1.32 M nodes, mortar faces), Taylor-Green state "
                         "with 1e-3 velocity noise, Mach 0.2, Roe + BR1, 30 timed steps, no CPU baseline")
    ap.add_argument("--batch", action="store_true",
                    help="only run the batching comparison: 64 independent 4^3 x N=3 periodic boxes (the "
                         "reference's CPU case size; seeded Taylor-Green phases; Roe + BR1, Mach 0.1, Re 1600) "
                         "for 100 fixed-dt steps, one engine over all of them against one engine per case")
    ap.add_argument("--serve", action="store_true",
                    help="only run the serving comparison: a Poisson stream of independent periodic-box "
                         "cases stepped together (continuous batching) against the same stream one case "
                         "at a time on the plain engine; prints one JSON line of both")
    ap.add_argument("--serve-cases", type=int, default=256)
    ap.add_argument("--serve-nmax", type=int, default=24, help="largest box side n of the serving stream")
    ap.add_argument("--serve-gap", type=float, default=0.2, help="mean arrival gap (s) of the serving stream")
    ap.add_argument("--serve-seed", type=int, default=4)
    ap.add_argument("--serve-chunk", type=int, default=50, help="steps per chunk of a serving engine")
    ap.add_argument("--serve-running", type=int, default=3, help="several-case engines stepping at once")
    ap.add_argument("--serve-workers", type=int, default=8, help="host threads that set up and retire cases")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has one MI355X and the "
              "multi-GPU path is not part of this build (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from horses3d_amd import mesh
    from horses3d_amd.engine import Engine

    assert torch.cuda.is_available(), "no GPU visible"
    torch.cuda.set_device(0)
    if args.copy_probe:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        ms = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            b.copy_(a)
            e1.record()
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        torch.cuda.synchronize()
        # read + write bytes over the best of the five copies: the HBM rate a plain stream reaches
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copy_probe_bytes_written": b.numel() * 8,
                          "copy_probe_best_GBs": round(2 * a.numel() * 8 / (min(ms) * 1e-3) / 1e9, 1)}))
        return
    if args.serve:
        serve_bench(args)
        return
    if args.mixed:
        mixed_bench(args)
        return
    if args.batch:
        batch_bench(args)
        return
    if args.gauss:
        gauss_bench(args)
        return
    if args.largest:
        args.n, args.order, args.no_cpu_baseline = 48, 5, True
        if args.steps == 200:
            args.steps = 20
    print("[bench] building case %d^3 N=%d" % (args.n, args.order), file=sys.stderr, flush=True)
    if args.largest:
        case = mesh.periodic_box(args.n, args.order, amplitude=0.15, phase_seed=2)
        kw = dict(mach=0.1, reynolds=1600.0, riemann="roe", inviscid="split-form", averaging="kennedy-gruber",
                  les="wale")
        form = "split-form Kennedy-Gruber + Roe + BR1 + WALE"
    else:
        case = mesh.periodic_box(args.n, args.order, amplitude=args.amplitude, phase_seed=2)
        kw = dict(mach=0.1, reynolds=1600.0, riemann="lxf", inviscid="split-form", averaging="pirozzoli")
        form = "split-form Pirozzoli + LxF + BR1"
    Q0 = mesh.taylor_green(case, kw["mach"])
    nd = mesh.ndof(case)
    eng = Engine(case, **kw)
    eng.set_state(Q0)
    # The GPU's clocks ramp up under sustained load over some tens of milliseconds: a short timed
    # run (20 steps of this case take 6 ms) would otherwise measure the ramp (20 steps measured
    # 0.318 ms/step right after start-up, 0.281 ms/step once the clocks had settled).  Untimed
    # steps for clock_warmup_s seconds come first; the state is then reset to the initial
    # condition, so the W warm-up and K timed steps are exactly those of a run without them.
    cw_steps, cw_t0 = 0, time.perf_counter()
    while time.perf_counter() - cw_t0 < args.clock_warmup_s:
        eng.advance(25, 0.4, 0.4, t=0.0)
        eng.synchronize()
        cw_steps += 25
    cw_s = time.perf_counter() - cw_t0
    eng.set_state(Q0)
    print("[bench] clock warm-up %d steps (%.2f s), then warmup %d steps" % (cw_steps, cw_s, args.warmup),
          file=sys.stderr, flush=True)
    eng.advance(args.warmup, 0.4, 0.4, t=0.0)
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.advance(args.steps, 0.4, 0.4)
    eng.synchronize()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = nd * args.steps / el
    print("[bench] timed %d steps in %.4f s" % (args.steps, el), file=sys.stderr, flush=True)
    assert not eng.has_nan()

    # per-kernel HIP-event timing over a second pass of the same steps
    eng.set_timing(True)
    eng.advance(args.steps, 0.4, 0.4)
    kt = eng.kernel_times()   # ms per launch: grad, face, volume, stage
    eng.set_timing(False)
    B = alg_bytes(args.order)
    FS = case["F"] * (args.order + 1) ** 2
    kern = {
        "grad": dict(ms=kt[0], bytes=B["grad"] * nd),
        "face": dict(ms=kt[1], bytes=376.0 * FS),
        "volume": dict(ms=kt[2], bytes=B["volume"] * nd),
    }
    for k in kern.values():
        k["GBs"] = k["bytes"] / (k["ms"] * 1e-3) / 1e9
    dom = max(kern, key=lambda k: kern[k]["ms"])
    traffic, fp64 = None, None
    tpath = os.path.join(REPO, "profiles", "traffic_tgv16_n4.json")
    if os.path.exists(tpath) and (args.n, args.order) == (16, 4):
        tj = json.load(open(tpath))
        traffic = tj.get("k_%s_bytes_per_launch" % dom)
        # FP64 vector FLOP per launch from the committed SQ counter run of the same workload
        # (64 x ADD/MUL/TRANS + 128 x FMA wave instructions), over this run's launch time
        flop = tj.get("k_%s_fp64_flop_per_launch" % dom)
        if flop:
            tf = flop / (kern[dom]["ms"] * 1e-3) / 1e12
            fp64 = dict(flop_per_launch=flop, achieved=round(tf, 2), peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                        frac=round(tf / FP64_PEAK_TFLOPS, 4))
    roof = dict(bound="hbm", kernel="k_" + dom, achieved=round(kern[dom]["GBs"], 1), peak=HBM_PEAK_GBS,
                unit="GB/s", frac=round(kern[dom]["GBs"] / HBM_PEAK_GBS, 4), traffic=traffic,
                bytes_per_launch=kern[dom]["bytes"], avg_launch_ms=round(kern[dom]["ms"], 5),
                stage_ms=round(kt[3], 5),
                stage_achieved_GBs=round(B["stage"] * nd / (kt[3] * 1e-3) / 1e9, 1),
                stage_frac=round(B["stage"] * nd / (kt[3] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                # the same figure from the timed region's wall clock: a step is three stages, and
                # the step time also holds the launch gaps that the per-launch events leave out
                stage_ms_timed_region=round(el / args.steps * 1e3 / 3.0, 5),
                stage_frac_timed_region=round(B["stage"] * nd / (el / args.steps / 3.0) / 1e9 / HBM_PEAK_GBS, 4),
                kernels={k: dict(avg_launch_ms=round(v["ms"], 5), GBs=round(v["GBs"], 1)) for k, v in kern.items()},
                fp64_vector=fp64)
    eng.close()
    single_batch = None
    if args.largest:
        # the same case through the batched engine as a single-case batch (block table, per-case
        # constants and dt slot): the same steps from the same state
        from horses3d_amd.engine import MultiEngine
        print("[bench] single-case batch", file=sys.stderr, flush=True)
        me = MultiEngine([case], [kw])
        me.set_states([Q0])
        me.advance(args.warmup, 0.4, 0.4)
        me.synchronize()
        t0 = time.perf_counter()
        me.advance(args.steps, 0.4, 0.4)
        me.synchronize()
        elb = time.perf_counter() - t0
        times, bad = me.status()
        me.close()
        single_batch = dict(value=nd * args.steps / elb, ms_per_step=elb / args.steps * 1e3, diverged=bool(bad[0]))

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        try:
            cpu = cpu_baseline_reference(args.n, args.order, args.cpu_steps, args.cpu_threads or cpu_threads())
        except Exception as e:   # the reference build is optional on a box
            print("[bench] reference CPU baseline failed: %s" % e, file=sys.stderr, flush=True)
        if cpu is None:
            small = mesh.periodic_box(8, args.order, amplitude=args.amplitude, phase_seed=2)
            cpu = cpu_baseline_port(small, kw, 3)

    line = {
        "metric": "DOF-updates/s (FP64)",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "clock_warmup": {"steps": cw_steps, "seconds": round(cw_s, 3),
                         "note": "untimed steps before the warm-up, then the initial state again"},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Taylor-Green vortex on a seeded curvilinear periodic box)",
        "config": {"workload": "plain DGSEM compressible NS, %d^3 %s hexes x N=%d (%d nodes), "
                               "%s, LS-RK3, CFL 0.4"
                               % (args.n, "curvilinear (amplitude %g)" % (0.15 if args.largest else args.amplitude),
                                  args.order, nd, form),
                   "elements": case["K"], "order": args.order, "ndof": nd, "mach": 0.1,
                   "reynolds": 1600.0, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if single_batch:
        line["single_case_batch"] = single_batch
    print(json.dumps(line))


if __name__ == "__main__":
    main()
