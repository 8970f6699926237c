""".control fixtures: hand-written control files that exercise the reader's rules (comments,
quotes, repeated keys, upper case, '#define' blocks, boundary name lists, the split-form node
switch, the three ways to give polynomial orders, Euler flows), each run through the reference
build (oracle/_ref/bin/horses3d.ns) for one step with the ProblemFile dump at step 0.

Kept per case under tests/golden/control: the control file itself, the orders file it names
(when it names one), and an .npz with the mesh arrays, state and time derivative the reference
built from it, plus the zone table and scheme names the reference printed.

    python tests/golden/make_control.py
"""
import os
import re
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

OUT = os.path.join(HERE, "control")

RUN_KEYS = '''solution file name    = "RESULTS/case.hsol"
restart               = .false.
restart file name     = "RESULTS/case.rst"
number of time steps  = 1
Output Interval       = 1
convergence tolerance = 1.d-30
compute time derivative after timestep = .true.
'''

MONITOR = '''#define volume monitor 1
   name = KinEn
   variable = kinetic energy
   prandtl number = 0.5
   lambda stabilization = 3.0
#end
'''

CASES = {
    # comments and quotes, repeated keys (the first one counts), keys in upper case, a
    # d exponent, values without quotes, keys inside '#define' blocks (not main keys), a boundary
    # block listing two names (one not in the mesh), a second block for a zone (not read), a
    # repeated key inside a block, split form on "Gauss" nodes (switched to Gauss-Lobatto)
    "syntax_split_kg_lxf": dict(n=2, amp=0.1, seed=6, N=3, orders=None, text='''!  hand-written control file: reader rules
/  a line that starts with a slash is blank
Flow equations        = "NS"        ! trailing comment
   MESH FILE NAME     = box.mesh
''' + RUN_KEYS + '''cfl                   = 0.3
DCFL                  = 0.25
Mach number           = "2.5d-1"    ! was "0.9"
mach number           = 0.9
Reynolds Number       = 1.2E3
riemann solver        = Lax-Friedrichs
Lambda stabilization  = 0.5
inviscid discretization = "Split-Form"
averaging             = "Kennedy-Gruber"
discretization nodes  = Gauss
polynomial order      = 3
''' + MONITOR + '''
#define boundary LEFT__nowhere
   type = periodic
   coupled boundary = "Right"
#end
#define boundary right
   TYPE = Periodic
   coupled boundary = left
#end
#define boundary right
   type = inflow
#end
#define boundary top__bottom
   type = periodic
   coupled boundary = bottom
#end
#define boundary back__front
   coupled boundary = front
   type = periodic
   coupled boundary = back
#end
'''),
    # one order per direction (polynomial order i/j/k), Smagorinsky LES with its intensity,
    # Roe by default, the standard form, a Prandtl number, dt from cfl/dcfl
    "orders_ijk_smagorinsky": dict(n=2, amp=0.0, seed=0, N=4, orders=None, text='''Flow equations = "NS"
mesh file name = "box.mesh"
''' + RUN_KEYS + '''cfl = 0.4
dcfl = 0.4
mach number = 0.3
reynolds number = 900
prandtl number = 0.8
polynomial order i = 2
polynomial order j = 4
polynomial order k = 3
les model = "Smagorinsky"
les model intensity = 0.15
discretization nodes = "Gauss-Lobatto"
#define boundary left__right
   type = periodic
   coupled boundary = right
#end
#define boundary top__bottom
   type = periodic
   coupled boundary = bottom
#end
#define boundary back__front
   type = periodic
   coupled boundary = front
#end
'''),
    # Euler flow (no Reynolds number), per-element orders from a file, split form with
    # Pirozzoli averaging, Roe
    "order_file_euler_pirozzoli": dict(n=2, amp=0.1, seed=11, N=4, orders=(2, 4, 55), text='''flow equations = "Euler"
mesh file name = "box.mesh"
''' + RUN_KEYS + '''cfl = 0.4
dcfl = 0.4
mach number = 0.2
polynomial order file = "orders.txt"
inviscid discretization = "split-form"
averaging = "Pirozzoli"
riemann solver = "Roe"
discretization nodes = "Gauss-Lobatto"
#define boundary left
   type = periodic
   coupled boundary = right
#end
#define boundary right
   type = periodic
   coupled boundary = left
#end
#define boundary top
   type = periodic
   coupled boundary = bottom
#end
#define boundary bottom
   type = periodic
   coupled boundary = top
#end
#define boundary front
   type = periodic
   coupled boundary = back
#end
#define boundary back
   type = periodic
   coupled boundary = front
#end
'''),
}


def parse_stdout(text):
    """Zone table and scheme names the reference printed while setting up."""
    zones = re.findall(r"Zone\s+\d+ for boundary: (\S+)\s*\n[-\s]*\n\s*->\s*Number of faces:\s*\d+\s*\n"
                       r"\s*->\s*Boundary condition type: (\S+)\s*\n(?:\s*->\s*Coupled boundary: (\S+))?", text)
    def item(label):
        m = re.search(r"->\s*%s:\s*(.+?)\s*$" % re.escape(label), text, re.M)
        return m.group(1) if m else ""
    return dict(zones=[(a, b, c) for a, b, c in zones],
                scheme=item("Numerical scheme"), averaging=item("Averaging function"),
                riemann=item("Riemann solver"), lambda_stab=item("Lambda stabilization"),
                mach=item("Mach number"), reynolds=item("Reynolds number"), prandtl=item("Prandtl number"))


def run(name, spec, work):
    d = os.path.join(work, name)
    os.makedirs(d)
    ph = np.zeros(9) if not spec["seed"] else np.random.default_rng(spec["seed"]).uniform(0.0, 2.0 * np.pi, 9)
    rc.write_specmesh(os.path.join(d, "box.mesh"), spec["n"], spec["amp"], ph, bface_order=max(4, spec["N"]))
    os.makedirs(OUT, exist_ok=True)
    if spec["orders"]:
        lo, hi, seed = spec["orders"]
        orders = np.random.default_rng(seed).integers(lo, hi + 1, (spec["n"] ** 3, 3))
        rc.write_orders(os.path.join(d, "orders.txt"), orders)
        shutil.copy(os.path.join(d, "orders.txt"), os.path.join(OUT, name + "_orders.txt"))
    with open(os.path.join(d, "case.control"), "w") as fh:
        fh.write(spec["text"])
    shutil.copy(os.path.join(d, "case.control"), os.path.join(OUT, name + ".control"))
    out = rc.run_reference(d, env={"HX3D_DUMP_PREFIX": os.path.join(d, "dump"), "HX3D_DUMP_STEPS": "0"})
    g = po.from_reference_dump(rc.read_dump(os.path.join(d, "dump_step0.bin")))
    info = parse_stdout(out)
    keep = {k: g[k] for k in ("N", "off", "faceID", "faceSide", "Ja", "J", "fElem", "fESide", "fRot",
                              "fNf", "foff", "normal", "t1", "t2", "fjac", "x", "Q", "QDot")}
    keep["K"], keep["F"], keep["time"] = np.int64(g["K"]), np.int64(g["F"]), np.float64(g["time"])
    meta = dict(n=spec["n"], amp=spec["amp"], phase_seed=spec["seed"], bface_order=max(4, spec["N"]))
    keep["meta"] = np.array(repr(meta))
    keep["printed"] = np.array(repr(info))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **keep)
    return info


def main(argv):
    work = tempfile.mkdtemp(prefix="hx3d_control_")
    try:
        for name in argv or list(CASES):
            info = run(name, CASES[name], work)
            print(name, info, flush=True)
    finally:
        shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main(sys.argv[1:])
