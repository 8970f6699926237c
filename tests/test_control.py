""".control files (hx3d_control_read / hx3d_control_case) against the reference.

tests/golden/control holds hand-written control files and what the reference solver built from
each (make_control.py): the mesh arrays, the state and time derivative of its ProblemFile dump at
step 0, and the zone table and scheme names it printed.  Read by the C reader here, each file
must give:
  * the reference's boundary conditions for every zone (type and coupled zone),
  * the same mesh, orders, topology and geometry, bit for bit (periodic pairs from the
    boundary blocks, orders from "polynomial order", "polynomial order i/j/k" or an order file),
  * physics and discretisation settings with which the oracle reproduces the reference's time
    derivative to 1e-12 (a wrong Mach, Reynolds or Prandtl number, Riemann solver, lambda,
    averaging or LES constant would not).
Plus the reader's errors.  Host only."""
import ast
import os
import shutil

import numpy as np
import pytest

import pyoracle as po
from horses3d_amd import engine as E
from horses3d_amd import mesh, specmesh

CTL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "control")
CASES = sorted(f[:-8] for f in os.listdir(CTL) if f.endswith(".control"))
INTS = ("N", "faceID", "faceSide", "fElem", "fESide", "fRot")
REALS = ("x", "Ja", "J", "normal", "t1", "t2", "fjac")
ZONES = ("left", "top", "back", "bottom", "front", "right")


def stage(name, d):
    """The case directory the reference ran in: control file, mesh, orders file."""
    z = np.load(os.path.join(CTL, name + ".npz"), allow_pickle=False)
    meta = ast.literal_eval(str(z["meta"]))
    shutil.copy(os.path.join(CTL, name + ".control"), os.path.join(str(d), "case.control"))
    if os.path.exists(os.path.join(CTL, name + "_orders.txt")):
        shutil.copy(os.path.join(CTL, name + "_orders.txt"), os.path.join(str(d), "orders.txt"))
    specmesh.write_specmesh(os.path.join(str(d), "box.mesh"), meta["n"], meta["amp"],
                            mesh.box_phases(meta["phase_seed"]), bface_order=meta["bface_order"])
    return z, ast.literal_eval(str(z["printed"]))


@pytest.mark.parametrize("name", CASES)
def test_boundary_blocks_match_reference(name, tmp_path):
    z, printed = stage(name, tmp_path)
    c = E.Control(os.path.join(str(tmp_path), "case.control"))
    assert [zn for zn, _, _ in printed["zones"]] == list(ZONES)
    for zone, typ, coupled in printed["zones"]:
        assert c.boundary(zone) == (typ.lower(), coupled), zone
        assert c.boundary(zone.upper()) == (typ.lower(), coupled), zone


@pytest.mark.parametrize("name", CASES)
def test_case_matches_reference_bitwise(name, tmp_path):
    z, printed = stage(name, tmp_path)
    case, phys, run = E.read_control(os.path.join(str(tmp_path), "case.control"), root=str(tmp_path))
    assert case["K"] == int(z["K"]) and case["F"] == int(z["F"])
    for k in INTS:
        assert np.array_equal(np.asarray(case[k]).reshape(z[k].shape), z[k]), k
    for k in REALS:
        a = np.asarray(case[k]).reshape(z[k].shape)
        assert np.array_equal(a, z[k]), (k, np.abs(a - z[k]).max())
    # the settings the reference printed
    assert "%.3f" % phys["mach"] == printed["mach"]
    if printed["reynolds"]:
        assert "%.3E" % phys["reynolds"] == printed["reynolds"]
        assert "%.3f" % phys["prandtl"] == printed["prandtl"]
    else:
        assert phys["euler"] and phys["reynolds"] == 0.0
    assert phys["inviscid"] == printed["scheme"].lower()
    assert phys["riemann"] == printed["riemann"].lower()
    assert "%.3f" % phys["lambda_stab"] == printed["lambda_stab"]
    if phys["inviscid"] == "split-form":
        assert phys["averaging"] == printed["averaging"].lower()


@pytest.mark.parametrize("name", CASES)
def test_settings_reproduce_reference_time_derivative(name, tmp_path):
    z, _ = stage(name, tmp_path)
    case, phys, run = E.read_control(os.path.join(str(tmp_path), "case.control"), root=str(tmp_path))
    ref = {k: z[k] for k in z.files if k not in ("meta", "printed")}
    ref["K"], ref["F"] = int(z["K"]), int(z["F"])
    o = po.Oracle(ref, po.make_params(phys["mach"], phys["reynolds"], prandtl=phys["prandtl"],
                                      riemann=phys["riemann"], inviscid=phys["inviscid"],
                                      averaging=phys["averaging"], les=None if phys["les"] == "none" else phys["les"],
                                      lambda_stab=phys["lambda_stab"], les_c=phys["les_c"], euler=phys["euler"]))
    qd = o.qdot(z["Q"], float(z["time"]))
    assert po.normwise_error(qd, z["QDot"]).max() <= 1e-12


def test_settings_of_the_syntax_case(tmp_path):
    """What the syntax file states, read by the reference's rules: the first of two "mach number"
    lines (quoted, with a d exponent, a trailing comment), keys in upper case, a value without
    quotes, keys inside a '#define' block left out of the main dictionary (Prandtl number 0.72 by
    default, lambda 0.5 from the main key), split form on Gauss nodes."""
    stage("syntax_split_kg_lxf", tmp_path)
    c = E.Control(os.path.join(str(tmp_path), "case.control"))
    assert c["mach number"] == "2.5d-1"
    assert c["MACH NUMBER"] == "2.5d-1"
    assert c["mesh file name"] == "box.mesh"
    assert c["riemann solver"] == "Lax-Friedrichs"
    assert c["discretization nodes"] == "Gauss-Lobatto"
    assert c["viscous discretization"] == "BR1"          # CheckInputIntegrity default
    assert c.get("prandtl number") is None and c.get("name") is None
    case, phys, run = c.case(str(tmp_path))
    assert phys == dict(mach=0.25, reynolds=1200.0, prandtl=0.72, euler=False, riemann="lax-friedrichs",
                        inviscid="split-form", averaging="kennedy-gruber", lambda_stab=0.5, les="none", les_c=0.0,
                        nodes="gauss-lobatto")
    assert run == dict(cfl=0.3, dcfl=0.25, dt=0.0, tolerance=1e-30, steps=1, output_interval=1, restart=False)


def test_orders_and_les_settings(tmp_path):
    stage("orders_ijk_smagorinsky", tmp_path)
    case, phys, run = E.read_control(os.path.join(str(tmp_path), "case.control"), root=str(tmp_path))
    assert (case["N"] == [2, 4, 3]).all()
    assert phys["les"] == "smagorinsky" and phys["les_c"] == 0.15 and phys["prandtl"] == 0.8


def write(d, text, name="case.control"):
    p = os.path.join(str(d), name)
    with open(p, "w") as fh:
        fh.write(text)
    return p


BASE = '''flow equations = "NS"
mesh file name = "box.mesh"
solution file name = "s.hsol"
restart = .true.
restart file name = "r.hsol"
number of time steps = 10
output interval = 5
convergence tolerance = 1e-10
mach number = 0.1
reynolds number = 1600
dt = 1.0e-3
polynomial order = 2
discretization nodes = "Gauss-Lobatto"
'''
PERIODIC = "".join("#define boundary %s\n  type = periodic\n  coupled boundary = %s\n#end\n" % p
                   for p in specmesh.PERIODIC_PAIRS)


def case_of(d, text):
    specmesh.write_specmesh(os.path.join(str(d), "box.mesh"), 2, 0.0, None)
    return E.read_control(write(d, text), root=str(d))


def test_defaults_and_fixed_dt(tmp_path):
    case, phys, run = case_of(tmp_path, BASE + PERIODIC)
    assert phys == dict(mach=0.1, reynolds=1600.0, prandtl=0.72, euler=False, riemann="roe", inviscid="standard",
                        averaging="standard", lambda_stab=1.0, les="none", les_c=0.0, nodes="gauss-lobatto")
    assert run == dict(cfl=0.0, dcfl=0.0, dt=1e-3, tolerance=1e-10, steps=10, output_interval=5, restart=True)
    assert case["K"] == 8 and (case["N"] == 2).all()


@pytest.mark.parametrize("text,code,msg", [
    (BASE.replace('mesh file name = "box.mesh"\n', ""), 1, "missing entry for keyword: mesh file name"),
    (BASE.replace("reynolds number = 1600\n", ""), 1, "missing entry for keyword: reynolds number"),
    (BASE.replace("output interval = 5\n", ""), 1, "missing entry for keyword: output interval"),
    (BASE.replace("dt = 1.0e-3\n", ""), 1, '"cfl" or "dt"'),
    (BASE.replace("dt = 1.0e-3\n", "cfl = 0.5\n"), 1, '"dcfl"'),
    (BASE.replace("polynomial order = 2\n", ""), 1, "polynomial order(s) must be specified"),
    (BASE.replace("polynomial order = 2\n", "polynomial order = two\n"), 1, "not an integer"),
    (BASE.replace("mach number = 0.1\n", "mach number = 0.1x\n"), 1, "not a number"),
    (BASE + 'riemann solver = "rusanov"\n', 2, "rusanov"),
    (BASE + 'riemann solver = "hllc"\n', 1, "not recognized"),
    (BASE + 'averaging = "chandrasekar"\n', 2, "chandrasekar"),
    (BASE + 'averaging = "mean"\n', 1, "Averaging not recognized"),
    (BASE + 'inviscid discretization = "weno"\n', 1, "weno"),
    (BASE + 'viscous discretization = "IP"\n', 2, "BR1"),
    (BASE + 'les model = "vreman"\n', 2, "vreman"),
    (BASE + 'les model = "dsm"\n', 1, "not implemented"),
    (BASE.replace('"Gauss-Lobatto"', '"gauss-lobatto"'), 1, "Unknown discretization nodes"),
])
def test_errors(tmp_path, text, code, msg):
    with pytest.raises(E.HX3DError) as ei:
        case_of(tmp_path, text + PERIODIC)
    assert msg in str(ei.value) and "(code %d)" % code in str(ei.value)


def test_discretization_nodes(tmp_path):
    """Gauss nodes (the reference's default when the key is absent, main.f90:278-281) reach the
    descriptor and the front end builds the geometry on them: no node on the element boundary;
    a split-form case on Gauss nodes runs on Gauss-Lobatto nodes, as the reference switches it."""
    case, phys, _ = case_of(tmp_path, BASE.replace('"Gauss-Lobatto"', '"Gauss"') + PERIODIC)
    assert phys["nodes"] == "gauss"
    x = np.asarray(case["x"]).reshape(-1, 3)
    assert np.abs(x[:3, 0] + np.pi).min() > 0.05   # element 0's first nodes sit inside it
    case, phys, _ = case_of(tmp_path, BASE.replace('discretization nodes = "Gauss-Lobatto"\n', "") + PERIODIC)
    assert phys["nodes"] == "gauss"
    case, phys, _ = case_of(tmp_path, BASE.replace('discretization nodes = "Gauss-Lobatto"\n', "") +
                            'inviscid discretization = "split-form"\n' + PERIODIC)
    assert phys["nodes"] == "gauss-lobatto" and phys["inviscid"] == "split-form"
    x = np.asarray(case["x"]).reshape(-1, 3)
    assert x[0, 0] == -np.pi


def test_boundary_errors(tmp_path):
    # walls: read, but the engine takes periodic boxes only
    walls = PERIODIC
    for z in ("front", "back"):
        walls = walls.replace("#define boundary %s\n  type = periodic" % z, "#define boundary %s\n  type = noslipwall" % z)
    with pytest.raises(E.HX3DError, match="not periodic"):
        case_of(tmp_path, BASE + walls)
    # a wall whose zone is the coupled zone of a periodic one is paired all the same
    # (ConstructPeriodicFaces goes by the periodic zone's coupled name)
    one = PERIODIC.replace("#define boundary front\n  type = periodic", "#define boundary front\n  type = noslipwall")
    assert case_of(tmp_path, BASE + one)[0]["F"] == 24
    with pytest.raises(E.HX3DError, match="Boundary type slip not recognized"):
        case_of(tmp_path, BASE + PERIODIC.replace("type = periodic", "type = slip", 1))
    with pytest.raises(E.HX3DError, match="Missing boundary condition type for boundary top"):
        case_of(tmp_path, BASE + PERIODIC.replace("#define boundary top\n  type = periodic\n", "#define boundary top\n"))
    with pytest.raises(E.HX3DError, match='"coupled boundary" for boundary left'):
        case_of(tmp_path, BASE + PERIODIC.replace("  coupled boundary = right\n", "", 1))
    # a header indented or a name list with spaces around "__" does not name the zone
    c = E.Control(write(tmp_path, BASE + " #define boundary left\n type = periodic\n#end\n"
                                        "#define boundary top __ bottom\n type = periodic\n coupled boundary = x\n#end\n"))
    with pytest.raises(E.HX3DError, match="Missing boundary condition type for boundary left"):
        c.boundary("left")
    assert c.boundary("top") == ("periodic", "x")
    with pytest.raises(E.HX3DError, match="for boundary bottom"):
        c.boundary("bottom")
    with pytest.raises(E.HX3DError, match="Wrong control file"):
        E.Control(os.path.join(str(tmp_path), "none.control"))


def test_order_file_count_checked(tmp_path):
    specmesh.write_orders(os.path.join(str(tmp_path), "o.txt"), np.full((7, 3), 3))
    with pytest.raises(E.HX3DError, match="given for 7 elements, the mesh has 8"):
        case_of(tmp_path, BASE.replace("polynomial order = 2\n", 'polynomial order file = "o.txt"\n') + PERIODIC)
