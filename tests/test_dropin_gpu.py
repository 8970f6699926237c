"""The drop-in inside the reference solver: HORSES3D's own set-up and time loop with its RK3 step
taken by the hx3d engine through the Fortran adapter (oracle/_ref/bin/hx3d_dropin.ns, built by
oracle/ref/build_dropin.py from the reference build, horses3d_amd/fortran and
tests/fortran/hx3d_dropin.f90), against the reference's own run of the same case
(tests/golden/dropin, make_dropin.py):
  * every step's time (the reference's MaxTimeStep on the host, from the state the adapter
    leaves) to 1e-12;
  * every step's residuals (ComputeMaxResiduals of the QDot the adapter leaves: the last RK
    stage's, or the new state's with CTD_AFTER_STEPS) to 1e-9 relative: the engine's QDot
    matches per stage to 1e-12, and the states drift apart by rounding over the steps;
  * the kinetic-energy monitor after every step to 1e-10 relative.
The CPU test pins the fixtures' first residual row against the oracle's QDot of the initial
state."""
import os
import subprocess

import numpy as np
import pytest

import make_dropin as md
import refcase as rc
from conftest import REPO

EXE = os.path.join(REPO, "oracle", "_ref", "bin", "hx3d_dropin.ns")
FIX = os.path.join(REPO, "tests", "golden", "dropin")


def _fixture(name):
    return np.load(os.path.join(FIX, name + ".npz"), allow_pickle=False)


@pytest.mark.parametrize("name", sorted(md.CASES))
def test_fixture_initial_residuals_match_oracle(name, tmp_path):
    """Row 0 of the reference's residual file is max |QDot| of the initial state: the oracle's
    QDot of the same Taylor-Green state on the same SpecMesh gives it to 1e-12."""
    import pyoracle as po
    from horses3d_amd import mesh
    from horses3d_amd.engine import read_specmesh
    spec = md.CASES[name]
    md.write_case(str(tmp_path), spec)
    from horses3d_amd import specmesh
    ctl = spec["ctl"]
    nodes = ctl.get("nodes", "Gauss-Lobatto").lower()
    case = read_specmesh(str(tmp_path / "box.mesh"), N=spec["N"], periodic=dict(specmesh.PERIODIC_PAIRS), nodes=nodes)
    n1 = spec["N"] + 1
    case["fNf"] = np.full((case["F"], 2), spec["N"], dtype=np.int32)   # one order: every face of order N
    case["foff"] = np.arange(case["F"] + 1, dtype=np.int64) * n1 * n1
    o = po.Oracle(case, po.make_params(0.1, 1600.0, riemann=ctl["riemann"], inviscid=ctl["inviscid"],
                                       averaging=ctl["averaging"].lower(), nodes=nodes))
    Q0 = mesh.taylor_green(case, 0.1)
    got = np.abs(o.qdot(Q0)).max(axis=0)
    ref = _fixture(name)["residuals"][0, 2:]
    assert np.all(np.abs(got / ref - 1.0) < 1e-12), (got, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(md.CASES))
def test_dropin_in_reference_time_loop(name, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("drop-in driver not built (oracle/ref/build_dropin.py needs the reference build)")
    spec = md.CASES[name]
    z = _fixture(name)
    md.write_case(str(tmp_path), spec)
    print("drop-in run", name, flush=True)
    env = dict(os.environ, OMP_NUM_THREADS="1", OMP_SCHEDULE="static")
    r = subprocess.run([EXE, "case.control"], cwd=str(tmp_path), capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "hx3d_dropin: RK3 steps on the hx3d engine" in r.stdout
    res, ke = md.read_residuals(str(tmp_path)), rc.read_volume_monitor(str(tmp_path), "kinen")
    ref_res, ref_ke = z["residuals"], z["kinen"]
    assert res.shape == ref_res.shape and ke.shape == ref_ke.shape
    assert np.array_equal(res[:, 0], ref_res[:, 0])
    t_err = np.abs(res[1:, 1] / ref_res[1:, 1] - 1.0).max()
    r_err = np.abs(res[:, 2:] / ref_res[:, 2:] - 1.0).max(axis=0)
    k_err = np.abs(ke[:, 2] / ref_ke[:, 2] - 1.0).max()
    print("time", t_err, "residuals", r_err, "kinetic energy", k_err, flush=True)
    assert t_err < 1e-12 and np.all(r_err < 1e-9) and k_err < 1e-10
