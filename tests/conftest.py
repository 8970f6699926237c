import ast
import glob
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "oracle"), os.path.join(REPO, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = sorted(glob.glob(os.path.join(REPO, "tests", "golden", "*.npz")))
# reference runs without CTD_AFTER_STEPS: the final QDot is the last RK stage's
STAGE3 = sorted(glob.glob(os.path.join(REPO, "tests", "golden", "stage3", "*.npz")))

RK_A = (0.0, -5.0 / 9.0, -153.0 / 128.0)
RK_C = (1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0)


def rk3_by_stages(o, Q, G, dt):
    """One Williamson RK3 step written out stage by stage from the oracle's time derivative
    (TakeRK3Step, ExplicitMethods.f90:697-775), Q and G in place; returns the last stage's
    QDot."""
    for a, c in zip(RK_A, RK_C):
        qd = o.qdot(Q)
        G[:] = a * G + qd
        Q[:] = Q + c * dt * G
    return qd


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def load_golden(path):
    z = dict(np.load(path, allow_pickle=False))
    meta = ast.literal_eval(str(z.pop("meta")))
    case = {k: (int(v) if k in ("K", "F") else v) for k, v in z.items()}
    return case, meta


def physics_kwargs(meta):
    """Engine/oracle keyword arguments from a fixture's control settings."""
    les = meta.get("les")
    return dict(mach=meta["mach"], reynolds=meta["re"], riemann=meta.get("riemann", "roe"),
                inviscid=meta.get("inviscid", "standard"),
                averaging=meta.get("averaging", "standard").lower(),
                les=les.lower() if les else None, les_c=meta.get("les_intensity"),
                nodes=meta.get("nodes", "Gauss-Lobatto").lower())


@pytest.fixture(params=GOLDEN, ids=lambda p: os.path.basename(p)[:-4])
def golden(request):
    return load_golden(request.param)


def reference_initial_state(case, z):
    """The reference run's exact initial state for a case whose node coordinates are the
    reference's: mesh.taylor_green plus the fixture's table of the entries the Fortran runtime's
    sines and cosines round the other way."""
    from horses3d_amd import mesh
    Q0 = mesh.taylor_green(case, 0.1)
    flat = Q0.reshape(-1)
    flat[z["ic_fix_index"]] = z["ic_fix_value"]
    return Q0
