"""Mesh front end (hx3d_mesh_read_specmesh) against the reference.

Each golden fixture holds the topology and geometry the reference solver built from a SpecMesh
file (Read_SpecMesh.f90, HexMesh.f90, MappedGeometry.f90): element faces and sides, face
elements, slots and rotations, node coordinates, metrics Ja, Jacobians, face normals, tangents
and surface Jacobians.  The same SpecMesh file, written again from the fixture's parameters,
read by the C front end must give the same integers and the same doubles bit for bit (straight,
curved with face-patch projection, and p-anisotropic meshes with mortar faces; geometry on
Gauss-Lobatto and on Gauss nodes).  Host only."""
import ast
import os

import numpy as np
import pytest

from horses3d_amd import engine as E
from horses3d_amd import mesh, specmesh

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(f[:-4] for f in os.listdir(GOLD) if f.endswith(".npz"))
INTS = ("N", "faceID", "faceSide", "fElem", "fESide", "fRot")
REALS = ("x", "Ja", "J", "normal", "t1", "t2", "fjac")


def write_fixture_mesh(name, d):
    z = np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False)
    meta = ast.literal_eval(str(z["meta"]))
    path = os.path.join(str(d), "box.mesh")
    # the fixture generator writes bFaceOrder = max(4, N)
    specmesh.write_specmesh(path, meta["n"], meta["amp"], mesh.box_phases(meta["phase_seed"]),
                            bface_order=max(4, meta["order"]))
    mixed = len({tuple(r) for r in z["N"]}) > 1
    return z, meta, path, (np.asarray(z["N"]) if mixed else None)


@pytest.mark.parametrize("name", FIXTURES)
def test_front_end_matches_reference_bitwise(name, tmp_path):
    z, meta, path, orders = write_fixture_mesh(name, tmp_path)
    c = E.read_specmesh(path, N=meta["order"], orders=orders, periodic=dict(specmesh.PERIODIC_PAIRS),
                        nodes=meta.get("nodes", "Gauss-Lobatto"))
    assert c["K"] == int(z["K"]) and c["F"] == int(z["F"])
    for k in INTS:
        assert np.array_equal(np.asarray(c[k]).reshape(z[k].shape), z[k]), k
    for k in REALS:
        a = np.asarray(c[k]).reshape(z[k].shape)
        assert np.array_equal(a, z[k]), (k, np.abs(a - z[k]).max())


def test_orders_file_round_trip(tmp_path):
    o = mesh.random_orders(27, 2, 6, 5)
    p = os.path.join(str(tmp_path), "orders.txt")
    specmesh.write_orders(p, o)
    assert np.array_equal(specmesh.read_orders(p), o)


def test_front_end_errors(tmp_path):
    p = os.path.join(str(tmp_path), "box.mesh")
    specmesh.write_specmesh(p, 2, 0.0, None)
    with pytest.raises(E.HX3DError, match="cannot open"):
        E.read_specmesh(os.path.join(str(tmp_path), "missing.mesh"), N=3)
    # a boundary that is not periodic: the engine takes interior and periodic faces only
    with pytest.raises(E.HX3DError, match="not periodic"):
        E.read_specmesh(p, N=3, periodic={"front": "back", "back": "front"})
    with pytest.raises(E.HX3DError, match="not found"):
        E.read_specmesh(p, N=3, periodic={"front": "nowhere"})
    with pytest.raises(E.HX3DError, match="orders"):
        E.read_specmesh(p, N=9, periodic=dict(specmesh.PERIODIC_PAIRS))
    bad = os.path.join(str(tmp_path), "bad.mesh")
    with open(bad, "w") as fh:
        fh.write("8 1 1\n0 0 0\n")
    with pytest.raises(E.HX3DError, match="truncated"):
        E.read_specmesh(bad, N=3)


def test_box_front_end_agrees_with_analytic_box(tmp_path):
    """A straight box read from file has the connectivity counts, node coordinates and constant
    metrics of the analytic box builder (to round-off: the conservative curl form leaves
    O(1e-15) off-diagonal metric terms)."""
    p = os.path.join(str(tmp_path), "box.mesh")
    specmesh.write_specmesh(p, 3, 0.0, None)
    c = E.read_specmesh(p, N=4, periodic=dict(specmesh.PERIODIC_PAIRS))
    b = mesh.periodic_box(3, 4)
    assert c["K"] == b["K"] and c["F"] == b["F"]
    for k in ("x", "Ja", "J"):
        assert np.abs(c[k] - b[k]).max() <= 1e-13 * np.abs(b[k]).max(), k
