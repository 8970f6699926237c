"""The disjoint union of independent cases (horses3d_amd/batch.py): numbering offsets and the
consistency of the joined connectivity.  Host only; the GPU form is in test_engine_gpu.py."""
import numpy as np

from horses3d_amd import batch, mesh


def test_union_numbering():
    cases = [mesh.periodic_box(2, 3, amplitude=0.1, phase_seed=1), mesh.straight_box(3, 3),
             mesh.periodic_box_mixed(2, mesh.random_orders(8, 2, 4, 1))]
    u, slices = batch.union(cases)
    assert u["K"] == sum(c["K"] for c in cases) and u["F"] == sum(c["F"] for c in cases)
    assert u["off"][-1] == sum(int(c["off"][-1]) for c in cases) == len(u["J"]) == len(u["Ja"])
    assert len(u["normal"]) == sum(len(c["normal"]) for c in cases)
    # every element face slot names a face whose side lists that element and slot
    for e in range(u["K"]):
        for s in range(6):
            f, side = u["faceID"][e, s], u["faceSide"][e, s]
            assert u["fElem"][f, side - 1] == e and u["fESide"][f, side - 1] == s + 1
    # no face joins two cases: both elements of a face lie in one case's element range
    eoff = np.cumsum([0] + [c["K"] for c in cases])
    case_of = np.searchsorted(eoff, np.arange(u["K"]), side="right") - 1
    assert np.all(case_of[u["fElem"][:, 0]] == case_of[u["fElem"][:, 1]])
    # each slice holds that case's nodes
    for c, s in zip(cases, slices):
        assert np.array_equal(u["J"][s], c["J"]) and np.array_equal(u["Ja"][s], c["Ja"])
