"""Batching: many independent cases stepped by one engine.

`union(cases)` joins several case dicts (same physics; any orders) into one case whose mesh is
their disjoint union: element, face and node numbers are offset case by case and no face joins
two cases.  One engine then runs every case in each launch, so a stream of small cases fills
the GPU that each one alone leaves mostly idle.

Every element and face value the kernels compute depends only on its own case's arrays, so with
a step size that does not depend on the other cases -- a fixed dt, `Engine.advance_dt` -- each
case's slice of the batched state is the state it reaches alone, bit for bit when the cases
share one order N (the same kernels run), and within the parity tolerance otherwise (the batch
then runs the mixed-order kernels).  With a CFL-based dt the batch would take the smallest dt of
all its cases, which is not what each case takes alone; use one engine per case there
(`serve.py`)."""
import numpy as np

from . import engine as E

_PER_ELEMENT = ("N", "faceSide")
_PER_NODE = ("Ja", "J", "x")
_PER_FACE_NODE = ("normal", "t1", "t2", "fjac")


def union(cases):
    """(union case dict, node slices): the cases' meshes side by side.  slices[i] selects case
    i's nodes in the union's state arrays."""
    if not cases:
        raise ValueError("no cases to batch")
    K = F = nodes = 0
    parts = {k: [] for k in ("N", "faceID", "faceSide", "fElem", "fESide", "fRot") + _PER_NODE + _PER_FACE_NODE}
    slices = []
    face_geo = all(all(k in c for k in _PER_FACE_NODE) for c in cases)
    for c in cases:
        n = int(c["off"][-1]) if "off" in c else len(c["J"])
        parts["N"].append(np.asarray(c["N"], dtype=np.int32).reshape(-1, 3))
        parts["faceID"].append(np.asarray(c["faceID"], dtype=np.int32).reshape(-1, 6) + F)
        parts["faceSide"].append(np.asarray(c["faceSide"], dtype=np.int32).reshape(-1, 6))
        parts["fElem"].append(np.asarray(c["fElem"], dtype=np.int32).reshape(-1, 2) + K)
        parts["fESide"].append(np.asarray(c["fESide"], dtype=np.int32).reshape(-1, 2))
        parts["fRot"].append(np.asarray(c["fRot"], dtype=np.int32).reshape(-1))
        for k in _PER_NODE:
            if k in c:
                parts[k].append(np.asarray(c[k]))
        if face_geo:
            for k in _PER_FACE_NODE:
                parts[k].append(np.asarray(c[k]))
        slices.append(slice(nodes, nodes + n))
        K += int(c["K"])
        F += int(c["F"])
        nodes += n
    u = {"K": K, "F": F}
    for k, v in parts.items():
        if v and len(v) == len(cases):
            u[k] = np.ascontiguousarray(np.concatenate(v))
    Nk = u["N"]
    u["off"] = np.concatenate([[0], np.cumsum(np.prod(Nk + 1, axis=1))]).astype(np.int64)
    return u, slices


class Batch:
    """One engine over the disjoint union of independent cases (same physics keywords)."""

    def __init__(self, cases, **physics):
        self.case, self.slices = union(cases)
        self.engine = E.Engine(self.case, **physics)
        self.nodes = self.engine.nodes

    def set_states(self, states):
        Q = np.empty((self.nodes, 5))
        for s, q in zip(self.slices, states):
            Q[s] = q
        self.engine.set_state(Q)

    def get_states(self):
        Q = self.engine.get_state()
        return [Q[s].copy() for s in self.slices]

    def advance_dt(self, nsteps, dt, t=None):
        return self.engine.advance_dt(nsteps, dt, t)

    def close(self):
        self.engine.close()
