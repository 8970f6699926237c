"""SpecMesh files (the mesh format of Solver/src/libs/mesh/Read_SpecMesh.f90:36-348) for the
synthetic periodic boxes of the benchmarks and tests, and the polynomial-order file
(Solver/src/libs/io/ReadInputFile.f90:132-153).

A box written here is read back by the C front end (``engine.read_specmesh``, which follows the
reference's reader and geometry) or by the reference solver itself; its boundary zones are
periodic in pairs (``PERIODIC_PAIRS``, the .control file's "coupled boundary" entries).

Format: a header "nNodes nElements bFaceOrder"; the nodes "x y z"; per element its 8 corner node
IDs (1-based, the reference's hex8 corner order), 6 face flags (1 = curved), for each curved face
(bFaceOrder+1)^2 points at the Chebyshev-Gauss-Lobatto points of the face (u fastest), and the 6
boundary names ("---" for interior faces).
"""
import numpy as np

# hex8 corner order of the reference (HexElementConnectivityDefinitions.f90:64-77): corner c
# sits at reference coordinates CORNER_XI[c] in {-1,1}^3.
CORNER_XI = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
# Face k (1..6 = FRONT, BACK, BOTTOM, RIGHT, TOP, LEFT): which reference axis is fixed (and
# where) and which axes the patch coordinates (u, v) run along.
FACE_FIXED = [(1, -1.0), (1, 1.0), (2, -1.0), (0, 1.0), (2, 1.0), (0, -1.0)]
FACE_UV = [(0, 2), (0, 2), (0, 1), (1, 2), (0, 1), (1, 2)]
FACE_NAMES_BOX = ["front", "back", "bottom", "right", "top", "left"]
PERIODIC_PAIRS = [("front", "back"), ("back", "front"), ("bottom", "top"), ("top", "bottom"),
                  ("left", "right"), ("right", "left")]

L = np.pi   # the box is [-L, L]^3


def _wrap(c):
    """Canonical periodic representative of a box coordinate in [-pi, pi]."""
    c = np.asarray(c, dtype=float)
    return np.where(c > L - 1e-12, c - 2.0 * L, c)


def box_map(amp, phases):
    """x' = x + amp * prod_k sin(x_k + phase[d, k]) per coordinate d (periodic-safe)."""
    ph = np.asarray(phases, dtype=float).reshape(3, 3)

    def f(X):
        X = np.asarray(X, dtype=float)
        W = _wrap(X)
        out = X.copy()
        if amp != 0.0:
            for d in range(3):
                out[..., d] = X[..., d] + amp * (np.sin(W[..., 0] + ph[d, 0])
                                                 * np.sin(W[..., 1] + ph[d, 1])
                                                 * np.sin(W[..., 2] + ph[d, 2]))
        return out
    return f


def write_specmesh(path, n, amp=0.0, phases=None, bface_order=4):
    """Periodic box [-pi,pi]^3 split into n^3 hexes; element e = i + n*(j + n*k)."""
    phases = np.zeros(9) if phases is None else phases
    fmap = box_map(amp, phases)
    g = np.linspace(-L, L, n + 1)
    curved = amp != 0.0
    nid = lambda i, j, k: 1 + i + (n + 1) * (j + (n + 1) * k)
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1)   # (i,j,k,3)
    nodes = fmap(P)
    lines = ["%d %d %d" % ((n + 1) ** 3, n ** 3, bface_order if curved else 1)]
    for k in range(n + 1):
        for j in range(n + 1):
            for i in range(n + 1):
                lines.append("%.17e %.17e %.17e" % tuple(nodes[i, j, k]))
    cgl = -np.cos(np.arange(bface_order + 1) * np.pi / bface_order)
    h = 2 * L / n
    for k in range(n):
        for j in range(n):
            for i in range(n):
                ids = []
                for c in range(8):
                    a, b, cc = ((CORNER_XI[c] + 1) / 2).astype(int)
                    ids.append(nid(i + a, j + b, k + cc))
                lines.append(" ".join(str(v) for v in ids))
                lines.append(" ".join(["1" if curved else "0"] * 6))
                if curved:
                    lo = np.array([g[i], g[j], g[k]])
                    for f in range(6):
                        ax, val = FACE_FIXED[f]
                        ua, va = FACE_UV[f]
                        for vv in cgl:
                            for uu in cgl:
                                xi = np.zeros(3)
                                xi[ax] = val
                                xi[ua] = uu
                                xi[va] = vv
                                X = lo + (xi + 1.0) * 0.5 * h
                                lines.append("%.17e %.17e %.17e" % tuple(fmap(X)))
                names = []
                for f in range(6):
                    ax, val = FACE_FIXED[f]
                    idx = (i, j, k)[ax]
                    on_bnd = (val < 0 and idx == 0) or (val > 0 and idx == n - 1)
                    names.append(FACE_NAMES_BOX[f] if on_bnd else "---")
                lines.append(" ".join(names))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_orders(path, orders):
    """The polynomial-order file: the element count, then "Nx Ny Nz" per element."""
    with open(path, "w") as fh:
        fh.write("%d\n" % len(orders))
        for o in orders:
            fh.write("%d %d %d\n" % tuple(o))


def read_orders(path):
    with open(path) as fh:
        K = int(fh.readline().split()[0])
        return np.array([[int(v) for v in fh.readline().split()[:3]] for _ in range(K)], dtype=np.int32)
