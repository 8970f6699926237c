"""Case builders: periodic hexahedral boxes with the reference's storage contract.

A case is a dict of numpy arrays laid out as the engine's C ABI expects (include/hx3d.h):
element orders, the element->face and face->element connectivity of HexMesh
(Solver/src/libs/mesh/HexMesh.f90:71-180; faces as in FaceClass.f90:63-101), the metric terms
Ja, J at every node, and the face geometry of the left element (normal, t1, t2, |Ja|), which
the oracle consumes.

Geometry follows the reference's conservative ("invariant curl") metric construction
(computeMetricTermsConservativeForm, MappedGeometry.f90:174-386): the mapping and its
gradient are sampled at Chebyshev-Gauss-Lobatto points, Ja^i = -1/2 curl(X_l grad X_m -
X_m grad X_l) is formed there and interpolated to the Gauss-Lobatto nodes; J likewise.
The mapping is the periodic box [-pi, pi]^3 deformed by
    x'_d = x_d + A * prod_k sin(x_k + phi[d, k]),
which is 2*pi-periodic, so every periodic face pair matches.
"""
import numpy as np

L = np.pi

# element face slots 1..6 = EFRONT, EBACK, EBOTTOM, ERIGHT, ETOP, ELEFT (MeshTypes.f90:17-18)
EFRONT, EBACK, EBOTTOM, ERIGHT, ETOP, ELEFT = 1, 2, 3, 4, 5, 6
SLOT_AXIS = {EFRONT: 1, EBACK: 1, EBOTTOM: 2, ERIGHT: 0, ETOP: 2, ELEFT: 0}
SLOT_PLUS = {EFRONT: 0, EBACK: 1, EBOTTOM: 0, ERIGHT: 1, ETOP: 1, ELEFT: 0}
SLOT_UV = {EFRONT: (0, 2), EBACK: (0, 2), EBOTTOM: (0, 1), ERIGHT: (1, 2), ETOP: (0, 1), ELEFT: (1, 2)}


# ---------------------------------------------------------------------------------------
# 1-D operators (restated from libs/spectral: LegendreAlgorithms.f90:275-355,
# InterpolationAndDerivatives.f90:230-262, 413-461, 799-834)
# ---------------------------------------------------------------------------------------
def gl_nodes(N):
    eps4 = 4.0 * np.finfo(float).eps
    x = np.zeros(N + 1)
    w = np.zeros(N + 1)

    def q_and_L(xx):
        a2, da2, a1, da1 = 1.0, 0.0, xx, 1.0
        for k in range(2, N + 1):
            Lk = ((2 * k - 1) * xx * a1 - (k - 1) * a2) / k
            dLk = da2 + (2 * k - 1) * a1
            a2, a1, da2, da1 = a1, Lk, da1, dLk
        k = N + 1
        Lk = ((2 * k - 1) * xx * a1 - (k - 1) * a2) / k
        dLk = da2 + (2 * k - 1) * a1
        return Lk - a2, dLk - da2, a1

    if N == 1:
        return np.array([-1.0, 1.0]), np.array([1.0, 1.0])
    x[0], x[N] = -1.0, 1.0
    w[0] = w[N] = 2.0 / (N * (N + 1))
    for j in range(1, (N + 1) // 2):
        xj = -np.cos((j + 0.25) * np.pi / N - 3.0 / (8 * N * np.pi * (j + 0.25)))
        for _ in range(11):
            q, dq, _LN = q_and_L(xj)
            d = -q / dq
            xj += d
            if abs(d) <= eps4 * abs(xj):
                break
        q, dq, LN = q_and_L(xj)
        x[j], w[j] = xj, 2.0 / (N * (N + 1) * LN * LN)
        x[N - j], w[N - j] = -xj, w[j]
    if N % 2 == 0:
        a2, a1 = 1.0, 0.0
        for k in range(2, N + 1):
            a2, a1 = a1, ((2 * k - 1) * 0.0 * a1 - (k - 1) * a2) / k
        x[N // 2], w[N // 2] = 0.0, 2.0 / (N * (N + 1) * a1 * a1)
    return x, w


def bary_weights(x):
    n = len(x)
    wb = np.ones(n)
    for j in range(1, n):
        for k in range(j):
            wb[k] *= x[k] - x[j]
            wb[j] *= x[j] - x[k]
    return 1.0 / wb


def deriv_matrix(x):
    wb = bary_weights(x)
    n = len(x)
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if j != i:
                D[i, j] = wb[j] / (wb[i] * (x[i] - x[j]))
                D[i, i] -= D[i, j]
    return D


def interp_matrix(xo, xn):
    wb = bary_weights(xo)
    T = np.zeros((len(xn), len(xo)))
    tol = 2.0 * np.finfo(float).eps
    for k, xk in enumerate(xn):
        hit = False
        for j, xj in enumerate(xo):
            a, b = xk, xj
            same = abs(a - b) <= tol if (a == 0.0 or b == 0.0) else abs(b - a) <= tol * max(abs(a), abs(b))
            if same:
                T[k, j] = 1.0
                hit = True
        if not hit:
            t = wb / (xk - xo)
            T[k] = t / t.sum()
    return T


# ---------------------------------------------------------------------------------------
# mapping
# ---------------------------------------------------------------------------------------
def box_phases(seed):
    if not seed:
        return np.zeros(9)
    return np.random.default_rng(seed).uniform(0.0, 2.0 * np.pi, 9)


def _map_and_grad(X0, h, amp, ph):
    """Physical coordinates and d x / d xi of the deformed box at base points X0 (..., 3)."""
    ph = np.asarray(ph).reshape(3, 3)
    x = X0.copy()
    G = np.zeros(X0.shape[:-1] + (3, 3))
    for m in range(3):
        G[..., m, m] = 0.5 * h
    if amp != 0.0:
        s = [[np.sin(X0[..., k] + ph[d, k]) for k in range(3)] for d in range(3)]
        c = [[np.cos(X0[..., k] + ph[d, k]) for k in range(3)] for d in range(3)]
        for d in range(3):
            x[..., d] = X0[..., d] + amp * s[d][0] * s[d][1] * s[d][2]
            for m in range(3):
                o = [k for k in range(3) if k != m]
                G[..., d, m] += 0.5 * h * amp * c[d][m] * s[d][o[0]] * s[d][o[1]]
    return x, G


def _along(M, A, axis):
    """M applied along one axis of A: out[.., i, ..] = sum_l M[i, l] A[.., l, ..] (BLAS)."""
    return np.moveaxis(np.tensordot(M, A, axes=(1, axis)), 0, axis)


def periodic_box(n, N, amplitude=0.0, phase_seed=0, phases=None):
    """Periodic box [-pi,pi]^3 of n^3 hexahedra of order N (element e = i + n*(j + n*k))."""
    if phases is None:
        phases = box_phases(phase_seed)
    K = n ** 3
    n1 = N + 1
    ne = n1 ** 3
    h = 2.0 * L / n
    x, w = gl_nodes(N)
    D = deriv_matrix(x)
    xc = -np.cos(np.arange(n1) * np.pi / N)
    Dc = deriv_matrix(xc)
    T = interp_matrix(xc, x)
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    lo = np.stack([(-L + h * ii).ravel(order="F"), (-L + h * jj).ravel(order="F"),
                   (-L + h * kk).ravel(order="F")], axis=-1)           # (K,3), e = i + n(j + n k)

    def base(pts):
        # pts (n1,n1,n1) reference points per axis -> (K, n1, n1, n1, 3) base coordinates, [e, i, j, k]
        P = np.stack(np.meshgrid(pts, pts, pts, indexing="ij"), axis=-1)
        return lo[:, None, None, None, :] + (P[None] + 1.0) * 0.5 * h

    # node coordinates (transfiniteMapAt at the GL nodes)
    Xn, _ = _map_and_grad(base(x), h, amplitude, phases)
    # conservative metrics at CGL points.  As in the reference, the element map is a
    # polynomial (there: the transfinite blend of polynomial face patches, TransfiniteMaps3D.f90);
    # here: the interpolant of the deformation at the CGL points, whose gradient is taken with
    # the CGL derivative matrix.  Tensor-product derivatives commute, so the curl form gives
    # watertight face metrics and exact discrete metric identities.
    Xc, _ = _map_and_grad(base(xc), h, amplitude, phases)
    Gc = np.stack([_along(Dc, Xc, 1), _along(Dc, Xc, 2), _along(Dc, Xc, 3)], axis=-1)   # [..., d, m] = d x_d / d xi_m

    def dcgl(P):
        # P (K, i, j, k, 3 comps over xi_m) -> aux[..., m, l] = d P_m / d xi_l
        a, b, c = _along(Dc, P, 1), _along(Dc, P, 2), _along(Dc, P, 3)
        return np.stack([a, b, c], axis=-1)

    Ja = np.zeros(Xc.shape[:-1] + (3, 3))    # [..., d (xi,eta,zeta), c (x,y,z)]
    for comp, (a_, b_) in enumerate([(2, 1), (0, 2), (1, 0)]):
        # coordsProduct = x_a grad x_b - x_b grad x_a (MappedGeometry.f90:199-203, 241-245, 283-287)
        P = Xc[..., a_, None] * Gc[..., b_, :] - Xc[..., b_, None] * Gc[..., a_, :]
        aux = dcgl(P)
        Ja[..., 0, comp] = -0.5 * (aux[..., 2, 1] - aux[..., 1, 2])
        Ja[..., 1, comp] = -0.5 * (aux[..., 0, 2] - aux[..., 2, 0])
        Ja[..., 2, comp] = -0.5 * (aux[..., 1, 0] - aux[..., 0, 1])
    Jc = np.einsum("...i,...i->...", Gc[..., :, 0], np.cross(Gc[..., :, 1], Gc[..., :, 2]))
    # CGL -> GL interpolation, one direction at a time
    Ja = _along(T, _along(T, _along(T, Ja, 1), 2), 3)
    J = _along(T, _along(T, _along(T, Jc, 1), 2), 3)
    # node-major packing: g = e*ne + i + n1*(j + n1*k)
    pack = lambda A: np.ascontiguousarray(A.transpose(0, 3, 2, 1, *range(4, A.ndim)).reshape(K * ne, *A.shape[4:]))
    return _box_case(n, N, pack(Xn), pack(Ja).reshape(K * ne, 9), pack(J).reshape(K * ne), D, amplitude, phases)


def straight_box(n, N):
    """Straight periodic box [-pi,pi]^3 of n^3 cubes of order N, with the exact constant metrics
    Ja = (h/2)^2 I, J = (h/2)^3 (no curl-form computation: fast at any size)."""
    K, n1 = n ** 3, N + 1
    ne = n1 ** 3
    h = 2.0 * L / n
    x, _ = gl_nodes(N)
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    lo = np.stack([(-L + h * ii).ravel(order="F"), (-L + h * jj).ravel(order="F"),
                   (-L + h * kk).ravel(order="F")], axis=-1)
    # node g = e*ne + i + n1*(j + n1*k): x from i, y from j, z from k
    P = np.stack(np.meshgrid(x, x, x, indexing="ij"), axis=-1).transpose(2, 1, 0, 3).reshape(ne, 3)
    X = (lo[:, None, :] + (P[None] + 1.0) * 0.5 * h).reshape(K * ne, 3)
    Ja = np.zeros((K * ne, 9))
    Ja[:, [0, 4, 8]] = (0.5 * h) ** 2
    J = np.full(K * ne, (0.5 * h) ** 3)
    return _box_case(n, N, X, Ja, J, deriv_matrix(x), 0.0, np.zeros(9))


def _box_case(n, N, Xp, Jap, Jp, D, amplitude, phases):
    """Case dict of an n^3 periodic box from packed node coordinates and metrics."""
    K, n1 = n ** 3, N + 1
    ne = n1 ** 3
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    # connectivity: face 3e+a joins element e's +axis face to its neighbour's -axis face (rotation 0)
    eid = lambda a, b, c: (a % n) + n * ((b % n) + n * (c % n))
    e_i, e_j, e_k = ii.ravel(order="F"), jj.ravel(order="F"), kk.ravel(order="F")
    E = np.arange(K)
    faceID = np.zeros((K, 6), np.int32)
    faceSide = np.zeros((K, 6), np.int32)
    fElem = np.zeros((3 * K, 2), np.int32)
    fESide = np.zeros((3 * K, 2), np.int32)
    for a, (plus, minus) in enumerate([(ERIGHT, ELEFT), (EBACK, EFRONT), (ETOP, EBOTTOM)]):
        nb = eid(e_i + (a == 0), e_j + (a == 1), e_k + (a == 2))
        f = 3 * E + a
        fElem[f, 0], fElem[f, 1] = E, nb
        fESide[f, 0], fESide[f, 1] = plus, minus
        faceID[E, plus - 1], faceSide[E, plus - 1] = f, 1
        faceID[nb, minus - 1], faceSide[nb, minus - 1] = f, 2
    F = 3 * K
    fRot = np.zeros(F, np.int32)
    case = dict(K=K, N=np.full((K, 3), N, np.int32), off=np.arange(K + 1, dtype=np.int64) * ne,
                faceID=faceID, faceSide=faceSide, F=F, fElem=fElem, fESide=fESide, fRot=fRot,
                fNf=np.full((F, 2), N, np.int32), foff=np.arange(F + 1, dtype=np.int64) * n1 * n1,
                Ja=Jap, J=Jp, x=Xp, n=n, order=N, amplitude=amplitude, phases=np.asarray(phases))
    case.update(face_geometry(case, D))
    return case


def random_orders(K, lo, hi, seed):
    """Per-element orders (Nx, Ny, Nz) drawn uniformly on {lo..hi} from a seed, shape (K, 3)."""
    return np.random.default_rng(seed).integers(lo, hi + 1, (K, 3)).astype(np.int32)


def periodic_box_mixed(n, orders):
    """Straight periodic box [-pi,pi]^3 of n^3 hexahedra with per-element orders
    orders[e] = (Nx, Ny, Nz) (e = i + n*(j + n*k)).  Faces join element e's +axis face to its
    neighbour's -axis face with rotation 0; a face between mismatched orders has the larger
    order in each direction (Face_LinkWithElements, FaceClass.f90:187-278), its frame
    n = +e_axis, t1 = e_u, t2 = n x t1 and |Ja| = (h/2)^2 at every face node.  Metrics of the
    affine map: Ja^d_c = (h/2)^2 delta_dc, J = (h/2)^3."""
    orders = np.asarray(orders, dtype=np.int32).reshape(-1, 3)
    K = n ** 3
    assert orders.shape[0] == K
    h = 2.0 * L / n
    ne = np.prod(orders + 1, axis=1)
    off = np.concatenate([[0], np.cumsum(ne)]).astype(np.int64)
    nodes = int(off[-1])
    X = np.zeros((nodes, 3))
    Ja = np.zeros((nodes, 9))
    for c in range(3):
        Ja[:, 4 * c] = (0.5 * h) ** 2
    J = np.full(nodes, (0.5 * h) ** 3)
    E = np.arange(K)
    ei, ej, ek = E % n, (E // n) % n, E // (n * n)
    lo = np.stack([-L + h * ei, -L + h * ej, -L + h * ek], axis=1)
    for e in range(K):
        Nx, Ny, Nz = orders[e]
        xs = [gl_nodes(int(m))[0] for m in (Nx, Ny, Nz)]
        P = np.stack(np.meshgrid(*xs, indexing="ij"), axis=-1)          # [i, j, k, 3]
        pts = lo[e] + (P + 1.0) * 0.5 * h
        X[off[e]:off[e + 1]] = pts.transpose(2, 1, 0, 3).reshape(-1, 3)  # i fastest
    eid = lambda a, b, c: (a % n) + n * ((b % n) + n * (c % n))
    faceID = np.zeros((K, 6), np.int32)
    faceSide = np.zeros((K, 6), np.int32)
    F = 3 * K
    fElem = np.zeros((F, 2), np.int32)
    fESide = np.zeros((F, 2), np.int32)
    fNf = np.zeros((F, 2), np.int32)
    uv = {0: (1, 2), 1: (0, 2), 2: (0, 1)}      # face coordinates (a, b) of an axis's faces
    for a, (plus, minus) in enumerate([(ERIGHT, ELEFT), (EBACK, EFRONT), (ETOP, EBOTTOM)]):
        nb = eid(ei + (a == 0), ej + (a == 1), ek + (a == 2))
        f = 3 * E + a
        fElem[f, 0], fElem[f, 1] = E, nb
        fESide[f, 0], fESide[f, 1] = plus, minus
        faceID[E, plus - 1], faceSide[E, plus - 1] = f, 1
        faceID[nb, minus - 1], faceSide[nb, minus - 1] = f, 2
        u, v = uv[a]
        fNf[f, 0] = np.maximum(orders[E, u], orders[nb, u])
        fNf[f, 1] = np.maximum(orders[E, v], orders[nb, v])
    fn = np.prod(fNf + 1, axis=1)
    foff = np.concatenate([[0], np.cumsum(fn)]).astype(np.int64)
    FS = int(foff[-1])
    normal = np.zeros((FS, 3))
    t1 = np.zeros((FS, 3))
    for f in range(F):
        a = f % 3
        normal[foff[f]:foff[f + 1], a] = 1.0
        t1[foff[f]:foff[f + 1], uv[a][0]] = 1.0
    t2 = np.cross(normal, t1)
    fjac = np.full(FS, (0.5 * h) ** 2)
    return dict(K=K, N=orders, off=off, faceID=faceID, faceSide=faceSide, F=F, fElem=fElem,
                fESide=fESide, fRot=np.zeros(F, np.int32), fNf=fNf, foff=foff, Ja=Ja, J=J, x=X,
                normal=normal, t1=t1, t2=t2, fjac=fjac, n=n, amplitude=0.0)


def element_face_nodes(N, slot):
    """Node indices n = i + (N+1)(j + (N+1)k) of an element face slot, ordered (a + (N+1) b)."""
    n1 = N + 1
    ax, plus, (ua, va) = SLOT_AXIS[slot], SLOT_PLUS[slot], SLOT_UV[slot]
    out = np.zeros(n1 * n1, np.int64)
    for b in range(n1):
        for a in range(n1):
            idx = [0, 0, 0]
            idx[ax] = N if plus else 0
            idx[ua], idx[va] = a, b
            out[a + n1 * b] = idx[0] + n1 * (idx[1] + n1 * idx[2])
    return out


def face_geometry(case, D=None):
    """Face normal, |Ja|, t1, t2 from the left element (ConstructMappedGeometryFace,
    MappedGeometry.f90:482-760): GL traces are the boundary nodes, dS = +-Ja^d there."""
    N = int(case["N"][0, 0])
    n1 = N + 1
    ne = n1 ** 3
    if D is None:
        D = deriv_matrix(gl_nodes(N)[0])
    F = case["F"]
    eL = case["fElem"][:, 0]
    sL = case["fESide"][:, 0]
    normal = np.zeros((F, n1 * n1, 3))
    t1 = np.zeros_like(normal)
    for slot in range(1, 7):
        m = sL == slot
        if not m.any():
            continue
        idx = element_face_nodes(N, slot)
        g = eL[m, None] * ne + idx[None, :]
        ax = SLOT_AXIS[slot]
        dS = case["Ja"][g][..., 3 * ax:3 * ax + 3]
        if not SLOT_PLUS[slot]:
            dS = -dS
        normal[m] = dS
        xf = case["x"][g].reshape(-1, n1, n1, 3)                 # [f, b, a, 3]
        t1[m] = np.einsum("il,fblc->fbic", D, xf).reshape(-1, n1 * n1, 3)
    jac = np.linalg.norm(normal, axis=-1)
    normal = normal / jac[..., None]
    t1 = t1 - np.sum(t1 * normal, axis=-1, keepdims=True) * normal
    t1 = t1 / np.linalg.norm(t1, axis=-1, keepdims=True)
    t2 = np.cross(normal, t1)
    return dict(normal=normal.reshape(-1, 3), t1=t1.reshape(-1, 3), t2=t2.reshape(-1, 3),
                fjac=jac.reshape(-1))


def taylor_green(case, mach, phase=(0.0, 0.0, 0.0), gamma=1.4, noise=0.0, seed=0):
    """Taylor-Green vortex state (nodes, 5): rho = 1, u = sin x cos y cos z,
    v = -cos x sin y cos z, w = 0, p = 1/(gamma M^2) + (cos 2x + cos 2y)(cos 2z + 2)/16;
    optional Gaussian velocity noise of the given amplitude."""
    X = case["x"] + np.asarray(phase)[None, :]
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    if noise:
        r = np.random.default_rng(seed)
        u = u + noise * r.standard_normal(u.shape)
        v = v + noise * r.standard_normal(u.shape)
        w = w + noise * r.standard_normal(u.shape)
    # 1/(gamma M^2) with gamma M^2 formed as gamma * (M * M), as the reference forms it
    # (PhysicsStorage_NS.f90:250) for its initial conditions
    p = 1.0 / (gamma * (mach * mach)) + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0) / 16.0
    rho = np.ones_like(u)
    Q = np.stack([rho, rho * u, rho * v, rho * w, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)], axis=1)
    return np.ascontiguousarray(Q)


def ndof(case):
    return int(case["off"][-1])
