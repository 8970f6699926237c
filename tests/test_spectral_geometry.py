"""Closed-form pins for the spectral operators (NodalStorageClass.f90:170-337,
InterpolationMatrices.f90:42-108) and for the case builder's geometry
(MappedGeometry.f90:174-386), in the oracle and in the product's host code."""
import numpy as np
import numpy.polynomial.legendre as npl
import pytest

import pyoracle as po
from horses3d_amd import mesh
from conftest import load_golden, GOLDEN

ORDERS = [1, 2, 3, 4, 5, 6, 7]


@pytest.mark.parametrize("N", ORDERS)
def test_gl_nodes_weights_closed_form(N):
    # GL nodes: -1, 1 and the roots of P_N'; weights 2 / (N (N+1) P_N(x)^2)
    cN = np.zeros(N + 1); cN[N] = 1.0
    roots = np.sort(npl.legroots(npl.legder(cN))) if N > 1 else np.array([])
    x_ref = np.concatenate([[-1.0], roots, [1.0]])
    w_ref = 2.0 / (N * (N + 1) * npl.legval(x_ref, cN) ** 2)
    for x, w in (mesh.gl_nodes(N), (po.nodal_gl(N)["x"], po.nodal_gl(N)["w"])):
        assert np.allclose(x, x_ref, atol=4e-15, rtol=0)
        assert np.allclose(w, w_ref, atol=4e-15, rtol=0)
        assert abs(w.sum() - 2.0) < 1e-15


@pytest.mark.parametrize("N", ORDERS)
def test_derivative_matrices(N):
    op = po.nodal_gl(N)
    x, w, D = op["x"], op["w"], op["D"]
    # exact on polynomials of degree <= N
    for p in range(N + 1):
        assert np.allclose(D @ x ** p, p * x ** max(p - 1, 0) if p else 0 * x, atol=1e-12)
    # summation by parts: W D + D^T W = B = diag(-1, 0, ..., 0, 1)
    W = np.diag(w)
    B = np.zeros_like(D); B[0, 0], B[-1, -1] = -1.0, 1.0
    assert np.allclose(W @ D + D.T @ W, B, atol=1e-13)
    # hatD = W^-1 D^T W, sharpD = 2D + e0 e0/w0 - eN eN/wN
    assert np.allclose(op["hatD"], np.diag(1 / w) @ D.T @ W, atol=1e-13)
    S = 2 * D; S[0, 0] += 1 / w[0]; S[-1, -1] -= 1 / w[-1]
    assert np.allclose(op["sharpD"], S, atol=1e-13)
    # GL boundary interpolation vectors are exact unit vectors, b = v / w
    assert np.array_equal(op["vL"], np.eye(N + 1)[0]) and np.array_equal(op["vR"], np.eye(N + 1)[-1])
    assert np.allclose(op["bL"], op["vL"] / w) and np.allclose(op["bR"], op["vR"] / w)
    assert np.allclose(mesh.deriv_matrix(x), D, atol=1e-13)


@pytest.mark.parametrize("N", ORDERS)
def test_gauss_nodes_and_operators_closed_form(N):
    """Gauss(-Legendre) NodalStorage (GaussLegendreNodesAndWeights, LegendreAlgorithms.f90:137-214):
    nodes the roots of P_{N+1}, weights 2 / ((1 - x^2) P'_{N+1}(x)^2), the rule exact for degree
    2N + 1; D exact on degree <= N, hatD = W^-1 D^T W; the boundary vectors the Lagrange basis at
    -1 and +1 (exact on degree <= N, dense: no node on the boundary), b = v / w; the same nodes in
    the product's host code (through the front end's geometry pins)."""
    op = po.nodal_gauss(N)
    x, w, D = op["x"], op["w"], op["D"]
    c = np.zeros(N + 2); c[N + 1] = 1.0
    x_ref = np.sort(npl.legroots(c))
    w_ref = 2.0 / ((1.0 - x_ref ** 2) * npl.legval(x_ref, npl.legder(c)) ** 2)
    assert np.allclose(x, x_ref, atol=4e-15, rtol=0) and np.allclose(w, w_ref, atol=4e-15, rtol=0)
    assert np.array_equal(x, -x[::-1]) and np.array_equal(w, w[::-1])
    for p in range(2 * N + 2):
        exact = 2.0 / (p + 1) if p % 2 == 0 else 0.0
        assert abs(w @ x ** p - exact) < 2e-14, p
    for p in range(N + 1):
        assert np.allclose(D @ x ** p, p * x ** max(p - 1, 0) if p else 0 * x, atol=1e-12)
        assert abs(op["vL"] @ x ** p - (-1.0) ** p) < 1e-13 and abs(op["vR"] @ x ** p - 1.0) < 1e-13
    assert np.allclose(op["hatD"], np.diag(1 / w) @ D.T @ np.diag(w), atol=1e-13)
    assert np.count_nonzero(op["vL"]) == N + 1 and np.count_nonzero(op["vR"]) == N + 1
    assert np.allclose(op["bL"], op["vL"] / w, rtol=1e-15) and np.allclose(op["bR"], op["vR"] / w, rtol=1e-15)
    # summation by parts on Gauss nodes holds with the boundary vectors: W D + D^T W = vR vR^T - vL vL^T
    W = np.diag(w)
    assert np.allclose(W @ D + D.T @ W, np.outer(op["vR"], op["vR"]) - np.outer(op["vL"], op["vL"]), atol=1e-12)


@pytest.mark.parametrize("N,M", [(2, 5), (4, 7), (5, 3), (7, 2)])
def test_tset_gauss_interpolation_and_projection(N, M):
    """Mortar matrices on Gauss nodes (InterpolationMatrix_Construct on the case's nodes)."""
    T = po.tset(N, M, nodes="gauss")
    gN, gM = po.nodal_gauss(N), po.nodal_gauss(M)
    xN, wN, xM, wM = gN["x"], gN["w"], gM["x"], gM["w"]
    if N <= M:
        for p in range(N + 1):
            assert np.allclose(T @ xN ** p, xM ** p, atol=1e-13)
    else:
        Tin = po.tset(M, N, nodes="gauss")
        assert np.allclose(T, (Tin * wN[:, None]).T / wM[:, None], atol=1e-13)
        for p in range(M + 1):
            assert abs(wM @ (T @ xN ** p) - wN @ xN ** p) < 1e-12
    assert not np.allclose(T, po.tset(N, M))   # not the Gauss-Lobatto matrix


def test_oracle_rejects_split_form_on_gauss_nodes():
    case, _ = load_golden([p for p in GOLDEN if "g1_tgv" in p][0])
    with pytest.raises(ValueError, match="Gauss-Lobatto"):
        po.Oracle(case, po.make_params(0.1, 1600.0, inviscid="split-form", averaging="pirozzoli", nodes="gauss"))


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_chebyshev_gauss_lobatto_operators(N):
    """The Chebyshev-Gauss-Lobatto framework of NodalStorage (NodalStorageClass.f90:282-296):
    nodes -cos(j pi / N), a derivative matrix exact on polynomials of degree <= N, and an
    interpolation to the GL nodes exact on the same polynomials."""
    op = po.nodal_cgl(N)
    x, D, T = op["xcgl"], op["DCGL"], op["Tcheb2gauss"]
    assert np.allclose(x, -np.cos(np.arange(N + 1) * np.pi / N), atol=1e-15, rtol=0)
    assert x[0] == -1.0 and abs(x[-1] - 1.0) < 1e-15
    xg = po.nodal_gl(N)["x"]
    for p in range(N + 1):
        dref = p * x ** (p - 1) if p else 0 * x
        assert np.allclose(D @ x ** p, dref, atol=1e-12), p
        assert np.allclose(T @ x ** p, xg ** p, atol=1e-13), p
    # one degree more is not reproduced where the two node sets differ (N >= 3; for N <= 2 the
    # CGL and GL nodes coincide and T is the identity)
    if N >= 3:
        assert not np.allclose(T @ x ** (N + 1), xg ** (N + 1), atol=1e-6)
    else:
        assert np.allclose(T, np.eye(N + 1), atol=1e-15)
    # rows of an interpolation sum to one, rows of a derivative to zero
    assert np.allclose(T.sum(axis=1), 1.0, atol=1e-14)
    assert np.allclose(D.sum(axis=1), 0.0, atol=1e-12)


@pytest.mark.parametrize("N,M", [(2, 5), (3, 4), (4, 7), (5, 3), (7, 2), (6, 6)])
def test_tset_interpolation_and_projection(N, M):
    T = po.tset(N, M)
    xN = po.nodal_gl(N)["x"]; wN = po.nodal_gl(N)["w"]
    xM = po.nodal_gl(M)["x"]; wM = po.nodal_gl(M)["w"]
    if N <= M:   # interpolation: exact on degree <= N
        for p in range(N + 1):
            assert np.allclose(T @ xN ** p, xM ** p, atol=1e-13)
    else:        # projection: transpose of the M->N interpolation, weighted w_N(j)/w_M(i)
        Tin = po.tset(M, N)
        assert np.allclose(T, (Tin * wN[:, None]).T / wM[:, None], atol=1e-13)
        # preserves the quadrature of (projection) * degree <= M polynomials
        for p in range(M + 1):
            assert abs(wM @ (T @ xN ** p) - wN @ xN ** p) < 1e-12


def test_box_geometry_matches_reference_straight():
    """Straight 4^3 box, N=3: the host case builder reproduces the reference's own
    metrics (conservative form on the hex8 map) and node coordinates."""
    path = [p for p in GOLDEN if "c1_tgv" in p][0]
    ref, _ = load_golden(path)
    case = mesh.periodic_box(4, 3)
    # both sides carry round-off from the CGL curl construction (reference: ~1e-15 noise on
    # exactly constant metrics)
    assert np.allclose(case["x"], ref["x"], atol=1e-14, rtol=0)
    assert np.allclose(case["J"], ref["J"], atol=1e-13, rtol=0)
    assert np.allclose(case["Ja"], ref["Ja"], atol=1e-13, rtol=0)


@pytest.mark.parametrize("n,N,amp", [(3, 3, 0.1), (2, 5, 0.15), (2, 4, 0.0)])
def test_box_geometry_invariants(n, N, amp):
    """Discrete metric identities (free-stream preservation) and the box volume."""
    case = mesh.periodic_box(n, N, amplitude=amp, phase_seed=5)
    K, n1 = case["K"], N + 1
    D = mesh.deriv_matrix(mesh.gl_nodes(N)[0])
    Ja = case["Ja"].reshape(K, n1, n1, n1, 3, 3)   # [e, k, j, i, d, c]
    div = (np.einsum("il,ekjlc->ekjic", D, Ja[..., 0, :]) + np.einsum("jl,eklic->ekjic", D, Ja[..., 1, :])
           + np.einsum("kl,eljic->ekjic", D, Ja[..., 2, :]))
    assert np.abs(div).max() < 1e-12
    w = mesh.gl_nodes(N)[1]
    W = np.einsum("k,j,i->kji", w, w, w).ravel()
    vol = (case["J"].reshape(K, -1) * W).sum()
    # a periodic deformation maps the box onto a fundamental cell of the same lattice, so the
    # exact volume is (2 pi)^3; the quadrature of the curved map is exact only when straight
    assert abs(vol / (2 * np.pi) ** 3 - 1.0) < (1e-12 if amp == 0 else 1e-5)


def test_free_stream_preservation_oracle():
    """A uniform state has QDot = 0 to round-off on a curved periodic mesh."""
    case = mesh.periodic_box(3, 4, amplitude=0.15, phase_seed=3)
    Q = np.tile([1.0, 0.3, -0.2, 0.1, 1.0 / (1.4 * 0.3 ** 2 * 0.4) + 0.07], (mesh.ndof(case), 1))
    for inv, avg, rie in [("standard", "standard", "roe"), ("split-form", "pirozzoli", "lxf"),
                          ("split-form", "kennedy-gruber", "roe"), ("split-form", "ducros", "lxf")]:
        o = po.Oracle(case, po.make_params(0.3, 1000.0, riemann=rie, inviscid=inv, averaging=avg))
        qd = o.qdot(Q)
        # round-off scale: p/(gamma-1) ~ 20 times |D| ~ 10 times 1/J
        assert np.abs(qd).max() < 1e-10, (inv, avg, np.abs(qd).max())
