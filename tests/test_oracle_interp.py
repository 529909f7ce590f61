"""Pin the oracle's generic transfers (G2P / P2G value, gradient and
divergence, Picasso_ParticleInterpolation.hpp:110-279) by mathematics that
fixes them independently of its formulas:

- G2P reproduces linear fields exactly (B-splines of every order have
  linear precision): the value is f(x_p), the gradient the constant slope,
  the divergence its trace;
- P2G conserves the zeroth and first moments of what it spreads;
- each P2G operation is the adjoint (transpose) of a G2P one:
  <P2G value(v), f> = <v, G2P value(f)>, <P2G gradient(s), F> =
  <s, G2P divergence(F)>, <P2G divergence(V), g> = <V, G2P gradient(g)>,
  and the tensor divergence against the vector gradient component by
  component.  With the G2P operations pinned, this fixes the P2G ones.
"""
import numpy as np
import pytest

from oracle import orpic

LOW = np.array([-0.25, -0.5, 0.125])
H = 0.125
LO = np.array([0, 0, 0], np.int32)
N = np.array([16, 15, 17], np.int32)
VALUE, GRADIENT, DIVERGENCE = orpic.INTERP_VALUE, orpic.INTERP_GRADIENT, orpic.INTERP_DIVERGENCE
ENTITIES = [orpic.NODE, orpic.CELL, 2, 7]  # Node, Cell, Face<I>, Edge<K>


def _particles(seed, n=40):
    rng = np.random.default_rng(seed)
    return LOW + H * rng.uniform(5.0, 10.0, size=(n, 3)), rng


def _node_x(entity):
    """Physical node positions of the entity's lattice over the window."""
    off = [orpic.lib().orc_entity_offset(entity, d, H) for d in range(3)]
    ax = [LOW[d] + off[d] + H * (LO[d] + np.arange(N[d])) for d in range(3)]
    return np.stack(np.meshgrid(*ax, indexing="ij"), -1)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("entity", ENTITIES)
def test_g2p_reproduces_linear_fields(order, entity):
    x, rng = _particles(order * 31 + entity)
    xn = _node_x(entity)
    a = rng.normal(size=3)
    B = rng.normal(size=(3, 3))
    f = a + xn @ B.T  # f_c(x) = a_c + sum_d B[c, d] x_d
    val = orpic.g2p_interp(VALUE, order, entity, LOW, H, LO, N, x, f, 3)
    np.testing.assert_allclose(val, a + x @ B.T, rtol=0, atol=1e-12)
    grad = orpic.g2p_interp(GRADIENT, order, entity, LOW, H, LO, N, x, f, 3)
    np.testing.assert_allclose(grad, np.broadcast_to(B.reshape(9), grad.shape), atol=1e-10)
    div = orpic.g2p_interp(DIVERGENCE, order, entity, LOW, H, LO, N, x, f, 3)
    np.testing.assert_allclose(div[:, 0], np.trace(B), atol=1e-10)
    s = f[..., :1]
    np.testing.assert_allclose(orpic.g2p_interp(VALUE, order, entity, LOW, H, LO, N, x, s, 1),
                               (a[0] + x @ B[0])[:, None], atol=1e-12)
    g1 = orpic.g2p_interp(GRADIENT, order, entity, LOW, H, LO, N, x, s, 1)
    np.testing.assert_allclose(g1, np.broadcast_to(B[0], g1.shape), atol=1e-10)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("entity", ENTITIES)
def test_p2g_moments(order, entity):
    x, rng = _particles(order * 17 + entity + 100)
    xn = _node_x(entity).reshape(-1, 3)
    v = rng.normal(size=(len(x), 3))
    f = orpic.p2g_interp(VALUE, order, entity, LOW, H, LO, N, x, v, 3).reshape(-1, 3)
    np.testing.assert_allclose(f.sum(0), v.sum(0), atol=1e-12)
    np.testing.assert_allclose(xn.T @ f, x.T @ v, atol=1e-11)          # sum_n x_n f_n
    s = v[:, :1]
    g = orpic.p2g_interp(GRADIENT, order, entity, LOW, H, LO, N, x, s, 1).reshape(-1, 3)
    np.testing.assert_allclose(g.sum(0), 0.0, atol=1e-10)
    np.testing.assert_allclose(xn.T @ g, s.sum() * np.eye(3), atol=1e-10)  # sum_n x_n g_n^T
    dv = orpic.p2g_interp(DIVERGENCE, order, entity, LOW, H, LO, N, x, v, 3).reshape(-1)
    assert abs(dv.sum()) < 1e-10
    np.testing.assert_allclose(xn.T @ dv, v.sum(0), atol=1e-10)
    T = rng.normal(size=(len(x), 9))
    dt = orpic.p2g_interp(DIVERGENCE, order, entity, LOW, H, LO, N, x, T, 9).reshape(-1, 3)
    np.testing.assert_allclose(xn.T @ dt, T.sum(0).reshape(3, 3).T, atol=1e-10)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("entity", ENTITIES)
def test_p2g_is_adjoint_of_g2p(order, entity):
    x, rng = _particles(order * 7 + entity + 200)
    shape = tuple(int(k) for k in N)
    # value (vector)
    v = rng.normal(size=(len(x), 3))
    F3 = rng.normal(size=shape + (3,))
    lhs = np.sum(orpic.p2g_interp(VALUE, order, entity, LOW, H, LO, N, x, v, 3) * F3)
    rhs = np.sum(v * orpic.g2p_interp(VALUE, order, entity, LOW, H, LO, N, x, F3, 3))
    assert abs(lhs - rhs) < 1e-11 * max(1.0, abs(rhs))
    # P2G gradient of a scalar <-> G2P divergence of a vector
    s = rng.normal(size=(len(x), 1))
    lhs = np.sum(orpic.p2g_interp(GRADIENT, order, entity, LOW, H, LO, N, x, s, 1) * F3)
    rhs = np.sum(s[:, 0] * orpic.g2p_interp(DIVERGENCE, order, entity, LOW, H, LO, N, x, F3,
                                            3)[:, 0])
    assert abs(lhs - rhs) < 1e-9 * max(1.0, abs(rhs))
    # P2G divergence of a vector <-> G2P gradient of a scalar
    F1 = rng.normal(size=shape + (1,))
    lhs = np.sum(orpic.p2g_interp(DIVERGENCE, order, entity, LOW, H, LO, N, x, v, 3) * F1)
    rhs = np.sum(v * orpic.g2p_interp(GRADIENT, order, entity, LOW, H, LO, N, x, F1, 1))
    assert abs(lhs - rhs) < 1e-9 * max(1.0, abs(rhs))
    # P2G divergence of a tensor <-> G2P gradient of a vector (component 3i + j)
    T = rng.normal(size=(len(x), 9))
    lhs = np.sum(orpic.p2g_interp(DIVERGENCE, order, entity, LOW, H, LO, N, x, T, 9) * F3)
    rhs = np.sum(T * orpic.g2p_interp(GRADIENT, order, entity, LOW, H, LO, N, x, F3, 3))
    assert abs(lhs - rhs) < 1e-9 * max(1.0, abs(rhs))


def test_interp_rejects_bad_shapes():
    x, _ = _particles(1, 3)
    f = np.zeros(tuple(int(k) for k in N) + (2,))
    with pytest.raises(orpic.OracleError):
        orpic.g2p_interp(DIVERGENCE, 2, 0, LOW, H, LO, N, x, f, 2)
    with pytest.raises(orpic.OracleError):
        orpic.p2g_interp(GRADIENT, 2, 0, LOW, H, LO, N, x, np.zeros((3, 3)), 3)
