"""Oracle BLAS / reduction / layout functions against the reference unit
tests' semantics (CPU only).

The reference's BLAS tests compare strided per-RHS kernels with CPU loops on
data drawn as ``Complex(rand()%16, rand()%16+1)`` (tests/public_complex_vector.h:
30-35), R=3 and n=1024^2+1 (tests/test_blas/test_caxpbypcz.cu:12-45,64-66) with
|diff|^2 <= 1e-12.  The same shapes and data range are used here, with exact
integer-valued expectations computed independently by numpy on a reshaped
[n][R] view.
"""
import numpy as np

from oracle import oracle
from qcu_mrhs_amd import fields


def _ref_data(shape, g):
    return g.integers(0, 16, shape).astype(np.float64) + 1j * (g.integers(0, 16, shape) + 1)


def test_strided_blas_reference_semantics():
    g = np.random.default_rng(12)
    R, n = 3, 1024 * 1024 + 1
    x, y, z = (_ref_data(n * R, g) for _ in range(3))
    a, b, c = (_ref_data(R, g) for _ in range(3))
    X, Y, Z = x.reshape(n, R), y.reshape(n, R), z.reshape(n, R)
    np.testing.assert_array_equal(oracle.cxsay(x, a, y, R).reshape(n, R), X - a * Y)
    np.testing.assert_array_equal(oracle.cxpay(x, a, y, R).reshape(n, R), X + a * Y)
    np.testing.assert_array_equal(oracle.caxpby(a, x, b, y, R).reshape(n, R), a * X + b * Y)
    np.testing.assert_array_equal(oracle.caxpbypcz(a, x, b, y, c, z, R).reshape(n, R),
                                  a * X + b * Y + c * Z)


def test_dot_and_norm_reference_semantics():
    # tests/test_inner_prod/test_inner_prod.cu: R=8 over 16*16*16*32*12 entries
    g = np.random.default_rng(4)
    R, n = 8, 16 * 16 * 16 * 32 * 12 // 8
    a = _ref_data(n * R, g)
    b = _ref_data(n * R, g)
    A, B = a.reshape(n, R), b.reshape(n, R)
    expect = (np.conj(A) * B).sum(axis=0)          # integers < 2^53: exact
    np.testing.assert_array_equal(oracle.dotc(a, b, R), expect)
    # tests/test_norm/test_norm.cu:47-68: per-RHS sqrt of sum |a|^2
    nrm = oracle.norm(a, R)
    np.testing.assert_allclose(nrm, np.sqrt((np.abs(A) ** 2).sum(axis=0)), rtol=1e-15)
    # conjugate-linearity in the first argument
    s = 2.0 - 3.0j
    np.testing.assert_allclose(oracle.dotc(s * a, b, R), np.conj(s) * expect, rtol=1e-15)


def test_gather_identity_and_round_trip():
    # tests/test_gather/test_gather.cu:32-85: 4^4, Nc=3, R=1 gather is the identity
    L, Nc = (4, 4, 4, 4), 3
    Vh = fields.half_volume(L)
    v = np.arange(Vh * 4 * Nc, dtype=np.float64).astype(np.complex128)
    np.testing.assert_array_equal(oracle.gather([v], Nc), v)
    g = np.random.default_rng(1)
    R = 5
    vecs = g.standard_normal((R, Vh * 4 * Nc)) + 1j * g.standard_normal((R, Vh * 4 * Nc))
    m = oracle.gather(list(vecs), Nc)
    np.testing.assert_array_equal(m, fields.to_mrhs(vecs, Nc))
    np.testing.assert_array_equal(oracle.scatter(m, Nc, R), vecs)


def test_gauge_eo_round_trip():
    """Parity unpinned (the reference's converter is in the absent qcu_io):
    round trip and agreement with the even/odd site convention only."""
    L, Nc = (4, 2, 6, 4), 3
    g = np.random.default_rng(2)
    V = int(np.prod(L))
    lex = g.standard_normal((4, V, Nc, Nc)) + 1j * g.standard_normal((4, V, Nc, Nc))
    eo = oracle.gauge_eo(lex, L, Nc)
    np.testing.assert_array_equal(eo, fields.gauge_lex_to_eo(lex, L, Nc).reshape(-1))
    np.testing.assert_array_equal(oracle.gauge_eo_reverse(eo, L, Nc), lex.reshape(-1))
    # site (x,y,z,t) = (1,0,0,0) is odd, checkerboard index 0
    np.testing.assert_array_equal(eo.reshape(4, 2, V // 2, Nc, Nc)[2, 1, 0], lex[2, 1])
