"""Pins of the oracle's full operator, Schur operator and BiCGStab (CPU only)."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import dense, oracle
from qcu_mrhs_amd import fields


def _dense_M(gauge, L, Nc, kappa):
    D = dense.hopping_matrix(gauge, L, Nc)
    return sp.identity(D.shape[0], dtype=np.complex128, format="csr") - kappa * D


def test_mat_matches_dense():
    L, Nc, R, kappa = (4, 4, 4, 4), 3, 2, 0.12
    gauge = fields.gauge_field(L, Nc, 5, kind="haar")
    M = _dense_M(gauge, L, Nc, kappa)
    b = fields.gaussian_sources(L, Nc, R, 6)
    out = fields.from_mrhs(oracle.mat(fields.to_mrhs(b, Nc), gauge, L, Nc, R, kappa), R)
    for r in range(R):
        np.testing.assert_allclose(out[r], M @ b[r], rtol=0, atol=1e-12)


def test_schur_is_dense_schur_complement():
    L, Nc, R, kappa = (4, 4, 2, 4), 3, 1, 0.13
    gauge = fields.gauge_field(L, Nc, 2, kind="haar")
    M = _dense_M(gauge, L, Nc, kappa).toarray()
    n = M.shape[0] // 2
    Mee, Meo, Moe, Moo = M[:n, :n], M[:n, n:], M[n:, :n], M[n:, n:]
    np.testing.assert_allclose(Mee, np.eye(n))
    S = Moo - Moe @ np.linalg.solve(Mee, Meo)
    v = fields.gaussian_sources(L, Nc, 1, 1)[0][:n]
    np.testing.assert_allclose(oracle.schur(v, gauge, L, Nc, R, kappa), S @ v, rtol=0, atol=1e-12)


@pytest.mark.parametrize("Nc,kind,kappa", [(3, "haar", 0.12), (2, "smooth", 0.13), (4, "haar", 0.11)])
def test_bicgstab_vs_dense_solve(Nc, kind, kappa):
    L, R, tol = (4, 4, 4, 4), 3, 1e-10
    gauge = fields.gauge_field(L, Nc, 31, kind=kind, eps=0.3)
    M = _dense_M(gauge, L, Nc, kappa).tocsc()
    b = fields.gaussian_sources(L, Nc, R, 32)
    x, ok, iters, res = oracle.bicgstab(fields.to_mrhs(b, Nc), gauge, L, Nc, R, kappa, 1000, tol)
    assert ok and 0 < iters < 1000 and np.all(res <= tol)
    xs = fields.from_mrhs(x, R)
    lu = sp.linalg.splu(M)
    for r in range(R):
        true_res = np.linalg.norm(b[r] - M @ xs[r]) / np.linalg.norm(b[r])
        assert true_res < 50 * tol
        xd = lu.solve(b[r])
        assert np.linalg.norm(xs[r] - xd) / np.linalg.norm(xd) < 1e-8


def test_bicgstab_small_config_true_residual():
    """8^3x8, SU(3) Haar, R=4, kappa=0.12, tol=1e-10 (the small plain case)."""
    L, Nc, R, kappa, tol = (8, 8, 8, 8), 3, 4, 0.12, 1e-10
    gauge = fields.gauge_field(L, Nc, 1001, kind="haar")
    b = fields.to_mrhs(fields.gaussian_sources(L, Nc, R, 1002), Nc)
    x, ok, iters, res = oracle.bicgstab(b, gauge, L, Nc, R, kappa, 2000, tol)
    assert ok
    r = b - oracle.mat(x, gauge, L, Nc, R, kappa)
    true_rel = oracle.norm(r, R) / oracle.norm(b, R)
    assert np.all(true_rel < 20 * tol), true_rel


def test_bicgstab_iteration_cap_reports_failure():
    L, Nc, R = (4, 4, 4, 4), 3, 2
    gauge = fields.gauge_field(L, Nc, 3, kind="haar")
    b = fields.to_mrhs(fields.gaussian_sources(L, Nc, R, 4), Nc)
    x, ok, iters, res = oracle.bicgstab(b, gauge, L, Nc, R, 0.12, 3, 1e-12)
    assert not ok and iters == 3 and np.all(res > 1e-12)


def _textbook_bicgstab(A, b, tol, max_iter):
    """van der Vorst (1992) BiCGStab, shadow residual r0 = b, x0 = 0,
    written from the published algorithm (not from the reference)."""
    x = np.zeros_like(b)
    r = b.copy()
    rhat = b.copy()
    rho_old = alpha = omega = 1.0
    v = np.zeros_like(b)
    p = np.zeros_like(b)
    nb = np.linalg.norm(b)
    for i in range(1, max_iter + 1):
        rho = np.vdot(rhat, r)
        beta = (rho / rho_old) * (alpha / omega)
        p = r + beta * (p - omega * v)
        v = A @ p
        alpha = rho / np.vdot(rhat, v)
        s = r - alpha * v
        t = A @ s
        omega = np.vdot(t, s) / np.vdot(t, t)
        x = x + alpha * p + omega * s
        r = s - omega * t
        if np.linalg.norm(r) / nb <= tol:
            return x, i
        rho_old = rho
    return x, max_iter


@pytest.mark.parametrize("kappa", [0.11, 0.125])
def test_bicgstab_recurrence_matches_textbook(kappa):
    """Same iterates as the textbook recurrence: iteration count within 1 and
    the odd-parity solution agrees, per RHS (a wrong beta/omega/alpha step
    still converges but changes the iteration count)."""
    L, Nc, R, tol = (4, 4, 4, 4), 3, 2, 1e-11
    gauge = fields.gauge_field(L, Nc, 77, kind="haar")
    M = _dense_M(gauge, L, Nc, kappa).toarray()
    n = M.shape[0] // 2
    S = M[n:, n:] - M[n:, :n] @ M[:n, n:]          # M_ee = 1
    b = fields.gaussian_sources(L, Nc, R, 78)
    x, ok, iters, res = oracle.bicgstab(fields.to_mrhs(b, Nc), gauge, L, Nc, R, kappa, 500, tol)
    xs = fields.from_mrhs(x, R)
    counts = []
    for r in range(R):
        bp = b[r][n:] - M[n:, :n] @ b[r][:n]         # b_o + kappa D_oe b_e
        xo, it_r = _textbook_bicgstab(S, bp, tol, 500)
        counts.append(it_r)
        np.testing.assert_allclose(xs[r][n:], xo, rtol=0, atol=1e-8 * np.abs(xo).max())
    # all-RHS stopping rule: the joint solve stops at the slowest RHS
    assert abs(iters - max(counts)) <= 1, (iters, counts)
