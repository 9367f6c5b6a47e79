"""Krylov drivers (Arnoldi + DGKS, implicit restarts, BiCGStab(L), damped
Newton) on dense host problems, where numpy's direct solvers pin the answers."""
import numpy as np
import pytest

from ns3dperiodic_amd import krylov

OPS = krylov.NumpyOps()


def _mv(A):
    def mv(x, y):
        y[...] = A @ x
        return y
    return mv


def test_arnoldi_factorisation():
    rng = np.random.default_rng(1)
    n, m = 120, 20
    A = rng.standard_normal((n, n))
    V = OPS.zeros(n, m)
    H = np.zeros((m, m))
    f = rng.standard_normal(n)
    assert krylov.arnoldi_extend(OPS, _mv(A), V, H, f, 0, m) == 0
    assert np.allclose(V @ V.T, np.eye(m), atol=1e-13)
    em = np.zeros(m)
    em[-1] = 1
    assert np.allclose(A @ V.T, V.T @ H + np.outer(f, em), atol=1e-11)
    assert np.allclose(np.tril(H, -2), 0)


@pytest.mark.parametrize("which", ["LM", "LR"])
def test_eigs_matches_dense(which):
    rng = np.random.default_rng(2)
    n = 300
    # nonsymmetric, with a few well separated eigenvalues (real and complex) on top of a bulk
    A = rng.standard_normal((n, n)) / np.sqrt(n)
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    D = np.zeros((n, n))
    D[:2, :2] = [[3.0, 1.5], [-1.5, 3.0]]      # 3 +- 1.5i
    D[2, 2], D[3, 3], D[4, 4], D[5, 5] = 2.6, -2.8, 2.2, 1.9
    A = A + Qm @ D @ Qm.T
    out = krylov.eigs(OPS, _mv(A), n, k=6, m=30, which=which, tol=1e-12)
    ref = np.linalg.eigvals(A)
    ref = ref[krylov._order(ref, which)][:6]
    got = out["values"]
    for lam in ref:
        assert np.min(np.abs(got - lam)) < 1e-9 * max(1, abs(lam)), (lam, got)
    assert np.all(out["residuals"] < 1e-8)


@pytest.mark.parametrize("L", [1, 2, 4, 6])
def test_bicgstab_l_solves(L):
    rng = np.random.default_rng(3)
    n = 200
    A = 3.0 * np.eye(n) + rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    x = np.zeros(n)
    flag, it, res = krylov.bicgstab_l(OPS, _mv(A), b, x, L=L, tol=1e-12)
    assert flag == 0 and res < 1e-11
    assert np.allclose(x, np.linalg.solve(A, b), atol=1e-10)


def test_newton_nonlinear_system():
    rng = np.random.default_rng(4)
    n = 60
    A = 4.0 * np.eye(n) + rng.standard_normal((n, n)) / np.sqrt(n)
    xs = rng.standard_normal(n)
    b = A @ xs + 0.1 * xs ** 3          # F(x) = A x + 0.1 x^3 - b, root xs

    def rhs(x, out):
        out[...] = A @ x + 0.1 * x ** 3 - b
        return out

    def jac(x):
        d = 0.3 * x ** 2
        x0 = x.copy()

        def mv(v, y):
            y[...] = A @ v + 0.3 * x0 ** 2 * v
            return y
        return mv

    x = np.zeros(n)
    out = krylov.newton(OPS, rhs, jac, x, tol=1e-11, max_iter=100)
    assert out["converged"] and out["iterations"] >= 3
    assert np.allclose(x, xs, atol=1e-10)
    # damping 0.8: the residual falls at least geometrically
    h = out["history"]
    assert h[-1] < 1e-11 and h[1] < h[0]


def test_shift_application_keeps_hessenberg_similarity():
    rng = np.random.default_rng(5)
    m = 12
    H = np.triu(rng.standard_normal((m, m)), -1)
    vals = np.linalg.eigvals(H)
    H2, Q = krylov._apply_shifts(H.copy(), vals[:4])
    assert np.allclose(Q.T @ Q, np.eye(m), atol=1e-12)
    assert np.allclose(Q.T @ H @ Q, H2, atol=1e-10)
    assert np.allclose(np.sort_complex(np.linalg.eigvals(H2)), np.sort_complex(vals), atol=1e-9)


def test_eigs_close_magnitudes_need_the_restart_buffer():
    """Spectrum modelled on the tangent map of the laminar Kolmogorov state:
    leading multipliers in complex pairs whose magnitudes differ by < 1e-3."""
    rng = np.random.default_rng(6)
    n = 240
    mags = [1.005614, 1.000034, 0.981687, 0.981096, 0.980199, 0.9795]
    angs = [0.3323, 0.2990, 0.2263, 0.3298, 0.0, 0.41]
    blocks = []
    for r, t in zip(mags, angs):
        if t == 0.0:
            blocks.append(np.array([[r]]))
            blocks.append(np.array([[r * 0.9999]]))
        else:
            blocks.append(r * np.array([[np.cos(t), -np.sin(t)], [np.sin(t), np.cos(t)]]))
    d = sum(b.shape[0] for b in blocks)
    bulk = rng.uniform(0.2, 0.9, n - d)
    D = np.zeros((n, n))
    i = 0
    for b in blocks:
        s = b.shape[0]
        D[i:i + s, i:i + s] = b
        i += s
    D[i:, i:] = np.diag(bulk)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ D @ Q.T
    ref = np.linalg.eigvals(A)
    ref = ref[krylov._order(ref, "LM")][:6]
    out = krylov.eigs(OPS, _mv(A), n, k=6, m=30, which="LM", tol=1e-11)
    for lam in ref:
        assert np.min(np.abs(out["values"] - lam)) < 1e-9, (lam, out["values"])


def _rotated_reference(E, J, shift, angle, k, which="LR"):
    """Dense restatement of the rotated shift-invert eigen mode: the k
    eigenvalues of the doubled real form M of E - sigma that are smallest in
    magnitude (largest of M^-1), an orthonormal real basis W of their
    invariant subspace, R = W_r^T J W_r + W_i^T J W_i, eig(R) sorted."""
    n = E.shape[0]
    sr, si = shift * np.cos(angle), shift * np.sin(angle)
    M = np.block([[E - sr * np.eye(n), si * np.eye(n)], [-si * np.eye(n), E - sr * np.eye(n)]])
    lam, X = np.linalg.eig(M)
    idx = np.argsort(np.abs(lam))[:k]
    Z = np.concatenate([X[:, idx].real, X[:, idx].imag], axis=1)
    U, s, _ = np.linalg.svd(Z, full_matrices=False)
    W = U[:, s > 1e-10 * s[0]]
    Wr, Wi = W[:n], W[n:]
    R = Wr.T @ J @ Wr + Wi.T @ J @ Wi
    ev = np.linalg.eigvals(R)
    return ev[krylov._order(ev, which)][:k]


def test_rotated_shift_invert_recovers_eigenvalues_of_j():
    """With E a function of J (here (I + tau J)^K), E's invariant subspaces are
    J's, so the Galerkin step returns J's eigenvalues exactly: the ones whose
    image (1 + tau lambda)^K lies closest to sigma.  Pins the doubled-vector
    operator (the signs of the rotation), the basis and the recovery."""
    rng = np.random.default_rng(7)
    n, tau, K = 160, 0.01, 10
    Qm, _ = np.linalg.qr(rng.standard_normal((n, n)))
    D = -np.diag(np.linspace(0.5, 40.0, n))
    # 0.05 +- 0.45i: E maps it near sigma = 1.01 exp(0.05i) (a complex pair of
    # J gives one conjugate pair of the doubled operator), real eigenvalues give
    # two eigenvalues of the doubled operator each (mu - sigma, mu - conj sigma),
    # so the recovery lists each of them twice, as the reference's does
    D[:2, :2] = [[0.05, 0.45], [-0.45, 0.05]]
    D[2, 2], D[3, 3], D[4, 4], D[5, 5] = 0.12, -0.2, -0.6, -0.9
    J = Qm @ D @ np.linalg.inv(Qm)
    E = np.linalg.matrix_power(np.eye(n) + tau * J, K)
    shift, angle = 1.01, 0.05
    out = krylov.eigs_rotated_shift_invert(OPS, _mv(J), _mv(E), n, k=6, m=30, shift=shift, angle=angle, L=6,
                                           lin_tol=1e-10, tol=1e-12)
    ref = _rotated_reference(E, J, shift, angle, 6)
    lamJ = np.linalg.eigvals(J)
    for lam in ref:
        assert np.min(np.abs(lamJ - lam)) < 1e-9            # the reference construction returns J's eigenvalues
    got = out["values"]
    assert len(got) == 6
    for lam in ref:
        assert np.min(np.abs(got - lam)) < 1e-8 * max(1, abs(lam)), (lam, got)
    assert np.all(out["residuals"] < 1e-7)
    for lam in [0.05 + 0.45j, 0.05 - 0.45j, 0.12]:
        assert np.min(np.abs(got - lam)) < 1e-8
    assert np.allclose(sorted(out["distinct"], key=lambda z: (z.real, z.imag)),
                       sorted([0.05 + 0.45j, 0.05 - 0.45j, 0.12, -0.2], key=lambda z: (z.real, z.imag))[-len(out["distinct"]):],
                       atol=1e-8)


def test_rotated_operator_is_the_real_form_of_e_minus_sigma():
    """The doubled-vector matvec, applied through its inverse, solves
    (E - sigma) z = b for complex z = x_r + i x_i."""
    rng = np.random.default_rng(8)
    n = 40
    E = rng.standard_normal((n, n)) / np.sqrt(n)
    shift, angle = 1.3, 0.4
    sigma = shift * np.exp(1j * angle)
    inv = krylov.rotated_shift_invert(OPS, _mv(E), n, shift, angle, L=4, tol=1e-13)
    b = rng.standard_normal(2 * n)
    x = np.zeros(2 * n)
    inv(b, x)
    z = np.linalg.solve(E - sigma * np.eye(n), b[:n] + 1j * b[n:])
    assert np.allclose(x[:n] + 1j * x[n:], z, atol=1e-10)
