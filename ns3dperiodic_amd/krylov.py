"""Krylov drivers over the J.v operators: implicitly restarted Arnoldi,
BiCGStab(L) and damped Newton (the Source/Arnoldi stack).

* ``arnoldi_extend`` -- Arnoldi_driver (Source/Arnoldi/Arnoldi_Driver.cu:29-123):
  grow A V = V H + f e^T from p to m columns with classical Gram-Schmidt and
  DGKS re-orthogonalisation until the correction norm is below 5e-12.
* ``eigs`` -- Implicit_restart_Arnoldi_GPU_data
  (Source/Arnoldi/Implicit_restart_Arnoldi.cu:443-741, shifts from
  Select_Shifts.cpp / QR_Shifts.cpp:7-72): exact shifts (the m-k unwanted
  Ritz values, complex pairs as one real double shift), restart to k,
  stop when the Ritz residual norm is below tol; eigenpairs are the wanted
  Ritz pairs of the final m-step factorisation, residuals ||A x - lambda x||
  evaluated with A.
* ``bicgstab_l`` -- BiCGStabL (Source/Arnoldi/BiCGStabL.cu:18-184), the
  BiCGStab(L) of Sleijpen and Fokkema (ETNA 1, 1993).
* ``newton`` -- Newton (Source/Arnoldi/Newton.cu:9-98): x += 0.8 delta with
  J(x) delta = -F(x) solved by BiCGStab(L), until ||F(x)|| < tol.
* ``fd_jv_batch`` -- finite-difference J.v of the time-T map for many
  perturbations at once (build_Jacobian's difference quotient,
  Source/Jacobian.cu:230-318, applied to the RK3 flow map), all members
  advanced together by ns3d_rk3_ssp_batch.

Vectors live in FP64 on the GPU; every inner product, projection and basis
update runs in the library's deterministic FP64 kernels (``GpuOps``: the
ns3d_vec_* entry points).  Only the m x m Hessenberg algebra (eigenvalues,
shifted QR, m <= 128) runs on the host in numpy, where the reference used
LAPACK.  ``NumpyOps`` runs the same drivers on host arrays (dense test
problems).  The Newton-Krylov split of precisions is the caller's choice of
contexts: the Jacobian may come from an FP32 context while the residual F(x)
is evaluated in FP64.
"""
import ctypes
import math

import numpy as np

# ----------------------------------------------------------------------------- vector backends


class NumpyOps:
    """Host backend: vectors are 1D float64 arrays, a basis is a 2D array [mmax, n]."""

    def zeros(self, n, m=None):
        return np.zeros(n if m is None else (m, n))

    def dot(self, x, y):
        return float(np.dot(x, y))

    def axpby(self, a, x, b, y):
        y *= b
        y += a * x
        return y

    def copy(self, x, y):
        y[...] = x
        return y

    def gemv_t(self, V, m, w):
        return V[:m] @ w

    def gemv(self, V, m, h, alpha, beta, y):
        y *= beta
        y += alpha * (np.asarray(h, float)[:m] @ V[:m])
        return y

    def gemm(self, V, m, Q, W):
        k = Q.shape[1]
        W[:k] = Q[:m, :k].T @ V[:m]
        return W

    def from_host(self, a):
        return np.array(a, dtype=np.float64)

    def to_host(self, x):
        return np.array(x)


class GpuOps:
    """Device backend over an NS3D context: vectors are contiguous float64
    torch tensors; every operation is one of the ns3d_vec_* kernels on the
    context's stream."""

    def __init__(self, g):
        import torch
        self.g = g
        self.lib = g.lib
        self.ctx = g.ctx
        self.torch = torch
        self.dev = g.device

    @staticmethod
    def _p(t):
        return ctypes.c_void_p(t.data_ptr())

    def _chk(self, st, what):
        from ._lib import check
        check(st, what)

    def zeros(self, n, m=None):
        return self.torch.zeros(n if m is None else (m, n), dtype=self.torch.float64, device=self.dev)

    def dot(self, x, y):
        out = ctypes.c_double()
        self._chk(self.lib.ns3d_vec_dot(self.ctx, x.numel(), self._p(x), self._p(y), ctypes.byref(out)), "ns3d_vec_dot")
        return out.value

    def axpby(self, a, x, b, y):
        self._chk(self.lib.ns3d_vec_axpby(self.ctx, y.numel(), float(a), self._p(x), float(b), self._p(y)),
                  "ns3d_vec_axpby")
        return y

    def copy(self, x, y):
        return self.axpby(1.0, x, 0.0, y)

    def gemv_t(self, V, m, w):
        h = (ctypes.c_double * m)()
        n = w.numel()
        self._chk(self.lib.ns3d_vec_gemv_t(self.ctx, n, m, self._p(V), V.shape[1], self._p(w), h), "ns3d_vec_gemv_t")
        return np.array(h[:], dtype=np.float64)

    def gemv(self, V, m, h, alpha, beta, y):
        hh = (ctypes.c_double * m)(*[float(v) for v in np.asarray(h)[:m]])
        self._chk(self.lib.ns3d_vec_gemv(self.ctx, y.numel(), m, self._p(V), V.shape[1], hh, float(alpha), float(beta),
                                         self._p(y)), "ns3d_vec_gemv")
        return y

    def gemm(self, V, m, Q, W):
        k = Q.shape[1]
        q = np.asfortranarray(np.asarray(Q, dtype=np.float64)[:m, :k])
        qq = q.ravel(order="F").ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self._chk(self.lib.ns3d_vec_gemm(self.ctx, V.shape[1], m, k, self._p(V), V.shape[1], qq, self._p(W), W.shape[1]),
                  "ns3d_vec_gemm")
        return W

    def from_host(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(self.dev)

    def to_host(self, x):
        self.torch.cuda.synchronize(self.dev)
        return x.detach().cpu().numpy()


# ----------------------------------------------------------------------------- Arnoldi

DGKS_TOL = 5e-12


def arnoldi_extend(ops, matvec, V, H, f, p, m, tol=DGKS_TOL, w=None):
    """Extend A V_p = V_p H_p + f e_p^T to m columns (p = 0 starts from f).
    V: basis storage [>= m, n]; H: host (m, m) array; f: residual vector
    (updated).  Returns the number of DGKS passes that did not converge."""
    n = f.shape[-1]
    w = ops.zeros(n) if w is None else w
    failed = 0
    j = p
    while j < m:
        beta = math.sqrt(max(ops.dot(f, f), 0.0))
        if j == 0:
            if beta < 1e-15:
                raise ValueError("Arnoldi: starting vector has norm < 1e-15")
        else:
            H[j, j - 1] = beta
        if beta == 0.0:
            raise ValueError(f"Arnoldi: invariant subspace found at column {j} (||f|| = 0)")
        ops.axpby(1.0 / beta, f, 0.0, V[j])
        matvec(V[j], w)
        h = ops.gemv_t(V, j + 1, w)
        ops.copy(w, f)
        ops.gemv(V, j + 1, h, -1.0, 1.0, f)
        for _ in range(12):
            cvec = ops.gemv_t(V, j + 1, f)
            ops.gemv(V, j + 1, cvec, -1.0, 1.0, f)
            h = h + cvec
            if np.linalg.norm(cvec) <= tol:
                break
        else:
            failed += 1
        H[: j + 1, j] = h
        j += 1
    return failed


def _order(vals, which):
    if which == "LM":
        key = -np.abs(vals)
    elif which == "LR":
        key = -vals.real
    elif which == "SR":
        key = vals.real
    elif which == "SM":
        key = np.abs(vals)
    else:
        raise ValueError(which)
    return np.lexsort((np.abs(vals.imag), key))


def _givens(a, b):
    """(c, s) with [c s; -s c] [a; b] = [r; 0]."""
    r = math.hypot(a, b)
    if r == 0.0:
        return 1.0, 0.0
    return a / r, b / r


def _house(v):
    """Householder reflector P = I - 2 w w^T (|w| = 1) with P v = -+|v| e_1."""
    w = np.array(v, dtype=np.float64)
    nv = np.linalg.norm(w)
    if nv == 0.0:
        return np.eye(len(w))
    w[0] += math.copysign(nv, w[0]) if w[0] != 0 else nv
    w /= np.linalg.norm(w)
    return np.eye(len(w)) - 2.0 * np.outer(w, w)


def _shift_single(H, Q, mu):
    """One implicit single-shift QR step (bulge chase with Givens rotations)."""
    m = H.shape[0]
    x, y = H[0, 0] - mu, H[1, 0]
    for k in range(m - 1):
        c, sn = _givens(x, y)
        G = np.array([[c, sn], [-sn, c]])
        q = max(k - 1, 0)
        H[k:k + 2, q:] = G @ H[k:k + 2, q:]
        r = min(k + 3, m)
        H[:r, k:k + 2] = H[:r, k:k + 2] @ G.T
        Q[:, k:k + 2] = Q[:, k:k + 2] @ G.T
        if k < m - 2:
            x, y = H[k + 1, k], H[k + 2, k]


def _shift_double(H, Q, mu):
    """One implicit Francis double-shift step for the pair mu, conj(mu)
    (Golub and Van Loan, Algorithm 7.5.1), in real arithmetic."""
    m = H.shape[0]
    if m < 3:
        raise ValueError("double shift needs m >= 3")
    s, t = 2.0 * mu.real, abs(mu) ** 2
    x = H[0, 0] * H[0, 0] + H[0, 1] * H[1, 0] - s * H[0, 0] + t
    y = H[1, 0] * (H[0, 0] + H[1, 1] - s)
    z = H[1, 0] * H[2, 1]
    for k in range(m - 2):
        P = _house([x, y, z])
        q = max(k - 1, 0)
        H[k:k + 3, q:] = P @ H[k:k + 3, q:]
        r = min(k + 4, m)
        H[:r, k:k + 3] = H[:r, k:k + 3] @ P
        Q[:, k:k + 3] = Q[:, k:k + 3] @ P
        x, y = H[k + 1, k], H[k + 2, k]
        if k < m - 3:
            z = H[k + 3, k]
    c, sn = _givens(x, y)
    G = np.array([[c, sn], [-sn, c]])
    H[m - 2:, m - 3:] = G @ H[m - 2:, m - 3:]
    H[:, m - 2:] = H[:, m - 2:] @ G.T
    Q[:, m - 2:] = Q[:, m - 2:] @ G.T


def _apply_shifts(H, shifts):
    """Apply the shifts to the host Hessenberg H (m x m) by implicit QR steps
    (bulge chasing keeps H Hessenberg also for exact shifts); complex
    conjugate pairs are applied together as one real double shift.
    Returns (H', Q) with H' = Q^T H Q."""
    H = np.array(H, dtype=np.float64)
    m = H.shape[0]
    Q = np.eye(m)
    done = np.zeros(len(shifts), bool)
    for i, mu in enumerate(shifts):
        if done[i]:
            continue
        done[i] = True
        if abs(np.imag(mu)) > 0:
            for j in range(i + 1, len(shifts)):
                if not done[j] and abs(shifts[j] - np.conj(mu)) <= 1e-12 * max(1.0, abs(mu)):
                    done[j] = True
                    break
            _shift_double(H, Q, complex(mu))
        else:
            _shift_single(H, Q, float(np.real(mu)))
    H = np.triu(H, -1)  # entries below the subdiagonal are rounding residue
    return H, Q


def eigs(ops, matvec, n, k=6, m=30, which="LM", tol=1e-10, max_iter=300, seed=0, v0=None, vectors=False, buffer=None,
         residuals=True):
    """k eigenvalues of the operator `matvec(x, y)` (y = A x, length-n FP64
    vectors) with an implicitly restarted Arnoldi of m vectors.  Each restart
    keeps k + buffer Ritz vectors (default buffer max(2, k/2)) and converges
    the leading k: without the buffer an eigenvalue just below the k-th in the
    `which` order can stay outside the subspace while a slightly smaller one
    converges in its place.

    Returns dict(values, residuals, iterations, ritz_norm[, vectors_real,
    vectors_imag]): eigenvalues sorted by `which`, residuals ||A x - lambda x||
    of the unit-norm eigenvectors computed with A (skipped, NaN, with
    residuals=False), the number of restarts.
    """
    if not (0 < k < m <= 128 and m <= n):
        raise ValueError("need 0 < k < m <= min(n, 128)")
    buffer = max(2, k // 2) if buffer is None else int(buffer)
    kw = max(k, min(k + buffer, m - 3))
    V = ops.zeros(n, m)
    V1 = ops.zeros(n, m)
    H = np.zeros((m, m))
    if v0 is None:
        v0 = np.random.default_rng(seed).normal(0.0, 2.0, n)
    f = ops.from_host(v0)
    w = ops.zeros(n)
    p = 0
    it = 0
    ritz_norm = np.inf
    kk = k
    while True:
        arnoldi_extend(ops, matvec, V, H, f, p, m, w=w)
        vals = np.linalg.eigvals(H)
        order = _order(vals, which)
        kk = kw
        # keep a complex pair together
        if abs(vals[order[kk - 1]].imag) > 0 and kk < m - 1 and abs(vals[order[kk]] - np.conj(vals[order[kk - 1]])) <= 1e-10 * max(1.0, abs(vals[order[kk - 1]])):
            kk = kk + 1
        # Ritz residual estimates |beta_m e_m^T y| of the leading k (plus a split pair)
        kc = k + 1 if (k < kk and abs(vals[order[k - 1]].imag) > 0
                       and abs(vals[order[k]] - np.conj(vals[order[k - 1]])) <= 1e-10 * max(1.0, abs(vals[order[k - 1]]))) else k
        ev, Y = np.linalg.eig(H)
        fnorm = math.sqrt(max(ops.dot(f, f), 0.0))
        want = order[:kc]
        idx = [int(np.argmin(np.abs(ev - vals[i]))) for i in want]
        ritz = np.array([fnorm * abs(Y[m - 1, i]) / max(np.linalg.norm(Y[:, i]), 1e-300) for i in idx])
        scale = np.maximum(np.abs(vals[want]), 1e-300)
        ritz_norm = float(np.linalg.norm(ritz / scale))
        it += 1
        if ritz_norm <= tol or it >= max_iter:
            break
        shifts = vals[order[kk:]]
        H, Q = _apply_shifts(H, shifts)
        # f = V Q[:, kk] * H[kk, kk-1] + f * Q[m-1, kk-1]
        ops.gemv(V, m, Q[:, kk], H[kk, kk - 1], Q[m - 1, kk - 1], f)
        ops.gemm(V, m, Q[:, :kk], V1)
        for j in range(kk):
            ops.copy(V1[j], V[j])
        H[kk:, :] = 0.0
        H[:, kk:] = 0.0
        p = kk
    # Ritz pairs of the final m-step factorisation: x = V_m y for the wanted eigenpairs of H_m
    lam, Y = np.linalg.eig(H)
    sel = _order(lam, which)[:k]
    lam, Y = lam[sel], Y[:, sel]
    Yr = np.ascontiguousarray(Y.real)
    Yi = np.ascontiguousarray(Y.imag)
    Xr = ops.zeros(n, k)
    Xi = ops.zeros(n, k)
    ops.gemm(V, m, Yr, Xr)
    ops.gemm(V, m, Yi, Xi)
    res = []
    ar, ai = ops.zeros(n), ops.zeros(n)
    for i in range(k if residuals else 0):
        nr2 = ops.dot(Xr[i], Xr[i]) + ops.dot(Xi[i], Xi[i])
        nrm = math.sqrt(nr2) if nr2 > 0 else 1.0
        matvec(Xr[i], ar)
        matvec(Xi[i], ai)
        # A x - lam x with x = xr + i xi, lam = a + i b
        a, b = lam[i].real, lam[i].imag
        ops.axpby(-a, Xr[i], 1.0, ar)
        ops.axpby(b, Xi[i], 1.0, ar)
        ops.axpby(-a, Xi[i], 1.0, ai)
        ops.axpby(-b, Xr[i], 1.0, ai)
        res.append(math.sqrt(max(ops.dot(ar, ar) + ops.dot(ai, ai), 0.0)) / nrm)
    out = {"values": lam, "residuals": np.array(res) if residuals else np.full(k, np.nan), "iterations": it,
           "ritz_norm": ritz_norm}
    if vectors:
        out["vectors_real"], out["vectors_imag"] = Xr, Xi
    return out


# ----------------------------------------------------------------------------- BiCGStab(L)


def bicgstab_l(ops, matvec, b, x, L=2, tol=1e-10, max_iter=500, rshadow=None):
    """Solve A x = b (x: initial guess, updated in place) with BiCGStab(L).
    Returns (flag, iterations, relative residual): flag 0 converged,
    1 not converged, -1 rho = 0, -2 omega = 0, -3 non-finite (the return
    codes of BiCGStabL, Source/Arnoldi/BiCGStabL.cu)."""
    n = b.shape[-1]
    r = ops.zeros(n, L + 1)
    u = ops.zeros(n, L + 1)
    tmp = ops.zeros(n)
    matvec(x, tmp)
    ops.copy(b, r[0])
    ops.axpby(-1.0, tmp, 1.0, r[0])
    rt = ops.zeros(n)
    ops.copy(r[0] if rshadow is None else rshadow, rt)
    bnorm = math.sqrt(max(ops.dot(b, b), 0.0)) or 1.0
    rho0, alpha, omega = 1.0, 0.0, 1.0
    res = math.sqrt(max(ops.dot(r[0], r[0]), 0.0)) / bnorm
    it = 0
    while res > tol and it < max_iter:
        it += 1
        rho0 = -omega * rho0
        # BiCG part
        for j in range(L):
            rho1 = ops.dot(r[j], rt)
            if rho0 == 0.0:
                return -1, it, res
            beta = alpha * rho1 / rho0
            rho0 = rho1
            for i in range(j + 1):
                ops.axpby(1.0, r[i], -beta, u[i])
            matvec(u[j], u[j + 1])
            gamma = ops.dot(u[j + 1], rt)
            if gamma == 0.0:
                return -1, it, res
            alpha = rho0 / gamma
            for i in range(j + 1):
                ops.axpby(-alpha, u[i + 1], 1.0, r[i])
            matvec(r[j], r[j + 1])
            ops.axpby(alpha, u[0], 1.0, x)
        # minimal-residual part (modified Gram-Schmidt on r_1..r_L)
        tau = np.zeros((L + 1, L + 1))
        sigma = np.zeros(L + 1)
        gp = np.zeros(L + 1)
        for j in range(1, L + 1):
            for i in range(1, j):
                tau[i, j] = ops.dot(r[j], r[i]) / sigma[i]
                ops.axpby(-tau[i, j], r[i], 1.0, r[j])
            sigma[j] = ops.dot(r[j], r[j])
            if sigma[j] == 0.0:
                return -2, it, res
            gp[j] = ops.dot(r[0], r[j]) / sigma[j]
        g = np.zeros(L + 1)
        g[L] = gp[L]
        omega = g[L]
        for j in range(L - 1, 0, -1):
            g[j] = gp[j] - sum(tau[j, i] * g[i] for i in range(j + 1, L + 1))
        gpp = np.zeros(L + 1)
        for j in range(1, L):
            gpp[j] = g[j + 1] + sum(tau[j, i] * g[i + 1] for i in range(j + 1, L))
        if omega == 0.0:
            return -2, it, res
        ops.axpby(g[1], r[0], 1.0, x)
        ops.axpby(-gp[L], r[L], 1.0, r[0])
        ops.axpby(-g[L], u[L], 1.0, u[0])
        for j in range(1, L):
            ops.axpby(-g[j], u[j], 1.0, u[0])
            ops.axpby(gpp[j], r[j], 1.0, x)
            ops.axpby(-gp[j], r[j], 1.0, r[0])
        res = math.sqrt(max(ops.dot(r[0], r[0]), 0.0)) / bnorm
        if not math.isfinite(res):
            return -3, it, res
    # true residual
    matvec(x, tmp)
    ops.axpby(1.0, b, -1.0, tmp)
    res = math.sqrt(max(ops.dot(tmp, tmp), 0.0)) / bnorm
    return (0 if res <= 10 * tol else 1), it, res


# ----------------------------------------------------------------------------- Newton


def newton(ops, rhs, jacobian, x, tol=1e-10, max_iter=50, alpha=0.8, L=6, lin_tol=1e-8, lin_iter=500, min_iter=3):
    """Damped Newton for F(x) = 0 (x updated in place).  rhs(x, out) writes
    F(x); jacobian(x) returns a matvec v -> J(x) v for the current x.  Each
    step solves J delta = -F(x) by BiCGStab(L) and sets x += alpha*delta; stop
    once ||F(x)|| < tol after at least `min_iter` steps (Newton.cu:9-98).
    A breakdown of the linear solve (BiCGStab(L) flag < 0) ends the
    iteration unconverged, x left at the last accepted step.
    Returns dict(converged, iterations, residual, history[, breakdown: the
    BiCGStab flag, lin_history: (iterations, relative residual) per solve])."""
    n = x.shape[-1]
    F = ops.zeros(n)
    delta = ops.zeros(n)
    rhs(x, F)
    fn = math.sqrt(max(ops.dot(F, F), 0.0))
    hist = [fn]
    if fn < tol:
        return {"converged": True, "iterations": 0, "residual": fn, "history": hist}
    it = 0
    lin = []
    for it in range(1, max_iter + 1):
        ops.axpby(0.0, F, -1.0, F)              # b = -F(x)
        mv = jacobian(x)
        ops.axpby(0.0, F, 0.0, delta)
        flag, lit, lres = bicgstab_l(ops, mv, F, delta, L=L, tol=lin_tol, max_iter=lin_iter)
        lin.append((lit, lres))
        if flag < 0:
            return {"converged": False, "iterations": it - 1, "residual": fn, "history": hist, "breakdown": flag,
                    "lin_history": lin}
        ops.axpby(alpha, delta, 1.0, x)
        rhs(x, F)
        fn = math.sqrt(max(ops.dot(F, F), 0.0))
        hist.append(fn)
        if fn < tol and it >= min_iter:
            return {"converged": True, "iterations": it, "residual": fn, "history": hist, "lin_history": lin}
        if not math.isfinite(fn):
            break
    return {"converged": False, "iterations": it, "residual": fn, "history": hist, "lin_history": lin}


# ----------------------------------------------------------------------------- batched FD J.v


def fd_jv_batch(g, u, V, eps=1e-7, steps=1, f=None, central=False):
    """Finite-difference J.v of the steps-step RK3 map Phi at u for every
    perturbation in V ([m, 3, Nx, Ny, Mz] complex128 on the GPU):
    forward (Phi(u + eps v) - Phi(u)) / eps, all m + 1 fields in one batch,
    or central (Phi(u + eps v) - Phi(u - eps v)) / (2 eps) (2m fields), the
    difference quotient of build_Jacobian (Source/Jacobian.cu:230-318).
    u: three spectral fields; returns [m, 3, Nx, Ny, Mz]."""
    import torch
    u3 = torch.stack([c.to(torch.complex128) for c in u])
    m = V.shape[0]
    if central:
        ub = torch.empty((2 * m,) + tuple(u3.shape), dtype=torch.complex128, device=u3.device)
        ub[0::2] = u3 + eps * V
        ub[1::2] = u3 - eps * V
        g.RK3_SSP_batch(ub, f, steps=steps)
        return (ub[0::2] - ub[1::2]) / (2.0 * eps)
    ub = torch.empty((m + 1,) + tuple(u3.shape), dtype=torch.complex128, device=u3.device)
    ub[0] = u3
    ub[1:] = u3 + eps * V
    g.RK3_SSP_batch(ub, f, steps=steps)
    return (ub[1:] - ub[0]) / eps


# ----------------------------------------------------------------------------- rotated shift-invert eigen mode


def rotated_shift_invert(ops, expo, n, shift, angle, L=6, tol=1e-9, max_iter=None, stats=None):
    """x -> (E - sigma)^-1 x, sigma = shift * exp(i angle), on the doubled real
    vector [x_r; x_i] (length 2n) of a real operator E (`expo(x, y)`: y = E x,
    length n): Axb_rotated_exponent_invert (Source/NS3D_periodic.cu:590-630),
    whose matvec NSCallMatrixVector_exponential_rotated_reduced (:541-586) is
        [y_r; y_i] = [E x_r - sr x_r + si x_i ; E x_i - sr x_i - si x_r]
    (sr, si = shift cos(angle), shift sin(angle)): the real form of the complex
    operator E - sigma acting on x_r + i x_i.  Each application solves that
    system by BiCGStab(L) from a zero start and raises if it does not converge
    (the reference exits).  `stats`, if a list, receives the iterations of
    every solve.  Returns the matvec of the inverse."""
    sr, si = shift * math.cos(angle), shift * math.sin(angle)
    max_iter = 2 * n if max_iter is None else max_iter

    def rot(x, y):
        xr, xi, yr, yi = x[:n], x[n:], y[:n], y[n:]
        expo(xr, yr)
        ops.axpby(-sr, xr, 1.0, yr)
        ops.axpby(si, xi, 1.0, yr)
        expo(xi, yi)
        ops.axpby(-sr, xi, 1.0, yi)
        ops.axpby(-si, xr, 1.0, yi)
        return y

    def inv(b, x):
        ops.axpby(0.0, b, 0.0, x)
        flag, it, res = bicgstab_l(ops, rot, b, x, L=L, tol=tol, max_iter=max_iter)
        if flag != 0:
            raise RuntimeError(f"rotated shift-invert: BiCGStab({L}) flag {flag} after {it} iterations, residual {res:.2e}")
        if stats is not None:
            stats.append(it)
        return x

    return inv


def _real_basis(ops, Xr, Xi, k, rank_tol=1e-10):
    """Orthonormal real basis (rows of an ops basis) of span{Xr_i, Xi_i, i < k}:
    the Schur basis of the invariant subspace spanned by k Ritz vectors (a
    complex pair contributes its real and imaginary parts)."""
    n2 = Xr.shape[-1]
    Z = ops.zeros(n2, 2 * k)
    for i in range(k):
        ops.copy(Xr[i], Z[i])
        ops.copy(Xi[i], Z[k + i])
    G = np.zeros((2 * k, 2 * k))
    for i in range(2 * k):
        G[i, : i + 1] = ops.gemv_t(Z, i + 1, Z[i])
    G = G + np.tril(G, -1).T
    w, U = np.linalg.eigh(G)
    keep = w > rank_tol * max(w.max(), 1e-300)
    w, U = w[keep][::-1], U[:, keep][:, ::-1]
    C = U / np.sqrt(w)                      # Z^T C: orthonormal columns
    W = ops.zeros(n2, C.shape[1])
    ops.gemm(Z, 2 * k, C, W)
    return W, C.shape[1]


def eigs_rotated_shift_invert(ops, jv, expo, n, k=6, m=30, shift=1.05, angle=0.05, L=6, lin_tol=1e-9,
                              lin_iter=None, tol=5e-8, max_iter=3000, which="LR", seed=0):
    """The reference's live eigen mode: Implicit_restart_Arnoldi_GPU_data_Matrix_Exponent
    with is_rotated (Source/Arnoldi/Implicit_restart_Arnoldi.cu:456-741), as
    main calls it (Source/NS3D_periodic.cu:1710: operator Axb_rotated_exponent_invert,
    "LM" on the inverse, "LR" for the original map).

    1. IRA ("LM", m vectors, k wanted) on (E - sigma)^-1 over doubled vectors
       of length 2n (``rotated_shift_invert``): the eigenvalues mu of E closest
       to sigma = shift * exp(i angle).
    2. An orthonormal real basis W = [W_r; W_i] of the converged invariant
       subspace (the reference's Schur vectors V(:,1:k) Q).
    3. compute_eigensystem_of_original_mapping_rotation (:334-430): the
       Galerkin matrix R = W_r^T J W_r + W_i^T J W_i of the original operator J
       (`jv(x, y)`: y = J x, length n), eig(R) sorted by `which`, eigenvectors
       (W_r + W_i) y.

    A real eigenvalue mu of E gives two eigenvalues of the doubled operator
    (mu - sigma and mu - conj(sigma)), so, as in the reference, each real
    eigenvalue of J appears twice among `values`; `distinct` merges them.

    Returns dict(values, distinct, residuals ||J x - lambda x|| / ||x||,
    inverse_values (the Ritz values of the inverse), iterations (restarts),
    ritz_norm, solves (BiCGStab iterations of every inner solve), basis_rank)."""
    solves = []
    inv = rotated_shift_invert(ops, expo, n, shift, angle, L=L, tol=lin_tol, max_iter=lin_iter, stats=solves)
    v0 = np.random.default_rng(seed).normal(0.0, 2.0, 2 * n)
    r = eigs(ops, inv, 2 * n, k=k, m=m, which="LM", tol=tol, max_iter=max_iter, v0=v0, vectors=True, residuals=False)
    W, rank = _real_basis(ops, r["vectors_real"], r["vectors_imag"], k)
    Wr, Wi = ops.zeros(n, rank), ops.zeros(n, rank)
    for i in range(rank):
        ops.copy(W[i][:n], Wr[i])
        ops.copy(W[i][n:], Wi[i])
    JW = ops.zeros(n)
    R = np.zeros((rank, rank))
    for B in (Wr, Wi):
        for j in range(rank):
            jv(B[j], JW)
            R[:, j] += ops.gemv_t(B, rank, JW)
    lam, Y = np.linalg.eig(R)
    sel = _order(lam, which)[:min(k, rank)]
    lam, Y = lam[sel], Y[:, sel]
    S = ops.zeros(n, rank)
    for i in range(rank):
        ops.copy(Wr[i], S[i])
        ops.axpby(1.0, Wi[i], 1.0, S[i])
    kk = len(sel)
    Xr, Xi = ops.zeros(n, kk), ops.zeros(n, kk)
    ops.gemm(S, rank, np.ascontiguousarray(Y.real), Xr)
    ops.gemm(S, rank, np.ascontiguousarray(Y.imag), Xi)
    res = []
    ar, ai = ops.zeros(n), ops.zeros(n)
    for i in range(kk):
        nrm = math.sqrt(max(ops.dot(Xr[i], Xr[i]) + ops.dot(Xi[i], Xi[i]), 0.0)) or 1.0
        jv(Xr[i], ar)
        jv(Xi[i], ai)
        a, b = lam[i].real, lam[i].imag
        ops.axpby(-a, Xr[i], 1.0, ar)
        ops.axpby(b, Xi[i], 1.0, ar)
        ops.axpby(-a, Xi[i], 1.0, ai)
        ops.axpby(-b, Xr[i], 1.0, ai)
        res.append(math.sqrt(max(ops.dot(ar, ar) + ops.dot(ai, ai), 0.0)) / nrm)
    distinct = []
    for z in lam:
        if all(abs(z - d) > 1e-8 * max(1.0, abs(z)) for d in distinct):
            distinct.append(z)
    return {"values": lam, "distinct": np.array(distinct), "residuals": np.array(res), "inverse_values": r["values"],
            "iterations": r["iterations"], "ritz_norm": r["ritz_norm"], "solves": solves, "basis_rank": rank}


# ----------------------------------------------------------------------------- NS operators


class NSOperators:
    """The Navier-Stokes maps of the reference's Krylov callbacks on one NS3D
    context, as matvecs on reduced vectors (length 3(NxNyMz - 1), the
    imaginary parts of the three components without mode 0,
    Source/NS3D_periodic.cu:238-250):

    * ``jv``: J(u) v = -P B(u, v) - nu |k|^2 v around the base state u
      (NSCallMatrixVector_reduced, :340-361);
    * ``rhs``: F(x) = L(x), the RK right-hand side of the state x
      (NSCallRHS_reduced, :414-440);
    * ``jacobian(x)``: J(x), the base set to x (Newton);
    * ``tangent(steps)``: v -> dPhi(u0) v over `steps` tangent-map steps
      RK3_SSP_UV (Source/RK_time_step.cu:342-355) from the base state u0,
      restarted from u0 at every call: the time-T return-map linearisation
      whose eigenvalues are Floquet multipliers (NSCallMatrixVector_exponential,
      Source/NS3D_periodic.cu:444-470, which instead keeps advancing the base).
    """

    def __init__(self, g, u_base=None, f=None, tau=None):
        self.g = g
        self.f = f
        self.tau = tau
        self.n = g.vector_length_reduced()
        self.ops = GpuOps(g)
        if u_base is not None:
            self.set_base(u_base)

    def set_base(self, u):
        self.u = [c.clone() for c in u]
        self.g.set_base_state(self.u, self.f, tau=1.0 if self.tau is None else self.tau)

    def jv(self, vin, vout):
        """J v, or with `tau` set the globalised J v - v/tau of Newton's
        pseudo-time step (NSCallMatrixVector_reduced_Newton_Globalization,
        Source/NS3D_periodic.cu:364-389; the reference's tau is 5)."""
        import ctypes as ct
        fn = self.g.lib.ns3d_jv_reduced if self.tau is None else self.g.lib.ns3d_jv_reduced_newton
        fn(self.g.ctx, ct.c_void_p(vin.data_ptr()), ct.c_void_p(vout.data_ptr()))
        from ._lib import check
        check(self.g.lib.ns3d_last_error(self.g.ctx), "ns3d_jv_reduced")
        return vout

    def rhs(self, x, out):
        u = self.g.unpack_reduced(x)
        R = self.g.return_RHS(u, self.f)
        out.copy_(self.g.pack_reduced(R))
        return out

    def jacobian(self, x):
        self.set_base(self.g.unpack_reduced(x))
        return self.jv

    def implicit_exponent(self, steps):
        """v -> E v with E = (single_time_step_iUV_RHS)^steps around the fixed
        base state u, time step the context's dt: NSCallMatrixVector_exponential_reduced
        (Source/NS3D_periodic.cu:474-500) without its shift; the operator of
        the rotated shift-invert eigen mode."""
        def mv(vin, vout):
            v = self.g.unpack_reduced(vin)
            for _ in range(steps):
                self.g.single_time_step_iUV_RHS(self.u, v)
            vout.copy_(self.g.pack_reduced(v))
            return vout
        return mv

    def tangent(self, steps):
        def mv(vin, vout):
            u = [c.clone() for c in self.u]
            v = self.g.unpack_reduced(vin)
            for _ in range(steps):
                self.g.RK3_SSP_UV(u, v, self.f)
            vout.copy_(self.g.pack_reduced(v))
            return vout
        return mv
