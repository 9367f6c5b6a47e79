"""Write the golden fixtures tests/golden/golden_*.npz.

The vectors come from an independent numpy restatement of the reference
numerics (numpy.fft / pocketfft, vectorised numpy), written separately from
the C oracle, so the oracle and the fixtures check each other.  Inputs are
seeded; outputs are FP64.  Reference functions restated (relative to the
reference root): Source/math_support.cu:393-519 (tables), Source/advection_2_3.cu
:146-525 (nonlinear terms), Source/diffusion.cu:96-365 (update),
Source/RK_time_step.cu:248-355 (stages, RK3), Source/divergence.cu:8-30,
Source/NS3D_periodic.cu:173-250,340-361 (Krylov packing, J.v).

    python tests/golden/make_golden.py      # rewrites the .npz files
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PI = math.pi


def tables(Nx, Ny, Nz, Lx, Ly, Lz):
    Mz = Nz // 2 + 1
    m = np.arange(Nx); m = np.where(m >= Nx // 2, m - Nx, m).astype(float)
    n = np.arange(Ny); n = np.where(n >= Ny // 2, n - Ny, n).astype(float)
    q = np.arange(Mz).astype(float)
    kx, ky, kz = 2 * PI / Lx * m, 2 * PI / Ly * n, 2 * PI / Lz * q
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    den = np.where(k2 == 0, 1.0, k2)
    K = [KX, KY, KZ]
    AM = [[K[a] * K[b] / den for b in range(3)] for a in range(3)]
    for a in range(3):
        for b in range(3):
            AM[a][b][0, 0, 0] = 0.0
    # 2/3 ellipsoid normalised by (Nx, Ny, Mz) in index space, mode 0 removed
    M, N_, Q = np.meshgrid(m, n, q, indexing="ij")
    s = np.sqrt((M * 2 * PI / Lx) ** 2 / (2 * Nx * PI / Lx) ** 2 + (N_ * 2 * PI / Ly) ** 2 / (2 * Ny * PI / Ly) ** 2
                + (Q * 2 * PI / Lz) ** 2 / (2 * Mz * PI / Lz) ** 2)
    mask = (s < 2.0 / 3.0).astype(float)
    mask[0, 0, 0] = 0.0
    return dict(K=K, k2=k2, AM=AM, mask=mask)


def c2r(X, shape):
    return np.fft.irfftn(X, s=shape, axes=(0, 1, 2))


def r2c(x):
    return np.fft.rfftn(x, axes=(0, 1, 2))


def nonlinear(u, v, T, shape, ipq=True):
    """Q_b = (v.grad) u_b + (u.grad) v_b on masked fields (v = None: (u.grad) u)."""
    K, M = T["K"], T["mask"]
    phys = lambda X: c2r(M * X, shape)
    U = [phys(c) for c in u]
    dU = [[phys(1j * K[a] * u[b]) for a in range(3)] for b in range(3)]
    if v is None:
        Q = [sum(U[a] * dU[b][a] for a in range(3)) for b in range(3)]
    else:
        V = [phys(c) for c in v]
        dV = [[phys(1j * K[a] * v[b]) for a in range(3)] for b in range(3)]
        Q = [sum(V[a] * dU[b][a] for a in range(3)) + sum(U[a] * dV[b][a] for a in range(3)) for b in range(3)]
    Qh = [r2c(q) for q in Q]
    if ipq:
        Qh = [1j * c.imag for c in Qh]
        for c in Qh:
            c[0, 0, 0] = 0.0
    return Qh


def proj(Q, T):
    AM = T["AM"]
    return [sum(AM[a][b] * Q[b] for b in range(3)) for a in range(3)]


def stage(u, f, T, shape, dt, Re):
    Q = nonlinear(u, None, T, shape)
    AQ = proj(Q, T)
    G = 1.0 - T["k2"] * dt / Re
    return [dt * AQ[c] + G * u[c] + dt * f[c] - dt * Q[c] for c in range(3)]


def rk3(u, f, T, shape, dt, Re):
    u1 = stage(u, f, T, shape, dt, Re)
    u1 = stage(u1, f, T, shape, dt, Re)
    u2 = [0.75 * a + 0.25 * b for a, b in zip(u, u1)]
    u2 = stage(u2, f, T, shape, dt, Re)
    return [a / 3.0 + 2.0 * b / 3.0 for a, b in zip(u, u2)]


def jv(u, v, T, shape, Re):
    B = nonlinear(u, v, T, shape)
    AB = proj(B, T)
    return [AB[c] - B[c] - T["k2"] / Re * v[c] for c in range(3)]


def random_field(rng, shape):
    """Real random field with a decaying spectrum, as spectral arrays (Hermitian by construction)."""
    Nx, Ny, Nz = shape
    out = []
    for _ in range(3):
        x = rng.standard_normal(shape)
        X = r2c(x)
        T = tables(Nx, Ny, Nz, 2 * PI, 2 * PI, 2 * PI)
        X *= np.exp(-0.1 * T["k2"])
        out.append(X)
    return out


def make(shape, seed, Re=100.0, dt=2e-3):
    Nx, Ny, Nz = shape
    rng = np.random.default_rng(seed)
    T = tables(Nx, Ny, Nz, 2 * PI, 2 * PI, 2 * PI)
    u = random_field(rng, shape)
    v = [0.1 * c for c in random_field(rng, shape)]
    f = [0.01 * c for c in random_field(rng, shape)]
    Q = nonlinear(u, None, T, shape)
    B = nonlinear(u, v, T, shape)
    S = stage(u, f, T, shape, dt, Re)
    R = rk3(u, f, T, shape, dt, Re)
    J = jv(u, v, T, shape, Re)
    div = (1j * (T["K"][0] * u[0] + T["K"][1] * u[1] + T["K"][2] * u[2])) / dt
    return dict(shape=np.array(shape), Re=Re, dt=dt,
                u=np.stack(u), v=np.stack(v), f=np.stack(f), Q=np.stack(Q), B=np.stack(B),
                stage=np.stack(S), rk3=np.stack(R), jv=np.stack(J), div=div, mask=T["mask"])


if __name__ == "__main__":
    for shape, seed in [((8, 8, 8), 1), ((12, 8, 16), 2)]:
        d = make(shape, seed)
        name = os.path.join(HERE, "golden_%dx%dx%d.npz" % shape)
        np.savez_compressed(name, **d)
        print("wrote", name, os.path.getsize(name), "bytes")
