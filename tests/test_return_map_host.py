"""Return-map crossing detection and bisection on an analytic trajectory."""
import math

import numpy as np

from ns3dperiodic_amd import return_map


class Circle:
    """Probe p(t) = (cos t, sin t, t/10); the state is t."""

    def __init__(self, t0):
        self.t = t0

    def save(self):
        return self.t

    def restore(self, s):
        self.t = s

    def step(self, dt):
        self.t += dt

    def probe(self):
        return np.array([math.cos(self.t), math.sin(self.t), 0.1 * self.t])


def test_crossing_time_and_direction():
    sysm = Circle(0.1)
    out = return_map.find_crossing(sysm, x0=[0, 0, 0], normal=[0, 1, 0], dt=0.05, rho=2.0)
    assert out["found"]
    # sin t crosses 0 upwards at t = 2 pi (the downward crossing at pi is skipped)
    assert abs(sysm.t - 2 * math.pi) < 1e-11
    assert abs(out["time"] - (2 * math.pi - 0.1)) < 1e-11
    assert out["error"] <= 1e-12 and out["bisections"] > 0
    assert abs(sum(out["step_sizes"]) - out["time"]) < 1e-12
    assert all(s == 0.05 for s in out["step_sizes"][:-1])


def test_ball_radius_gates_crossings():
    # the plane through (0, 0, 0.5) with normal z is crossed once, at t = 5; a ball of radius
    # 0.5 around x0 excludes it (the probe is ~1 away), a ball of radius 2 admits it
    sysm = Circle(0.0)
    out = return_map.find_crossing(sysm, x0=[0, 0, 0.5], normal=[0, 0, 1], dt=0.1, rho=0.5, max_steps=200)
    assert not out["found"]
    sysm = Circle(0.0)
    out = return_map.find_crossing(sysm, x0=[0, 0, 0.5], normal=[0, 0, 1], dt=0.1, rho=2.0, max_steps=200)
    assert out["found"] and abs(sysm.t - 5.0) < 1e-10


def test_min_steps():
    # a crossing within the first 4 steps is not taken
    sysm = Circle(2 * math.pi - 0.12)
    out = return_map.find_crossing(sysm, x0=[0, 0, 0], normal=[0, 1, 0], dt=0.05, rho=2.0, max_steps=300)
    assert out["found"] and abs(sysm.t - 4 * math.pi) < 1e-11


def test_plane_points_follow_the_reference_grid():
    """construct_plane_rectangular: x_j = j*2eps/n - eps, y_k = k*2eps/n - eps,
    z = 0, index j*n_phi + k (the grid stops one spacing short of +eps)."""
    pts = return_map.plane_points(7, 7, 0.01)
    assert pts.shape == (49, 3)
    assert np.allclose(pts[0], [-0.01, -0.01, 0.0])
    assert np.allclose(pts[1], [-0.01, -0.01 + 0.02 / 7, 0.0])
    assert np.allclose(pts[7], [-0.01 + 0.02 / 7, -0.01, 0.0])
    assert np.allclose(pts[48], [0.01 - 0.02 / 7, 0.01 - 0.02 / 7, 0.0])
    assert np.all(pts[:, 2] == 0.0)


def test_rotation_takes_z_onto_the_normal():
    rng = np.random.default_rng(4)
    for n in [rng.standard_normal(3) for _ in range(20)] + [np.array([0, 0, 2.0]), np.array([0, 0, -1.0]),
                                                              np.array([1.0, 0, 0])]:
        R = return_map.rotation_onto(n)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-13)
        assert abs(np.linalg.det(R) - 1.0) < 1e-13
        assert np.allclose(R @ [0, 0, 1.0], n / np.linalg.norm(n), atol=1e-13)
    # the reference's matrix entries (create_matrix3_direct, Source/return_map.cu:146-175)
    n = np.array([0.3, -0.5, 0.8])
    n = n / np.linalg.norm(n)
    u = np.cross([0, 0, 1.0], n)
    c = n[2]
    a = 1.0 / (1.0 + c)
    ref = np.array([[1 - a * (u[1] ** 2 + u[2] ** 2), a * u[0] * u[1] - u[2], u[1] + a * u[0] * u[2]],
                    [u[2] + a * u[0] * u[1], 1 - a * (u[0] ** 2 + u[2] ** 2), a * u[1] * u[2] - u[0]],
                    [a * u[0] * u[2] - u[1], u[0] + a * u[1] * u[2], 1 - a * (u[0] ** 2 + u[1] ** 2)]])
    assert np.allclose(return_map.rotation_onto(n), ref, atol=1e-15)


def test_seed_disc_lies_in_the_plane():
    """Seeds rotated onto the normal and translated lie in the plane through x0."""
    x0, n = np.array([0.2, -0.1, 0.4]), np.array([1.0, 2.0, -0.5])
    seeds = return_map.plane_points() @ return_map.rotation_onto(n).T + x0
    assert np.allclose([return_map.plane_value(p, x0, n) for p in seeds], 0.0, atol=1e-15)
    assert np.max(np.linalg.norm(seeds - x0, axis=1)) <= 0.01 * np.sqrt(2) + 1e-15
