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
