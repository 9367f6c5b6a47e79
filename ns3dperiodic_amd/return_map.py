"""Return map: crossings of a probe point's trajectory through a plane
(Source/return_map.cu:380-690).

The probe is the velocity (ux, uy, uz) at one grid point, a point in R^3
that moves as the flow is stepped.  A crossing is detected as in
find_intersection (Source/return_map.cu:425-480): more than 3 steps taken,
the previous point within distance rho of the plane's origin x0, the plane
function s(p) = n.(p - x0) changing sign across the step, and the step
moving along n.  The crossing is then located by bisection on the length of
the last step (restart from the state before it, |s| < 1e-12 or 500
iterations).  This build bisects the absolute bracket [a, b] of the step
length; the reference updates b with the half-width of the bracket, which
makes later sub-steps negative once the base point has moved.

The probe readout is the one-pass spectral sum ``ns3d_probe`` (the
reference inverse-transforms all three fields every step to read one point).
``find_crossing`` works on any system with ``save() / restore(s) / step(dt)
/ probe()``; ``NSSystem`` is the Navier-Stokes one.
"""
import numpy as np


def plane_value(p, x0, normal):
    """test_plane_location: n.(p - x0)."""
    return float(np.dot(normal, np.asarray(p) - np.asarray(x0)))


class CrossingSearch:
    """The crossing search one step at a time (for callers that interleave it
    with other work).  ``advance()`` takes one step and returns True once the
    crossing has been located; ``result`` then holds what find_crossing returns."""

    def __init__(self, system, x0, normal, dt, rho=0.5, tol=1e-12, max_bisect=500, min_steps=3):
        self.system = system
        self.x0 = np.asarray(x0, float)
        self.normal = np.asarray(normal, float)
        self.dt, self.rho, self.tol, self.max_bisect, self.min_steps = dt, rho, tol, max_bisect, min_steps
        self.prev = np.asarray(system.probe(), float)
        self.t = 0.0
        self.steps = 0
        self.sizes = []
        self.result = None

    def advance(self):
        sysm, x0, normal, dt = self.system, self.x0, self.normal, self.dt
        back = sysm.save()
        sysm.step(dt)
        nxt = np.asarray(sysm.probe(), float)
        s1, s2 = plane_value(self.prev, x0, normal), plane_value(nxt, x0, normal)
        moving = float(np.dot(normal, nxt - self.prev))
        if (self.steps > self.min_steps and np.linalg.norm(self.prev - x0) < self.rho and s1 * s2 < 0.0
                and moving > 0.0):
            # the crossing lies inside this step: bisect its length from the saved state
            a, b = 0.0, dt
            err, it, point, h = s2, 0, nxt, dt
            while abs(err) > self.tol and it < self.max_bisect:
                it += 1
                h = 0.5 * (a + b)
                sysm.restore(back)
                sysm.step(h)
                point = np.asarray(sysm.probe(), float)
                err = plane_value(point, x0, normal)
                if err > 0.0:
                    b = h
                else:
                    a = h
            # the system is left at the last sub-step evaluated
            self.sizes.append(h)
            self.result = {"found": True, "steps": self.steps + 1, "time": self.t + h, "point": point,
                           "error": abs(err), "bisections": it, "step_sizes": self.sizes}
            return True
        self.sizes.append(dt)
        self.t += dt
        self.steps += 1
        self.prev = nxt
        return False

    def not_found(self):
        return {"found": False, "steps": self.steps, "time": self.t, "point": self.prev, "error": float("nan"),
                "bisections": 0, "step_sizes": self.sizes}


def find_crossing(system, x0, normal, dt, rho=0.5, max_steps=100000, tol=1e-12, max_bisect=500, min_steps=3):
    """Step `system` with dt until its probe crosses the plane (x0, normal) in
    the direction of `normal`; the system is left at the crossing.  Returns
    dict(found, steps, time, point, error, bisections, step_sizes): time is
    measured from the call, step_sizes lists every step length applied to the
    state that ends at the crossing (for replaying the solve)."""
    search = CrossingSearch(system, x0, normal, dt, rho, tol, max_bisect, min_steps)
    for _ in range(max_steps):
        if search.advance():
            return search.result
    return search.not_found()


class NSSystem:
    """A Navier-Stokes state on an NS3D context as a return-map system:
    RK3_SSP steps of any length, probe at grid point (j, k, l)."""

    def __init__(self, g, u, f=None, point=(0, 0, 0)):
        self.g = g
        self.u = u
        self.f = f
        self.point = tuple(int(v) for v in point)
        self.dt = g.dt

    def save(self):
        return [c.clone() for c in self.u]

    def restore(self, s):
        for d, c in zip(self.u, s):
            d.copy_(c)

    def step(self, dt):
        if dt != self.dt:
            self.g.set_params(dt, None)
            self.dt = dt
        self.g.RK3_SSP(self.u, self.f, steps=1)

    def probe(self):
        return self.g.probe(self.u, *self.point)

    def probe_rhs(self):
        """return_vector3_RHS (Source/return_map.cu:162-200): the RHS at the probe point."""
        return self.g.probe(self.g.return_RHS(self.u, self.f), *self.point)
