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


def find_crossing(system, x0, normal, dt, rho=0.5, max_steps=100000, tol=1e-12, max_bisect=500, min_steps=3):
    """Step `system` with dt until its probe crosses the plane (x0, normal) in
    the direction of `normal`; the system is left at the crossing.  Returns
    dict(found, steps, time, point, error, bisections, step_sizes): time is
    measured from the call, step_sizes lists every step length applied to the
    state that ends at the crossing (for replaying the solve)."""
    x0 = np.asarray(x0, float)
    normal = np.asarray(normal, float)
    prev = np.asarray(system.probe(), float)
    t = 0.0
    sizes = []
    for step in range(max_steps):
        back = system.save()
        system.step(dt)
        nxt = np.asarray(system.probe(), float)
        s1, s2 = plane_value(prev, x0, normal), plane_value(nxt, x0, normal)
        moving = float(np.dot(normal, nxt - prev))
        if step > min_steps and np.linalg.norm(prev - x0) < rho and s1 * s2 < 0.0 and moving > 0.0:
            # the crossing lies inside this step: bisect its length from the saved state
            a, b = 0.0, dt
            err, it, point = s2, 0, nxt
            while abs(err) > tol and it < max_bisect:
                it += 1
                mid = 0.5 * (a + b)
                system.restore(back)
                system.step(mid)
                point = np.asarray(system.probe(), float)
                err = plane_value(point, x0, normal)
                if err > 0.0:
                    b = mid
                else:
                    a = mid
            # the system is left at the last sub-step evaluated
            sizes.append(mid if it else dt)
            return {"found": True, "steps": step + 1, "time": t + (mid if it else dt), "point": point,
                    "error": abs(err), "bisections": it, "step_sizes": sizes}
        sizes.append(dt)
        t += dt
        prev = nxt
    return {"found": False, "steps": max_steps, "time": t, "point": prev, "error": float("nan"), "bisections": 0,
            "step_sizes": sizes}


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
