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


def bisect_crossing(system, back, x0, normal, dt, err, point, tol=1e-12, max_bisect=500):
    """Locate the crossing inside a step of length dt taken from the saved
    state `back` (find_intersection's loop, Source/return_map.cu:449-476):
    bisect the absolute bracket [a, b] of the sub-step length, restarting from
    `back` each time, until |n.(p - x0)| <= tol or max_bisect sub-steps.
    `err`, `point`: plane value and probe after the full step.  The system is
    left at the last sub-step evaluated.  Returns (h, point, err, iterations)."""
    a, b = 0.0, dt
    it, h = 0, dt
    while abs(err) > tol and it < max_bisect:
        it += 1
        h = 0.5 * (a + b)
        system.restore(back)
        system.step(h)
        point = np.asarray(system.probe(), float)
        err = plane_value(point, x0, normal)
        if err > 0.0:
            b = h
        else:
            a = h
    return h, point, err, it


def crossing_condition(steps, prev, nxt, x0, normal, rho, min_steps=3):
    """find_intersection's test (Source/return_map.cu:438-447): more than
    min_steps steps, the previous point inside the ball of radius rho around
    x0, a sign change of n.(p - x0) across the step, motion along n."""
    s1, s2 = plane_value(prev, x0, normal), plane_value(nxt, x0, normal)
    moving = float(np.dot(normal, np.asarray(nxt) - np.asarray(prev)))
    return steps > min_steps and np.linalg.norm(np.asarray(prev) - x0) < rho and s1 * s2 < 0.0 and moving > 0.0


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
        if crossing_condition(self.steps, self.prev, nxt, x0, normal, self.rho, self.min_steps):
            # the crossing lies inside this step: bisect its length from the saved state
            h, point, err, it = bisect_crossing(sysm, back, x0, normal, dt, plane_value(nxt, x0, normal), nxt,
                                                self.tol, self.max_bisect)
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


# ----------------------------------------------------------------------------- execute_return_map


def plane_points(n_rad=7, n_phi=7, radius=0.01):
    """construct_plane_rectangular (Source/return_map.cu:123-143): the grid
    x_j = j dx - eps, y_k = k dy - eps (dx = 2 eps / n_rad, dy = 2 eps / n_phi),
    z = 0, point (j, k) at index j*n_phi + k.  Returns [n_rad*n_phi, 3]."""
    dx, dy = 2.0 * radius / n_rad, 2.0 * radius / n_phi
    pts = np.zeros((n_rad * n_phi, 3))
    for j in range(n_rad):
        for k in range(n_phi):
            pts[j * n_phi + k] = (j * dx - radius, k * dy - radius, 0.0)
    return pts


def rotation_onto(n, m=(0.0, 0.0, 1.0)):
    """create_matrix3_direct / rotate_plane (Source/return_map.cu:146-242): the
    rotation taking the unit vector of m onto that of n, R = I + [u]x +
    [u]x^2 / (1 + cos a) with u = m x n, cos a = m.n.  For n = -m (where that
    formula divides by zero) the half turn about the x axis."""
    n = np.asarray(n, float)
    m = np.asarray(m, float)
    n = n / (np.linalg.norm(n) or 1.0)
    m = m / (np.linalg.norm(m) or 1.0)
    u = np.cross(m, n)
    c = float(np.dot(m, n))
    if 1.0 + c < 1e-14:
        axis = np.array([1.0, 0.0, 0.0]) if abs(m[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        axis = axis - np.dot(axis, m) * m
        axis /= np.linalg.norm(axis)
        return 2.0 * np.outer(axis, axis) - np.eye(3)
    ux = np.array([[0.0, -u[2], u[1]], [u[2], 0.0, -u[0]], [-u[1], u[0], 0.0]])
    return np.eye(3) + ux + ux @ ux / (1.0 + c)


def seed_state(g, u, point, value):
    """construct_physical_vector (Source/return_map.cu:407-423): the physical
    velocity of u with the probe point's value replaced by `value`, transformed
    back to the half spectrum.  Returns three new spectral fields."""
    out = []
    for c, v in zip(u, value):
        x = g.iFFTN(c)
        x[point] = float(v)
        out.append(g.FFTN(x))
    return out


def execute_return_map(g, u, f=None, point=(0, 0, 0), n_rad=7, n_phi=7, radius=0.01, shift_steps=930, rho=0.5,
                       max_steps=100000, tol=1e-12, max_bisect=500, min_steps=3, chain=False, trajectories=False):
    """execute_return_map (Source/return_map.cu:502-690): the section map of a
    disc of seed points around the probe.

    1. x0 = probe(u), n0 = the RHS at the probe (return_vector3_RHS).
    2. The section plane: u advanced shift_steps steps gives x0s = probe and
       ns = RHS there (the reference's "another cut plane", 930 steps).
    3. n_rad x n_phi seeds on the plane through x0 normal to n0
       (``plane_points`` rotated onto n0, translated to x0).
    4. Each seed state is u with the probe value replaced by the seed
       (``seed_state``); it is stepped until its probe crosses (x0s, ns) as
       find_intersection decides, and the crossing is located by bisection.

    chain=False (default): every seed starts from u, and all seeds advance
    together through ns3d_rk3_ssp_batch (one launch per pass for the whole
    disc; bitwise the steps each seed takes alone), the crossing members
    leaving the batch.  chain=True follows the reference literally: the seeds
    run one after another and each starts from the state where the previous
    one crossed ("advance solution", :681).

    Returns dict(x0, normal, x0_shift, normal_shift, seeds [S,3], points [S,3]
    (crossing probes; the last probe for seeds that did not cross), found [S],
    steps [S], time [S], errors [S], bisections [S][, trajectories: per seed
    the probe after every step, as written to test_point_<j>.dat])."""
    import torch
    point = tuple(int(v) for v in point)
    dt = g.dt
    base = NSSystem(g, [c.clone() for c in u], f, point)
    x0 = np.asarray(base.probe(), float)
    n0 = np.asarray(base.probe_rhs(), float)
    shifted = NSSystem(g, [c.clone() for c in u], f, point)
    if shift_steps:
        g.RK3_SSP(shifted.u, f, steps=shift_steps)
    x0s = np.asarray(shifted.probe(), float)
    ns = np.asarray(shifted.probe_rhs(), float)
    seeds = plane_points(n_rad, n_phi, radius) @ rotation_onto(n0).T + x0
    S = len(seeds)
    res = {"x0": x0, "normal": n0, "x0_shift": x0s, "normal_shift": ns, "seeds": seeds,
           "points": np.zeros((S, 3)), "found": np.zeros(S, bool), "steps": np.zeros(S, int),
           "time": np.zeros(S), "errors": np.full(S, np.nan), "bisections": np.zeros(S, int)}
    traj = [[] for _ in range(S)]

    def record(j, r):
        res["points"][j] = r["point"]
        res["found"][j] = r["found"]
        res["steps"][j] = r["steps"]
        res["time"][j] = r["time"]
        res["errors"][j] = r["error"]
        res["bisections"][j] = r["bisections"]

    if chain:
        cur = [c.clone() for c in u]
        for j in range(S):
            sysm = NSSystem(g, seed_state(g, cur, point, seeds[j]), f, point)
            search = CrossingSearch(sysm, x0s, ns, dt, rho, tol, max_bisect, min_steps)
            traj[j].append(np.asarray(sysm.probe(), float))
            for _ in range(max_steps):
                done = search.advance()
                traj[j].append(np.asarray(sysm.probe(), float) if done else search.prev.copy())
                if done:
                    break
            record(j, search.result if search.result is not None else search.not_found())
            g.set_params(dt, None)
            cur = sysm.u
    else:
        ub = torch.empty((S, 3) + tuple(u[0].shape), dtype=torch.complex128, device=u[0].device)
        for j in range(S):
            for b, c in enumerate(seed_state(g, u, point, seeds[j])):
                ub[j, b].copy_(c)
        members = list(range(S))
        prev = g.probe_batch(ub, *point)
        for j in range(S):
            traj[j].append(prev[j].copy())
        nsteps = 0
        while members and nsteps < max_steps:
            back = ub.clone()
            g.RK3_SSP_batch(ub, f, steps=1)
            nxt = g.probe_batch(ub, *point)
            leave = []
            for i, j in enumerate(members):
                if crossing_condition(nsteps, prev[i], nxt[i], x0s, ns, rho, min_steps):
                    sysm = NSSystem(g, [ub[i, 0], ub[i, 1], ub[i, 2]], f, point)
                    h, pt, err, it = bisect_crossing(sysm, [back[i, b] for b in range(3)], x0s, ns, dt,
                                                     plane_value(nxt[i], x0s, ns), nxt[i], tol, max_bisect)
                    g.set_params(dt, None)
                    record(j, {"point": pt, "found": True, "steps": nsteps + 1, "time": nsteps * dt + h,
                               "error": abs(err), "bisections": it})
                    traj[j].append(np.asarray(pt, float))
                    leave.append(i)
                else:
                    traj[j].append(nxt[i].copy())
            nsteps += 1
            if leave:
                keep = [i for i in range(len(members)) if i not in leave]
                members = [members[i] for i in keep]
                if not members:
                    break
                idx = torch.tensor(keep, device=ub.device)
                ub = ub.index_select(0, idx).contiguous()
                nxt = nxt[keep]
            prev = nxt
        for i, j in enumerate(members):
            record(j, {"point": prev[i], "found": False, "steps": nsteps, "time": nsteps * dt, "error": float("nan"),
                       "bisections": 0})
    if trajectories:
        res["trajectories"] = [np.array(t) for t in traj]
    return res


def write_return_map(res, directory):
    """The reference's text outputs of execute_return_map (Source/return_map.cu:
    586-690): test_point_<j>.dat (one "%.16le %.16le %.16le" line per probe
    value along seed j's trajectory; needs trajectories=True), res_3D.dat (the
    seeds) and res_3D_1.dat (the crossings)."""
    import os
    os.makedirs(directory, exist_ok=True)
    fmt = "%.16le %.16le %.16le\n"
    for j, t in enumerate(res.get("trajectories", [])):
        with open(os.path.join(directory, f"test_point_{j}.dat"), "w") as fh:
            for p in t:
                fh.write(fmt % tuple(p))
    for name, arr in (("res_3D.dat", res["seeds"]), ("res_3D_1.dat", res["points"])):
        with open(os.path.join(directory, name), "w") as fh:
            for p in arr:
                fh.write(fmt % tuple(p))
