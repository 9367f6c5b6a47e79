"""GPU test of ns3d_rk3_ssp_steps_noalloc, the entry point that allocates
nothing: its results must be bitwise those of ns3d_rk3_ssp_steps (same
kernels, same order) on two calls in a row, the second with a smaller grid,
in every precision; the context's workspace must not grow; grids beyond the
fixed scratch are refused before any launch.  Calls go through the C ABI."""
import ctypes

import numpy as np
import pytest

from ns3dperiodic_amd import fields
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _steps(g, fn, u, f, nsteps):
    return fn(g.ctx, nsteps, *g._c3(u), *g._f3(f))


def _fields(g, N, seed, forced):
    from tests.gpu_util import to_dev
    u0 = fields.random_solenoidal(N, seed=seed)
    f = fields.kolmogorov_forcing_sin(N, n=2, amp=0.3) if forced else None
    return u0, to_dev(u0), (to_dev(f) if forced else None)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32s", "fp32i"])
def test_noalloc_bitwise_equals_rk3_ssp_steps(precision):
    import torch
    from ns3dperiodic_amd import NS3D, _lib
    lib = _lib.load()
    # the largest grid the fixed scratch holds, then a smaller one
    for N, forced, seed in ((64, True, 21), (32, False, 22)):
        print(f"\n[case] noalloc vs rk3_ssp_steps {precision} {N}^3 forced={forced}", flush=True)
        kw = dict(Re=200.0, dt=2e-3, precision=precision)
        gn, gr = NS3D(N, **kw), NS3D(N, **kw)
        _, un, fd = _fields(gn, N, seed, forced)
        ur = [t.clone() for t in un]
        ws0 = lib.ns3d_workspace_bytes(gn.ctx)
        _lib.check(_steps(gn, lib.ns3d_rk3_ssp_steps_noalloc, un, fd, 3), "noalloc")
        _lib.check(_steps(gr, lib.ns3d_rk3_ssp_steps, ur, fd, 3), "rk3_ssp_steps")
        torch.cuda.synchronize()
        for a, b in zip(un, ur):
            assert torch.equal(a, b)
        assert bool(torch.isfinite(torch.view_as_real(un[0])).all())
        # the context's own workspace was neither used nor grown
        assert lib.ns3d_workspace_bytes(gn.ctx) == ws0
        gn.close()
        gr.close()


def test_noalloc_matches_oracle_and_refuses_oversize():
    import torch
    from ns3dperiodic_amd import NS3D, _lib
    from oracle.oracle import Oracle
    from tests.gpu_util import to_dev, to_host
    lib = _lib.load()
    print("\n[case] noalloc 32^3 FP64 vs oracle, 2 steps", flush=True)
    N = 32
    g, o = NS3D(N, Re=100.0, dt=2e-3), Oracle(N, Re=100.0, dt=2e-3)
    u0 = fields.random_solenoidal(N, seed=31)
    f = fields.kolmogorov_forcing_sin(N, n=2, amp=0.3)
    ud, fd = to_dev(u0), to_dev(f)
    _lib.check(_steps(g, lib.ns3d_rk3_ssp_steps_noalloc, ud, fd, 2), "noalloc")
    ref = o.rk3(u0, f, steps=2)
    assert rel_l2(to_host(ud), ref) < 1e-12
    g.close()
    print("[case] noalloc refuses 96^3 (scratch beyond 64 MiB)", flush=True)
    big = NS3D(96, Re=100.0, dt=1e-3)
    ub = big.cfield(3)
    for t in ub:
        t.fill_(1.0 + 0.5j)
    st = _steps(big, lib.ns3d_rk3_ssp_steps_noalloc, ub, None, 1)
    assert st == _lib.NS3D_ERR_CAPACITY
    torch.cuda.synchronize()
    assert all(bool((t == (1.0 + 0.5j)).all()) for t in ub)   # nothing launched
    big.close()
