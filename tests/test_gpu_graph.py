"""Whole-step graph replay of the solid step (sph_graphs) against the eager step.

Two particle systems on the same deck: one stepped eagerly, one through captured
graphs on a side stream (a steady run alternates between two captured argument sets, so
six steps replay each graph twice after its capture).  After every step, with the
program's own synchronisation and host reads in between, every field and every scalar of
the graph run must equal the eager run's bit for bit: the replayed kernels are the eager
step's, and the step has no host-side branch on device data.
"""
import ctypes

import numpy as np
import pytest
import torch

from sigmasph_amd import _native as N
from sigmasph_amd.config import SimConfig
from sigmasph_amd.particle_system import ParticleSystem
from sigmasph_amd.scenes import solid_block

pytestmark = pytest.mark.gpu


def _state(ps):
    order = np.argsort(ps.pid.to_numpy(), kind="stable")
    return [ps.x.to_numpy()[order], ps.v.to_numpy()[order], ps.sigma_dev.to_numpy()[order],
            ps.e_int.to_numpy()[order], ps.pressure.to_numpy()[order], ps.scal.cpu().numpy().copy(),
            ps.cons.cpu().numpy().copy()]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_equals_eager(gpu, precision):
    sc = solid_block(n=(24, 8, 1), two_d="axisymmetric", grips=0.05, ramp_time=1e-4,
                     div_velocity=(300.0, 0.0, 0.0))
    a = ParticleSystem(SimConfig(sc), precision=precision)
    b = ParticleSystem(SimConfig(sc), precision=precision)
    sa, sb = a.build_solver(), b.build_solver()
    sa.initialize()
    sb.initialize()
    N.check(b._lib.sph_graphs(b.ctx, 1), b.ctx, "sph_graphs")
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    for k in range(6):
        print("step %d" % k, flush=True)
        sa.step()
        with torch.cuda.stream(side):
            sb.step()
        torch.cuda.synchronize()
        for u, v in zip(_state(a), _state(b)):
            assert np.array_equal(u, v), "step %d differs" % k
    assert b.state.cur == a.state.cur and b.solid.hcur == a.solid.hcur
    N.check(b._lib.sph_graphs(b.ctx, 0), b.ctx, "sph_graphs")
    a.close()
    b.close()
