"""Step graphs: fs2d_hip_step_n / fs2d_hip_solve replaying captured hipGraphs
must give exactly what the same launches issued one by one give.

Each replay is checked against an eager handle, with downloads, syncs, source
changes and a change of dt between replays (a new dt is a new captured state).
"""
from __future__ import annotations

import numpy as np
import pytest

import fs2d_amd
from conftest import bits_equal, describe
from fs2d_amd import workloads as W

pytestmark = pytest.mark.gpu

GS, JAC = fs2d_amd.GS_LEX, fs2d_amd.JACOBI


def pair(R, order, gsb, block, kd=12, kp=24):
    kw = dict(diffuse_iterations=kd, pressure_iterations=kp, order=order, gs_block=gsb, jacobi_block=block)
    return (fs2d_amd.Solver(R, launch_mode=fs2d_amd.LAUNCH_EAGER, **kw),
            fs2d_amd.Solver(R, launch_mode=fs2d_amd.LAUNCH_GRAPH, **kw))


@pytest.mark.parametrize("order,gsb,block", [(GS, 8, 0), (GS, 2, 0), (GS, 1, 0), (JAC, 0, 0), (JAC, 0, 3)])
@pytest.mark.parametrize("R", [34, 130])
def test_graph_replays_match_eager(R, order, gsb, block):
    st = W.vortex_state(R, seed=9, n_vortices=16)
    rs, us, vs = W.disc_sources(R, radius=R / 16)
    eager, graph = pair(R, order, gsb, block)
    try:
        for s in (eager, graph):
            s.upload(*st)
            s.set_sources(rs, us, vs)
        dt = 0.05
        for i, chunk in enumerate((1, 1, 3, 13, 2, 5)):
            if i == 3:
                for s in (eager, graph):
                    s.set_sources(None, us, None)
            if i == 4:
                dt = 0.03
            eager.step_n(chunk, 1e-4, 1e-5, dt)
            graph.step_n(chunk, 1e-4, 1e-5, dt)
            if i % 2:
                graph.sync()
            ge, gg = eager.download(), graph.download()
            for f, name in enumerate(("density", "u", "v")):
                assert bits_equal(gg[f], ge[f]), f"{name} after chunk {i}: {describe(gg[f], ge[f])}"
        assert eager.graph_count() == 0
        assert 1 <= graph.graph_count() <= 32
    finally:
        eager.close()
        graph.close()


def test_solve_host_arrays_with_graphs():
    """The synchronous drop-in (fluid_solver::solve) also replays graphs."""
    R = 66
    st = W.vortex_state(R, seed=4, n_vortices=8)
    rs, us, vs = W.disc_sources(R, radius=5)
    eager, graph = pair(R, GS, 8, 0, kd=20, kp=20)
    try:
        a = [x.copy() for x in st]
        b = [x.copy() for x in st]
        for _ in range(6):
            eager.solve(a[0], rs, 1e-4, a[1], a[2], us, vs, 1e-5, 0.1)
            graph.solve(b[0], rs, 1e-4, b[1], b[2], us, vs, 1e-5, 0.1)
            for f in range(3):
                assert bits_equal(b[f], a[f]), describe(b[f], a[f])
        assert graph.graph_count() >= 1
    finally:
        eager.close()
        graph.close()
