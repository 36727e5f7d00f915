"""fs2d_hip_project_noalloc, the projection that allocates nothing: its output
must equal fs2d_hip_stage_project's bit for bit (which the parity tests pin to
the oracle), on two calls in a row with the second on a smaller grid, for both
sweep orders.  The arrays are the test's own torch device tensors."""
from __future__ import annotations

import numpy as np
import pytest

import fs2d_amd
from conftest import bits_equal, describe
from fs2d_amd import workloads as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def handle_projection(R, st, K, order, **kw):
    with fs2d_amd.Solver(R, pressure_iterations=K, order=order, **kw) as s:
        s.upload(*st[:3])
        s.project(K)
        _, u, v = s.download()
        p = s.pressure()
    return u, v, p


def noalloc_projection(R, st, K, order, **kw):
    dev = torch.device("cuda", 0)
    tu = torch.from_numpy(st[1].copy()).to(dev)
    tv = torch.from_numpy(st[2].copy()).to(dev)
    tp = torch.full((R, R), float("nan"), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev).cuda_stream
    fs2d_amd.project_noalloc(R, tu.data_ptr(), tv.data_ptr(), tp.data_ptr(), stream, pressure_iterations=K,
                             order=order, **kw)
    torch.cuda.synchronize()
    return tu.cpu().numpy(), tv.cpu().numpy(), tp.cpu().numpy()


CASES = [(fs2d_amd.GS_LEX, {}), (fs2d_amd.GS_LEX, {"gs_block": 4, "gs_lag": 1, "gs_tile_cols": 128}),
         (fs2d_amd.JACOBI, {"jacobi_block": 4}), (fs2d_amd.JACOBI, {"jacobi_block": 3})]


@pytest.mark.parametrize("order,kw", CASES)
@pytest.mark.parametrize("K", [0, 1, 20])
def test_noalloc_projection_equals_stage_project(order, kw, K):
    for call, R in enumerate((258, 67)):  # two calls in a row, the second smaller
        st = W.random_state(R, seed=31 * R + K, vel_scale=1.0)
        want = handle_projection(R, st, K, order, **kw)
        got = noalloc_projection(R, st, K, order, **kw)
        for name, g, w in zip(("u", "v", "pressure"), got, want):
            assert bits_equal(g, w), f"call {call} R={R} K={K} order={order} {kw} {name}: {describe(g, w)}"


def test_noalloc_projection_at_its_maximum_size():
    R = 1026
    st = W.vortex_state(R, seed=12, n_vortices=16) + (None, None, None)
    want = handle_projection(R, st, 40, fs2d_amd.GS_LEX)
    got = noalloc_projection(R, st, 40, fs2d_amd.GS_LEX)
    for name, g, w in zip(("u", "v", "pressure"), got, want):
        assert bits_equal(g, w), f"R={R} {name}: {describe(g, w)}"
