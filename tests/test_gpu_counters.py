"""The list counters of the right-hand side: eager calls take a ring of
counter sets cleared once per round, captured calls and the integrator's
stages a set of their own that each binning call clears.

* 3 rounds and more of eager calls on alternating states, in FP64 and
  FP32-mixed, reproduce the first call's residual and active flags bit for
  bit, also with integrator attempts and a PostStep between them;
* cipher_rhs captured into a HIP graph and replayed several times, with
  eager right-hand sides on other states between the replays, gives the
  eager result on every replay.
"""
import numpy as np
import pytest

from cipher_amd import abi
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _ctx(wl, precision=abi.FP64):
    from cipher_amd.runtime import Context
    ctx = Context(wl.cfg, precision=precision)
    ctx.upload_mesh(wl.mesh)
    return ctx


def _run(ctx, X):
    import torch
    act = torch.zeros(ctx.ncells, dtype=torch.int32, device=ctx.device)
    F = ctx.rhs(X, active=act)
    torch.cuda.synchronize()
    return F.cpu().numpy(), act.cpu().numpy()


@pytest.mark.parametrize("name", ["gg3d_amr", "ternary"])
@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_MIXED], ids=["fp64", "fp32mixed"])
def test_eager_calls_over_several_ring_rounds(name, precision):
    import torch
    wl = getattr(H, name)()
    XA = H.perturb(wl, H.evolve(wl, H.initial_state(wl), steps=4))
    XB = H.initial_state(wl)
    ctx = _ctx(wl, precision)
    ctx.ts_setup(abi.ts_params(1e-4, 1e-4, 1e-9, 1.0, use_graph=False))
    dA, dB = (torch.as_tensor(X).to(ctx.device) for X in (XA, XB))
    refA, refB = _run(ctx, dA), _run(ctx, dB)
    print(f"[counters] {name} precision={precision}: 200 eager calls", flush=True)
    for i in range(200):
        X, ref = (dA, refA) if i % 2 else (dB, refB)
        if i % 3 == 0:
            F, act = _run(ctx, X)
            assert np.array_equal(F, ref[0]) and np.array_equal(act, ref[1]), f"call {i}"
        else:
            ctx.rhs(X)
        if i in (70, 150):
            # integrator stages (their own counter set) and a PostStep
            # between eager calls
            Xs = dB.clone()
            ctx.ts_attempt(Xs, 1e-9)
            ctx.poststep(Xs)
    ctx.close()


def test_captured_rhs_replays_match_eager():
    import torch
    wl = H.gg3d_amr()
    XA = H.perturb(wl, H.evolve(wl, H.initial_state(wl), steps=4))
    XB = H.initial_state(wl)
    ctx = _ctx(wl)
    dA, dB = (torch.as_tensor(X).to(ctx.device) for X in (XA, XB))
    refA = _run(ctx, dA)[0]
    F = torch.empty_like(dA)
    Fb = torch.empty_like(dB)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        ctx.rhs(dA, F=F)                       # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        ctx.rhs(dA, F=F)
    torch.cuda.synchronize()
    print("[counters] captured cipher_rhs: 5 replays", flush=True)
    for rep in range(5):
        F.fill_(float("nan"))
        for _ in range(3):
            ctx.rhs(dB, F=Fb)                  # eager calls on the ring between replays
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(F.cpu().numpy(), refA), f"replay {rep}"
    del g
    ctx.close()
