"""Launch shapes of the per-size cell lists (lanes per cell, grid cap,
per-XCD block order) and the autotuner that picks them: every shape gives
bit-identical residuals and active flags, in FP64 and FP32-mixed."""
import numpy as np
import pytest

from cipher_amd import abi
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _run(ctx, X):
    import torch
    act = torch.zeros(ctx.ncells, dtype=torch.int32, device=ctx.device)
    F = ctx.rhs(X, active=act)
    torch.cuda.synchronize()
    return F.cpu().numpy(), act.cpu().numpy()


SHAPES = [
    # (list, lanes, cap, map); map 2 and 3: the single-phase list of a mesh
    # with chemistry in a launch of its own, in launch order and per-XCD order
    (1, 4, 16384, 0), (1, 8, 4096, 1), (2, 8, 65536, 1), (0, 1, 16, 1), (3, 8, 1, 1), (2, 4, 3, 0),
    (0, 1, 16384, 2), (0, 1, 64, 3), (0, 1, 16384, 0),
]


@pytest.mark.parametrize("name", ["gg3d_amr", "ternary", "chem4", "aniso3d"])
@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_MIXED], ids=["fp64", "fp32mixed"])
def test_every_shape_gives_the_same_residual(name, precision):
    import torch
    from cipher_amd.runtime import Context
    from cipher_amd import workload as W
    if name == "aniso3d":
        wl = H._cache.setdefault(("aniso3d",), W.Workload(H.aniso_pair_yaml(res=(16, 16, 16), minlevel=2), "aniso3d"))
    elif name == "chem4":
        wl = H.chem()
    else:
        wl = getattr(H, name)()
    X = H.perturb(wl, H.evolve(wl, H.initial_state(wl), steps=4))
    ctx = Context(wl.cfg, precision=precision)
    ctx.upload_mesh(wl.mesh)
    Xd = torch.as_tensor(X).to(ctx.device)
    F0, A0 = _run(ctx, Xd)
    print(f"[shapes] {name} precision {precision} default {ctx.shapes()}", flush=True)
    P = wl.layout["pf_size"]
    capl = 4 if P <= 4 else (8 if P <= 8 else 16)
    chem = bool(wl.layout["dp_size"])
    for q, lanes, cap, mp in SHAPES:
        if mp & 2 and not chem:
            continue       # a launch of its own is for chemistry meshes only
        if q == 0 and chem and mp < 2 and cap != 16384:
            lanes = 2      # chemistry cells of list 0 may take more lanes
        if lanes > capl:
            continue
        ctx.set_shape(q, lanes, cap, mp)
        F, A = _run(ctx, Xd)
        assert np.array_equal(F, F0), f"{name}: list {q} shape {(lanes, cap, mp)} changes F"
        assert np.array_equal(A, A0)
    ms, shapes = ctx.autotune(Xd, reps=2)
    print(f"[shapes] {name} tuned {ms:.4f} ms {shapes}", flush=True)
    F, A = _run(ctx, Xd)
    assert np.array_equal(F, F0) and np.array_equal(A, A0)


def test_set_shape_rejects_bad_shapes():
    from cipher_amd.runtime import Context
    wl = H.gg2d()
    ctx = Context(wl.cfg)
    ctx.upload_mesh(wl.mesh)
    for q, lanes, cap, mp in [(1, 1, 16, 0), (2, 16, 16, 0), (0, 3, 16, 0), (0, 1, 0, 0), (0, 1, 16, 2),
                              (5, 1, 16, 0), (1, 2, 16, 2), (0, 1, 16, 4)]:
        with pytest.raises(abi.CipherError) as e:
            ctx.set_shape(q, lanes, cap, mp)
        assert e.value.code == abi.ERR_ARG
