"""CIPHER's time loop with remeshing (src/cipher.c:1589-1830) through
cipher_amd.timeloop.solve on the device against its CPU restatement
(tests/helpers.py cpu_timeloop: oracle RHS and PostStep, the same host
remesh), on 2D periodic grain growth on an adaptive quadtree (levels 3-6,
1660 leaves at the start) with two load cases of different temperature
rates and a remesh every 10 accepted steps (37 remeshes, 1660 -> ~2200
leaves): every chunk's case, accepted and rejected counts and leaf count
equal, times and next step sizes within 1e-10, the same final forest, the
same active sets and phase fractions within 1e-9, and the temperature
T0 + sum_k rate_k * time_k in every leaf (refinement copies it, coarsening
averages equal values).  With outputfreq 10 the output step writes a VTU
file at every 10th step: each file's leaves and phase ids equal the oracle's
output fields (src/cipher.c:1005-1080) of the CPU run's state at that step.
"""
import numpy as np
import pytest

from cipher_amd import abi
from tests import helpers as H

pytestmark = pytest.mark.gpu

TIMES = [20.0, 10.0]
RATES = [0.5, -0.5]
INTERVAL = 10


def _params(wl, graph):
    return abi.ts_params(1e-4, 1e-4, wl.cfg.timestepmin, 100.0, use_graph=graph)


def test_timeloop_with_remeshing_matches_cpu(tmp_path):
    import torch
    import oracle
    from cipher_amd.runtime import Context
    from cipher_amd import timeloop as TL
    from cipher_amd import vtu
    wl = H.gg2d_amr(amrinterval=INTERVAL)
    print(f"[timeloop] {wl.ncells} leaves, cases {TIMES} rates {RATES}, remesh every {INTERVAL} steps", flush=True)
    fc, Xc, chunks_c, rem_c, outs_c = H.cpu_timeloop(wl.cfg, wl.forest, H.initial_state(wl), _params(wl, False),
                                                     TIMES, RATES, INTERVAL, ofreq=INTERVAL)
    print(f"   cpu: {len(chunks_c)} chunks, {rem_c} remeshes, {fc.ncells} leaves at the end", flush=True)
    ctx = Context(wl.cfg)
    ctx.upload_mesh(wl.mesh)
    ctx.ts_setup(_params(wl, True))
    X = torch.as_tensor(H.initial_state(wl)).to(ctx.device)
    saved = wl.cfg.outputfreq
    wl.cfg.outputfreq = INTERVAL
    try:
        run = TL.solve(ctx, wl.cfg, wl.forest, X, _params(wl, True), times=TIMES, rates=RATES,
                       amrinterval=INTERVAL, output_dir=str(tmp_path))
    finally:
        wl.cfg.outputfreq = saved
    torch.cuda.synchronize()
    print(f"   gpu: {len(run.chunks)} chunks, {run.remeshes} remeshes, {run.forest.ncells} leaves at the end",
          flush=True)
    assert run.remeshes == rem_c > 10
    assert len(run.chunks) == len(chunks_c)
    for g, c in zip(run.chunks, chunks_c):
        assert (g[0], g[1], g[2], g[5]) == (c[0], c[1], c[2], c[5]), (g, c)
        assert g[3] == pytest.approx(c[3], rel=1e-10, abs=0.0)
        assert g[4] == pytest.approx(c[4], rel=1e-8, abs=0.0)
    assert run.t == sum(TIMES)
    assert np.array_equal(run.forest.level, fc.level) and np.array_equal(run.forest.anchor, fc.anchor)
    Xg = run.X.cpu().numpy()
    L = wl.layout
    m = H.state_mask(Xc, L, wl.cfg)
    assert np.array_equal(H.state_mask(Xg, L, wl.cfg), m), "active sets differ"
    errs = H.field_errors(np.where(m, Xg, 0.0), np.where(m, Xc, 0.0), L)
    print("   final state errors", errs, flush=True)
    assert errs["PF"] <= 1e-9
    T0 = H.initial_state(wl).reshape(-1, L["ndof"])[0, L["tm_offset"]]
    T = Xg.reshape(-1, L["ndof"])[:, L["tm_offset"]]
    np.testing.assert_allclose(T, T0 + RATES[0] * TIMES[0] + RATES[1] * TIMES[1], rtol=1e-12)
    # output steps
    assert [p.rsplit("_", 1)[1] for p in run.outputs] == [f"{st}.vtu" for st, _, _ in outs_c]
    D = L["ndof"]
    for path, (st, f, Xs) in zip(run.outputs, outs_c):
        _, conn, _, data = vtu.read(path)
        want = oracle.output(wl.cfg, f.flatten(), Xs, [(abi.OUT_PHASEID, 0)])[:, 0]
        assert conn.shape[0] == f.ncells, path
        # equal ids, except where the two largest interpolants of the CPU
        # state tie within the states' 1e-9 agreement
        for c in np.flatnonzero(data["phaseid"] != want):
            n = int(round(Xs[c * D]))
            h = np.sort(oracle.interpolant(wl.cfg.interpolation, Xs[c * D + L["pf_offset"]:c * D + L["pf_offset"] + n]))
            assert n > 1 and h[-1] - h[-2] < 1e-9, (path, c)
    print(f"   {len(run.outputs)} output files checked", flush=True)
    ctx.close()


def test_reference_example_with_rk2a_matches_cpu():
    """A configuration derived from the reference's GrainGrowth/256x256_N50
    example, which asks for '-ts_rk_type 2a' (Heun 2(1)): the time loop
    runs it with that method (the integrator set up from the configuration's
    TSRK type) to the time the CPU restatement reaches in 25 accepted steps
    from SetUpProblem; step counts, end time and next step size agree and
    the phase fields agree to 1e-9."""
    import torch
    from cipher_amd.runtime import Context
    from cipher_amd import timeloop as TL
    wl = H.ref_case("grain_growth")
    cfg = wl.cfg
    assert cfg.ts_rk_type == "2a"
    p = abi.ts_params(cfg.abstol, cfg.reltol, cfg.timestepmin, cfg.timestepmax, use_graph=True,
                      method=abi.rk_method(cfg.ts_rk_type))
    X0 = H.initial_state(wl)
    _, info = H.cpu_solve(wl, X0, 0.0, 1e30, cfg.timestepmin, p, max_steps=25)
    t_end = info["t"]
    fc, Xc, chunks_c, rem_c = H.cpu_timeloop(cfg, wl.forest, X0, p, [t_end], None, cfg.amrinterval)
    print(f"[timeloop 2a] {wl.ncells} cells, t_end {t_end:.6e}, cpu chunks {chunks_c}", flush=True)
    ctx = Context(cfg)
    ctx.upload_mesh(wl.mesh)
    ctx.ts_setup(p)
    X = torch.as_tensor(X0).to(ctx.device)
    run = TL.solve(ctx, cfg, wl.forest, X, p, times=[t_end], rates=None)
    torch.cuda.synchronize()
    print(f"   gpu chunks {run.chunks}", flush=True)
    assert len(run.chunks) == len(chunks_c) == 1
    g, c = run.chunks[0], chunks_c[0]
    assert (g[1], g[2]) == (c[1], c[2]) and c[1] >= 25
    assert g[3] == c[3] == t_end
    assert g[4] == pytest.approx(c[4], rel=1e-8, abs=0.0)
    Xg = run.X.cpu().numpy()
    L = wl.layout
    m = H.state_mask(Xc, L, cfg)
    assert np.array_equal(H.state_mask(Xg, L, cfg), m), "active sets differ"
    errs = H.field_errors(np.where(m, Xg, 0.0), np.where(m, Xc, 0.0), L)
    print("   final state errors", errs, flush=True)
    assert errs["PF"] <= 1e-9
