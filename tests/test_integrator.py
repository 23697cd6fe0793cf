"""The adaptive Runge-Kutta integrator: host step controller (CPU), the
Bogacki-Shampine tableau (order of the CPU restatement), and the device
integrator against the CPU restatement (GPU)."""
import math

import numpy as np
import pytest

import oracle
from cipher_amd import abi
from tests import helpers as H


def _params(**kw):
    d = dict(atol=1e-4, rtol=1e-4, dt_min=1e-6, dt_max=10.0)
    d.update(kw)
    return abi.ts_params(**d)


@pytest.fixture(scope="module")
def lib():
    import os
    if not os.path.exists(abi.LIB_PATH):
        from cipher_amd.csrc import build
        build.build()
    return abi.load_library()


def test_controller_accepts_grows_rejects_shrinks(lib):
    from cipher_amd.runtime import ts_adapt
    p = _params()
    acc, nxt = ts_adapt(p, abi.TsCtrl(), 0.1, 0.25)
    assert acc and nxt > 0.1
    acc, nxt = ts_adapt(p, abi.TsCtrl(), 0.1, 4.0)
    assert not acc and nxt < 0.1
    # smaller error, larger next step (elementary controller, no history)
    prev = 0.0
    for e in (8.0, 2.0, 1.0, 0.5, 0.1, 0.01):
        _, nxt = ts_adapt(p, abi.TsCtrl(), 0.1, e)
        assert nxt >= prev
        prev = nxt
    # the arctan limiter bounds a shrink to 1 + atan(-1) (times the safety);
    # a zero error takes the largest ratio, then the limits clamp
    _, nxt = ts_adapt(p, abi.TsCtrl(), 0.1, 0.0)
    assert nxt == pytest.approx(0.1 * p.safety * p.clip_hi)
    _, nxt = ts_adapt(p, abi.TsCtrl(), 0.1, 1e300)
    assert nxt == pytest.approx(0.1 * p.safety * (1.0 + math.atan(-1.0)), rel=1e-9)
    _, nxt = ts_adapt(_params(dt_max=0.15), abi.TsCtrl(), 0.1, 1e-8)
    assert nxt == 0.15


def test_controller_minimum_step_and_nonfinite(lib):
    from cipher_amd.runtime import ts_adapt
    p = _params(dt_min=1e-3)
    acc, nxt = ts_adapt(p, abi.TsCtrl(), 1e-3, 50.0)
    assert acc and nxt == 1e-3                     # accepted at the minimum step
    acc, nxt = ts_adapt(p, abi.TsCtrl(), 0.1, float("nan"))
    assert not acc and nxt == pytest.approx(0.01)
    with pytest.raises(abi.CipherError) as e:
        ts_adapt(p, abi.TsCtrl(), 1e-3, float("inf"))
    assert e.value.code == 4


def test_controller_history_and_consecutive_rejections(lib):
    from cipher_amd.runtime import ts_adapt
    p = _params()
    # with an accepted history, a rising error damps the growth (PI action)
    c1 = abi.TsCtrl()
    ts_adapt(p, c1, 0.1, 0.1)
    _, n_hist = ts_adapt(p, c1, 0.1, 0.5)
    _, n_fresh = ts_adapt(p, abi.TsCtrl(), 0.1, 0.5)
    assert n_hist < n_fresh
    # a second rejection in a row shortens more aggressively
    c2 = abi.TsCtrl()
    _, first = ts_adapt(p, c2, 0.1, 3.0)
    _, second = ts_adapt(p, c2, 0.1, 3.0)
    assert second == pytest.approx(first * p.reject_safety, rel=1e-12)
    assert c2.rejected_last == 1


@pytest.mark.parametrize("method", [abi.RK_3BS, abi.RK_2A], ids=["3bs", "2a"])
def test_controller_arithmetic_matches_libm(lib, method):
    """The controller evaluates pow and atan from their series in basic IEEE
    operations, so that the host and the device take bit-identical
    decisions; against libm's pow and atan the next step agrees to a few
    ulps over errors from 1e-8 to 1e4, with and without history.  The
    exponents divide by the method's order k: 3 for Bogacki-Shampine 3(2),
    2 for Heun 2(1)."""
    from cipher_amd.runtime import ts_adapt
    p = _params(method=method)
    k = 3.0 if method == abi.RK_3BS else 2.0
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(3000):
        err = 10.0 ** rng.uniform(-8, 4)
        prev = 10.0 ** rng.uniform(-6, 0) if rng.random() < 0.5 else 0.0
        c = abi.TsCtrl()
        c.err_prev = prev
        acc, nxt = ts_adapt(p, c, 0.1, err)
        r = (1.0 / err) ** ((0.6 if prev > 0 else 1.0) / k)
        if prev > 0:
            r *= (1.0 / prev) ** (-0.2 / k)
        ratio = min(max(p.safety * (1.0 + math.atan(r - 1.0)), p.clip_lo), p.clip_hi)
        ref = min(max(0.1 * ratio, p.dt_min), p.dt_max)
        assert acc == (err <= 1.0)
        worst = max(worst, abs(nxt - ref) / ref)
    assert worst <= 1e-14


def test_bogacki_shampine_orders():
    """On the oracle RHS with frozen active sets: the embedded error estimate
    is O(dt^3) and the step's local error O(dt^4) (halving dt divides them by
    ~8 and ~16), which fixes the tableau."""
    wl = H.gg2d()
    X = H.evolve(wl, H.initial_state(wl), steps=4)
    F = oracle.rhs(wl.cfg, wl.mesh, X)
    dt0 = 0.01 / np.abs(F).max()
    est, loc = [], []
    for k in range(3):
        dt = dt0 / 2 ** k
        Xn, Xh, _ = H.cpu_attempt(wl, X, dt, 1e-4, 1e-4)
        # reference: 16 substeps of the same scheme
        Y = X.copy()
        for _ in range(16):
            Y, _, _ = H.cpu_attempt(wl, Y, dt / 16, 1e-4, 1e-4)
        est.append(np.abs(Xn - Xh).max())
        loc.append(np.abs(Xn - Y).max())
    r_est = est[1] / est[2]
    r_loc = loc[1] / loc[2]
    assert 6.0 < r_est < 10.0, est
    assert 12.0 < r_loc < 20.0, loc


def test_heun_orders():
    """Heun 2(1) ('-ts_rk_type 2a') on the oracle RHS with frozen active
    sets: the embedded solution is the explicit Euler step (its difference
    from the step is O(dt^2): halving dt divides it by ~4) and the step's
    local error is O(dt^3) (~8), which fixes the tableau; the embedded
    solution is exactly one Euler step."""
    wl = H.gg2d()
    X = H.evolve(wl, H.initial_state(wl), steps=4)
    F = oracle.rhs(wl.cfg, wl.mesh, X)
    dt0 = 0.01 / np.abs(F).max()
    est, loc = [], []
    for k in range(3):
        dt = dt0 / 2 ** k
        Xn, Xh, _ = H.cpu_attempt(wl, X, dt, 1e-4, 1e-4, method=abi.RK_2A)
        assert np.array_equal(Xh, X + dt * F)
        Y = X.copy()
        for _ in range(16):
            Y, _, _ = H.cpu_attempt(wl, Y, dt / 16, 1e-4, 1e-4, method=abi.RK_2A)
        est.append(np.abs(Xn - Xh).max())
        loc.append(np.abs(Xn - Y).max())
    r_est = est[1] / est[2]
    r_loc = loc[1] / loc[2]
    assert 3.0 < r_est < 5.0, est
    assert 6.0 < r_loc < 10.0, loc


def test_rk_method_names():
    assert abi.rk_method("3bs") == abi.RK_3BS and abi.rk_method("2a") == abi.RK_2A
    with pytest.raises(abi.CipherError) as e:
        abi.rk_method("5dp")
    assert e.value.code == abi.ERR_UNSUPPORTED


def test_loadcases_integrate_the_temperature_rate():
    """Load cases (src/cipher.c:1589-1606) on the oracle: with no conduction
    the temperature integrates the piecewise constant rate exactly, each
    case ends at the running sum of the case lengths, and a case set to the
    configuration's own state (no rate) follows cpu_solve step for step."""
    wl = H.chem()
    p = abi.ts_params(1e-4, 1e-4, 1e-6, 1e-3, use_graph=False)
    X0 = H.initial_state(wl)
    L = wl.layout
    times, rates = [4e-4, 3e-4], [1.0e4, -2.0e4]
    X, infos = H.cpu_loadcases(wl, X0, times, rates, p)
    assert [i["t"] for i in infos] == [4e-4, 4e-4 + 3e-4]
    T0 = X0.reshape(-1, L["ndof"])[:, L["tm_offset"]]
    T = X.reshape(-1, L["ndof"])[:, L["tm_offset"]]
    np.testing.assert_allclose(T, T0 + 1.0e4 * 4e-4 - 2.0e4 * 3e-4, rtol=1e-12)
    assert wl.cfg.struct.has_temperature_rate == 0      # restored
    X1, i1 = H.cpu_loadcases(wl, X0, [7e-4], None, p)
    X2, i2 = H.cpu_solve(wl, X0, 0.0, 7e-4, p.dt_min, p)
    assert np.array_equal(X1, X2) and i1[0]["steps"] == i2["steps"]


pytestmark_gpu = pytest.mark.gpu


def _ctx(wl, params):
    from cipher_amd.runtime import Context
    ctx = Context(wl.cfg)
    ctx.upload_mesh(wl.mesh)
    ctx.ts_setup(params)
    return ctx


def _dev(ctx, X):
    import torch
    return torch.as_tensor(np.ascontiguousarray(X)).to(ctx.device)


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("method", [abi.RK_3BS, abi.RK_2A], ids=["3bs", "2a"])
@pytest.mark.parametrize("name", ["gg2d", "ternary"])
def test_attempt_matches_cpu(name, method):
    wl = getattr(H, name)()
    X = H.evolve(wl, H.initial_state(wl), steps=3)
    F = oracle.rhs(wl.cfg, wl.mesh, X)
    L = wl.layout
    Fv = F.reshape(-1, L["ndof"])
    dt = 0.02 / np.abs(Fv[:, L["pf_offset"]:L["pf_offset"] + L["pf_size"]]).max()
    dt = min(dt, 0.02 / (np.abs(Fv) / np.maximum(np.abs(X.reshape(-1, L["ndof"])), 1.0)).max())
    p = _params(method=method)
    ctx = _ctx(wl, p)
    print(f"[ts attempt] {name} method {method} dt={dt:.3e}", flush=True)
    Xd = _dev(ctx, X)
    err_g = ctx.ts_attempt(Xd, dt)
    Xn, _, err_c = H.cpu_attempt(wl, X, dt, p.atol, p.rtol, method=method)
    assert err_g == pytest.approx(err_c, rel=1e-8, abs=0.0)
    ctx.ts_accept(Xd)
    Xo, rc = oracle.poststep(wl.cfg, wl.mesh, Xn)
    m = H.state_mask(Xo, L, wl.cfg)
    Xg = _host(Xd)
    assert np.array_equal(H.state_mask(Xg, L, wl.cfg), m)
    errs = H.field_errors(np.where(m, Xg, 0.0), np.where(m, Xo, 0.0), L)
    print("   state errors", errs, flush=True)
    for f, e in errs.items():
        assert e <= 1e-10, f


@pytest.mark.gpu
@pytest.mark.parametrize("name,method", [("gg2d", abi.RK_3BS), ("gg3d_amr", abi.RK_3BS), ("gg2d", abi.RK_2A),
                                         ("gg3d_amr", abi.RK_2A)], ids=["gg2d-3bs", "gg3d_amr-3bs", "gg2d-2a",
                                                                        "gg3d_amr-2a"])
def test_trajectory_matches_cpu(name, method):
    """Full adaptive trajectory: accepted/rejected counts, step sizes and the
    final phase fields agree with the CPU restatement (same controller); the
    Heun 2(1) runs replay each step from a captured HIP graph."""
    wl = getattr(H, name)()
    X0 = H.initial_state(wl)
    p = _params(atol=1e-3, rtol=1e-3, dt_min=1e-4, dt_max=2.0, method=method, use_graph=method == abi.RK_2A)
    t_end = 4.0 if name == "gg2d" else 1.5
    Xc, info = H.cpu_solve(wl, X0, 0.0, t_end, 1e-4, p)
    ctx = _ctx(wl, p)
    Xd = _dev(ctx, X0)
    print(f"[ts solve] {name} method {method} cpu steps={info['steps']} rejected={info['rejected']}", flush=True)
    st, dt_next = ctx.ts_solve(Xd, 0.0, t_end, 1e-4)
    print(f"   gpu steps={st.steps} rejected={st.rejected} t={st.t} dt={st.dt}", flush=True)
    assert (st.steps, st.rejected) == (info["steps"], info["rejected"])
    assert st.t == info["t"]
    # the same controller decisions (the error norms differ by rounding only)
    assert dt_next == pytest.approx(info["dt"], rel=1e-8, abs=0.0)
    # several attempts per host synchronisation
    attempts = st.steps + st.rejected
    print(f"   {attempts} attempts, {st.host_syncs} host syncs", flush=True)
    assert st.host_syncs <= attempts // 3 + 6
    Xg = _host(Xd)
    L = wl.layout
    m = H.state_mask(Xc, L, wl.cfg)
    assert np.array_equal(H.state_mask(Xg, L, wl.cfg), m)
    errs = H.field_errors(np.where(m, Xg, 0.0), np.where(m, Xc, 0.0), L)
    print("   final state errors", errs, flush=True)
    assert errs["PF"] <= 1e-9


@pytest.mark.gpu
def test_graph_replays_match_eager():
    """The captured attempt, replayed over many steps with eager work between
    the replays, gives the same trajectory bits as eager launches."""
    wl = H.gg3d_amr()
    X0 = H.initial_state(wl)
    pe = _params(atol=1e-3, rtol=1e-3, dt_min=1e-4, dt_max=2.0)
    pg = _params(atol=1e-3, rtol=1e-3, dt_min=1e-4, dt_max=2.0, use_graph=True)
    ce, cg = _ctx(wl, pe), _ctx(wl, pg)
    Xe, Xg = _dev(ce, X0), _dev(cg, X0)
    se, de = ce.ts_solve(Xe, 0.0, 1.0, 1e-4)
    sg, dg = cg.ts_solve(Xg, 0.0, 1.0, 1e-4)
    print(f"[ts graph] eager {se.steps}/{se.rejected} graph {sg.steps}/{sg.rejected}", flush=True)
    assert (se.steps, se.rejected, se.t, de) == (sg.steps, sg.rejected, sg.t, dg)
    assert np.array_equal(_host(Xe), _host(Xg))
    # the FP32-mixed path changes the captured launches: recaptured
    cg.set_precision(abi.FP32_MIXED)
    ce.set_precision(abi.FP32_MIXED)
    se2, _ = ce.ts_solve(Xe, 1.0, 1.5, de)
    sg2, _ = cg.ts_solve(Xg, 1.0, 1.5, dg)
    assert (se2.steps, se2.rejected) == (sg2.steps, sg2.rejected)
    assert np.array_equal(_host(Xe), _host(Xg))
