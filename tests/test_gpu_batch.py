"""Batched ensembles (cipher_batch_*) against the members run alone and the oracle.

The batch runs the single-mesh kernel bodies behind a per-block member
lookup, so every member's right-hand side, active flags, PostStep and error
norm must equal (bit for bit) what its own context gives on the same slice,
in FP64 and in FP32-mixed; the FP64 results also meet the oracle bar of the
single-mesh tests.  Trajectories of the per-member adaptive integrator equal
the single-context integrator's, step counts included.
"""
import numpy as np
import pytest

import oracle
from cipher_amd import abi
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _contexts(wls, precision=abi.FP64):
    from cipher_amd.runtime import Context
    out = []
    for wl in wls:
        c = Context(wl.cfg, precision=precision)
        c.upload_mesh(wl.mesh)
        out.append(c)
    return out


def _cat(states):
    import torch
    return torch.as_tensor(np.ascontiguousarray(np.concatenate(states))).to("cuda")


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _members_alone(ctxs, states, precision):
    """Each member's F and active flags from its own context."""
    import torch
    Fs, As = [], []
    for c, X in zip(ctxs, states):
        c.set_precision(precision)
        Xd = torch.as_tensor(X).to(c.device)
        act = torch.zeros(c.ncells, dtype=torch.int32, device=c.device)
        Fs.append(_host(c.rhs(Xd, active=act)))
        As.append(_host(act))
    return np.concatenate(Fs), np.concatenate(As)


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_MIXED], ids=["fp64", "fp32mixed"])
def test_batch_rhs_equals_members_alone(precision):
    import torch
    from cipher_amd.runtime import Batch
    wls = H.ensemble()
    states = H.ensemble_states()
    ctxs = _contexts(wls)
    Fref, Aref = _members_alone(ctxs, states, precision)
    b = Batch(ctxs, precision=precision)
    print(f"[batch-rhs] {len(wls)} members, {b.cell_offsets[-1]} cells, precision {precision}", flush=True)
    X = _cat(states)
    act = torch.zeros(b.cell_offsets[-1], dtype=torch.int32, device=X.device)
    F = _host(b.rhs(X, active=act))
    for i, wl in enumerate(wls):
        s = slice(b.offsets[i], b.offsets[i + 1])
        assert np.array_equal(F[s], Fref[s]), f"member {i} ({wl.name}): F differs"
    assert np.array_equal(_host(act), Aref)
    # twice in a row: counters and lists are rebuilt per call
    F2 = _host(b.rhs(X))
    assert np.array_equal(F2, F)
    b.close()


def test_batch_rhs_fp64_meets_the_oracle_bar():
    from cipher_amd.runtime import Batch
    wls = H.ensemble()
    states = H.ensemble_states()
    b = Batch(_contexts(wls))
    F = _host(b.rhs(_cat(states)))
    for i, (wl, X) in enumerate(zip(wls, states)):
        Fo = oracle.rhs(wl.cfg, wl.mesh, X)
        errs = H.field_errors(F[b.offsets[i]:b.offsets[i + 1]], Fo, wl.layout)
        spread = H.ulp_spread(wl, X, Fo)
        print(f"[batch-oracle] {wl.name} {errs} spread {spread}", flush=True)
        for f, e in errs.items():
            # grain growth is bit-exact; with chemistry the single-mesh bar:
            # 1e-12 or 4x the oracle's own one-ulp spread
            bar = 0.0 if wl.layout["dp_size"] == 0 else max(1e-12, 4.0 * spread[f])
            assert e <= bar, f"{wl.name} {f}: {e:.3e} > {bar:.3e}"
    b.close()


def test_batch_poststep_and_norms_equal_members_alone():
    import torch
    from cipher_amd.runtime import Batch
    wls = H.ensemble()
    states = H.ensemble_states()
    ctxs = _contexts(wls)
    stage, ref, norms = [], [], []
    for c, wl, X in zip(ctxs, wls, states):
        F = oracle.rhs(wl.cfg, wl.mesh, X)
        L = wl.layout
        D = L["ndof"]
        Fv = F.reshape(-1, D)
        m = np.abs(Fv[:, L["pf_offset"]:L["pf_offset"] + L["pf_size"]]).max()
        r = (np.abs(Fv) / np.maximum(np.abs(X.reshape(-1, D)), 1.0)).max()
        Xs = X + min(0.05 / m, 0.05 / r) * F
        stage.append(Xs)
        ref.append(_host(c.poststep(torch.as_tensor(Xs).to(c.device))))
        norms.append(c.wrms_norm(torch.as_tensor(Xs).to(c.device), torch.as_tensor(X).to(c.device), 1e-4, 1e-4))
    b = Batch(ctxs)
    U = _cat(stage)
    Y = _cat(states)
    got = b.wrms_norms(U, Y, 1e-4, 1e-4)
    assert got == norms
    Xp = _host(b.poststep(U))
    assert np.array_equal(Xp, np.concatenate(ref))
    b.close()


def test_batch_poststep_names_the_overflowing_member():
    from cipher_amd import workload as W
    from cipher_amd.runtime import Batch, Context
    wl_ok = H.ensemble(0)
    wl_bad = W.Workload(W.grain_growth_yaml(res=(16, 16), ngrains=24, capacity=2), "cap")
    ctxs = _contexts([wl_ok, wl_bad])
    X0 = H.initial_state(wl_ok)
    # a state of the crowded member whose union overflows: every cell lists
    # its own label only, so the union over neighbours exceeds 2
    L = wl_bad.layout
    D = L["ndof"]
    Xb = np.zeros(wl_bad.ncells * D)
    Xv = Xb.reshape(-1, D)
    Xv[:, L["as_offset"]] = 1.0
    Xv[:, L["as_offset"] + 1] = wl_bad.phase
    Xv[:, L["pf_offset"]] = 1.0
    b = Batch(ctxs)
    with pytest.raises(abi.CipherError) as e:
        b.poststep(_cat([X0, Xb]))
    assert e.value.code == abi.ERR_CAPACITY and e.value.member == 1
    b.close()


@pytest.mark.parametrize("precision,method", [(abi.FP64, abi.RK_3BS), (abi.FP32_MIXED, abi.RK_3BS),
                                              (abi.FP64, abi.RK_2A)], ids=["fp64", "fp32mixed", "fp64-2a"])
def test_batch_ts_solve_equals_members_alone(precision, method):
    import torch
    from cipher_amd.runtime import Batch
    wls = H.ensemble()
    states = H.ensemble_states()
    # per-member step sizes and end times; the phase-field-only members use
    # nondimensional time, the chemical ones seconds
    t_end, dt0 = [], []
    for i, wl in enumerate(wls):
        scale = 1.0 if wl.layout["dp_size"] == 0 else 1e-9
        t_end.append(scale * (0.6 + 0.2 * i))
        dt0.append(scale * 0.05)
    params = abi.ts_params(1e-4, 1e-4, 1e-16, 1.0, use_graph=True, method=method)
    ref_X, ref_st = [], []
    ctxs = _contexts(wls, precision)
    for i, (c, X) in enumerate(zip(ctxs, states)):
        c.ts_setup(params)
        Xd = torch.as_tensor(X).to(c.device)
        st, _ = c.ts_solve(Xd, 0.0, t_end[i], dt0[i], max_steps=8)
        ref_X.append(_host(Xd))
        ref_st.append((st.steps, st.rejected, st.t, st.dt))
    b = Batch(ctxs, precision=precision)
    b.ts_setup(params)
    X = _cat(states)
    print(f"[batch-ts] {len(wls)} members precision {precision} method {method}", flush=True)
    stats, dts = b.ts_solve(X, [0.0] * len(wls), t_end, dt0, max_steps=8)
    got = _host(X)
    for i, wl in enumerate(wls):
        st = stats[i]
        print(f"   {wl.name}: steps {st.steps} rejected {st.rejected} t {st.t:.3e}", flush=True)
        assert (st.steps, st.rejected, st.t, st.dt) == ref_st[i], f"member {i} ({wl.name})"
        assert np.array_equal(got[b.offsets[i]:b.offsets[i + 1]], ref_X[i]), f"member {i} ({wl.name}): state differs"
    b.close()


def test_batch_autotune_keeps_results():
    from cipher_amd.runtime import Batch
    wls = H.ensemble()
    b = Batch(_contexts(wls), precision=abi.FP32_MIXED)
    X = _cat(H.ensemble_states())
    F0 = _host(b.rhs(X))
    ms, shapes = b.autotune(X, reps=3)
    print(f"[batch-tune] {ms:.4f} ms, shapes {shapes}", flush=True)
    assert ms > 0.0 and shapes
    for ncp, capl, q, lanes, cap, mp in shapes:
        # map 2 and 3: the single-phase list of a chemistry class in a launch of its own
        assert (1 << q) <= lanes <= capl and cap >= 1
        assert mp in (0, 1) or (q == 0 and ncp > 1 and mp in (2, 3))
    assert np.array_equal(_host(b.rhs(X)), F0)
    b.close()


def test_batch_loadcases_equal_members_alone():
    """Load cases through the batch (per-member case lengths and temperature
    rates, cipher_batch_set_temperature_rate) give each member the counts,
    times and state of cipher_ts_loadcases on its own context, bit for bit."""
    import torch
    from cipher_amd.runtime import Batch
    wls = H.ensemble()
    states = H.ensemble_states()
    times, rates = [], []
    for i, wl in enumerate(wls):
        scale = 1.0 if wl.layout["dp_size"] == 0 else 1e-9
        times.append([scale * 0.3, 0.0, scale * (0.2 + 0.1 * i)])
        rates.append([5.0e3 * (i + 1), 1.0e9, -2.0e4])
    params = abi.ts_params(1e-4, 1e-4, 1e-16, 1.0, use_graph=True)
    ctxs = _contexts(wls)
    ref_X, ref_st = [], []
    for i, (c, X) in enumerate(zip(ctxs, states)):
        c.ts_setup(params)
        Xd = torch.as_tensor(X).to(c.device)
        st = c.solve_loadcases(Xd, times[i], rates[i])
        ref_X.append(_host(Xd))
        ref_st.append([(s.steps, s.rejected, s.t, s.dt) for s in st])
    b = Batch(ctxs)
    b.ts_setup(params)
    X = _cat(states)
    print(f"[batch-loadcases] {len(wls)} members, 3 cases each", flush=True)
    per_case = b.solve_loadcases(X, times, rates)
    got = _host(X)
    for i, wl in enumerate(wls):
        mine = [(cs[i].steps, cs[i].rejected, cs[i].t, cs[i].dt) for cs in per_case]
        print(f"   {wl.name}: {mine}", flush=True)
        assert [m[:3] for m in mine] == [r[:3] for r in ref_st[i]], f"member {i} ({wl.name})"
        # the next step size of a skipped (zero-length) case is dt_min in both
        assert [m[3] for m in mine if m[0]] == [r[3] for r in ref_st[i] if r[0]]
        assert np.array_equal(got[b.offsets[i]:b.offsets[i + 1]], ref_X[i]), f"member {i} ({wl.name}): state differs"
    b.close()
