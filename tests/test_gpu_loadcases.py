"""CIPHER's load-case loop (src/cipher.c:1589-1606) through cipher_ts_loadcases
on the 4-component constitutive-model configuration (64 cells, calphaddis /
calphad2sl / quadratic phases, temperature-dependent thermodynamics):

* against its CPU restatement (tests/helpers.py cpu_loadcases: oracle RHS,
  the library's host controller, one controller memory for the run): per
  case accepted/rejected counts and the time reached equal, next step sizes
  within 1e-8, final fields within 1e-9;
* the temperature, which has no conduction here, integrates the piecewise
  constant rate exactly: T = T0 + sum_k rate_k * time_k in every cell;
* no rates: T stays T0 bit for bit; a zero-length case takes no step;
  a negative case length is an argument error.
"""
import numpy as np
import pytest

from cipher_amd import abi
from tests import helpers as H

pytestmark = pytest.mark.gpu

TIMES = [4e-4, 3e-4, 3e-4]
RATES = [1.0e4, -2.0e4, 0.0]


def _params():
    return abi.ts_params(1e-4, 1e-4, 1e-6, 1e-3, use_graph=True)


def _gpu(wl, times, rates):
    import torch
    from cipher_amd.runtime import Context
    ctx = Context(wl.cfg)
    ctx.upload_mesh(wl.mesh)
    ctx.ts_setup(_params())
    X = torch.as_tensor(H.initial_state(wl)).to(ctx.device)
    st = ctx.solve_loadcases(X, times, rates)
    torch.cuda.synchronize()
    out = X.cpu().numpy()
    ctx.close()
    return out, st


@pytest.mark.parametrize("times,rates", [(TIMES, RATES), ([4e-4, 0.0, 3e-4], [5.0e3, 1.0e6, -5.0e3]),
                                         (TIMES, None)])
def test_loadcases_match_cpu(times, rates):
    wl = H.chem()
    print(f"[loadcases] times={times} rates={rates}", flush=True)
    Xc, infos = H.cpu_loadcases(wl, H.initial_state(wl), times, rates, _params())
    Xg, st = _gpu(wl, times, rates)
    for k, (s, i) in enumerate(zip(st, infos)):
        print(f"   case {k}: gpu {s.steps}/{s.rejected} t={s.t} dt={s.dt:.6e}; "
              f"cpu {i['steps']}/{i['rejected']} t={i['t']} dt={i['dt']:.6e}", flush=True)
        assert (s.steps, s.rejected) == (i["steps"], i["rejected"])
        assert s.t == i["t"]
        assert s.dt == pytest.approx(i["dt"], rel=1e-8, abs=0.0)
    assert st[-1].t == sum(times)
    if 0.0 in times:
        assert st[times.index(0.0)].steps == 0
    L = wl.layout
    m = H.state_mask(Xc, L, wl.cfg)
    assert np.array_equal(H.state_mask(Xg, L, wl.cfg), m), "active sets differ"
    errs = H.field_errors(np.where(m, Xg, 0.0), np.where(m, Xc, 0.0), L)
    print("   final state errors", errs, flush=True)
    assert max(errs.values()) <= 1e-9
    T0 = H.initial_state(wl).reshape(-1, L["ndof"])[:, L["tm_offset"]]
    T = Xg.reshape(-1, L["ndof"])[:, L["tm_offset"]]
    if rates is None:
        assert np.array_equal(T, T0)
    else:
        want = T0 + sum(r * tau for r, tau in zip(rates, times))
        np.testing.assert_allclose(T, want, rtol=1e-12)


def test_loadcases_reject_negative_length():
    import torch
    from cipher_amd.runtime import Context
    wl = H.chem()
    ctx = Context(wl.cfg)
    ctx.upload_mesh(wl.mesh)
    ctx.ts_setup(_params())
    X = torch.as_tensor(H.initial_state(wl)).to(ctx.device)
    with pytest.raises(abi.CipherError):
        ctx.solve_loadcases(X, [1e-4, -1e-4], None)
    with pytest.raises(ValueError):
        ctx.solve_loadcases(X, [1e-4, 1e-4], [1.0])
    ctx.close()
