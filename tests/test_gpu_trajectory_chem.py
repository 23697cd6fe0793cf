"""Adaptive trajectories of the chemistry configurations against the CPU
restatement (tests/helpers.py cpu_solve: oracle RHS and PostStep, the same
controller): 3D AMR ternary precipitation (2 500 leaves, calphaddis matrix,
calphad2sl precipitates, stiff: step sizes near 1e-12) for 60 accepted
steps, and the 4-component configuration with every constitutive model for
a fixed end time.  Accepted and rejected counts equal, final time and next
step size within 1e-8, active sets equal, the final phase fractions within
1e-9 and the final compositions (the output step's "<component>_c" fields,
src/cipher.c:1040-1060) within 1e-8 relative.  The FP64 right-hand sides
differ from the oracle's by an ulp of exp/log, so the two trajectories are
compared to these bounds, not bit for bit.
"""
import numpy as np
import pytest

import oracle
from cipher_amd import abi
from tests import helpers as H

pytestmark = pytest.mark.gpu

CASES = {
    # name: (workload, dt_min, t_end, max_steps)
    "ternary": (lambda: H.ternary(), 1e-12, 1.0, 60),
    "chem4": (lambda: H.chem(), 1e-6, 7e-4, 0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_chemistry_trajectory_matches_cpu(name):
    import torch
    from cipher_amd.runtime import Context
    mk, dt_min, t_end, max_steps = CASES[name]
    wl = mk()
    p = abi.ts_params(1e-4, 1e-4, dt_min, 1e-3, use_graph=True)
    X0 = H.initial_state(wl)
    Xc, info = H.cpu_solve(wl, X0, 0.0, t_end, dt_min, p, max_steps=max_steps)
    print(f"[chem trajectory] {name} {wl.ncells} cells: cpu steps={info['steps']} rejected={info['rejected']} "
          f"t={info['t']:.6e}", flush=True)
    ctx = Context(wl.cfg)
    ctx.upload_mesh(wl.mesh)
    ctx.ts_setup(p)
    Xd = torch.as_tensor(X0).to(ctx.device)
    st, dt = ctx.ts_solve(Xd, 0.0, t_end, dt_min, max_steps=max_steps)
    comp_fields = [(abi.OUT_COMP, k) for k in range(wl.cfg.ncp)]
    cg = ctx.output(Xd, comp_fields).cpu().numpy()
    Xg = Xd.cpu().numpy()
    print(f"   gpu steps={st.steps} rejected={st.rejected} t={st.t:.6e}", flush=True)
    assert (st.steps, st.rejected) == (info["steps"], info["rejected"])
    assert st.t == pytest.approx(info["t"], rel=1e-8, abs=0.0)
    assert dt == pytest.approx(info["dt"], rel=1e-8, abs=0.0)
    L = wl.layout
    m = H.state_mask(Xc, L, wl.cfg)
    assert np.array_equal(H.state_mask(Xg, L, wl.cfg), m), "active sets differ"
    errs = H.field_errors(np.where(m, Xg, 0.0), np.where(m, Xc, 0.0), L)
    cc = oracle.output(wl.cfg, wl.mesh, Xc, comp_fields)
    cerr = float(np.abs(cg - cc).max() / np.abs(cc).max())
    print(f"   final state errors {errs}, composition {cerr:.3e}", flush=True)
    assert errs["PF"] <= 1e-9
    assert cerr <= 1e-8
    ctx.close()
