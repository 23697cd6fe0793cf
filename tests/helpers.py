"""Shared builders for the parity tests (small workloads, evolved states)."""
import numpy as np

import oracle
from cipher_amd import workload as W

_cache = {}


def gg2d(res=(32, 32), ngrains=8, periodic=True):
    key = ("gg2d", res, ngrains, periodic)
    if key not in _cache:
        _cache[key] = W.Workload(W.grain_growth_yaml(res=res, ngrains=ngrains, periodic=periodic), "gg2d")
    return _cache[key]


def gg3d_amr(res=(16, 16, 16), ngrains=5, minlevel=2):
    key = ("gg3d", res, ngrains, minlevel)
    if key not in _cache:
        txt = W.grain_growth_yaml(res=res, ngrains=ngrains, periodic=False, energy_range=(0.8, 1.2),
                                  ntypes=16, minlevel=minlevel)
        _cache[key] = W.Workload(txt, "gg3d")
    return _cache[key]


def gg2d_amr(res=(64, 64), ngrains=6, minlevel=3, amrinterval=None):
    """2D periodic grain growth on an adaptive quadtree (levels minlevel..6)."""
    key = ("gg2d_amr", res, ngrains, minlevel, amrinterval)
    if key not in _cache:
        txt = W.grain_growth_yaml(res=res, ngrains=ngrains, periodic=True, minlevel=minlevel)
        if amrinterval is not None:
            assert "amrinterval : 1000000\n" in txt
            txt = txt.replace("amrinterval : 1000000\n", f"amrinterval : {amrinterval}\n")
        _cache[key] = W.Workload(txt, "gg2d_amr")
    return _cache[key]


def ternary(res=(16, 16, 16), ngrains=4, nppt=2, radius=2, minlevel=2):
    key = ("tern", res, ngrains, nppt, radius, minlevel)
    if key not in _cache:
        txt = W.ternary_ppt_yaml(res=res, ngrains=ngrains, nppt=nppt, radius=radius, minlevel=minlevel)
        _cache[key] = W.Workload(txt, "ternary")
    return _cache[key]


def with_interpolation(wl, kind):
    """The same workload with the linear or quadratic interpolant
    (src/utility.c:470-518) in place of the cubic one."""
    key = ("interp", wl.name, kind)
    if key not in _cache:
        assert "interpolation: cubic\n" in wl.text
        _cache[key] = W.Workload(wl.text.replace("interpolation: cubic\n", f"interpolation: {kind}\n"),
                                 f"{wl.name}_{kind}")
    return _cache[key]


def initial_state(wl):
    X = oracle.setup_problem(wl.cfg, wl.mesh, wl.phase)
    if wl.layout["dp_size"]:
        W.uniform_mu_override(X, wl.layout, wl.phase, wl.cfg.phase_material)
    return X


def evolve(wl, X, steps=10, frac=0.02, seed=0):
    """Explicit Euler + active-set update with the oracle; dt picked so the
    largest phase-fraction change per step is `frac` and no entry changes by
    more than 5% of max(|x|, 1)."""
    L = wl.layout
    D = L["ndof"]
    P = L["pf_size"]
    for _ in range(steps):
        F = oracle.rhs(wl.cfg, wl.mesh, X)
        Fv = F.reshape(-1, D)
        m = np.abs(Fv[:, L["pf_offset"]:L["pf_offset"] + P]).max()
        dt = frac / m if m > 0 else 1.0
        # and no field changes by more than 5% of its size (stiff chemistry)
        r = (np.abs(Fv) / np.maximum(np.abs(X.reshape(-1, D)), 1.0)).max()
        if r > 0:
            dt = min(dt, 0.05 / r)
        X = X + dt * F
        X, rc = oracle.poststep(wl.cfg, wl.mesh, X)
        assert rc == 0
    return X


def perturb(wl, X, seed=1, amp=0.3):
    """Random phase fractions in multi-phase cells (then projected), small
    random mu/EX offsets: a generic state exercising every branch."""
    rng = np.random.default_rng(seed)
    L = wl.layout
    D = L["ndof"]
    Xv = X.reshape(-1, D).copy()
    n = np.rint(Xv[:, 0]).astype(int)
    for k in range(L["pf_size"]):
        sel = n > k
        Xv[sel, L["pf_offset"] + k] += amp * rng.random(sel.sum()) * (n[sel] > 1)
    if L["dp_size"]:
        mu = Xv[:, L["dp_offset"]:L["dp_offset"] + L["dp_size"]]
        mu += 0.02 * np.abs(mu).max() * rng.standard_normal(mu.shape)
        ex = Xv[:, L["ex_offset"]:L["ex_offset"] + L["ex_size"]]
        ex += 0.02 * (np.abs(ex).max() + 1.0) * rng.standard_normal(ex.shape)
    X2, rc = oracle.poststep(wl.cfg, wl.mesh, Xv.ravel())
    assert rc == 0
    return X2


def field_slices(L):
    return {"PF": slice(L["pf_offset"], L["pf_offset"] + L["pf_size"]),
            "DP": slice(L["dp_offset"], L["dp_offset"] + L["dp_size"]),
            "EX": slice(L["ex_offset"], L["ex_offset"] + L["ex_size"]),
            "TM": slice(L["tm_offset"], L["tm_offset"] + 1)}


def field_errors(F, Fref, L):
    """Per field: max|F - Fref| / max|Fref| (normwise relative)."""
    D = L["ndof"]
    A, B = F.reshape(-1, D), Fref.reshape(-1, D)
    out = {}
    for name, sl in field_slices(L).items():
        a, b = A[:, sl], B[:, sl]
        if a.size == 0:
            continue
        scale = np.abs(b).max()
        err = np.abs(a - b).max()
        out[name] = (err / scale) if scale > 0 else err
    return out


def state_mask(X, L, cfg):
    """Entries of a state that carry information (stale slots excluded)."""
    D = L["ndof"]
    Xv = X.reshape(-1, D)
    n = np.rint(Xv[:, 0]).astype(int)
    mask = np.zeros_like(Xv, dtype=bool)
    mask[:, 0] = True
    for k in range(L["pf_size"]):
        mask[:, 1 + k] = n > k
        mask[:, L["pf_offset"] + k] = n > k
    mask[:, L["dp_offset"]:L["dp_offset"] + L["dp_size"]] = True
    mask[:, L["tm_offset"]] = True
    if L["sp_size"]:
        ids = np.rint(Xv[:, 1:1 + L["pf_size"]]).astype(int)
        nd = L["dp_size"]
        for k in range(L["pf_size"]):
            ok = n > k
            mats = cfg.phase_material[np.where(ok, ids[:, k], 0)]
            nsites = np.array([cfg.materials[m].nsites for m in range(len(cfg.materials))])[mats]
            for s in range(L["maxsites"]):
                for c in range(nd):
                    mask[:, L["ex_offset"] + k * L["sp_size"] + s * nd + c] = ok & (nsites > s)
    return mask.ravel()


def chem_test_yaml(ncomp=4, sources=True, thermal=False):
    """A small 2D config with every constitutive model: calphaddis with RK
    binaries (orders 0-2) and a ternary term, calphad2sl with unary/binary/
    ternary tables on both sublattices, quadratic, none; interfaces with a
    segregation potential and solute mobility factors; a sink source."""
    comps = ["al", "cu", "mg", "zn", "si"][:ncomp]
    rng = np.random.default_rng(7)

    def ts(ind, coeffs, exps, tlnt=None):
        s = " " * ind + "t_coefficient : [" + ", ".join(repr(float(c)) for c in coeffs) + "]\n"
        s += " " * ind + "t_exponent : [" + ", ".join(str(e) for e in exps) + "]\n"
        if tlnt is not None:
            s += " " * ind + f"tlnt_coefficient : {tlnt!r}\n"
        return s

    def mob(ind):
        s = " " * ind + "mobilityc :\n"
        for c in comps:
            s += " " * (ind + 2) + f"{c} :\n" + " " * (ind + 4) + "mobility0 : 1.0e-5\n"
            s += " " * (ind + 4) + "unary_migration :\n"
            for cj in comps:
                s += " " * (ind + 6) + f"{cj} :\n" + ts(ind + 8, [-1.0e4 * rng.random(), -5.0], [0, 1])
            s += " " * (ind + 4) + "binary_migration :\n"
            s += " " * (ind + 6) + f"{comps[0]} :\n" + " " * (ind + 8) + f"{comps[1]} :\n"
            s += " " * (ind + 10) + "nrk : 2\n"
            s += " " * (ind + 10) + "rk_0 :\n" + ts(ind + 12, [3000.0], [0])
            s += " " * (ind + 10) + "rk_1 :\n" + ts(ind + 12, [-1500.0, 2.0], [0, 1])
        return s

    n = len(comps)
    txt = ("---\nheader:\n  grid : [8, 8]\n  size : [8.0, 8.0]\n  n_phases : 4\n"
           "  materials : [dis, sl, quad, none]\n  interfaces : [ia, ib]\n"
           f"  components : [{', '.join(comps)}]\n  outputs : [phaseid]\n"
           + ("  boundaries : [bl, br]\n" if thermal else "") + "\n"
           "solution_parameters:\n  time : [1.0]\n  initblocksize : [1, 1]\n  initrefine : 3\n"
           "  maxnrefine : 3\n  minnrefine : 3\n  initcoarsen : 0\n  amrinterval : 100000\n"
           "  interpolation: cubic\n  max_phase_neighbours : 4\n"
           + ("  temperature_rate : [2.5]\n  petscoptions : -dm_p4est_brick_periodicity 0,1\n\n" if thermal
              else "  petscoptions : -dm_p4est_brick_periodicity 1,1\n\n"))
    txt += "material:\n  dis:\n    chemicalenergy : calphaddis\n    molarvolume : 1.0e-5\n"
    txt += "    temperature0 : 650.0\n    chempot_ex_kineticcoeff : 2.0\n"
    c0 = rng.dirichlet(np.ones(n) * 3)
    txt += "    c0 : [" + ", ".join(repr(float(v)) for v in c0) + "]\n" + mob(4)
    txt += "    ref_enthalpy :\n" + ts(6, [100.0, -1.0], [0, 1])
    if thermal:
        txt += ("    tconductivity :\n" + ts(6, [2.0, 1.0e-3], [0, 1]) + "    specific_heat :\n"
                + ts(6, [3.0e6], [0]) + "    latent_heat :\n" + ts(6, [-1.0e8], [0]))
    txt += "    unary_enthalpy :\n"
    for c in comps:
        txt += f"      {c} :\n" + ts(8, [-8000.0 * rng.random(), 130.0, -2.0e-3, 5.0e4], [0, 1, 2, -1], tlnt=-24.0)
    txt += "    binary_enthalpy :\n"
    for i in range(n):
        txt += f"      {comps[i]} :\n"
        for j in range(i + 1, n):
            txt += f"        {comps[j]} :\n          nrk : 3\n"
            for o in range(3):
                txt += f"          rk_{o} :\n" + ts(12, [2.0e4 * (rng.random() - 0.5), 3.0], [0, 1])
    if n >= 3:
        txt += (f"    ternary_enthalpy :\n      {comps[0]} :\n        {comps[1]} :\n          {comps[2]} :\n"
                "            nrk : 2\n            rk_0 :\n" + ts(14, [5000.0], [0]) + "              index : 1\n"
                "            rk_1 :\n" + ts(14, [-3000.0], [0]) + "              index : 2\n")
    txt += "  sl:\n    chemicalenergy : calphad2sl\n    molarvolume : 1.0e-5\n"
    txt += "    temperature0 : 650.0\n    chempot_ex_kineticcoeff : 1.0\n"
    cp, cq = rng.dirichlet(np.ones(n) * 3), rng.dirichlet(np.ones(n) * 3)
    txt += "    c0_p : [" + ", ".join(repr(float(v)) for v in cp) + "]\n"
    txt += "    c0_q : [" + ", ".join(repr(float(v)) for v in cq) + "]\n"
    txt += "    stochiometry : [0.75, 0.25]\n" + mob(4)
    if thermal:
        txt += ("    tconductivity :\n" + ts(6, [3.0], [0]) + "    specific_heat :\n"
                + ts(6, [2.5e6, 100.0], [0, 1]) + "    latent_heat :\n" + ts(6, [2.0e8], [0]))
    txt += "    unary_enthalpy :\n"
    for a in comps:
        for b in comps:
            txt += f"      ({a}:{b}) :\n" + ts(8, [-2.0e4 * rng.random(), 1.0], [0, 1], tlnt=-1.5)
    txt += "    binary_enthalpy :\n"
    txt += f"      ({comps[0]},{comps[1]}:{comps[0]}) :\n        nrk : 2\n        rk_0 :\n" + ts(10, [4000.0], [0])
    txt += "        rk_1 :\n" + ts(10, [-700.0], [0])
    txt += f"      ({comps[1]}:{comps[0]},{comps[2 % n]}) :\n        nrk : 2\n        rk_0 :\n" + ts(10, [-2500.0], [0])
    txt += "        rk_1 :\n" + ts(10, [900.0], [0])
    if n >= 3:
        txt += f"    ternary_enthalpy :\n      ({comps[0]},{comps[1]},{comps[2]}:{comps[0]}) :\n        nrk : 1\n        rk_0 :\n" + ts(10, [1200.0], [0])
        txt += f"      ({comps[1]}:{comps[0]},{comps[1]},{comps[2]}) :\n        nrk : 1\n        rk_0 :\n" + ts(10, [-800.0], [0])
    txt += "  quad:\n    chemicalenergy : quadratic\n    molarvolume : 1.0e-5\n    temperature0 : 650.0\n"
    txt += "    c0 : [" + ", ".join(repr(float(v)) for v in rng.dirichlet(np.ones(n) * 3)) + "]\n" + mob(4)
    txt += "    ref_enthalpy :\n" + ts(6, [10.0], [0])
    for key, scale in (("ceq", None), ("unary_enthalpy", 1000.0), ("binary_enthalpy", 5.0e4)):
        txt += f"    {key} :\n"
        vals = rng.dirichlet(np.ones(n) * 4) if scale is None else scale * (1 + rng.random(n))
        for c, v in zip(comps, vals):
            # binary coefficients constant in T: the temperature tangent
            # (src/material.c:924-970) holds the inverted matrix fixed
            tdep = key != "binary_enthalpy"
            txt += f"      {c} :\n" + (ts(8, [0.9 * float(v), 0.1 * float(v) / 650.0], [0, 1]) if tdep
                                      else ts(8, [float(v)], [0]))
    if sources:
        txt += "    sources :\n      activesources : [sk]\n      sk :\n        source : sink\n"
        txt += "        rate : [" + ", ".join("1.0e3" for _ in comps) + "]\n"
        txt += "        ceq : [" + ", ".join(repr(float(v)) for v in rng.dirichlet(np.ones(n))) + "]\n"
    txt += "  none:\n    chemicalenergy : none\n    molarvolume : 1.0e-5\n    temperature0 : 650.0\n\n"
    txt += ("interface:\n  ia:\n    width : 3.0\n    energy :\n      e0: 0.5\n    mobility :\n      m0: 1.0e-3\n"
            "    potential : [" + ", ".join("-500.0" for _ in comps[:-1]) + "]\n"
            "    mobilityc : [" + ", ".join("3000.0" for _ in comps) + "]\n"
            "  ib:\n    width : 3.0\n    energy :\n      e0: 0.7\n      activation_energy :\n"
            + ts(8, [-2000.0], [0]) + "    mobility :\n      m0: 2.0e-3\n      activation_energy :\n"
            + ts(8, [-1000.0], [0]) + "\n")
    if thermal:
        # x low: chemical Neumann on two components, thermal Dirichlet;
        # x high: chemical Dirichlet on one component, thermal Neumann
        stars = lambda vals: "[" + ", ".join(vals) + "]"
        v1 = ["1.0e3"] * (n - 1)
        v1[1 % (n - 1)] = "'*'"
        v2 = ["'*'"] * (n - 1)
        v2[0] = "-1.0e7"
        txt += ("boundary:\n  bl:\n    boundary_id : 1\n    chem :\n      type : neumann\n"
                f"      value : {stars(v1)}\n    thermal :\n      type : dirichlet\n      value : [640.0]\n"
                "  br:\n    boundary_id : 2\n    chem :\n      type : dirichlet\n"
                f"      value : {stars(v2)}\n    thermal :\n      type : neumann\n      value : [-3.0]\n\n")
    labels = np.arange(64).reshape(8, 8) // 16          # 4 stripes of phases 0..3
    # phase 3 reuses 'dis': a cell whose phases all have no chemistry has a zero
    # composition, and the harmonic face mean of two zero compositions is 0/0
    # (src/cipher.c:280); 'none' stays in the material table for the
    # per-function checks
    txt += "mappings:\n  phase_material_mapping : |-\n    1\n    2\n    3\n    1\n"
    from cipher_amd import config as cfgmod
    txt += "  voxel_phase_mapping: |-\n" + "\n".join("    " + l for l in cfgmod.encode_rle(labels.T.ravel() + 1).split("\n")) + "\n"
    txt += "  interface_mapping: |-\n    1\n    1\n    2\n    2\n    1\n    1\n    2\n    2\n    1\n    1\n    1\n    2\n    1\n    1\n    1\n    1\n"
    return txt


def chem(thermal=False, sources=True):
    key = ("chem", thermal, sources)
    if key not in _cache:
        _cache[key] = W.Workload(chem_test_yaml(sources=sources, thermal=thermal), "chem")
    return _cache[key]


# ---- CPU restatement of the adaptive RK driver (for trajectory parity) ----
BS3_A2, BS3_A3 = 1.0 / 2.0, 3.0 / 4.0
BS3_B = (2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0)
BS3_BH = (7.0 / 24.0, 1.0 / 4.0, 1.0 / 3.0, 1.0 / 8.0)


def cpu_attempt(wl, X, dt, atol, rtol, rhs=None, method=0):
    """One Bogacki-Shampine 3(2) attempt (method 0, abi.RK_3BS) or Heun 2(1)
    attempt (method 1, abi.RK_2A: Xh = X + dt K1 the explicit Euler step,
    Xn = X + dt (K1 + K2)/2 with K2 = f(Xh)) with the oracle RHS, in the
    stage arithmetic of the device integrator: Y = X + ((a0 K0 + a1 K1) +
    ...), a_j = dt * A_j.  Returns (new solution, embedded solution, error
    norm)."""
    f = rhs or (lambda Y: oracle.rhs(wl.cfg, wl.mesh, Y))
    K1 = f(X)
    if method == 1:
        Xh = X + (dt * 1.0) * K1
        K2 = f(Xh)
        Xn = X + ((dt * 0.5) * K1 + (dt * 0.5) * K2)
        return Xn, Xh, oracle.wrms_norm(Xn, Xh, atol, rtol)
    K2 = f(X + (dt * BS3_A2) * K1)
    K3 = f(X + (dt * BS3_A3) * K2)
    a = [dt * b for b in BS3_B]
    Xn = X + ((a[0] * K1 + a[1] * K2) + a[2] * K3)
    K4 = f(Xn)
    ah = [dt * b for b in BS3_BH]
    Xh = X + (((ah[0] * K1 + ah[1] * K2) + ah[2] * K3) + ah[3] * K4)
    return Xn, Xh, oracle.wrms_norm(Xn, Xh, atol, rtol)


def cpu_solve(wl, X, t0, t_end, dt, params, max_steps=0, ctrl=None):
    """The device integrator's loop (exact final time, PostStep after each
    accepted step) on the CPU, with the library's host step controller;
    ctrl carries the controller's memory in and out (fresh if None)."""
    from cipher_amd import abi
    from cipher_amd.runtime import ts_adapt
    ctrl = abi.TsCtrl() if ctrl is None else ctrl
    t, steps, rejected, dts = t0, 0, 0, []
    X = np.array(X, dtype=np.float64)
    while t < t_end and (max_steps <= 0 or steps < max_steps):
        dt_try = t_end - t if t + dt > t_end else dt
        if not dt_try > 0:
            break
        Xn, _, err = cpu_attempt(wl, X, dt_try, params.atol, params.rtol, method=params.method)
        acc, dt_next = ts_adapt(params, ctrl, dt_try, err)
        dts.append((dt_try, err, acc))
        if acc:
            X, rc = oracle.poststep(wl.cfg, wl.mesh, Xn)
            assert rc == 0
            t = t_end if dt_try == t_end - t else t + dt_try
            steps += 1
        else:
            rejected += 1
        dt = dt_next
    return X, dict(steps=steps, rejected=rejected, t=t, dt=dt, history=dts)


def cpu_loadcases(wl, X, times, rates, params, t0=0.0):
    """CIPHER's load-case loop (src/cipher.c:1589-1606) over cpu_solve: case
    k sets the temperature rate rates[k] (none if rates is None), restarts
    from dt = params.dt_min and runs to the running sum t0 + times[0] + ...
    + times[k]; one controller memory for the whole run.  Restores the
    configuration's rate afterwards."""
    from cipher_amd import abi
    s = wl.cfg.struct
    saved = (s.has_temperature_rate, s.temperature_rate)
    ctrl, out, start, t = abi.TsCtrl(), [], t0, t0
    try:
        for k, tau in enumerate(times):
            s.has_temperature_rate = 0 if rates is None else 1
            s.temperature_rate = 0.0 if rates is None else rates[k]
            end = start + tau
            info = dict(steps=0, rejected=0, t=t, dt=params.dt_min)
            if t < end:
                X, info = cpu_solve(wl, X, t, end, params.dt_min, params, ctrl=ctrl)
            t = info["t"]
            out.append(info)
            start += tau
    finally:
        s.has_temperature_rate, s.temperature_rate = saved
    return X, out


def cpu_timeloop(cfg, forest, X, params, times, rates, interval, t0=0.0, ofreq=0):
    """CIPHER's time loop with remeshing (src/cipher.c:1589-1830) over
    cpu_solve, in the arithmetic of cipher_amd.timeloop.solve: chunks end at
    multiples of interval accepted steps or at the case's end; after a chunk
    ending on a multiple the host remesh (timeloop.adapt_state) runs and, if
    the forest changed, the controller memory starts afresh (a new TS).
    With ofreq, chunks also end at multiples of ofreq steps, where the
    state is recorded.  Returns (forest, X, chunks, remeshes) and, with
    ofreq, the list of (step, forest, X) at the output steps."""
    import types
    from cipher_amd import abi
    from cipher_amd import timeloop as TL
    s = cfg.struct
    saved = (s.has_temperature_rate, s.temperature_rate)
    wl = types.SimpleNamespace(cfg=cfg, mesh=forest.flatten())
    ctrl, chunks, remeshes, outputs = abi.TsCtrl(), [], 0, []
    current = case_t = t0
    nsteps = 0
    X = np.array(X, dtype=np.float64)
    try:
        for k, tau in enumerate(times):
            s.has_temperature_rate = 0 if rates is None else 1
            s.temperature_rate = 0.0 if rates is None else rates[k]
            dt = params.dt_min
            end = case_t + tau
            while current < end:
                chunk = interval - nsteps % interval
                if ofreq:
                    chunk = min(chunk, ofreq - nsteps % ofreq)
                X, info = cpu_solve(wl, X, current, end, dt, params, max_steps=chunk, ctrl=ctrl)
                assert info["steps"] > 0
                current, dt = info["t"], info["dt"]
                nsteps += info["steps"]
                chunks.append((k, info["steps"], info["rejected"], current, dt, wl.mesh.ncells))
                if ofreq and nsteps % ofreq == 0:
                    outputs.append((nsteps, forest, X.copy()))
                if nsteps % interval == 0:
                    f2, X2 = TL.adapt_state(forest, X, cfg.layout(), cfg.minnrefine, cfg.maxnrefine)
                    if f2 is not forest:
                        forest, X = f2, np.array(X2)
                        wl.mesh = forest.flatten()
                        ctrl = abi.TsCtrl()
                        remeshes += 1
            case_t += tau
    finally:
        s.has_temperature_rate, s.temperature_rate = saved
    if ofreq:
        return forest, X, chunks, remeshes, outputs
    return forest, X, chunks, remeshes


GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)), "golden")
REF_CASES = ["anisotropy", "bc_chem", "bc_thermal", "gb_precipitate", "grain_growth", "nucleation_7050",
             "nucleation_crti", "segregation", "sl2_t250", "solidification"]


def ref_case(name):
    """A fixture derived from one of the reference's example configurations
    (tests/golden/make_reference_cases.py)."""
    import os
    key = ("ref", name)
    if key not in _cache:
        with open(os.path.join(GOLDEN, f"ref_{name}.yaml")) as f:
            _cache[key] = W.Workload(f.read(), name)
    return _cache[key]


def crowded(npf=24, cap=16, nmin=6, nmax=16, res=(16, 16), seed=3):
    """A synthetic state with large active sets (n up to the capacity) on a
    2D periodic grid: every cell holds a random sorted subset of the phases
    with random fractions on the simplex; it exercises the widest lane groups
    and long neighbour-list searches."""
    key = ("crowded", npf, cap, nmin, nmax, res, seed)
    if key not in _cache:
        txt = W.grain_growth_yaml(res=res, ngrains=npf, periodic=True, capacity=cap, energy_range=(0.5, 1.5),
                                  ntypes=12)
        wl = W.Workload(txt, f"crowded{cap}")
        rng = np.random.default_rng(seed)
        L = wl.layout
        X = np.zeros((wl.ncells, L["ndof"]))
        for c in range(wl.ncells):
            n = int(rng.integers(nmin, min(nmax, cap) + 1))
            ids = np.sort(rng.choice(npf, size=n, replace=False))
            X[c, 0] = n
            X[c, 1:1 + n] = ids
            X[c, L["pf_offset"]:L["pf_offset"] + n] = rng.dirichlet(np.ones(n))
        X[:, L["tm_offset"]] = 1.0
        _cache[key] = (wl, X.ravel())
    return _cache[key]


def aniso_pair_yaml(res=(16, 16), e_aniso=(0.3, -0.2), m_aniso=(0.5, 0.25), gradient=True, scale_iso=False,
                    minlevel=None):
    """Two phases, one interface type; with gradient, anisotropic energy and
    mobility along the axes (values e_aniso / m_aniso for directions x, y);
    with scale_iso, the isotropic interface whose constants are scaled by
    (1 + value along x) instead."""
    txt = W.grain_growth_yaml(res=res, ngrains=2, periodic=False, minlevel=minlevel)
    e0, m0 = 1.0, 1.0
    if scale_iso:
        e0, m0 = e0 * (1.0 + e_aniso[0]), m0 * (1.0 + m_aniso[0])
    dirs = "[1.0, 0.0, 0.0, 1.0]" if len(res) == 2 else "[1.0, 0.0, 0.0, 0.0, 1.0, 0.0]"
    old = "    energy :\n      e0: 1.0\n    mobility :\n      m0: 1.0\n"
    assert old in txt
    new = f"    energy :\n      e0: {e0!r}\n"
    if gradient and not scale_iso:
        new += f"      anisotropy_values: [{e_aniso[0]!r}, {e_aniso[1]!r}]\n      anisotropy_directions: {dirs}\n"
    new += f"    mobility :\n      m0: {m0!r}\n"
    if gradient and not scale_iso:
        new += f"      anisotropy_values: [{m_aniso[0]!r}, {m_aniso[1]!r}]\n      anisotropy_directions: {dirs}\n"
    txt = txt.replace(old, new)
    if gradient:
        txt = txt.replace("interpolation: cubic\n", "interpolation: cubic\n  gradient_calculation: 1\n")
    return txt


# small members of a batched ensemble: every component count, all three
# phase types, lane capacities 4, 8 and 16, uniform and AMR meshes
ENSEMBLE_SPECS = [
    {"side": 64, "grains": 12, "components": 1, "types": 1, "capacity": 8, "minlevel": 3, "seed": 11},
    {"side": 32, "grains": 10, "components": 2, "types": 2, "capacity": 8, "minlevel": 3, "seed": 12},
    {"side": 32, "grains": 12, "components": 3, "types": 3, "capacity": 8, "minlevel": 3, "seed": 13},
    {"side": 32, "grains": 10, "components": 4, "types": 3, "capacity": 8, "minlevel": 5, "seed": 14},
    {"side": 64, "grains": 6, "components": 3, "types": 2, "capacity": 4, "minlevel": 4, "seed": 15},
    {"side": 64, "grains": 24, "components": 1, "types": 1, "capacity": 16, "minlevel": 6, "seed": 16},
    {"side": 16, "grains": 4, "components": 2, "types": 1, "capacity": 8, "minlevel": 4, "seed": 17},
]


def ensemble(idx=None):
    """Workloads of ENSEMBLE_SPECS (cached), or the one at idx."""
    specs = ENSEMBLE_SPECS if idx is None else [ENSEMBLE_SPECS[idx]]
    out = []
    for sp in specs:
        key = ("ens",) + tuple(sorted(sp.items()))
        if key not in _cache:
            _cache[key] = W.ensemble_member(sp)
        out.append(_cache[key])
    return out if idx is None else out[0]


def ensemble_states(steps=3):
    """Evolved states of the ensemble members (oracle), cached."""
    key = ("ens_states", steps)
    if key not in _cache:
        _cache[key] = [evolve(wl, initial_state(wl), steps=steps) for wl in ensemble()]
    return _cache[key]


def ulp_spread(wl, X, F, trials=3):
    """Largest per-field change of the oracle's residual under random one-ulp
    perturbations of the state (active-set entries kept): the FP64 bar of a
    state on which the oracle itself is that ill-conditioned."""
    L = wl.layout
    D = L["ndof"]
    rng = np.random.default_rng(0)
    worst = {}
    for _ in range(trials):
        Xp = X * (1.0 + 2.0 ** -52 * rng.choice([-1.0, 1.0], X.size))
        Xp.reshape(-1, D)[:, :L["as_size"]] = X.reshape(-1, D)[:, :L["as_size"]]
        for f, e in field_errors(oracle.rhs(wl.cfg, wl.mesh, Xp), F, L).items():
            worst[f] = max(worst.get(f, 0.0), e)
    return worst
