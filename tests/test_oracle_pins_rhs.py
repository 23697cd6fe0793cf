"""Pins of the oracle's right-hand-side sections that the per-function pins
do not reach: boundary faces, the thermal face flux, the temperature rate
with specific and latent heat, and the chemical driving-force assembly.

Each check states the quantity independently of the oracle's code: a
conservation law or balance of the semi-discrete system evaluated with the
already-pinned per-phase functions (site fractions, chemical energy,
mobility, interpolants), on states built so that the expected value has a
closed form.  A planted sign, factor or index error in the corresponding
oracle section makes its check fail (the mutations are listed in
oracle/mutate.py, which runs them on the CPU).

Reference sections: boundary faces src/cipher.c:320-389; thermal face flux
:310-316; temperature rate :558-568; chemical driving force :499-522 with
Chemenergy src/material.c:480-490.
"""
import math

import numpy as np
import pytest

import oracle
from cipher_amd import abi
from cipher_amd import workload as W
from tests import helpers as H

RG = abi.R_GAS_CONST
PI = math.pi


def _yaml(thermal, size=None, minlevel=None, periodic=None, mobility0=None):
    txt = H.chem_test_yaml(sources=False, thermal=thermal)
    if mobility0 is not None:
        txt = txt.replace("mobility0 : 1.0e-5\n", f"mobility0 : {mobility0!r}\n")
    if size is not None:
        txt = txt.replace("  size : [8.0, 8.0]\n", f"  size : [{size!r}, {size!r}]\n")
    if minlevel is not None:
        txt = txt.replace("  minnrefine : 3\n", f"  minnrefine : {minlevel}\n")
    if periodic is not None:
        txt = txt.replace("-dm_p4est_brick_periodicity 0,1", f"-dm_p4est_brick_periodicity {periodic}")
    return txt


@pytest.fixture(scope="module")
def thermal_fine():
    # h = 0.25: a factor h, 1/h or 1/h^2 in the wrong place changes the
    # result by 4x or more
    # component mobilities raised so that a Dirichlet flux M (mu_bc - mu)
    # is of the size of the Neumann fluxes
    wl = W.Workload(_yaml(True, size=2.0, mobility0=1.0e3), "thermal_fine")
    assert np.allclose(wl.mesh.cellgeom, 0.25)
    return wl


def _amr_yaml():
    """The thermal configuration on a 16x16 periodic grid of two phase
    regions (levels 2-4, refined along the region borders)."""
    import re
    from cipher_amd import config as cfgmod
    txt = _yaml(True, size=4.0, periodic="1,1")
    txt = txt.replace("  grid : [8, 8]\n", "  grid : [16, 16]\n")
    txt = txt.replace("  initrefine : 3\n  maxnrefine : 3\n  minnrefine : 3\n",
                      "  initrefine : 4\n  maxnrefine : 4\n  minnrefine : 2\n")
    x, y = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    labels = np.where((x - 7.5) ** 2 + (y - 7.5) ** 2 < 25.0, 1, 0)
    rle = "\n".join("    " + l for l in cfgmod.encode_rle(labels.T.ravel() + 1).split("\n"))
    txt, n = re.subn(r"  voxel_phase_mapping: \|-\n(?:    .*\n)+", "  voxel_phase_mapping: |-\n" + rle + "\n", txt)
    assert n == 1
    return txt


@pytest.fixture(scope="module")
def thermal_amr():
    wl = W.Workload(_amr_yaml(), "thermal_amr")
    assert np.unique(wl.mesh.cellgeom).size > 1 and wl.mesh.bface_cell.size == 0
    return wl


def _single_phase_state(wl, phase, T, mu_scale=1.0, seed=0, ex_noise=0.05):
    """Every cell holds one phase; mu uniform; EX perturbed; T per cell."""
    cfg, L = wl.cfg, wl.layout
    rng = np.random.default_rng(seed)
    X0 = oracle.setup_problem(cfg, wl.mesh, np.full(wl.ncells, phase, np.int32)).reshape(wl.ncells, -1)
    X = X0.copy()
    nd = cfg.ncp - 1
    dpo, exo = L["dp_offset"], L["ex_offset"]
    X[:, dpo:dpo + nd] = X0[0, dpo:dpo + nd] * mu_scale
    ex = X[:, exo:exo + L["sp_size"]]
    ex += ex_noise * (np.abs(ex).max() + 1.0) * rng.standard_normal(ex.shape)
    X[:, L["tm_offset"]] = T
    return X


def _phase_y(cfg, L, x, g, mui=None):
    """Site fractions of active slot g of a cell record (Sitefrac of the
    implicit site potentials nu_s (mu - mu_int) - EX, src/cipher.c:180-189)."""
    nd = cfg.ncp - 1
    ids = np.rint(x[1:1 + int(round(x[0]))]).astype(int)
    mat = cfg.phase_material[ids[g]]
    m = cfg.materials[mat]
    mu = x[L["dp_offset"]:L["dp_offset"] + nd]
    mui = np.zeros(nd) if mui is None else mui
    ex = x[L["ex_offset"] + g * L["sp_size"]:L["ex_offset"] + (g + 1) * L["sp_size"]]
    spim = np.zeros(2 * nd)
    for s in range(m.nsites):
        spim[s * nd:(s + 1) * nd] = m.stoichiometry[s] * (mu - mui) - ex[s * nd:(s + 1) * nd]
    return mat, oracle.sitefrac(cfg, mat, spim, x[L["tm_offset"]])


def _ts(cfg, idx, T):
    t = cfg.pools.ts[idx]
    return oracle.sum_tseries(T, t) if t.n else None


def _without_rate(cfg):
    class _Ctx:
        def __enter__(self):
            self.saved = (cfg.struct.has_temperature_rate, cfg.struct.temperature_rate)
            cfg.struct.has_temperature_rate = 0
            return self

        def __exit__(self, *a):
            cfg.struct.has_temperature_rate, cfg.struct.temperature_rate = self.saved
    return _Ctx()


# --------------------------------------------------------------------------
# boundary faces, src/cipher.c:320-389
# --------------------------------------------------------------------------
def test_boundary_chemical_fluxes_closed_form(thermal_fine):
    """A uniform chemical potential has no interior flux, so every cell's
    composition changes only through its boundary faces (the potential rate
    balances the phase, explicit-potential and temperature terms cell by
    cell, pinned by the conservation tests): dc/dt = sum over the cell's
    boundary faces of a Neumann flux value*h or a Dirichlet flux
    M_c (mu_bc - mu_c), each over h^2.  M_c is the component mobility
    (CompositionMobilityComponent) of the cell's single phase."""
    wl = thermal_fine
    cfg, L = wl.cfg, wl.layout
    nd = cfg.ncp - 1
    X = _single_phase_state(wl, 0, 650.0, ex_noise=0.0)
    with _without_rate(cfg):
        F = oracle.rhs(cfg, wl.mesh, X.ravel()).reshape(X.shape)
    assert np.isfinite(F).all()
    expect = np.zeros((wl.ncells, nd))
    nbc = np.zeros((wl.ncells, nd), dtype=int)
    h = wl.mesh.cellgeom
    bcs = {int(cfg.bcs[b].boundary_id): cfg.bcs[b] for b in range(cfg.struct.nbcs)}
    seen = set()
    for cell, bid in zip(wl.mesh.bface_cell, wl.mesh.bface_id):
        bc = bcs.get(int(bid))
        if bc is None or bc.chem_type == abi.BC_NONE:
            continue
        seen.add(int(bc.chem_type))
        x = X[cell]
        mat, y = _phase_y(cfg, L, x, 0)
        mob = oracle.mobility_component(cfg, mat, y, x[L["tm_offset"]])
        for c in range(nd):
            if not (bc.chem_mask >> c) & 1:
                continue
            if bc.chem_type == abi.BC_NEUMANN:
                flux = bc.chem_val[c] * h[cell]
            else:
                flux = mob[c] * (bc.chem_val[c] - x[L["dp_offset"] + c])
            expect[cell, c] += flux / h[cell] ** 2
            nbc[cell, c] += 1
    assert seen == {abi.BC_NEUMANN, abi.BC_DIRICHLET}
    dc = np.array([H_dcdt(wl, X[c], F[c]) for c in range(wl.ncells)])
    # each entry against its own value (the Neumann and Dirichlet fluxes
    # differ in size); entries without a constrained face must stay 0
    on = nbc > 0
    assert on.sum() >= 3 * 8 and (~on).sum() > 0
    assert np.all(np.abs(dc[on] - expect[on]) <= 1e-6 * np.abs(expect[on]))
    cell_scale = np.abs(expect).max(axis=1)
    cell_scale[cell_scale == 0] = np.abs(expect[on]).min()
    assert np.all(np.abs(dc) * ~on <= 1e-6 * cell_scale[:, None])


def test_boundary_thermal_fluxes_closed_form(thermal_fine):
    """Uniform temperature and single-phase cells: the temperature rate is
    zero inside and, at a boundary face, k (T_bc - T)/h^2 (Dirichlet) or
    k q h/h^2 (Neumann) over the specific heat, k and c_p the series of the
    cell's material at its temperature."""
    wl = thermal_fine
    cfg, L = wl.cfg, wl.layout
    T0 = 655.0
    X = _single_phase_state(wl, 0, T0)
    with _without_rate(cfg):
        F = oracle.rhs(cfg, wl.mesh, X.ravel()).reshape(X.shape)
    m = cfg.materials[cfg.phase_material[0]]
    k = _ts(cfg, m.tconductivity, T0)
    cp = _ts(cfg, m.specific_heat, T0)
    assert k and cp
    h = wl.mesh.cellgeom
    bcs = {int(cfg.bcs[b].boundary_id): cfg.bcs[b] for b in range(cfg.struct.nbcs)}
    expect = np.zeros(wl.ncells)
    kinds = set()
    for cell, bid in zip(wl.mesh.bface_cell, wl.mesh.bface_id):
        bc = bcs.get(int(bid))
        if bc is None or bc.thermal_type == abi.BC_NONE:
            continue
        kinds.add(int(bc.thermal_type))
        if bc.thermal_type == abi.BC_NEUMANN:
            expect[cell] += k * bc.thermal_val * h[cell] / h[cell] ** 2
        else:
            expect[cell] += k * (bc.thermal_val - T0) / h[cell] ** 2
    assert kinds == {abi.BC_NEUMANN, abi.BC_DIRICHLET}
    expect /= cp
    got = F[:, L["tm_offset"]]
    assert np.abs(expect).max() > 0
    assert np.allclose(got, expect, rtol=1e-12, atol=1e-14 * np.abs(expect).max())


# --------------------------------------------------------------------------
# thermal face flux, src/cipher.c:310-316
# --------------------------------------------------------------------------
def test_thermal_flux_hot_cell_closed_form(thermal_fine):
    """One interior cell hotter by dT: it loses, and each face neighbour
    gains, the heat 2 k1 k2/(k1 + k2) dT/h^2 per face (harmonic mean of the
    two cells' conductivities at their own temperatures), over the
    receiving cell's specific heat; cells further away do not change."""
    wl = thermal_fine
    cfg, L = wl.cfg, wl.layout
    T0, dT = 650.0, 7.0
    xy = wl.forest.anchor
    hot = int(np.flatnonzero((xy[:, 0] == 4) & (xy[:, 1] == 3))[0])
    T = np.full(wl.ncells, T0)
    T[hot] += dT
    X = _single_phase_state(wl, 0, T)
    with _without_rate(cfg):
        F = oracle.rhs(cfg, wl.mesh, X.ravel()).reshape(X.shape)
    m = cfg.materials[cfg.phase_material[0]]
    k0, k1 = _ts(cfg, m.tconductivity, T0), _ts(cfg, m.tconductivity, T0 + dT)
    cp0, cp1 = _ts(cfg, m.specific_heat, T0), _ts(cfg, m.specific_heat, T0 + dT)
    kbar = 2.0 * k0 * k1 / (k0 + k1)
    h = 0.25
    nbrs = {int(b) for a, b in wl.mesh.faces if a == hot} | {int(a) for a, b in wl.mesh.faces if b == hot}
    assert len(nbrs) == 4
    interior = (xy[:, 0] > 0) & (xy[:, 0] < 7)
    got = F[:, L["tm_offset"]]
    for c in np.flatnonzero(interior):
        if c == hot:
            want = -4 * kbar * dT / h ** 2 / cp1
        elif c in nbrs:
            want = kbar * dT / h ** 2 / cp0
        else:
            want = 0.0
        assert got[c] == pytest.approx(want, rel=1e-12, abs=1e-12 * kbar * dT / h ** 2 / cp0), c


def test_thermal_flux_steady_profile_has_no_divergence(thermal_fine):
    """T(x) built so that the discrete flux 2 k_i k_j/(k_i + k_j)(T_j - T_i)/h
    is the same through every x face (k depends on T): every interior cell's
    temperature rate vanishes to rounding, although each face carries a
    large flux.  An arithmetic mean, or a flux scaled by the wrong cell's
    volume, leaves a residual."""
    wl = thermal_fine
    cfg, L = wl.cfg, wl.layout
    m = cfg.materials[cfg.phase_material[0]]
    kf = lambda T: _ts(cfg, m.tconductivity, T)
    h, J = 0.25, 400.0
    Tx = [600.0]
    for _ in range(7):
        a = Tx[-1]
        b = a + J * h / kf(a)
        for _ in range(60):           # Newton on 2 k(a) k(b)/(k(a)+k(b)) (b - a) = J h
            ka, kb = kf(a), kf(b)
            r = 2 * ka * kb / (ka + kb) * (b - a) - J * h
            e = 1e-4
            kbe = kf(b + e)
            dr = (2 * ka * kbe / (ka + kbe) * (b + e - a) - J * h - r) / e
            b -= r / dr
        Tx.append(b)
    Tx = np.array(Tx)
    X = _single_phase_state(wl, 0, Tx[wl.forest.anchor[:, 0]])
    with _without_rate(cfg):
        F = oracle.rhs(cfg, wl.mesh, X.ravel()).reshape(X.shape)
    cp = np.array([_ts(cfg, m.specific_heat, t) for t in X[:, L["tm_offset"]]])
    face = J / h / cp.min()               # one face's contribution to the rate
    inner = (wl.forest.anchor[:, 0] > 0) & (wl.forest.anchor[:, 0] < 7)
    assert np.abs(F[inner, L["tm_offset"]]).max() <= 1e-10 * face


def test_thermal_flux_conserves_heat_on_amr(thermal_amr):
    """Periodic adaptive mesh with hanging faces, random temperatures, no
    latent heat (single-phase cells): the face fluxes telescope, so
    sum(vol * c_p * dT/dt) = 0 over the mesh."""
    wl = thermal_amr
    cfg, L = wl.cfg, wl.layout
    rng = np.random.default_rng(5)
    T = 650.0 + rng.uniform(-30, 30, wl.ncells)
    X = _single_phase_state(wl, 0, T)
    with _without_rate(cfg):
        F = oracle.rhs(cfg, wl.mesh, X.ravel()).reshape(X.shape)
    m = cfg.materials[cfg.phase_material[0]]
    cp = np.array([_ts(cfg, m.specific_heat, t) for t in T])
    vol = wl.mesh.cellgeom ** cfg.dim
    heat = vol * cp * F[:, L["tm_offset"]]
    assert np.abs(heat).max() > 0
    assert abs(heat.sum()) <= 1e-12 * np.abs(heat).sum()


# --------------------------------------------------------------------------
# temperature rate with specific and latent heat, src/cipher.c:558-568
# --------------------------------------------------------------------------
def test_temperature_rate_latent_heat_balance():
    """No conduction (uniform temperature, no boundary conditions):
    each cell's energy balance c_p (dT/dt - rate) = sum_g L_g dphi_g/dt,
    c_p = sum_g h_g c_p,g over the phases whose material has a specific
    heat, the rate added after the division."""
    wl = W.Workload(_yaml(True, periodic="1,1"), "latent")
    cfg, L = wl.cfg, wl.layout
    assert wl.mesh.bface_cell.size == 0
    X = H.perturb(wl, H.initial_state(wl), seed=4).reshape(wl.ncells, -1)
    checked = 0
    saved = (cfg.struct.has_temperature_rate, cfg.struct.temperature_rate)
    try:
        for T0, rate in ((620.0, 0.0), (700.0, 2.5)):
            X[:, L["tm_offset"]] = T0
            cfg.struct.has_temperature_rate = 1 if rate else 0
            cfg.struct.temperature_rate = rate
            F = oracle.rhs(cfg, wl.mesh, X.ravel()).reshape(X.shape)
            for cell in range(wl.ncells):
                x, f = X[cell], F[cell]
                n = int(round(x[0]))
                ids = np.rint(x[1:1 + n]).astype(int)
                hg = oracle.interpolant(cfg.interpolation, x[L["pf_offset"]:L["pf_offset"] + n])
                cp, lat = 0.0, 0.0
                for g in range(n):
                    mm = cfg.materials[cfg.phase_material[ids[g]]]
                    c_g = _ts(cfg, mm.specific_heat, T0)
                    l_g = _ts(cfg, mm.latent_heat, T0)
                    if c_g is not None:
                        cp += hg[g] * c_g
                    if l_g is not None:
                        lat += f[L["pf_offset"] + g] * l_g
                want = (lat / cp if cp else lat) + rate
                if lat:
                    checked += 1
                assert f[L["tm_offset"]] == pytest.approx(want, rel=1e-12, abs=1e-12 * (abs(rate) + 1e-300)), cell
    finally:
        cfg.struct.has_temperature_rate, cfg.struct.temperature_rate = saved
    assert checked > wl.ncells // 4


# --------------------------------------------------------------------------
# chemical driving force, src/cipher.c:499-522 (Chemenergy src/material.c:480-490)
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def chem_periodic():
    return W.Workload(_yaml(False), "chem_periodic")


def _activation(cfg, idx, T):
    v = _ts(cfg, idx, T)
    return 1.0 if v is None else math.exp(v / RG / T)


@pytest.mark.parametrize("pair", [(0, 1), (0, 2), (1, 2)])
@pytest.mark.parametrize("p", [0.3, 0.62])
def test_chemical_driving_force_two_phase_closed_form(chem_periodic, pair, p):
    """Every cell holds the same two phases at phi = (p, 1-p), uniform mu,
    EX and T: the Laplacians vanish and the phase rate is the two-phase
    Allen-Cahn form with the grand-potential driving force,

      dphi_0/dt = M/2 [(cap_0 - cap_1) - d/dp(h_0 w_0 + h_1 w_1) 8 sqrt(h_0 h_1)/pi],

    cap_k = -8 e phi_other/eta, and for the cubic interpolant
    d/dp(h_0 w_0 + h_1 w_1) = 6 p (1-p) (w_0 - w_1) along the simplex.  The
    grand potential of phase g is w_g = G_g(y_g)/V_m - sum_s nu_s mu . y_gs
    plus, with a segregation potential V on the interface, V . c h_other
    (c = sum_g h_g sum_s nu_s y_gs); y_g, G_g and h come from the pinned
    Sitefrac, Chemenergy and interpolant."""
    wl = chem_periodic
    cfg, L = wl.cfg, wl.layout
    nd = cfg.ncp - 1
    assert cfg.interpolation == abi.INTERP_CUBIC
    a, b = pair
    rng = np.random.default_rng(a * 7 + b)
    X0 = H.initial_state(wl).reshape(wl.ncells, -1)
    x = np.zeros(L["ndof"])
    x[0], x[1], x[2] = 2, a, b
    x[L["pf_offset"]], x[L["pf_offset"] + 1] = p, 1.0 - p
    x[L["dp_offset"]:L["dp_offset"] + nd] = X0[0, L["dp_offset"]:L["dp_offset"] + nd] * 0.9
    x[L["ex_offset"]:L["ex_offset"] + 2 * L["sp_size"]] = 300.0 * rng.standard_normal(2 * L["sp_size"])
    T = 640.0
    x[L["tm_offset"]] = T
    X = np.tile(x, (wl.ncells, 1))
    F = oracle.rhs(cfg, wl.mesh, X.ravel()).reshape(X.shape)
    it = cfg.interfaces[cfg.interface_list[a * cfg.npf + b]]
    h0 = p * p * (3 - 2 * p)
    h = np.array([h0, 1.0 - h0])
    mui = h[0] * h[1] * np.array(it.potential[:nd]) if it.has_potential else np.zeros(nd)
    mu = x[L["dp_offset"]:L["dp_offset"] + nd]
    w = np.zeros(2)
    cmix = np.zeros(nd)
    for g in range(2):
        mat, y = _phase_y(cfg, L, x, g, mui)
        m = cfg.materials[mat]
        w[g] = oracle.chemenergy_phase(cfg, mat, y, T) / m.molarvolume
        for s in range(m.nsites):
            w[g] -= m.stoichiometry[s] * (mu @ y[s * cfg.ncp:s * cfg.ncp + nd])
            cmix += h[g] * m.stoichiometry[s] * y[s * cfg.ncp:s * cfg.ncp + nd]
    if it.has_potential:
        rv = np.array(it.potential[:nd]) @ cmix
        w = w + rv * h[::-1]
    e = it.e0 * _activation(cfg, it.e_act, T)
    M = it.m0 * _activation(cfg, it.m_act, T)
    eta = it.width
    phi = np.array([p, 1.0 - p])
    cap = -8.0 * e * phi[::-1] / eta
    dw = 6.0 * p * (1.0 - p) * (w[0] - w[1])
    rate0 = M * ((cap[0] - cap[1]) - dw * 8.0 * math.sqrt(h[0] * h[1]) / PI) / 2.0
    got = F[:, L["pf_offset"]:L["pf_offset"] + 2]
    assert abs(dw * 8.0 * math.sqrt(h[0] * h[1]) / PI) > 0.1 * abs(cap[0] - cap[1])   # chemistry matters
    assert np.allclose(got[:, 0], rate0, rtol=1e-9, atol=0.0)
    assert np.allclose(got[:, 1], -rate0, rtol=1e-9, atol=0.0)


def H_dcdt(wl, x, f):
    """Composition rate of one cell: central differences of the composition
    (from the pinned per-phase functions) along the state's rate."""
    from tests.test_oracle_pins import _dcdt
    return _dcdt(wl, x, f)
