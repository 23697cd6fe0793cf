"""Output fields of the oracle (src/cipher.c:1005-1080) and the YAML name
parser, pinned on the CPU.

* phase id / material / interface id / phi / mu / T: recomputed from the
  state with the definitions (largest interpolant, first on ties; the last
  active pair's interface type; -1 for one phase);
* composition: sums to sum_s nu_s in every cell (each sublattice's site
  fractions sum to 1), and in the SetUpProblem state of a calphaddis phase
  (mu and the explicit potential set from the material's c0) a single-phase
  cell's composition is c0: the site potentials, explicit potential and
  site fractions have to invert one another for that to hold.
"""
import numpy as np
import pytest

import oracle
from cipher_amd import abi
from cipher_amd import config as cfgmod
from tests import helpers as H

ALL = [(abi.OUT_PHASEID, 0), (abi.OUT_MATID, 0), (abi.OUT_INTERFACEID, 0), (abi.OUT_TEMPERATURE, 0)]


def _fields(wl):
    f = list(ALL)
    f += [(abi.OUT_CHEMPOT, c) for c in range(wl.cfg.ncp - 1)]
    f += [(abi.OUT_COMP, c) for c in range(wl.cfg.ncp)]
    f += [(abi.OUT_PHI, p) for p in range(min(wl.cfg.npf, 4))]
    return f


@pytest.mark.parametrize("name", ["gg2d", "ternary", "chem"])
def test_output_definitions(name):
    wl = getattr(H, name)()
    X = H.perturb(wl, H.evolve(wl, H.initial_state(wl), steps=2))
    fields = _fields(wl)
    out = oracle.output(wl.cfg, wl.mesh, X, fields)
    L = wl.layout
    D = L["ndof"]
    Xv = X.reshape(-1, D)
    npf = wl.cfg.npf
    ilist = np.ctypeslib.as_array(wl.cfg.struct.interface_list, shape=(npf * npf,))
    pmat = np.ctypeslib.as_array(wl.cfg.struct.phase_material, shape=(npf,))
    for cell in range(0, wl.ncells, max(1, wl.ncells // 200)):
        x = Xv[cell]
        n = int(round(x[L["as_offset"]]))
        ids = [int(round(v)) for v in x[L["as_offset"] + 1:L["as_offset"] + 1 + n]]
        phi = x[L["pf_offset"]:L["pf_offset"] + n]
        h = oracle.interpolant(wl.cfg.interpolation, phi)
        g = int(np.argmax(h))          # first maximum
        row = dict(zip(fields, out[cell]))
        assert row[(abi.OUT_PHASEID, 0)] == ids[g]
        assert row[(abi.OUT_MATID, 0)] == pmat[ids[g]]
        assert row[(abi.OUT_INTERFACEID, 0)] == (ilist[ids[n - 2] * npf + ids[n - 1]] if n > 1 else -1)
        assert row[(abi.OUT_TEMPERATURE, 0)] == x[L["tm_offset"]]
        for c in range(wl.cfg.ncp - 1):
            assert row[(abi.OUT_CHEMPOT, c)] == x[L["dp_offset"] + c]
        for p in range(min(npf, 4)):
            assert row[(abi.OUT_PHI, p)] == (phi[ids.index(p)] if p in ids else 0.0)


@pytest.mark.parametrize("name", ["ternary", "chem"])
def test_composition_sums_to_the_stoichiometry(name):
    wl = getattr(H, name)()
    X = H.perturb(wl, H.evolve(wl, H.initial_state(wl), steps=2))
    nc = wl.cfg.ncp
    out = oracle.output(wl.cfg, wl.mesh, X, [(abi.OUT_COMP, c) for c in range(nc)])
    L = wl.layout
    Xv = X.reshape(-1, L["ndof"])
    pmat = np.ctypeslib.as_array(wl.cfg.struct.phase_material, shape=(wl.cfg.npf,))
    for cell in range(wl.ncells):
        n = int(round(Xv[cell, L["as_offset"]]))
        ids = [int(round(v)) for v in Xv[cell, L["as_offset"] + 1:L["as_offset"] + 1 + n]]
        h = oracle.interpolant(wl.cfg.interpolation, Xv[cell, L["pf_offset"]:L["pf_offset"] + n])
        want = sum(h[g] * sum(wl.cfg.struct.material[pmat[ids[g]]].stoichiometry[s]
                              for s in range(wl.cfg.struct.material[pmat[ids[g]]].nsites)) for g in range(n))
        assert out[cell].sum() == pytest.approx(want, rel=1e-12, abs=1e-12)


def test_setup_composition_is_c0_in_calphaddis_cells():
    wl = H.ternary()
    X = oracle.setup_problem(wl.cfg, wl.mesh, wl.phase)
    nc = wl.cfg.ncp
    out = oracle.output(wl.cfg, wl.mesh, X, [(abi.OUT_COMP, c) for c in range(nc)])
    L = wl.layout
    Xv = X.reshape(-1, L["ndof"])
    pmat = np.ctypeslib.as_array(wl.cfg.struct.phase_material, shape=(wl.cfg.npf,))
    checked = 0
    for cell in range(wl.ncells):
        n = int(round(Xv[cell, L["as_offset"]]))
        if n != 1:
            continue
        m = wl.cfg.struct.material[pmat[int(round(Xv[cell, L["as_offset"] + 1]))]]
        if m.model != abi.CHEM_CALPHADDIS:
            continue
        np.testing.assert_allclose(out[cell], [m.c0[c] for c in range(nc)], rtol=1e-12, atol=1e-15)
        checked += 1
    assert checked > 0


def test_output_names_parse_like_the_reference():
    wl = H.chem()
    comps = wl.cfg.component_names
    names = ["phaseid", "matid", "interfaceid", "temperature", f"{comps[0]}_chempot", f"{comps[-1]}_chempot",
             f"{comps[1]}_c", f"{comps[-1]}_c", "3_phi", "XX_c"]
    f = cfgmod.output_fields(wl.cfg, names)
    assert f == [(abi.OUT_PHASEID, 0), (abi.OUT_MATID, 0), (abi.OUT_INTERFACEID, 0), (abi.OUT_TEMPERATURE, 0),
                 (abi.OUT_CHEMPOT, 0), (abi.OUT_CHEMPOT, -1),   # the last component has no chemical potential
                 (abi.OUT_COMP, 1), (abi.OUT_COMP, len(comps) - 1), (abi.OUT_PHI, 3), (abi.OUT_COMP, -1)]
    with pytest.raises(ValueError):
        cfgmod.output_fields(wl.cfg, ["nonsense"])
    # an unmatched component leaves the output at 0
    X = H.initial_state(wl)
    out = oracle.output(wl.cfg, wl.mesh, X, [(abi.OUT_COMP, -1), (abi.OUT_CHEMPOT, len(comps) - 1)])
    assert not out.any()
    # the reference examples name their outputs in the header
    assert wl.cfg.output_names == ["phaseid"]
