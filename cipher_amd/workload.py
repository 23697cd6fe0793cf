"""Synthetic workloads in CIPHER's YAML schema.

Each builder writes a configuration text in the reference's schema (header,
solution_parameters, material, interface, RLE mappings) from a seed, so that
the same text drives the oracle and the HIP path.  Draws use numpy's PCG64
generator seeded with the examples' random_seed 1579993586.

Workloads (sizes are the defaults; tests pass smaller ones):
  grain_growth_2d    128x128 periodic, 20 Voronoi grains, material 'none',
                     width 4, e0 = m0 = 1, capacity 8, cubic, t_end 100.
  grain_growth_3d    64^3, AMR levels 3-6 at interfaces, 500 grains, interface
                     energies U[0.8,1.2] per pair quantised to 255 types.
  ternary_gb_ppt     128^3, AMR levels 4-7, T = 700 K, 100 calphaddis grains
                     (L0 ~ U[-20,20] kJ/mol, L1 ~ U[-5,5] kJ/mol), 40 calphad2sl
                     (A,B)3(C)1 precipitates of radius 3 on grain boundaries,
                     mobilities log-uniform in [1e-17,1e-15], mu uniform from
                     the nominal composition 0.90/0.07/0.03.
"""
import math

import numpy as np
from scipy import ndimage
from scipy.spatial import cKDTree

from . import config as cfgmod
from . import mesh as meshmod

SEED = 1579993586


def voronoi_labels(res, nseeds, rng, periodic):
    """Nearest-seed labels on the voxel grid, indexed [x,y(,z)]; 0-based."""
    res = np.asarray(res)
    dim = res.size
    seeds = rng.uniform(0.0, 1.0, size=(nseeds, dim)) * res
    grids = np.meshgrid(*[np.arange(int(r)) + 0.5 for r in res], indexing="ij")
    pts = np.stack([g.ravel() for g in grids], axis=1)
    if all(periodic):
        tree = cKDTree(np.mod(seeds, res), boxsize=res.astype(float))
    else:
        tree = cKDTree(seeds)
    _, lab = tree.query(pts, k=1)
    return lab.reshape(tuple(int(r) for r in res)).astype(np.int32)


def _flow(vals):
    return "[" + ", ".join(str(v) for v in vals) + "]"


def _block(text, indent):
    pad = " " * indent
    return "\n".join(pad + line for line in text.split("\n"))


def _voxel_rle(labels):
    # src/init.c: voxel offset = x + res0*(y + res1*z): x fastest
    flat = np.asarray(labels).transpose().ravel() + 1
    return cfgmod.encode_rle(flat)


def _interface_rle(npf, pair_type):
    """interface_mapping over npf*npf entries; pair_type(i, j) (1-based) for
    j > i, 1 elsewhere."""
    vals = np.ones((npf, npf), dtype=np.int64)
    iu = np.triu_indices(npf, k=1)
    vals[iu] = pair_type(iu[0], iu[1])
    return cfgmod.encode_rle(vals.ravel())


def _header(grid, size, npf, materials, interfaces, components):
    return (f"---\nheader:\n  grid : {_flow(grid)}\n  size : {_flow(size)}\n"
            f"  n_phases : {npf}\n  materials : {_flow(materials)}\n"
            f"  interfaces : {_flow(interfaces)}\n  components : {_flow(components)}\n"
            f"  outputs : [phaseid]\n\n")


def _solution(times, blocks, initrefine, minlevel, maxlevel, capacity, tmin, tmax,
              periodic, reltol=1e-4, abstol=1e-4, amrinterval=1000000):
    opts = "-ts_adapt_monitor"
    if any(periodic):
        opts += " -dm_p4est_brick_periodicity " + ",".join("1" if p else "0" for p in periodic)
    return (f"solution_parameters:\n  time : {_flow(times)}\n"
            f"  initblocksize : {_flow(blocks)}\n  initrefine : {initrefine}\n"
            f"  maxnrefine : {maxlevel}\n  minnrefine : {minlevel}\n  initcoarsen : 0\n"
            f"  amrinterval : {amrinterval}\n  outputfreq : 1000000\n  interpolation: cubic\n"
            f"  reltol: {reltol}\n  abstol: {abstol}\n  timestepmin: {tmin}\n"
            f"  timestepmax: {tmax}\n  max_phase_neighbours : {capacity}\n"
            f"  petscoptions : {opts}\n  random_seed : {SEED}\n\n")


def grain_growth_yaml(res=(128, 128), ngrains=20, periodic=True, capacity=8, t_end=100.0,
                      energy_range=None, ntypes=255, width=4.0, seed=SEED, minlevel=None,
                      tmin=1e-3):
    """Isotropic grain growth (material 'none', single component)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dim = len(res)
    per = [periodic] * dim
    labels = voronoi_labels(res, ngrains, rng, per)
    lmax = int(round(math.log2(res[0])))
    minlevel = lmax if minlevel is None else minlevel
    if energy_range is None:
        interfaces = ["grainboundary"]
        inter = ("interface:\n  grainboundary:\n    width : %r\n    energy :\n      e0: 1.0\n"
                 "    mobility :\n      m0: 1.0\n\n" % width)
        imap = _interface_rle(ngrains, lambda i, j: np.ones_like(i))
    else:
        lo, hi = energy_range
        e_pair = rng.uniform(lo, hi, size=(ngrains, ngrains))
        types = np.clip(((e_pair - lo) / (hi - lo) * ntypes).astype(np.int64), 0, ntypes - 1)
        centres = lo + (np.arange(ntypes) + 0.5) * (hi - lo) / ntypes
        interfaces = [f"gb{t}" for t in range(ntypes)]
        inter = "interface:\n" + "".join(
            f"  gb{t}:\n    width : {float(width)!r}\n    energy :\n      e0: {float(centres[t])!r}\n"
            f"    mobility :\n      m0: 1.0\n" for t in range(ntypes)) + "\n"
        imap = _interface_rle(ngrains, lambda i, j: types[i, j] + 1)
    text = _header(list(res), [float(r) for r in res], ngrains, ["matrix"], interfaces, ["ti"])
    text += _solution([t_end], [1] * dim, lmax, minlevel, lmax, capacity, tmin, t_end, per)
    text += ("material:\n  matrix:\n    chemicalenergy : none\n    molarvolume : 1.0\n"
             "    temperature0: 1.0\n\n")
    text += inter
    text += "mappings:\n  phase_material_mapping : |-\n" + _block(f"{ngrains} of 1", 4) + "\n"
    text += "  voxel_phase_mapping: |-\n" + _block(_voxel_rle(labels), 4) + "\n"
    text += "  interface_mapping: |-\n" + _block(imap, 4) + "\n"
    return text


def ternary_ppt_yaml(res=(128, 128, 128), ngrains=100, nppt=40, radius=3, capacity=8,
                     minlevel=4, seed=SEED, temperature=700.0, t_end=1.0e-3, dx=1.0e-9):
    """Ternary grain-boundary precipitation (calphaddis matrix, calphad2sl ppt)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dim = len(res)
    per = [False] * dim
    labels = voronoi_labels(res, ngrains, rng, per)
    gb = ndimage.maximum_filter(labels, size=3, mode="nearest") != ndimage.minimum_filter(
        labels, size=3, mode="nearest")
    cand = np.argwhere(gb)
    pick = cand[rng.choice(cand.shape[0], size=nppt, replace=False)]
    grids = np.meshgrid(*[np.arange(int(r)) for r in res], indexing="ij")
    for k, c in enumerate(pick):
        d2 = sum((grids[d] - c[d]) ** 2 for d in range(dim))
        labels[d2 <= radius * radius] = ngrains + k
    npf = ngrains + nppt
    comps = ["A", "B", "C"]
    RT = 8.314462 * temperature
    L0 = rng.uniform(-20e3, 20e3, size=3)
    L1 = rng.uniform(-5e3, 5e3, size=3)
    M = 10.0 ** rng.uniform(-17.0, -15.0, size=3)
    G_em = rng.uniform(-30e3, -10e3, size=2)
    penalty = 20e3

    def mob_block(ind):
        s = " " * ind + "mobilityc :\n"
        for c, nm in enumerate(comps):
            s += " " * (ind + 2) + f"{nm} :\n" + " " * (ind + 4) + f"mobility0 : {float(M[c] * RT)!r}\n"
        return s

    lmax = int(round(math.log2(res[0])))
    size = [r * dx for r in res]
    text = _header(list(res), size, npf, ["matrix", "precipitate"],
                   ["grainboundary", "pptboundary", "pptppt"], comps)
    text += _solution([t_end], [1] * dim, lmax, minlevel, lmax, capacity, 1e-12, t_end, per)
    text += "material:\n  matrix:\n    chemicalenergy : calphaddis\n    molarvolume : 1e-5\n"
    text += f"    temperature0: {float(temperature)!r}\n    chempot_ex_kineticcoeff : 1.0\n"
    text += "    c0 : [0.90, 0.07, 0.03]\n" + mob_block(4)
    text += "    binary_enthalpy :\n"
    pairs = [(0, 1), (0, 2), (1, 2)]
    for k, (i, j) in enumerate(pairs):
        text += (f"      {comps[i]} :\n        {comps[j]} :\n          nrk : 2\n"
                 f"          rk_0 :\n            t_coefficient : {float(L0[k])!r}\n            t_exponent : 0\n"
                 f"          rk_1 :\n            t_coefficient : {float(L1[k])!r}\n            t_exponent : 0\n")
    text += "  precipitate:\n    chemicalenergy : calphad2sl\n    molarvolume : 1e-5\n"
    text += f"    temperature0: {float(temperature)!r}\n    chempot_ex_kineticcoeff : 1.0\n"
    text += "    stochiometry : [0.75, 0.25]\n"
    text += "    c0_p : [0.70, 0.29, 0.01]\n    c0_q : [0.01, 0.01, 0.98]\n" + mob_block(4)
    text += "    unary_enthalpy :\n"
    for cp in range(3):
        for cq in range(3):
            if cq == 2 and cp < 2:
                g = G_em[cp]
            else:
                g = penalty
            text += f"      ({comps[cp]}:{comps[cq]}) :\n        t_coefficient : {float(g)!r}\n        t_exponent : 0\n"
    text += "\n"
    width = 4.0 * dx
    text += "interface:\n"
    for nm, e0 in (("grainboundary", 0.5), ("pptboundary", 0.3), ("pptppt", 0.5)):
        text += (f"  {nm}:\n    width : {float(width)!r}\n    energy :\n      e0: {e0}\n"
                 f"    mobility :\n      m0: 1.0e-13\n")
    text += "\n"

    def ptype(i, j):
        ti = (i >= ngrains).astype(np.int64) + (j >= ngrains).astype(np.int64)
        return ti + 1

    text += "mappings:\n  phase_material_mapping : |-\n" + _block(f"{ngrains} of 1\n{nppt} of 2", 4) + "\n"
    text += "  voxel_phase_mapping: |-\n" + _block(_voxel_rle(labels), 4) + "\n"
    text += "  interface_mapping: |-\n" + _block(_interface_rle(npf, ptype), 4) + "\n"
    return text


class Workload:
    """Parsed config + leaf forest + flattened mesh + per-cell phase labels."""

    def __init__(self, text, name=""):
        self.name = name
        self.text = text
        self.cfg = cfgmod.load_config(text)
        c = self.cfg
        res = c.grid
        labels = c.voxel_phase.reshape(tuple(res[::-1])).transpose()
        self.labels = labels
        if c.minnrefine >= c.maxnrefine or c.initrefine == c.minnrefine:
            self.forest = meshmod.uniform_forest(res, c.initblocksize, c.size, c.periodic,
                                                 level=c.maxnrefine)
        else:
            self.forest = meshmod.adapt_to_labels(labels, c.initblocksize, c.size, c.periodic,
                                                  c.minnrefine, c.maxnrefine)
        self.mesh = self.forest.flatten()
        self.phase = meshmod.cell_labels(self.forest, labels)
        self.layout = c.layout()

    @property
    def ncells(self):
        return self.mesh.ncells


def uniform_mu_override(X, layout, phase, phase_material, matrix_material=0):
    """Set every cell's chemical potential to that of a single-phase matrix
    cell (the 'mu uniform from the nominal composition' initial state)."""
    L = layout
    D = L["ndof"]
    Xv = X.reshape(-1, D)
    n = np.rint(Xv[:, L["as_offset"]]).astype(np.int64)
    cand = np.flatnonzero((n == 1) & (phase_material[phase] == matrix_material))
    if cand.size == 0 or L["dp_size"] == 0:
        return X
    mu = Xv[cand[0], L["dp_offset"]:L["dp_offset"] + L["dp_size"]].copy()
    Xv[:, L["dp_offset"]:L["dp_offset"] + L["dp_size"]] = mu
    return X
