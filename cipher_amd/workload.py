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


def solution_2d_yaml(res=(128, 128), ngrains=50, ncomp=3, ntypes=3, capacity=8, minlevel=None, seed=SEED,
                     temperature=700.0, t_end=1.0e-3, dx=1.0e-9):
    """2D grand-canonical microstructure with ncomp components and ntypes
    phase types: 'sol0' and 'sol1' are calphaddis solutions with
    Redlich-Kister binaries (L0 ~ U[-20,20] kJ/mol, L1 ~ U[-5,5] kJ/mol per
    pair), 'cef' a calphad2sl (.)3(.)1 phase whose end members with the last
    component on the second sublattice are drawn in [-30,-10] kJ/mol (others
    carry a 20 kJ/mol penalty).  Mobilities log-uniform in [1e-17,1e-15];
    grains are Voronoi cells of a periodic domain, assigned to the phase
    types with weights 0.7/0.15/0.15; mu starts uniform (uniform_mu_override)."""
    if not 2 <= ncomp <= 4 or not 1 <= ntypes <= 3:
        raise ValueError("2 <= ncomp <= 4 and 1 <= ntypes <= 3")
    rng = np.random.Generator(np.random.PCG64(seed))
    dim = len(res)
    per = [True] * dim
    labels = voronoi_labels(res, ngrains, rng, per)
    comps = ["A", "B", "C", "D"][:ncomp]
    RT = 8.314462 * temperature
    lmax = int(round(math.log2(res[0])))
    minlevel = lmax if minlevel is None else minlevel
    w = np.array([0.7, 0.15, 0.15][:ntypes])
    gtype = rng.choice(ntypes, size=ngrains, p=w / w.sum())
    gtype[0] = 0
    names = ["sol0", "sol1", "cef"][:ntypes]
    nominal = {2: [0.9, 0.1], 3: [0.90, 0.07, 0.03], 4: [0.88, 0.07, 0.03, 0.02]}[ncomp]

    def mob_block(ind):
        M = 10.0 ** rng.uniform(-17.0, -15.0, size=ncomp)
        out = " " * ind + "mobilityc :\n"
        for c, nm in enumerate(comps):
            out += " " * (ind + 2) + f"{nm} :\n" + " " * (ind + 4) + f"mobility0 : {float(M[c] * RT)!r}\n"
        return out

    def head(name, model):
        return (f"  {name}:\n    chemicalenergy : {model}\n    molarvolume : 1e-5\n"
                f"    temperature0: {float(temperature)!r}\n    chempot_ex_kineticcoeff : 1.0\n")

    text = _header(list(res), [r * dx for r in res], ngrains, names,
                   [f"if{a}{b}" for a in range(ntypes) for b in range(a, ntypes)], comps)
    text += _solution([t_end], [1] * dim, lmax, minlevel, lmax, capacity, 1e-12, t_end, per)
    text += "material:\n"
    for t in range(min(ntypes, 2)):
        c0 = list(nominal) if t == 0 else [nominal[0] - 0.1] + [v + 0.1 / (ncomp - 1) for v in nominal[1:]]
        text += head(names[t], "calphaddis") + f"    c0 : {_flow([round(v, 6) for v in c0])}\n" + mob_block(4)
        text += "    binary_enthalpy :\n"
        for i in range(ncomp):
            for j in range(i + 1, ncomp):
                L0 = rng.uniform(-20e3, 20e3)
                L1 = rng.uniform(-5e3, 5e3)
                text += (f"      {comps[i]} :\n        {comps[j]} :\n          nrk : 2\n"
                         f"          rk_0 :\n            t_coefficient : {float(L0)!r}\n            t_exponent : 0\n"
                         f"          rk_1 :\n            t_coefficient : {float(L1)!r}\n            t_exponent : 0\n")
    if ntypes == 3:
        last = ncomp - 1
        c0p = [0.01] * ncomp
        c0p[0] = 1.0 - 0.01 * (ncomp - 1)
        c0q = [0.01] * ncomp
        c0q[last] = 1.0 - 0.01 * (ncomp - 1)
        text += head("cef", "calphad2sl") + "    stochiometry : [0.75, 0.25]\n"
        text += f"    c0_p : {_flow([round(v, 6) for v in c0p])}\n    c0_q : {_flow([round(v, 6) for v in c0q])}\n"
        text += mob_block(4) + "    unary_enthalpy :\n"
        for cp in range(ncomp):
            for cq in range(ncomp):
                g = rng.uniform(-30e3, -10e3) if (cq == last and cp < last) else 20e3
                text += (f"      ({comps[cp]}:{comps[cq]}) :\n        t_coefficient : {float(g)!r}\n"
                         f"        t_exponent : 0\n")
    text += "\ninterface:\n"
    for a in range(ntypes):
        for b in range(a, ntypes):
            e0 = 0.5 if a == b else 0.3
            text += (f"  if{a}{b}:\n    width : {float(4.0 * dx)!r}\n    energy :\n      e0: {e0}\n"
                     f"    mobility :\n      m0: 1.0e-13\n")
    text += "\n"
    pair_id = {}
    k = 0
    for a in range(ntypes):
        for b in range(a, ntypes):
            pair_id[(a, b)] = k
            k += 1
    lut = np.zeros((ntypes, ntypes), dtype=np.int64)
    for (a, b), v in pair_id.items():
        lut[a, b] = lut[b, a] = v

    def ptype(i, j):
        return lut[gtype[i], gtype[j]] + 1

    pm = "\n".join(f"1 of {int(t) + 1}" for t in gtype)
    text += "mappings:\n  phase_material_mapping : |-\n" + _block(pm, 4) + "\n"
    text += "  voxel_phase_mapping: |-\n" + _block(_voxel_rle(labels), 4) + "\n"
    text += "  interface_mapping: |-\n" + _block(_interface_rle(ngrains, ptype), 4) + "\n"
    return text


def ensemble_specs(n=96, seed=SEED, min_side=64, max_side=1024, max_grains=2000):
    """Instances of the batched ensemble: side log-uniform (a power of two
    in [min_side, max_side]; AMR from three levels below the finest),
    grains log-uniform in [10, max_grains] (at most one per 64 voxels),
    components uniform in 1..4, phase types uniform in 1..3 (1 without
    chemistry), capacity 8 (every eighth instance 4, with at most one grain
    per 64x64 voxels so that no union of neighbouring sets exceeds 4; every
    twelfth 16)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = int(round(math.log2(min_side))), int(round(math.log2(max_side)))
    out = []
    for i in range(n):
        k = int(rng.integers(lo, hi + 1))
        side = 1 << k
        g = int(round(math.exp(rng.uniform(math.log(10), math.log(max_grains)))))
        g = max(10, min(g, side * side // 64))
        nc = int(rng.integers(1, 5))
        nt = int(rng.integers(1, 4)) if nc > 1 else 1
        cap = 4 if i % 8 == 7 else (16 if i % 12 == 11 else 8)
        if cap == 4:
            g = max(10, min(g, side * side // 4096))
        out.append({"side": side, "grains": g, "components": nc, "types": nt, "capacity": cap,
                    "minlevel": max(k - 3, 3), "seed": seed + 7919 * (i + 1)})
    return out


def ensemble_member(spec):
    """Workload of one ensemble instance (see ensemble_specs)."""
    res = (spec["side"], spec["side"])
    if spec["components"] == 1:
        txt = grain_growth_yaml(res=res, ngrains=spec["grains"], periodic=True, capacity=spec["capacity"],
                                minlevel=spec["minlevel"], seed=spec["seed"])
    else:
        txt = solution_2d_yaml(res=res, ngrains=spec["grains"], ncomp=spec["components"], ntypes=spec["types"],
                               capacity=spec["capacity"], minlevel=spec["minlevel"], seed=spec["seed"])
    return Workload(txt, "ensemble_{side}_{grains}g_{components}c_{types}t".format(**spec))


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
