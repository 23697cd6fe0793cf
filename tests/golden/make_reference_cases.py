"""Derive small parity fixtures from the reference's example configurations.

Reads micmog/CIPHER's examples/*/*.yaml (data: material databases, interface
tables, boundary conditions) and writes tests/golden/ref_<name>.yaml: the same
header, materials, interfaces, boundaries and solver options on a 16^2 (2D) or
8^3 (3D) uniform grid with the same cell size, and a seeded Voronoi
microstructure over a few phases chosen so that every material appears.
Run from the repository root in a checkout that has the reference next to it:

    python tests/golden/make_reference_cases.py /root/reference/examples
"""
import glob
import math
import os
import sys

import numpy as np
import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cipher_amd import config as C  # noqa: E402

CASES = {
    "2SL/4x1024_2SL_T250.yaml": "sl2_t250",
    "BCs/Cr15-Ti-V10-1557K_chemBC.yaml": "bc_chem",
    "BCs/Cr15-Ti-V10-1557K_thermalBC.yaml": "bc_thermal",
    "GrainBoundaryPrecipitate/GrainBoundaryPrecipitate.yaml": "gb_precipitate",
    "GrainGrowth/256x256_N50.yaml": "grain_growth",
    "Nucleation/1024x1024N128_7050.yaml": "nucleation_7050",
    "Nucleation/256x256N225_Cr10-Ti.yaml": "nucleation_crti",
    "Segregation/BiXtal_1024x1024.yaml": "segregation",
    "Solidification/Cr15-Ti-V10-1557K.yaml": "solidification",
    "Anisotropy/1024x1024_Cr10-Ti_cubic.yaml": "anisotropy",
}
SEED = 1579993586


def reduce_case(text, seed):
    root = yaml.safe_load(text)
    cfg = C.load_config(text)
    hdr = root["header"]
    grid_old = [int(v) for v in hdr["grid"]]
    size_old = [float(v) for v in hdr["size"]]
    dim = len(grid_old)
    n = 16 if dim == 2 else 8
    level = int(round(math.log2(n)))
    grid = [n] * dim
    # keep the cell size of the finest level of the original grid
    hdr["grid"] = grid
    hdr["size"] = [size_old[d] / grid_old[d] * n for d in range(dim)]
    sp = root["solution_parameters"]
    sp["initblocksize"] = [1] * dim
    sp["initrefine"] = level
    sp["maxnrefine"] = level
    sp["minnrefine"] = level
    sp["initcoarsen"] = 0
    sp["amrinterval"] = 1000000
    # microstructure: one seed per chosen phase, every material represented
    npf = cfg.npf
    pm = np.asarray(cfg.phase_material)
    rng = np.random.Generator(np.random.PCG64(seed))
    chosen = []
    for m in range(len(cfg.materials)):
        ph = np.flatnonzero(pm == m)
        if ph.size:
            chosen.append(int(rng.choice(ph)))
    rest = [p for p in range(npf) if p not in chosen]
    extra = min(len(rest), max(0, 5 - len(chosen)))
    if extra:
        chosen += [int(p) for p in rng.choice(rest, size=extra, replace=False)]
    seeds = rng.uniform(0, n, size=(len(chosen), dim))
    grids = np.meshgrid(*[np.arange(n) + 0.5] * dim, indexing="ij")
    pts = np.stack([g.ravel() for g in grids], axis=1)
    d2 = ((pts[:, None, :] - seeds[None, :, :]) ** 2).sum(axis=2)
    labels = np.asarray(chosen)[np.argmin(d2, axis=1)].reshape([n] * dim)
    flat = labels.transpose().ravel() + 1                  # x fastest (src/init.c)
    root["mappings"]["voxel_phase_mapping"] = C.encode_rle(flat)
    return root


def main(src):
    for rel, name in CASES.items():
        path = os.path.join(src, rel)
        text = open(path).read()
        root = reduce_case(text, SEED)
        out = os.path.join(HERE, f"ref_{name}.yaml")
        with open(out, "w") as f:
            f.write(f"# derived from micmog/CIPHER examples/{rel} by make_reference_cases.py\n")
            yaml.safe_dump(root, f, sort_keys=False, default_flow_style=None, width=120)
        cfg = C.load_config(open(out).read())
        print(f"{name}: npf={cfg.npf} ncp={cfg.ncp} ndof={cfg.layout()['ndof']} -> {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/examples")
