"""Planted-error check of the oracle's pins (CPU only, test infrastructure).

Each mutation below changes one line of oracle.c the way a plausible slip in
restating the reference would (a sign, a factor h, a wrong index, an
operation order), builds that variant into a temporary directory, and runs
the oracle pins against it (tests/test_oracle_pins.py and
tests/test_oracle_pins_rhs.py, through CIPHER_ORACLE_LIB).  A mutation is
"killed" when at least one pin fails.  Exit status 1 if any survives.

    python oracle/mutate.py [-k NAME] [-v]
"""
import argparse
import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# (name, reference section, original text, replacement)
MUTATIONS = [
    ("bc_neumann_chem_h", "src/cipher.c:355-358",
     "fd[c] = bc->chem_val[c] * dL;", "fd[c] = bc->chem_val[c] / dL;"),
    ("bc_dirichlet_chem_sign", "src/cipher.c:361-364",
     "fd[c] = mL[c] * (bc->chem_val[c] - xL[L.dpo + c]);", "fd[c] = mL[c] * (xL[L.dpo + c] - bc->chem_val[c]);"),
    ("bc_dirichlet_chem_mobility", "src/cipher.c:361-364",
     "fd[c] = mL[c] * (bc->chem_val[c] - xL[L.dpo + c]);", "fd[c] = mL[c + 1] * (bc->chem_val[c] - xL[L.dpo + c]);"),
    ("bc_mask", "src/cipher.c:361-362",
     "if (bc->chem_mask & (1 << c)) fd[c] = mL[c]", "fd[c] = mL[c]"),
    ("bc_factor_h2", "src/cipher.c:379",
     "double cf = 1.0 / orc_fastpow(dL, 2);", "double cf = 1.0 / orc_fastpow(dL, 1);"),
    ("bc_thermal_neumann_h", "src/cipher.c:368-369",
     "ft = tcond[cL] * bc->thermal_val * dL;", "ft = tcond[cL] * bc->thermal_val;"),
    ("bc_thermal_dirichlet_sign", "src/cipher.c:370-371",
     "ft = tcond[cL] * (bc->thermal_val - xL[L.tmo]);", "ft = tcond[cL] * (xL[L.tmo] - bc->thermal_val);"),
    ("thermal_arithmetic_mean", "src/cipher.c:311-313",
     "double ft = 2.0 * tcond[cL] * tcond[cR] / (tcond[cL] + tcond[cR]) * (xR[L.tmo] - xL[L.tmo]);",
     "double ft = 0.5 * (tcond[cL] + tcond[cR]) * (xR[L.tmo] - xL[L.tmo]);"),
    ("thermal_sign", "src/cipher.c:314-315",
     "lL[L.tmo] += ff * ft / vL;\n            lR[L.tmo] -= ff * ft / vR;",
     "lL[L.tmo] -= ff * ft / vL;\n            lR[L.tmo] += ff * ft / vR;"),
    ("thermal_volume_R", "src/cipher.c:315",
     "lR[L.tmo] -= ff * ft / vR;", "lR[L.tmo] -= ff * ft / vL;"),
    ("tdot_cp_from_phi", "src/cipher.c:562-563",
     "cp_ += h[g] * orc_sum_tseries(T, TS(m->specific_heat));",
     "cp_ += pc[g] * orc_sum_tseries(T, TS(m->specific_heat));"),
    ("tdot_latent_from_h", "src/cipher.c:564-565",
     "*tmrhs += pfrhs[g] * orc_sum_tseries(T, TS(m->latent_heat));",
     "*tmrhs += h[g] * orc_sum_tseries(T, TS(m->latent_heat));"),
    ("tdot_rate_before_division", "src/cipher.c:567-568",
     "if (cp_ != 0.0) *tmrhs /= cp_;\n            if (cfg->has_temperature_rate) *tmrhs += cfg->temperature_rate;",
     "if (cfg->has_temperature_rate) *tmrhs += cfg->temperature_rate;\n            if (cp_ != 0.0) *tmrhs /= cp_;"),
    ("drive_molar_volume", "src/material.c:484",
     "chs[g] /= m->molarvolume;", "chs[g] /= 1.0;"),
    ("drive_mu_sign", "src/material.c:485-487",
     "chs[g] -= mu[c] * m->stoichiometry[s] * sf[g * SF + s * nc + c];",
     "chs[g] += mu[c] * m->stoichiometry[s] * sf[g * SF + s * nc + c];"),
    ("drive_stoichiometry", "src/material.c:485-487",
     "chs[g] -= mu[c] * m->stoichiometry[s] * sf[g * SF + s * nc + c];",
     "chs[g] -= mu[c] * sf[g * SF + s * nc + c];"),
    ("drive_site_index", "src/material.c:485-487",
     "chs[g] -= mu[c] * m->stoichiometry[s] * sf[g * SF + s * nc + c];",
     "chs[g] -= mu[c] * m->stoichiometry[s] * sf[g * SF + c];"),
    ("drive_potential_partner", "src/cipher.c:518",
     "chs[gk] += rv * h[gj];", "chs[gk] += rv * h[gk];"),
    ("drive_sign_in_rate", "src/cipher.c:529-531,546-548",
     "((cap[gk] - cap[gj]) - (chs[gk] - chs[gj]) * 8.0",
     "((cap[gk] - cap[gj]) + (chs[gk] - chs[gj]) * 8.0"),
    ("mobility_factor_2", "src/material.c:1070-1072",
     "mob[k] = y[k] * mb->m0 / RG / T * exp(mig / RG / T);",
     "mob[k] = 2.0 * y[k] * mb->m0 / RG / T * exp(mig / RG / T);"),
    ("mobility_second_sublattice", "src/material.c:1023-1046",
     "mig += y[j] * orc_sum_tseries(T, TS(mb->unary + j));",
     "mig += y[(m->nsites > 1 ? nc : 0) + j] * orc_sum_tseries(T, TS(mb->unary + j));"),
    ("mobility_no_volume", "src/material.c:1117-1119",
     "for (int c = 0; c < nc; c++) mob[c] *= m->molarvolume;", "for (int c = 0; c < nc; c++) mob[c] *= 1.0;"),
]

PIN_FILES = ["tests/test_oracle_pins.py", "tests/test_oracle_pins_rhs.py"]


def run_one(src, mut, tmp, verbose):
    name, ref, old, new = mut
    n = src.count(old)
    if n == 0:
        return name, ref, "NOT FOUND", 0.0
    text = src.replace(old, new)
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "oracle.c"), "w") as f:
        f.write(text)
    shutil.copy(os.path.join(HERE, "oracle.h"), d)
    so = os.path.join(d, "liboracle.so")
    cc = ["gcc", "-O2", "-std=c99", "-fPIC", "-ffp-contract=off", "-shared", f"-I{ROOT}/include", "-o", so,
          os.path.join(d, "oracle.c"), "-lm"]
    # oracle.h includes ../include/cipher_amd.h relative to itself
    os.makedirs(os.path.join(tmp, "include"), exist_ok=True)
    shutil.copy(os.path.join(ROOT, "include", "cipher_amd.h"), os.path.join(tmp, "include"))
    subprocess.run(cc, check=True)
    env = dict(os.environ, CIPHER_ORACLE_LIB=so)
    t0 = time.time()
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] + PIN_FILES,
                       cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    dt = time.time() - t0
    if verbose:
        print(r.stdout[-1500:])
    failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED") or l.startswith("ERROR")]
    status = "killed" if r.returncode != 0 else "SURVIVED"
    if failed:
        status += " by " + failed[0].split("::", 1)[-1].split(" ")[0]
    return name, ref, status, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-k", default="", help="run the mutations whose name contains this")
    ap.add_argument("-v", action="store_true")
    args = ap.parse_args()
    src = open(os.path.join(HERE, "oracle.c")).read()
    bad = 0
    with tempfile.TemporaryDirectory() as tmp:
        for mut in MUTATIONS:
            if args.k and args.k not in mut[0]:
                continue
            name, ref, status, dt = run_one(src, mut, tmp, args.v)
            print(f"{name:30s} {ref:32s} {status} ({dt:.0f}s)", flush=True)
            bad += not status.startswith("killed")
    print("all mutations killed" if not bad else f"{bad} mutation(s) not killed")
    sys.exit(1 if bad else 0)


if __name__ == "__main__":
    main()
