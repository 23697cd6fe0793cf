"""Reader of CIPHER's YAML configuration schema (src/init.c:104-1716).

``load_config(text)`` parses a configuration in the reference's schema and
returns a :class:`CipherConfig` whose ``struct`` is the ``cipher_config`` the
library consumes.  Values follow the reference's parsing rules: scalars and
flow sequences through ``atof``/``atoi``, RLE mapping blocks (``N of P``,
``A to B``, bare values; src/init.c:1346-1712), interface ids cast to
``unsigned char`` (src/init.c:1501), layout sizes of src/init.c:1330-1343.
"""
import ctypes as C
import math
import re

import numpy as np
import yaml

from . import abi


def _atof(v):
    if isinstance(v, bool):
        return float(v)
    if isinstance(v, (int, float)):
        return float(v)
    s = str(v).strip()
    try:
        return float(s)
    except ValueError:
        # atof: longest numeric prefix, 0.0 if none
        for k in range(len(s), 0, -1):
            try:
                return float(s[:k])
            except ValueError:
                continue
        return 0.0


def _atoi(v):
    if isinstance(v, (int,)) and not isinstance(v, bool):
        return int(v)
    s = str(v).strip()
    if s.isdigit():
        return int(s)
    digits = ""
    for ch in s:
        if ch in "+-" and not digits:
            digits += ch
        elif ch.isdigit():
            digits += ch
        else:
            break
    try:
        return int(digits)
    except ValueError:
        return 0


class _Doc:
    """Path lookup in the parsed YAML, returning lists like GetProperty."""

    def __init__(self, root):
        self.root = root

    def node(self, path):
        cur = self.root
        for part in [p for p in path.split("/") if p]:
            if not isinstance(cur, dict):
                return None
            if part in cur:
                cur = cur[part]
            else:
                # keys are compared as text (e.g. "rk_0", "(al:cu)", numeric keys)
                hit = None
                for k in cur:
                    if str(k) == part:
                        hit = k
                        break
                if hit is None:
                    return None
                cur = cur[hit]
        return cur

    def prop(self, path, key):
        n = self.node(path)
        if not isinstance(n, dict):
            return []
        v = None
        found = False
        for k in n:
            if str(k) == key:
                v = n[k]
                found = True
                break
        if not found or v is None:
            return []
        if isinstance(v, list):
            return v
        if isinstance(v, dict):
            return []
        return [v]


def parse_rle(text, count):
    """RLE mapping lines (src/init.c:1372-1406): 'N of P', 'A to B', 'P'.
    Returns a list of 1-based values (before the reference's '-1')."""
    out = []
    for tok in str(text).split("\n"):
        if len(out) >= count:
            break
        if "of" in tok:
            a, b = tok.split("of")[0], tok.split("of")[1]
            n, v = _atoi(a.split()[0]), _atoi(b.split()[0])
            out.extend([v] * n)
        elif "to" in tok:
            a, b = tok.split("to")[0], tok.split("to")[1]
            lo, hi = _atoi(a.split()[0]), _atoi(b.split()[0])
            step = 1 if lo < hi else -1
            out.extend(range(lo, hi + step, step))
        elif _atoi(tok):
            out.append(_atoi(tok))
    return out


def encode_rle(values):
    """Inverse of parse_rle for 1-based values: 'N of P' runs."""
    lines = []
    vals = np.asarray(values, dtype=np.int64).ravel()
    if vals.size == 0:
        return ""
    change = np.flatnonzero(np.diff(vals)) + 1
    starts = np.concatenate([[0], change])
    ends = np.concatenate([change, [vals.size]])
    for s, e in zip(starts, ends):
        lines.append(f"{e - s} of {vals[s]}")
    return "\n".join(lines)


class _Pools:
    def __init__(self):
        self.ts = [self._ts([], [], 0.0)]   # index 0: the empty series
        self.rk = []
        self.rkindex = []
        self.mob = []
        self.src = []
        self.aniso = []

    @staticmethod
    def _ts(coeff, exps, logc):
        t = abi.TSeries()
        n = min(len(coeff), abi.MAXTSER)
        t.n = n
        for i in range(n):
            t.coeff[i] = coeff[i]
            t.exp[i] = exps[i]
        t.logcoeff = logc
        return t

    def add_ts(self, doc, path, with_log=False):
        coeff = [_atof(v) for v in doc.prop(path, "t_coefficient")]
        exps = [_atoi(v) for v in doc.prop(path, "t_exponent")]
        if len(exps) != len(coeff):
            raise ValueError(f"{path}: t_exponent/t_coefficient size mismatch")
        logc = 0.0
        if with_log:
            lc = doc.prop(path, "tlnt_coefficient")
            logc = _atof(lc[0]) if lc else 0.0
        self.ts.append(self._ts(coeff, exps, logc))
        return len(self.ts) - 1

    def add_rk(self, doc, path, ternary=False):
        nrk = doc.prop(path, "nrk")
        n = _atoi(nrk[0]) if nrk else 0
        ts0 = len(self.ts)
        idx0 = len(self.rkindex)
        for o in range(n):
            self.add_ts(doc, f"{path}/rk_{o}")
            if ternary:
                ix = doc.prop(f"{path}/rk_{o}", "index")
                if len(ix) != 1:
                    raise ValueError(f"{path}/rk_{o}: ternary 'index' required")
                self.rkindex.append(_atoi(ix[0]))
        r = abi.RK()
        r.n, r.ts, r.idx = n, ts0 if n else 0, idx0
        self.rk.append(r)
        return len(self.rk) - 1

    def add_sources(self, doc, path, ncp):
        names = [str(v) for v in doc.prop(path, "activesources")]
        first = len(self.src)
        for nm in names:
            p = f"{path}/{nm}"
            model = str(doc.prop(p, "source")[0])
            s = abi.Source()
            if model == "sink":
                s.model = abi.SOURCE_SINK
                for key, arr in (("rate", s.rate), ("ceq", s.ceq)):
                    vals = doc.prop(p, key)
                    for c in range(ncp):
                        arr[c] = _atof(vals[c]) if vals else 0.0
            elif model == "random":
                s.model = abi.SOURCE_RANDOM
                vals = doc.prop(p, "fluctuation_amplitute")
                for c in range(min(ncp, len(vals))):
                    s.amplitude[c] = _atof(vals[c])
            else:
                s.model = -1
            self.src.append(s)
        return len(names), first


def _arr(ctype, items):
    n = max(len(items), 1)
    a = (ctype * n)()
    for i, it in enumerate(items):
        a[i] = it
    return a


class CipherConfig:
    """Parsed configuration: ``struct`` (cipher_config), layout and run data."""

    def __init__(self):
        self.struct = abi.Config()
        self._keep = []

    # layout of src/init.c:1330-1343
    def layout(self):
        mpn = self.max_phase_neighbours
        L = {}
        L["maxsites"] = max(m.nsites for m in self.materials)
        L["sf_size"] = L["maxsites"] * self.ncp
        L["sp_size"] = L["maxsites"] * (self.ncp - 1)
        L["as_offset"] = 0
        L["as_size"] = min(mpn, self.npf) + 1
        L["pf_offset"] = L["as_size"]
        L["pf_size"] = min(mpn, self.npf)
        L["dp_offset"] = L["pf_offset"] + L["pf_size"]
        L["dp_size"] = self.ncp - 1
        L["ex_offset"] = L["dp_offset"] + L["dp_size"]
        L["ex_size"] = L["pf_size"] * L["sp_size"]
        L["tm_offset"] = L["ex_offset"] + L["ex_size"]
        L["tm_size"] = 1
        L["ndof"] = L["tm_offset"] + 1
        return L


def load_config(text, loadcase=0):
    """Parse reference-schema YAML text into a CipherConfig."""
    # libyaml's safe loader when present (same documents, much faster on the
    # long RLE mapping blocks); it constructs no arbitrary objects either
    loader = getattr(yaml, "CSafeLoader", yaml.SafeLoader)
    root = yaml.load(text, Loader=loader)
    doc = _Doc(root)
    cfg = CipherConfig()
    pools = _Pools()

    grid = [_atoi(v) for v in doc.prop("header", "grid")]
    if not 1 <= len(grid) <= 3:
        raise ValueError("header/grid must have 1..3 entries")
    dim = len(grid)
    size = [_atof(v) for v in doc.prop("header", "size")]
    if len(size) != dim:
        raise ValueError("header/size must match header/grid")
    npf = _atoi(doc.prop("header", "n_phases")[0])
    matnames = [str(v) for v in doc.prop("header", "materials")]
    ifnames = [str(v) for v in doc.prop("header", "interfaces")]
    comps = [str(v) for v in doc.prop("header", "components")]
    bcnames = [str(v) for v in doc.prop("header", "boundaries")]
    ncp = len(comps)
    if not 0 < ncp <= abi.MAXCP:
        raise ValueError("header/components: 1..5 components")
    nsites_hdr = doc.prop("header", "n_sites")
    cfg.n_sites = _atoi(nsites_hdr[0]) if nsites_hdr else 0
    cfg.dim, cfg.grid, cfg.size, cfg.npf, cfg.ncp = dim, grid, size, npf, ncp
    cfg.component_names, cfg.material_names, cfg.interface_names = comps, matnames, ifnames
    cfg.output_names = [str(v) for v in doc.prop("header", "outputs")]
    ndp = ncp - 1

    mats = []
    for mname in matnames:
        mp = f"material/{mname}"
        m = abi.Material()
        mv = doc.prop(mp, "molarvolume")
        if len(mv) != 1:
            raise ValueError(f"{mp}/molarvolume required")
        m.molarvolume = _atof(mv[0])
        kx = doc.prop(mp, "chempot_ex_kineticcoeff")
        m.chempot_ex_kineticcoeff = _atof(kx[0]) if kx else 0.0
        m.nsources, m.source = pools.add_sources(doc, f"{mp}/sources", ncp)
        model = str(doc.prop(mp, "chemicalenergy")[0])
        m.ref = m.unary = m.binary = m.ternary = m.binaryq = m.ternaryq = 0
        m.quad_binary = m.quad_ceq = m.mobility = 0

        def mobility_tables():
            first = len(pools.mob)
            for c in range(ncp):
                cp_ = f"{mp}/mobilityc/{comps[c]}"
                mb = abi.Mobility()
                m0 = doc.prop(cp_, "mobility0")
                if len(m0) != 1:
                    raise ValueError(f"{cp_}/mobility0 required")
                mb.m0 = _atof(m0[0])
                mb.unary = len(pools.ts)
                unary_ts = []
                bin_paths = []
                for cj in range(ncp):
                    unary_ts.append(f"{cp_}/unary_migration/{comps[cj]}")
                    for ci in range(cj + 1, ncp):
                        bin_paths.append(f"{cp_}/binary_migration/{comps[cj]}/{comps[ci]}")
                for p in unary_ts:
                    pools.add_ts(doc, p)
                mb.binary = len(pools.rk)
                for p in bin_paths:
                    pools.add_rk(doc, p)
                pools.mob.append(mb)
            return first

        if model == "none":
            m.model, m.nsites = abi.CHEM_NONE, 0
            m.ref = pools.add_ts(doc, f"{mp}/ref_enthalpy")
        elif model == "quadratic":
            m.model, m.nsites = abi.CHEM_QUADRATIC, 1
            c0 = doc.prop(mp, "c0")
            for c in range(ncp):
                m.c0[c] = _atof(c0[c])
            m.mobility = mobility_tables()
            m.quad_ceq = len(pools.ts)
            for c in range(ncp):
                pools.add_ts(doc, f"{mp}/ceq/{comps[c]}")
            m.ref = pools.add_ts(doc, f"{mp}/ref_enthalpy")
            m.unary = len(pools.ts)
            for c in range(ncp):
                pools.add_ts(doc, f"{mp}/unary_enthalpy/{comps[c]}")
            m.quad_binary = len(pools.ts)
            for c in range(ncp):
                pools.add_ts(doc, f"{mp}/binary_enthalpy/{comps[c]}")
            m.stoichiometry[0] = 1.0
        elif model == "calphaddis":
            m.model, m.nsites = abi.CHEM_CALPHADDIS, 1
            c0 = doc.prop(mp, "c0")
            for c in range(ncp):
                m.c0[c] = _atof(c0[c])
            m.mobility = mobility_tables()
            m.ref = pools.add_ts(doc, f"{mp}/ref_enthalpy")
            m.unary = len(pools.ts)
            for c in range(ncp):
                pools.add_ts(doc, f"{mp}/unary_enthalpy/{comps[c]}", with_log=True)
            m.binary = len(pools.rk)
            for k in range(ncp):
                for j in range(k + 1, ncp):
                    pools.add_rk(doc, f"{mp}/binary_enthalpy/{comps[k]}/{comps[j]}")
            m.ternary = len(pools.rk)
            for k in range(ncp):
                for j in range(k + 1, ncp):
                    for i in range(j + 1, ncp):
                        pools.add_rk(doc, f"{mp}/ternary_enthalpy/{comps[k]}/{comps[j]}/{comps[i]}",
                                     ternary=True)
            m.stoichiometry[0] = 1.0
        elif model == "calphad2sl":
            m.model, m.nsites = abi.CHEM_CALPHAD2SL, 2
            c0p, c0q = doc.prop(mp, "c0_p"), doc.prop(mp, "c0_q")
            for c in range(ncp):
                m.c0[c] = _atof(c0p[c])
                m.c0[ncp + c] = _atof(c0q[c])
            st = doc.prop(mp, "stochiometry")
            if len(st) != 2:
                raise ValueError(f"{mp}/stochiometry needs two entries")
            p, q = _atof(st[0]), _atof(st[1])
            if abs(p + q - 1.0) >= abi.TOL:
                raise ValueError(f"{mp}/stochiometry must sum to 1 (src/init.c:641)")
            m.mobility = mobility_tables()
            m.unary = len(pools.ts)
            for cp in range(ncp):
                for cq in range(ncp):
                    pools.add_ts(doc, f"{mp}/unary_enthalpy/({comps[cp]}:{comps[cq]})", with_log=True)
            m.binary = len(pools.rk)
            for i in range(ncp):
                for j in range(i + 1, ncp):
                    for cq in range(ncp):
                        pools.add_rk(doc, f"{mp}/binary_enthalpy/({comps[i]},{comps[j]}:{comps[cq]})")
            m.binaryq = len(pools.rk)
            for cp in range(ncp):
                for i in range(ncp):
                    for j in range(i + 1, ncp):
                        pools.add_rk(doc, f"{mp}/binary_enthalpy/({comps[cp]}:{comps[i]},{comps[j]})")
            m.ternary = len(pools.rk)
            for i in range(ncp):
                for j in range(i + 1, ncp):
                    for k in range(j + 1, ncp):
                        for cq in range(ncp):
                            pools.add_rk(doc, f"{mp}/ternary_enthalpy/({comps[i]},{comps[j]},{comps[k]}:{comps[cq]})")
            m.ternaryq = len(pools.rk)
            for cp in range(ncp):
                for i in range(ncp):
                    for j in range(i + 1, ncp):
                        for k in range(j + 1, ncp):
                            pools.add_rk(doc, f"{mp}/ternary_enthalpy/({comps[cp]}:{comps[i]},{comps[j]},{comps[k]})")
            m.stoichiometry[0], m.stoichiometry[1] = p, q
        else:
            m.model, m.nsites = abi.CHEM_NONE, 0
        t0 = doc.prop(mp, "temperature0")
        m.temperature0 = _atof(t0[0]) if t0 else 0.0
        m.specific_heat = pools.add_ts(doc, f"{mp}/specific_heat")
        m.latent_heat = pools.add_ts(doc, f"{mp}/latent_heat")
        m.tconductivity = pools.add_ts(doc, f"{mp}/tconductivity")
        mats.append(m)

    ifaces = []
    for iname in ifnames:
        ip = f"interface/{iname}"
        it = abi.Interface()
        it.nsources, it.source = pools.add_sources(doc, f"{ip}/sources", ncp)
        m0 = doc.prop(f"{ip}/mobility", "m0")
        e0 = doc.prop(f"{ip}/energy", "e0")
        if len(m0) != 1 or len(e0) != 1:
            raise ValueError(f"{ip}: mobility/m0 and energy/e0 required")
        it.m0, it.e0 = _atof(m0[0]), _atof(e0[0])
        it.m_act = pools.add_ts(doc, f"{ip}/mobility/activation_energy")
        it.e_act = pools.add_ts(doc, f"{ip}/energy/activation_energy")
        for which in ("mobility", "energy"):
            vals = [_atof(v) for v in doc.prop(f"{ip}/{which}", "anisotropy_values")]
            dirs = [_atof(v) for v in doc.prop(f"{ip}/{which}", "anisotropy_directions")]
            start = len(pools.aniso)
            if vals:
                if len(dirs) != dim * len(vals):
                    raise ValueError(f"{ip}/{which}: anisotropy_directions size")
                for k in range(len(vals)):
                    d = dirs[k * dim:(k + 1) * dim]
                    nrm = math.sqrt(sum(x * x for x in d))
                    pools.aniso.extend([x / nrm for x in d])
                pools.aniso.extend(vals)
            if which == "mobility":
                it.m_naniso, it.m_aniso = len(vals), start
            else:
                it.e_naniso, it.e_aniso = len(vals), start
        pot = doc.prop(ip, "potential")
        if pot:
            it.has_potential = 1
            for c in range(ndp):
                it.potential[c] = _atof(pot[c])
        mc = doc.prop(ip, "mobilityc")
        if mc:
            it.has_mobilityc = 1
            for c in range(ncp):
                it.mobilityc[c] = _atof(mc[c])
        w = doc.prop(ip, "width")
        if len(w) != 1:
            raise ValueError(f"{ip}/width required")
        it.width = _atof(w[0])
        ifaces.append(it)

    bcs = []
    for bname in bcnames:
        bp = f"boundary/{bname}"
        b = abi.BC()
        b.boundary_id = _atoi(doc.prop(bp, "boundary_id")[0])
        b.chem_type = b.thermal_type = abi.BC_NONE
        for which in ("chem", "thermal"):
            tp = doc.prop(f"{bp}/{which}", "type")
            kind = abi.BC_NONE
            if tp:
                kind = {"neumann": abi.BC_NEUMANN, "dirichlet": abi.BC_DIRICHLET}.get(str(tp[0]), abi.BC_NONE)
            vals = doc.prop(f"{bp}/{which}", "value")
            if which == "chem":
                b.chem_type = kind
                if kind != abi.BC_NONE:
                    mask = 0
                    for c, v in enumerate(vals[:ncp]):
                        if str(v).strip() != "*":
                            mask |= 1 << c
                            b.chem_val[c] = _atof(v)
                    b.chem_mask = mask
            else:
                b.thermal_type = kind
                if kind != abi.BC_NONE:
                    b.thermal_val = _atof(vals[0])
        bcs.append(b)

    sp = "solution_parameters"
    interp = str(doc.prop(sp, "interpolation")[0])
    interp_id = {"linear": abi.INTERP_LINEAR, "quadratic": abi.INTERP_QUADRATIC,
                 "cubic": abi.INTERP_CUBIC}.get(interp)
    if interp_id is None:
        raise ValueError("solution_parameters/interpolation must be linear/quadratic/cubic")
    times = [_atof(v) for v in doc.prop(sp, "time")]
    trate = [_atof(v) for v in doc.prop(sp, "temperature_rate")]

    def opt(key, default, conv=_atof):
        v = doc.prop(sp, key)
        return conv(v[0]) if v else default

    cfg.times = times
    cfg.temperature_rates = trate or None
    cfg.timestep0 = opt("timestep0", 1.0e-18)
    cfg.timestepmin = opt("timestepmin", 1.0e-18)
    cfg.timestepmax = opt("timestepmax", times[0])
    cfg.junctionpenalty = opt("junctionpenalty", 1.0)
    cfg.initblocksize = [_atoi(v) for v in doc.prop(sp, "initblocksize")]
    cfg.initrefine = opt("initrefine", 0, _atoi)
    cfg.initcoarsen = opt("initcoarsen", 0, _atoi)
    cfg.maxnrefine = opt("maxnrefine", 0, _atoi)
    cfg.minnrefine = opt("minnrefine", 0, _atoi)
    cfg.amrinterval = opt("amrinterval", 1 << 30, _atoi)
    cfg.outputfreq = opt("outputfreq", 1, _atoi)             # src/init.c:1297-1299
    cfg.outfile = opt("outfile", "output", str)              # src/init.c:1300-1302
    cfg.reltol = opt("reltol", 1.0e-6)
    cfg.abstol = opt("abstol", 1.0e-6)
    cfg.random_seed = opt("random_seed", 0, _atoi)
    cfg.petscoptions = str(doc.prop(sp, "petscoptions")[0]) if doc.prop(sp, "petscoptions") else ""
    cfg.gradient_calculation = opt("gradient_calculation", 0, _atoi)
    cfg.max_phase_neighbours = opt("max_phase_neighbours", abi.MAXAP, _atoi)
    cfg.interpolation = interp_id
    cfg.materials, cfg.interfaces, cfg.bcs = mats, ifaces, bcs
    cfg.periodic = [False] * dim
    # TSRK scheme of the time loop (src/cipher.c:1104 sets 3bs; a YAML
    # petscoptions '-ts_rk_type <t>' overrides it through TSSetFromOptions)
    mo = re.search(r"-ts_rk_type\s+(\S+)", cfg.petscoptions)
    cfg.ts_rk_type = mo.group(1) if mo else "3bs"
    mo = re.search(r"-dm_p4est_brick_periodicity\s+([0-9,]+)", cfg.petscoptions)
    if mo:
        flags = [_atoi(v) for v in mo.group(1).split(",") if v]
        for d in range(min(dim, len(flags))):
            cfg.periodic[d] = bool(flags[d])
    if len(cfg.initblocksize) != dim:
        raise ValueError("solution_parameters/initblocksize must have dim entries")
    for d in range(dim):
        if cfg.initblocksize[d] * (2 ** cfg.initrefine) != grid[d]:
            raise ValueError("initblocksize*2^initrefine must equal grid (src/init.c:1321)")

    # mappings
    pm = parse_rle(doc.prop("mappings", "phase_material_mapping")[0], npf)
    if len(pm) != npf:
        raise ValueError("phase_material_mapping must cover n_phases")
    cfg.phase_material = np.array([v - 1 for v in pm], dtype=np.int32)
    total = int(np.prod(grid))
    vm = parse_rle(doc.prop("mappings", "voxel_phase_mapping")[0], total)
    if len(vm) != total:
        raise ValueError("voxel_phase_mapping must cover the grid")
    cfg.voxel_phase = np.array([v - 1 for v in vm], dtype=np.int32)
    il = parse_rle(doc.prop("mappings", "interface_mapping")[0], npf * npf)
    if len(il) != npf * npf:
        raise ValueError("interface_mapping must cover n_phases^2")
    ilist = np.zeros(npf * npf, dtype=np.int32)
    for j, v in enumerate(il):
        row, col = divmod(j, npf)
        if col > row:
            ilist[j] = (v & 0xFF) - 1           # (unsigned char) interface - 1
    cfg.interface_list = ilist

    s = cfg.struct
    tsa = _arr(abi.TSeries, pools.ts)
    rka = _arr(abi.RK, pools.rk)
    rki = _arr(C.c_int32, pools.rkindex)
    moa = _arr(abi.Mobility, pools.mob)
    maa = _arr(abi.Material, mats)
    ifa = _arr(abi.Interface, ifaces)
    sra = _arr(abi.Source, pools.src)
    bca = _arr(abi.BC, bcs)
    ana = _arr(C.c_double, pools.aniso)
    pma = np.ascontiguousarray(cfg.phase_material)
    ila = np.ascontiguousarray(cfg.interface_list)
    cfg._keep = [tsa, rka, rki, moa, maa, ifa, sra, bca, ana, pma, ila]
    s.tseries, s.rk, s.rkindex, s.mobility = tsa, rka, rki, moa
    s.material, s.interface, s.source, s.bc, s.aniso = maa, ifa, sra, bca, ana
    s.phase_material = pma.ctypes.data_as(C.POINTER(C.c_int32))
    s.interface_list = ila.ctypes.data_as(C.POINTER(C.c_int32))
    s.junctionpenalty = cfg.junctionpenalty
    s.has_temperature_rate = 1 if trate else 0
    s.temperature_rate = trate[loadcase] if trate else 0.0
    s.dim, s.npf, s.ncp, s.nmat, s.nf, s.nbcs = dim, npf, ncp, len(mats), len(ifaces), len(bcs)
    s.ntseries, s.nrk, s.nrkindex = len(pools.ts), len(pools.rk), len(pools.rkindex)
    s.nmobility, s.nsource, s.naniso = len(pools.mob), len(pools.src), len(pools.aniso)
    s.max_phase_neighbours = cfg.max_phase_neighbours
    s.interpolation = interp_id
    s.gradient_calculation = cfg.gradient_calculation
    cfg.pools = pools
    return cfg


def output_fields(cfg, names=None):
    """Output names of the YAML header (src/init.c:258-265) as
    (kind, index) pairs, matched the way the output step matches them
    (src/cipher.c:1006-1076): exact "phaseid", "matid", "interfaceid",
    "temperature"; "<component>_chempot" before "<component>_c" (the first
    '_'-separated token names the component; no match gives index -1, an
    output left at 0); "<id>_phi" takes atoi of the first token."""
    names = cfg.output_names if names is None else names
    comps = cfg.component_names
    out = []
    for name in names:
        tok = name.split("_")[0]
        if name == "phaseid":
            out.append((abi.OUT_PHASEID, 0))
        elif name == "matid":
            out.append((abi.OUT_MATID, 0))
        elif name == "interfaceid":
            out.append((abi.OUT_INTERFACEID, 0))
        elif "_chempot" in name:
            idx = [c for c in range(cfg.ncp - 1) if comps[c] == tok]
            out.append((abi.OUT_CHEMPOT, idx[-1] if idx else -1))
        elif "_c" in name:
            idx = [c for c in range(cfg.ncp) if comps[c] == tok]
            out.append((abi.OUT_COMP, idx[-1] if idx else -1))
        elif "_phi" in name:
            out.append((abi.OUT_PHI, _atoi(tok)))
        elif name == "temperature":
            out.append((abi.OUT_TEMPERATURE, 0))
        else:
            raise ValueError(f"unknown output field {name!r}")
    return out
