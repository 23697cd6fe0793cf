"""Benchmark: RHS cell-updates/s and HBM roofline of the phase-field RHS.

One "step" is one evaluation of the right-hand side over every leaf cell of
the workload, inputs resident in HBM.  Default workload: the batched
ensemble of 96 independent 2D instances (side log-uniform 64..1024 with AMR,
10..2000 Voronoi grains, 1..4 components, 1..3 phase types, capacities
4/8/16) run through one batch; the measured value is the batched FP32-mixed
right-hand side, with the batched FP64 one as its control in the same line
and the FP32-mixed error against FP64 per field; whole adaptive steps of
every member (four right-hand sides, stage combinations, error norm, the
device step controller and PostStep, --whole-steps of them from each
member's timestepmin) are timed separately.  Other workloads:
--workload ternary (3D AMR ternary precipitation, 128^3, levels 4-7),
gg3d (3D AMR grain growth, 64^3, 500 grains), gg2d (128^2 periodic, 20
grains).

States: the SetUpProblem initial condition (chemical potential uniform
from the matrix composition) advanced by --evolve explicit phase-field steps
on the GPU with PostStep after each; a step whose active-set union outgrows
the capacity is undone and the evolution stops there (the reference asserts
at that point), and the number of steps taken is reported.

The CPU baseline is the C oracle (the reference's algorithm, FP64) on the
same states, run on parallel threads over the host's cores (the oracle
holds no global state; ctypes releases the interpreter lock), for a bounded
sample of about --cpu-seconds.

--counters runs only set-up, the autotuner, warm-up and the timed calls of
one precision (for rocprofv3 --pmc runs, which count every dispatch): a
small torch kernel marks the start of the timed region.

Prints one JSON line.  Run: python bench.py [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def build_workload(name):
    from cipher_amd import workload as W
    if name == "gg3d":
        txt = W.grain_growth_yaml(res=(64, 64, 64), ngrains=500, periodic=False, energy_range=(0.8, 1.2),
                                  ntypes=255, minlevel=3, t_end=50.0)
        return W.Workload(txt, "grain_growth_3d_amr_64"), {
            "workload": "grain_growth_3d_amr_64", "domain": "64^3", "amr_levels": "3-6",
            "grains": 500, "capacity": 8, "components": 1}
    if name == "gg2d":
        txt = W.grain_growth_yaml(res=(128, 128), ngrains=20, periodic=True)
        return W.Workload(txt, "grain_growth_2d_128"), {
            "workload": "grain_growth_2d_128", "domain": "128^2 periodic", "grains": 20, "capacity": 8}
    if name == "ternary":
        txt = W.ternary_ppt_yaml(res=(128, 128, 128), ngrains=100, nppt=40, radius=3, minlevel=4)
        return W.Workload(txt, "ternary_gb_precipitation_128"), {
            "workload": "ternary_gb_precipitation_128", "domain": "128^3", "amr_levels": "4-7",
            "grains": 100, "precipitates": 40, "components": 3, "capacity": 8, "T": 700.0}
    raise SystemExit(f"unknown workload {name}")


def algorithmic_bytes(ncells, nfaces, ndof):
    """B_RHS = N_c*(16*D + 8) + 8*N_f: read X and write F once (8 B/dof each),
    one FP64 cell size per cell, two int32 cell ids per interior face."""
    return ncells * (16 * ndof + 8) + 8 * nfaces


def copy_probe(nbytes, reps=5):
    """A plain device-to-device copy of known size: what FETCH_SIZE and
    WRITE_SIZE read for it calibrates the counters against bytes moved."""
    import torch
    n = nbytes // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8, "reps": reps}))


def measured_traffic(workload, key="fp64"):
    """HBM bytes per right-hand side of the timed calls (FETCH_SIZE +
    WRITE_SIZE over every kernel of one RHS, calibrated on a plain copy),
    from the committed counter summary of this workload's --counters runs
    (profiles/traffic_<workload>.md or profiles/k_rhs_<workload>.md), or
    None."""
    import re
    for name in (f"traffic_{workload}.md", f"k_rhs_{workload}.md"):
        path = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(path):
            continue
        m = re.search(rf"^traffic_bytes_per_rhs_{key}\s*=\s*([0-9.eE+]+)", open(path).read(), re.M)
        if m:
            return float(m.group(1))
    return None


def host_threads():
    """Threads for the CPU baseline: the cores this process may run on, at
    most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_threads(jobs, seconds, threads):
    """Run oracle right-hand sides on `threads` threads until `seconds` have
    passed (each thread takes the next job, round robin); jobs are
    (callable, cells).  Returns (cells/s, cells, calls, wall)."""
    import itertools
    import threading
    nxt = itertools.count()
    lock = threading.Lock()
    done = [0, 0]
    t0 = time.perf_counter()

    def worker():
        while time.perf_counter() - t0 < seconds:
            with lock:
                k = next(nxt)
            fn, cells = jobs[k % len(jobs)]
            fn()
            with lock:
                done[0] += cells
                done[1] += 1

    th = [threading.Thread(target=worker) for _ in range(threads)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    return done[0] / wall, done[0], done[1], wall


def timed_region_marker():
    """A tiny torch kernel that marks, in a counter trace, where the timed
    right-hand sides start."""
    import torch
    torch.ones(1, device="cuda").mul_(2.0)
    torch.cuda.synchronize()


def prepare_state(ctx, wl, steps):
    """SetUpProblem state (mu uniform from the matrix composition) whose
    phase fields are advanced by `steps` explicit steps on the GPU; returns
    (X, steps done, histogram of active-set sizes)."""
    import torch
    from cipher_amd import abi
    from cipher_amd import workload as W
    L = wl.layout
    D = L["ndof"]
    P = L["pf_size"]
    X = ctx.setup_problem(wl.phase)
    if L["dp_size"]:
        Xh = X.cpu().numpy()
        W.uniform_mu_override(Xh, L, wl.phase, wl.cfg.phase_material)
        X = torch.as_tensor(Xh).to(ctx.device)
    F = torch.empty_like(X)
    pf = slice(L["pf_offset"], L["pf_offset"] + P)
    chem = slice(L["dp_offset"], L["tm_offset"])
    good = X.clone()
    evolved = 0
    for _ in range(steps):
        ctx.rhs(X, F)
        # phase fields only: an explicit step sized for the phase fields is far
        # beyond the stability limit of the stiff chemical-potential dynamics,
        # so mu and the explicit potentials keep their initial values
        F.view(-1, D)[:, chem] = 0.0
        m = F.view(-1, D)[:, pf].abs().max().item()
        dt = 0.02 / m if m > 0 else 0.0
        X.add_(F, alpha=dt)
        try:
            ctx.poststep(X)
        except abi.CipherError as e:
            if e.code != abi.ERR_CAPACITY:
                raise
            # the union outgrew the capacity (the reference asserts here):
            # keep the last state whose active sets fit
            X.copy_(good)
            break
        good.copy_(X)
        evolved += 1
    del good
    torch.cuda.synchronize()
    nact = np.bincount(np.rint(X.view(-1, D)[:, 0].cpu().numpy()).astype(int), minlength=P + 1)
    return X, evolved, nact


def trajectory_fp32_vs_fp64(ctx, wl, X0, steps):
    """The FP32-mixed adaptive trajectory against the FP64 one from the same
    state: FP64 runs `steps` accepted steps to a time t1, FP32-mixed runs to
    the same t1 (exact final time); reported: accepted/rejected step counts
    and the max and RMS differences of phase fraction (where both runs hold
    the phase), composition (output fields) and chemical potential."""
    import torch
    from cipher_amd import abi
    L = wl.layout
    D = L["ndof"]
    nc = wl.cfg.ncp
    params = abi.ts_params(1e-4, 1e-4, 1e-12, 1e-3, use_graph=True)
    res = {}
    for prec in (abi.FP64, abi.FP32_MIXED):
        ctx.set_precision(prec)
        ctx.ts_setup(params)
        X = X0.clone()
        if prec == abi.FP64:
            st, _ = ctx.ts_solve(X, 0.0, 1e-3, 1e-12, max_steps=steps)
            t1 = st.t
        else:
            st, _ = ctx.ts_solve(X, 0.0, t1, 1e-12)
        comp = ctx.output(X, [(abi.OUT_COMP, c) for c in range(nc)]).cpu().numpy()
        res[prec] = (st, X.cpu().numpy().reshape(-1, D), comp)
    ctx.set_precision(abi.FP64)
    s64, X64, c64 = res[abi.FP64]
    s32, X32, c32 = res[abi.FP32_MIXED]
    pf = slice(L["pf_offset"], L["pf_offset"] + L["pf_size"])
    n64 = np.rint(X64[:, 0]).astype(int)
    same = np.all(X64[:, :L["as_size"]] == X32[:, :L["as_size"]], axis=1)
    slot = np.arange(L["pf_size"])[None, :] < n64[:, None]
    dphi = np.abs(X32[:, pf] - X64[:, pf])[same[:, None] & slot]
    dc = np.abs(c32 - c64).ravel()
    dmu = np.abs(X32[:, L["dp_offset"]:L["dp_offset"] + L["dp_size"]] - X64[:, L["dp_offset"]:L["dp_offset"] + L["dp_size"]])
    mus = np.abs(X64[:, L["dp_offset"]:L["dp_offset"] + L["dp_size"]]).max()

    def mr(a):
        return {"max": float(a.max()) if a.size else 0.0, "rms": float(np.sqrt(np.mean(a * a))) if a.size else 0.0}

    return {"t_end": t1, "fp64_steps": [s64.steps, s64.rejected], "fp32_mixed_steps": [s32.steps, s32.rejected],
            "accepted_step_difference": int(s32.steps - s64.steps),
            "cells_with_equal_active_sets": float(same.mean()),
            "phase_fraction_abs": mr(dphi), "composition_abs": mr(dc),
            "chemical_potential_rel": {k: v / mus for k, v in mr(dmu.ravel()).items()}}


def rel_errors(A, B, L):
    """Normwise relative error of A against B per field (max|A-B| / max|B|)."""
    pf = slice(L["pf_offset"], L["pf_offset"] + L["pf_size"])
    rel = {}
    for name, sl in (("phase_fraction", pf), ("chemical_potential", slice(L["dp_offset"], L["dp_offset"] + L["dp_size"])),
                     ("explicit_potential", slice(L["ex_offset"], L["ex_offset"] + L["ex_size"]))):
        if A[:, sl].size:
            s = np.abs(B[:, sl]).max()
            rel[name] = float(np.abs(A[:, sl] - B[:, sl]).max() / s) if s > 0 else float(np.abs(A[:, sl]).max())
    return rel


def whole_steps(args, b, ctxs, wls):
    """Whole adaptive steps of every member (four right-hand sides, stage
    combinations, error norm, the device step controller and PostStep),
    each member from its SetUpProblem state and its own timestepmin (the
    evolved states sit at the edge of their capacity), timed separately from
    the right-hand side alone.  A member whose active sets outgrow its
    capacity within the steps (the reference asserts there) is left out and
    the rest run again in a batch of their own, autotuned like the first."""
    import re
    import torch
    from cipher_amd import abi
    from cipher_amd.runtime import Batch
    X0 = [prepare_state(c, w, 0)[0] for c, w in zip(ctxs, wls)]
    members = list(range(len(wls)))
    left_out = []
    out = {}
    bw = b
    for _ in range(8):
        Xw = torch.cat([X0[i] for i in members])
        if bw is not b:
            bw.autotune(Xw, reps=2)
        n = len(members)
        dt0 = [wls[i].cfg.timestepmin for i in members]
        failed = None
        for prec in (abi.FP32_MIXED, abi.FP64):
            key = "fp32_mixed" if prec else "fp64"
            bw.set_precision(prec)
            bw.ts_setup(abi.ts_params(1e-4, 1e-4, 1e-16, 1.0, use_graph=True))
            Xs = Xw.clone()
            try:
                bw.ts_solve(Xs, [0.0] * n, [1e30] * n, dt0, max_steps=1)    # capture and warm-up
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                stats, _ = bw.ts_solve(Xs, [0.0] * n, [1e30] * n, dt0, max_steps=args.whole_steps)
                torch.cuda.synchronize()
                el = time.perf_counter() - t0
            except abi.CipherError as e:
                m = re.search(r"in member (\d+)", str(e))
                if e.code != abi.ERR_CAPACITY or not m:
                    raise
                failed = members[int(m.group(1))]
                break
            steps = max(s.steps for s in stats)
            out[key] = {"ms_per_step": 1e3 * el / max(steps, 1), "steps_per_member": steps,
                        "attempts_max": max(s.steps + s.rejected for s in stats), "rhs_per_attempt": 4}
            print(f"[bench] whole steps {key}: {steps} steps per member, {1e3 * el / max(steps, 1):.3f} ms per step",
                  file=sys.stderr, flush=True)
        if failed is None:
            break
        print(f"[bench] whole steps: member {failed} ({wls[failed].name}) outgrows its capacity, left out",
              file=sys.stderr, flush=True)
        left_out.append(failed)
        members.remove(failed)
        out = {}
        if bw is not b:
            bw.close()
        bw = Batch([ctxs[i] for i in members])
    if bw is not b:
        bw.close()
    out["members"] = len(members)
    out["cells"] = int(sum(wls[i].ncells for i in members))
    out["left_out"] = [wls[i].name for i in left_out]
    return out


def run_ensemble(args):
    """Batched ensemble: `--members` independent 2D instances (side
    log-uniform 64..1024 with AMR, 10..2000 grains, 1..4 components, 1..3
    phase types) in one batch; the batched FP32-mixed right-hand side is the
    measured value, the batched FP64 one its control, and the same members
    run one context at a time show what batching gains."""
    import torch
    from cipher_amd import abi
    from cipher_amd import workload as W
    from cipher_amd.runtime import Batch, Context

    t0 = time.time()
    specs = W.ensemble_specs(n=args.members)
    wls, ctxs, states = [], [], []
    nact_tot = np.zeros(17, dtype=np.int64)
    evolved = []
    for i, sp in enumerate(specs):
        wl = W.ensemble_member(sp)
        ctx = Context(wl.cfg, precision=abi.FP64)
        ctx.upload_mesh(wl.mesh)
        X, ev, nact = prepare_state(ctx, wl, args.evolve)
        nact_tot[:nact.size] += nact
        evolved.append(ev)
        wls.append(wl)
        ctxs.append(ctx)
        states.append(X)
        if i % 8 == 7 or i == len(specs) - 1:
            print(f"[bench] {i + 1}/{len(specs)} members ready ({time.time() - t0:.0f}s)", file=sys.stderr, flush=True)
    X = torch.cat(states)
    del states
    b = Batch(ctxs)
    F = torch.empty_like(X)
    N = sum(w.ncells for w in wls)
    bytes_alg = sum(algorithmic_bytes(w.ncells, w.mesh.nfaces, w.layout["ndof"]) for w in wls)
    stream = torch.cuda.current_stream()

    def timed(fn, steps):
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(steps):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / steps

    modes = (abi.FP32_MIXED, abi.FP64)
    if args.counters:
        modes = (abi.FP32_MIXED,) if args.counters == "fp32" else (abi.FP64,)
    res = {}
    for prec in modes:
        b.set_precision(prec)
        tune_ms, shapes = b.autotune(X, reps=3) if not args.no_tune else (0.0, b.shapes())
        for _ in range(args.warmup):
            b.rhs(X, F)
        torch.cuda.synchronize()
        if args.counters:
            timed_region_marker()
        ms = timed(lambda: b.rhs(X, F), args.steps)
        res[prec] = (ms, None if args.counters else F.clone(), shapes)
        print(f"[bench] batched {'fp32-mixed' if prec else 'fp64'}: {ms:.3f} ms/RHS "
              f"(autotuned {tune_ms:.3f})", file=sys.stderr, flush=True)
    if args.counters:
        print(json.dumps({"counters_run": "batched_ensemble_2d", "precision": args.counters, "steps": args.steps,
                          "ms_per_step": res[modes[0]][0], "shapes": res[modes[0]][2],
                          "algorithmic_bytes_per_rhs": bytes_alg}))
        return
    # the same members one context at a time (FP64, default launch shapes)
    views = [(c, b.slice(X, i), b.slice(F, i)) for i, c in enumerate(ctxs)]
    for c in ctxs:
        c.set_precision(abi.FP64)
    for _ in range(max(1, args.warmup // 4)):
        for c, x, f in views:
            c.rhs(x, f)
    ms_seq = timed(lambda: [c.rhs(x, f) for c, x, f in views], max(1, args.steps // 4))
    print(f"[bench] one context at a time fp64: {ms_seq:.3f} ms per ensemble RHS", file=sys.stderr, flush=True)
    whole = whole_steps(args, b, ctxs, wls)
    ms32, F32, shapes32 = res[abi.FP32_MIXED]
    ms64, F64, shapes64 = res[abi.FP64]
    A, B = F32.cpu().numpy(), F64.cpu().numpy()
    # FP32-mixed against FP64 per field and member: max and RMS of the
    # difference over the largest FP64 value of the field (normwise); the bar
    # of the metric is on the phase fractions' rate
    per_field = {}
    worst = {}
    for i, w in enumerate(wls):
        D = w.layout["ndof"]
        sl = slice(b.offsets[i], b.offsets[i + 1])
        for k, v in field_error_stats(A[sl].reshape(-1, D), B[sl].reshape(-1, D), w.layout).items():
            cur = per_field.setdefault(k, {"max": 0.0, "rms_max": 0.0})
            if v["max"] >= cur["max"]:
                worst[k] = w.name
            cur["max"] = max(cur["max"], v["max"])
            cur["rms_max"] = max(cur["rms_max"], v["rms"])
    # conditioning of each member's state: the FP64 residual's change under
    # random one-ulp perturbations of the state (active sets kept); a field
    # that moves by s under 2^-53 input changes moves by ~s 2^29 under FP32
    # rounding, so its FP32-mixed bar is max(1e-6, 4 s 2^29) (the phase
    # fractions' rate keeps the metric's 1e-6)
    gen = torch.Generator(device=X.device).manual_seed(0)
    sign = torch.randint(0, 2, X.shape, device=X.device, generator=gen, dtype=torch.int8).to(torch.float64) * 2 - 1
    Xp = X * (1.0 + 2.0 ** -52 * sign)
    del sign
    for i, w in enumerate(wls):
        D = w.layout["ndof"]
        nas = w.layout["as_size"]
        b.slice(Xp, i).view(-1, D)[:, :nas] = b.slice(X, i).view(-1, D)[:, :nas]
    b.set_precision(abi.FP64)
    Bp = b.rhs(Xp).cpu().numpy()
    del Xp
    ratio = {}
    for i, w in enumerate(wls):
        D = w.layout["ndof"]
        sl = slice(b.offsets[i], b.offsets[i + 1])
        e32 = field_error_stats(A[sl].reshape(-1, D), B[sl].reshape(-1, D), w.layout)
        e64 = field_error_stats(Bp[sl].reshape(-1, D), B[sl].reshape(-1, D), w.layout)
        for k, v in e32.items():
            bar = 1e-6 if k == "PF" else max(1e-6, 4.0 * e64[k]["max"] * 2.0 ** 29)
            ratio[k] = max(ratio.get(k, 0.0), v["max"] / bar)
    del Bp
    flags_equal = True
    act32 = torch.zeros(N, dtype=torch.int32, device=X.device)
    act64 = torch.zeros(N, dtype=torch.int32, device=X.device)
    b.set_precision(abi.FP32_MIXED)
    b.rhs(X, F, active=act32)
    b.set_precision(abi.FP64)
    b.rhs(X, F, active=act64)
    flags_equal = bool(torch.equal(act32, act64))

    cpu = None
    if not args.no_cpu:
        import oracle
        thr = host_threads()
        jobs = []
        for i, w in enumerate(wls):
            Xi = X[b.offsets[i]:b.offsets[i + 1]].cpu().numpy()
            jobs.append((lambda w=w, Xi=Xi: oracle.rhs(w.cfg, w.mesh, Xi), w.ncells))
        rate, cells, calls, el = cpu_baseline_threads(jobs, args.cpu_seconds, thr)
        cpu = {"value": rate, "unit": "cell-updates/s", "cores": thr, "kind": "port",
               "sample": f"{calls} oracle RHS calls over the members in order, round robin on {thr} threads "
                         f"({cells} cell-updates, {el:.1f} s wall)"}

    rate32 = N / (ms32 * 1e-3)
    rate64 = N / (ms64 * 1e-3)
    gbs32 = bytes_alg / (ms32 * 1e-3) / 1e9
    gbs64 = bytes_alg / (ms64 * 1e-3) / 1e9
    comps = [s["components"] for s in specs]
    out = {
        "metric": "RHS cell-updates/s + %HBM peak: FP32-mixed vs FP64, <=1e-6 rel err",
        "value": rate32,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed (f64 state, thermodynamics and active-set decisions)",
        "data": "synthetic (seeded Voronoi microstructures, reference YAML schema)",
        "config": {"workload": "batched_ensemble_2d", "members": len(specs), "cells": N,
                   "sides": [s["side"] for s in specs], "components_hist": np.bincount(comps, minlength=5)[1:].tolist(),
                   "capacities": sorted(set(s["capacity"] for s in specs)),
                   "classes": len({(sh[0], sh[1]) for sh in shapes32}),
                   "evolved_steps": [min(evolved), max(evolved)], "active_set_sizes": nact_tot.tolist(),
                   "parallelism": "1 GPU, one batch"},
        "roofline": {"bound": "hbm", "achieved": gbs32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs32 / HBM_PEAK_GBS, "traffic": measured_traffic("batched_ensemble_2d", "fp32"),
                     "algorithmic_bytes_per_launch": bytes_alg},
        "fp64_control": {"value": rate64, "ms_per_step": ms64, "achieved_gbs": gbs64, "frac": gbs64 / HBM_PEAK_GBS,
                         "traffic": measured_traffic("batched_ensemble_2d", "fp64"), "fp32_speedup": ms64 / ms32},
        "fp32_vs_fp64": {"normwise_rel_err_per_field": per_field, "worst_member": worst,
                         "worst_err_over_bar": ratio, "all_fields_within_bar": all(v <= 1.0 for v in ratio.values()),
                         "active_flags_equal": flags_equal,
                         "phase_fraction_rate_within_1e-6": per_field.get("PF", {}).get("max", 0.0) <= 1e-6},
        "one_context_at_a_time_fp64": {"ms_per_step": ms_seq, "batching_speedup_fp64": ms_seq / ms64},
        "whole_step": whole,
        "autotuned_shapes_fp32": shapes32,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def field_error_stats(A, B, L):
    """Per field: max and RMS of |A - B| over max|B| (entries of B's active
    slots and unused slots alike; both are 0 where a slot is unused)."""
    out = {}
    for name, sl in (("PF", slice(L["pf_offset"], L["pf_offset"] + L["pf_size"])),
                     ("DP", slice(L["dp_offset"], L["dp_offset"] + L["dp_size"])),
                     ("EX", slice(L["ex_offset"], L["ex_offset"] + L["ex_size"])),
                     ("TM", slice(L["tm_offset"], L["tm_offset"] + 1))):
        a, bb = A[:, sl], B[:, sl]
        if a.size == 0:
            continue
        scale = np.abs(bb).max()
        d = np.abs(a - bb)
        if scale == 0:
            out[name] = {"max": float(d.max()), "rms": float(np.sqrt(np.mean(d * d)))}
        else:
            out[name] = {"max": float(d.max() / scale), "rms": float(np.sqrt(np.mean(d * d)) / scale)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="ensemble", help="ensemble (default), ternary, gg3d or gg2d")
    ap.add_argument("--members", type=int, default=96, help="instances of the ensemble workload")
    ap.add_argument("--evolve", type=int, default=20)
    ap.add_argument("--whole-steps", type=int, default=10,
                    help="adaptive steps of every ensemble member timed whole (stages, norm, controller, PostStep)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes instead of autotuned ones")
    ap.add_argument("--traj-steps", type=int, default=20,
                    help="accepted steps of the FP32-mixed vs FP64 trajectory comparison (workloads with chemistry)")
    ap.add_argument("--counters", choices=("fp32", "fp64"), default=None,
                    help="only set-up, autotuner, warm-up and the timed calls of one precision (counter runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by Morton ranges with a one-cell halo; see the design notes)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "ensemble":
        run_ensemble(args)
        return

    import torch
    from cipher_amd import abi
    from cipher_amd import workload as W
    from cipher_amd.runtime import Context

    t0 = time.time()
    wl, cfgdesc = build_workload(args.workload)
    L = wl.layout
    D = L["ndof"]
    P = L["pf_size"]
    N, NF = wl.ncells, wl.mesh.nfaces
    print(f"[bench] workload {wl.name}: {N} cells, {NF} faces, D={D}, built in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)

    ctx = Context(wl.cfg, precision=abi.FP64)
    ctx.upload_mesh(wl.mesh)
    X, evolved, nact = prepare_state(ctx, wl, args.evolve)
    F = torch.empty_like(X)
    pf = slice(L["pf_offset"], L["pf_offset"] + P)
    print(f"[bench] state ready after {evolved} steps: active-set sizes {nact.tolist()}", file=sys.stderr, flush=True)

    shapes = {}

    def time_mode(precision):
        ctx.set_precision(precision)
        if not args.no_tune:
            # launch shapes picked by the library's autotuner on this state
            # (untimed set-up, as a user would run it once per mesh)
            shapes[precision] = ctx.autotune(X, reps=5)[1]
        for _ in range(args.warmup):
            ctx.rhs(X, F)
        torch.cuda.synchronize()
        if args.counters:
            timed_region_marker()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        stream = torch.cuda.current_stream()
        ev0.record(stream)
        t_wall = time.perf_counter()
        for _ in range(args.steps):
            ctx.rhs(X, F)
        ev1.record(stream)
        torch.cuda.synchronize()
        t_wall = time.perf_counter() - t_wall
        ms = ev0.elapsed_time(ev1) / args.steps
        return ms, t_wall, None if args.counters else F.clone()

    bytes_alg = algorithmic_bytes(N, NF, D)
    if args.counters:
        prec = abi.FP32_MIXED if args.counters == "fp32" else abi.FP64
        ms, _, _ = time_mode(prec)
        print(json.dumps({"counters_run": wl.name, "precision": args.counters, "steps": args.steps,
                          "ms_per_step": ms, "shapes": shapes.get(prec), "algorithmic_bytes_per_rhs": bytes_alg}))
        return
    ms64, wall64, F64 = time_mode(abi.FP64)
    print(f"[bench] fp64: {ms64:.4f} ms/RHS", file=sys.stderr, flush=True)
    ms32, wall32, F32 = time_mode(abi.FP32_MIXED)
    print(f"[bench] fp32-mixed: {ms32:.4f} ms/RHS", file=sys.stderr, flush=True)
    ctx.set_precision(abi.FP64)

    # FP32-mixed error against FP64, per field (normwise relative)
    rel = rel_errors(F32.view(-1, D).cpu().numpy(), F64.view(-1, D).cpu().numpy(), L)
    traj = trajectory_fp32_vs_fp64(ctx, wl, X, args.traj_steps) if L["dp_size"] and args.traj_steps > 0 else None

    rate64 = N / (ms64 * 1e-3)
    rate32 = N / (ms32 * 1e-3)
    gbs64 = bytes_alg / (ms64 * 1e-3) / 1e9
    gbs32 = bytes_alg / (ms32 * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        import oracle
        Xh = X.cpu().numpy()
        thr = host_threads()
        # every thread evaluates whole right-hand sides of the same state
        # (the aggregate rate of the host's cores; a partition with halos
        # would add exchange work on top)
        rate, cells, calls, el = cpu_baseline_threads([(lambda: oracle.rhs(wl.cfg, wl.mesh, Xh), N)],
                                                      args.cpu_seconds, thr)
        cpu = {"value": rate, "unit": "cell-updates/s", "cores": thr, "kind": "port",
               "sample": f"{calls} oracle RHS calls on the same {N}-cell state on {thr} threads ({el:.1f} s wall)"}

    # configurations 1 and 2 are the FP64 drop-in; the ternary workload
    # (configurations 3/4) is measured in FP32-mixed with FP64 beside it
    mixed = args.workload == "ternary"
    out = {
        "metric": "RHS cell-updates/s + %HBM peak: FP32-mixed vs FP64, <=1e-6 rel err",
        "value": rate32 if mixed else rate64,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32 if mixed else ms64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed (f64 state, thermodynamics and active-set decisions)" if mixed else "f64",
        "data": "synthetic (seeded Voronoi microstructure, reference YAML schema)",
        "config": dict(cfgdesc, cells=N, faces=NF, dofs_per_cell=D, evolved_steps=evolved,
                       active_set_sizes=nact.tolist(), parallelism="1 GPU"),
        "roofline": {"bound": "hbm", "achieved": gbs32 if mixed else gbs64, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (gbs32 if mixed else gbs64) / HBM_PEAK_GBS,
                     "traffic": measured_traffic(wl.name, "fp32" if mixed else "fp64"),
                     "algorithmic_bytes_per_launch": bytes_alg},
        "fp64": {"value": rate64, "ms_per_step": ms64, "achieved_gbs": gbs64, "frac": gbs64 / HBM_PEAK_GBS,
                 "traffic": measured_traffic(wl.name, "fp64")},
        "fp32_mixed": {"value": rate32, "ms_per_step": ms32, "achieved_gbs": gbs32,
                       "frac": gbs32 / HBM_PEAK_GBS, "speedup_vs_fp64": ms64 / ms32, "rel_err_vs_fp64": rel},
        "trajectory_fp32_vs_fp64": traj,
        "list_shapes": {"autotuned": not args.no_tune,
                        "fp64": shapes.get(abi.FP64), "fp32_mixed": shapes.get(abi.FP32_MIXED)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
