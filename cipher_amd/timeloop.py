"""CIPHER's outer time loop with remeshing (src/cipher.c:1589-1830), on the
host around the device integrator.

For each load case (solution_parameters time / temperature_rate) the loop
restarts from timestepmin and calls the integrator in chunks that end at the
next multiple of amrinterval accepted steps or at the case's end time (exact
final time).  After a chunk that ends on a multiple of amrinterval it
remeshes: every leaf with more than one active phase is labelled for
refinement, every other one for coarsening (src/cipher.c:1616-1629), the
forest is adapted and 2:1 balanced (``mesh.adapt``), the state is transferred
(``mesh.transfer``: copies into refined leaves, means into coarsened ones),
the mesh is flattened and uploaded, and the integrator is set up afresh (the
reference destroys and re-creates its TS, src/cipher.c:1826-1827); an adapt
that changes nothing keeps everything (DMAdaptLabel returns no forest).
Nucleation sites, which also mark cells for refinement in the reference, are
out of scope.  With an output directory, chunks also end at every
outputfreq-th accepted step, where the output step of PostStep
(src/cipher.c:976-1091) writes "<outfile>_<step>.vtu" (``vtu.write``) with
the header's output names computed on the device (cipher_output).

The state moves through host memory at a remesh (the forest and its
flattening are host data structures, as p4est's are).
"""
import ctypes as C

import numpy as np

from . import abi
from . import mesh as meshmod


def refine_flags(X, layout):
    """Cells labelled for refinement: round(AS count) > 1 (src/cipher.c:1622-1628)."""
    D = layout["ndof"]
    return np.rint(np.asarray(X).reshape(-1, D)[:, layout["as_offset"]]) > 1


def adapt_state(forest, X, layout, minlevel, maxlevel):
    """One remesh on host arrays: (new forest, new X) or (forest, X) when the
    adapt changes nothing."""
    D = layout["ndof"]
    Xv = np.asarray(X).reshape(-1, D)
    new = meshmod.adapt(forest, refine_flags(Xv, layout), minlevel, maxlevel)
    if new.ncells == forest.ncells and np.array_equal(new.level, forest.level) and \
            np.array_equal(new.anchor, forest.anchor):
        return forest, X
    return new, meshmod.transfer(forest, new, Xv).reshape(-1)


class Run:
    """State of a run: the forest, its flattened mesh, the device state and
    per-chunk statistics."""

    def __init__(self, forest, mesh, X):
        self.forest = forest
        self.mesh = mesh
        self.X = X
        self.chunks = []        # (load case, steps, rejected, t, dt, ncells)
        self.outputs = []       # VTU files written
        self.remeshes = 0
        self.nsteps = 0
        self.t = 0.0


def write_output(ctx, cfg, forest, X, step, output_dir):
    """The output step of PostStep (src/cipher.c:976-1091): the header's
    output names computed on the device (cipher_output) and written to
    "<outfile>_<step>.vtu" in output_dir."""
    import os
    from . import vtu
    names = list(cfg.output_names)
    out = ctx.output(X, names).cpu().numpy() if names else np.zeros((ctx.ncells, 0))
    path = os.path.join(output_dir, f"{cfg.outfile}_{step}.vtu")
    vtu.write(path, forest, {nm: out[:, i] for i, nm in enumerate(names)})
    return path


def solve(ctx, cfg, forest, X, params, t0=0.0, times=None, rates=(), amrinterval=None, on_remesh=None,
          output_dir=None):
    """The loop of src/cipher.c:1589-1606 with the remesh of :1607-1830 on a
    Context whose mesh is forest's flattening and whose integrator is set up
    with params (params.dt_min is timestepmin).  X (device tensor) is
    replaced at each remesh; returns a Run holding the final forest, mesh
    and state.  times / rates / amrinterval default to the configuration's.
    With output_dir, chunks also end at every outputfreq-th step, where the
    output step writes its VTU file (before a remesh at the same step, as
    PostStep runs inside the solve); Run.outputs lists the files.  The
    configuration's TSRK type (petscoptions -ts_rk_type, 3bs by default)
    must be params.method (abi.rk_method): integrating a configuration with
    another scheme would change its trajectory and step counts without
    notice."""
    method = abi.rk_method(getattr(cfg, "ts_rk_type", "3bs"))
    if params is None or params.method != method:
        raise abi.CipherError(abi.ERR_ARG, f"integrator method {getattr(params, 'method', None)} does not match "
                              f"petscoptions -ts_rk_type {getattr(cfg, 'ts_rk_type', '3bs')} (method {method})")
    times = list(cfg.times if times is None else times)
    if rates == ():
        rates = cfg.temperature_rates
    if rates is not None and len(rates) != len(times):
        raise ValueError("one temperature rate per load case")
    interval = int(cfg.amrinterval if amrinterval is None else amrinterval)
    if interval < 1:
        raise ValueError("amrinterval must be >= 1")
    ofreq = int(cfg.outputfreq) if output_dir is not None else 0
    if output_dir is not None and ofreq < 1:
        raise ValueError("outputfreq must be >= 1")
    torch = ctx.torch
    run = Run(forest, ctx.mesh, X)
    current = loadcase_time = t0
    nsteps = 0
    for k, tau in enumerate(times):
        abi.check(ctx.lib.cipher_set_temperature_rate(ctx.h, 0 if rates is None else 1,
                                                      C.c_double(0.0 if rates is None else rates[k])),
                  "cipher_set_temperature_rate")
        dt = params.dt_min
        end = loadcase_time + tau
        while current < end:
            chunk = interval - nsteps % interval
            if ofreq:
                chunk = min(chunk, ofreq - nsteps % ofreq)
            st, dt = ctx.ts_solve(run.X, current, end, dt, max_steps=chunk)
            if st.steps == 0 and st.t == current:
                raise RuntimeError(f"integrator made no progress at t={current}")
            current = st.t
            nsteps += st.steps
            run.chunks.append((k, st.steps, st.rejected, st.t, dt, ctx.ncells))
            if ofreq and nsteps % ofreq == 0:
                run.outputs.append(write_output(ctx, cfg, run.forest, run.X, nsteps, output_dir))
            if nsteps % interval == 0:
                torch.cuda.synchronize()
                Xh = run.X.cpu().numpy()
                new, Xn = adapt_state(run.forest, Xh, cfg.layout(), cfg.minnrefine, cfg.maxnrefine)
                if new is not run.forest:
                    run.forest = new
                    run.mesh = new.flatten()
                    ctx.upload_mesh(run.mesh)
                    ctx.ts_setup(params)
                    run.X = torch.as_tensor(np.ascontiguousarray(Xn)).to(ctx.device)
                    run.remeshes += 1
                    if on_remesh is not None:
                        on_remesh(run)
        loadcase_time += tau
    run.nsteps = nsteps
    run.t = current
    return run
