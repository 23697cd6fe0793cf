"""Python host mirror of CIPHER's right-hand-side interface over the C ABI.

``Context`` owns a ``cipher_ctx`` (configuration tables + mesh lists in HBM)
and calls the HIP library on torch-allocated device buffers, on torch's
current stream.  ``RHSFunction`` and ``PostStep`` keep the reference's names
and argument meaning (src/cipher.c:20-21): the solution/residual vectors are
device tensors in the per-cell AoS layout.  There is no CPU fallback: the
library must be built and a GPU present.
"""
import ctypes as C

import numpy as np

from . import abi


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class Context:
    def __init__(self, cfg, precision=abi.FP64, device=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("cipher_amd needs a GPU: no HIP device is visible")
        self.torch = torch
        self.lib = abi.load_library()
        self.cfg = cfg
        self.device = torch.device("cuda", 0 if device is None else device)
        torch.cuda.set_device(self.device)
        h = C.c_void_p()
        abi.check(self.lib.cipher_ctx_create(C.byref(cfg.struct), precision, C.byref(h)), "cipher_ctx_create")
        self.h = h
        lay = abi.Layout()
        abi.check(self.lib.cipher_ctx_layout(h, C.byref(lay)), "cipher_ctx_layout")
        self.layout = {f: getattr(lay, f) for f, _ in abi.Layout._fields_}
        self.ndof = self.layout["ndof"]
        self.ncells = 0
        self.precision = precision
        self._stream()

    def _stream(self):
        s = self.torch.cuda.current_stream(self.device).cuda_stream
        abi.check(self.lib.cipher_set_stream(self.h, C.c_void_p(s)), "cipher_set_stream")

    def set_precision(self, precision):
        abi.check(self.lib.cipher_set_precision(self.h, precision), "cipher_set_precision")
        self.precision = precision

    def upload_mesh(self, mesh):
        ms = mesh.struct()
        abi.check(self.lib.cipher_upload_mesh(self.h, C.byref(ms)), "cipher_upload_mesh")
        self.ncells = mesh.ncells
        self.mesh = mesh

    def empty_state(self):
        return self.torch.zeros(self.ncells * self.ndof, dtype=self.torch.float64, device=self.device)

    def _check_state(self, X, what):
        if X.dtype != self.torch.float64 or not X.is_cuda or not X.is_contiguous():
            raise ValueError(f"{what}: expected a contiguous float64 device tensor")
        if X.numel() != self.ncells * self.ndof:
            raise ValueError(f"{what}: expected {self.ncells * self.ndof} doubles, got {X.numel()}")

    def rhs(self, X, F=None, t=0.0, active=None):
        self._check_state(X, "X")
        if F is None:
            F = self.torch.empty_like(X)
        self._check_state(F, "F")
        self._stream()
        if active is None:
            rc = self.lib.cipher_rhs(self.h, t, _ptr(X), _ptr(F))
        else:
            if active.dtype != self.torch.int32 or active.numel() != self.ncells:
                raise ValueError("active: expected int32 tensor of ncells")
            rc = self.lib.cipher_rhs_active(self.h, t, _ptr(X), _ptr(F), _ptr(active))
        abi.check(rc, "cipher_rhs")
        return F

    def poststep(self, X):
        self._check_state(X, "X")
        self._stream()
        abi.check(self.lib.cipher_poststep(self.h, _ptr(X)), "cipher_poststep")
        return X

    def setup_problem(self, phase):
        ph = self.torch.as_tensor(np.ascontiguousarray(phase, dtype=np.int32)).to(self.device)
        X = self.empty_state()
        self._stream()
        abi.check(self.lib.cipher_setup_problem(self.h, _ptr(ph), _ptr(X)), "cipher_setup_problem")
        return X

    def wrms_norm(self, U, Y, atol, rtol):
        out = C.c_double()
        self._stream()
        abi.check(self.lib.cipher_wrms_norm(self.h, _ptr(U), _ptr(Y), U.numel(), atol, rtol, C.byref(out)),
                  "cipher_wrms_norm")
        return out.value

    def axpy_n(self, Y, X, coeffs, Ks):
        nk = len(coeffs)
        a = (C.c_double * max(nk, 1))(*coeffs)
        ks = (C.c_void_p * max(nk, 1))(*[K.data_ptr() for K in Ks])
        self._stream()
        abi.check(self.lib.cipher_axpy_n(self.h, X.numel(), _ptr(Y), _ptr(X), nk, a, ks), "cipher_axpy_n")
        return Y

    def output(self, X, fields=None):
        """Derived fields of the output step (src/cipher.c:1005-1080) as an
        (ncells, nfields) device tensor; fields: (kind, index) pairs or
        output names, default the config's header/outputs."""
        from . import config as cfgmod
        self._check_state(X, "X")
        if fields is None or (fields and isinstance(fields[0], str)):
            fields = cfgmod.output_fields(self.cfg, fields)
        n = len(fields)
        arr = (C.c_int32 * (2 * max(n, 1)))(*[v for kv in fields for v in kv])
        out = self.torch.zeros((self.ncells, n), dtype=self.torch.float64, device=self.device)
        self._stream()
        abi.check(self.lib.cipher_output(self.h, _ptr(X), n, arr, _ptr(out)), "cipher_output")
        return out

    # ---- launch shapes ------------------------------------------------------
    def shapes(self):
        """(lanes, cap, map) of the five per-size lists."""
        out = []
        sh = abi.ListShape()
        for q in range(5):
            abi.check(self.lib.cipher_get_shape(self.h, q, C.byref(sh)), "cipher_get_shape")
            out.append(sh.tuple())
        return out

    def set_shape(self, q, lanes, cap=16384, map=0):
        sh = abi.ListShape(lanes, cap, map, 0)
        abi.check(self.lib.cipher_set_shape(self.h, q, C.byref(sh)), "cipher_set_shape")

    def autotune(self, X, reps=5):
        """Pick the list shapes by timing on X (current precision); returns
        (ms per RHS, shapes)."""
        self._check_state(X, "X")
        F = self.torch.empty_like(X)
        self._stream()
        ms = C.c_double()
        abi.check(self.lib.cipher_autotune(self.h, _ptr(X), _ptr(F), reps, C.byref(ms)), "cipher_autotune")
        return ms.value, self.shapes()

    # ---- adaptive RK (TSRK 3bs + step controller, src/cipher.c:1097-1115) ----
    def ts_setup(self, params):
        self.ts_params = params
        abi.check(self.lib.cipher_ts_setup(self.h, C.byref(params)), "cipher_ts_setup")

    def ts_attempt(self, X, dt):
        self._check_state(X, "X")
        self._stream()
        err = C.c_double()
        abi.check(self.lib.cipher_ts_attempt(self.h, _ptr(X), dt, C.byref(err)), "cipher_ts_attempt")
        return err.value

    def ts_accept(self, X):
        self._check_state(X, "X")
        self._stream()
        abi.check(self.lib.cipher_ts_accept(self.h, _ptr(X)), "cipher_ts_accept")

    def ts_solve(self, X, t0, t_end, dt, max_steps=0):
        """Integrate X in place; returns (stats, next dt)."""
        self._check_state(X, "X")
        self._stream()
        st = abi.TsStats()
        d = C.c_double(dt)
        abi.check(self.lib.cipher_ts_solve(self.h, _ptr(X), t0, t_end, C.byref(d), max_steps, C.byref(st)),
                  "cipher_ts_solve")
        return st, d.value

    def solve_loadcases(self, X, times=None, rates=(), t0=0.0):
        """CIPHER's load-case loop (src/cipher.c:1589-1606) on X in place:
        times and rates default to the configuration's solution_parameters
        time and temperature_rate lists (rates None: no temperature rate).
        Returns one TsStats per load case."""
        self._check_state(X, "X")
        times = list(self.cfg.times if times is None else times)
        if rates == ():
            rates = self.cfg.temperature_rates
        if rates is not None and len(rates) != len(times):
            raise ValueError("one temperature rate per load case")
        n = len(times)
        ta = (C.c_double * max(n, 1))(*times)
        ra = (C.c_double * max(n, 1))(*rates) if rates is not None else None
        st = (abi.TsStats * max(n, 1))()
        self._stream()
        abi.check(self.lib.cipher_ts_loadcases(self.h, _ptr(X), n, ta, ra, t0, st), "cipher_ts_loadcases")
        return list(st)[:n]

    def close(self):
        if getattr(self, "h", None):
            self.lib.cipher_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many contexts run through shared launches (``cipher_batch``).  Batch
    state tensors hold the members' records back to back; ``offsets[i]`` is
    member i's first double, ``cell_offsets[i]`` its first cell.  The members
    must stay alive and unused on their own while the batch exists."""

    def __init__(self, members, precision=abi.FP64):
        import torch
        self.torch = torch
        self.lib = abi.load_library()
        self.members = list(members)
        n = len(self.members)
        hs = (C.c_void_p * n)(*[m.h.value for m in self.members])
        h = C.c_void_p()
        abi.check(self.lib.cipher_batch_create(hs, n, C.byref(h)), "cipher_batch_create")
        self.h = h
        self.device = self.members[0].device
        off = (C.c_int64 * (n + 1))()
        coff = (C.c_int64 * (n + 1))()
        abi.check(self.lib.cipher_batch_offsets(h, off, coff), "cipher_batch_offsets")
        self.offsets = list(off)
        self.cell_offsets = list(coff)
        self.set_precision(precision)

    def __len__(self):
        return len(self.members)

    def _stream(self):
        s = self.torch.cuda.current_stream(self.device).cuda_stream
        abi.check(self.lib.cipher_batch_set_stream(self.h, C.c_void_p(s)), "cipher_batch_set_stream")

    def set_precision(self, precision):
        abi.check(self.lib.cipher_batch_set_precision(self.h, precision), "cipher_batch_set_precision")
        self.precision = precision

    def _check(self, X, what):
        if X.dtype != self.torch.float64 or not X.is_cuda or not X.is_contiguous():
            raise ValueError(f"{what}: expected a contiguous float64 device tensor")
        if X.numel() != self.offsets[-1]:
            raise ValueError(f"{what}: expected {self.offsets[-1]} doubles, got {X.numel()}")

    def empty_state(self):
        return self.torch.zeros(self.offsets[-1], dtype=self.torch.float64, device=self.device)

    def slice(self, X, i):
        return X[self.offsets[i]:self.offsets[i + 1]]

    def rhs(self, X, F=None, t=0.0, active=None):
        self._check(X, "X")
        if F is None:
            F = self.torch.empty_like(X)
        self._check(F, "F")
        self._stream()
        if active is None:
            rc = self.lib.cipher_batch_rhs(self.h, t, _ptr(X), _ptr(F))
        else:
            if active.dtype != self.torch.int32 or active.numel() != self.cell_offsets[-1]:
                raise ValueError("active: expected int32 tensor of the batch's cells")
            rc = self.lib.cipher_batch_rhs_active(self.h, t, _ptr(X), _ptr(F), _ptr(active))
        abi.check(rc, "cipher_batch_rhs")
        return F

    def poststep(self, X):
        self._check(X, "X")
        self._stream()
        bad = C.c_int32(-1)
        rc = self.lib.cipher_batch_poststep(self.h, _ptr(X), C.byref(bad))
        if rc:
            err = abi.CipherError(rc, "cipher_batch_poststep")
            err.member = bad.value
            raise err
        return X

    def wrms_norms(self, U, Y, atol, rtol):
        n = len(self.members)
        out = (C.c_double * n)()
        self._stream()
        abi.check(self.lib.cipher_batch_wrms_norm(self.h, _ptr(U), _ptr(Y), atol, rtol, out),
                  "cipher_batch_wrms_norm")
        return list(out)

    def autotune(self, X, reps=5):
        """Pick launch shapes by timing on X; returns (ms per batched RHS,
        [(components, lane capacity, list, lanes, cap, map), ...])."""
        self._check(X, "X")
        F = self.torch.empty_like(X)
        self._stream()
        ms = C.c_double()
        abi.check(self.lib.cipher_batch_autotune(self.h, _ptr(X), _ptr(F), reps, C.byref(ms)),
                  "cipher_batch_autotune")
        return ms.value, self.shapes()

    def shapes(self):
        """[(components, lane capacity, list, lanes, cap, map), ...]"""
        out = []
        k = 0
        ncp, lanes, sh = C.c_int32(), C.c_int32(), abi.ListShape()
        while self.lib.cipher_batch_shape(self.h, k, 0, C.byref(ncp), C.byref(lanes), C.byref(sh)) == 0:
            nl = 5 if lanes.value >= 16 else (4 if lanes.value >= 8 else 3)
            for q in range(nl):
                abi.check(self.lib.cipher_batch_shape(self.h, k, q, C.byref(ncp), C.byref(lanes), C.byref(sh)),
                          "cipher_batch_shape")
                out.append((ncp.value, lanes.value, q) + sh.tuple())
            k += 1
        return out

    def ts_setup(self, params):
        self.ts_params = params
        abi.check(self.lib.cipher_batch_ts_setup(self.h, C.byref(params)), "cipher_batch_ts_setup")

    def ts_solve(self, X, t0, t_end, dt, max_steps=0):
        """Integrate every member in place; t0, t_end, dt are per-member
        sequences.  Returns (list of stats, list of next dt)."""
        self._check(X, "X")
        self._stream()
        n = len(self.members)
        a0 = (C.c_double * n)(*t0)
        a1 = (C.c_double * n)(*t_end)
        d = (C.c_double * n)(*dt)
        st = (abi.TsStats * n)()
        abi.check(self.lib.cipher_batch_ts_solve(self.h, _ptr(X), a0, a1, d, max_steps, st), "cipher_batch_ts_solve")
        return list(st), list(d)

    def set_temperature_rate(self, member, rate):
        """Member's load-case temperature rate (None: no rate)."""
        abi.check(self.lib.cipher_batch_set_temperature_rate(self.h, member, 0 if rate is None else 1,
                                                             0.0 if rate is None else float(rate)),
                  "cipher_batch_set_temperature_rate")

    def solve_loadcases(self, X, times, rates=None, t0=0.0):
        """Every member through its load cases at once, each as
        cipher_ts_loadcases does on its own context (src/cipher.c:1589-1606):
        times[i][k] and rates[i][k] (rates None: no temperature rate) are
        member i's case lengths and rates, all members with the same number
        of cases; case k restarts every member at the integrator's dt_min
        and ends it at its running sum of case lengths.  Returns, per case,
        the list of per-member TsStats."""
        n = len(self.members)
        if len(times) != n or (rates is not None and len(rates) != n):
            raise ValueError("one list of case lengths (and rates) per member")
        ncase = len(times[0])
        if any(len(tm) != ncase for tm in times) or (rates is not None and any(len(r) != ncase for r in rates)):
            raise ValueError("every member needs the same number of load cases")
        if any(tau < 0.0 for tm in times for tau in tm):
            raise ValueError("load-case time must be >= 0")
        if getattr(self, "ts_params", None) is None:
            raise RuntimeError("Batch.ts_setup must be called before solve_loadcases")
        dt_min = self.ts_params.dt_min
        start = [t0] * n
        cur = [t0] * n
        out = []
        for k in range(ncase):
            for i in range(n):
                self.set_temperature_rate(i, None if rates is None else rates[i][k])
            end = [start[i] + times[i][k] for i in range(n)]
            stats, _ = self.ts_solve(X, cur, end, [dt_min] * n)
            cur = [st.t if cur[i] < end[i] else cur[i] for i, st in enumerate(stats)]
            out.append(stats)
            start = end
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.cipher_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def RHSFunction(ts, ftime, X, F, user):
    """Mirror of RHSFunction(TS,PetscReal,Vec,Vec,void*) (src/cipher.c:27):
    user is a Context; X and F are device state tensors."""
    user.rhs(X, F, t=ftime)
    return 0


def ts_adapt(params, ctrl, dt, err):
    """The library's step-size decision (host only): (accept, dt_next)."""
    lib = abi.load_library()
    acc = C.c_int()
    nxt = C.c_double()
    abi.check(lib.cipher_ts_adapt(C.byref(params), C.byref(ctrl), dt, err, C.byref(acc), C.byref(nxt)),
              "cipher_ts_adapt")
    return bool(acc.value), nxt.value


def PostStep(ts):
    """Mirror of PostStep(TS) (src/cipher.c:686), active-set block: ts must
    expose .solution (device tensor) and .user (Context)."""
    ts.user.poststep(ts.solution)
    return 0
