"""The C ABI: header, ctypes mirror and the built library agree (no GPU).

No compute entry point is called here; only the host-only layout query and
the error path of context creation on a machine without a HIP device.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

from cipher_amd import abi
from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cipher_amd.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(cipher_[a-z_0-9]+)\(", txt, re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(abi.LIB_PATH):
        from cipher_amd.csrc import build
        build.build()
    return abi.load_library()


def test_header_and_ctypes_declare_the_same_entry_points():
    assert _declared() == sorted(abi.EXPORTED)


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    for name in _declared():
        assert name in syms, name
        assert getattr(lib, name) is not None


def test_struct_layouts_match_the_c_compiler(tmp_path):
    names = ["cipher_tseries", "cipher_rk", "cipher_mobility", "cipher_source", "cipher_material",
             "cipher_interface", "cipher_bc", "cipher_config", "cipher_layout", "cipher_mesh",
             "cipher_ts_params", "cipher_ts_ctrl", "cipher_ts_stats", "cipher_list_shape"]
    mirrors = [abi.TSeries, abi.RK, abi.Mobility, abi.Source, abi.Material, abi.Interface, abi.BC,
               abi.Config, abi.Layout, abi.Mesh, abi.TsParams, abi.TsCtrl, abi.TsStats, abi.ListShape]
    fields = [("cipher_config", "junctionpenalty", abi.Config), ("cipher_config", "gradient_calculation", abi.Config),
              ("cipher_material", "model", abi.Material), ("cipher_material", "source", abi.Material),
              ("cipher_interface", "e_act", abi.Interface), ("cipher_bc", "boundary_id", abi.BC),
              ("cipher_mesh", "ncells", abi.Mesh), ("cipher_tseries", "exp", abi.TSeries),
              ("cipher_ts_params", "use_graph", abi.TsParams), ("cipher_ts_params", "method", abi.TsParams),
              ("cipher_ts_stats", "t", abi.TsStats),
              ("cipher_list_shape", "map", abi.ListShape)]
    src = tmp_path / "sz.c"
    body = "".join(f'printf("%zu\\n", sizeof({n}));' for n in names)
    body += "".join(f'printf("%zu\\n", offsetof({s}, {f}));' for s, f, _ in fields)
    src.write_text(f'#include <stdio.h>\n#include <stddef.h>\n#include "{HEADER}"\nint main(void){{{body}return 0;}}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-o", str(exe), str(src)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(m) for m in mirrors] + [getattr(t, f).offset for _, f, t in fields]
    assert vals == want


def test_constants_match_header():
    txt = open(HEADER).read()
    tol = re.search(r"#define CIPHER_TOL (\S+)", txt).group(1)
    assert float(tol) == abi.TOL
    assert f"#define CIPHER_MAXCP {abi.MAXCP}" in txt and f"#define CIPHER_MAXAP {abi.MAXAP}" in txt
    assert f"#define CIPHER_R_GAS_CONST {abi.R_GAS_CONST}" in txt
    for k, v in abi.STATUS.items():
        assert f"{v} = {k}" in txt


@pytest.mark.parametrize("name", ["gg2d", "ternary", "chem"])
def test_layout_of_matches_host_layout(lib, name):
    wl = getattr(H, name)()
    out = abi.Layout()
    assert lib.cipher_layout_of(C.byref(wl.cfg.struct), C.byref(out)) == 0
    want = wl.layout
    for k, _ in abi.Layout._fields_:
        assert getattr(out, k) == want[k], k


def test_layout_of_rejects_null(lib):
    assert lib.cipher_layout_of(None, None) == 1
    assert lib.cipher_last_error()


def test_context_creation_fails_cleanly_without_a_device(lib):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is visible")
    ctx = C.c_void_p()
    wl = H.gg2d()
    rc = lib.cipher_ctx_create(C.byref(wl.cfg.struct), abi.FP64, C.byref(ctx))
    assert rc != 0 and not ctx.value
    assert lib.cipher_last_error()


def test_runtime_refuses_to_run_without_the_library(monkeypatch, tmp_path):
    monkeypatch.setattr(abi, "_lib", None)
    with pytest.raises(ImportError):
        abi.load_library(str(tmp_path / "missing.so"))


def test_config_validation_accepts_every_workload_and_rejects_bad_rk_blocks(lib):
    """cipher_ctx_create validates the tables before any device call: every
    workload config of the tests passes (without a device the call then fails
    with a HIP error, not an argument error), and a material whose
    Redlich-Kister block runs past the table is refused."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is visible")
    from cipher_amd import config as cfgmod
    from cipher_amd import workload as W
    ctx = C.c_void_p()
    cfgs = [H.chem().cfg, H.chem(thermal=True).cfg, H.ternary().cfg]
    cfgs += [H.ref_case(n).cfg for n in H.REF_CASES]
    for sp in W.ensemble_specs(96)[:12]:
        if sp["components"] > 1:
            cfgs.append(cfgmod.load_config(W.solution_2d_yaml(res=(64, 64), ngrains=10, ncomp=sp["components"],
                                                              ntypes=sp["types"], seed=sp["seed"])))
    for cfg in cfgs:
        rc = lib.cipher_ctx_create(C.byref(cfg.struct), abi.FP64, C.byref(ctx))
        assert rc == 2, (rc, lib.cipher_last_error())
    cfg = H.ternary().cfg
    m = [i for i, mt in enumerate(cfg.materials) if mt.model == abi.CHEM_CALPHAD2SL][0]
    mats = (abi.Material * cfg.struct.nmat)()
    for i in range(cfg.struct.nmat):
        mats[i] = cfg.struct.material[i]
    mats[m].ternaryq = cfg.struct.nrk - 1
    bad = abi.Config()
    C.pointer(bad)[0] = cfg.struct
    bad.material = C.cast(mats, C.POINTER(abi.Material))
    rc = lib.cipher_ctx_create(C.byref(bad), abi.FP64, C.byref(ctx))
    assert rc == abi.ERR_ARG and b"Redlich-Kister" in lib.cipher_last_error()
