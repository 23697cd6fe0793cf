"""Build libcipher_amd.so for gfx950 with hipcc (one object per source).

FP contraction is off so the FP64 kernels follow the reference's rounding
sequence; packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
SOURCES = ["rhs_nc1.hip", "rhs_nc2.hip", "rhs_nc3.hip", "rhs_nc4.hip", "rhs_nc5.hip",
           "rhs_batch_nc1.hip", "rhs_batch_nc2.hip", "rhs_batch_nc3.hip", "rhs_batch_nc4.hip", "rhs_batch_nc5.hip",
           "rhs.hip", "post.hip", "ts.hip", "api.hip", "integrator.hip", "batch.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-variable", "-Wno-unused-function"]


def _stale(src, obj):
    if not os.path.exists(obj):
        return True
    deps = [src] + [os.path.join(HERE, h) for h in os.listdir(HERE) if h.endswith(".h")]
    deps.append(os.path.join(os.path.dirname(PKG), "include", "cipher_amd.h"))
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose=False, jobs=8, force=False):
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for s in SOURCES:
        src = os.path.join(HERE, s)
        obj = os.path.join(objdir, s + ".o")
        objs.append(obj)
        if not force and not _stale(src, obj):
            continue
        cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    out = os.path.join(PKG, "libcipher_amd.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, out)
    return out


def _wait(item):
    name, p = item
    out, _ = p.communicate()
    txt = out.decode(errors="replace")
    if p.returncode:
        raise RuntimeError(f"hipcc {name} failed:\n{txt}")
    warn = [l for l in txt.splitlines() if "warning" in l and "packed-fp32" not in l]
    if warn:
        sys.stderr.write(f"[{name}]\n" + "\n".join(warn[:20]) + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, force="-f" in sys.argv))
