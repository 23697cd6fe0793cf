"""Build libcipher_amd.so for gfx950 with hipcc (one object per source).

FP contraction is off so the FP64 kernels follow the reference's rounding
sequence; packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
# (source, defines): the right-hand-side kernels are instantiated per
# component count and lane capacity in translation units of their own
SOURCES = [(src, {"CIPHER_NC": nc, "CIPHER_G": g})
           for src in ("rhs_inst.hip", "rhs_batch_inst.hip") for nc in (1, 2, 3, 4, 5) for g in (16, 8, 4, 0)]
SOURCES += [(s, {}) for s in ("rhs.hip", "post.hip", "ts.hip", "api.hip", "integrator.hip", "batch.hip")]
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
    import concurrent.futures as cf
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    todo = []
    objs = []
    for s, defs in SOURCES:
        src = os.path.join(HERE, s)
        tag = "".join(f"_{k.split('_')[-1].lower()}{v}" for k, v in defs.items())
        obj = os.path.join(objdir, s + tag + ".o")
        objs.append(obj)
        if not force and not _stale(src, obj):
            continue
        cmd = [HIPCC] + FLAGS + [f"-D{k}={v}" for k, v in defs.items()] + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        todo.append((s + tag, cmd))
    # a new compile starts as soon as any finishes (the instantiation units
    # differ several-fold in compile time)
    with cf.ThreadPoolExecutor(max_workers=jobs) as pool:
        for name, r in zip([t[0] for t in todo], pool.map(
                lambda t: subprocess.run(t[1], stdout=subprocess.PIPE, stderr=subprocess.STDOUT), todo)):
            _report(name, r)
    out = os.path.join(PKG, "libcipher_amd.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, out)
    return out


def _report(name, r):
    txt = r.stdout.decode(errors="replace")
    if r.returncode:
        raise RuntimeError(f"hipcc {name} failed:\n{txt}")
    warn = [l for l in txt.splitlines() if "warning" in l and "packed-fp32" not in l]
    if warn:
        sys.stderr.write(f"[{name}]\n" + "\n".join(warn[:20]) + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, force="-f" in sys.argv))
