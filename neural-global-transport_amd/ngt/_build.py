"""Build libngt_transport.so (all HIP sources, gfx950) in-tree.

Usage: ``python -m ngt._build`` or ``ngt._build.build()``.  The library is
written next to this file so it travels with the source tree.
"""
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(ROOT, "csrc")
INCLUDE = os.path.join(os.path.dirname(ROOT), "include")
LIB_NAME = "libngt_transport.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)

SOURCES = ["ngt_host.hip", "ngt_resample.hip", "ngt_blend.hip", "ngt_raymarch.hip", "ngt_advect.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    # no packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def _finish(proc):
    out, _ = proc.communicate()
    msgs = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    return msgs


def build(verbose=False, out=LIB_PATH, jobs=None):
    """Compile every source to an object in parallel, then link the shared
    library (no relocatable device code: kernels never call across files)."""
    import tempfile
    jobs = jobs or min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    with tempfile.TemporaryDirectory(prefix="ngt_build_") as tmpdir:
        objs, procs = [], []
        for src in SOURCES:
            obj = os.path.join(tmpdir, src.replace(".hip", ".o"))
            objs.append(obj)
            cmd = [hipcc()] + [f for f in HIPCC_FLAGS if f != "-shared"] + [
                "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
            procs.append(_run(cmd, verbose))
            if len(procs) >= jobs:
                _finish(procs.pop(0))
        for p in procs:
            _finish(p)
        tmp = out + ".tmp"
        msgs = _finish(_run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose))
        if verbose and msgs:
            print("\n".join(msgs))
        os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
