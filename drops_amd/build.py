"""Build recipe of libdhip (gfx950) and of the CPU oracle.

    python -m drops_amd.build          # compiles drops_amd/lib/libdhip.so
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib", "libdhip.so")
OBJ_DIR = os.path.join(HERE, "lib", "obj")
# translation units: the library proper, and the persistent solver kernels in fifteen shares
# (the longest compilations first, so that eight parallel jobs finish together)
UNITS = [(f"dhip_solvers_{k}", os.path.join(CSRC, "dhip_solvers.hip"), [f"-DDHIP_SOLVER_PART={k}"])
         for k in (7, 10, 13)] + [("dhip", os.path.join(CSRC, "dhip.hip"), [])] + [
    (f"dhip_solvers_{k}", os.path.join(CSRC, "dhip_solvers.hip"), [f"-DDHIP_SOLVER_PART={k}"])
    for k in (12, 0, 1, 2, 8, 14, 9, 11, 6, 5, 4, 3)]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # products and sums are rounded one by one, in the reference's order
    "-ffp-contract=off",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _sources_mtime():
    files = [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [os.path.join(ROOT, "include", "dhip.h")]
    return max(os.path.getmtime(f) for f in files)


def build_lib(verbose=False, jobs=8, incremental=False):
    """Compiles every translation unit (in parallel, `jobs` at a time) and links libdhip.so.
    incremental=True skips units whose object is newer than every source and header."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    newest = _sources_mtime()
    procs, pending = [], []
    for name, src, defs in UNITS:
        obj = os.path.join(OBJ_DIR, name + ".o")
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > newest:
            continue
        pending.append([HIPCC] + HIP_FLAGS + defs + ["-c", "-o", obj + ".tmp", src])
    failed = []
    while pending or procs:
        while pending and len(procs) < jobs:
            cmd = pending.pop(0)
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        cmd, p = procs.pop(0)
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append(out)
        else:
            os.replace(cmd[cmd.index("-o") + 1], cmd[cmd.index("-o") + 1][:-4])
    if failed:
        sys.stderr.write("\n".join(failed))
        raise RuntimeError("hipcc failed for libdhip")
    tmp = LIB + ".tmp"
    objs = [os.path.join(OBJ_DIR, name + ".o") for name, _, _ in UNITS]
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed for libdhip")
    os.replace(tmp, LIB)
    return LIB


HOST_EXAMPLES = {"tracefem_p1": os.path.join(HERE, "host", "examples", "tracefem_p1.cpp")}


def build_host_examples(verbose=False):
    """C++ example programs written against the host mirror (drops_amd/host/dropship.hpp), linked to libdhip."""
    out = {}
    for name, src in HOST_EXAMPLES.items():
        exe = os.path.join(HERE, "lib", name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-fopenmp", "-o", exe + ".tmp", src, "-L" + os.path.dirname(LIB), "-ldhip",
               "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"g++ failed for {name}")
        os.replace(exe + ".tmp", exe)
        out[name] = exe
    return out


def build_oracle(verbose=False):
    """The oracle is test infrastructure; it is compiled so that tests and the CPU
    baseline can run, never linked into libdhip."""
    odir = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-C", odir, "-B", "-j2"], capture_output=True, text=True)
    if verbose:
        print(r.stdout, flush=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")


if __name__ == "__main__":
    print(build_lib(verbose=True))
    build_host_examples(verbose=True)
    build_oracle(verbose=True)
