"""Build recipe of libdhip (gfx950) and of the CPU oracle.

    python -m drops_amd.build          # compiles drops_amd/lib/libdhip.so
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = [os.path.join(HERE, "csrc", "dhip.hip")]
LIB = os.path.join(HERE, "lib", "libdhip.so")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # products and sums are rounded one by one, in the reference's order
    "-ffp-contract=off",
    # no packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_lib(verbose=False):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + HIP_FLAGS + ["-o", tmp] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libdhip")
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
