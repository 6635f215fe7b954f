"""Build step for the aphpois library (HIP, gfx950) and the CPU checkers.

    python -m aphros_amd.build            # library + oracle (+ reference harness)

Outputs (git-ignored, carried to the GPU box with the tree):
    aphros_amd/libaphpois.so   product: HIP kernels + C ABI (include/aphpois.h)
    oracle/liboracle.so        test-only CPU oracle
    oracle/_ref/ref_driver     test-only harness over the reference's sources
                               (only where the reference checkout exists)
"""
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "aphros_amd", "csrc")
LIB = os.path.join(ROOT, "aphros_amd", "libaphpois.so")
ORACLE_SRC = os.path.join(ROOT, "oracle", "oracle.c")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SRC = "/root/reference/src"

HIP_SOURCES = ["kernels.hip", "engine.hip", "api.hip", "server.hip", "noalloc.hip", "synth.cpp",
               "sysio.cpp"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd))


def build_library():
    tmp = LIB + ".tmp"
    _run([_hipcc(), *HIPCC_FLAGS, *[os.path.join(CSRC, s) for s in HIP_SOURCES], "-o", tmp])
    os.replace(tmp, LIB)


def build_oracle():
    tmp = ORACLE_LIB + ".tmp"
    _run(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c99", ORACLE_SRC, "-lm", "-o", tmp])
    os.replace(tmp, ORACLE_LIB)


def build_reference_harness():
    """Compiles oracle/_ref/ref_driver from the reference's own sources (only
    where the reference checkout exists; the GPU box uses the prebuilt file)."""
    if not os.path.isdir(REF_SRC):
        return False
    script = os.path.join(ROOT, "oracle", "build_ref.sh")
    _run(["bash", script])
    return True


def main():
    build_library()
    build_oracle()
    build_reference_harness()


if __name__ == "__main__":
    sys.exit(main())
