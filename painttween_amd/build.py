"""Build the gfx950 HIP library and the CPU oracle in place.

``build()`` always recompiles:
  painttween_amd/libpainttween.so  -- HIP kernels + host runtime + multi-job executor (hipcc, gfx950)
  oracle/libptoracle.so           -- CPU oracle (gcc, OpenMP), test infrastructure
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpainttween.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "libptoracle.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_SOURCES = ["pt_kernels.hip", "pt_runtime.hip", "pt_engine.hip"]

# -ffp-contract=off: only the explicit fmaf calls fuse, so the device runs the
# same IEEE operation sequence as the CPU build of the algorithm.
# -packed-fp32-ops is disabled: packed FP32 forms are generated by nothing here.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
ORACLE_FLAGS = ["-std=c11", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                "-mfma", "-fPIC", "-shared", "-Wall", "-Wextra"]


def _run(cmd):
    proc = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return proc.stdout


def build_hip():
    """Compiles the translation units in parallel, then links them."""
    inc = ["-I", CSRC, "-I", os.path.join(ROOT, "include")]
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    objs, procs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        procs.append(subprocess.Popen([HIPCC] + cflags + inc + ["-c", os.path.join(CSRC, src), "-o", obj],
                                      cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    failed = []
    for src, p in zip(HIP_SOURCES, procs):
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(src)
    if failed:
        raise RuntimeError("build step failed: " + ", ".join(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    for obj in objs:
        os.remove(obj)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    tmp = ORACLE_LIB + ".tmp"
    _run(["gcc"] + ORACLE_FLAGS + ["-o", tmp, os.path.join(ORACLE_DIR, "pt_oracle.c"), "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def build():
    build_oracle()
    build_hip()


if __name__ == "__main__":
    build()
    print("built", LIB, ORACLE_LIB)
