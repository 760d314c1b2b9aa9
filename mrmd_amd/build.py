"""Build recipe of the gfx950 library (libmrmd_hip.so) and the C++ facade library.

Every HIP translation unit is compiled by hipcc for gfx950 only, with FMA contraction off
(-ffp-contract=off: every expression is evaluated as written, the convention that makes
binning, neighbour tests, integrators and the noise bit-reproducible against the CPU
restatement) and without packed-FP32 instructions.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmrmd_hip.so")
FACADE_LIB = os.path.join(HERE, "libmrmd.so")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ, src.replace(".hip", ".o"))
    cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build_hip(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for src, obj, rc, out in ex.map(_compile, sources()):
            msgs = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if rc != 0:
                sys.stderr.write("\n".join(msgs) + "\n")
                raise RuntimeError(f"hipcc failed on {src}")
            if verbose and msgs:
                sys.stderr.write("\n".join(msgs) + "\n")
            objs.append(obj)
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build(verbose=False):
    build_hip(verbose=verbose)
    facade = os.path.join(HERE, "facade", "build.py")
    if os.path.exists(facade):
        subprocess.check_call([sys.executable, facade])


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
