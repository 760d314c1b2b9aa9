"""Builds the C++ drop-in programs (tests/cpp/*.cpp) against the facade headers and
libmrmd_hip.so into mrmd_amd/facade/bin/."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
BIN = os.path.join(HERE, "bin")
SRC = os.path.join(ROOT, "tests", "cpp")


def build():
    os.makedirs(BIN, exist_ok=True)
    for f in sorted(os.listdir(SRC)):
        if not f.endswith(".cpp"):
            continue
        out = os.path.join(BIN, f[:-4])
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Wno-unused-parameter",
               f"-I{os.path.join(HERE, 'include')}", f"-I{os.path.join(ROOT, 'include')}", f"-I{SRC}",
               os.path.join(SRC, f), "-o", out, f"-L{PKG}", "-lmrmd_hip", "-Wl,-rpath,$ORIGIN/../.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"building {f} failed")


if __name__ == "__main__":
    build()
