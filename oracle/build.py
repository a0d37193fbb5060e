"""Build the CPU oracle (liboracle.so) with gcc.

TEST INFRASTRUCTURE ONLY: the oracle is the checker of the HIP product path.
Usage: python oracle/build.py
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    """Compile oracle.c into oracle/liboracle.so. Raises on failure."""
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared", "-Wall",
           os.path.join(HERE, "oracle.c"), "-o", tmp, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
