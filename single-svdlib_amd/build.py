"""Build the svdh shared library (libsvdh.so) for gfx950 with hipcc.

Usage: python single-svdlib_amd/build.py [--jobs N]

Objects go to single-svdlib_amd/_build/, the library to
single-svdlib_amd/libsvdh.so (in-tree, so it travels with the source tree).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsvdh.so")
SOURCES = ["api.cpp", "hostla.cpp", "las2.cpp", "rsvd.cpp", "adjoint.cpp", "sparse.hip", "dense.hip", "ingest.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _source_hash():
    import importlib.util
    spec = importlib.util.spec_from_file_location("svdh_srchash", os.path.join(HERE, "single_svdlib", "srchash.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.source_hash()


def _compile(src, defs=()):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *defs, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(jobs=None, verbose=False):
    """Compile every source and link libsvdh.so. Raises on failure."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    # api.cpp carries the hash of the sources as svdh_build_id()
    bid = [f'-DSVDH_BUILD_ID="{_source_hash()}"']
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, bid if s == "api.cpp" else ()), SOURCES))
    errors = []
    for src, _obj, rc, out in results:
        # the host pass does not know the packed-fp32 feature flag; that warning is expected
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if rc != 0:
            errors.append(f"--- {src} ---\n" + "\n".join(lines))
        elif verbose and lines:
            print(f"--- {src} ---\n" + "\n".join(lines))
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [r[1] for r in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_synth()
    build_probe()
    return LIB


PROBE_LIB = os.path.join(HERE, "libsvdh_probe.so")


def build_probe():
    """The SpMM gather-ceiling probe (measurement only, used by bench.py)."""
    tmp = PROBE_LIB + ".tmp"
    cmd = [HIPCC, *COMMON, "-shared", "-x", "hip", os.path.join(CSRC, "probe_gather.hip"), "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed (probe):\n" + r.stdout + r.stderr)
    os.replace(tmp, PROBE_LIB)
    return PROBE_LIB


SYNTH_LIB = os.path.join(HERE, "libsynth.so")


def build_synth():
    """The workload generator (host C, gcc)."""
    tmp = SYNTH_LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-D_DEFAULT_SOURCE", "-fopenmp", "-fPIC", "-shared", "-Wall",
           os.path.join(HERE, "synth", "synth.c"), "-o", tmp, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("gcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


if __name__ == "__main__":
    j = None
    if "--jobs" in sys.argv:
        j = int(sys.argv[sys.argv.index("--jobs") + 1])
    print(build(jobs=j, verbose=True))
