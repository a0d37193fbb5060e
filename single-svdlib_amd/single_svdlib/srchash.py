"""SHA-256 of the sources libsvdh.so is built from (csrc/*.hip, *.cpp, *.h
and include/svdh.h). build.py compiles it into the library
(svdh_build_id); the loader refuses a library whose id differs from the
sources beside it, so a stale build cannot run silently."""
import hashlib
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROOT = os.path.dirname(PKG_DIR)


def source_files():
    csrc = os.path.join(PKG_DIR, "csrc")
    names = sorted(f for f in os.listdir(csrc) if f.endswith((".hip", ".cpp", ".h")))
    return [os.path.join(csrc, f) for f in names] + [os.path.join(ROOT, "include", "svdh.h")]


def source_hash():
    """Hex digest over (file name, contents) of every source, in a fixed order;
    None when the sources are not present."""
    h = hashlib.sha256()
    try:
        for p in source_files():
            with open(p, "rb") as f:
                data = f.read()
            h.update(os.path.basename(p).encode() + b"\0" + len(data).to_bytes(8, "little") + data)
    except OSError:
        return None
    return h.hexdigest()
