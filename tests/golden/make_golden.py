"""Write tests/golden/legacy_kats.json from the known answers documented in
the reference (read as text; nothing of the reference is run or imported).

Sources (paths relative to the reference tree):
  src/legacy.rs:137-167   R's svd() of [[1,16,49],[4,25,64],[9,36,81]] to 12 digits
  src/legacy.rs:230-232   asserted singular values of the same matrix
  src/legacy.rs:254-283   svdlibrs' full SvdRec dump for that matrix, seed 3141
                          (ut, s, vt and every Diagnostics field)
  src/legacy.rs:2184-2213 [[4,0],[3,-5]]: d = 2, s = 6.3245553203368, 3.1622776601684
  src/legacy.rs:2217-2235 I_3: d = 1, s = 1
Usage: python tests/golden/make_golden.py [/root/reference]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "legacy_kats.json")
NUM = r"-?\d+\.\d+(?:e-?\d+)?"


def lines(path, a, b):
    with open(os.path.join(REF, path)) as f:
        return f.read().splitlines()[a - 1:b]


def nums(text):
    return [float(x) for x in re.findall(NUM, text)]


def main():
    # R output block, legacy.rs:146-161
    blk = lines("src/legacy.rs", 145, 162)
    s_line = [ln for ln in blk if "[1] 123.676578742544" in ln][0]
    r_s = nums(s_line.split("]", 1)[1])
    u_rows = [ln for ln in blk[blk.index([x for x in blk if "$u" in x][0]):] if re.search(r"\[\d,\]", ln)][:3]
    v_start = blk.index([x for x in blk if "$v" in x][0])
    v_rows = [ln for ln in blk[v_start:] if re.search(r"\[\d,\]", ln)][:3]
    r_u = [nums(ln.split("]", 1)[1]) for ln in u_rows]
    r_v = [nums(ln.split("]", 1)[1]) for ln in v_rows]
    # asserted values, legacy.rs:230-232
    asserted = nums("\n".join(lines("src/legacy.rs", 230, 232)))
    asserted = [x for x in asserted if x > 1e-3 and x != 1.0e-12]
    # svdlibrs dump, legacy.rs:256-283
    dump = "\n".join(lines("src/legacy.rs", 256, 283))
    ut_txt = dump[dump.index("ut:"):dump.index("shape=[3, 3]")]
    s_txt = dump[dump.index("s: ["):]
    s_txt = s_txt[:s_txt.index("]")]
    vt_txt = dump[dump.index("vt:"):]
    vt_txt = vt_txt[:vt_txt.index("shape=")]
    ut = nums(ut_txt)
    vt = nums(vt_txt)
    diag = {}
    for key in ("non_zero", "dimensions", "iterations", "lanczos_steps", "ritz_values_stabilized",
                "significant_values", "singular_values", "random_seed"):
        diag[key] = int(re.search(key + r": (\d+)", dump).group(1))
    diag["transposed"] = re.search(r"transposed: (\w+)", dump).group(1) == "true"
    diag["kappa"] = float(re.search(r"kappa: ([0-9.e-]+)", dump).group(1))
    two = nums("\n".join(lines("src/legacy.rs", 2211, 2212)))
    two = [x for x in two if x > 1.0]
    out = {
        "matrix_3x3": [[1.0, 16.0, 49.0], [4.0, 25.0, 64.0], [9.0, 36.0, 81.0]],
        "r_svd_3x3": {"s": r_s, "u": r_u, "v": r_v, "digits": 12, "source": "src/legacy.rs:146-161"},
        "asserted_s_3x3": {"s": asserted, "tol": 1e-12, "source": "src/legacy.rs:230-232"},
        "svdlibrs_3x3_seed3141": {"ut": [ut[0:3], ut[3:6], ut[6:9]], "s": nums(s_txt),
                                  "vt": [vt[0:3], vt[3:6], vt[6:9]], "diagnostics": diag,
                                  "source": "src/legacy.rs:256-283"},
        "basic_2x2": {"matrix": [[4.0, 0.0], [3.0, -5.0]], "d": 2, "s": two, "tol": 1e-12,
                      "source": "src/legacy.rs:2184-2213"},
        "identity_3x3": {"d": 1, "s": [1.0], "tol": 1e-12, "source": "src/legacy.rs:2217-2235"},
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(OUT)


if __name__ == "__main__":
    main()
