"""Prints the 64-entry table of the library's logarithm (glog in gh_device.h; the oracle's
ref_log holds the same literals). x = 2^e m1 with m1 in [1, 2); i = the top six mantissa
bits of x. Entries 0..31 cover m = m1 in [1, 1.5), entries 32..63 cover m = m1 / 2 in
[0.75, 1): c_i = 1 + (i + 1/2) / 64 (halved for i >= 32), except that the two entries
next to 1 (i = 0 and i = 63) take c = 1 exactly, so that log x near 1 is log1p of an exact
r = m - 1. Entry i holds rcp_i = the float nearest 1 / c_i and lg_i = the float nearest
-log(rcp_i). Usage: python gen_log_table.py"""
import numpy as np


def table():
    out = []
    for i in range(64):
        c = 1.0 + (i + 0.5) / 64.0
        if i >= 32:
            c *= 0.5
        if i in (0, 63):
            c = 1.0
        rcp = np.float32(1.0 / c)
        lg = np.float32(-np.log(np.float64(rcp)))
        out.append((rcp, lg))
    return out


def literal(v):
    # C hex-float literal of an FP32 value (exact)
    h = float(np.float32(v)).hex()  # e.g. 0x1.f81f820000000p-1
    mant, exp = h.split("p")
    mant = mant.rstrip("0").rstrip(".")
    return f"{mant}p{exp}f"


if __name__ == "__main__":
    rows = [f"{{{literal(r)}, {literal(g)}}}" for r, g in table()]
    for k in range(0, 64, 4):
        print("    " + ", ".join(rows[k:k + 4]) + ("," if k + 4 < 64 else ""))
