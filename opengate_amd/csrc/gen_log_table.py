"""Prints the 64-entry table of the library's logarithm (glog in gh_device.h; the oracle's
ref_log holds the same literals). x = 2^e m with m in [0.5, 1) (frexp); i = the top six
mantissa bits of x, so entry i covers m in [1/2 + i/128, 1/2 + (i+1)/128) and takes
c_i = 1/2 + (i + 1/2)/128, the middle of that interval, except the last entry (next to
1), which takes c = 1 exactly, so that log x just below 1 is log1p of an exact r = m - 1.
Entry i holds rcp_i = the float nearest 1 / c_i and lg_i = the float nearest -log(rcp_i).
Usage: python gen_log_table.py"""
import numpy as np


def table():
    out = []
    for i in range(64):
        c = 0.5 + (i + 0.5) / 128.0
        if i == 63:
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
