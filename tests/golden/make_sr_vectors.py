"""Writes sr_vectors.npz: input/output vectors of gfx950's stochastic
conversions (v_cvt_sr_f16_f32, v_cvt_sr_bf16_f32) recorded on an MI355X.

sr_probe.hip (beside this script) converts 357 crafted (x, r) pairs (17
random words for each of 21 values: exact, half-way, subnormal, near the f16
maximum, beyond it, zeros) and 200000 pseudo-random pairs (exponents spread
over the f16 normal, subnormal and underflow ranges) into a binary file.
Built and run on the GPU:

    hipcc --offload-arch=gfx950 -O2 tests/golden/sr_probe.hip -o sr_probe && ./sr_probe srprobe.bin

This script keeps the crafted pairs and the first 4096 random ones:

    python tests/golden/make_sr_vectors.py srprobe.bin
"""

import os
import sys

import numpy as np


def main(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    a = np.frombuffer(raw[4:], np.uint32)
    x = a[:n].view(np.float32)
    r = a[n : 2 * n]
    f16 = (a[2 * n : 3 * n] & 0xFFFF).astype(np.uint16)
    bf16 = (a[3 * n : 4 * n] & 0xFFFF).astype(np.uint16)
    keep = np.r_[0 : 357 + 4096]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sr_vectors.npz")
    np.savez_compressed(out, x=x[keep], r=r[keep], f16=f16[keep], bf16=bf16[keep])
    print(f"wrote {out}: {len(keep)} vectors")


if __name__ == "__main__":
    main(sys.argv[1])
