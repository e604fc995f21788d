"""Writes philox_vectors.json: Philox4x32-10 blocks computed by an independent
implementation, ATen's at::Philox4_32 (PhiloxRNGEngine.h, shipped as a source
header with torch), compiled for the host.  The engine with seed s,
subsequence q and offset o returns philox(ctr={o_lo, o_hi, q_lo, q_hi},
key={s_lo, s_hi}); inputs are drawn from numpy with a fixed seed.

    python tests/golden/make_philox_vectors.py
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(os.path.dirname(torch.__file__), "include")

SRC = r"""
#include <ATen/core/PhiloxRNGEngine.h>
#include <cstdio>
#include <cstdint>
int main() {
    unsigned long long s, q, o;
    while (std::scanf("%llu %llu %llu", &s, &q, &o) == 3) {
        at::Philox4_32 e(s, q, o);
        uint32_t a = e(), b = e(), c = e(), d = e();
        std::printf("%u %u %u %u\n", a, b, c, d);
    }
}
"""


def main():
    rng = np.random.default_rng(20261203)
    n = 256
    seeds = rng.integers(0, 2**64, n, dtype=np.uint64)
    subs = rng.integers(0, 2**64, n, dtype=np.uint64)
    offs = rng.integers(0, 2**64, n, dtype=np.uint64)
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "p.cpp")
        exe = os.path.join(td, "p")
        open(src, "w").write(SRC)
        subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", INC, src, "-o", exe])
        inp = "".join(f"{int(s)} {int(q)} {int(o)}\n" for s, q, o in zip(seeds, subs, offs))
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    rows = [list(map(int, l.split())) for l in out.strip().splitlines()]
    vec = []
    for s, q, o, r in zip(seeds, subs, offs, rows):
        s, q, o = int(s), int(q), int(o)
        ctr = [o & 0xFFFFFFFF, o >> 32, q & 0xFFFFFFFF, q >> 32]
        key = [s & 0xFFFFFFFF, s >> 32]
        vec.append({"ctr": ctr, "key": key, "out": r})
    with open(os.path.join(HERE, "philox_vectors.json"), "w") as f:
        json.dump({"source": "ATen at::Philox4_32 (torch " + torch.__version__ + ")",
                   "vectors": vec}, f, indent=0)


if __name__ == "__main__":
    main()
