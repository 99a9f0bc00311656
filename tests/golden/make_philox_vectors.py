"""Write tests/golden/philox_vectors.json (run in the build container only).

Two independent sources pin the Philox4x32-10 restatement in oracle/philox.py:

1. The published Random123 known-answer vectors for philox4x32 with 10
   rounds (Salmon et al., SC'11; kat_vectors of the Random123 distribution).
2. Vectors from PyTorch's own Philox4_32_10 engine
   (torch/include/ATen/core/PhiloxRNGEngine.h, an independent
   implementation), compiled here with g++ from that header. With
   ``philox_engine(seed, subsequence, offset)`` the first four outputs are
   the block for key = seed and counter = (offset_lo, offset_hi,
   subsequence_lo, subsequence_hi).

Usage: python tests/golden/make_philox_vectors.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

RANDOM123_KAT = [
    # counter (4 words), key (2 words), expected output (4 words)
    {"ctr": ["00000000", "00000000", "00000000", "00000000"], "key": ["00000000", "00000000"],
     "out": ["6627e8d5", "e169c58d", "bc57ac4c", "9b00dbd8"]},
    {"ctr": ["ffffffff", "ffffffff", "ffffffff", "ffffffff"], "key": ["ffffffff", "ffffffff"],
     "out": ["408f276d", "41c83b0e", "a20bc7c6", "6d5451fd"]},
    {"ctr": ["243f6a88", "85a308d3", "13198a2e", "03707344"], "key": ["a4093822", "299f31d0"],
     "out": ["d16cfe09", "94fdcceb", "5001e420", "24126ea1"]},
]

CPP = r"""
#include <cstdio>
#include <cstdint>
#include <ATen/core/PhiloxRNGEngine.h>
int main(int argc, char** argv) {
  uint64_t seed, subseq, offset;
  while (std::scanf("%lu %lu %lu", &seed, &subseq, &offset) == 3) {
    at::Philox4_32 eng(seed, subseq, 0);
    eng.incr_n(offset);
    uint32_t a = eng(), b = eng(), c = eng(), d = eng();
    std::printf("%08x %08x %08x %08x\n", a, b, c, d);
  }
  return 0;
}
"""


def torch_vectors(cases):
    import torch

    inc = os.path.join(os.path.dirname(torch.__file__), "include")
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "p.cpp")
        exe = os.path.join(td, "p")
        with open(src, "w") as f:
            f.write(CPP)
        subprocess.check_call(["g++", "-O1", "-std=c++17", "-I", inc, src, "-o", exe])
        inp = "".join(f"{s} {q} {o}\n" for s, q, o in cases)
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    return [line.split() for line in out.strip().splitlines()]


def main():
    rng = np.random.default_rng(20260)
    cases = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2**64 - 1, 2**64 - 1, 2**64 - 1)]
    for _ in range(27):
        cases.append(tuple(int(v) for v in rng.integers(0, 2**63, size=3, dtype=np.uint64)))
    outs = torch_vectors(cases)
    torch_cases = []
    for (seed, subseq, offset), out in zip(cases, outs):
        torch_cases.append({
            "ctr": [f"{offset & 0xffffffff:08x}", f"{offset >> 32:08x}",
                    f"{subseq & 0xffffffff:08x}", f"{subseq >> 32:08x}"],
            "key": [f"{seed & 0xffffffff:08x}", f"{seed >> 32:08x}"],
            "out": out,
        })
    doc = {
        "generator": "philox4x32, 10 rounds",
        "random123_kat": RANDOM123_KAT,
        "torch_engine": torch_cases,
        "made_by": "tests/golden/make_philox_vectors.py",
    }
    with open(os.path.join(HERE, "philox_vectors.json"), "w") as f:
        json.dump(doc, f, indent=1)
    print("wrote", len(torch_cases), "torch-engine vectors")


if __name__ == "__main__":
    sys.exit(main())
