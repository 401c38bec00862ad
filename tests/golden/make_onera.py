#!/usr/bin/env python3
"""Writes tests/golden/onera_m6_120b.npz: the reference's own multi-block
ONERA M6 mesh (Meshes/oneraM6_120B: topology file and 120 SU2 block files,
582,752 tets, written by the reference's partitioner) and its input file
(input_onera120), stored byte for byte (deflate-compressed) so that a machine
without the reference tree can read exactly the files the reference reads.

Container-only (reads /root/reference). The arrays are the files' bytes:
  file:<name>   each file of Meshes/oneraM6_120B
  input         input_onera120

usage: python tests/golden/make_onera.py
"""
import os

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    d = os.path.join(REF, "Meshes", "oneraM6_120B")
    arrays = {"file:" + f: np.frombuffer(open(os.path.join(d, f), "rb").read(), np.uint8) for f in sorted(os.listdir(d))}
    arrays["input"] = np.frombuffer(open(os.path.join(REF, "input_onera120"), "rb").read(), np.uint8)
    out = os.path.join(HERE, "onera_m6_120b.npz")
    np.savez_compressed(out, **arrays)
    print(f"{out}: {len(arrays) - 1} mesh files, {os.path.getsize(out) / 1e6:.1f} MB")


if __name__ == "__main__":
    main()
