"""Writes the committed fixtures of tests/golden/ from the oracle:
  small_pair.npz      a 128x96 seeded pair (disp 0..31) with the oracle's
                      D1, D2, support points, meshes and final support lattice;
  synth_sha256.txt    SHA-256 of the 450x375 left/right images of seed 0.
The reference itself cannot run here (CUDA only, no GPU), so these pin the
oracle against regressions; parity with the reference rests on the closed
forms and brute-force checks in tests/test_oracle_*.py.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402
import oracle  # noqa: E402


def main():
    synth = __graft_entry__.load_synth()
    l, r, _ = synth.make_pair(128, 96, 31, regions=4, seed=42)
    out = oracle.process(l, r, oracle.params(disp_max=31))
    np.savez_compressed(os.path.join(HERE, "small_pair.npz"), left=l, right=r, disp_max=31,
                        D1=out["D1"], D2=out["D2"], sp0=out["sp0"], tri0=out["tri0"], tri1=out["tri1"],
                        can_final=out["can_final"])
    l0, r0, _, _ = synth.config_pair("450x375_d64")
    with open(os.path.join(HERE, "synth_sha256.txt"), "w") as f:
        f.write(hashlib.sha256(l0.tobytes()).hexdigest() + "\n" + hashlib.sha256(r0.tobytes()).hexdigest() + "\n")


if __name__ == "__main__":
    main()
