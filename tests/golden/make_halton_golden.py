"""Writes tests/golden/halton_golden.npz from the reference's own Halton
library (Aurora/external/halton, header-only), compiled with the driver
oracle/halton_ref_main.cpp by `make -C oracle ref`.

The fixture holds inputs and outputs only: Halton_sampler::sample (after
init_faure) for 256 dimensions at 40 indices, and Halton_enum's scales and
get_index for 7 resolutions.  Run in a container that has the reference:

    python tests/golden/make_halton_golden.py
"""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "halton_ref")], check=True, stdout=subprocess.PIPE,
                         text=True).stdout.split("\n")
    S, R, E = [], [], []
    for line in out:
        f = line.split()
        if not f:
            continue
        if f[0] == "S":
            S.append([int(x) for x in f[1:]])
        elif f[0] == "R":
            R.append([int(x) for x in f[1:]])
        elif f[0] == "E":
            E.append([int(x) for x in f[1:]])
    S = np.array(S, np.uint64)
    np.savez_compressed(os.path.join(HERE, "halton_golden.npz"),
                        sample_dim=S[:, 0].astype(np.uint32), sample_index=S[:, 1].astype(np.uint32),
                        sample_bits=S[:, 2].astype(np.uint32), res=np.array(R, np.uint32),
                        enum=np.array(E, np.uint32))


if __name__ == "__main__":
    main()
