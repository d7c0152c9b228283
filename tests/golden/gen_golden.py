"""Write the reference golden vectors used by the oracle tests.

Builds the reference's own ch4/v3 sources (in place, from /root/reference)
together with our harness main (oracle/ref_harness.cpp) into oracle/_ref/,
runs its `golden` mode and packs the raw float64 dumps into
tests/golden/ref_v3.npz and the text files its Output writers produced into
tests/golden/ref_outputs.npz.  The fixtures hold only data: inputs and the
reference's outputs.  Run from the repository root:

    python tests/golden/gen_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
TABLE = os.path.join(ROOT, "engineering-degree-in-plasma-simulations_amd", "data", "sigma_el_O.dat")


def main():
    subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    with tempfile.TemporaryDirectory() as tmp:
        subprocess.check_call([harness, "golden", tmp, TABLE], stdout=subprocess.DEVNULL,
                              stderr=subprocess.DEVNULL)
        arrays = {}
        with open(os.path.join(tmp, "manifest.txt")) as f:
            for line in f:
                name, count = line.split()
                data = np.fromfile(os.path.join(tmp, name + ".bin"), dtype="<f8")
                assert data.size == int(count), name
                arrays[name] = data
        # the text files the reference's Output writers produced
        texts = {}
        for name in sorted(os.listdir(os.path.join(tmp, "results"))):
            with open(os.path.join(tmp, "results", name), "rb") as f:
                texts[name] = np.frombuffer(f.read(), dtype=np.uint8)
    out = os.path.join(ROOT, "tests", "golden", "ref_v3.npz")
    np.savez_compressed(out, **arrays)
    print("wrote %s (%d arrays, %d bytes)" % (out, len(arrays), os.path.getsize(out)))
    out = os.path.join(ROOT, "tests", "golden", "ref_outputs.npz")
    np.savez_compressed(out, **texts)
    print("wrote %s (%s)" % (out, ", ".join(sorted(texts))))


if __name__ == "__main__":
    sys.exit(main())
