"""Writes tests/golden/kitten_vectors.npz from the reference's own header-only math.

Builds oracle/_ref/kitten_probe from the reference headers under
/root/reference (``make -C oracle ref``), runs it, and stores each section's
inputs and outputs.  Run in the build container only; the .npz is the
committed fixture that tests/test_oracle_primitives.py reads.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    probe = os.path.join(REPO, "oracle", "_ref", "kitten_probe")
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "v.bin")
        subprocess.run([probe, out], check=True)
        raw = open(out, "rb").read()
    arrays = {}
    off = 0
    while off < len(raw):
        name = raw[off:off + 16].split(b"\0")[0].decode()
        count, nin, nout = np.frombuffer(raw, np.uint32, 3, off + 16)
        off += 28
        a_in = np.frombuffer(raw, np.float32, count * nin, off).reshape(count, nin)
        off += 4 * count * nin
        a_out = np.frombuffer(raw, np.float32, count * nout, off).reshape(count, nout)
        off += 4 * count * nout
        arrays[name + "_in"] = a_in.copy()
        arrays[name + "_out"] = a_out.copy()
    path = os.path.join(HERE, "kitten_vectors.npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, sorted(arrays))


if __name__ == "__main__":
    sys.exit(main())
