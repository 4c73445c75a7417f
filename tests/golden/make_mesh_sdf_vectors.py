"""Writes tests/golden/mesh_sdf_vectors.npz from the reference's own body SDF
construction.

Builds oracle/_ref/discregrid_probe from the reference sources under
/root/reference (``make -C oracle ref``: src/utilities/Geometry.cpp with the
vendored Discregrid, Eigen and glm), runs it on the meshes of
tests/mesh_cases.py and stores, per mesh, the inputs with the cubic grid's
node values, the sampled body grid (phi, grad) and TriangleMeshDistance's
signed distance, nearest point and triangle at query points.  Run in the
build container only; the .npz is the committed fixture that
tests/test_oracle_meshsdf.py reads.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))

import mesh_cases  # noqa: E402


def run_probe(probe, v, t, dim, q, td):
    src, dst = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
    with open(src, "wb") as f:
        f.write(np.array([len(v), len(t), *dim, len(q)], np.int32).tobytes())
        f.write(np.ascontiguousarray(v, np.float32).tobytes())
        f.write(np.ascontiguousarray(t, np.int32).tobytes())
        f.write(np.ascontiguousarray(q, np.float64).tobytes())
    subprocess.run([probe, src, dst], check=True, stdout=subprocess.DEVNULL)
    raw = open(dst, "rb").read()
    nn = int(np.frombuffer(raw, np.uint64, 1, 0)[0])
    off = 8
    n = int(np.prod(dim))
    out = {}
    for name, dt, count in (("coef", np.float64, nn), ("phi", np.float32, n), ("grad", np.float32, 3 * n),
                            ("sd", np.float64, len(q)), ("near", np.float64, 3 * len(q)), ("tri", np.int32, len(q))):
        out[name] = np.frombuffer(raw, dt, count, off).copy()
        off += np.dtype(dt).itemsize * count
    assert off == len(raw)
    return out


def main():
    subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    probe = os.path.join(REPO, "oracle", "_ref", "discregrid_probe")
    arrays = {}
    with tempfile.TemporaryDirectory() as td:
        for k, (name, v, t, dim) in enumerate(mesh_cases.cases()):
            q = mesh_cases.query_points(v, 400, 100 + k)
            out = run_probe(probe, v, t, dim, q, td)
            arrays[f"{name}_vertices"] = v
            arrays[f"{name}_triangles"] = t
            arrays[f"{name}_dim"] = np.array(dim, np.int32)
            arrays[f"{name}_query"] = q
            for key, a in out.items():
                arrays[f"{name}_{key}"] = a
    path = os.path.join(HERE, "mesh_sdf_vectors.npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
