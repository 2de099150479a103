"""Writes golden vectors from the reference's own binning and kd-tree code.

Runs oracle/_ref/ref_harness (built by oracle/build_ref.sh from the unmodified
reference sources src/constants.f90, src/utils.f90, src/kdtree.f90) on seeded
inputs and stores inputs and outputs as .npz fixtures next to this script:

  ref_bins.npz    create_bins edges / centres and find_bin_index for lin, log,
                  ihs and sqrt spacing (src/utils.f90:101-230)
  ref_kdtree.npz  search_within_r index sets, tree and brute force
                  (src/kdtree.f90:207-375)

Run in the build container only (the reference tree is not on the GPU host):
    python tests/golden/make_ref_vectors.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

BIN_CASES = [  # spacing, x_min, x_max, n_bin, p
    (1, 1.0, 100.0, 20, 1.0),
    (1, 0.5, 250.0, 30, 1.0),
    (1, 1.0, 1024.0, 10, 1.0),
    (0, 0.0, 10.0, 10, 1.0),
    (0, 2.5, 40.0, 15, 1.0),
    (2, 0.5, 20.0, 12, 0.7),
    (2, 1.0, 300.0, 25, 2.0),
    (3, 0.0, 20.0, 10, 1.0),
    (3, 1.0, 200.0, 20, 1.0),
]


def run(mode, payload: bytes) -> bytes:
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(payload)
        subprocess.run([HARNESS, mode, fin, fout], check=True)
        with open(fout, "rb") as f:
            return f.read()


def make_bins(rng):
    out = {}
    for c, (sp, lo, hi, nb, p) in enumerate(BIN_CASES):
        xs = [rng.uniform(lo * 0.8 if lo > 0 else -1.0, hi * 1.1, 4000)]
        xs.append(np.array([lo, hi, np.nextafter(np.float32(hi), np.float32(np.inf))]))
        x = np.concatenate(xs).astype(np.float32)
        payload = (np.array([sp, nb, len(x)], np.int32).tobytes() + np.array([lo, hi, p], np.float32).tobytes()
                   + x.tobytes())
        raw = run("bins", payload)
        edges = np.frombuffer(raw, np.float32, nb + 1, 0)
        centers = np.frombuffer(raw, np.float32, nb, 4 * (nb + 1))
        idx = np.frombuffer(raw, np.int32, len(x), 4 * (2 * nb + 1))
        out[f"case{c}_params"] = np.array([sp, lo, hi, nb, p], np.float64)
        out[f"case{c}_x"] = x
        out[f"case{c}_edges"] = edges.copy()
        out[f"case{c}_centers"] = centers.copy()
        out[f"case{c}_idx"] = idx.copy()
    out["n_cases"] = np.array(len(BIN_CASES))
    np.savez_compressed(os.path.join(HERE, "ref_bins.npz"), **out)


def make_tree(rng):
    n, nq, r = 4000, 40, np.float32(7.5)
    px = rng.uniform(0, 100, n).astype(np.float32)
    py = rng.uniform(0, 100, n).astype(np.float32)
    qx = rng.uniform(-5, 105, nq).astype(np.float32)
    qy = rng.uniform(-5, 105, nq).astype(np.float32)
    payload = np.array([n, nq], np.int32).tobytes() + np.array([r], np.float32).tobytes() + \
        px.tobytes() + py.tobytes() + qx.tobytes() + qy.tobytes()
    raw = np.frombuffer(run("tree", payload), np.int32)
    pos, tree_sets, brute_sets = 0, [], []
    for _ in range(nq):
        for dst in (tree_sets, brute_sets):
            k = raw[pos]
            dst.append(raw[pos + 1:pos + 1 + k].copy())
            pos += 1 + k
    counts = np.array([[len(t), len(b)] for t, b in zip(tree_sets, brute_sets)], np.int32)
    np.savez_compressed(os.path.join(HERE, "ref_kdtree.npz"), px=px, py=py, qx=qx, qy=qy, r=np.array([r]),
                        counts=counts, tree_idx=np.concatenate(tree_sets), brute_idx=np.concatenate(brute_sets))


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref/ref_harness missing: run oracle/build_ref.sh")
    rng = np.random.default_rng(20240611)
    make_bins(rng)
    make_tree(rng)
    print("wrote ref_bins.npz, ref_kdtree.npz")
