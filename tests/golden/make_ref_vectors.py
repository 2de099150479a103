"""Writes golden vectors from the reference's own binning and kd-tree code.

Runs oracle/_ref/ref_harness (built by oracle/build_ref.sh from the unmodified
reference sources src/constants.f90, src/utils.f90, src/kdtree.f90) on seeded
inputs and stores inputs and outputs as .npz fixtures next to this script:

  ref_bins.npz    create_bins edges / centres and find_bin_index for lin, log,
                  ihs and sqrt spacing (src/utils.f90:101-230)
  ref_kdtree.npz  search_within_r index sets, tree and brute force
                  (src/kdtree.f90:207-375)
  ref_corr.npz    tangential_shear_tpcf / scalar_kappa_tpcf outputs
                  (src/tpcf.f90:190-545, kd-tree path, marks on) through
                  oracle/_ref/libref2pcf.so, with their FP32 inputs

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
        # dense FP32 values around every edge: where FP32 and FP64 binning can differ
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        e = oracle.edges(sp, lo, hi, nb, p)
        near = (e[:, None] * (1.0 + np.arange(-64, 65)[None, :] * 2.0 ** -24)).ravel()
        xs.append(near[near >= 0])
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


def blobs(rng, n, side, n_blob=12, sigma=2.0):
    """Clustered positions: Gaussian blobs plus a uniform half."""
    c = rng.uniform(0, side, (n_blob, 2))
    k = rng.integers(0, n_blob, n // 2)
    p = c[k] + rng.normal(0, sigma, (n // 2, 2))
    u = rng.uniform(0, side, (n - n // 2, 2))
    q = np.clip(np.concatenate([p, u]), 0, side)
    return q[:, 0].astype(np.float32), q[:, 1].astype(np.float32)


def shear_case(rng, n_a, n_b, side, clustered=False, masks=True, m=True):
    f32 = np.float32
    xa, ya = blobs(rng, n_a, side) if clustered else (rng.uniform(0, side, n_a).astype(f32),
                                                      rng.uniform(0, side, n_a).astype(f32))
    xb, yb = blobs(rng, n_b, side) if clustered else (rng.uniform(0, side, n_b).astype(f32),
                                                      rng.uniform(0, side, n_b).astype(f32))
    lens = dict(x=xa, y=ya, z=np.zeros(n_a, f32), val=rng.normal(0, .02, n_a).astype(f32),
                w=rng.uniform(.5, 1.5, n_a).astype(f32))
    src = dict(x=xb, y=yb, z=np.ones(n_b, f32), e1=rng.normal(0, .27, n_b).astype(f32),
               e2=rng.normal(0, .27, n_b).astype(f32), w=rng.uniform(.5, 1.5, n_b).astype(f32),
               m=(rng.normal(0, .02, n_b) if m else np.zeros(n_b)).astype(f32))
    if masks:
        lens["z"] = rng.uniform(-0.2, 1.0, n_a).astype(f32)       # z < 0 masked, z_A > z_B skipped
        src["z"] = rng.uniform(-0.2, 1.2, n_b).astype(f32)
        lens["w"][rng.random(n_a) < 0.05] = 0.0                    # skipped lenses
        src["w"][rng.random(n_b) < 0.05] = 0.0
    return lens, src


def kappa_case(rng, n_a, n_b, side, auto=False):
    f32 = np.float32
    xa, ya = rng.uniform(0, side, n_a).astype(f32), rng.uniform(0, side, n_a).astype(f32)
    lens = dict(x=xa, y=ya, val=rng.normal(0, .02, n_a).astype(f32), w=rng.uniform(.5, 1.5, n_a).astype(f32))
    if auto:
        src = dict(x=xa, y=ya, kappa_E=lens["val"], kappa_B=rng.normal(0, .02, n_a).astype(f32), w=lens["w"])
    else:
        src = dict(x=rng.uniform(0, side, n_b).astype(f32), y=rng.uniform(0, side, n_b).astype(f32),
                   kappa_E=rng.normal(0, .02, n_b).astype(f32), kappa_B=rng.normal(0, .02, n_b).astype(f32),
                   w=rng.uniform(.5, 1.5, n_b).astype(f32))
    return lens, src


def make_corr(rng):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    cases = []
    #              kind, spacing, x_min, x_max, n_bin, p, (lens, src)
    cases.append((0, 1, 0.5, 20.0, 12, 1.0, shear_case(rng, 1500, 2500, 80.0)))
    cases.append((1, 0, 0.0, 10.0, 10, 1.0, kappa_case(rng, 2000, 0, 60.0, auto=True)))
    cases.append((0, 2, 0.5, 15.0, 10, 0.7, shear_case(rng, 1500, 1500, 60.0, clustered=True, masks=False)))
    cases.append((1, 3, 1.0, 30.0, 15, 1.0, kappa_case(rng, 1200, 2000, 90.0)))
    lens, src = shear_case(rng, 600, 600, 40.0, masks=False, m=False)
    src["x"][:20], src["y"][:20] = lens["x"][:20], lens["y"][:20]  # zero separations with x_min = 0
    cases.append((0, 0, 0.0, 8.0, 8, 1.0, (lens, src)))
    cases.append((0, 1, 1.0, 25.0, 8, 1.0, shear_case(rng, 1000, 3000, 100.0, clustered=True)))
    out = {"n_cases": np.array(len(cases))}
    for c, (kind, sp, lo, hi, nb, p, (lens, src)) in enumerate(cases):
        r = oracle.reference(kind, lens, src, lo, hi, nb, sp, p, n_threads=4, use_tree=True)
        out[f"case{c}_params"] = np.array([kind, sp, lo, hi, nb, p], np.float64)
        for k, v in lens.items():
            out[f"case{c}_lens_{k}"] = v
        for k, v in src.items():
            out[f"case{c}_src_{k}"] = v
        for k in ("mean_theta", "g_t", "g_x", "weights", "n_pairs", "xi_e", "xi_b", "k_marks", "w_marks",
                  "marks_weight"):
            out[f"case{c}_ref_{k}"] = r[k]
    np.savez_compressed(os.path.join(HERE, "ref_corr.npz"), **out)


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref/ref_harness missing: run oracle/build_ref.sh")
    rng = np.random.default_rng(20240611)
    make_bins(rng)
    make_tree(rng)
    make_corr(np.random.default_rng(20240612))
    print("wrote ref_bins.npz, ref_kdtree.npz, ref_corr.npz")
