"""Writes the golden fixtures of tests/golden/ from LBM-Bench's own sources
compiled into oracle/_ref/ (make -C oracle ref).  Runs only where the
reference tree is present; the fixtures are committed and the tests read
them everywhere else.

Fixtures (all numpy .npz, float64/int32, no pickles):
  geometry.npz      lattice flags of src/lbm.c:52-187 for several types
  runs.npz          logical PDFs after n steps of the reference's pull-soa
                    kernel (-DVERIFICATION: body force on), from seeded
                    inputs stored in the same file; every variant must
                    reproduce them (aa-soa runs an odd n as n+1)
  runs_sp.npz       the same for the single-precision build, one case
  verification.json the -V error norm of every kernel, dp and sp
                    (cli.c:84-99, lbm.c:964-1023)
  manifest.json     parameters and sha256 of each file
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from oracle import KERNELS, Reference  # noqa: E402

GEOMETRY_CASES = [
    ("channel", (12, 12, 12), (0, 0, 1)),
    ("fluid", (12, 12, 12), (0, 0, 0)),
    ("blocks-4", (12, 12, 12), (0, 0, 0)),
    ("box", (10, 11, 12), (1, 0, 1)),
    ("pipe", (12, 12, 12), (0, 0, 0)),
    ("blocks", (20, 20, 20), (0, 0, 0)),
    ("channel", (9, 10, 11), (0, 0, 0)),
]

RUN_CASES = [
    ("channel", (6, 8, 10), (0, 0, 1)),
    ("fluid", (6, 7, 8), (0, 0, 0)),
    ("blocks-4", (9, 10, 11), (0, 0, 0)),
]
RUN_STEPS = (1, 2, 3, 4, 50)
OMEGA = 1.25
XFORCE = 1e-4
AMP = 1e-2


def splitmix_init(dims, flags, seed, amp):
    """f_i = w_i (1 + eps) with eps from a numpy restatement of the
    splitmix64 chain; only used to make inputs (stored with the outputs)."""
    w = np.array([1 / 18] * 4 + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 18] + [1 / 36] * 4 + [1 / 3])
    rng = np.random.default_rng(seed)
    eps = rng.uniform(-amp, amp, size=tuple(dims) + (19,))
    out = w * (1.0 + eps)
    fluid = flags.transpose(2, 1, 0) != 0
    out[~fluid] = 0.0
    return out


def sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def main():
    ref = Reference()
    ref_sp = Reference(sp=True)
    manifest = {"source": "LBM-Bench src/ compiled by oracle/Makefile (gcc -O3 -ffast-math "
                          "-march=x86-64-v3 -DVERIFICATION)",
                "omega": OMEGA, "xforce": XFORCE, "amplitude": AMP, "magic": 1.0 / 12.0,
                "steps": list(RUN_STEPS)}

    geo = {}
    for i, (g, dims, per) in enumerate(GEOMETRY_CASES):
        lat, per_out, counts = ref.geometry(g, dims, per)
        geo[f"g{i}_flags"] = lat.astype(np.int32)
        geo[f"g{i}_periodic"] = np.array(per_out, dtype=np.int32)
        geo[f"g{i}_counts"] = np.array(counts, dtype=np.int32)
    manifest["geometry_cases"] = [[g, list(d), list(p)] for g, d, p in GEOMETRY_CASES]
    np.savez_compressed(os.path.join(HERE, "geometry.npz"), **geo)

    runs = {}
    spread = {}
    for i, (g, dims, per) in enumerate(RUN_CASES):
        lat, per_out, _ = ref.geometry(g, dims, per)
        init = splitmix_init(dims, lat, 100 + i, AMP)
        runs[f"r{i}_init"] = init
        runs[f"r{i}_flags"] = lat.astype(np.int32)
        runs[f"r{i}_periodic"] = np.array(per_out, dtype=np.int32)
        for n in RUN_STEPS:
            base, stats = ref.run("pull-soa", g, dims, per, init, n, OMEGA, XFORCE)
            runs[f"r{i}_n{n}"] = base
            runs[f"r{i}_n{n}_density"] = np.array([stats[0]])
            fluid = lat.transpose(2, 1, 0) != 0
            for k in KERNELS:
                out, _ = ref.run(k, g, dims, per, init, n, OMEGA, XFORCE)
                expect = base if (k != "aa-soa" or n % 2 == 0) else None
                if expect is None:
                    continue
                dev = float(np.abs(out - expect)[fluid].max() / np.abs(expect[fluid]).max())
                spread[f"r{i}_n{n}_{k}"] = dev
    manifest["run_cases"] = [[g, list(d), list(p)] for g, d, p in RUN_CASES]
    manifest["reference_variant_spread"] = spread
    np.savez_compressed(os.path.join(HERE, "runs.npz"), **runs)

    lat, per_out, _ = ref_sp.geometry("channel", (6, 8, 10), (0, 0, 1))
    init = runs["r0_init"]
    sp = {"init": init, "flags": lat.astype(np.int32)}
    for k in KERNELS:
        out, _ = ref_sp.run(k, "channel", (6, 8, 10), (0, 0, 1), init, 10, OMEGA, XFORCE)
        sp[f"n10_{k}"] = out
    np.savez_compressed(os.path.join(HERE, "runs_sp.npz"), **sp)

    ver = {"dp": {k: ref.verification_norm(k) for k in KERNELS},
           "sp": {k: ref_sp.verification_norm(k) for k in KERNELS}}
    with open(os.path.join(HERE, "verification.json"), "w") as f:
        json.dump(ver, f, indent=1)

    manifest["sha256"] = {n: sha(os.path.join(HERE, n)) for n in
                          ("geometry.npz", "runs.npz", "runs_sp.npz", "verification.json")}
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
