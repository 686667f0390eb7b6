"""The oracle against the committed golden fixtures of the reference
(tests/golden/, written by make_golden.py from LBM-Bench's own sources)."""
import json
import os

import numpy as np
import pytest
from conftest import GOLDEN, WEIGHTS, rel_err

from oracle import KERNELS


def test_fixture_hashes(manifest):
    import hashlib
    for name, digest in manifest["sha256"].items():
        with open(os.path.join(GOLDEN, name), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == digest, name


def test_geometry_fixtures(oracle, manifest):
    geo = np.load(os.path.join(GOLDEN, "geometry.npz"))
    for i, (g, dims, per) in enumerate(manifest["geometry_cases"]):
        lat, per_o, counts = oracle.geometry(g, tuple(dims), tuple(per))
        assert (lat == geo[f"g{i}_flags"]).all(), g
        assert tuple(geo[f"g{i}_periodic"]) == per_o
        assert tuple(geo[f"g{i}_counts"]) == counts


@pytest.mark.parametrize("case", [0, 1, 2])
@pytest.mark.parametrize("kernel", KERNELS)
def test_run_fixtures(oracle, golden_runs, manifest, case, kernel):
    g, dims, per = manifest["run_cases"][case]
    dims = tuple(dims)
    lat = golden_runs[f"r{case}_flags"]
    per_o = tuple(golden_runs[f"r{case}_periodic"])
    init = golden_runs[f"r{case}_init"]
    fluid = lat.transpose(2, 1, 0) != 0
    for n in manifest["steps"]:
        got, done = oracle.run(kernel, dims, lat, per_o, init, n, manifest["omega"],
                               manifest["magic"], manifest["xforce"], True)
        key = f"r{case}_n{done}"
        if key not in golden_runs:
            continue
        assert rel_err(got, golden_runs[key], fluid) < 1e-13, (kernel, n)


def test_sp_fixture_is_float(manifest):
    sp = np.load(os.path.join(GOLDEN, "runs_sp.npz"))
    for k in KERNELS:
        a = sp[f"n10_{k}"]
        assert np.array_equal(a, a.astype(np.float32).astype(np.float64))


def test_verification_fixture_within_reference_tolerance():
    with open(os.path.join(GOLDEN, "verification.json")) as f:
        ver = json.load(f)
    for prec in ("dp", "sp"):
        for k in KERNELS:
            assert 0.0 < ver[prec][k] < 0.1


def test_density_fixture(oracle, golden_runs, manifest):
    lat = golden_runs["r0_flags"]
    dims = tuple(manifest["run_cases"][0][1])
    for n in manifest["steps"]:
        d = oracle.density(dims, lat, golden_runs[f"r0_n{n}"])
        assert abs(d - golden_runs[f"r0_n{n}_density"][0]) < 1e-14
    assert abs(WEIGHTS.sum() - 1.0) < 1e-15
