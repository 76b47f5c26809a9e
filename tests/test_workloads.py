"""The benchmark's workload generators (paxpro_amd/workloads.py, plain numpy)
build the same CSG arrays as the oracle's transform builders, whose
translation is pinned by the reference's known answers, and draw the same
SplitMix64 streams as the parity tests."""
import numpy as np

import pyoracle as O
from paxpro_amd import workloads as W
from scenes import random_spheres as scene_spheres
from scenes import spheres_scene, splitmix64


def test_splitmix64_matches_published_first_values():
    # SplitMix64 with seed 0: first output 0xE220A8397B1DCDAF (Vigna's reference code)
    g = W.splitmix64(0)
    assert next(g) == (0xE220A8397B1DCDAF >> 11) * (1.0 / (1 << 53))
    a, b = W.splitmix64(0x5EED0002), splitmix64(0x5EED0002)
    assert [next(a) for _ in range(50)] == [next(b) for _ in range(50)]


def test_spheres_csg_equals_oracle_builders():
    sph = W.adi_phantom()
    assert sph == scene_spheres(0x5EED0002, 64, (-17000, 17000), (-2000, 2000), (500, 3000), (1e-6, 1e-5))
    s1, p1, m1 = W.spheres_csg(sph)
    s2, p2, m2 = spheres_scene(sph)
    assert np.array_equal(s1["T"], s2["T"]) and np.array_equal(s1["p"], s2["p"])
    assert np.array_equal(s1["type"], s2["type"]) and np.array_equal(s1["op"], s2["op"])
    for f in ("first_shape", "n_shapes", "op", "delta", "beta"):
        assert np.array_equal(p1[f], p2[f])
    assert np.array_equal(m1, m2)
    # parts are scanned last to first: a sphere's centre takes its material
    # unless a later sphere also holds it
    mat = O.CSGMaterial(s1, p1, m1)
    checked = 0
    for i, (cx, cy, cz, r, d, b) in enumerate(sph):
        later = [s for s in sph[i + 1:] if (cx - s[0]) ** 2 + (cy - s[1]) ** 2 + (cz - s[2]) ** 2 < s[3] ** 2]
        if not later:
            assert mat.eval(cx, cy, cz) == (d, b)
            checked += 1
    assert checked > 32


def test_curved_guide_jobs_and_trace():
    jobs = W.curved_guide_jobs()
    assert len(jobs) == 256
    nx = np.array([j["nx"] for j in jobs])
    assert nx.min() >= 256 and nx.max() <= 8192
    for j in jobs[:8]:
        hi = j["xmin"] + (j["nx"] - 1) * j["dx"]
        assert abs(hi - (j["width"] + 200.0)) < 1e-6 * abs(hi)
    tr = W.served_trace(300)
    t = np.array([r["t"] for r in tr])
    assert np.all(np.diff(t) > 0)
    frac = np.mean([r["kind"] == "fft3d" for r in tr])
    assert 0.6 < frac < 0.8
    assert {r["n"] for r in tr if r["kind"] == "fft3d"} == {256, 512, 1024, 2048}
    # mean arrival rate within 20 % of 200 / s
    assert 160 < len(tr) / t[-1] < 240


def test_generated_workloads_match_fixtures():
    # the committed inputs of the ADI and batched-CN benchmarks (floats round-trip
    # exactly through JSON)
    import json
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    ph = json.load(open(os.path.join(here, "adi_phantom_64.json")))
    assert [list(s) for s in W.adi_phantom()] == ph["spheres"]
    jobs = json.load(open(os.path.join(here, "cn_jobs_256.json")))["jobs"]
    gen = W.curved_guide_jobs()
    assert len(gen) == len(jobs) == 256
    for g, j in zip(gen, jobs):
        for k, v in j.items():
            assert g[k] == v, k
    nx = [j["nx"] for j in jobs]
    assert min(nx) >= 256 and max(nx) <= 8192
