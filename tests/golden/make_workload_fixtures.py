"""Writes the benchmark workloads' generated inputs as JSON fixtures: the 64-sphere
ADI phantom (SplitMix64 seed 0x5EED0002) and the 256 curved-guide CN jobs (seed
0x5EED0003).  tests/test_workloads.py checks that the generators still produce
them, so a change to a generator cannot silently change a benchmark.
Run from the repository root: python tests/golden/make_workload_fixtures.py"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from paxpro_amd import workloads as W  # noqa: E402


def main():
    phantom = [list(s) for s in W.adi_phantom()]
    jobs = [{k: j[k] for k in ("nx", "lam", "R", "width", "xmin", "dx", "dz", "nsteps")}
            for j in W.curved_guide_jobs()]
    with open(os.path.join(HERE, "adi_phantom_64.json"), "w") as f:
        json.dump({"seed": "0x5EED0002", "fields": ["cx", "cy", "cz", "r", "delta", "beta"],
                   "spheres": phantom}, f, indent=0)
    with open(os.path.join(HERE, "cn_jobs_256.json"), "w") as f:
        json.dump({"seed": "0x5EED0003", "jobs": jobs}, f, indent=0)


if __name__ == "__main__":
    main()
