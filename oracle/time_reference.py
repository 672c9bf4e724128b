"""Writes a sample of the 4096x4096 large-grid bed (the benchmark's seeded
recipe at a smaller size) and times the reference JS solver on it with
time_reference.mjs.  Build-container only: the reference does not travel.

    python oracle/time_reference.py [--n 512] [--steps 5] [--reps 3]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "water-simulator_amd"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args()
    import workloads
    (b, h, hu, hv), p = workloads.large_grid(nx=a.n, ny=a.n, dx=10.0)
    with tempfile.NamedTemporaryFile(suffix=".f64", delete=False) as f:
        f.write(b.astype("<f8").tobytes())
        bed = f.name
    try:
        r = subprocess.run(["node", os.path.join(HERE, "time_reference.mjs"), bed, str(a.n), str(a.n),
                            str(a.steps), str(a.reps)], capture_output=True, text=True, check=True)
    finally:
        os.unlink(bed)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
