"""Pack the reference's own module-test case (2D dam break, 23,030 particles)
into one compressed fixture.

Source (data files of the reference's tests, read in this container only):
  sources/RRSPH_COMMON/tests/caseDamBreak/{ModelParams.json, ParticleParams.json,
  dump/0.00.csv, test/OMP/*.csv}
Output: tests/golden/case_dambreak.npz -- the parsed values, unchanged.

Run: python tests/golden/make_dambreak_fixture.py [/root/reference]
"""
import csv
import json
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
CASE = os.path.join(REF, "sources", "RRSPH_COMMON", "tests", "caseDamBreak")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "case_dambreak.npz")


def read_csv(path):
    with open(path, newline="") as f:
        rd = csv.reader(f)
        header = next(rd)
        rows = [r for r in rd if r]
    return header, np.array(rows, dtype=np.float64)


def main():
    out = {}
    with open(os.path.join(CASE, "ModelParams.json")) as f:
        out["model_params_json"] = np.frombuffer(json.dumps(json.load(f), sort_keys=True).encode(), dtype=np.uint8)
    with open(os.path.join(CASE, "ParticleParams.json")) as f:
        out["particle_params_json"] = np.frombuffer(json.dumps(json.load(f), sort_keys=True).encode(),
                                                    dtype=np.uint8)
    header, dump = read_csv(os.path.join(CASE, "dump", "0.00.csv"))
    for k, name in enumerate(header):
        out["dump_" + name] = dump[:, k]
    for backend in ("OMP",):
        d = os.path.join(CASE, "test", backend)
        for fn in sorted(os.listdir(d)):
            if not fn.endswith(".csv"):
                continue
            header, vals = read_csv(os.path.join(d, fn))
            stem = fn[:-4]
            assert header[0] == "particle"
            assert np.array_equal(vals[:, 0], np.arange(len(vals)))
            for k, name in enumerate(header[1:], start=1):
                out[f"{backend}_{stem}__{name}"] = vals[:, k]
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
