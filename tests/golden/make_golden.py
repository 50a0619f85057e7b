"""Writes tests/golden/siddon_test.json: the reference's test/SiddonTest.cpp
input set and the Siddon lists the CPU oracle gives for it.

Inputs: tests/golden/siddon_test_inputs.cpp, compiled with g++, makes the
10000 segments with the same libstdc++ calls as SiddonTest.cpp:8-34; their
SHA-256 is stored, and tests/siddon_inputs.py (a numpy restatement of
minstd_rand0 + generate_canonical) must reproduce the same bytes.
Outputs: the oracle's CSR lists on the test's grid (100^3 voxels of 2 mm from
-100 mm, SiddonTest.cpp:11-14), stored as SHA-256 digests of the counts,
indices and length bits, plus the full lists of the first non-empty segments.
The reference prints these lists (SiddonTest.cpp:36-45) but never stored them,
so the fixture is a regression record of the oracle, not reference output.

Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

N = 10000
GRID = dict(n=100, d=2.0, lo=-100.0)


def cpp_inputs(n=N):
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "gen")
        subprocess.run(["g++", "-O2", "-o", exe, os.path.join(HERE, "siddon_test_inputs.cpp")],
                       check=True)
        return subprocess.run([exe, str(n)], check=True, capture_output=True).stdout


def digests(row, x, y, z, ln):
    h = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    return {"row_ptr_u64": h(row.astype("<u8")), "x_u32": h(x.astype("<u4")),
            "y_u32": h(y.astype("<u4")), "z_u32": h(z.astype("<u4")),
            "length_f64_bits": h(ln.astype("<f8"))}


def main():
    import oracle
    import siddon_inputs
    from muontomography_amd._abi import make_grid
    raw = cpp_inputs()
    p1, p2 = siddon_inputs.siddon_test_segments(N)
    mine = np.concatenate([p1, p2], 1).astype("<f8").tobytes()
    assert mine == raw, "numpy restatement differs from libstdc++"
    g = make_grid(GRID["n"], GRID["d"], GRID["lo"])
    row, x, y, z, ln = oracle.siddon_csr(g, p1, p2)
    cnt = np.diff(row.astype(np.int64))
    first = []
    for i in np.nonzero(cnt)[0][:3]:
        a, b = int(row[i]), int(row[i + 1])
        first.append({"segment": int(i), "p1": [v.hex() for v in p1[i]],
                      "p2": [v.hex() for v in p2[i]],
                      "voxels": [[int(x[k]), int(y[k]), int(z[k])] for k in range(a, b)],
                      "lengths": [float(ln[k]).hex() for k in range(a, b)]})
    out = {"source": "reference test/SiddonTest.cpp:8-45 inputs; CPU oracle outputs",
           "n_segments": N, "grid": GRID,
           "inputs_sha256": hashlib.sha256(raw).hexdigest(),
           "nonempty_segments": int((cnt > 0).sum()), "entries": int(row[-1]),
           "digests": digests(row, x, y, z, ln), "first_nonempty": first}
    with open(os.path.join(HERE, "siddon_test.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(out["nonempty_segments"], out["entries"])


if __name__ == "__main__":
    main()
