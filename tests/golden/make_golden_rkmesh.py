"""Writes tests/golden/rkmesh.json from the reference's own rkmesh / rkmesh2D
(subroutines/mhkpack_subs.f90:1-52), run through oracle/_ref/ref_rkmesh.x.

    make -C oracle && python tests/golden/make_golden_rkmesh.py

Cases: seeded random lattices (2D hexagonal and oblique with a vacuum axis, cubic,
random 3D cells) at k-point densities rk spread over 0.5..200, plus densities placed
within 1e-6 of a rounding boundary of rk |b_i| / (2 pi) + 1/2, where a wrong value
of pi (double instead of the reference's single-precision acos(-1.0)) or a dropped
term moves the grid by one.  The fixture is plain data: inputs and the reference's
outputs.
"""
from __future__ import annotations

import json
import math
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
EXE = ROOT / "oracle" / "_ref" / "ref_rkmesh.x"
OUT = Path(__file__).resolve().parent / "rkmesh.json"


def lattices(rng):
    yield np.array([[3.2, 0, 0], [-1.6, 1.6 * math.sqrt(3.0), 0], [0, 0, 20.0]])
    yield np.diag([5.4, 5.4, 5.4])
    for _ in range(8):
        a = rng.uniform(2.5, 6.0)
        b = rng.uniform(2.5, 6.0)
        g = rng.uniform(math.radians(60), math.radians(120))
        yield np.array([[a, 0, 0], [b * math.cos(g), b * math.sin(g), 0], [0, 0, rng.uniform(12, 30)]])
    for _ in range(8):
        m = np.diag(rng.uniform(3, 7, 3)) + rng.uniform(-0.8, 0.8, (3, 3))
        yield m


def main():
    rng = np.random.default_rng(20240607)
    cases = []
    for L in lattices(rng):
        for rk in [0.5, 3.0, 17.0, 40.0, 120.0, 200.0, float(rng.uniform(1, 150))]:
            cases.append((rk, L))
        # densities next to a rounding boundary of the first axis: rk |b1|/(2 pi) + 1/2 = n
        # with |b1|/(2 pi) = area of the other two vectors' cross product / volume
        v = abs(np.linalg.det(L))
        s1 = np.linalg.norm(np.cross(L[1], L[2])) / v
        n = int(rng.integers(4, 60))
        base = (n - 0.5) / s1
        for d in (-1e-6, 1e-6):
            cases.append((base * (1 + d), L))
    text = "".join(f"{float(rk)!r} " + " ".join(repr(float(x)) for x in L.reshape(-1)) + "\n" for rk, L in cases)
    out = subprocess.run([str(EXE)], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    rows = [list(map(int, l.split())) for l in out if l.strip()]
    assert len(rows) == len(cases)
    data = [{"rk": float(rk), "rlat": L.reshape(-1).tolist(), "rk3d": r[:3], "rk2d": r[3:]} for (rk, L), r in zip(cases, rows)]
    OUT.write_text(json.dumps({"source": "subroutines/mhkpack_subs.f90:1-52 (rkmesh, rkmesh2D)", "cases": data},
                              separators=(",", ":")))
    print(f"wrote {len(data)} cases to {OUT}")


if __name__ == "__main__":
    main()
