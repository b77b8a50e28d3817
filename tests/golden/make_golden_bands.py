"""Writes the finite-Q exciton band fixtures under tests/golden/bands/.

Runs the unmodified reference program (oracle/_ref/wtb.x, built by oracle/Makefile
from the reference sources) with BSE_BND= T (BSET_BND= T for the finite-temperature
cases) on seeded synthetic TB models and a
KPATH_BSE file, and stores what it writes: bands_bse.dat (path coordinate as a
REAL, then the exciton energy, for every exciton and q-point) and
KLABELS-BSE.dat (the segment ends), beside the inputs (TB model arrays, run
parameters, the k-path).

    make -C oracle && python tests/golden/make_golden_bands.py

Needs the reference build (this container only); the fixtures are plain data.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from wantibexos_amd import tbmodel  # noqa: E402

PROGRAM = ROOT / "oracle" / "_ref" / "wtb.x"
OUTDIR = Path(__file__).resolve().parent / "bands"

BASE = dict(nc=2, nv=2, coultype="V2DRK", ediel=(1.0, 1.0, 1.0), lc=1.0, ez=1.0, w=0.0,
            r0=40.0, ktol=0.001, cshift=0.05, shift=(0.0, 0.0, 0.0), ebse0=0.0, ebsef=6.0,
            numbse=1001, renorm=True, dtfull=True)
# hexagonal path Gamma - K - K' - Gamma in fractional coordinates (segment pairs)
HEX_PATH = [(0.0, 0.0, 0.0), (0.6667, -0.3333, 0.0), (0.6667, -0.3333, 0.0), (0.3333, 0.3333, 0.0),
            (0.3333, 0.3333, 0.0), (0.0, 0.0, 0.0)]
CUB_PATH = [(0.0, 0.0, 0.0), (0.5, 0.0, 0.0), (0.5, 0.0, 0.0), (0.5, 0.5, 0.5)]


def cases():
    """(name, model factory, params, path (fractional), points per segment)"""
    return [
        ("hex6_g4_v2drk", lambda: tbmodel.hex2d(41, 6, (4, 4, 1)), dict(BASE, ngrid=(4, 4, 1)), HEX_PATH, 4),
        # V2DK: V(q -> 0) finite, so the exchange term is large near Gamma
        ("hex6_g3_v2dk_c1v2", lambda: tbmodel.hex2d(42, 6, (3, 3, 1)),
         dict(BASE, ngrid=(3, 3, 1), coultype="V2DK", ediel=(1.5, 4.0, 2.0), lc=6.0, nc=1, nv=2), HEX_PATH, 5),
        ("hex8_g5_v2dt2_shift", lambda: tbmodel.hex2d(43, 8, (5, 5, 1)),
         dict(BASE, ngrid=(5, 5, 1), coultype="V2DT2", lc=8.0, shift=(0.1, 0.2, 0.0)), HEX_PATH[:4], 3),
        ("cub6_g3_v3d", lambda: tbmodel.cubic3d(44, 6, (3, 3, 3)),
         dict(BASE, ngrid=(3, 3, 3), coultype="V3D", nc=1, nv=1), CUB_PATH, 3),
        # bsebndstemp (BSET_BND= T): gap correction on both grids, row factors f_v - f_c
        ("hex6_g4_v2dk_t5000_ve", lambda: tbmodel.hex2d(45, 6, (4, 4, 1)),
         dict(BASE, ngrid=(4, 4, 1), coultype="V2DK", ediel=(1.5, 4.0, 2.0), lc=6.0, bset=True, temp=5000.0,
              ta="VE", st=0.4, phavg=0.04), HEX_PATH[:4], 4),
        ("hex6_g3_v2drk_t3000_be", lambda: tbmodel.hex2d(46, 6, (3, 3, 1)),
         dict(BASE, ngrid=(3, 3, 1), bset=True, temp=3000.0, ta="BE", st=0.05, phavg=1500.0), HEX_PATH, 3),
    ]


def path_text(path, npts):
    lines = [str(len(path)), str(npts)] + [f"{a!r} {b!r} {c!r}" for a, b, c in path]
    return "\n".join(lines) + "\n"


def table(path):
    rows = [l.split() for l in Path(path).read_text().splitlines() if l.strip()]
    return np.array([[float(x) for x in r] for r in rows])


def run_case(name, factory, p, path, npts):
    model = factory()
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        (td / "out").mkdir()
        tbmodel.write_tb(model, td / "tb.dat")
        (td / "kpath.dat").write_text(path_text(path, npts))
        (td / "input.dat").write_text(tbmodel.input_dat(dict(p, kpath_bse="kpath.dat"), "tb.dat", nthreads=4))
        env = dict(os.environ, OMP_NUM_THREADS="4")
        r = subprocess.run([str(PROGRAM)], stdin=open(td / "input.dat"), cwd=td, env=env,
                           capture_output=True, text=True, timeout=600)
        bands = td / "out" / "bands_bse.dat"
        if not bands.exists():
            raise RuntimeError(f"{name}: no bands_bse.dat\n{r.stdout}\n{r.stderr}")
        t = table(bands)
        kl = table(td / "out" / "KLABELS-BSE.dat")[:, 0]
    nq = (len(path) // 2) * npts
    N = int(np.prod(p["ngrid"])) * p["nc"] * p["nv"]
    assert t.shape == (N * nq, 2), t.shape
    t = t.reshape(N, nq, 2)
    fx = dict(
        model_systype=np.array(model.systype), model_scs=np.float64(model.scs),
        model_efermi=np.float64(model.efermi), model_rlat=model.rlat, model_rvec=model.rvec,
        model_ndeg=model.ndeg, model_hop=model.hop, model_ihop=model.ihop,
        params=np.array(json.dumps(dict(p, ngrid=list(p["ngrid"]))), dtype=object).astype(str),
        kpath=np.array(path, dtype=np.float64), npts=np.int32(npts),
        qcoord=t[0, :, 0],          # REAL(qauxv(:,1)) as printed
        exk=t[:, :, 1],             # exk(N, nq)
        klabels=kl,
    )
    OUTDIR.mkdir(exist_ok=True)
    np.savez_compressed(OUTDIR / f"{name}.npz", **fx)
    return N, nq


def main():
    if not PROGRAM.exists():
        sys.exit("build the reference program first: make -C oracle")
    only = set(sys.argv[1:])
    for name, factory, p, path, npts in cases():
        if only and name not in only:
            continue
        N, nq = run_case(name, factory, p, path, npts)
        print(f"{name}: N={N}, {nq} q-points -> {(OUTDIR / (name + '.npz')).stat().st_size} bytes")


if __name__ == "__main__":
    main()
