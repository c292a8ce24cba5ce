"""Writes the chemistry-table fixtures tests/golden/chem_<name>.npz by running
the reference's own ChemtableCartesianLinear_single (oracle/_ref/chem_ref_dump,
made by `make -C oracle ref` from the mounted reference sources) on tables
written by myjoe_amd.chemtable.write_table.

    python tests/golden/make_chem_golden.py

Each fixture holds the inputs (axes X1, X2, X3, float32 DATA[n1][n2][n3][nvar],
NAMES, the points ZM, ZV, PC, and the sha256 of the table file the reference
read) and the reference's outputs: LOOKUP[n][nvar] (Lookup of every variable
by name), COEFF[n][10] (LookupCoeff) and SELECTED[n][5] (LookupSelectedCoeff)
where the table has that many variables.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = os.path.join(ROOT, "oracle", "_ref", "chem_ref_dump")

from myjoe_amd import chemtable  # noqa: E402


def points(x1, x2, x3, n_in, seed):
    """Random interior points, exact nodes, points with one exact coordinate,
    points outside the table on either side, and NaN / +-inf / -0.0."""
    rng = np.random.default_rng(seed)
    ax = (x1, x2, x3)
    cols = [[], [], []]

    def add(a, b, c):
        for col, v in zip(cols, (a, b, c)):
            col.append(np.asarray(v, np.float64).reshape(-1))

    add(*(rng.uniform(x[0], x[-1], n_in) for x in ax))
    add(*(rng.choice(x, 150) for x in ax))
    for d in range(3):
        p = [rng.uniform(x[0], x[-1], 80) for x in ax]
        p[d] = rng.choice(ax[d], 80)
        add(*p)
    for d in range(3):
        span = ax[d][-1] - ax[d][0]
        p = [rng.uniform(x[0], x[-1], 40) for x in ax]
        p[d] = np.concatenate([ax[d][0] - span * rng.uniform(0, 2, 20), ax[d][-1] + span * rng.uniform(0, 2, 20)])
        add(*p)
    mid = [0.5 * (x[0] + x[-1]) for x in ax]
    special = []
    for d in range(3):
        for v in (np.nan, np.inf, -np.inf, -0.0):
            q = list(mid)
            q[d] = v
            special.append(q)
    add(*np.array(special).T)
    return tuple(np.concatenate(c) for c in cols)


def table_b():
    """Six variables over a 4 x 3 x 2 table (the smallest third axis), with
    names that differ only by a space."""
    rng = np.random.default_rng(7)
    x1 = np.array([0.0, 0.1, 0.35, 1.0])
    x2 = np.array([-0.5, 0.0, 2.0])
    x3 = np.array([1e-3, 4.0])
    data = rng.standard_normal((4, 3, 2, 6)).astype(np.float32) * np.float32(1e3)
    data[1, 1, 0, 2] = np.float32(-0.0)
    data[2, 0, 1, :] = np.float32(0.0)
    return x1, x2, x3, data, ["rho", "T", "T ", "Y OH", "src", "gam"]


def main():
    if not os.path.exists(REF):
        raise SystemExit("oracle/_ref/chem_ref_dump is not built (make -C oracle ref, needs the reference tree)")
    x1, x2, x3, data, names = chemtable.synthetic_table(9, 5, 7, seed=1)
    data[0, 0, 0, 4] = np.float32(-0.0)
    data[3, 2, 1, :] = np.float32(-0.0)
    cases = {"fpva9x5x7": (x1, x2, x3, data, names), "six4x3x2": table_b()}
    for seed, (name, (x1, x2, x3, data, names)) in enumerate(cases.items()):
        zm, zv, pc = points(x1, x2, x3, 700, seed)
        with tempfile.TemporaryDirectory() as tmp:
            tab, pts, out = (os.path.join(tmp, s) for s in ("table.bin", "points.bin", "out.bin"))
            chemtable.write_table(tab, x1, x2, x3, data, names)
            with open(tab, "rb") as f:
                sha = hashlib.sha256(f.read()).hexdigest()
            with open(pts, "wb") as f:
                f.write(np.int64(zm.size).tobytes())
                f.write(np.concatenate([zm, zv, pc]).tobytes())
            r = subprocess.run([REF, tab, pts, out], capture_output=True, text=True)
            if r.returncode != 0:
                raise SystemExit(f"chem_ref_dump failed on {name}: {r.stderr}")
            got = np.fromfile(out, np.float64)
        n, nvar = zm.size, data.shape[3]
        fix = dict(X1=x1, X2=x2, X3=x3, DATA=data, NAMES=np.array("\n".join(names).encode()), ZM=zm, ZV=zv, PC=pc,
                   TABLE_SHA256=np.array(sha.encode()))
        k = n * nvar
        fix["LOOKUP"] = got[:k].reshape(n, nvar)
        if nvar >= 10:
            fix["COEFF"] = got[k:k + 10 * n].reshape(n, 10)
            k += 10 * n
        if nvar >= 5:
            fix["SELECTED"] = got[k:k + 5 * n].reshape(n, 5)
            k += 5 * n
        assert k == got.size, (name, k, got.size)
        np.savez_compressed(os.path.join(HERE, f"chem_{name}.npz"), **fix)
        print(f"chem_{name}: {n} points, {nvar} variables")


if __name__ == "__main__":
    main()
