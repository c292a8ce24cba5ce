"""FPVA chemistry tables in the reference's file format, and a synthetic
table of the reference's coefficient layout for tests and timing.

The file is what ChemtableCartesianLinear_single::Load reads
(mixing_combustion/ChemtableCartesianLinear_single.cpp:19-246): version 2.0
(double), combustion model and table type (64 characters each), reference
pressure, oxidizer and fuel temperatures (doubles), n1, n2, n3, nvar (int32),
the axes x1 (Zmean), x2 (Zvar), x3 (progress variable) as doubles, nvar
variable names (64 characters each), then the nodes as float32 in i, j, k,
variable order. The device lookup is joemi.Chemtable.
"""
from __future__ import annotations

import struct

import numpy as np

# LookupCoeff's outputs, in table order (ChemtableCartesianLinear_single.cpp:438-492)
COEFF_NAMES = ("RoM", "T0", "E0", "Gam0", "a_Gam", "mu0", "a_mu", "lamOcp0", "a_lamOcp", "src_prog")


def _name(s: str) -> bytes:
    b = s.encode()
    if len(b) > 63:
        raise ValueError(f"name {s!r} is longer than 63 bytes")
    return b + b"\0" * (64 - len(b))


def write_table(path, x1, x2, x3, data, names, model="FPVA", table_type="COEFF", pref=101325.0, toxi=300.0,
                tfuel=300.0):
    """Writes data[n1][n2][n3][nvar] (float32) over the axes in the reference's format."""
    x1, x2, x3 = (np.ascontiguousarray(x, np.float64) for x in (x1, x2, x3))
    data = np.ascontiguousarray(data, np.float32)
    n1, n2, n3, nvar = data.shape
    if (len(x1), len(x2), len(x3), len(names)) != (n1, n2, n3, nvar):
        raise ValueError("axes and names must match the data's shape")
    with open(path, "wb") as f:
        f.write(struct.pack("<d", 2.0))
        f.write(_name(model))
        f.write(_name(table_type))
        f.write(struct.pack("<3d", pref, toxi, tfuel))
        f.write(struct.pack("<4i", n1, n2, n3, nvar))
        for x in (x1, x2, x3):
            f.write(x.tobytes())
        for s in names:
            f.write(_name(s))
        f.write(data.tobytes())


def synthetic_table(n1=128, n2=32, n3=64, seed=0):
    """A table of the reference's ten coefficients over (Zmean, Zvar, C):
    Burke-Schumann-like temperature and composition profiles in Zmean
    (stoichiometric at 0.055), flattened by Zvar, scaled by the progress
    variable, with a seeded relative perturbation of 1e-3. Axes are
    non-uniform (Zmean clustered at the stoichiometric value, Zvar
    quadratic). Returns (x1, x2, x3, data[n1][n2][n3][10] float32, names)."""
    rng = np.random.default_rng(seed)
    zst = 0.055
    s = np.linspace(0.0, 1.0, n1)
    x1 = np.where(s < 0.5, zst * (2 * s) ** 2, zst + (1 - zst) * np.clip(2 * s - 1, 0.0, None) ** 1.5)
    x1[0], x1[-1] = 0.0, 1.0
    x2 = 0.25 * np.linspace(0.0, 1.0, n2) ** 2
    x3 = np.linspace(0.0, 1.0, n3)
    Z, V, C = np.meshgrid(x1, x2, x3, indexing="ij")
    width = 1.0 + 40.0 * V
    flame = np.where(Z <= zst, Z / zst, (1 - Z) / (1 - zst)) / width
    T = 300.0 + 1900.0 * C * flame
    W = 28.9 - 6.0 * Z + 2.0 * C * flame  # mean molecular weight
    RoM = 8314.46 / W
    gam = 1.4 - 0.1 * C * flame - 0.02 * Z
    out = np.empty(Z.shape + (10,))
    out[..., 0] = RoM
    out[..., 1] = T
    out[..., 2] = RoM * T / (gam - 1.0)
    out[..., 3] = gam
    out[..., 4] = -1e-4 * C * flame
    out[..., 5] = 1.8e-5 * (T / 300.0) ** 0.7
    out[..., 6] = 0.7
    out[..., 7] = 2.6e-5 * (T / 300.0) ** 0.7
    out[..., 8] = 0.7
    out[..., 9] = 50.0 * C * (1 - C) * flame * width
    out *= 1.0 + 1e-3 * rng.standard_normal(out.shape)
    return x1, x2, x3, out.astype(np.float32), list(COEFF_NAMES)
