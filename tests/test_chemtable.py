"""The FPVA chemistry-table lookup (the reference's
ChemtableCartesianLinear_single): the oracle against the reference's own
outputs, the table file and its checks, and the device lookup against both.

Fixtures (tests/golden/chem_*.npz, written by make_chem_golden.py from the
reference's class built from its sources): a 9 x 5 x 7 table of the ten
LookupCoeff coefficients and a 4 x 3 x 2 table of six variables, each with
1222 points (interior, exact nodes, one exact coordinate, outside the table on
both sides, NaN, +-inf, -0.0) and the reference's Lookup of every variable by
name, LookupCoeff and LookupSelectedCoeff. The bar is bit-exact (NaN
included): the lookup is FP32 nodes with FP64 weights and sums in the
reference's order.
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from myjoe_amd import chemtable, joemi
from oracle.oracle import chemtable_lookup
from tests.test_oracle_golden import load

FIXTURES = ["fpva9x5x7", "six4x3x2"]


def same_bits(a, b):
    """Equal bit for bit where the values are numbers, NaN where the other is
    NaN (a NaN's sign and payload are the processor's, not the lookup's)."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.shape != b.shape:
        print("shape", a.shape, b.shape)
        return False
    na, nb = np.isnan(a), np.isnan(b)
    bad = (na != nb) | (~na & (a.view(np.int64) != b.view(np.int64)))
    if bad.any():
        idx = np.argwhere(bad)
        print(f"{bad.sum()} of {bad.size} differ; first {idx[:4].tolist()}: {a[tuple(idx[0])]!r} vs {b[tuple(idx[0])]!r}")
    return not bad.any()


def fixture(name):
    g = load("chem_" + name)
    names = bytes(g["NAMES"]).decode().split("\n")
    return g, names


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_matches_reference(name):
    g, names = fixture(name)
    nvar = len(names)
    got = chemtable_lookup(g["X1"], g["X2"], g["X3"], g["DATA"], g["ZM"], g["ZV"], g["PC"], range(nvar))
    assert same_bits(got.T, g["LOOKUP"])
    if "COEFF" in g:
        assert same_bits(got[:10].T, g["COEFF"])
    if "SELECTED" in g:
        assert same_bits(got[:5].T, g["SELECTED"])
    assert np.isnan(g["LOOKUP"]).any()  # the NaN points; -0.0 nodes give +0.0 (the sums start at 0.0)


def test_oracle_closed_forms():
    # pins the restatement beyond the fixtures: at a node the value is the
    # node's float; outside the table the clipped point's value; a table
    # linear in each coordinate is reproduced to rounding
    x1, x2, x3, data, _ = chemtable.synthetic_table(12, 6, 5, seed=3)
    rng = np.random.default_rng(0)
    i, j, k = rng.integers(0, 12, 300), rng.integers(0, 6, 300), rng.integers(0, 5, 300)
    got = chemtable_lookup(x1, x2, x3, data, x1[i], x2[j], x3[k], range(10))
    assert np.array_equal(got.T, data[i, j, k].astype(np.float64))
    zm, zv, pc = rng.uniform(-1, 2, 300), rng.uniform(-1, 1, 300), rng.uniform(-1, 2, 300)
    clip = chemtable_lookup(x1, x2, x3, data, np.clip(zm, 0, 1), np.clip(zv, 0, 0.25), np.clip(pc, 0, 1), range(10))
    assert np.array_equal(chemtable_lookup(x1, x2, x3, data, zm, zv, pc, range(10)), clip)
    Z, V, C = np.meshgrid(x1, x2, x3, indexing="ij")
    lin = (1.0 + 2.0 * Z) * (3.0 - V) * (0.5 + C)  # trilinear in the coordinates: interpolated exactly
    lin32 = lin.astype(np.float32)[..., None]
    zm, zv, pc = rng.uniform(0, 1, 500), rng.uniform(0, 0.25, 500), rng.uniform(0, 1, 500)
    want = (1.0 + 2.0 * zm) * (3.0 - zv) * (0.5 + pc)
    got = chemtable_lookup(x1, x2, x3, lin32, zm, zv, pc, [0])[0]
    assert np.max(np.abs(got - want) / want) < 1e-6  # the nodes' float rounding


@pytest.mark.parametrize("name", FIXTURES)
def test_table_file_is_what_the_reference_read(name, tmp_path):
    g, names = fixture(name)
    path = tmp_path / "table.bin"
    chemtable.write_table(path, g["X1"], g["X2"], g["X3"], g["DATA"], names)
    assert hashlib.sha256(path.read_bytes()).hexdigest() == bytes(g["TABLE_SHA256"]).decode()


def _write(tmp_path, **kw):
    x1, x2, x3, data, names = chemtable.synthetic_table(4, 3, 3, seed=0)
    args = dict(x1=x1, x2=x2, x3=x3, data=data, names=names)
    args.update(kw)
    path = tmp_path / "t.bin"
    chemtable.write_table(path, **args)
    return path


def test_table_checked_before_device(tmp_path):
    # a file that is not a version-2.0 table is a parameter error, a bad axis
    # an argument error; both before any device call (no GPU needed)
    path = _write(tmp_path)
    raw = bytearray(path.read_bytes())
    bad = tmp_path / "v1.bin"
    bad.write_bytes(np.float64(1.0).tobytes() + bytes(raw[8:]))
    for p, status, needle in [(bad, 3, "version"), (tmp_path / "none.bin", 3, "cannot open")]:
        with pytest.raises(joemi.JoemiError) as e:
            joemi.Chemtable.load(p)
        assert e.value.status == status and needle in str(e.value)
    short = tmp_path / "short.bin"
    short.write_bytes(bytes(raw[:-40]))
    with pytest.raises(joemi.JoemiError) as e:
        joemi.Chemtable.load(short)
    assert e.value.status == 3 and "shorter" in str(e.value)
    x1, x2, x3, data, names = chemtable.synthetic_table(4, 3, 3, seed=0)
    for args, needle in [((x1[::-1].copy(), x2, x3, data), "strictly increasing"),
                         ((x1, x2, x3[:1].copy(), data[:, :, :1].copy()), "at least two"),
                         ((x1, x2, np.array([0.0, np.nan, 1.0]), data), "finite")]:
        with pytest.raises(joemi.JoemiError) as e:
            joemi.Chemtable.create(*args)
        assert e.value.status == 1 and needle in str(e.value), str(e.value)
    with pytest.raises(joemi.JoemiError) as e:
        joemi.Chemtable.create(x1, x2, x3, data, names[:-1] + [names[0]])
    assert e.value.status == 1 and "twice" in str(e.value)


# ---------------------------------------------------------------------------
# the device lookup
# ---------------------------------------------------------------------------


@pytest.mark.gpu
@pytest.mark.parametrize("name", FIXTURES)
def test_device_lookup_matches_reference(name, tmp_path):
    print(f"chemtable golden {name}", flush=True)
    g, names = fixture(name)
    path = tmp_path / "table.bin"
    chemtable.write_table(path, g["X1"], g["X2"], g["X3"], g["DATA"], names)
    t = joemi.Chemtable.load(path)
    assert (t.n1, t.n2, t.n3, t.nvar) == g["DATA"].shape
    pts = (g["ZM"], g["ZV"], g["PC"])
    by_name = t.lookup(*pts, names)  # Lookup(tag) of each variable
    assert same_bits(by_name.T, g["LOOKUP"])
    if "T " in names:  # names match exactly, trailing space included
        assert t.var("T ") == names.index("T ") and t.var("T") == names.index("T")
    if "COEFF" in g:
        assert same_bits(t.lookup(*pts, t.COEFF).T, g["COEFF"])
    if "SELECTED" in g:
        assert same_bits(t.lookup(*pts, t.SELECTED).T, g["SELECTED"])
    # any list of variables (the one-at-a-time form), and each single one
    nvar = len(names)
    order = [nvar - 1, 0, nvar // 2]
    assert same_bits(t.lookup(*pts, order).T, g["LOOKUP"][:, order])
    for v in range(nvar):
        assert same_bits(t.lookup(*pts, [v])[0], g["LOOKUP"][:, v])
    with pytest.raises(joemi.JoemiError) as e:
        t.var("no such variable")
    assert e.value.status == 1


@pytest.mark.gpu
def test_device_lookup_full_size_table():
    # the 128 x 32 x 64 table of ten coefficients: 2M points against the
    # oracle bit for bit; the nodes return their floats
    print("chemtable 128x32x64, 2M points", flush=True)
    x1, x2, x3, data, names = chemtable.synthetic_table(128, 32, 64, seed=0)
    t = joemi.Chemtable.create(x1, x2, x3, data, names)
    rng = np.random.default_rng(5)
    n = 1 << 21
    zm, zv, pc = rng.uniform(-0.05, 1.05, n), rng.uniform(0, 0.26, n), rng.uniform(0, 1, n)
    got = t.lookup(zm, zv, pc, t.COEFF)
    want = chemtable_lookup(x1, x2, x3, data, zm, zv, pc, range(10))
    assert same_bits(got, want)
    i, j, k = rng.integers(0, 128, n), rng.integers(0, 32, n), rng.integers(0, 64, n)
    nodes = t.lookup(x1[i], x2[j], x3[k], t.COEFF)
    assert np.array_equal(nodes.T, data[i, j, k].astype(np.float64))


@pytest.mark.gpu
def test_device_lookup_long_axes_and_wide_rows():
    # axes too long for the LDS copy (n1 + n2 + n3 > 4096) and a row of 19
    # variables (a selection wider than four float4 vectors runs per variable)
    print("chemtable long axes", flush=True)
    rng = np.random.default_rng(11)
    x1 = np.cumsum(rng.uniform(0.5, 1.5, 4000))
    x2 = np.array([0.0, 1.0, 3.0])
    x3 = np.array([-1.0, 0.0, 0.5, 2.0])
    data = rng.standard_normal((4000, 3, 4, 19)).astype(np.float32)
    t = joemi.Chemtable.create(x1, x2, x3, data)
    n = 50000
    zm = rng.uniform(x1[0] - 10, x1[-1] + 10, n)
    zv, pc = rng.uniform(-0.5, 3.5, n), rng.uniform(-1.5, 2.5, n)
    for vars in (range(19), range(3, 19), [18, 2, 7], [5]):
        vars = list(vars)
        assert same_bits(t.lookup(zm, zv, pc, vars), chemtable_lookup(x1, x2, x3, data, zm, zv, pc, vars)), vars


# ---------------------------------------------------------------------------
# the C++ mirror of the reference's class (myjoe_amd/host/ChemtableCartesianLinear_single.hpp)
# ---------------------------------------------------------------------------

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MIRROR_PROGRAM = r"""
#include <cstdio>
#include <vector>
#include "myjoe_amd/host/ChemtableCartesianLinear_single.hpp"

// a caller written against the reference's class: Load, the accessors, one
// point at a time by name and LookupCoeff, then the batched form
int main(int argc, char **argv) {
  ChemtableCartesianLinear_single tab;
  try {
    tab.Load(argv[1]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 3;
  }
  FILE *fp = std::fopen(argv[2], "rb");
  long n = 0;
  if (std::fread(&n, sizeof n, 1, fp) != 1) return 4;
  std::vector<double> a(3 * n);
  if (std::fread(a.data(), sizeof(double), 3 * n, fp) != (size_t)(3 * n)) return 4;
  std::fclose(fp);
  const int nvar = tab.GetChemtableDimension4(), np = n < 64 ? n : 64;
  std::vector<double> one, coeff(10 * (size_t)n);
  for (int p = 0; p < np; p++)
    for (int l = 0; l < nvar; l++) one.push_back(tab.Lookup(a[p], a[n + p], a[2 * n + p], tab.GetChemtableVariables(l)));
  for (int p = 0; p < np; p++) {
    double c[10];
    tab.LookupCoeff(c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], a[p], a[n + p], a[2 * n + p]);
    one.insert(one.end(), c, c + 10);
  }
  tab.LookupCoeff((int)n, a.data(), a.data() + n, a.data() + 2 * n, coeff.data());
  fp = std::fopen(argv[3], "wb");
  double meta[5] = {(double)tab.GetChemtableDimension1(), tab.GetChemtableCoordinate1(1),
                    tab.GetChemtableCoordinate3(tab.GetChemtableDimension3()), tab.GetReferencePressure(),
                    (double)(tab.GetCombustionModel() == "FPVA")};
  std::fwrite(meta, sizeof(double), 5, fp);
  std::fwrite(one.data(), sizeof(double), one.size(), fp);
  std::fwrite(coeff.data(), sizeof(double), coeff.size(), fp);
  std::fclose(fp);
  return 0;
}
"""


def _mirror_program(tmp_path):
    src = tmp_path / "mirror.cpp"
    src.write_text(MIRROR_PROGRAM)
    exe = tmp_path / "mirror"
    lib = os.path.join(ROOT, "myjoe_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", f"-I{ROOT}", str(src), f"-L{lib}", "-ljoemi",
                        f"-Wl,-rpath,{lib}", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _mirror_inputs(tmp_path, name):
    g, names = fixture(name)
    table = tmp_path / "table.bin"
    chemtable.write_table(table, g["X1"], g["X2"], g["X3"], g["DATA"], names)
    pts = tmp_path / "points.bin"
    pts.write_bytes(np.int64(g["ZM"].size).tobytes() + np.concatenate([g["ZM"], g["ZV"], g["PC"]]).tobytes())
    return g, names, table, pts


def test_mirror_compiles_and_fails_loudly_without_device(tmp_path):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    exe = _mirror_program(tmp_path)
    _, _, table, pts = _mirror_inputs(tmp_path, "fpva9x5x7")
    r = subprocess.run([str(exe), str(table), str(pts), str(tmp_path / "out.bin")], capture_output=True, text=True)
    assert r.returncode == 3 and "HIP" in r.stderr, (r.returncode, r.stderr)


@pytest.mark.gpu
def test_mirror_class_matches_reference(tmp_path):
    print("chemtable C++ mirror", flush=True)
    exe = _mirror_program(tmp_path)
    g, names, table, pts = _mirror_inputs(tmp_path, "fpva9x5x7")
    out = tmp_path / "out.bin"
    r = subprocess.run([str(exe), str(table), str(pts), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.float64)
    n, nvar, np_ = g["ZM"].size, len(names), min(64, g["ZM"].size)
    meta, got = got[:5], got[5:]
    assert meta.tolist() == [9.0, g["X1"][0], g["X3"][-1], 101325.0, 1.0]
    one = got[:np_ * nvar].reshape(np_, nvar)
    coeff1 = got[np_ * nvar:np_ * (nvar + 10)].reshape(np_, 10)
    batch = got[np_ * (nvar + 10):].reshape(10, n)
    assert same_bits(one, g["LOOKUP"][:np_])
    assert same_bits(coeff1, g["COEFF"][:np_])
    assert same_bits(batch.T, g["COEFF"])
