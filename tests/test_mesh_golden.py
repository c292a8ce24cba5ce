"""libjoemi's host mesh pipeline against the reference's numbering and geometry.

The reference's Ugp/UgpWithCv init (redistReconnectReorder, buildNbocv,
calcGeometry, calcLsgCoeff) was run on the same faces by
tests/golden/make_golden.py; libjoemi must reproduce every array bit for
bit: node/face/cell numbering, zone order and ranges, connectivity, face and
cell geometry, and the least-squares coefficients. No GPU is used.
"""
import os

import numpy as np
import pytest

from myjoe_amd import joemi, meshgen
from tests.test_oracle_golden import load

INT_KEYS = ["COUNTS", "NOOFA_I", "NOOFA_V", "CVOFA", "FAOCV_I", "FAOCV_V", "NBOCV_I", "NBOCV_V", "ZONES"]
F64_KEYS = ["X_NO", "X_FA", "FA_NORMAL", "X_CV", "CV_VOLUME", "NBOCV_LSG", "BFA_LSG"]

MESHES = {
    "bump24": lambda: meshgen.bump_channel(24, 8),
    "bump96": lambda: meshgen.bump_channel(96, 32),
    "box4": lambda: meshgen.split_box(4, 0.5, 0.2, 3),
}


def check_same(mesh, g, name):
    d = mesh.as_dump()
    assert bytes(d["ZONE_NAMES"]) == bytes(g["ZONE_NAMES"]), name
    for k in INT_KEYS:
        a, b = np.asarray(d[k]).reshape(-1), np.asarray(g[k]).reshape(-1)
        if k == "COUNTS":
            # nno nfa nfa_b ncv ncv_g nbocv_s faocv_s noofa_s
            a, b = a[:8], b[:8]
        assert np.array_equal(a, b), f"{name}:{k}"
    for k in F64_KEYS:
        a, b = d[k].reshape(-1), g[k].reshape(-1)
        assert a.shape == b.shape, f"{name}:{k}"
        nbad = int(np.sum(a != b))
        assert nbad == 0, f"{name}:{k}: {nbad} of {a.size} differ, max |d| {np.max(np.abs(a - b)):.3e}"


@pytest.mark.parametrize("name", sorted(MESHES))
def test_from_faces_matches_reference(name):
    mesh = joemi.Mesh.from_faces(MESHES[name]())
    check_same(mesh, load("mesh_" + name), name)


@pytest.mark.parametrize("name", ["bump24", "box4"])
def test_read_msh_matches_reference(name, tmp_path):
    path = os.path.join(tmp_path, name + ".msh")
    meshgen.write_msh(MESHES[name](), path)
    mesh = joemi.Mesh.read_msh(path)
    check_same(mesh, load("mesh_" + name), name)


def test_zone_table():
    mesh = joemi.Mesh.from_faces(MESHES["bump24"]())
    zones = mesh.zones()
    # boundary zones sorted by name (FaZone::operator<, Ugp.h), then the fluid zone
    names = [z[0] for z in zones if z[1] == 1]
    assert names == sorted(names)
    covered = sum(z[3] - z[2] + 1 for z in zones if z[1] == 1)
    assert covered == mesh.nfa_b
    # consecutive boundary ranges starting at face 0
    f = 0
    for z in zones:
        if z[1] == 1:
            assert z[2] == f
            f = z[3] + 1


def test_geometry_closes():
    # sum of outward area vectors of every cell is zero; volumes add up
    mm = MESHES["box4"]()
    mesh = joemi.Mesh.from_faces(mm)
    d = mesh.as_dump()
    cvofa = d["CVOFA"].reshape(-1, 2)
    n = d["FA_NORMAL"].reshape(-1, 3)
    acc = np.zeros((mesh.ncv, 3))
    np.add.at(acc, cvofa[:, 0], n)
    inner = cvofa[:, 1] >= 0
    np.add.at(acc, cvofa[inner, 1], -n[inner])
    assert np.max(np.abs(acc)) < 1e-13
    assert abs(d["CV_VOLUME"].sum() - 1.0) < 1e-12  # the unit box
    sdn, neg = mesh.quality()
    assert sdn > 0.0 and neg == 0


def test_ls_coefficients_reproduce_linear_fields():
    # least-squares weights are exact for linear data: sum_k w_k (phi_k - phi_c) = grad phi
    mesh = joemi.Mesh.from_faces(MESHES["box4"]())
    d = mesh.as_dump()
    xcv = d["X_CV"].reshape(-1, 3)
    xfa = d["X_FA"].reshape(-1, 3)
    nbi, nbv = d["NBOCV_I"], d["NBOCV_V"]
    lsg = d["NBOCV_LSG"].reshape(-1, 3)
    blsg = d["BFA_LSG"].reshape(-1, 3)
    cvofa = d["CVOFA"].reshape(-1, 2)
    gvec = np.array([0.3, -1.7, 2.2])
    phi = xcv @ gvec + 0.5
    phib = xfa[:mesh.nfa_b] @ gvec + 0.5
    g = np.zeros((mesh.ncv, 3))
    for c in range(mesh.ncv):
        for k in range(nbi[c], nbi[c + 1]):
            g[c] += lsg[k] * phi[nbv[k]]
    for f in range(mesh.nfa_b):
        g[cvofa[f, 0]] += blsg[f] * phib[f]
    assert np.max(np.abs(g - gvec)) < 1e-10


@pytest.mark.parametrize("name", sorted(MESHES))
def test_from_ugp_keeps_reference_arrays(name):
    # jm_mesh_from_ugp takes the reference's own arrays over unchanged
    g = load("mesh_" + name)
    mesh = joemi.Mesh.from_ugp(g)
    assert (mesh.ncv, mesh.nfa, mesh.nfa_b) == tuple(int(x) for x in g["COUNTS"][[3, 1, 2]])
    for k in ["CVOFA", "FAOCV_I", "FAOCV_V", "NBOCV_I", "NBOCV_V"]:
        assert np.array_equal(mesh.array(k).reshape(-1), g[k].reshape(-1)), f"{name}:{k}"
    ncv = mesh.ncv
    for k in ["X_FA", "FA_NORMAL", "NBOCV_LSG", "BFA_LSG"]:
        assert np.array_equal(mesh.array(k).reshape(-1), g[k].reshape(-1)), f"{name}:{k}"
    assert np.array_equal(mesh.array("X_CV").reshape(-1), g["X_CV"].reshape(-1)[:3 * ncv])
    assert np.array_equal(mesh.array("CV_VOLUME").reshape(-1), g["CV_VOLUME"].reshape(-1)[:ncv])
    zt = g["ZONES"].reshape(-1, 3)
    assert [(z[1], z[2], z[3]) for z in mesh.zones()] == [tuple(int(v) for v in r) for r in zt]


def test_from_ugp_rejects_inconsistent_arrays():
    g = dict(load("mesh_box4"))
    cv = g["CVOFA"].copy().reshape(-1)
    cv[1] = 0  # a boundary face given a second cell
    g["CVOFA"] = cv
    with pytest.raises(joemi.JoemiError):
        joemi.Mesh.from_ugp(g)
