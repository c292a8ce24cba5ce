"""Synthetic unstructured meshes for the steady-solve workloads.

Every generator returns a :class:`MshMesh`: node coordinates, faces with node
lists and the two adjacent cells, and face zones, all in the order and
orientation a Fluent ASCII ``.msh`` file would hold them *after* the Joe reader
has applied its flip rule (``MshFilter.cpp:453-472``): the face normal given by
the node order (right-hand rule) points from ``cells[f, 0]`` to ``cells[f, 1]``
and boundary faces carry ``-1`` in the second slot.

``write_msh`` emits the file the reference reads; ``myjoe_amd.mesh`` builds the
solver numbering from the same arrays without going through text.

Workloads
---------
* ``bump_channel``: 2D Euler channel with a 10 %-thick circular-arc bump on the
  lower wall, one cell thick in z (96x32x1 in the bump-channel steady solve).
* ``naca_omesh``: O-mesh of ni x nj hexes around a NACA 4-digit airfoil (closed
  trailing edge, chord 1) out to a circle of radius 20, one hex layer in z,
  geometric radial stretching (the serving workload's airfoils).
* ``flat_plate``: zero-pressure-gradient flat plate, x in [-0.33, 2] with the
  plate from x = 0, height 1, cells stretched geometrically from the wall
  (ratio 1.1 in the implicit RANS flat-plate solve) and towards the leading
  edge, nz hex layers in z.
* ``split_box``: [0,1]^3 box of n^3 hexes, a seeded fraction split into six
  Kuhn tetrahedra each (globally consistent diagonals), interior nodes
  jittered uniformly by +-jitter*h (160^3, 50 %, 0.2 in the single-residual
  throughput test).
"""
from __future__ import annotations

import dataclasses
import numpy as np

__all__ = ["MshMesh", "bump_channel", "flat_plate", "plate_spacing", "split_box", "naca_omesh", "write_msh"]


@dataclasses.dataclass
class MshMesh:
    x_no: np.ndarray          # (nno, 3) float64
    face_ptr: np.ndarray      # (nfa+1,) int64 CSR into face_nodes
    face_nodes: np.ndarray    # (sum nno_f,) int32, 0-based
    face_cells: np.ndarray    # (nfa, 2) int32, [icv0, icv1]; icv1 = -1 on boundary
    face_zone: np.ndarray     # (nfa,) int32 zone id, faces of a zone contiguous
    zones: list               # [(zone_id, fluent_type, name)], file order
    ncv: int

    @property
    def nfa(self) -> int:
        return int(self.face_cells.shape[0])

    @property
    def nno(self) -> int:
        return int(self.x_no.shape[0])


# hex corner order: v000, v100, v110, v010, v001, v101, v111, v011
_HEX_FACES = {  # outward-normal node order, by (axis, side)
    (0, 0): [0, 4, 7, 3], (0, 1): [1, 2, 6, 5],
    (1, 0): [0, 1, 5, 4], (1, 1): [3, 7, 6, 2],
    (2, 0): [0, 3, 2, 1], (2, 1): [4, 5, 6, 7],
}


def _assemble(x_no, faces, ncv, zones):
    """faces: list of (nodes (m,k) int, cells (m,2) int, zone_id) blocks in file order."""
    counts = np.concatenate([np.full(f[0].shape[0], f[0].shape[1], np.int64) for f in faces])
    ptr = np.zeros(counts.size + 1, np.int64)
    np.cumsum(counts, out=ptr[1:])
    nodes = np.concatenate([f[0].reshape(-1) for f in faces]).astype(np.int32)
    cells = np.concatenate([f[1] for f in faces]).astype(np.int32)
    zone = np.concatenate([np.full(f[0].shape[0], f[2], np.int32) for f in faces])
    return MshMesh(np.ascontiguousarray(x_no, np.float64), ptr, nodes, cells, zone, zones, int(ncv))


def bump_channel(nx: int = 96, ny: int = 32, thickness: float = 0.1, dz: float | None = None) -> MshMesh:
    """Channel x in [0,3], y in [0,1] with a circular-arc bump of the given
    thickness between x=1 and x=2 on the lower wall; one hex layer in z.

    Zones: inlet (x=0), outlet (x=3), bump_wall (lower wall), top, zsym (both
    z faces)."""
    if dz is None:
        dz = 3.0 / nx
    chord = 1.0
    R = (0.25 * chord * chord + thickness * thickness) / (2.0 * thickness)
    xs = np.linspace(0.0, 3.0, nx + 1)
    ywall = np.zeros_like(xs)
    inb = (xs > 1.0) & (xs < 2.0)
    ywall[inb] = np.sqrt(R * R - (xs[inb] - 1.5) ** 2) - (R - thickness)
    jj = np.arange(ny + 1) / ny
    X = np.empty((2, ny + 1, nx + 1, 3))
    for k in range(2):
        X[k, :, :, 0] = xs[None, :]
        X[k, :, :, 1] = ywall[None, :] + (1.0 - ywall[None, :]) * jj[:, None]
        X[k, :, :, 2] = k * dz
    x_no = X.reshape(-1, 3)

    def nid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i

    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")  # cell (j, i)
    I = I.reshape(-1)
    J = J.reshape(-1)
    cid = J * nx + I
    corners = np.stack([nid(I, J, 0), nid(I + 1, J, 0), nid(I + 1, J + 1, 0), nid(I, J + 1, 0),
                        nid(I, J, 1), nid(I + 1, J, 1), nid(I + 1, J + 1, 1), nid(I, J + 1, 1)], 1)

    # interior faces, cell-major: +x then +y face of each cell
    blocks_nodes, blocks_cells, order_key = [], [], []
    m = I < nx - 1
    blocks_nodes.append(corners[m][:, _HEX_FACES[(0, 1)]])
    blocks_cells.append(np.stack([cid[m], cid[m] + 1], 1))
    order_key.append(cid[m] * 2)
    m = J < ny - 1
    blocks_nodes.append(corners[m][:, _HEX_FACES[(1, 1)]])
    blocks_cells.append(np.stack([cid[m], cid[m] + nx], 1))
    order_key.append(cid[m] * 2 + 1)
    key = np.concatenate(order_key)
    o = np.argsort(key, kind="stable")
    int_nodes = np.concatenate(blocks_nodes)[o]
    int_cells = np.concatenate(blocks_cells)[o]

    def bnd(mask, axis, side):
        return corners[mask][:, _HEX_FACES[(axis, side)]], np.stack([cid[mask], -np.ones(mask.sum(), np.int64)], 1)

    zones = [(3, "interior", "int_fluid"), (4, "velocity-inlet", "inlet"), (5, "pressure-outlet", "outlet"),
             (6, "wall", "bump_wall"), (7, "wall", "top"), (8, "symmetry", "zsym")]
    zn_i, zc_i = bnd(I == 0, 0, 0)
    zn_o, zc_o = bnd(I == nx - 1, 0, 1)
    zn_w, zc_w = bnd(J == 0, 1, 0)
    zn_t, zc_t = bnd(J == ny - 1, 1, 1)
    all_c = np.ones_like(I, bool)
    zn_z0, zc_z0 = bnd(all_c, 2, 0)
    zn_z1, zc_z1 = bnd(all_c, 2, 1)
    faces = [(int_nodes, int_cells, 3), (zn_i, zc_i, 4), (zn_o, zc_o, 5), (zn_w, zc_w, 6),
             (zn_t, zc_t, 7), (np.concatenate([zn_z0, zn_z1]), np.concatenate([zc_z0, zc_z1]), 8)]
    return _assemble(x_no, faces, nx * ny, zones)


def _geometric(n: int, length: float, ratio: float) -> np.ndarray:
    """n + 1 coordinates from 0 to length whose spacing grows by ratio."""
    if abs(ratio - 1.0) < 1e-14:
        return np.linspace(0.0, length, n + 1)
    d0 = length * (ratio - 1.0) / (ratio ** n - 1.0)
    x = np.concatenate([[0.0], np.cumsum(d0 * ratio ** np.arange(n))])
    x[-1] = length
    return x


def plate_spacing(ny: int, ratio: float = 1.1, height: float = 1.0) -> float:
    """Height of the first cell above the plate."""
    return float(_geometric(ny, height, ratio)[1])


def flat_plate(nx_up: int, nx: int, ny: int, nz: int = 1, ratio: float = 1.1, x_up: float = 0.33,
               length: float = 2.0, height: float = 1.0, ratio_x: float = 1.05, dz: float | None = None) -> MshMesh:
    """Zero-pressure-gradient flat plate: x in [-x_up, length] with the plate
    from x = 0, y in [0, height], nz hex layers in z. Cells grow from the
    wall by ratio in y, from the leading edge by ratio_x along the plate and
    towards the inlet upstream of it.

    Zones: inlet (x = -x_up), outlet (x = length), symm_up (y = 0, x < 0),
    plate (y = 0, x > 0), top (y = height), zsym (both z faces)."""
    xp = _geometric(nx, length, ratio_x)
    xu = -_geometric(nx_up, x_up, ratio_x)[::-1]
    xs = np.concatenate([xu[:-1], xp])
    nxt = nx_up + nx
    ys = _geometric(ny, height, ratio)
    if dz is None:
        dz = length / nx
    zs = dz * np.arange(nz + 1)
    X = np.empty((nz + 1, ny + 1, nxt + 1, 3))
    X[..., 0] = xs[None, None, :]
    X[..., 1] = ys[None, :, None]
    X[..., 2] = zs[:, None, None]
    x_no = X.reshape(-1, 3)

    def nid(i, j, k):
        return (k * (ny + 1) + j) * (nxt + 1) + i

    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nxt), indexing="ij")
    I, J, K = I.reshape(-1), J.reshape(-1), K.reshape(-1)
    cid = (K * ny + J) * nxt + I
    corners = np.stack([nid(I, J, K), nid(I + 1, J, K), nid(I + 1, J + 1, K), nid(I, J + 1, K),
                        nid(I, J, K + 1), nid(I + 1, J, K + 1), nid(I + 1, J + 1, K + 1), nid(I, J + 1, K + 1)], 1)
    # interior faces, cell-major: +x, +y, +z face of each cell
    blocks_nodes, blocks_cells, order_key = [], [], []
    for axis, m, step in ((0, I < nxt - 1, 1), (1, J < ny - 1, nxt), (2, K < nz - 1, nxt * ny)):
        blocks_nodes.append(corners[m][:, _HEX_FACES[(axis, 1)]])
        blocks_cells.append(np.stack([cid[m], cid[m] + step], 1))
        order_key.append(cid[m] * 3 + axis)
    key = np.concatenate(order_key)
    o = np.argsort(key, kind="stable")
    int_nodes = np.concatenate(blocks_nodes)[o]
    int_cells = np.concatenate(blocks_cells)[o]

    def bnd(mask, axis, side):
        return corners[mask][:, _HEX_FACES[(axis, side)]], np.stack([cid[mask], -np.ones(mask.sum(), np.int64)], 1)

    zones = [(3, "interior", "int_fluid"), (4, "velocity-inlet", "inlet"), (5, "pressure-outlet", "outlet"),
             (6, "symmetry", "symm_up"), (7, "wall", "plate"), (8, "pressure-far-field", "top"),
             (9, "symmetry", "zsym")]
    zi = bnd(I == 0, 0, 0)
    zo = bnd(I == nxt - 1, 0, 1)
    zu = bnd((J == 0) & (I < nx_up), 1, 0)
    zp = bnd((J == 0) & (I >= nx_up), 1, 0)
    zt = bnd(J == ny - 1, 1, 1)
    z0 = bnd(K == 0, 2, 0)
    z1 = bnd(K == nz - 1, 2, 1)
    faces = [(int_nodes, int_cells, 3), (zi[0], zi[1], 4), (zo[0], zo[1], 5), (zu[0], zu[1], 6),
             (zp[0], zp[1], 7), (zt[0], zt[1], 8), (np.concatenate([z0[0], z1[0]]), np.concatenate([z0[1], z1[1]]), 9)]
    return _assemble(x_no, faces, nxt * ny * nz, zones)


def naca_surface(psi: np.ndarray, thickness: float, camber: float, camber_pos: float = 0.4,
                 closed_te: bool = False):
    """Points of a NACA 4-digit section at the angle parameter psi in
    [0, 2 pi): x = (1 + cos psi) / 2, the lower surface for psi in [0, pi],
    the upper one after it (clockwise seen from +z). The standard open
    (blunt) trailing edge by default: its base is the O-mesh edge between the
    last upper and the first lower point, which avoids the cusp cells of a
    closed one."""
    xc = 0.5 * (1.0 + np.cos(psi))
    t, m, p = thickness, camber, camber_pos
    c4 = -0.1036 if closed_te else -0.1015
    yt = 5.0 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2 + 0.2843 * xc ** 3 + c4 * xc ** 4)
    if m > 0.0:
        yc = np.where(xc < p, m / p ** 2 * (2 * p * xc - xc ** 2), m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * xc - xc ** 2))
        dy = np.where(xc < p, 2 * m / p ** 2 * (p - xc), 2 * m / (1 - p) ** 2 * (p - xc))
    else:
        yc = np.zeros_like(xc)
        dy = np.zeros_like(xc)
    th = np.arctan(dy)
    upper = psi > np.pi
    sgn = np.where(upper, 1.0, -1.0)
    return xc - sgn * yt * np.sin(th), yc + sgn * yt * np.cos(th)


def naca_omesh(ni: int = 200, nj: int = 100, thickness: float = 0.12, camber: float = 0.0,
               camber_pos: float = 0.4, radius: float = 20.0, first: float = 2e-3, dz: float = 0.05) -> MshMesh:
    """O-mesh around a NACA 4-digit airfoil, ni cells around, nj cells out,
    one hex layer in z. Zones: airfoil (the section), farfield (circle of the
    given radius about x = 0.5), zsym (both z faces). ncv = ni * nj."""
    psi = 2.0 * np.pi * np.arange(ni) / ni
    xs, ys = naca_surface(psi, thickness, camber, camber_pos)
    xf = 0.5 + radius * np.cos(-psi)
    yf = radius * np.sin(-psi)
    # geometric radial spacing: s_1 * radius ~ first
    target = first / radius
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        if (r - 1.0) / (r ** nj - 1.0) > target:
            lo = r
        else:
            hi = r
    r = 0.5 * (lo + hi)
    sj = (r ** np.arange(nj + 1) - 1.0) / (r ** nj - 1.0)
    X = np.empty((2, nj + 1, ni, 3))
    for k in range(2):
        X[k, :, :, 0] = (1.0 - sj[:, None]) * xs[None, :] + sj[:, None] * xf[None, :]
        X[k, :, :, 1] = (1.0 - sj[:, None]) * ys[None, :] + sj[:, None] * yf[None, :]
        X[k, :, :, 2] = k * dz
    x_no = X.reshape(-1, 3)

    def nid(i, j, k):
        return (k * (nj + 1) + j) * ni + (i % ni)

    I, J = np.meshgrid(np.arange(ni), np.arange(nj), indexing="xy")
    I = I.reshape(-1)
    J = J.reshape(-1)
    cid = J * ni + I
    corners = np.stack([nid(I, J, 0), nid(I + 1, J, 0), nid(I + 1, J + 1, 0), nid(I, J + 1, 0),
                        nid(I, J, 1), nid(I + 1, J, 1), nid(I + 1, J + 1, 1), nid(I, J + 1, 1)], 1)
    blocks_nodes, blocks_cells, order_key = [], [], []
    # +i face of every cell (periodic around the section)
    blocks_nodes.append(corners[:, _HEX_FACES[(0, 1)]])
    blocks_cells.append(np.stack([cid, J * ni + (I + 1) % ni], 1))
    order_key.append(cid * 2)
    m = J < nj - 1
    blocks_nodes.append(corners[m][:, _HEX_FACES[(1, 1)]])
    blocks_cells.append(np.stack([cid[m], cid[m] + ni], 1))
    order_key.append(cid[m] * 2 + 1)
    key = np.concatenate(order_key)
    o = np.argsort(key, kind="stable")
    int_nodes = np.concatenate(blocks_nodes)[o]
    int_cells = np.concatenate(blocks_cells)[o]

    def bnd(mask, axis, side):
        return corners[mask][:, _HEX_FACES[(axis, side)]], np.stack([cid[mask], -np.ones(mask.sum(), np.int64)], 1)

    zones = [(3, "interior", "int_fluid"), (4, "wall", "airfoil"), (5, "pressure-far-field", "farfield"),
             (6, "symmetry", "zsym")]
    zn_w, zc_w = bnd(J == 0, 1, 0)
    zn_f, zc_f = bnd(J == nj - 1, 1, 1)
    all_c = np.ones_like(I, bool)
    zn_z0, zc_z0 = bnd(all_c, 2, 0)
    zn_z1, zc_z1 = bnd(all_c, 2, 1)
    faces = [(int_nodes, int_cells, 3), (zn_w, zc_w, 4), (zn_f, zc_f, 5),
             (np.concatenate([zn_z0, zn_z1]), np.concatenate([zc_z0, zc_z1]), 6)]
    return _assemble(x_no, faces, ni * nj, zones)


# Kuhn tetrahedra of the unit cube: permutation of the axes visited on the
# path 000 -> 111.
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
_PERM_INDEX = {p: i for i, p in enumerate(_PERMS)}


def _kuhn_tet_vertices(p):
    v = [np.zeros(3, np.int64)]
    for a in p:
        w = v[-1].copy()
        w[a] = 1
        v.append(w)
    return v


def _tri_tables():
    """For each cube face (axis d, side s) and triangle t in {0,1}: local corner
    offsets of the 3 vertices and the Kuhn permutation index owning it."""
    tab = {}
    for d in range(3):
        a, b = [x for x in range(3) if x != d]
        for s in range(2):
            base = np.zeros(3, np.int64)
            base[d] = s
            tris = []
            for t in range(2):
                first, second = (a, b) if t == 0 else (b, a)
                v0 = base.copy()
                v1 = base.copy()
                v1[first] = 1
                v2 = base.copy()
                v2[a] = 1
                v2[b] = 1
                perm = (first, second, d) if s == 0 else (d, first, second)
                tris.append((np.stack([v0, v1, v2]), _PERM_INDEX[perm]))
            tab[(d, s)] = tris
    return tab


def _inner_tri_table():
    """The six faces shared by Kuhn tets of one cube: (vertices, perm_a, perm_b)."""
    out = []
    seen = set()
    for i, p in enumerate(_PERMS):
        for pos in range(2):
            q = list(p)
            q[pos], q[pos + 1] = q[pos + 1], q[pos]
            j = _PERM_INDEX[tuple(q)]
            key = (min(i, j), max(i, j))
            if key in seen:
                continue
            seen.add(key)
            vp = _kuhn_tet_vertices(p)
            vq = {tuple(x) for x in _kuhn_tet_vertices(tuple(q))}
            shared = [v for v in vp if tuple(v) in vq]
            out.append((np.stack(shared), key[0], key[1]))
    return out


def split_box(n: int, frac: float = 0.5, jitter: float = 0.2, seed: int = 0) -> MshMesh:
    """[0,1]^3 box of n^3 hexes; a seeded ``frac`` of them split into 6 Kuhn
    tets, interior nodes jittered by U(-jitter*h, jitter*h) per coordinate.

    Zones: x0 'inlet', x1 'outlet', y0 'wall', y1 'farfield', z0/z1 'symm'.
    Face file order: interior faces grouped by the lower cube (its inner tet
    faces, then its +x, +y, +z faces), then each boundary zone in cube order.
    """
    rng = np.random.default_rng(seed)
    N = n + 1
    h = 1.0 / n
    g = np.arange(N, dtype=np.float64) * h
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    x_no = np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], 1)
    if jitter > 0:
        ijk = np.stack(np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij"), -1).reshape(-1, 3)
        interior = np.all((ijk > 0) & (ijk < n), axis=1)
        d = rng.uniform(-jitter * h, jitter * h, size=(int(interior.sum()), 3))
        x_no[interior] += d
    ncube = n ** 3
    split = rng.random(ncube) < frac
    cpc = np.where(split, 6, 1)
    cstart = np.zeros(ncube + 1, np.int64)
    np.cumsum(cpc, out=cstart[1:])
    ncv = int(cstart[-1])
    cube = np.arange(ncube, dtype=np.int64)
    ci = cube % n
    cj = (cube // n) % n
    ck = cube // (n * n)

    def nid(i, j, k):
        return (k * N + j) * N + i

    def cell_of(cb, perm):
        return np.where(split[cb], cstart[cb] + perm, cstart[cb])

    corner_off = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                           [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    tri_tab = _tri_tables()
    inner_tab = _inner_tri_table()
    pos_int = np.stack([ci, cj, ck], 1)

    def local_nodes(cb, offs):
        # offs: (k,3) corner offsets -> (m,k) global node ids
        return nid(ci[cb][:, None] + offs[None, :, 0], cj[cb][:, None] + offs[None, :, 1],
                   ck[cb][:, None] + offs[None, :, 2])

    def orient(nodes, frm, to):
        """Flip node order where the right-hand normal (on the unjittered grid)
        does not point from point `frm` to point `to` (both (m,3), grid units)."""
        p = np.stack([nodes % N, (nodes // N) % N, nodes // (N * N)], -1).astype(np.float64)
        nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        flip = np.einsum("ij,ij->i", nrm, to - frm) < 0
        out = nodes.copy()
        out[flip] = nodes[flip][:, ::-1]
        return out

    def tet_centroid(cb, perm):
        v = np.stack(_kuhn_tet_vertices(_PERMS[perm])).mean(0)
        return pos_int[cb] + v[None, :]

    blocks = []  # (nodes, cells, key)
    # 1) inner tet faces of split cubes
    sc = cube[split]
    for slot, (verts, pa, pb) in enumerate(inner_tab):
        nodes = local_nodes(sc, verts)
        nodes = orient(nodes, tet_centroid(sc, pa), tet_centroid(sc, pb))
        cells = np.stack([cstart[sc] + pa, cstart[sc] + pb], 1)
        blocks.append((nodes, cells, sc * 32 + slot))
    # 2) faces to the +d neighbour
    for d in range(3):
        coord = (ci, cj, ck)[d]
        m = coord < n - 1
        ca = cube[m]
        cb_ = ca + (1, n, n * n)[d]
        both_hex = ~split[ca] & ~split[cb_]
        # quads
        qa = ca[both_hex]
        qb = cb_[both_hex]
        nodes = local_nodes(qa, corner_off[_HEX_FACES[(d, 1)]])
        blocks.append((nodes, np.stack([cstart[qa], cstart[qb]], 1), qa * 32 + 8 + 3 * d))
        # triangles
        ta = ca[~both_hex]
        tb = cb_[~both_hex]
        for t, (verts, perm_a) in enumerate(tri_tab[(d, 1)]):
            perm_b = tri_tab[(d, 0)][t][1]
            nodes = local_nodes(ta, verts)
            ctr_a = np.where(split[ta][:, None], tet_centroid(ta, perm_a), pos_int[ta] + 0.5)
            ctr_b = np.where(split[tb][:, None], tet_centroid(tb, perm_b), pos_int[tb] + 0.5)
            nodes = orient(nodes, ctr_a, ctr_b)
            cells = np.stack([cell_of(ta, perm_a), cell_of(tb, perm_b)], 1)
            blocks.append((nodes, cells, ta * 32 + 8 + 3 * d + 1 + t))
    keys = np.concatenate([b[2] for b in blocks])
    o = np.argsort(keys, kind="stable")
    # pad triangles to 4 with -1 so blocks stack, then split back
    def pad4(a):
        if a.shape[1] == 4:
            return a
        return np.concatenate([a, -np.ones((a.shape[0], 1), a.dtype)], 1)
    int_nodes = np.concatenate([pad4(b[0]) for b in blocks])[o]
    int_cells = np.concatenate([b[1] for b in blocks])[o]

    zones = [(3, "interior", "int_fluid"), (4, "velocity-inlet", "inlet"), (5, "pressure-outlet", "outlet"),
             (6, "wall", "wall"), (7, "pressure-far-field", "farfield"), (8, "symmetry", "symm")]
    faces = [(int_nodes, int_cells, 3)]
    zone_sides = [(4, [(0, 0)]), (5, [(0, 1)]), (6, [(1, 0)]), (7, [(1, 1)]), (8, [(2, 0), (2, 1)])]
    for zid, sides in zone_sides:
        zb = []
        for d, s in sides:
            coord = (ci, cj, ck)[d]
            m = coord == (0 if s == 0 else n - 1)
            cb = cube[m]
            hx = cb[~split[cb]]
            tt = cb[split[cb]]
            out_dir = np.zeros(3)
            out_dir[d] = 1.0 if s == 1 else -1.0
            nodes = local_nodes(hx, corner_off[_HEX_FACES[(d, s)]])
            zb.append((nodes, np.stack([cstart[hx], -np.ones(hx.size, np.int64)], 1), hx * 4))
            for t, (verts, perm) in enumerate(tri_tab[(d, s)]):
                nodes = local_nodes(tt, verts)
                ctr = tet_centroid(tt, perm)
                nodes = orient(nodes, ctr, ctr + out_dir[None, :])
                zb.append((pad4(nodes), np.stack([cstart[tt] + perm, -np.ones(tt.size, np.int64)], 1),
                           tt * 4 + 1 + t))
        k = np.concatenate([b[2] for b in zb])
        o = np.argsort(k, kind="stable")
        zn = np.concatenate([pad4(b[0]) for b in zb])[o]
        zc = np.concatenate([b[1] for b in zb])[o]
        faces.append((zn, zc, zid))

    # unpad: build CSR directly
    allnodes = np.concatenate([f[0] for f in faces])
    allcells = np.concatenate([f[1] for f in faces]).astype(np.int32)
    allzone = np.concatenate([np.full(f[0].shape[0], f[2], np.int32) for f in faces])
    valid = allnodes >= 0
    counts = valid.sum(1)
    ptr = np.zeros(counts.size + 1, np.int64)
    np.cumsum(counts, out=ptr[1:])
    return MshMesh(x_no, ptr, allnodes[valid].astype(np.int32), allcells, allzone, zones, ncv)


def write_msh(mesh: MshMesh, path: str) -> None:
    """Fluent ASCII mesh as read by MshFilter (hex indices, mixed face
    sections with per-face node counts). Faces are written with c0 = icv0 and
    the node order reversed, which the reader flips back (MshFilter.cpp:453-462)."""
    nno, nfa, ncv = mesh.nno, mesh.nfa, mesh.ncv
    with open(path, "w") as f:
        f.write('(0 "myjoe_amd meshgen")\n(2 3)\n')
        f.write(f"(10 (0 1 {nno:x} 0 3))\n")
        f.write(f"(10 (1 1 {nno:x} 1 3)(\n")
        np.savetxt(f, mesh.x_no, fmt="%.17e")
        f.write("))\n")
        f.write(f"(12 (0 1 {ncv:x} 0))\n(12 (2 1 {ncv:x} 1 7))\n")
        f.write(f"(13 (0 1 {nfa:x} 0))\n")
        ptr = mesh.face_ptr
        start = 0
        for zid, ztype, _name in mesh.zones:
            idx = np.nonzero(mesh.face_zone == zid)[0]
            if idx.size == 0:
                continue
            assert idx[0] == start and idx[-1] == start + idx.size - 1, "zone faces must be contiguous"
            bctype = 2 if ztype == "interior" else 3
            f.write(f"(13 ({zid:x} {start + 1:x} {start + idx.size:x} {bctype:x} 0)(\n")
            lines = []
            for ifa in idx:
                nodes = mesh.face_nodes[ptr[ifa]:ptr[ifa + 1]][::-1] + 1
                c0, c1 = mesh.face_cells[ifa] + 1
                lines.append(f"{nodes.size:x} " + " ".join(f"{v:x}" for v in nodes) + f" {c0:x} {c1:x}")
            f.write("\n".join(lines))
            f.write("\n))\n")
            start += idx.size
        for zid, ztype, name in [(2, "fluid", "fluid")] + list(mesh.zones):
            f.write(f"(45 ({zid} {ztype} {name})())\n")
