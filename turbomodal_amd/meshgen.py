"""Seeded structured TET10 sector meshes for the cyclic modal workloads.

The reference ships no generator for its benchmark-sized sectors, so the
workloads are defined here from their published descriptions:

* ``flat_disk`` - annular disk sector on an (r, theta, z) hex lattice;
* ``bladed_sector`` - disk sector plus a twisted flat-plate blade whose root
  face is node-merged with the disk rim.

Every hex cell is split into six tetrahedra along its v000->v111 diagonal
(Kuhn / Freudenthal split: one tet per axis permutation), which gives a
conforming mesh.  TET10 mid-edge nodes are the points of the half-step
lattice: every half-lattice point is the midpoint of exactly one lattice edge
(axis edge, face diagonal or body diagonal), so node ids are simply indices
into the doubled lattice.  Interior lattice vertices are jittered uniformly by
+-5 % of the local cell size from a fixed seed; boundary vertices are not, so
the left and right cyclic faces stay exact rotations of each other and the
hub bore stays on r = r_min.  Mid-edge nodes are placed after jittering, so
every element stays straight-edged.

Node ordering inside each element follows the gmsh TET10 convention the
reference uses (``src/element.cpp:71-88``): corners 0-3, then mid-edge nodes
4:(0,1) 5:(1,2) 6:(0,2) 7:(0,3) 8:(1,3) 9:(2,3).  Node sets are named as the
reference expects (``left_boundary``, ``right_boundary``, ``hub_constraint``;
``src/mesh.cpp:231-258``, ``src/cyclic_solver.cpp:88-105``).  The rotation
axis is z.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

# Corner pairs for the six mid-edge nodes, gmsh TET10 order.
TET10_EDGES = ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))

STEEL = dict(E=200e9, nu=0.3, rho=7800.0)
TITANIUM = dict(E=114e9, nu=0.34, rho=4430.0)


@dataclass
class SectorMesh:
    """Plain-array sector mesh (what ``Mesh.load_from_arrays`` consumes)."""

    nodes: np.ndarray                     # (N, 3) float64
    elements: np.ndarray                  # (E, 10) int32, gmsh TET10 order
    node_sets: dict = field(default_factory=dict)   # name -> sorted int array
    num_sectors: int = 36
    rotation_axis: int = 2
    seed: int = 0
    description: str = ""

    @property
    def num_nodes(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def num_dof(self) -> int:
        return 3 * self.num_nodes

    @property
    def num_elements(self) -> int:
        return int(self.elements.shape[0])


def _kuhn_cells(shape):
    """Yield the 6 Kuhn tets of every cell as doubled-lattice corner coords.

    ``shape`` = number of cells per axis.  Returns an array (ncell*6, 4, 3) of
    corner coordinates in the doubled lattice (vertex (i,j,k) -> (2i,2j,2k)).
    """
    n0, n1, n2 = shape
    base = np.stack(np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2),
                                indexing="ij"), axis=-1).reshape(-1, 3)
    tets = []
    for perm in itertools.permutations(range(3)):
        p = np.zeros((4, 3), dtype=np.int64)
        p[1, perm[0]] = 1
        p[2] = p[1]
        p[2, perm[1]] = 1
        p[3] = 1
        tets.append(p)
    tets = np.asarray(tets)                              # (6, 4, 3)
    corners = base[:, None, None, :] + tets[None, :, :, :]  # (ncell, 6, 4, 3)
    return 2 * corners.reshape(-1, 4, 3)


def _fill_midpoints(dcoords, vertex_xyz, dshape):
    """Coordinates of every doubled-lattice point.

    ``vertex_xyz`` holds jittered vertex positions indexed by (i, j, k).
    A point with odd coordinates on axis set S is the midpoint of the edge
    p - d .. p + d with d = sum of unit vectors over S.
    """
    odd = dcoords % 2
    a = (dcoords - odd) // 2
    b = (dcoords + odd) // 2
    pa = vertex_xyz[a[:, 0], a[:, 1], a[:, 2]]
    pb = vertex_xyz[b[:, 0], b[:, 1], b[:, 2]]
    return 0.5 * (pa + pb)


def _lattice_points(dshape):
    d0, d1, d2 = dshape
    return np.stack(np.meshgrid(np.arange(d0), np.arange(d1), np.arange(d2),
                                indexing="ij"), axis=-1).reshape(-1, 3)


def _jitter(rng, vshape, frac=0.05):
    """Uniform jitter in units of the local cell size, zero on the boundary."""
    j = rng.uniform(-frac, frac, size=tuple(vshape) + (3,))
    interior = np.zeros(vshape, dtype=bool)
    interior[1:-1, 1:-1, 1:-1] = True
    j[~interior] = 0.0
    return j


def _build_elements(corner_ids_fn, tets_dc, xyz):
    """Turn doubled-coordinate Kuhn tets into positively oriented TET10 rows."""
    c = tets_dc  # (T, 4, 3) doubled coords
    cid = corner_ids_fn(c.reshape(-1, 3)).reshape(-1, 4)
    p = xyz[cid]                                          # (T, 4, 3)
    vol = np.einsum("ti,ti->t", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]),
                    p[:, 3] - p[:, 0])
    neg = vol < 0
    c = c.copy()
    c[neg, 1], c[neg, 2] = c[neg, 2].copy(), c[neg, 1].copy()
    conn = np.empty((c.shape[0], 10), dtype=np.int64)
    conn[:, :4] = corner_ids_fn(c.reshape(-1, 3)).reshape(-1, 4)
    for m, (u, v) in enumerate(TET10_EDGES):
        conn[:, 4 + m] = corner_ids_fn((c[:, u] + c[:, v]) // 2)
    return conn


def _disk_vertices(rng, r0, r1, thick, cells, alpha, jitter_frac):
    nr, nt, nz = cells
    vshape = (nr + 1, nt + 1, nz + 1)
    ir, it, iz = np.meshgrid(np.arange(nr + 1), np.arange(nt + 1),
                             np.arange(nz + 1), indexing="ij")
    jit = _jitter(rng, vshape, jitter_frac)
    r = r0 + (ir + jit[..., 0]) * (r1 - r0) / nr
    th = (it + jit[..., 1]) * alpha / nt
    z = (iz + jit[..., 2]) * thick / nz
    return np.stack([r * np.cos(th), r * np.sin(th), z], axis=-1)


def flat_disk(num_sectors=36, r_inner=0.10, r_outer=0.30, thickness=0.010,
              cells=(12, 6, 2), seed=1, jitter=0.05) -> SectorMesh:
    """Annular disk sector on a Kuhn-split (r, theta, z) lattice."""
    rng = np.random.default_rng(seed)
    alpha = 2.0 * np.pi / num_sectors
    nr, nt, nz = cells
    dshape = (2 * nr + 1, 2 * nt + 1, 2 * nz + 1)
    vxyz = _disk_vertices(rng, r_inner, r_outer, thickness, cells, alpha, jitter)
    pts = _lattice_points(dshape)
    xyz = _fill_midpoints(pts, vxyz, dshape)

    def ids(dc):
        return (dc[:, 0] * dshape[1] + dc[:, 1]) * dshape[2] + dc[:, 2]

    conn = _build_elements(ids, _kuhn_cells(cells), xyz)
    I, J = pts[:, 0], pts[:, 1]
    sets = {
        "left_boundary": np.flatnonzero(J == 0),
        "right_boundary": np.flatnonzero(J == dshape[1] - 1),
        "hub_constraint": np.flatnonzero(I == 0),
    }
    return SectorMesh(xyz, conn.astype(np.int32), sets, num_sectors, 2, seed,
                      f"flat disk N={num_sectors} cells={cells}")


def bladed_sector(num_sectors=36, r_inner=0.10, r_outer=0.25, disk_thickness=0.020,
                  disk_cells=(24, 12, 4), span=0.15, chord=0.04, blade_thickness=0.004,
                  blade_cells=(24, 8, 2), twist_deg=35.0, blend_layers=4,
                  seed=2, jitter=0.05) -> SectorMesh:
    """Disk sector plus a twisted flat-plate blade merged at the rim.

    The blade lattice is (span, chord, thickness).  Its root face is the disk
    rim sub-lattice of ``blade_cells[1]`` theta cells x ``blade_cells[2]`` z
    cells centred on the sector, so the root nodes are disk nodes (merged by
    index) and the two Kuhn splits conform.  Above the root the plate section
    (``chord`` along the local tangential direction, ``blade_thickness``
    axial) is twisted linearly from 0 at the root to ``twist_deg`` at the tip
    about the radial line through the sector centre.  The root section is
    whatever the rim sub-lattice spans, so the section size is blended
    smoothly from the root section to the nominal plate over the first
    ``blend_layers`` span layers.
    """
    rng = np.random.default_rng(seed)
    alpha = 2.0 * np.pi / num_sectors
    nr, nt, nz = disk_cells
    ns, nc, nh = blade_cells
    if nc > nt or nh > nz or (nt - nc) % 2 or (nz - nh) % 2:
        raise ValueError("blade root lattice must be a centred sub-lattice of the rim")
    ddshape = (2 * nr + 1, 2 * nt + 1, 2 * nz + 1)
    vxyz = _disk_vertices(rng, r_inner, r_outer, disk_thickness, disk_cells, alpha, jitter)
    dpts = _lattice_points(ddshape)
    dxyz = _fill_midpoints(dpts, vxyz, ddshape)
    n_disk = dxyz.shape[0]

    def disk_ids(dc):
        return (dc[:, 0] * ddshape[1] + dc[:, 1]) * ddshape[2] + dc[:, 2]

    disk_conn = _build_elements(disk_ids, _kuhn_cells(disk_cells), dxyz)

    # --- blade vertices (span l, chord u, thickness v) ---
    j0 = (nt - nc) // 2          # first rim theta cell under the blade
    k0 = (nz - nh) // 2
    bvshape = (ns + 1, nc + 1, nh + 1)
    il, iu, iv = np.meshgrid(np.arange(ns + 1), np.arange(nc + 1), np.arange(nh + 1),
                             indexing="ij")
    jit = _jitter(rng, bvshape, jitter)
    lf = (il + jit[..., 0]) / ns                     # span fraction 0..1
    uf = (iu + jit[..., 1]) / nc - 0.5               # -0.5..0.5 along chord
    vf = (iv + jit[..., 2]) / nh - 0.5               # -0.5..0.5 through thickness
    th_c = 0.5 * alpha
    er = np.array([np.cos(th_c), np.sin(th_c), 0.0])
    et = np.array([-np.sin(th_c), np.cos(th_c), 0.0])
    ez = np.array([0.0, 0.0, 1.0])
    zc = 0.5 * disk_thickness
    # root section extents from the rim sub-lattice (arc length, axial)
    root_chord = r_outer * alpha * nc / nt
    root_thick = disk_thickness * nh / nz
    blend = np.clip(lf * ns / max(blend_layers, 1), 0.0, 1.0)
    blend = blend * blend * (3.0 - 2.0 * blend)      # smoothstep
    c_eff = root_chord + (chord - root_chord) * blend
    t_eff = root_thick + (blade_thickness - root_thick) * blend
    phi = np.deg2rad(twist_deg) * lf
    u_loc = uf * c_eff
    v_loc = vf * t_eff
    rad = r_outer + lf * span
    d_t = u_loc * np.cos(phi) - v_loc * np.sin(phi)
    d_z = u_loc * np.sin(phi) + v_loc * np.cos(phi)
    bv = (rad[..., None] * er + d_t[..., None] * et + d_z[..., None] * ez
          + zc * ez)
    # root vertices sit exactly on the disk rim vertices
    bv[0] = vxyz[nr, j0:j0 + nc + 1, k0:k0 + nh + 1]

    bdshape = (2 * ns + 1, 2 * nc + 1, 2 * nh + 1)
    bpts = _lattice_points(bdshape)
    bxyz_all = _fill_midpoints(bpts, bv, bdshape)
    # doubled-lattice blade points with l == 0 are disk rim points
    root = bpts[:, 0] == 0
    nonroot_idx = np.flatnonzero(~root)
    blade_id = np.full(bpts.shape[0], -1, dtype=np.int64)
    blade_id[nonroot_idx] = n_disk + np.arange(nonroot_idx.size)
    rdc = bpts[root]
    blade_id[np.flatnonzero(root)] = disk_ids(
        np.stack([np.full(rdc.shape[0], 2 * nr), 2 * j0 + rdc[:, 1], 2 * k0 + rdc[:, 2]], axis=1))

    def blade_ids(dc):
        return blade_id[(dc[:, 0] * bdshape[1] + dc[:, 1]) * bdshape[2] + dc[:, 2]]

    xyz = np.concatenate([dxyz, bxyz_all[nonroot_idx]], axis=0)
    blade_conn = _build_elements(blade_ids, _kuhn_cells(blade_cells), xyz)
    conn = np.concatenate([disk_conn, blade_conn], axis=0)
    I, J = dpts[:, 0], dpts[:, 1]
    sets = {
        "left_boundary": np.flatnonzero(J == 0),
        "right_boundary": np.flatnonzero(J == ddshape[1] - 1),
        "hub_constraint": np.flatnonzero(I == 0),
    }
    return SectorMesh(xyz, conn.astype(np.int32), sets, num_sectors, 2, seed,
                      f"bladed sector N={num_sectors} disk={disk_cells} blade={blade_cells}")


def config_mesh(name: str, num_sectors: int | None = None) -> SectorMesh:
    """Meshes of the named workloads.

    ``flat_disk_c1``      - (12,6,2) flat disk, 1625 nodes / 4875 DOF
    ``bladed_c2``         - (24,12,4) disk + (24,8,2) blade, 15105 nodes / 45315 DOF
    ``bladed_c3``         - every lattice dimension doubled, 109313 nodes / 327939 DOF
    """
    if name == "flat_disk_c1":
        return flat_disk(num_sectors=num_sectors or 36)
    if name == "bladed_c2":
        return bladed_sector(num_sectors=num_sectors or 36)
    if name == "bladed_c3":
        return bladed_sector(num_sectors=num_sectors or 36, disk_cells=(48, 24, 8),
                             blade_cells=(48, 16, 4), seed=3)
    raise KeyError(name)


def to_core_mesh(sm: SectorMesh, core):
    """Build a ``Mesh`` of the ``_core`` module from a :class:`SectorMesh`."""
    mesh = core.Mesh()
    sets = []
    for name, ids in sm.node_sets.items():
        ns = core.NodeSet()
        ns.name = name
        ns.node_ids = [int(i) for i in ids]
        sets.append(ns)
    mesh.load_from_arrays(sm.nodes, sm.elements, sets, sm.num_sectors, sm.rotation_axis)
    return mesh
