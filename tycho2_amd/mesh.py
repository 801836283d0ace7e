"""Tetrahedral meshes for the sweep engine.

* ``SerialMesh`` holds the reference's serial-mesh data model: cells with four
  bounding faces and nodes, faces with two bounding cells (the second is
  ``-1`` on the domain boundary) and three nodes, and node coordinates
  (reference: src/SerialMesh.hh:43-68).
* ``read_smesh`` / ``write_smesh`` use the reference's ``.smesh`` v2 binary
  layout (reference: src/SerialMesh.cc:209-293 for the reader).
* ``write_pmesh`` / ``read_pmesh`` use its partitioned ``.pmesh`` v2 layout
  (src/ParallelMesh.hh:43-81, src/ParallelMesh.cc:167-392): a serial mesh cut
  into partitions by a per-cell partition number, and any such file read
  back into one serial mesh (one GPU holds every partition).
* ``kuhn_cube`` generates the unit-cube meshes of the benchmark problems: an
  ``n x n x n`` grid of hexahedra, each split into six tetrahedra around the
  (0,0,0)-(1,1,1) diagonal, with optional seeded jitter of interior nodes.
* ``tycho_connectivity`` derives the per-cell face adjacency and the face/cell
  vertex maps the sweep needs, with the reference's conventions: local face k
  is opposite cell vertex k (src/TychoMeshIO.cc:129-145, 226-246), face vertex
  j is the face record's node j (src/TychoMeshIO.cc:275-296), and the
  neighbour vertex of a face vertex is the neighbour's vertex holding the same
  node (src/TychoMesh.cc:245-259).

Indices are int64; ``-1`` stands for the reference's UINT64_MAX sentinels.
"""
from __future__ import annotations

import dataclasses
import itertools
import struct

import numpy as np

SMESH_NAME = b"Tycho 2 Serial Mesh"
SMESH_VERSION = 2
PMESH_NAME = b"Tycho 2 Parallel Mesh"
PMESH_VERSION = 2
INVALID = np.uint64(0xFFFFFFFFFFFFFFFF)
# face boundary types of a parallel mesh (src/ParallelMesh.cc:88-90)
MESH_BOUNDARY, PARTITION_BOUNDARY, NOT_BOUNDARY = 0, 1, 2


@dataclasses.dataclass
class SerialMesh:
    cell_faces: np.ndarray   # (nc, 4) int64
    cell_nodes: np.ndarray   # (nc, 4) int64
    material: np.ndarray     # (nc,)   int64
    face_cells: np.ndarray   # (nf, 2) int64, [1] == -1 on the boundary
    face_nodes: np.ndarray   # (nf, 3) int64
    coords: np.ndarray       # (nn, 3) float64

    @property
    def n_cells(self) -> int:
        return int(self.cell_nodes.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.face_nodes.shape[0])

    @property
    def n_nodes(self) -> int:
        return int(self.coords.shape[0])


# ---------------------------------------------------------------- file format
def write_smesh(mesh: SerialMesh, path: str) -> None:
    """Write ``mesh`` as a ``.smesh`` v2 file (32-byte name, 4 x uint64
    header, 9 uint64 per cell, 5 uint64 per face, 3 doubles per node)."""
    name = SMESH_NAME.ljust(32, b"\0")
    cells = np.concatenate([mesh.cell_faces, mesh.cell_nodes,
                            mesh.material[:, None]], axis=1).astype(np.int64)
    faces = np.concatenate([mesh.face_cells, mesh.face_nodes],
                           axis=1).astype(np.int64)
    with open(path, "wb") as f:
        f.write(name)
        f.write(struct.pack("<4Q", SMESH_VERSION, mesh.n_cells, mesh.n_faces,
                            mesh.n_nodes))
        f.write(cells.view(np.uint64).tobytes())
        f.write(faces.view(np.uint64).tobytes())
        f.write(np.ascontiguousarray(mesh.coords, dtype=np.float64).tobytes())


def read_smesh(path: str) -> SerialMesh:
    """Read a ``.smesh`` v2 file; raises ``ValueError`` on a bad header."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:32].rstrip(b"\0") != SMESH_NAME:
        raise ValueError(f"{path}: not a Tycho 2 serial mesh")
    version, nc, nf, nn = struct.unpack_from("<4Q", raw, 32)
    if version != SMESH_VERSION:
        raise ValueError(f"{path}: serial mesh version {version} != 2")
    off = 64
    cells = np.frombuffer(raw, np.int64, nc * 9, off).reshape(nc, 9)
    off += nc * 9 * 8
    faces = np.frombuffer(raw, np.int64, nf * 5, off).reshape(nf, 5)
    off += nf * 5 * 8
    coords = np.frombuffer(raw, np.float64, nn * 3, off).reshape(nn, 3)
    if off + nn * 24 != len(raw):
        raise ValueError(f"{path}: size mismatch")
    return SerialMesh(cell_faces=cells[:, 0:4].copy(),
                      cell_nodes=cells[:, 4:8].copy(),
                      material=cells[:, 8].copy(),
                      face_cells=faces[:, 0:2].copy(),
                      face_nodes=faces[:, 2:5].copy(),
                      coords=coords.copy())


def write_pmesh(mesh: SerialMesh, path: str, partition=None,
                n_parts: int | None = None) -> None:
    """Write ``mesh`` as a ``.pmesh`` v2 file, cell c in partition
    ``partition[c]`` (default: one partition).  Within a partition, cells,
    faces and nodes keep ascending global order; a face's bounding cell in
    another partition is that partition's local index, its partition entry
    the other partition, its type PartitionBoundary; a face on the domain
    boundary has INVALID for its missing cell and partition, type
    MeshBoundary; any other face NotBoundary."""
    nc = mesh.n_cells
    part = (np.zeros(nc, np.int64) if partition is None
            else np.asarray(partition, np.int64))
    n_parts = int(part.max()) + 1 if n_parts is None else int(n_parts)
    cells_of = [np.flatnonzero(part == p) for p in range(n_parts)]
    local = np.empty(nc, np.int64)           # cell -> index in its partition
    for cs in cells_of:
        local[cs] = np.arange(cs.size)
    blobs = []
    for p, cs in enumerate(cells_of):
        faces = np.unique(mesh.cell_faces[cs])
        nodes = np.unique(mesh.cell_nodes[cs])
        cell = np.empty((cs.size, 10), np.uint64)
        cell[:, 0:4] = np.searchsorted(faces, mesh.cell_faces[cs])
        cell[:, 4:8] = np.searchsorted(nodes, mesh.cell_nodes[cs])
        cell[:, 8] = cs
        cell[:, 9] = mesh.material[cs]
        face = np.empty((faces.size, 9), np.uint64)
        btype = np.full(faces.size, NOT_BOUNDARY, np.uint64)
        for side in range(2):
            gc = mesh.face_cells[faces, side]
            edge = gc < 0
            other = ~edge & (part[np.maximum(gc, 0)] != p)
            face[:, side] = np.where(edge, INVALID,
                                     local[np.maximum(gc, 0)].astype(np.uint64))
            face[:, 6 + side] = np.where(
                edge, INVALID, part[np.maximum(gc, 0)].astype(np.uint64))
            btype = np.where(edge, np.uint64(MESH_BOUNDARY), btype)
            btype = np.where(other, np.uint64(PARTITION_BOUNDARY), btype)
        face[:, 2:5] = np.searchsorted(nodes, mesh.face_nodes[faces])
        face[:, 5] = faces
        face[:, 8] = btype
        node = np.empty((nodes.size, 4), np.uint64)
        node[:, 0:3] = np.ascontiguousarray(
            mesh.coords[nodes], np.float64).view(np.uint64)
        node[:, 3] = nodes
        blobs.append(struct.pack("<3Q", cs.size, faces.size, nodes.size) +
                     cell.tobytes() + face.tobytes() + node.tobytes())
    head = 32 + (3 + n_parts) * 8
    offsets = [head]
    for b in blobs:
        offsets.append(offsets[-1] + len(b))
    with open(path, "wb") as f:
        f.write(PMESH_NAME.ljust(32, b"\0"))
        f.write(struct.pack(f"<{2 + n_parts + 1}Q", PMESH_VERSION, n_parts,
                            *offsets))
        for b in blobs:
            f.write(b)


def read_pmesh(path: str) -> SerialMesh:
    """Read a ``.pmesh`` v2 file into one serial mesh numbered by the global
    IDs it stores; raises ``ValueError`` on a bad header or inconsistent
    partitions."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:32].rstrip(b"\0") != PMESH_NAME:
        raise ValueError(f"{path}: not a Tycho 2 parallel mesh")
    version, n_parts = struct.unpack_from("<2Q", raw, 32)
    if version != PMESH_VERSION:
        raise ValueError(f"{path}: parallel mesh version {version} != 2")
    offsets = struct.unpack_from(f"<{n_parts + 1}Q", raw, 48)
    if offsets[-1] != len(raw):
        raise ValueError(f"{path}: size mismatch")
    parts = []
    for p in range(n_parts):
        off = offsets[p]
        nc, nf, nn = struct.unpack_from("<3Q", raw, off)
        off += 24
        cell = np.frombuffer(raw, np.uint64, nc * 10, off).reshape(nc, 10)
        off += nc * 80
        face = np.frombuffer(raw, np.uint64, nf * 9, off).reshape(nf, 9)
        off += nf * 72
        node = np.frombuffer(raw, np.uint64, nn * 4, off).reshape(nn, 4)
        if off + nn * 32 != offsets[p + 1]:
            raise ValueError(f"{path}: partition {p} size mismatch")
        parts.append((cell, face, node))
    n_cells = sum(c.shape[0] for c, _, _ in parts)
    n_faces = int(max(f[:, 5].max() for _, f, _ in parts if f.size)) + 1
    n_nodes = int(max(n[:, 3].max() for _, _, n in parts if n.size)) + 1
    cell_faces = np.full((n_cells, 4), -1, np.int64)
    cell_nodes = np.full((n_cells, 4), -1, np.int64)
    material = np.zeros(n_cells, np.int64)
    face_cells = np.full((n_faces, 2), -2, np.int64)
    face_nodes = np.full((n_faces, 3), -1, np.int64)
    coords = np.full((n_nodes, 3), np.nan)
    cell_gid = [c[:, 8].astype(np.int64) for c, _, _ in parts]
    for p, (cell, face, node) in enumerate(parts):
        fg = face[:, 5].astype(np.int64)
        ng = node[:, 3].astype(np.int64)
        g = cell_gid[p]
        cell_faces[g] = fg[cell[:, 0:4].astype(np.int64)]
        cell_nodes[g] = ng[cell[:, 4:8].astype(np.int64)]
        material[g] = cell[:, 9].astype(np.int64)
        coords[ng] = node[:, 0:3].copy().view(np.float64)
        face_nodes[fg] = ng[face[:, 2:5].astype(np.int64)]
        for side in range(2):
            lc = face[:, side]
            owner = face[:, 6 + side]
            gc = np.full(fg.size, -1, np.int64)
            for q in range(n_parts):
                sel = (lc != INVALID) & (owner == np.uint64(q))
                gc[sel] = cell_gid[q][lc[sel].astype(np.int64)]
            face_cells[fg, side] = gc
    if (cell_faces < 0).any() or (face_cells == -2).any() or \
            np.isnan(coords).any():
        raise ValueError(f"{path}: partitions do not cover the mesh")
    return SerialMesh(cell_faces=cell_faces, cell_nodes=cell_nodes,
                      material=material, face_cells=face_cells,
                      face_nodes=face_nodes, coords=coords)


# ---------------------------------------------------------------- generator
# The six tetrahedra of a hexahedron sharing the (0,0,0)-(1,1,1) diagonal:
# one per ordering of the three axes.  Hex corner index = x + 2y + 4z.
def _kuhn_local_tets() -> np.ndarray:
    tets = []
    for perm in itertools.permutations(range(3)):
        corner = [0, 0, 0]
        verts = [0]
        for axis in perm:
            corner[axis] = 1
            verts.append(corner[0] + 2 * corner[1] + 4 * corner[2])
        tets.append(verts)
    return np.array(tets, dtype=np.int64)


KUHN_LOCAL_TETS = _kuhn_local_tets()


def _signed_volumes(coords: np.ndarray, cell_nodes: np.ndarray) -> np.ndarray:
    p = coords[cell_nodes]                      # (nc, 4, 3)
    a = p[:, 1] - p[:, 0]
    b = p[:, 2] - p[:, 0]
    c = p[:, 3] - p[:, 0]
    return np.einsum("ij,ij->i", np.cross(a, b), c) / 6.0


def kuhn_cube(n: int, jitter: float = 0.0, seed: int = 0,
              length: float = 1.0) -> SerialMesh:
    """Cube ``[0, length]^3`` cut into ``n^3`` hexahedra x 6 Kuhn tetrahedra.

    ``jitter`` is the half-width, as a fraction of the hex size h, of the
    uniform displacement drawn per axis for every interior node from
    ``numpy.random.default_rng(seed)``.  A node of any tetrahedron whose
    volume would become non-positive (orientation relative to the unjittered
    tetrahedron) is drawn again until none is left.
    """
    if n < 1:
        raise ValueError("n must be >= 1")
    h = length / n
    m = n + 1
    idx = np.arange(m)
    kk, jj, ii = np.meshgrid(idx, idx, idx, indexing="ij")
    coords = np.stack([ii.ravel() * h, jj.ravel() * h, kk.ravel() * h],
                      axis=1).astype(np.float64)
    # boundary nodes keep exact grid coordinates
    coords[:, 0][ii.ravel() == n] = length
    coords[:, 1][jj.ravel() == n] = length
    coords[:, 2][kk.ravel() == n] = length

    # cells: hex (i, j, k) -> 6 tets; node id = i + m*(j + m*k)
    hi, hj, hk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n),
                             indexing="ij")
    base = (hi + m * (hj + m * hk)).transpose(2, 1, 0).ravel()  # x fastest
    corner_off = np.array([(c & 1) + m * ((c >> 1) & 1) + m * m * (c >> 2)
                           for c in range(8)], dtype=np.int64)
    cell_nodes = (base[:, None, None] + corner_off[KUHN_LOCAL_TETS][None])
    cell_nodes = cell_nodes.reshape(-1, 4)
    nc = cell_nodes.shape[0]

    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        interior = ((ii.ravel() > 0) & (ii.ravel() < n) & (jj.ravel() > 0)
                    & (jj.ravel() < n) & (kk.ravel() > 0) & (kk.ravel() < n))
        int_nodes = np.nonzero(interior)[0]
        base_coords = coords.copy()
        sign0 = np.sign(_signed_volumes(base_coords, cell_nodes))
        amp = jitter * h
        coords[int_nodes] += rng.uniform(-amp, amp, (int_nodes.size, 3))
        for _ in range(1000):
            bad = np.nonzero(_signed_volumes(coords, cell_nodes) * sign0
                             <= 0.0)[0]
            if bad.size == 0:
                break
            redraw = np.unique(cell_nodes[bad].ravel())
            redraw = redraw[interior[redraw]]
            coords[redraw] = base_coords[redraw] + rng.uniform(
                -amp, amp, (redraw.size, 3))
        else:
            raise RuntimeError("jitter redraw did not converge")

    # faces: face k of a cell is opposite vertex k
    opp = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    slots = np.sort(cell_nodes[:, opp], axis=2).reshape(-1, 3)  # (4nc, 3)
    bits = max(1, int(coords.shape[0] - 1).bit_length())
    if 3 * bits <= 63:
        # one int64 key per sorted node triple, ordered as the triples are
        # lexicographically (the same face numbering as a row-wise unique)
        keys = (slots[:, 0] << (2 * bits)) | (slots[:, 1] << bits) | \
            slots[:, 2]
        ukeys, inverse = np.unique(keys, return_inverse=True)
        mask = (1 << bits) - 1
        face_nodes = np.stack([ukeys >> (2 * bits), (ukeys >> bits) & mask,
                               ukeys & mask], axis=1)
    else:
        face_nodes, inverse = np.unique(slots, axis=0, return_inverse=True)
    inverse = inverse.ravel()
    nf = face_nodes.shape[0]
    cell_faces = inverse.reshape(nc, 4).astype(np.int64)
    owner = np.repeat(np.arange(nc, dtype=np.int64), 4)
    order = np.argsort(inverse, kind="stable")
    inv_sorted = inverse[order]
    first = np.ones(order.size, dtype=bool)
    first[1:] = inv_sorted[1:] != inv_sorted[:-1]
    face_cells = np.full((nf, 2), -1, dtype=np.int64)
    face_cells[inv_sorted[first], 0] = owner[order[first]]
    face_cells[inv_sorted[~first], 1] = owner[order[~first]]
    return SerialMesh(cell_faces=cell_faces,
                      cell_nodes=cell_nodes.astype(np.int64),
                      material=np.ones(nc, dtype=np.int64),
                      face_cells=face_cells,
                      face_nodes=face_nodes.astype(np.int64),
                      coords=coords)


# ---------------------------------------------------------------- topology
@dataclasses.dataclass
class Connectivity:
    adj_cell: np.ndarray   # (nc, 4) int64, -1 on the boundary
    f2c: np.ndarray        # (nc, 4, 3) int64: face vertex -> cell vertex
    nbr_vrtx: np.ndarray   # (nc, 4, 3) int64: face vertex -> neighbour vertex


def tycho_connectivity(mesh: SerialMesh) -> Connectivity:
    """Per-cell adjacency and vertex maps of a one-partition mesh."""
    cn = mesh.cell_nodes
    nc = cn.shape[0]
    fn = mesh.face_nodes[mesh.cell_faces]                   # (nc, 4, 3)
    # lface = index of the cell vertex not on the face
    on_face = (cn[:, None, :, None] == fn[:, :, None, :]).any(axis=3)
    if not (on_face.sum(axis=2) == 3).all():
        raise ValueError("face nodes are not a subset of cell nodes")
    lface = np.argmin(on_face, axis=2)                      # (nc, 4) by slot
    if not (np.sort(lface, axis=1) == np.arange(4)).all():
        raise ValueError("cell faces do not cover the four local faces")
    rows = np.arange(nc)[:, None]
    fc = mesh.face_cells[mesh.cell_faces]                   # (nc, 4, 2)
    other = np.where(fc[:, :, 0] == rows, fc[:, :, 1], fc[:, :, 0])
    adj = np.empty((nc, 4), dtype=np.int64)
    adj[rows, lface] = other
    # face vertex j -> cell vertex holding face node j
    match = fn[:, :, :, None] == cn[:, None, None, :]       # (nc, 4, 3, 4)
    f2c_slot = np.argmax(match, axis=3)
    f2c = np.empty((nc, 4, 3), dtype=np.int64)
    f2c[rows, lface] = f2c_slot
    nbr = np.full((nc, 4, 3), -1, dtype=np.int64)
    interior = adj >= 0
    cidx, fidx = np.nonzero(interior)
    nodes = cn[cidx[:, None], f2c[cidx, fidx]]              # (k, 3)
    ncn = cn[adj[cidx, fidx]]                               # (k, 4)
    nbr[cidx, fidx] = np.argmax(nodes[:, :, None] == ncn[:, None, :], axis=2)
    return Connectivity(adj_cell=adj, f2c=f2c, nbr_vrtx=nbr)
