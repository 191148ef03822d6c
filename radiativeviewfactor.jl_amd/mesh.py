"""Mesh data model and deterministic synthetic meshes.

The data model mirrors the reference's ``MeshData`` / ``SurfaceElement``
(src/MeshIO.jl:36-64) in a flat, array-of-structure-free form that can be
handed to the C ABI without copies:

* ``coords``   float64 (Nn, 3), C order — the same bytes as Julia's
  column-major ``3 x Nn`` matrix.
* ``family``   int8 (N,) — 0 = Quad8 (``:quad``), 1 = Tri6 (``:tri``),
  2 = Quad4, 3 = Tri3 (the codes of src/GPUKernels.jl:344-375).
* ``nodes``    int32 (N, 8) — 1-based node ids, unused slots 0.
* ``group``    int32 (N,) — physical group tag per element.
* ``group_tags``  dict tag -> name (radiating groups only, like the reference).
* ``group_tri_soup`` dict tag -> float64 (T, 3, 3) array whose bytes equal the
  reference's ``(3, 3, T)`` column-major soup: ``soup[t, b, a]`` is the
  reference's ``soup[a, b, t]``.  ``soup_layout`` says how the producer filled
  it: ``"xyz_vertex"`` (what src/MeshIO.jl:660-669 writes: ``soup[t, v, c]``)
  or ``"vertex_xyz"`` (what test/runtests.jl:104-107 builds: ``soup[t, c, v]``).

Generators produce the five workloads the view-factor path is measured on
(parallel plates, unit-cube enclosure, occluded plates, curved cylinder,
random box enclosures) plus scaled-down variants for tests.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

FAMILY_CODE = {"quad": 0, "tri": 1, "quad4": 2, "tri3": 3, "line3": 4, "line2": 5}
FAMILY_NAME = {v: k for k, v in FAMILY_CODE.items()}
FAMILY_NNODES = {0: 8, 1: 6, 2: 4, 3: 3, 4: 3, 5: 2}


@dataclass
class SurfaceElement:
    """Mirror of src/MeshIO.jl:36-40 (nodes are 1-based)."""
    nodes: List[int]
    group: int
    family: str  # "quad" (Quad8), "tri" (Tri6), "quad4", "tri3"


@dataclass
class MeshData:
    coords: np.ndarray
    family: np.ndarray
    nodes: np.ndarray
    group: np.ndarray
    group_tags: Dict[int, str]
    group_tri_soup: Dict[int, np.ndarray] = field(default_factory=dict)
    soup_layout: str = "xyz_vertex"
    mesh_dim: int = 2

    def __post_init__(self):
        self.coords = np.ascontiguousarray(self.coords, dtype=np.float64).reshape(-1, 3)
        self.family = np.ascontiguousarray(self.family, dtype=np.int8)
        self.nodes = np.ascontiguousarray(self.nodes, dtype=np.int32).reshape(-1, 8)
        self.group = np.ascontiguousarray(self.group, dtype=np.int32)
        curve = self.family >= 4
        if curve.size and curve.all():
            self.mesh_dim = 1
        elif curve.any():
            raise ValueError("a mesh holds either surface or curve (Line2/Line3) elements, not both")
        nv = self.soup_vertices
        self.group_tri_soup = {int(k): np.ascontiguousarray(v, dtype=np.float64).reshape(-1, nv, 3)
                               for k, v in self.group_tri_soup.items()}

    @property
    def soup_vertices(self) -> int:
        """Vertices per obstruction item: 3 (triangles), 2 for curve meshes
        (segment soups, src/MeshIO.jl:641-645)."""
        return 2 if self.mesh_dim == 1 else 3

    # -- reference-shaped views -------------------------------------------------
    @property
    def n_elems(self) -> int:
        return int(self.family.shape[0])

    @property
    def n_nodes(self) -> int:
        return int(self.coords.shape[0])

    @property
    def surface_elems(self) -> List[SurfaceElement]:
        out = []
        for e in range(self.n_elems):
            f = int(self.family[e])
            out.append(SurfaceElement([int(x) for x in self.nodes[e, :FAMILY_NNODES[f]]],
                                      int(self.group[e]), FAMILY_NAME[f]))
        return out

    @property
    def group_elems(self) -> Dict[int, np.ndarray]:
        """tag -> 0-based element indices, in element order (MeshIO.jl:57-64)."""
        return {t: np.nonzero(self.group == t)[0] for t in self.group_tags}

    @staticmethod
    def from_elements(coords, elems: Sequence[SurfaceElement], group_tags: Dict[int, str],
                      group_tri_soup: Optional[Dict[int, np.ndarray]] = None,
                      soup_layout: str = "xyz_vertex") -> "MeshData":
        n = len(elems)
        fam = np.zeros(n, np.int8)
        nod = np.zeros((n, 8), np.int32)
        grp = np.zeros(n, np.int32)
        for e, el in enumerate(elems):
            fam[e] = FAMILY_CODE[el.family]
            nod[e, :len(el.nodes)] = el.nodes
            grp[e] = el.group
        return MeshData(np.asarray(coords, np.float64).reshape(-1, 3), fam, nod, grp, dict(group_tags),
                        dict(group_tri_soup or {}), soup_layout)

    def merged_soup(self):
        """All triangle soups concatenated in ascending group-tag order, with
        a per-triangle group tag.  The per-active-set merge order of the
        reference (src/Assembly.jl:123-141) is derived from this on the host."""
        tags = sorted(self.group_tri_soup)
        if not tags:
            return np.zeros((0, self.soup_vertices, 3)), np.zeros(0, np.int32)
        soup = np.concatenate([self.group_tri_soup[t] for t in tags], axis=0)
        grp = np.concatenate([np.full(self.group_tri_soup[t].shape[0], t, np.int32) for t in tags])
        return np.ascontiguousarray(soup), grp


def build_group_obs_soups(mesh: MeshData, groups: Optional[Sequence[int]] = None) -> Dict[int, np.ndarray]:
    """Per-group obstruction triangle soups from element corners, restating
    src/MeshIO.jl:648-685 (Tri -> (c1,c2,c3); quads -> (c1,c2,c3) + (c1,c3,c4)),
    written in the loader's ``[xyz, vertex, tri]`` layout, element by element."""
    soups = {}
    tags = groups if groups is not None else sorted(mesh.group_tags)
    X = mesh.coords
    if mesh.mesh_dim == 1:  # segment soups from the first two corners (src/MeshIO.jl:672-680)
        for g in tags:
            c = mesh.nodes[np.nonzero(mesh.group == g)[0], :2].astype(np.int64) - 1
            soups[int(g)] = np.ascontiguousarray(np.stack([X[c[:, 0]], X[c[:, 1]]], axis=1))
        return soups
    for g in tags:
        idx = np.nonzero(mesh.group == g)[0]
        c = mesh.nodes[idx, :4].astype(np.int64) - 1
        quad = np.isin(mesh.family[idx], (0, 2))
        first = np.concatenate([[0], np.cumsum(1 + quad.astype(np.int64))[:-1]]).astype(np.int64)
        out = np.empty((len(idx) + int(quad.sum()), 3, 3), np.float64)
        out[first] = np.stack([X[c[:, 0]], X[c[:, 1]], X[c[:, 2]]], axis=1)
        q = np.nonzero(quad)[0]
        out[first[q] + 1] = np.stack([X[c[q, 0]], X[c[q, 2]], X[c[q, 3]]], axis=1)
        soups[int(g)] = out
    return soups


# ----------------------------------------------------------------------------
# Mesh files
# ----------------------------------------------------------------------------
def load_mesh(filename: str, surface_dim: int = 2, reverse_normals: bool = False, verbose: bool = True,
              group_field: Optional[str] = None) -> MeshData:
    """Mirror of load_mesh (src/MeshIO.jl:94-145): a Gmsh .msh (format 2.2 or
    4.1, ASCII or binary) or an XML VTK unstructured grid, read by the
    library's native reader (rvf_meshfile_read), with the obstruction soups
    of every group built as _build_group_obs_soups does.  surface_dim 2 reads
    surfaces; 1 reads curves (Line2, Line3), whose elements in a .msh file are
    turned to face the surface next to their curve (_orient_line3_normals!)."""
    import ctypes as C
    from . import _lib
    L = _lib.lib()
    h = C.c_void_p()
    _lib.check(L.rvf_meshfile_read(str(filename).encode(), surface_dim, 1 if reverse_normals else 0,
                                   group_field.encode() if group_field else None, C.byref(h)))
    try:
        nn, ne, ng, fl = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        _lib.check(L.rvf_meshfile_info(h, C.byref(nn), C.byref(ne), C.byref(ng), C.byref(fl)))
        coords = np.zeros((nn.value, 3), np.float64)
        fam = np.zeros(ne.value, np.int8)
        nodes = np.zeros((ne.value, 8), np.int32)
        grp = np.zeros(ne.value, np.int32)
        _lib.check(L.rvf_meshfile_arrays(h, _lib.ptr(coords), _lib.ptr(fam), _lib.ptr(nodes), _lib.ptr(grp)))
        tags = {}
        buf = C.create_string_buffer(4096)
        for k in range(ng.value):
            t = C.c_int32()
            _lib.check(L.rvf_meshfile_group(h, k, C.byref(t), buf, 4096))
            tags[int(t.value)] = buf.value.decode()
        flags = fl.value
    finally:
        L.rvf_meshfile_free(h)
    m = MeshData(coords, fam, nodes, grp, tags)
    m.group_tri_soup = build_group_obs_soups(m)
    if verbose and surface_dim == 1:
        if flags & 1:
            print('No physical groups of dimension 1 found; created a synthetic "default" group.')
        if flags & 16:
            print("  Reoriented line element(s) to point toward the adjacent surface interior.")
        if flags & 8:
            print("  Warning: some line elements have no curve entity or adjacent surface; their normals "
                  "follow the node order.")
        print(f"Loaded {m.n_elems} curve elements (line3: {int(np.sum(fam == 4))}, line2: "
              f"{int(np.sum(fam == 5))}) in {len(tags)} physical group(s).")
    elif verbose:
        counts = {n: int(np.sum(fam == c)) for n, c in (("quad8", 0), ("quad4", 2), ("tri6", 1), ("tri3", 3))}
        if flags & 1:
            print('No physical groups of dimension 2 found; created a synthetic "default" group.')
        if flags & 2:
            print("Quad9 (type 10) found — centre node dropped, treated as Quad8.")
        if flags & 4:
            print(f'Requested group_field "{group_field}" not found as a per-cell array; using a single '
                  '"default" group.')
        print(f"Loaded {m.n_elems} surface elements (quad8: {counts['quad8']}, quad4: {counts['quad4']}, "
              f"tri6: {counts['tri6']}, tri3: {counts['tri3']}) in {len(tags)} physical group(s).")
        if reverse_normals:
            print("  All normals reversed.")
    return m


def load_vtu(filename: str, surface_dim: int = 2, reverse_normals: bool = False, verbose: bool = True,
             group_field: Optional[str] = None) -> MeshData:
    """Mirror of load_vtu (src/MeshIO.jl:190-206): XML VTK through the same
    native reader."""
    return load_mesh(filename, surface_dim, reverse_normals, verbose, group_field)


# ----------------------------------------------------------------------------
# Generators
# ----------------------------------------------------------------------------

def _plate(n: int, z: float, flip: bool, x0=0.0, y0=0.0, size=1.0, node_base=0):
    xs = x0 + size * np.arange(n + 1) / n
    ys = y0 + size * np.arange(n + 1) / n
    X, Y = np.meshgrid(xs, ys, indexing="xy")  # node (ix, iy) -> iy*(n+1)+ix
    coords = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], axis=1)
    iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = iy * (n + 1) + ix
    b = a + 1
    c = a + n + 2
    d = a + n + 1
    if flip:  # mirrored winding (test/runtests.jl:279-284): normal -z
        quad = np.stack([b, a, d, c], axis=-1).reshape(-1, 4)
    else:
        quad = np.stack([a, b, c, d], axis=-1).reshape(-1, 4)
    return coords, quad + 1 + node_base


def parallel_plates(n: int = 16, gap: float = 1.0) -> MeshData:
    """Two coaxial unit squares ``gap`` apart, each an ``n x n`` Quad4 grid,
    normals facing each other; groups 1 (bottom, +z) and 2 (top, -z)."""
    c1, q1 = _plate(n, 0.0, False)
    c2, q2 = _plate(n, gap, True, node_base=c1.shape[0])
    coords = np.concatenate([c1, c2])
    quads = np.concatenate([q1, q2])
    N = quads.shape[0]
    nodes = np.zeros((N, 8), np.int32)
    nodes[:, :4] = quads
    group = np.concatenate([np.full(q1.shape[0], 1), np.full(q2.shape[0], 2)]).astype(np.int32)
    return MeshData(coords, np.full(N, 2, np.int8), nodes, group, {1: "bottom", 2: "top"})


def unit_cube(n: int = 48) -> MeshData:
    """Interior of the unit cube: 6 faces, each an ``n x n`` Quad4 grid, nodes
    shared along cube edges, inward normals, groups 1..6."""
    node_id = {}
    coords = []

    def nid(p):
        k = node_id.get(p)
        if k is None:
            k = len(coords) + 1
            node_id[p] = k
            coords.append(p)
        return k

    elems = []
    groups = []
    names = {}
    g = 0
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        for v in (0, n):
            g += 1
            names[g] = f"{'xyz'[a]}{'0' if v == 0 else '1'}"
            for w in range(n):
                for u in range(n):
                    def P(uu, ww):
                        p = [0, 0, 0]
                        p[a] = v; p[b] = uu; p[c] = ww
                        return tuple(p)
                    n1, n2, n3, n4 = nid(P(u, w)), nid(P(u + 1, w)), nid(P(u + 1, w + 1)), nid(P(u, w + 1))
                    if v == 0:
                        elems.append((n1, n2, n3, n4))
                    else:
                        elems.append((n2, n1, n4, n3))
                    groups.append(g)
    coords = np.array(coords, np.float64) / n
    N = len(elems)
    nodes = np.zeros((N, 8), np.int32)
    nodes[:, :4] = np.array(elems)
    return MeshData(coords, np.full(N, 2, np.int8), nodes, np.array(groups, np.int32), names)


def icosphere(subdiv: int, radius: float, center, seed: int, jitter: float = 0.02):
    """Tri3 icosphere with ``20 * 4**subdiv`` triangles; vertex radii are
    ``radius * (1 + U(-jitter, jitter))`` from a seeded generator.  Returns
    (vertices (V,3), triangles (T,3) 0-based, outward winding)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, np.float64) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
         (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = V[i] + V[j]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        nf = []
        for (i, j, k) in F:
            a, b, c = mid(i, j), mid(j, k), mid(k, i)
            nf += [(i, a, c), (j, b, a), (k, c, b), (a, b, c)]
        F = nf
    V = np.array(V)
    rng = np.random.default_rng(seed)
    r = radius * (1.0 + rng.uniform(-jitter, jitter, size=V.shape[0]))
    V = np.asarray(center, np.float64) + V * r[:, None]
    return V, np.array(F, np.int64)


def occluded_plates(n: int = 96, subdiv: int = 5, radius: float = 0.3, seed: int = 0x5EED0003,
                    soup_layout: str = "xyz_vertex") -> MeshData:
    """Two unit plates 1.0 apart (``n x n`` Quad4 each, groups 1, 2) with an
    obstruction-only Tri3 sphere (group 3) centred between them."""
    m = parallel_plates(n, 1.0)
    V, T = icosphere(subdiv, radius, (0.5, 0.5, 0.5), seed)
    tri = V[T]  # (T, vertex, xyz)
    if soup_layout == "xyz_vertex":
        soup = tri  # soup[t, v, c]
    else:
        soup = np.ascontiguousarray(np.transpose(tri, (0, 2, 1)))  # soup[t, c, v]
    m.group_tri_soup = {3: np.ascontiguousarray(soup)}
    m.soup_layout = soup_layout
    return m


def box_occluder(lo, hi, n_per_edge: int):
    """Axis-aligned box surface as Tri3 triangles (outward), (T, vertex, xyz)."""
    lo = np.asarray(lo, np.float64); hi = np.asarray(hi, np.float64)
    tris = []
    k = n_per_edge
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        for side in (0, 1):
            for w in range(k):
                for u in range(k):
                    def P(uu, ww):
                        p = np.empty(3)
                        p[a] = hi[a] if side else lo[a]
                        p[b] = lo[b] + (hi[b] - lo[b]) * uu / k
                        p[c] = lo[c] + (hi[c] - lo[c]) * ww / k
                        return p
                    p1, p2, p3, p4 = P(u, w), P(u + 1, w), P(u + 1, w + 1), P(u, w + 1)
                    if side:
                        tris += [(p1, p2, p3), (p1, p3, p4)]
                    else:
                        tris += [(p2, p1, p4), (p2, p4, p3)]
    return np.array(tris)


def cylinder(n_around: int = 128, n_along: int = 64, radius: float = 0.5, height: float = 2.0,
             cap_tris: int = 3000, seed: int = 0x5EED0004) -> MeshData:
    """Interior of a closed cylinder (axis z): Quad8 side with shared seam
    nodes and mid-side nodes on the exact surface, Tri6 Delaunay caps whose
    rim edges share the side's mid-side nodes; inward normals.  Groups:
    1 side, 2 bottom cap, 3 top cap."""
    from scipy.spatial import Delaunay

    coords = []
    node_id = {}

    def nid(key, p):
        k = node_id.get(key)
        if k is None:
            coords.append(p)
            k = len(coords)
            node_id[key] = k
        return k

    # side: angle index a in [0, 2*n_around) half-steps, z index h in [0, 2*n_along] half-steps
    def side_node(a2, h2):
        a2 %= 2 * n_around
        th = np.pi * a2 / n_around
        z = height * h2 / (2 * n_along)
        key = ("rim" if h2 in (0, 2 * n_along) else "side", a2, h2)
        return nid(key, (radius * np.cos(th), radius * np.sin(th), z))

    elems = []
    groups = []
    fams = []
    for h in range(n_along):
        for a in range(n_around):
            A, H = 2 * a, 2 * h
            # inward normal: dxi along -theta, deta along +z -> (-t) x z points to the axis
            c1 = side_node(A + 2, H); c2 = side_node(A, H); c3 = side_node(A, H + 2); c4 = side_node(A + 2, H + 2)
            m5 = side_node(A + 1, H); m6 = side_node(A, H + 1); m7 = side_node(A + 1, H + 2); m8 = side_node(A + 2, H + 1)
            elems.append([c1, c2, c3, c4, m5, m6, m7, m8]); groups.append(1); fams.append(0)

    rng = np.random.default_rng(seed)
    n_rim = n_around
    n_int = (cap_tris + 2 - n_rim) // 2
    for cap, z in ((2, 0.0), (3, height)):
        # seeded interior points, rejection-sampled inside 0.97 r
        pts = []
        while len(pts) < n_int:
            p = rng.uniform(-radius, radius, size=2)
            if p @ p < (0.97 * radius) ** 2:
                pts.append(p)
        pts = np.array(pts)
        th = 2 * np.pi * np.arange(n_rim) / n_rim
        rim = np.stack([radius * np.cos(th), radius * np.sin(th)], axis=1)
        P = np.concatenate([rim, pts])
        tri = Delaunay(P).simplices
        ids = []
        for k in range(P.shape[0]):
            if k < n_rim:
                ids.append(side_node(2 * k, 0 if z == 0.0 else 2 * n_along))
            else:
                ids.append(nid(("cap", cap, k), (P[k, 0], P[k, 1], z)))
        for t in tri:
            a, b, c = [int(x) for x in t]
            # orient: bottom cap normal +z, top cap normal -z
            ab = P[b] - P[a]; ac = P[c] - P[a]
            cross = ab[0] * ac[1] - ab[1] * ac[0]
            if (cross < 0) == (cap == 2):
                b, c = c, b
            corner = [a, b, c]
            mids = []
            for (u, v) in ((corner[0], corner[1]), (corner[1], corner[2]), (corner[0], corner[2])):
                if u < n_rim and v < n_rim and ((u - v) % n_rim in (1, n_rim - 1)):
                    # rim arc: share the side's mid-side node
                    k2 = (2 * u + 1) if (v - u) % n_rim == 1 else (2 * v + 1)
                    mids.append(side_node(k2 % (2 * n_around), 0 if z == 0.0 else 2 * n_along))
                else:
                    key = ("capmid", cap, min(u, v), max(u, v))
                    pm = 0.5 * (P[u] + P[v])
                    mids.append(nid(key, (pm[0], pm[1], z)))
            elems.append([ids[corner[0]], ids[corner[1]], ids[corner[2]], mids[0], mids[1], mids[2], 0, 0])
            groups.append(cap); fams.append(1)
    coords = np.array(coords, np.float64)
    return MeshData(coords, np.array(fams, np.int8), np.array(elems, np.int32), np.array(groups, np.int32),
                    {1: "side", 2: "bottom", 3: "top"})


def sphere_interior(n: int = 4, radius: float = 1.0, order: int = 2) -> MeshData:
    """Interior of a sphere as a cubed sphere: 6 faces of ``n x n`` Quad8
    (``order=2``) or Quad4 (``order=1``) elements whose nodes (mid-side nodes
    too) lie on the sphere, shared between elements, inward normals; groups
    1..6 by cube face.  Between two points of a sphere's interior the kernel
    cos(t_i) cos(t_j) / (pi r^2) is 1 / (4 pi R^2), so F_ij -> A_j / (4 pi R^2)
    for every pair, the element with itself included."""
    node_id = {}
    coords = []

    def nid(p):  # p: integer point of the doubled grid [0, 2n]^3
        k = node_id.get(p)
        if k is None:
            k = len(coords) + 1
            node_id[p] = k
            x = np.array(p, np.float64) / n - 1.0
            coords.append(radius * x / np.sqrt(x @ x))
        return k

    elems, groups, names = [], [], {}
    g = 0
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        for v in (0, 2 * n):
            g += 1
            names[g] = f"{'xyz'[a]}{'-' if v == 0 else '+'}"
            for w in range(0, 2 * n, 2):
                for u in range(0, 2 * n, 2):
                    def P(uu, ww):
                        q = [0, 0, 0]
                        q[a] = v; q[b] = uu; q[c] = ww
                        return nid(tuple(q))
                    c1, c2, c3, c4 = P(u, w), P(u + 2, w), P(u + 2, w + 2), P(u, w + 2)
                    m12, m23, m34, m41 = P(u + 1, w), P(u + 2, w + 1), P(u + 1, w + 2), P(u, w + 1)
                    if v == 0:  # face at -1 along a: inward normal +e_a
                        e = [c1, c2, c3, c4, m12, m23, m34, m41]
                    else:
                        e = [c2, c1, c4, c3, m12, m41, m34, m23]
                    elems.append(e if order == 2 else e[:4] + [0, 0, 0, 0])
                    groups.append(g)
    if order == 1:  # drop the mid-side nodes
        used = sorted({k for e in elems for k in e[:4]})
        remap = {k: r + 1 for r, k in enumerate(used)}
        coords = [coords[k - 1] for k in used]
        elems = [[remap[k] for k in e[:4]] + [0, 0, 0, 0] for e in elems]
    N = len(elems)
    return MeshData(np.array(coords, np.float64), np.full(N, 0 if order == 2 else 2, np.int8),
                    np.array(elems, np.int32), np.array(groups, np.int32), names)


# ----------------------------------------------------------------------------
# Curve meshes (mesh_dim 1, per unit depth): Line3 = [left, right, midpoint],
# Line2 = [left, right]; the normal is the tangent turned 90 degrees counter-
# clockwise in xy (src/Geometry.jl:219-221), so walking a closed curve
# counter-clockwise gives inward normals.
# ----------------------------------------------------------------------------
def _polyline(points, order: int, closed: bool, node_base: int, mid=None):
    """Elements along ``points`` (K, 3); ``mid(k)`` gives element k's midpoint
    node for Line3 (default: the chord midpoint).  Returns coords, nodes."""
    P = np.asarray(points, np.float64)
    K = P.shape[0]
    nel = K if closed else K - 1
    coords = [P]
    nodes = np.zeros((nel, 8), np.int32)
    for k in range(nel):
        a, b = k, (k + 1) % K
        nodes[k, 0], nodes[k, 1] = node_base + a + 1, node_base + b + 1
    if order == 2:
        M = np.array([mid(k) if mid is not None else 0.5 * (P[k] + P[(k + 1) % K]) for k in range(nel)])
        coords.append(M)
        nodes[:, 2] = node_base + K + 1 + np.arange(nel)
    return np.concatenate(coords), nodes


def parallel_lines(n: int = 16, gap: float = 1.0, width: float = 1.0, order: int = 1,
                   offset: float = 0.0) -> MeshData:
    """Two parallel strips of ``width`` (per unit depth), ``gap`` apart, each
    ``n`` Line2 (order 1) or Line3 (order 2) elements, normals facing each
    other; the top strip is shifted by ``offset`` along x.  Groups 1 (bottom,
    normal +y), 2 (top, normal -y)."""
    xs = np.linspace(0.0, width, n + 1)
    bot = np.stack([xs, np.zeros_like(xs), np.zeros_like(xs)], axis=1)           # walks +x: normal +y
    top = np.stack([offset + xs[::-1], np.full_like(xs, gap), np.zeros_like(xs)], axis=1)  # walks -x: normal -y
    c1, n1 = _polyline(bot, order, False, 0)
    c2, n2 = _polyline(top, order, False, c1.shape[0])
    fam = 5 if order == 1 else 4
    N = n1.shape[0] + n2.shape[0]
    return MeshData(np.concatenate([c1, c2]), np.full(N, fam, np.int8), np.concatenate([n1, n2]),
                    np.array([1] * n1.shape[0] + [2] * n2.shape[0], np.int32), {1: "bottom", 2: "top"})


def circle_interior(n: int = 32, radius: float = 1.0, order: int = 2, groups: int = 1) -> MeshData:
    """Interior of a circle as ``n`` arcs: Line3 with every node on the circle
    (order 2) or its inscribed polygon of Line2 chords (order 1); inward
    normals.  Element k spans angles [2 pi k / n, 2 pi (k+1) / n]; groups
    1..``groups`` split the circle into equal runs.  Between two arcs of a
    circle Hottel's crossed strings give F_ij exactly, and an arc sees itself
    with F_ii = 1 - chord / arc."""
    th = 2 * np.pi * np.arange(n) / n
    P = radius * np.stack([np.cos(th), np.sin(th), np.zeros(n)], axis=1)

    def mid(k):
        t = 2 * np.pi * (k + 0.5) / n
        return (radius * np.cos(t), radius * np.sin(t), 0.0)
    coords, nodes = _polyline(P, order, True, 0, mid)
    grp = (np.arange(n) * groups // n + 1).astype(np.int32)
    return MeshData(coords, np.full(n, 5 if order == 1 else 4, np.int8), nodes, grp,
                    {g: f"arc{g}" for g in range(1, groups + 1)})


def square_enclosure_2d(n: int = 16, order: int = 1, baffle: bool = True) -> MeshData:
    """Interior of the unit square, ``n`` elements per side (groups 1..4:
    bottom, right, top, left; inward normals), and optionally an
    obstruction-only baffle (group 5 segment soup): the vertical strip
    x = 0.5, y in [0.2, 0.8], 8 segments."""
    s = np.linspace(0.0, 1.0, n + 1)[:-1]
    z = np.zeros(n)
    sides = [np.stack([s, z, z], 1), np.stack([1 + z, s, z], 1), np.stack([1 - s, 1 + z, z], 1),
             np.stack([z, 1 - s, z], 1)]
    P = np.concatenate(sides)  # counter-clockwise: inward normals
    coords, nodes = _polyline(P, order, True, 0)
    grp = np.repeat(np.arange(1, 5), n).astype(np.int32)
    m = MeshData(coords, np.full(4 * n, 5 if order == 1 else 4, np.int8), nodes, grp,
                 {1: "bottom", 2: "right", 3: "top", 4: "left"})
    if baffle:
        y = np.linspace(0.2, 0.8, 9)
        seg = np.stack([np.stack([0.5 + 0 * y[:-1], y[:-1], 0 * y[:-1]], 1),
                        np.stack([0.5 + 0 * y[1:], y[1:], 0 * y[1:]], 1)], axis=1)
        m.group_tri_soup = {5: seg}
    return m


def random_box_enclosure(rng: np.random.Generator, n_elems: int, max_occluders: int = 3,
                         max_occluder_tris: int = 8000):
    """One serving request: a closed box (sides U[0.5, 2]) meshed with about
    ``n_elems`` elements, a random family per face from {Quad4, Tri3, Quad8},
    inward normals, groups 1..6; plus 0..3 axis-aligned Tri3 box occluders as
    obstruction-only groups 7.. (each <= ``max_occluder_tris`` triangles).
    Returns (mesh, obstruction_groups, nquad)."""
    L = rng.uniform(0.5, 2.0, size=3)
    fams = rng.integers(0, 3, size=6)  # 0 quad4, 1 tri3, 2 quad8
    areas = np.array([L[1] * L[2], L[1] * L[2], L[2] * L[0], L[2] * L[0], L[0] * L[1], L[0] * L[1]])
    per = np.maximum(1, np.round(n_elems * areas / areas.sum()))
    coords = []
    node_id = {}

    def nid(key, p):
        k = node_id.get(key)
        if k is None:
            coords.append(p)
            k = len(coords)
            node_id[key] = k
        return k

    elems, groups, famc = [], [], []
    g = 0
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        for side in (0, 1):
            g += 1
            fam = int(fams[g - 1])
            ne = per[g - 1] / (2 if fam == 1 else 1)
            nb = max(1, int(round(np.sqrt(ne * L[b] / L[c]))))
            nc = max(1, int(round(ne / nb)))

            def P(ub, uc):
                # ub, uc in half-steps
                p = [0.0, 0.0, 0.0]
                p[a] = L[a] if side else 0.0
                p[b] = L[b] * ub / (2 * nb)
                p[c] = L[c] * uc / (2 * nc)
                return nid((g, ub, uc), tuple(p))

            for w in range(nc):
                for u in range(nb):
                    U, W = 2 * u, 2 * w
                    c1, c2, c3, c4 = P(U, W), P(U + 2, W), P(U + 2, W + 2), P(U, W + 2)
                    if side:
                        c1, c2, c3, c4 = c2, c1, c4, c3
                        m5, m6, m7, m8 = P(U + 1, W), P(U, W + 1), P(U + 1, W + 2), P(U + 2, W + 1)
                    else:
                        m5, m6, m7, m8 = P(U + 1, W), P(U + 2, W + 1), P(U + 1, W + 2), P(U, W + 1)
                    if fam == 0:
                        elems.append([c1, c2, c3, c4, 0, 0, 0, 0]); famc.append(2); groups.append(g)
                    elif fam == 2:
                        elems.append([c1, c2, c3, c4, m5, m6, m7, m8]); famc.append(0); groups.append(g)
                    else:
                        elems.append([c1, c2, c3, 0, 0, 0, 0, 0]); famc.append(3); groups.append(g)
                        elems.append([c1, c3, c4, 0, 0, 0, 0, 0]); famc.append(3); groups.append(g)
    mesh = MeshData(np.array(coords, np.float64), np.array(famc, np.int8), np.array(elems, np.int32),
                    np.array(groups, np.int32), {k: f"face{k}" for k in range(1, 7)})
    n_occ = int(rng.integers(0, max_occluders + 1))
    obs = []
    soups = {}
    for k in range(n_occ):
        size = rng.uniform(0.1, 0.3, size=3) * L
        lo = rng.uniform(0.1 * L, 0.9 * L - size)
        kmax = max(1, int(np.sqrt(max_occluder_tris / 12)))
        kk = int(rng.integers(1, kmax + 1))
        soups[7 + k] = box_occluder(lo, lo + size, kk)
        obs.append(7 + k)
    mesh.group_tri_soup = {t: np.ascontiguousarray(s) for t, s in soups.items()}
    mesh.soup_layout = "xyz_vertex"
    nquad = int(rng.integers(2, 6))
    return mesh, obs, nquad
