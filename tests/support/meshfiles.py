"""Test-side writers of mesh files (Gmsh .msh 2.2 / 4.1, ASCII and binary;
XML VTK .vtu in every data encoding) and the surface meshes the reference's
load_mesh tests build with Gmsh (test/runtests.jl:303-395): a unit plate at
element order 1 and 2, the same without physical groups, and two unstructured
plates one apart.  The writers follow the published file-format
specifications and are independent of the library's reader, which the tests
check against them.
"""
from __future__ import annotations

import base64
import struct
import zlib
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

GMSH_NNODES = {1: 2, 2: 3, 3: 4, 8: 3, 9: 6, 10: 9, 15: 1, 16: 8, 4: 4}
GMSH_DIM = {1: 1, 2: 2, 3: 2, 8: 1, 9: 2, 10: 2, 15: 0, 16: 2, 4: 3}


@dataclass
class Entity:
    dim: int
    tag: int
    phys: List[int]
    blocks: List[Tuple[int, np.ndarray]] = field(default_factory=list)  # (gmsh type, (n, nn) node tags)
    node_tags: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))  # nodes classified here


@dataclass
class GmshModel:
    node_tags: np.ndarray                 # (Nn,) in file order
    xyz: np.ndarray                       # (Nn, 3)
    entities: List[Entity]
    names: Dict[Tuple[int, int], str] = field(default_factory=dict)


def _phys_section(m: GmshModel) -> str:
    if not m.names:
        return ""
    s = f"$PhysicalNames\n{len(m.names)}\n"
    for (d, t), n in sorted(m.names.items()):
        s += f'{d} {t} "{n}"\n'
    return s + "$EndPhysicalNames\n"


def write_msh41(m: GmshModel, path: str, binary: bool = False):
    out = bytearray()
    w = out.extend
    w(b"$MeshFormat\n4.1 %d 8\n" % (1 if binary else 0))
    if binary:
        w(struct.pack("<i", 1) + b"\n")
    w(b"$EndMeshFormat\n")
    w(_phys_section(m).encode())
    ents = [[e for e in m.entities if e.dim == d] for d in range(4)]
    w(b"$Entities\n")
    if binary:
        w(struct.pack("<4Q", *[len(x) for x in ents]))
    else:
        w(("%d %d %d %d\n" % tuple(len(x) for x in ents)).encode())
    for d in range(4):
        for e in ents[d]:
            pts = m.xyz[np.isin(m.node_tags, e.node_tags)] if len(e.node_tags) else np.zeros((1, 3))
            lo, hi = pts.min(0), pts.max(0)
            if binary:
                w(struct.pack("<i", e.tag))
                w(struct.pack("<3d", *lo) if d == 0 else struct.pack("<6d", *lo, *hi))
                w(struct.pack("<Q", len(e.phys)) + struct.pack("<%di" % len(e.phys), *e.phys))
                if d > 0:
                    w(struct.pack("<Q", 0))
            else:
                box = " ".join("%.17g" % v for v in (lo if d == 0 else np.concatenate([lo, hi])))
                line = f"{e.tag} {box} {len(e.phys)} " + " ".join(str(p) for p in e.phys)
                if d > 0:
                    line += " 0"
                w((line.strip() + "\n").encode())
    w(b"$EndEntities\n")
    # nodes, one block per entity that classifies nodes
    blocks = [e for e in m.entities if len(e.node_tags)]
    order = np.concatenate([e.node_tags for e in blocks]) if blocks else np.zeros(0, np.int64)
    assert len(order) == len(m.node_tags) and np.array_equal(np.sort(order), np.sort(m.node_tags))
    pos = {int(t): k for k, t in enumerate(m.node_tags)}
    w(b"$Nodes\n")
    hdr = (len(blocks), len(m.node_tags), int(m.node_tags.min()), int(m.node_tags.max()))
    w(struct.pack("<4Q", *hdr) if binary else ("%d %d %d %d\n" % hdr).encode())
    for e in blocks:
        if binary:
            w(struct.pack("<3i", e.dim, e.tag, 0) + struct.pack("<Q", len(e.node_tags)))
            w(np.asarray(e.node_tags, "<u8").tobytes())
            w(np.ascontiguousarray(m.xyz[[pos[int(t)] for t in e.node_tags]], "<f8").tobytes())
        else:
            w(("%d %d 0 %d\n" % (e.dim, e.tag, len(e.node_tags))).encode())
            w("".join("%d\n" % t for t in e.node_tags).encode())
            w("".join("%.17g %.17g %.17g\n" % tuple(m.xyz[pos[int(t)]]) for t in e.node_tags).encode())
    w(b"$EndNodes\n")
    eblocks = [(e, t, n) for e in m.entities for (t, n) in e.blocks]
    nel = sum(len(n) for _, _, n in eblocks)
    w(b"$Elements\n")
    hdr = (len(eblocks), nel, 1, nel)
    w(struct.pack("<4Q", *hdr) if binary else ("%d %d %d %d\n" % hdr).encode())
    tag = 1
    for e, t, n in eblocks:
        if binary:
            w(struct.pack("<3i", e.dim, e.tag, t) + struct.pack("<Q", len(n)))
            rows = np.concatenate([np.arange(tag, tag + len(n))[:, None], n], axis=1)
            w(np.asarray(rows, "<u8").tobytes())
        else:
            w(("%d %d %d %d\n" % (e.dim, e.tag, t, len(n))).encode())
            for k, row in enumerate(n):
                w((("%d " % (tag + k)) + " ".join(str(int(x)) for x in row) + "\n").encode())
        tag += len(n)
    w(b"$EndElements\n")
    with open(path, "wb") as f:
        f.write(bytes(out))


def write_msh22(m: GmshModel, path: str, binary: bool = False):
    """Format 2.2: one element line per physical tag of its entity (as Gmsh
    writes them), tags = (physical, elementary)."""
    out = bytearray()
    w = out.extend
    w(b"$MeshFormat\n2.2 %d 8\n" % (1 if binary else 0))
    if binary:
        w(struct.pack("<i", 1) + b"\n")
    w(b"$EndMeshFormat\n")
    w(_phys_section(m).encode())
    w(b"$Nodes\n%d\n" % len(m.node_tags))
    if binary:
        for t, p in zip(m.node_tags, m.xyz):
            w(struct.pack("<i3d", int(t), *p))
        w(b"\n")
    else:
        w("".join("%d %.17g %.17g %.17g\n" % (t, *p) for t, p in zip(m.node_tags, m.xyz)).encode())
    w(b"$EndNodes\n")
    rows = []
    for e in m.entities:
        for p in (e.phys or [0]):
            for t, n in e.blocks:
                for r in n:
                    rows.append((t, p, e.tag, r))
    w(b"$Elements\n%d\n" % len(rows))
    if binary:
        k = 0
        while k < len(rows):
            t = rows[k][0]
            j = k
            while j < len(rows) and rows[j][0] == t:
                j += 1
            w(struct.pack("<3i", t, j - k, 2))
            for q in range(k, j):
                _, p, ent, r = rows[q]
                w(struct.pack("<3i", q + 1, p, ent) + np.asarray(r, "<i4").tobytes())
            k = j
        w(b"\n")
    else:
        for q, (t, p, ent, r) in enumerate(rows):
            w(("%d %d 2 %d %d " % (q + 1, t, p, ent) + " ".join(str(int(x)) for x in r) + "\n").encode())
    w(b"$EndElements\n")
    with open(path, "wb") as f:
        f.write(bytes(out))


def write_msh(m: GmshModel, path: str, version: str = "4.1", binary: bool = False):
    (write_msh41 if version == "4.1" else write_msh22)(m, path, binary)


# ---------------------------------------------------------------------------
# XML VTK
# ---------------------------------------------------------------------------
VTK_TYPE = {3: 5, 1: 22, 2: 9, 0: 23, 5: 3, 4: 21}  # family code -> VTK cell type


def write_vtu(path: str, xyz: np.ndarray, cells: List[Tuple[int, List[int]]], cell_data: Dict[str, np.ndarray],
              encoding: str = "ascii", compress: bool = False, header64: bool = False, leading_zero: bool = False,
              joint: bool = False):
    """cells: (VTK type, 0-based point ids).  encoding: ascii | binary |
    appended-raw | appended-base64.  leading_zero: offsets with a leading 0
    (length ncells+1), which readers accept too.  joint: inline binary blocks
    encode header and data as one base64 stream instead of two."""
    conn = np.array([i for _, c in cells for i in c], np.int64)
    offs = np.cumsum([len(c) for _, c in cells]).astype(np.int64)
    if leading_zero:
        offs = np.concatenate([[0], offs])
    types = np.array([t for t, _ in cells], np.uint8)
    arrays = [("Points", None, np.asarray(xyz, np.float64).ravel(), "Float64", 3),
              ("Cells", "connectivity", conn, "Int64", 1), ("Cells", "offsets", offs, "Int64", 1),
              ("Cells", "types", types, "UInt8", 1)]
    for nm, v in cell_data.items():
        arrays.append(("CellData", nm, np.asarray(v, np.int32), "Int32", 1))
    hfmt = "<Q" if header64 else "<I"

    def block(raw: bytes):
        if not compress:
            return struct.pack(hfmt, len(raw)), raw
        bs = 1 << 12
        chunks = [raw[i:i + bs] for i in range(0, len(raw), bs)] or [b""]
        comp = [zlib.compress(c) for c in chunks]
        last = len(chunks[-1]) if len(chunks[-1]) != bs else 0
        hdr = struct.pack(hfmt, len(chunks)) + struct.pack(hfmt, bs) + struct.pack(hfmt, last)
        hdr += b"".join(struct.pack(hfmt, len(c)) for c in comp)
        return hdr, b"".join(comp)

    appended = bytearray()
    body = {}
    attrs = {}
    for sec, nm, v, ty, nc in arrays:
        raw = v.astype({"Float64": "<f8", "Int64": "<i8", "UInt8": "u1", "Int32": "<i4"}[ty]).tobytes()
        key = (sec, nm)
        if encoding == "ascii":
            body[key] = " ".join(repr(float(x)) if ty == "Float64" else str(int(x)) for x in v)
            attrs[key] = 'format="ascii"'
        elif encoding == "binary":
            h, d = block(raw)
            body[key] = (base64.b64encode(h + d) if joint else base64.b64encode(h) + base64.b64encode(d)).decode()
            attrs[key] = 'format="binary"'
        else:
            h, d = block(raw)
            piece = h + d if encoding == "appended-raw" else base64.b64encode(h) + base64.b64encode(d)
            attrs[key] = 'format="appended" offset="%d"' % len(appended)
            body[key] = None
            appended += piece
    hdr_attr = ' header_type="UInt64"' if header64 else ' header_type="UInt32"'
    comp_attr = ' compressor="vtkZLibDataCompressor"' if compress else ""
    x = ['<?xml version="1.0"?>',
         f'<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian"{hdr_attr}{comp_attr}>',
         "<UnstructuredGrid>", f'<Piece NumberOfPoints="{len(xyz)}" NumberOfCells="{len(cells)}">']

    def da(sec, nm, ty, nc):
        key = (sec, nm)
        name = f' Name="{nm}"' if nm else ""
        comps = f' NumberOfComponents="{nc}"' if nc > 1 else ""
        if body[key] is None:
            return f'<DataArray type="{ty}"{name}{comps} {attrs[key]}/>'
        return f'<DataArray type="{ty}"{name}{comps} {attrs[key]}>\n{body[key]}\n</DataArray>'
    for sec in ("Points", "Cells", "CellData"):
        items = [a for a in arrays if a[0] == sec]
        if not items:
            continue
        x.append(f"<{sec}>")
        x += [da(sec, nm, ty, nc) for _, nm, _, ty, nc in items]
        x.append(f"</{sec}>")
    x += ["</Piece>", "</UnstructuredGrid>"]
    data = ("\n".join(x) + "\n").encode()
    if encoding.startswith("appended"):
        enc = "raw" if encoding == "appended-raw" else "base64"
        data += f'<AppendedData encoding="{enc}">\n_'.encode() + bytes(appended) + b"\n</AppendedData>\n"
    data += b"</VTKFile>\n"
    with open(path, "wb") as f:
        f.write(data)


# ---------------------------------------------------------------------------
# Meshes of the reference's load_mesh tests
# ---------------------------------------------------------------------------
def _square_points(h: float, rng: np.random.Generator, z: float):
    """Boundary points every h plus jittered interior points of the unit square."""
    n = int(round(1.0 / h))
    t = np.linspace(0.0, 1.0, n + 1)
    edge = [(x, 0.0) for x in t[:-1]] + [(1.0, y) for y in t[:-1]] + [(x, 1.0) for x in t[::-1][:-1]] + \
           [(0.0, y) for y in t[::-1][:-1]]
    inner = []
    for i in range(1, n):
        for j in range(1, n):
            inner.append((t[i] + rng.uniform(-0.25, 0.25) * h, t[j] + rng.uniform(-0.25, 0.25) * h))
    P = np.array(edge + inner)
    return np.column_stack([P, np.full(len(P), z)]), len(edge)


def _triangulate(P2):
    from scipy.spatial import Delaunay
    T = Delaunay(P2).simplices.astype(np.int64)
    # counter-clockwise in the xy plane (normal +z)
    a, b, c = P2[T[:, 0]], P2[T[:, 1]], P2[T[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    T[cross < 0] = T[cross < 0][:, [0, 2, 1]]
    return T


def _order2(xyz: List, tris: np.ndarray, quads: np.ndarray):
    """Mid-side nodes shared between neighbours (Gmsh ordering: corners, then
    edge midpoints 1-2, 2-3, ...)."""
    mids = {}

    def mid(a, b):
        k = (min(a, b), max(a, b))
        if k not in mids:
            mids[k] = len(xyz)
            xyz.append(0.5 * (np.asarray(xyz[a]) + np.asarray(xyz[b])))
        return mids[k]
    t6 = [list(t) + [mid(t[0], t[1]), mid(t[1], t[2]), mid(t[2], t[0])] for t in tris]
    q8 = [list(q) + [mid(q[0], q[1]), mid(q[1], q[2]), mid(q[2], q[3]), mid(q[3], q[0])] for q in quads]
    return np.array(t6, np.int64).reshape(-1, 6), np.array(q8, np.int64).reshape(-1, 8)


def plate_model(order: int, physical: bool = True, seed: int = 0x5EED0101, tag_offset: int = 100) -> GmshModel:
    """Unit plate at z=0 meshed with triangles and quadrangles (left half
    quads, right half a triangulation), element order 1 or 2, node tags
    offset and shuffled to exercise the tag -> index map.  Points, curves and
    the surface are separate entities like a Gmsh model's."""
    rng = np.random.default_rng(seed)
    xyz = []
    quads = []
    n = 2
    idx = {}
    for i in range(n + 1):          # left half [0, 0.5] x [0, 1]: 1 x 2 quads of 0.5
        for j in range(3):
            idx[(i, j)] = len(xyz)
            xyz.append((0.25 * i, 0.5 * j, 0.0))
    for i in range(n):
        for j in range(2):
            quads.append([idx[(i, j)], idx[(i + 1, j)], idx[(i + 1, j + 1)], idx[(i, j + 1)]])
    right = [(0.5, 0.0), (0.5, 0.5), (0.5, 1.0), (1.0, 0.0), (1.0, 1.0), (1.0, 0.5), (0.75, 0.25 + 0.1 * rng.random()),
             (0.75, 0.75), (0.75, 0.0), (0.75, 1.0)]
    rid = []
    for p in right:
        k = next((q for q, v in enumerate(xyz) if abs(v[0] - p[0]) < 1e-12 and abs(v[1] - p[1]) < 1e-12), None)
        if k is None:
            k = len(xyz)
            xyz.append((p[0], p[1], 0.0))
        rid.append(k)
    T = _triangulate(np.array(right))
    tris = np.array([[rid[a] for a in t] for t in T], np.int64)
    quads = np.array(quads, np.int64)
    t_type, q_type = 2, 3
    if order == 2:
        tris, quads = _order2(xyz, tris, quads)
        t_type, q_type = 9, 16
    xyz = np.array(xyz, np.float64)
    ntag = np.arange(len(xyz), dtype=np.int64) + tag_offset
    rng.shuffle(ntag)
    # classify nodes: corners on points 1-4, other boundary nodes on curves, the rest on the surface
    corners = {(0.0, 0.0): 1, (1.0, 0.0): 2, (1.0, 1.0): 3, (0.0, 1.0): 4}
    pt_nodes = {k: [] for k in range(1, 5)}
    cv_nodes = {k: [] for k in range(1, 5)}
    sf_nodes = []
    for k, (x, y, _) in enumerate(xyz):
        key = (round(x, 12), round(y, 12))
        if key in corners:
            pt_nodes[corners[key]].append(ntag[k])
        elif abs(y) < 1e-12:
            cv_nodes[1].append(ntag[k])
        elif abs(x - 1) < 1e-12:
            cv_nodes[2].append(ntag[k])
        elif abs(y - 1) < 1e-12:
            cv_nodes[3].append(ntag[k])
        elif abs(x) < 1e-12:
            cv_nodes[4].append(ntag[k])
        else:
            sf_nodes.append(ntag[k])
    ents = [Entity(0, k, [], [(15, np.array([[v] for v in pt_nodes[k]], np.int64))], np.array(pt_nodes[k]))
            for k in range(1, 5)]
    ents += [Entity(1, k, [], [], np.array(cv_nodes[k], np.int64)) for k in range(1, 5)]
    surf = Entity(2, 1, [7] if physical else [],
                  [(q_type, ntag[quads]), (t_type, ntag[tris])], np.array(sf_nodes, np.int64))
    ents.append(surf)
    order_tags = np.concatenate([e.node_tags for e in ents if len(e.node_tags)])
    pos = {int(t): k for k, t in enumerate(ntag)}
    xyz_f = xyz[[pos[int(t)] for t in order_tags]]
    names = {(2, 7): "plate"} if physical else {}
    return GmshModel(order_tags, xyz_f, ents, names)


def two_plates_model(order: int, h: float = 0.2, seed: int = 0x5EED0102) -> GmshModel:
    """Two coaxial unit plates one apart, unstructured triangles, physical
    groups "bottom" (normal +z) and "top" (wound clockwise: normal -z)."""
    rng = np.random.default_rng(seed)
    xyz = []
    ents = []
    tag = 1
    for s, z in ((1, 0.0), (2, 1.0)):
        P, _ = _square_points(h, rng, z)
        T = _triangulate(P[:, :2])
        if z > 0:
            T = T[:, [0, 2, 1]]
        base = len(xyz)
        xyz += [tuple(p) for p in P]
        tris = T + base
        etype = 2
        if order == 2:
            t6, _ = _order2(xyz, tris, np.zeros((0, 4), np.int64))
            tris, etype = t6, 9
        ents.append((s, tris, etype))
    xyz = np.array(xyz, np.float64)
    ntag = np.arange(1, len(xyz) + 1, dtype=np.int64)
    model_ents = []
    used = set()
    for s, tris, etype in ents:
        nodes = sorted(set(int(x) for x in tris.ravel()) - used)
        used |= set(nodes)
        model_ents.append(Entity(2, s, [s], [(etype, ntag[tris])], ntag[nodes]))
    order_tags = np.concatenate([e.node_tags for e in model_ents])
    names = {(2, 1): "bottom", (2, 2): "top"}
    return GmshModel(order_tags, xyz[order_tags - 1], model_ents, names)


def square_curves_model(order: int, n: int = 4, seed: int = 0x5EED0103, physical_surface: bool = True,
                        tag_offset: int = 50) -> GmshModel:
    """The unit square as a 2-D Gmsh model: a structured n x n Quad4 / Quad8
    surface (entity 1, physical group 20 "domain" unless physical_surface is
    false) bounded by four curves (entities 1..4: bottom, right, top, left;
    physical groups 11..14) of n Line2 / Line3 elements, corners on point
    entities with point elements.  A seeded half of the line elements is wound
    backwards, so their counter-clockwise normals face out of the square until
    the reader reorients them."""
    rng = np.random.default_rng(seed)
    m2 = 2 * n  # doubled grid: (even, even) corners, one odd index: edge midpoints
    keys, xyz = {}, []

    def node(i, j):
        if (i, j) not in keys:
            keys[(i, j)] = len(xyz)
            xyz.append((i / m2, j / m2, 0.0))
        return keys[(i, j)]
    quads = []
    for a in range(0, m2, 2):
        for b in range(0, m2, 2):
            q = [node(a, b), node(a + 2, b), node(a + 2, b + 2), node(a, b + 2)]
            if order == 2:
                q += [node(a + 1, b), node(a + 2, b + 1), node(a + 1, b + 2), node(a, b + 1)]
            quads.append(q)
    sides = {1: lambda t: (t, 0), 2: lambda t: (m2, t), 3: lambda t: (m2 - t, m2), 4: lambda t: (0, m2 - t)}
    lines = {}
    for c, f in sides.items():  # walking counter-clockwise: the inward-facing order
        rows = []
        for t in range(0, m2, 2):
            a, b = node(*f(t)), node(*f(t + 2))
            r = [a, b] + ([node(*f(t + 1))] if order == 2 else [])
            if rng.random() < 0.5:
                r[0], r[1] = r[1], r[0]
            rows.append(r)
        lines[c] = np.array(rows, np.int64)
    xyz = np.array(xyz, np.float64)
    ntag = np.arange(len(xyz), dtype=np.int64) + tag_offset
    rng.shuffle(ntag)
    corner = {(0, 0): 1, (m2, 0): 2, (m2, m2): 3, (0, m2): 4}
    pt, cv, sf = {k: [] for k in range(1, 5)}, {k: [] for k in range(1, 5)}, []
    for (i, j), k in keys.items():
        if (i, j) in corner:
            pt[corner[(i, j)]].append(ntag[k])
        elif j == 0:
            cv[1].append(ntag[k])
        elif i == m2:
            cv[2].append(ntag[k])
        elif j == m2:
            cv[3].append(ntag[k])
        elif i == 0:
            cv[4].append(ntag[k])
        else:
            sf.append(ntag[k])
    lt = 8 if order == 2 else 1
    ents = [Entity(0, k, [], [(15, np.array([[v] for v in pt[k]], np.int64))], np.array(pt[k], np.int64))
            for k in range(1, 5)]
    ents += [Entity(1, k, [10 + k], [(lt, ntag[lines[k]])], np.array(cv[k], np.int64)) for k in range(1, 5)]
    ents.append(Entity(2, 1, [20] if physical_surface else [], [(16 if order == 2 else 3, ntag[np.array(quads)])],
                       np.array(sf, np.int64)))
    order_tags = np.concatenate([e.node_tags for e in ents if len(e.node_tags)])
    pos = {int(t): k for k, t in enumerate(ntag)}
    names = {(1, 11): "bottom", (1, 12): "right", (1, 13): "top", (1, 14): "left"}
    if physical_surface:
        names[(2, 20)] = "domain"
    return GmshModel(order_tags, xyz[[pos[int(t)] for t in order_tags]], ents, names)


def model_elements(m: GmshModel):
    """(family code, corner-first node coordinates, physical tag) of every
    surface element, in the reader's order: groups ascending, entities
    ascending, triangle blocks before quadrangle blocks."""
    fam = {2: 3, 3: 2, 9: 1, 16: 0, 10: 0}
    keep = {2: 3, 3: 4, 9: 6, 16: 8, 10: 8}
    pos = {int(t): k for k, t in enumerate(m.node_tags)}
    out = []
    phys = sorted({p for e in m.entities if e.dim == 2 for p in e.phys})
    groups = phys if phys else [None]
    for g in groups:
        for e in sorted((e for e in m.entities if e.dim == 2 and (g is None or g in e.phys)), key=lambda e: e.tag):
            blocks = sorted(e.blocks, key=lambda b: 0 if b[0] in (2, 9) else 1)
            for t, n in blocks:
                if t not in fam:
                    continue
                for row in n:
                    out.append((fam[t], m.xyz[[pos[int(x)] for x in row[:keep[t]]]], g))
    return out
