"""Concave test meshes for self view factors.

``sphere_groups`` is the Quad8 interior of a unit sphere (96 elements) with
groups chosen so that, in slot order, one column chunk holds element 0 alone
and another ends at slot 64: both make a pair tile whose only pair is an
element with itself.  ``occluders`` adds an obstruction-only group 9: a Tri3
icosphere at the centre and a small box just under the x- wall whose faces
cut chords between points of element 10 (geometric soup reading).
"""
import numpy as np

from rvf_amd import mesh as M


def sphere_groups(n: int = 4, occluders: bool = False) -> M.MeshData:
    m = M.sphere_interior(n, 1.0, 2)
    g = np.full(m.n_elems, 3, np.int32)
    g[0] = 1
    g[1:65] = 2
    m.group = g
    m.group_tags = {1: "first", 2: "band", 3: "rest"}
    if occluders:
        V, T = M.icosphere(2, 0.3, (0.0, 0.0, 0.0), 5)
        box = M.box_occluder((-0.99, 0.04, 0.04), (-0.95, 0.2, 0.2), 2)
        m.group_tri_soup = {9: np.ascontiguousarray(np.concatenate([V[T], box]))}
    return m
