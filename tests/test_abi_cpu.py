"""CPU-side checks of the C ABI: the library loads, exports every symbol the
header declares, and its host-side pieces (quadrature tables, BVH builder)
match the oracle bit for bit.  No compute is launched (no GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from rvf_amd import _lib, api
from rvf_amd import mesh as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_header_symbols():
    L = _lib.lib()
    hdr = open(os.path.join(ROOT, "include", "rvf.h")).read()
    names = set(re.findall(r"\b(rvf_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 20
    for n in sorted(names):
        assert hasattr(L, n), n
    assert L.rvf_version() >= 1


@pytest.mark.parametrize("n", range(1, 9))
def test_quad_tables_bitwise(orc, n):
    p, w = api.quad_rule(n)
    op, ow = orc.quad_rule(n)
    assert np.array_equal(p, op) and np.array_equal(w, ow)
    p, w = api.quad_rule(n, tri=True)
    op, ow = orc.tri_rule(n)
    assert np.array_equal(p, op) and np.array_equal(w, ow)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_bvh_builder_bitwise(orc, mode):
    V, T = M.icosphere(3, 0.3, (0.5, 0.5, 0.5), seed=11)
    soup = V[T]
    box, meta, tri = api.bvh_build_host(soup, mode)
    obox, ometa, otri = orc.bvh_export(soup, mode)
    assert np.array_equal(box, obox)
    assert np.array_equal(meta, ometa)
    assert np.array_equal(tri, otri)


def test_bvh_builder_ties_and_degenerate(orc):
    # many identical centroids (stable-sort ties) and flat boxes
    rng = np.random.default_rng(5)
    soup = np.repeat(rng.uniform(0, 1, (7, 3, 3)), 9, axis=0)
    soup[::2, :, 2] = 0.25
    box, meta, tri = api.bvh_build_host(soup, 1)
    obox, ometa, otri = orc.bvh_export(soup, 1)
    assert np.array_equal(box, obox) and np.array_equal(meta, ometa) and np.array_equal(tri, otri)


def test_errors_without_device():
    # invalid arguments are reported through the ABI, never abort
    L = _lib.lib()
    with pytest.raises(_lib.RVFError):
        api.quad_rule(9)
    rc = L.rvf_bvh_build_host(None, 0, 0, 0, None, None, None, None)
    assert rc < 0 and b"bad arguments" in L.rvf_last_error()


def test_family_matrices_layout():
    m = M.cylinder(n_around=8, n_along=2, cap_tris=12)
    mats, counts, idx = api.family_matrices(m)
    for e in range(m.n_elems):
        f = int(m.family[e])
        assert list(mats[f][idx[e] - 1]) == list(m.nodes[e, :(8, 6, 4, 3)[f]])
    assert counts.sum() == m.n_elems


def test_tune_table_round_trip(tmp_path):
    # the autotuner's persisted table: keys bucket sizes to powers of two,
    # entries survive a reload, plans without occluders have no key
    from types import SimpleNamespace as NS
    from rvf_amd import _lib as L
    o = L.Opts(); o.nquad = 3; o.precision = L.PRECISIONS["mixed"]; o.method = L.METHODS["quad"]
    plan = NS(opts=o, dmesh=NS(mesh=NS(n_elems=13824)), dobs=NS(n_tris=20480))
    k = api.TuneTable.key(plan)
    assert k.endswith("|mixed|quad|nq3|N16384|T32768")
    near = NS(opts=o, dmesh=NS(mesh=NS(n_elems=9000)), dobs=NS(n_tris=17000))
    assert api.TuneTable.key(near) == k
    assert api.TuneTable.key(NS(opts=o, dmesh=plan.dmesh, dobs=None)) is None
    t = api.TuneTable(str(tmp_path / "sub" / "t.json"))
    assert k not in t
    t[k] = (api.VARIANT_SPLIT(3), 16, 8, 3)
    t.save()
    t2 = api.TuneTable(str(tmp_path / "sub" / "t.json"))
    assert k in t2 and t2[k] == (api.VARIANT_SPLIT(3), 16, 8, 3)
    # entries written without a tile width or a build read as the kernel's
    # defaults (0)
    import json
    with open(tmp_path / "old.json", "w") as f:
        json.dump({k: [api.VARIANT_QUEUE(1), 16]}, f)
    assert api.TuneTable(str(tmp_path / "old.json"))[k] == (api.VARIANT_QUEUE(1), 16, 0, 0)
    t[k] = (api.VARIANT_QUEUE(2), 8, 4)
    assert t[k] == (api.VARIANT_QUEUE(2), 8, 4, 0)
