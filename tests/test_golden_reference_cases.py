"""The reference's own known answers (test/runtests.jl, carried as data in
tests/golden/reference_cases.json) checked three ways:

  * the oracle, on the CPU, for every case;
  * the library's host side, on the CPU: the quadrature tables the device
    integrates with (rvf_quad_rule) and check_reciprocity / check_closure;
  * the device, through the C-ABI: the element-pair cases as two-element
    solves (FP64 and mixed precision), and the ray / visibility cases as
    two-element occluded solves with one quadrature point per element placed
    at the ray's ends (a blocked pair has F = 0 and one blocked ray).

Tolerances are the reference's own, as stored with each case.
"""
import json
import os

import numpy as np
import pytest

from rvf_amd import api
from rvf_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "reference_cases.json")) as _f:
    CASES = json.load(_f)["cases"]
BY_ID = {c["id"]: c for c in CASES}


def of_kind(*kinds):
    return [pytest.param(c, id=c["id"]) for c in CASES if c["kind"] in kinds]


def pair_mesh(c):
    els = [M.SurfaceElement(e["nodes"], e["group"], e["family"]) for e in c["elements"]]
    return M.MeshData.from_elements(np.array(c["coords"], float), els, {1: "a", 2: "b"})


def check_pair(c, A, F):
    """Expectations of an element-pair case on areas A and view factors F."""
    e = c["expect"]
    if "A" in e:
        assert np.allclose(A, e["A"], rtol=0, atol=e["A_atol"])
    if "raw_symmetric_atol" in e:
        assert abs(A[0] * F[0, 1] - A[1] * F[1, 0]) <= e["raw_symmetric_atol"]
        assert A[0] * F[0, 1] > 0
    if "F01_range" in e:
        lo, hi = e["F01_range"]
        assert lo < F[0, 1] < hi
    if "F01" in e:
        assert abs(F[0, 1] - e["F01"]) <= e["F01_atol"], F[0, 1]


def soup_of(c):
    """One-triangle soup as the reference's test fills it: vertex-major
    (soup[t, c, v] in C order)."""
    return np.ascontiguousarray(np.array(c["soup_vertices"], float).T[None])


# ------------------------------------------------------------------------------------------------
# oracle
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("c", of_kind("gl_sums", "gl_moment", "tri_sums", "tri_moment"))
def test_oracle_quadrature(orc, c):
    k = c["kind"]
    if k == "gl_sums":
        for n in c["n"]:
            p, w = orc.gl1d(n)
            qp, qw = orc.quad_rule(n)
            assert len(p) == n and qp.shape == (n * n, 2)
            assert abs(w.sum() - c["expect"]["sum_1d"]) <= c["atol"]
            assert abs(qw.sum() - c["expect"]["sum_2d"]) <= c["atol"]
    elif k == "gl_moment":
        p, w = orc.gl1d(c["n"])
        assert abs(np.dot(w, p ** c["power"]) - c["expect"]) <= c["atol"]
    elif k == "tri_sums":
        for n in c["n"]:
            assert abs(orc.tri_rule(n)[1].sum() - c["expect"]) <= c["atol"]
    else:
        p, w = orc.tri_rule(c["n"])
        assert abs(np.sum(w * p[:, 0] ** c["a"] * p[:, 1] ** c["b"]) - c["expect"]) <= c["atol"]


@pytest.mark.parametrize("c", of_kind("shape"))
def test_oracle_shape(orc, c):
    for x, e in c["partition_points"]:
        assert abs(orc.shape(c["family"], float(x), float(e))[0].sum() - 1.0) <= c["atol_partition"]
    for b, (x, e) in enumerate(c["nodal_points"]):
        N = orc.shape(c["family"], float(x), float(e))[0]
        want = np.eye(len(c["nodal_points"]))[b]
        assert np.allclose(N, want, rtol=0, atol=c["atol_nodal"])


@pytest.mark.parametrize("c", of_kind("eval_point", "element_area"))
def test_oracle_geometry(orc, c):
    coords = np.array(c["coords"], float)
    nodes = np.array(c["nodes"], np.int32)
    if c["kind"] == "eval_point":
        x, n, dA = orc.eval_point(c["family"], coords, nodes, *c["point"])
        e = c["expect"]
        assert np.allclose(x, e["x"], rtol=0, atol=c["atol"]) and np.allclose(n, e["n"], rtol=0, atol=c["atol"])
        assert abs(dA - e["dA"]) <= c["atol"]
    else:
        qp, qw = orc.quad_rule(c["nquad"])
        A = sum(w * orc.eval_point(c["family"], coords, nodes, p[0], p[1])[2] for p, w in zip(qp, qw))
        assert abs(A - c["expect"]) <= c["atol"]


@pytest.mark.parametrize("c", of_kind("vf_kernel"))
def test_oracle_vf_kernel(orc, c):
    k = orc.vf_kernel(*(np.array(c[key], float) for key in ("xi", "ni", "xj", "nj")))
    assert abs(k - c["expect"]) <= c["atol"]


def oracle_pair(orc, c):
    m = pair_mesh(c)
    r = orc.compute(m, nquad=c["nquad"])
    return r["A"], r["raw"] / r["A"][:, None]


@pytest.mark.parametrize("c", of_kind("pair"))
def test_oracle_pairs(orc, c):
    A, F = oracle_pair(orc, c)
    check_pair(c, A, F)
    other = c["expect"].get("same_F01_as")
    if other:
        assert abs(F[0, 1] - oracle_pair(orc, BY_ID[other])[1][0, 1]) <= c["expect"]["F01_atol"]


@pytest.mark.parametrize("c", of_kind("ray", "visible"))
def test_oracle_rays(orc, c):
    soup = soup_of(c)
    # the reference's reading (mode 0) and the geometric reading of a
    # vertex-major soup (mode 2) give the reference's answers on these cases
    for mode in (0, 2):
        if c["kind"] == "ray":
            got = orc.ray_hits(soup, mode, c["origin"], c["direction"], c["tmax"])
        else:
            got = bool(orc.visible_batch(soup, mode, np.array([c["xi"]], float), np.array([c["xj"]], float))[0])
        assert got == c["expect"], (mode, got)


@pytest.mark.parametrize("c", of_kind("checks"))
def test_result_checks(orc, c):
    F, A = np.array(c["F"], float), np.array(c["A"], float)
    res = api.ViewFactorResult(F, A, F, A, [1, 2], ["surface1", "surface2"])
    if "reciprocity" in c["expect"]:
        assert api.check_reciprocity(res) == c["expect"]["reciprocity"]
        assert (orc.reciprocity_error(F, A) < 1e-4) == c["expect"]["reciprocity"]
    if "closure" in c["expect"]:
        assert api.check_closure(res) == c["expect"]["closure"]


# ------------------------------------------------------------------------------------------------
# the library's host-side quadrature tables (what the device integrates with)
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("c", of_kind("gl_sums", "gl_moment", "tri_sums", "tri_moment"))
def test_library_quadrature_tables(orc, c):
    k = c["kind"]
    if k == "gl_sums":
        for n in c["n"]:
            qp, qw = api.quad_rule(n)
            assert qp.shape == (n * n, 2) and abs(qw.sum() - c["expect"]["sum_2d"]) <= c["atol"]
            op, ow = orc.quad_rule(n)
            assert np.array_equal(qp, op) and np.array_equal(qw, ow)
    elif k == "gl_moment":
        # the tensor rule's moment in xi alone is the 1-D moment times sum(w) = 2
        qp, qw = api.quad_rule(c["n"])
        assert abs(np.dot(qw, qp[:, 0] ** c["power"]) - 2 * c["expect"]) <= 2 * c["atol"]
    elif k == "tri_sums":
        for n in c["n"]:
            qp, qw = api.quad_rule(n, tri=True)
            assert abs(qw.sum() - c["expect"]) <= c["atol"]
            op, ow = orc.tri_rule(n)
            assert np.array_equal(qp, op) and np.array_equal(qw, ow)
    else:
        qp, qw = api.quad_rule(c["n"], tri=True)
        assert abs(np.sum(qw * qp[:, 0] ** c["a"] * qp[:, 1] ** c["b"]) - c["expect"]) <= c["atol"]


# ------------------------------------------------------------------------------------------------
# the device, through the C-ABI
# ------------------------------------------------------------------------------------------------
def device_solve(mesh, nquad, precision, obs=(), parity="reference"):
    backend = api.MI355XBackend(precision=precision, parity=parity, audit=len(obs) > 0)
    return api.compute_view_factors(mesh, nquad, obs, backend=backend, verbose=False)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("c", of_kind("pair"))
def test_device_pairs(orc, c, precision):
    r = device_solve(pair_mesh(c), c["nquad"], precision)
    check_pair(c, r.A_elem, r.F_elem)
    oA, oF = oracle_pair(orc, c)
    if precision == "fp64":
        assert np.array_equal(r.A_elem, oA) and np.array_equal(r.F_elem, oF)
    else:
        assert np.abs(r.F_elem - oF).max() <= 1e-6
    other = c["expect"].get("same_F01_as")
    if other:
        o = BY_ID[other]
        F2 = device_solve(pair_mesh(o), o["nquad"], precision).F_elem
        assert abs(r.F_elem[0, 1] - F2[0, 1]) <= c["expect"]["F01_atol"]


def facing_pair(xi, xj, half=1e-3):
    """Two small Quad4 elements centred at xi and xj, each facing the other;
    with one quadrature point per element the integration points are xi and
    xj themselves."""
    xi, xj = np.array(xi, float), np.array(xj, float)
    d = (xj - xi) / np.linalg.norm(xj - xi)
    coords = []
    for c, n in ((xi, d), (xj, -d)):
        a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
        u = np.cross(n, a); u /= np.linalg.norm(u)
        v = np.cross(n, u)  # u x v = n: counter-clockwise about n
        coords += [c - half * u - half * v, c + half * u - half * v, c + half * u + half * v, c - half * u + half * v]
    els = [M.SurfaceElement([1, 2, 3, 4], 1, "quad4"), M.SurfaceElement([5, 6, 7, 8], 2, "quad4")]
    return M.MeshData.from_elements(np.array(coords), els, {1: "i", 2: "j"})


@pytest.mark.gpu
@pytest.mark.parametrize("parity", ["reference", "geometric"])
@pytest.mark.parametrize("c", of_kind("ray", "visible"))
def test_device_visibility(orc, c, parity):
    if c["kind"] == "ray":
        xi = np.array(c["origin"], float)
        xj = xi + c["tmax"] * np.array(c["direction"], float)
        visible = not c["expect"]
    else:
        xi, xj, visible = c["xi"], c["xj"], c["expect"]
    m = facing_pair(xi, xj)
    m.group_tri_soup = {3: soup_of(c)}
    m.soup_layout = "vertex_xyz"
    for precision in ("fp64", "mixed"):
        r = device_solve(m, 1, precision, obs=[3], parity=parity)
        free = device_solve(m, 1, precision)
        assert (r.F_elem[0, 1] > 0) == visible, (precision, r.F_elem[0, 1])
        assert int(r.audit_count[0, 1]) == (0 if visible else 1)
        if visible and precision == "fp64":
            assert r.F_elem[0, 1] == free.F_elem[0, 1]
        elif visible:
            assert abs(r.F_elem[0, 1] / free.F_elem[0, 1] - 1) <= 1e-6
        # the oracle's occluded solve agrees
        o = orc.compute(m, nquad=1, obstruction_groups=[3], parity_mode=parity)
        assert (o["raw"][0, 1] > 0) == visible


def test_fixture_matches_its_generator():
    # the committed fixture is what tests/golden/make_reference_cases.py writes
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_reference_cases",
                                                  os.path.join(HERE, "golden", "make_reference_cases.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    assert json.loads(json.dumps(gen.cases())) == CASES
