"""Meshes read from files through the HIP path: the two unstructured plates of
the reference's load_mesh test (test/runtests.jl:358-395), written as Gmsh
files, solved on the GPU.  FP64 F and A equal the oracle's bit for bit and
F(bottom -> top) = 0.19982 within 5e-3 (the reference's tolerance); mixed
precision within 1e-6 of FP64."""
import numpy as np
import pytest

from rvf_amd import api
from rvf_amd import mesh as M

from support import meshfiles as MF

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [1, 2])
def test_unstructured_plates_from_file(orc, tmp_path, order):
    p = str(tmp_path / "plates.msh")
    MF.write_msh(MF.two_plates_model(order), p, "4.1", True)
    m = M.load_mesh(p, verbose=False)
    r64 = api.compute_view_factors(m, nquad=4, backend=api.MI355XBackend(precision="fp64"), verbose=False)
    o = orc.compute(m, nquad=4)
    F = o["raw"] / o["A"][:, None]
    assert np.array_equal(r64.A_elem, o["A"]) and np.array_equal(r64.F_elem, F)
    b, t = r64.group_names.index("bottom"), r64.group_names.index("top")
    assert abs(r64.F_group[b, t] - 0.19982) <= 5e-3
    assert api.check_reciprocity(r64)
    rmx = api.compute_view_factors(m, nquad=4, backend=api.MI355XBackend(precision="mixed"), verbose=False)
    assert float(np.abs(rmx.F_elem - r64.F_elem).max()) <= 1e-6
