"""Writes reference_cases.json: the known answers of the reference's own test
suite (test/runtests.jl) as data, one record per assertion group, each with
the line range it comes from, its inputs, the expected result and the
tolerance the reference states.

The values are the ones written in that file (inputs and expectations); no
part of the reference is imported or run.  Run from the repository root:

    python tests/golden/make_reference_cases.py

Test sets not carried here, and where they are covered instead:
  * "load_mesh" (runtests.jl:303-359) and "Unstructured surface mesh
    end-to-end" (runtests.jl:361-416): the reference writes its meshes with
    Gmsh at test time; the same plates are written by tests/support/meshfiles.py
    and checked in tests/test_meshio_cpu.py and tests/test_gpu_meshio.py.
  * "VTK routing" (runtests.jl:439-462): file-format sniffing, checked in
    tests/test_meshio_cpu.py.
"""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "reference_cases.json")

# Quad8 unit square in z = 0, Gmsh node order (runtests.jl:70-86)
QUAD8_SQUARE = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0, 0], [1, 0.5, 0], [0.5, 1, 0], [0, 0.5, 0]]

# Quad8 plates 1 apart, top plate x-mirrored so that its normal is -z
# (runtests.jl:160-184)
PLATES8 = QUAD8_SQUARE + [[1, 0, 1], [0, 0, 1], [0, 1, 1], [1, 1, 1], [0.5, 0, 1], [0, 0.5, 1], [0.5, 1, 1],
                          [1, 0.5, 1]]

# Quad4 plates 1 apart (runtests.jl:280-284, 425-427)
PLATES4 = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [1, 0, 1], [0, 0, 1], [0, 1, 1], [1, 1, 1]]

# the flat Quad8 plates that share their corners with PLATES4 (runtests.jl:291-297)
PLATES4_AS_QUAD8 = PLATES4[0:4] + [[0.5, 0, 0], [1, 0.5, 0], [0.5, 1, 0], [0, 0.5, 0]] + PLATES4[4:8] + \
    [[0.5, 0, 1], [0, 0.5, 1], [0.5, 1, 1], [1, 0.5, 1]]


def cases():
    c = []
    # --- quadrature (runtests.jl:16-48) ---------------------------------------------
    c.append(dict(id="gl_weight_sums", lines="16-25,34-43", kind="gl_sums", n=list(range(1, 9)),
                  expect={"sum_1d": 2.0, "sum_2d": 4.0}, atol=1e-12))
    c.append(dict(id="gl3_x4", lines="27-30", kind="gl_moment", n=3, power=4, expect=2 / 5, atol=1e-12))
    c.append(dict(id="gl6_x10", lines="45-47", kind="gl_moment", n=6, power=10, expect=2 / 11, atol=1e-12))
    # --- Dunavant triangle rules (runtests.jl:258-275) ----------------------------------
    c.append(dict(id="tri_weight_sums", lines="258-268", kind="tri_sums", n=list(range(1, 9)), expect=0.5,
                  atol=1e-6))
    c.append(dict(id="tri4_x3y4", lines="269-275", kind="tri_moment", n=4, a=3, b=4,
                  expect=math.factorial(3) * math.factorial(4) / math.factorial(9), atol=1e-10))
    # --- shape functions (runtests.jl:51-67, 227-256) -----------------------------------
    c.append(dict(id="quad8_shape", lines="51-67", kind="shape", family=0,
                  partition_points=[[-0.5, 0.3], [0.0, 0.0], [1.0, 1.0], [-1.0, -1.0]],
                  nodal_points=[[-1, -1], [1, -1], [1, 1], [-1, 1], [0, -1], [1, 0], [0, 1], [-1, 0]],
                  atol_partition=1e-14, atol_nodal=1e-13))
    c.append(dict(id="quad4_shape", lines="228-239", kind="shape", family=2,
                  partition_points=[[-0.3, 0.7], [0.0, 0.0], [1.0, -1.0], [0.5, 0.5]],
                  nodal_points=[[-1, -1], [1, -1], [1, 1], [-1, 1]], atol_partition=1e-14, atol_nodal=1e-14))
    c.append(dict(id="line2_shape", lines="241-247", kind="shape", family=5,
                  partition_points=[[-0.4, 0.0], [0.0, 0.0], [0.8, 0.0]],
                  nodal_points=[[-1, 0], [1, 0]], atol_partition=1e-14, atol_nodal=1e-14))
    c.append(dict(id="tri3_shape", lines="249-255", kind="shape", family=3,
                  partition_points=[[0.2, 0.3], [0.0, 0.0], [1.0, 0.0], [0.0, 1.0]],
                  nodal_points=[], atol_partition=1e-14, atol_nodal=1e-14))
    # --- element geometry (runtests.jl:81-99) -------------------------------------------
    c.append(dict(id="quad8_square_centre", lines="81-95", kind="eval_point", family=0, coords=QUAD8_SQUARE,
                  nodes=list(range(1, 9)), point=[0.0, 0.0],
                  expect={"x": [0.5, 0.5, 0.0], "n": [0.0, 0.0, 1.0], "dA": 0.25}, atol=1e-14))
    c.append(dict(id="quad8_square_area", lines="96-98", kind="element_area", family=0, coords=QUAD8_SQUARE,
                  nodes=list(range(1, 9)), nquad=4, expect=1.0, atol=1e-12))
    # --- point kernel (runtests.jl:137-153) ---------------------------------------------
    c.append(dict(id="vf_kernel_facing", lines="137-144", kind="vf_kernel",
                  xi=[0, 0, 0], ni=[0, 0, 1], xj=[0, 0, 1], nj=[0, 0, -1], expect=1 / math.pi, atol=1e-14))
    c.append(dict(id="vf_kernel_back", lines="146-148", kind="vf_kernel",
                  xi=[0, 0, 0], ni=[0, 0, 1], xj=[0, 0, 1], nj=[0, 0, 1], expect=0.0, atol=0.0))
    c.append(dict(id="vf_kernel_same_point", lines="150-152", kind="vf_kernel",
                  xi=[0, 0, 0], ni=[0, 0, 1], xj=[0, 0, 0], nj=[0, 0, -1], expect=0.0, atol=0.0))
    # --- element pairs (runtests.jl:160-198, 278-300, 420-436) ---------------------------
    c.append(dict(id="plates_quad8", lines="160-198", kind="pair", coords=PLATES8,
                  elements=[{"nodes": list(range(1, 9)), "group": 1, "family": "quad"},
                            {"nodes": list(range(9, 17)), "group": 2, "family": "quad"}],
                  nquad=4, expect={"A": [1.0, 1.0], "A_atol": 1e-10, "raw_symmetric_atol": 1e-12,
                                   "F01_range": [0.1, 1.0]}))
    c.append(dict(id="plates_quad4", lines="278-289", kind="pair", coords=PLATES4,
                  elements=[{"nodes": [1, 2, 3, 4], "group": 1, "family": "quad4"},
                            {"nodes": [5, 6, 7, 8], "group": 2, "family": "quad4"}],
                  nquad=8, expect={"A": [1.0, 1.0], "A_atol": 1e-10, "F01": 0.19982, "F01_atol": 2e-4}))
    c.append(dict(id="plates_quad8_flat_vs_quad4", lines="291-299", kind="pair", coords=PLATES4_AS_QUAD8,
                  elements=[{"nodes": list(range(1, 9)), "group": 1, "family": "quad"},
                            {"nodes": list(range(9, 17)), "group": 2, "family": "quad"}],
                  nquad=8, expect={"same_F01_as": "plates_quad4", "F01_atol": 1e-6}))
    c.append(dict(id="gpu_hook_quad4", lines="420-436", kind="pair", coords=PLATES4,
                  elements=[{"nodes": [1, 2, 3, 4], "group": 1, "family": "quad4"},
                            {"nodes": [5, 6, 7, 8], "group": 2, "family": "quad4"}],
                  nquad=6, expect={"A": [1.0, 1.0], "A_atol": 1e-9, "F01": 0.19982, "F01_atol": 2e-4}))
    # --- ray casting and visibility (runtests.jl:102-134) ---------------------------------
    # soups as the test fills them, soup[vertex, xyz, tri] in Julia's column
    # order; stored here as rows = vertices of the one triangle
    c.append(dict(id="ray_hit", lines="102-113", kind="ray", soup_vertices=[[0, 0, 1], [1, 0, 1], [0, 1, 1]],
                  origin=[0.25, 0.25, 0.0], direction=[0, 0, 1], tmax=2.0, expect=True))
    c.append(dict(id="ray_miss", lines="115-116", kind="ray", soup_vertices=[[0, 0, 1], [1, 0, 1], [0, 1, 1]],
                  origin=[2.0, 2.0, 0.0], direction=[0, 0, 1], tmax=2.0, expect=False))
    c.append(dict(id="visibility_blocked", lines="120-130", kind="visible",
                  soup_vertices=[[-1, -1, 1], [1, -1, 1], [0, 1, 1]], xi=[0, 0, 0], xj=[0, 0, 2], expect=False))
    c.append(dict(id="visibility_free", lines="132-133", kind="visible",
                  soup_vertices=[[-1, -1, 1], [1, -1, 1], [0, 1, 1]], xi=[0, 0, 0], xj=[5, 0, 2], expect=True))
    # --- result checks (runtests.jl:201-224) --------------------------------------------
    c.append(dict(id="reciprocity_closure_two_surfaces", lines="201-212", kind="checks",
                  F=[[0, 1], [1, 0]], A=[1.0, 1.0], expect={"reciprocity": True, "closure": True}))
    c.append(dict(id="reciprocity_unequal_areas", lines="214-223", kind="checks",
                  F=[[0, 0.5], [1, 0]], A=[2.0, 1.0], expect={"reciprocity": True}))
    return c


def main():
    data = {"source": "test/runtests.jl of rot4te/RadiativeViewFactor.jl", "cases": cases()}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
        f.write("\n")
    print(f"wrote {len(data['cases'])} cases to {OUT}")


if __name__ == "__main__":
    main()
