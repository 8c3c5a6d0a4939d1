"""The colour classes and the colour-parallel accumulate() of the oracle's OpenMP build (the CPU
baseline's structure, num/accumulator.h:161-188 with geom/multigrid.cpp:1238-1319).

Pinned by the definition: every tet in exactly one class, no two tets of a class sharing a vertex,
classes ascending, the greedy choice reproduced by an independent Python colouring; and by the
serial assembly: the same pattern and load-vector support, every matrix entry and load entry equal
to the serial mesh-order sum up to the summation order (each entry has at most one contribution
per colour)."""
import numpy as np
import pytest

import cases
import oracle as O
from test_host_colors import greedy_sizes


@pytest.mark.parametrize("n", [3, 6])
def test_color_classes_against_definition(n):
    mesh = O.make_mesh((0.0, 0.0, 0.0), np.eye(3), (n, n, n))
    start, tets = O.color_classes(mesh)
    sizes = np.diff(start).tolist()
    assert sizes == greedy_sizes(n)
    assert np.array_equal(np.sort(tets), np.arange(6 * n ** 3))
    cell, k = tets // 6, tets % 6
    ijk = np.stack([cell % n, (cell // n) % n, cell // (n * n)], axis=-1)
    verts = ijk[:, None, :] + cases.KUHN[k]
    vid = (verts[..., 2] * (n + 1) + verts[..., 1]) * (n + 1) + verts[..., 0]
    for c in range(len(sizes)):
        sl = slice(start[c], start[c + 1])
        assert np.all(np.diff(tets[sl]) > 0)
        v = vid[sl].ravel()
        assert np.unique(v).size == v.size


@pytest.mark.parametrize("kind", ["parity", "omp"])
def test_colored_assembly_matches_serial(kind):
    mesh, phi = cases.c1_sphere(24)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    dof, nd = O.number_p1if(om, phi)
    ref = O.assemble_p1(om, phi, dof, nd, rhs_kind=O.RHS_3XY_MINUS_X3)
    col = O.assemble_p1_colored(om, phi, dof, nd, O.color_classes(om), rhs_kind=O.RHS_3XY_MINUS_X3, kind=kind)
    assert np.array_equal(col["rowptr"], ref["rowptr"]) and np.array_equal(col["colind"], ref["colind"])
    for key in ("mass", "lb", "stab", "rhs"):
        a, b = col[key], ref[key]
        # up to 24 contributions per entry, summed in another order
        assert np.all(np.abs(a - b) <= 1e-13 * np.abs(b).max()), key
