"""ColorClassesCL of the C++ host mirror (geom/multigrid.cpp:1238-1319) on the CPU: the example
program checks every class (sorted, no shared vertex, every tet once, the greedy choice), and the
class sizes equal those of an independent greedy colouring written here from the definition
(tets in mesh order, smallest colour not carried by one of the tet's four vertices)."""
import os
import re
import subprocess

import numpy as np
import pytest

import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "drops_amd", "lib", "tracefem_p1")


def greedy_sizes(n):
    used = {}
    sizes = {}
    for c3 in range(n):
        for c2 in range(n):
            for c1 in range(n):
                for k in range(6):
                    vs = [(c1 + a, c2 + b, c3 + c) for a, b, c in cases.KUHN[k]]
                    u = 0
                    for v in vs:
                        u |= used.get(v, 0)
                    col = 0
                    while u >> col & 1:
                        col += 1
                    sizes[col] = sizes.get(col, 0) + 1
                    for v in vs:
                        used[v] = used.get(v, 0) | (1 << col)
    return [sizes[c] for c in range(len(sizes))]


@pytest.mark.parametrize("n", [3, 7])
def test_color_classes_match_greedy_definition(n):
    r = subprocess.run([EXE, "colors", str(n)], capture_output=True, text=True, timeout=60,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert r.returncode == 0, r.stderr
    assert re.search(r"valid: 1", r.stdout)
    sizes = [int(x) for x in re.search(r"sizes:(.*)", r.stdout).group(1).split()]
    assert sizes == greedy_sizes(n)
    assert sum(sizes) == 6 * n ** 3
