"""Shared problem definitions for the tests (meshes + P2 level sets)."""
import numpy as np

from drops_amd.mesh import KuhnMesh, config_shift, ellipsoid, plane_x, sphere, sphere_dist, torus


def c1_sphere(n=32):
    """Configuration 1 geometry: [-2,2]^3, n^3 cells, phi = |x|^2 - 1, tie-break shift."""
    m = KuhnMesh.box(-2.0, 2.0, n, shift=config_shift(4.0 / n))
    return m, m.sample_p2(sphere())


def torus_case(n=24):
    m = KuhnMesh.box(-2.0, 2.0, n, shift=config_shift(4.0 / n))
    return m, m.sample_p2(torus(1.0, 0.3))


def ellipsoid_case(n=(17, 13, 11), seed=3):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    m = KuhnMesh.box(-2.0, 2.0, n, shift=(1e-4, 2e-4, 3e-4))
    return m, m.sample_p2(ellipsoid(rng.uniform(0.5, 1.5, 3), q))


def skew_case(n=(9, 8, 7)):
    """A sheared parallelepiped brick (general BrickBuilderCL edge vectors)."""
    m = KuhnMesh(orig=(-1.3, -1.1, -1.2), e=((2.6, 0.2, 0.0), (0.1, 2.4, 0.3), (0.0, -0.2, 2.5)), n=n)
    return m, m.sample_p2(sphere((0.05, -0.02, 0.03), 0.83))


def sphere20():
    """tests/interfaceP1FE.cpp:236-255: unit cube, 20^3, phi = |x - 0.5| - 0.31."""
    m = KuhnMesh.box(0.0, 1.0, 20)
    return m, m.sample_p2(sphere_dist())


def area_case():
    """tests/principallattice.cpp:296-330: [-1,1]^3, 8^3, sphere of radius 0.5 (exact zeros)."""
    m = KuhnMesh.box(-1.0, 1.0, 8)
    return m, m.sample_p2(lambda x, y, z: np.sqrt(x * x + y * y + z * z) - 0.5)


def planes():
    """tests/interfaceP1FE.cpp:112-135: unit cube 10x1x1, planes x=0.41 and x=0.4."""
    m = KuhnMesh.box(0.0, 1.0, (10, 1, 1))
    return m, m.sample_p2(plane_x(0.41)), m.sample_p2(plane_x(0.4))


PARITY_CASES = {
    "c1_sphere16": lambda: c1_sphere(16),
    "c1_sphere32": lambda: c1_sphere(32),
    "torus24": torus_case,
    "ellipsoid_ragged": ellipsoid_case,
    "skew_brick": skew_case,
    "sphere20": sphere20,
}


KUHN = np.array([[[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1]],
                 [[0, 0, 0], [1, 0, 0], [1, 0, 1], [1, 1, 1]],
                 [[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 1, 1]],
                 [[0, 0, 0], [0, 1, 0], [0, 1, 1], [1, 1, 1]],
                 [[0, 0, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1]],
                 [[0, 0, 0], [0, 0, 1], [0, 1, 1], [1, 1, 1]]])
P2_EDGES = ((0, 1), (0, 2), (1, 2), (0, 3), (1, 3), (2, 3))


def tet_p2_nodes(mesh, tets):
    """Half-grid indices (T, 10) of the P2 nodes of tets, in the order vertices 0..3, then
    edges (0,1), (0,2), (1,2), (0,3), (1,3), (2,3); an edge node sits at doubled coordinates
    v_a + v_b."""
    tets = np.asarray(tets, np.int64)
    n1, n2, _ = mesh.n
    h1, h2 = 2 * n1 + 1, 2 * n2 + 1
    cell, k = tets // 6, tets % 6
    base = np.stack([cell % n1, (cell // n1) % n2, cell // (n1 * n2)], axis=-1) * 2  # (T,3)
    off = KUHN[k]  # (T,4,3)
    rel = [2 * off[:, a] for a in range(4)] + [off[:, a] + off[:, b] for a, b in P2_EDGES]
    j = base[:, None, :] + np.stack(rel, axis=1)  # (T,10,3)
    return (j[..., 2] * h2 + j[..., 1]) * h1 + j[..., 0]
