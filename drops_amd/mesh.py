"""Kuhn brick meshes and P2 level sets on them.

A brick of n1 x n2 x n3 cells, each split into 6 Kuhn tetrahedra with the
vertex order of BrickBuilderCL (geom/builder.cpp:99-178).  Every Kuhn tet
edge joins two cell corners whose index offset is a 0/1 vector, so the P2
nodes of the whole mesh (vertices and edge midpoints) are exactly the points
of the half-spaced grid: index j in 0..2n along each axis, vertex for even
j, edge midpoint for odd j.  A P2 level set (LevelsetP2CL::Phi) is therefore
stored as one dense float64 array of shape (2n3+1, 2n2+1, 2n1+1), the layout
the GPU library consumes.
"""
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class KuhnMesh:
    orig: tuple
    e: tuple          # three edge vectors of the brick (BrickBuilderCL e1, e2, e3)
    n: tuple          # cells per direction (n1, n2, n3)

    @staticmethod
    def box(lo, hi, n, shift=(0.0, 0.0, 0.0)):
        """Axis-aligned brick [lo, hi]^3 moved by -shift (the mesh transform
        p -> p - shift of surfphasesep/surfphasesep.cpp:162-166)."""
        lo = np.broadcast_to(np.asarray(lo, np.float64), (3,))
        hi = np.broadcast_to(np.asarray(hi, np.float64), (3,))
        n = tuple(int(v) for v in np.broadcast_to(np.asarray(n), (3,)))
        orig = tuple(float(lo[c] - shift[c]) for c in range(3))
        e = tuple(tuple(float(hi[k] - lo[k]) if c == k else 0.0 for c in range(3)) for k in range(3))
        return KuhnMesh(orig, e, n)

    @property
    def h(self):
        return float(np.linalg.norm(self.e[0]) / self.n[0])

    @property
    def num_tets(self):
        return 6 * self.n[0] * self.n[1] * self.n[2]

    @property
    def num_vertices(self):
        return (self.n[0] + 1) * (self.n[1] + 1) * (self.n[2] + 1)

    @property
    def halfgrid_shape(self):
        return (2 * self.n[2] + 1, 2 * self.n[1] + 1, 2 * self.n[0] + 1)

    def vertex_coords(self):
        """(n3+1, n2+1, n1+1, 3): orig + i1*off1 + i2*off2 + i3*off3, off_k = 1/n_k * e_k,
        summed left to right as geom/builder.cpp:116-119."""
        n1, n2, n3 = self.n
        out = np.empty((n3 + 1, n2 + 1, n1 + 1, 3))
        i1 = np.arange(n1 + 1, dtype=np.float64)[None, None, :]
        i2 = np.arange(n2 + 1, dtype=np.float64)[None, :, None]
        i3 = np.arange(n3 + 1, dtype=np.float64)[:, None, None]
        for c in range(3):
            o1 = 1.0 / float(n1) * self.e[0][c]
            o2 = 1.0 / float(n2) * self.e[1][c]
            o3 = 1.0 / float(n3) * self.e[2][c]
            out[..., c] = ((self.orig[c] + i1 * o1) + i2 * o2) + i3 * o3
        return out

    def halfgrid_coords(self):
        """World coordinates of all P2 nodes: vertex, or the barycentre .5*(a+b) of the Kuhn
        edge a-b (misc/container.h:395-400).  Returns three arrays of halfgrid_shape."""
        V = self.vertex_coords()
        shp = self.halfgrid_shape
        ja = [np.arange(s) // 2 for s in shp]
        jb = [(np.arange(s) + 1) // 2 for s in shp]
        A = V[np.ix_(ja[0], ja[1], ja[2])]
        B = V[np.ix_(jb[0], jb[1], jb[2])]
        X = 0.5 * (A + B)  # exact for vertices (A == B)
        return X[..., 0], X[..., 1], X[..., 2]

    def halfgrid_axes(self):
        """For an axis-aligned brick: the 1-D half-grid coordinates (x along j1, y along j2, z along
        j3), equal to halfgrid_coords() broadcast (the other edge vectors add exact zeros)."""
        for k in range(3):
            for c in range(3):
                if c != k and self.e[k][c] != 0.0:
                    raise ValueError("halfgrid_axes needs an axis-aligned brick")
        axes = []
        for k in range(3):
            nk = self.n[k]
            i = np.arange(nk + 1, dtype=np.float64)
            v = self.orig[k] + i * (1.0 / float(nk) * self.e[k][k])
            j = np.arange(2 * nk + 1)
            axes.append(0.5 * (v[j // 2] + v[(j + 1) // 2]))
        return axes

    def sample_p2(self, fn):
        """P2 nodal interpolation of fn(x, y, z) (LevelsetP2CL::Init / VecDescCL::Interpolate)."""
        x, y, z = self.halfgrid_coords()
        return np.ascontiguousarray(fn(x, y, z), dtype=np.float64)


# ---- level-set functions used by the configurations and the reference's tests ----

def sphere(center=(0.0, 0.0, 0.0), radius=1.0):
    """phi = |x - c|^2 - r^2 (quadratic, reproduced exactly by P2)."""
    cx, cy, cz = center

    def f(x, y, z):
        dx, dy, dz = x - cx, y - cy, z - cz
        return (dx * dx + dy * dy + dz * dz) - radius * radius
    return f


def sphere_dist(center=(0.5, 0.5, 0.5), radius=0.31):
    """phi = |x - c| - r (tests/interfaceP1FE.cpp:79-85, sphere_2)."""
    cx, cy, cz = center

    def f(x, y, z):
        dx, dy, dz = x - cx, y - cy, z - cz
        return np.sqrt(dx * dx + dy * dy + dz * dz) - radius
    return f


def plane_x(offset):
    """phi = x - offset (tests/interfaceP1FE.cpp:63-71)."""
    return lambda x, y, z: x - offset


def torus(R=1.0, r=0.3):
    """phi = (sqrt(x^2 + y^2) - R)^2 + z^2 - r^2."""
    def f(x, y, z):
        s = np.sqrt(x * x + y * y) - R
        return (s * s + z * z) - r * r
    return f


def ellipsoid(axes, rotation=None, center=(0.0, 0.0, 0.0)):
    """phi = sum_k (q_k / a_k)^2 - 1 with q = R^T (x - c)."""
    a = np.asarray(axes, np.float64)
    Rm = np.eye(3) if rotation is None else np.asarray(rotation, np.float64)

    def f(x, y, z):
        d = (x - center[0], y - center[1], z - center[2])
        out = -1.0
        for k in range(3):
            q = Rm[0, k] * d[0] + Rm[1, k] * d[1] + Rm[2, k] * d[2]
            out = out + (q / a[k]) ** 2
        return out
    return f


def config_shift(h):
    """Tie-break mesh shift for meshes whose vertices lie exactly on the surface:
    s = 1e-3 * h * (1, 2, 3) / sqrt(14)."""
    return tuple(1e-3 * h * np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0))
