"""Problem sets for the batched solve: many uneven P1 trace-FEM problems on ellipsoids.

The batched configuration: ellipsoids with semi-axes U(0.5, 1.5) and uniformly random rotations,
each on its own Kuhn brick over [-2, 2]^3 with n^3 cells, n drawn uniformly from {nmin..nmax}
(seeded).  Each mesh carries the tie-break shift s = 1e-3 h (1, 2, 3)/sqrt(14) so that no level-set
node is exactly zero.  The operator is the configuration-2 one: mass + Laplace-Beltrami + h times
the normal-gradient stabilisation, load f = 3xy - x^3."""
import numpy as np

from .mesh import KuhnMesh, config_shift, ellipsoid


def random_rotation(rng):
    """Haar-uniform rotation: QR of a Gaussian matrix with the signs of R's diagonal moved into Q."""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def make_problems(nprob=512, seed=0, nmin=16, nmax=64):
    """[(mesh, axes, rotation)] of the batched configuration."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nprob):
        n = int(rng.integers(nmin, nmax + 1))
        axes = rng.uniform(0.5, 1.5, 3)
        rot = random_rotation(rng)
        out.append((KuhnMesh.box(-2.0, 2.0, n, shift=config_shift(4.0 / n)), axes, rot))
    return out


def level_sets(problems):
    """All half-grid level sets concatenated (float64) and each problem's offset."""
    sizes = [int(np.prod(m.halfgrid_shape)) for m, _, _ in problems]
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    phi = np.empty(int(sum(sizes)))
    for (m, axes, rot), o, sz in zip(problems, offsets, sizes):
        x, y, z = m.halfgrid_axes()
        f = ellipsoid(axes, rot)
        phi[o:o + sz] = np.broadcast_to(f(x[None, None, :], y[None, :, None], z[:, None, None]),
                                        m.halfgrid_shape).ravel()
    return phi, offsets
