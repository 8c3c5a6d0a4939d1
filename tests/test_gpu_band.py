"""GPU parity of the narrow-band assembly (dhip_assemble_p1_band) against the oracle's
restatement of SurfactantNarrowBandStblP1CL::Update: band tets, numbering, CSR pattern
(the stabilisation's), stabilisation values, mass / LB (the reference's own entries bit
for bit, 0.0 on the entries only band tets touch), load vector and the LinComb of the
three, all bit-exact."""
import numpy as np
import pytest

import cases
import oracle as O
from drops_amd import dhip as D
from drops_amd.mesh import KuhnMesh, plane_x

pytestmark = pytest.mark.gpu


def scatter(rp_to, ci_to, rp, ci, val):
    """values of CSR (rp, ci, val) placed on the pattern (rp_to, ci_to); entries absent there
    raise; entries of the target without a source stay 0.0 and are reported by the mask"""
    out = np.zeros(ci_to.size)
    hit = np.zeros(ci_to.size, bool)
    for r in range(rp.size - 1):
        lo, hi = rp_to[r], rp_to[r + 1]
        for q in range(rp[r], rp[r + 1]):
            k = lo + np.searchsorted(ci_to[lo:hi], ci[q])
            assert k < hi and ci_to[k] == ci[q]
            out[k] = val[q]
            hit[k] = True
    return out, hit


CASES = [("c1_sphere16", 0.0), ("c1_sphere16", 0.25), ("ellipsoid_ragged", 0.3), ("skew_brick", 0.12),
         ("torus24", 0.05), ("c1_sphere16", 100.0)]   # the last: every tet of the mesh is in the band


@pytest.mark.parametrize("name,width", CASES)
def test_band_assembly_bit_exact(ctx, name, width):
    mesh, phi = cases.PARITY_CASES[name]()
    coef, lc = (0.7, 1.0, 1.3), (1.0, 0.9, 0.1)
    print(f"[gpu band] {name} width={width}", flush=True)
    ctx.mesh(mesh)
    ctx.levelset(phi)
    nb, nd, nnz = ctx.assemble_p1_band(width, coef=coef, rhs_kind=D.RHS_3XY_MINUS_X3, lc=lc)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    t_ref, cut_ref = O.classify_band(om, phi, width)
    assert nb == t_ref.size and np.array_equal(ctx.cut_tets(), t_ref)
    _, flags = ctx.local()
    assert np.all(flags & 1)
    assert np.array_equal((flags >> 5) & 1, 1 - cut_ref)   # every cut tet here has a patch
    dof, ndr = O.number_p1band(om, phi, width)
    assert nd == ndr
    vod = ctx.dof_vertices()
    assert np.array_equal(dof[vod], np.arange(nd))
    ref = O.assemble_p1band(om, phi, dof, nd, width, coef=coef, rhs_kind=O.RHS_3XY_MINUS_X3)
    g = ctx.csr()
    rps, cs, vs = ref["stab"]
    assert np.array_equal(g["rowptr"], rps) and np.array_equal(g["colind"], cs)
    assert np.array_equal(g["stab"], vs)
    rpm, cm, vm, vl = ref["surf"]
    m_on, hit = scatter(rps, cs, rpm, cm, vm)
    l_on, _ = scatter(rps, cs, rpm, cm, vl)
    assert np.array_equal(g["mass"], m_on) and np.array_equal(g["lb"], l_on)
    assert np.all(g["mass"][~hit] == 0.0) and np.all(g["lb"][~hit] == 0.0)
    assert np.array_equal(g["rhs"], ref["rhs"])
    # LinComb(lc0, M, lc1, A, lc2, C) over the merged patterns (num/spmat.h:1458-1560): entries in
    # M and A get 1.0*(lc0 m + lc1 l) + lc2 s, the others lc2 s
    comb = np.where(hit, 1.0 * (lc[0] * m_on + lc[1] * l_on) + lc[2] * vs, lc[2] * vs)
    assert np.array_equal(g["comb"], comb)


def test_band_planar_closed_form(ctx):
    """plane x = 0.437 on [0,1]^3 (10^3 cells): u'Cu = (band volume)/6 for u = x through the GPU matrix"""
    m = KuhnMesh.box(0.0, 1.0, 10)
    phi = m.sample_p2(plane_x(0.437))
    print("[gpu band] plane", flush=True)
    ctx.mesh(m)
    ctx.levelset(phi)
    nb, nd, nnz = ctx.assemble_p1_band(0.15, lc=(0.0, 0.0, 1.0))
    g = ctx.csr()
    V = m.vertex_coords().reshape(-1, 3)
    u = V[ctx.dof_vertices(), 0]
    A = ctx.csr_from_assembly(D.OP_STAB)
    Cu = ctx.spmv(A, u)
    A.close()
    vol = nb * m.h ** 3 / 6.0
    assert abs(u @ Cu - vol / 6.0) <= 1e-12 * vol
    assert np.array_equal(g["comb"], g["stab"])


def test_band_zero_node_and_state(ctx):
    """an exact zero in a cut tet throws under ZERO_THROW; the P2 steps refuse a band listing"""
    m, p041, p040 = cases.planes()
    ctx.mesh(m)
    ctx.levelset(p040)
    with pytest.raises(D.DhipError) as e:
        ctx.assemble_p1_band(0.1)
    assert e.value.code == D.E_ZERO_NODE
    ctx.levelset(p041)
    ctx.assemble_p1_band(0.1)
    with pytest.raises(D.DhipError) as e:
        ctx.check(ctx.L.dhip_local_p2(ctx.h, None, D.RHS_NONE, None))
    assert e.value.code == D.E_STATE
    with pytest.raises(D.DhipError) as e:
        ctx.assemble_p1_band(-1.0)
    assert e.value.code == D.E_ARG


def test_band_empty(ctx):
    """no tet within the width of a level set that does not meet the mesh: nothing listed, no rows"""
    m = KuhnMesh.box(0.0, 1.0, 6)
    phi = m.sample_p2(plane_x(3.0))     # phi = x - 3 <= -2 on the mesh
    print("[gpu band] empty", flush=True)
    ctx.mesh(m)
    ctx.levelset(phi)
    nb, nd, nnz = ctx.assemble_p1_band(0.5)
    assert (nb, nd, nnz) == (0, 0, 0)
    om = O.make_mesh(m.orig, m.e, m.n)
    assert O.classify_band(om, phi, 0.5)[0].size == 0
    nb, nd, nnz = ctx.assemble_p1_band(2.5)      # the tets with a node at x >= 0.5
    assert nb == O.classify_band(om, phi, 2.5)[0].size and 0 < nb < 6 * 6 ** 3
    nb, nd, nnz = ctx.assemble_p1_band(3.5)      # |phi| <= 3: the whole mesh
    assert nb == 6 * 6 ** 3 and nd == 7 ** 3
