"""The C++ host mirror (drops_amd/host/dropship.hpp) driven like a DROPS application:
the example program assembles through accumulate() and solves with
PCGSolverCL<JACPcCL>; its matrices must equal the oracle's bit for bit and its
solution must meet the tolerance."""
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import cases
import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "drops_amd", "lib", "tracefem_p1")


def run(*args):
    print(f"[host mirror] {' '.join(args)}", flush=True)
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_known_answer_through_host_mirror():
    out = run("sphere20")
    assert re.search(r"\(sphere\) NumUnknowns: 1156\b", out)


def test_c1_through_host_mirror(tmp_path):
    n = 32
    out = run("c1", str(n), str(tmp_path))
    it = int(re.search(r"PCG iterations: (\d+)", out).group(1))
    true_res = float(re.search(r"true relative residual: (\S+)", out).group(1))
    assert true_res < 2e-10
    ld = lambda name, dt: np.fromfile(tmp_path / f"{name}.bin", dtype=dt)
    rp, col = ld("rowptr", np.int64), ld("colind", np.int64)
    mesh, phi = cases.c1_sphere(n)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    dof, nd = O.number_p1if(om, phi)
    ref = O.assemble_p1(om, phi, dof, nd, rhs_kind=O.RHS_3XY_MINUS_X3)
    assert np.array_equal(rp, ref["rowptr"]) and np.array_equal(col, ref["colind"].astype(np.int64))
    for k in ("mass", "lb", "stab", "rhs"):
        assert np.array_equal(ld(k, np.float64), ref[k]), k
    vod = ld("vertex_of_dof", np.int64)
    assert np.array_equal(dof[vod], np.arange(nd))
    # LinComb(1, M, 1, A, h, C) = 1*(M + A) + h*C, as the reference's variadic LinComb
    comb = (1.0 * (ref["mass"] + ref["lb"])) + mesh.h * ref["stab"]
    assert np.array_equal(ld("comb", np.float64), comb)
    xs, its, _, _ = O.pcg_jacobi(ref["rowptr"], ref["colind"], comb, ref["rhs"], tol=1e-13)
    x = ld("x", np.float64)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
    _, itr, _, _ = O.pcg_jacobi(ref["rowptr"], ref["colind"], comb, ref["rhs"], tol=1e-10)
    assert abs(it - itr) <= max(2, 0.02 * itr)
    # exportMTX, out() and WriteFEToFile of the mirror: the iostream restatement's bytes
    assert "formats round trip: ok" in out
    for name, fmt in (("L.mtx", "mtx"), ("L.txt", "plain")):
        O.fmt_matrix(str(tmp_path / ("ref_" + name)), ref["rowptr"], ref["colind"], comb, fmt=fmt)
        assert (tmp_path / name).read_bytes() == (tmp_path / ("ref_" + name)).read_bytes(), name
    for name, binary in (("rhs.dat", True), ("rhs.txt", False)):
        O.fmt_vector(str(tmp_path / ("ref_" + name)), ref["rhs"], binary=binary)
        assert (tmp_path / name).read_bytes() == (tmp_path / ("ref_" + name)).read_bytes(), name


def test_p2_through_host_mirror(tmp_path):
    """IdxDescCL(P2IF_FE) + the P2 accumulators through accumulate(): numbering, pattern and every value
    equal to the oracle's P2 assembly; PCG to 1e-10."""
    n = 16
    out = run("p2", str(n), str(tmp_path))
    ld = lambda name, dt: np.fromfile(tmp_path / f"{name}.bin", dtype=dt)
    mesh, phi = cases.c1_sphere(n)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    dof, nd = O.number_p2if(om, phi)
    ref = O.assemble_p2(om, phi, dof, nd, rhs_kind=O.RHS_3XY_MINUS_X3)
    assert np.array_equal(ld("rowptr", np.int64), ref["rowptr"])
    assert np.array_equal(ld("colind", np.int64), ref["colind"].astype(np.int64))
    for k in ("mass", "lb", "stab", "rhs"):
        assert np.array_equal(ld(k, np.float64), ref[k]), k
    nod = ld("node_of_dof", np.int64)
    assert np.array_equal(dof[nod], np.arange(nd))
    assert float(re.search(r"PCG residual: (\S+)", out).group(1)) <= 1e-10


def test_band_through_host_mirror(tmp_path):
    """IdxDescCL::CreateNumbering(..., width) + NarrowBandCommonDataP1CL + NarrowBandMatrixAccuP1CL
    through accumulate() (SurfactantNarrowBandStblP1CL::Update's tuple): numbering, stabilisation,
    surface forms, load and the LinComb bit-exact against the oracle's narrow band."""
    n, width = 16, 0.3
    out = run("band", str(n), str(width), str(tmp_path))
    ld = lambda name, dt: np.fromfile(tmp_path / f"{name}.bin", dtype=dt)
    mesh, phi = cases.c1_sphere(n)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    dof, nd = O.number_p1band(om, phi, width)
    assert re.search(rf"NumUnknowns: {nd}\b", out)
    assert np.array_equal(dof[ld("vertex_of_dof", np.int64)], np.arange(nd))
    ref = O.assemble_p1band(om, phi, dof, nd, width, rhs_kind=O.RHS_3XY_MINUS_X3)
    rps, cs, vs = ref["stab"]
    assert np.array_equal(ld("rowptr", np.int64), rps) and np.array_equal(ld("colind", np.int64), cs.astype(np.int64))
    assert np.array_equal(ld("stab", np.float64), vs)
    assert np.array_equal(ld("rhs", np.float64), ref["rhs"])
    # surface forms: the oracle's own entries, 0.0 elsewhere on the band pattern
    rpm, cm, vm, vl = ref["surf"]
    m_on, l_on, hit = np.zeros(cs.size), np.zeros(cs.size), np.zeros(cs.size, bool)
    for r in range(nd):
        lo, hi = rps[r], rps[r + 1]
        for q in range(rpm[r], rpm[r + 1]):
            k = lo + np.searchsorted(cs[lo:hi], cm[q])
            m_on[k], l_on[k], hit[k] = vm[q], vl[q], True
    assert np.array_equal(ld("mass", np.float64), m_on) and np.array_equal(ld("lb", np.float64), l_on)
    comb = np.where(hit, 1.0 * (m_on + l_on) + mesh.h * vs, mesh.h * vs)
    assert np.array_equal(ld("comb", np.float64), comb)


@pytest.mark.parametrize("space", ["p1", "p2"])
@pytest.mark.parametrize("threads", [1, 4])
def test_user_accumulator_next_to_device_accumulators(space, threads):
    """A user-written TetraAccumulatorCL (num/accumulator.h:42-66) in one tuple with the interface
    common data and two device-assembled matrix accumulators: accumulate() runs the device pass and
    visits every tet of the triangulation on the host exactly once -- in mesh order with one thread,
    over the colour classes with cloned accumulators with four -- with begin/finalize called once.
    Through the common data the user accumulator sees exactly the cut tets with a patch, and their
    patch areas sum to 1^T M 1 of the device mass matrix (the basis is a partition of unity)."""
    n = 16
    print(f"[host mirror] accus {n} {space} threads={threads}", flush=True)
    r = subprocess.run([EXE, "accus", str(n), space], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
    assert r.returncode == 0, r.stderr
    v = dict(re.findall(r"^([a-z ]+): (\S+)$", r.stdout, flags=re.M))
    tets = 6 * n ** 3
    assert int(v["tets"]) == tets and int(v["visited"]) == tets and int(v["visited once"]) == tets
    assert int(v["begin"]) == 1 and int(v["finalize"]) == 1
    assert int(v["threads"]) == threads
    if threads == 1:
        assert int(v["mesh order"]) == 1 and int(v["clones"]) == 0
    else:
        assert int(v["clones"]) == threads - 1 and int(v["colors"]) >= 27
    mesh, phi = cases.c1_sphere(n)
    om = O.make_mesh(mesh.orig, mesh.e, mesh.n)
    cut = O.classify(om, phi)
    # no exact zeros (tie-break shift): every cut tet has a sign change on a lattice edge, so a patch
    assert int(v["cut"]) == cut.size
    if space == "p1":
        dof, nd = O.number_p1if(om, phi)
        ref = O.assemble_p1(om, phi, dof, nd)
    else:
        dof, nd = O.number_p2if(om, phi)
        ref = O.assemble_p2(om, phi, dof, nd)
    assert int(v["unknowns"]) == nd and int(v["nnz"]) == ref["colind"].size
    area, msum = float(v["area"]), float(v["mass sum"])
    assert abs(msum - ref["mass"].sum()) <= 1e-13 * abs(msum)
    assert abs(area - msum) <= 1e-12 * abs(msum)


def test_ch_class_through_host_mirror(tmp_path):
    """CahnHilliardcGP1CL of the mirror (CreateNumbering, SetInitialValue, DoStep x 5 with dt 1e-4)
    against five oracle steps from the same interpolated initial state: numbering equal, trajectory
    within 1e-6 (GMRES tolerances 1e-13 absolute on the device -- the class's absolute residual --,
    1e-12 relative in the oracle)."""
    n, steps = 16, 5
    out = run("ch", str(n), str(steps), str(tmp_path))
    assert len(re.findall(r"step \d+: iter \d+", out)) == steps
    ld = lambda name, dt: np.fromfile(tmp_path / f"{name}.bin", dtype=dt)
    m, phi = cases.torus_case(n)
    om = O.make_mesh(m.orig, m.e, m.n)
    dof, nd = O.number_p1if(om, phi)
    vod = ld("vertex_of_dof", np.int64)
    assert vod.size == nd and np.array_equal(dof[vod], np.arange(nd))
    r = O.assemble_p1(om, phi, dof, nd, coef=(1.0, 1.0, 1.0))
    c, mu = ld("c0", np.float64), np.zeros(nd)
    for _ in range(steps):
        o = O.ch_step(om, phi, dof, nd, r["rowptr"], r["colind"], r["mass"], r["lb"], r["stab"], c, mu, dt=1e-4,
                      eps=0.05, rho=m.h, tol=1e-12, restart=50)
        c, mu = o["c"], o["mu"]
    assert np.linalg.norm(ld("c", np.float64) - c) <= 1e-6 * np.linalg.norm(c)
    assert np.linalg.norm(ld("mu", np.float64) - mu) <= 1e-6 * np.linalg.norm(mu)


def test_block_solver_through_host_mirror(tmp_path):
    """BlockMatrixSolverCL<GMResSolverCL<DiagBlockJacobiPcCL>>::Solve(A, B, C, D, ...) against a direct
    solve of scipy's bmat([[A, B], [C, D]]); transp_mul on the device equals transpose() + SpMV bit for
    bit and scipy's C^T x to rounding."""
    n = 16
    out = run("block", str(n), str(tmp_path))
    assert float(re.search(r"transp_mul vs transpose: (\S+)", out).group(1)) == 0.0
    ld = lambda name, dt: np.fromfile(tmp_path / f"{name}.bin", dtype=dt)
    mats = {}
    for k in "ABCD":
        rp, ci, v = ld(k + "_rowptr", np.int64), ld(k + "_colind", np.int64), ld(k + "_val", np.float64)
        mats[k] = sp.csr_matrix((v, ci, rp), shape=(rp.size - 1, rp.size - 1))
    K = sp.bmat([[mats["A"], mats["B"]], [mats["C"], mats["D"]]]).tocsc()
    b = np.concatenate([ld("rhs3", np.float64), ld("rhs4", np.float64)])
    xd = spla.spsolve(K, b)
    x = np.concatenate([ld("mu", np.float64), ld("c", np.float64)])
    assert np.linalg.norm(b - K @ x) <= 2e-10 * np.linalg.norm(b)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)
    yt = ld("yt", np.float64)
    yref = mats["C"].T @ ld("c", np.float64)
    bound = (abs(mats["C"]).T @ np.abs(ld("c", np.float64))).max()
    assert np.abs(yt - yref).max() <= 1e-14 * bound


@pytest.mark.parametrize("dist", [0.0, 0.1])
def test_distribute_dofs_through_host_mirror(tmp_path, dist):
    """IdxDescCL::DistributeDOFs (misc/problem.cpp:283-336) on the torus: the marked tets are those with
    distance() <= dist and the DOFs are every vertex of a marked tet, numbered on first encounter in tet
    order -- checked against a numpy restatement from the definition."""
    n = 16
    out = run("distribute", str(n), repr(dist), str(tmp_path))
    ntet = int(re.search(r"marked tets: (\d+)", out).group(1))
    vod = np.fromfile(tmp_path / "vertex_of_dof.bin", dtype=np.int64)
    m, phi = cases.torus_case(n)
    n1, n2, _ = m.n
    h1, h2 = 2 * n1 + 1, 2 * n2 + 1
    nodes = cases.tet_p2_nodes(m, np.arange(6 * n1 * n2 * m.n[2]))   # half-grid nodes of every tet
    vals = np.asarray(phi).reshape(-1)[nodes]
    j = nodes[:, :4]
    tets = (((j // (h1 * h2)) // 2) * (n2 + 1) + ((j // h1) % h2) // 2) * (n1 + 1) + (j % h1) // 2
    same = np.all(vals > 0, axis=1) | np.all(vals < 0, axis=1)
    dist_t = np.where(same, np.abs(vals).min(axis=1), 0.0)
    marked = np.flatnonzero(dist_t <= dist)
    assert ntet == marked.size
    seen, order = set(), []
    for t in marked:
        for v in tets[t]:
            if v not in seen:
                seen.add(v)
                order.append(v)
    assert np.array_equal(vod, np.array(order, np.int64))
