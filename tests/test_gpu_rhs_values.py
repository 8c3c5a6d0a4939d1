"""GPU tests of right-hand sides from any function (DHIP_RHS_VALUES: dhip_quad_points lists the patches'
composite Quad5 points, the caller evaluates f there, dhip_rhs_values hands the values back) through the
C ABI.  Bars: f = 3xy - x^3 given as values gives the load vector of the device kind bit for bit (P1 and
P2), a constant the kind DHIP_RHS_CONST's bit for bit, so the points are the ones the kind evaluates f at
and the values are summed in the same order; the points lie on the discrete interface (|x|^2 - 1 = O(h^2)
on the sphere); the load vector is linear in f; stale or mismatched values are refused.  The C++ mirror's
LocalVectorP1CL / LocalVectorP2CL with a host function give the kind's vector bit for bit."""
import os
import re
import subprocess

import numpy as np
import pytest

import cases
from drops_amd import dhip as D

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "drops_amd", "lib", "tracefem_p1")


def f_kind(xyz):
    """f = 3xy - x^3 with the device's evaluation order: (3 x) y - (x x) x"""
    x, y = xyz[:, 0], xyz[:, 1]
    return (3.0 * x) * y - (x * x) * x


def setup(ctx, n):
    m, phi = cases.c1_sphere(n)
    ctx.mesh(m)
    ctx.levelset(phi)
    return m, phi


def rhs_of(ctx, space, kind, param=None):
    if space == 2:
        ctx.assemble_p2(rhs_kind=kind, rhs_param=param)
    else:
        ctx.assemble_p1(rhs_kind=kind, rhs_param=param)
    return ctx.csr()["rhs"].copy()


@pytest.mark.parametrize("space", [1, 2])
def test_values_equal_device_kind(ctx, space):
    print(f"[gpu rhs values] space={space}", flush=True)
    m, phi = setup(ctx, 24)
    ref = rhs_of(ctx, space, D.RHS_3XY_MINUS_X3)
    cref = rhs_of(ctx, space, D.RHS_CONST, [2.5])
    ctx.classify()
    xyz = ctx.quad_points()
    assert xyz.shape[0] > 0 and xyz.shape[0] % 7 == 0
    # the points lie on the piecewise planar interface of the P2 level set: |x|^2 - 1 = O(h^2)
    assert np.abs((xyz ** 2).sum(axis=1) - 1.0).max() <= 0.5 * m.h ** 2
    ctx.rhs_values(f_kind(xyz))
    got = rhs_of(ctx, space, D.RHS_VALUES)
    assert np.array_equal(got, ref)
    ctx.rhs_values(np.full(xyz.shape[0], 2.5))
    assert np.array_equal(rhs_of(ctx, space, D.RHS_VALUES), cref)
    # linear in f
    g1, g2 = np.sin(xyz[:, 0]) * xyz[:, 2], np.cos(3.0 * xyz[:, 1])
    ctx.rhs_values(g1)
    b1 = rhs_of(ctx, space, D.RHS_VALUES)
    ctx.rhs_values(g2)
    b2 = rhs_of(ctx, space, D.RHS_VALUES)
    ctx.rhs_values(g1 + g2)
    b12 = rhs_of(ctx, space, D.RHS_VALUES)
    assert np.abs(b12 - (b1 + b2)).max() <= 1e-14 * np.abs(b1).max() + 1e-14 * np.abs(b2).max()


def test_values_refused_when_stale(ctx):
    print("[gpu rhs values] state checks", flush=True)
    m, phi = setup(ctx, 16)
    ctx.classify()
    xyz = ctx.quad_points()
    with pytest.raises(D.DhipError) as ei:
        ctx.rhs_values(np.zeros(xyz.shape[0] - 1))
    assert ei.value.code == D.E_ARG
    ctx.rhs_values(f_kind(xyz))
    ctx.levelset(phi)          # a new level set: the values no longer belong to the cut tets
    with pytest.raises(D.DhipError) as ei:
        ctx.assemble_p1(rhs_kind=D.RHS_VALUES)
    assert ei.value.code == D.E_STATE


@pytest.mark.parametrize("space", ["p1", "p2"])
def test_user_function_through_host_mirror(space):
    print(f"[gpu rhs values] host mirror {space}", flush=True)
    r = subprocess.run([EXE, "userrhs", "20", space], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert float(re.search(r"user rhs vs device rhs: (\S+)", r.stdout).group(1)) == 0.0
    sums = re.findall(r"sum \|b\| = (\S+)", r.stdout)
    assert len(sums) == 2 and sums[0] == sums[1] and float(sums[0]) > 0
