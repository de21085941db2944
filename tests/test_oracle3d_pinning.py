"""Pins the D3Q27 oracle (oracle/lbm_oracle3d.hpp) against what its code does not restate:
lattice isotropy, moment identities, closed forms, numpy.roll, conservation laws and two
analytic flows. CPU only."""
import ctypes
import itertools
import math

import numpy as np
import pytest

from lbm_amd import _abi
import oracle as O

W, CX, CY, CZ, OPP = O.tables3d()
C = np.stack([CX, CY, CZ], 1).astype(np.float64)


def cfg3(nx, ny, nz, omega=1.0, periodic=(1, 1, 1), force=(0.0, 0.0, 0.0), storage=_abi.LBM_F32, quirks=0):
    c = _abi.Lbm3dConfig()
    c.nx, c.ny, c.nz = nx, ny, nz
    c.collision = _abi.LBM_BGK
    c.omega = omega
    c.rho0 = 1.0
    c.storage = storage
    c.periodic_x, c.periodic_y, c.periodic_z = periodic
    for i in range(3):
        c.force[i] = force[i]
    c.quirks = quirks
    return c


def test_velocity_set_is_the_full_27_stencil_in_shell_order():
    assert len({tuple(v) for v in C.astype(int)}) == 27
    assert set(map(tuple, C.astype(int))) == set(itertools.product((-1, 0, 1), repeat=3))
    shells = (C ** 2).sum(1)
    assert shells[0] == 0 and all(shells[1:7] == 1) and all(shells[7:19] == 2) and all(shells[19:] == 3)
    assert np.array_equal(C[OPP], -C)
    assert all(OPP[OPP] == np.arange(27))


def test_weights_give_an_isotropic_fourth_order_lattice():
    assert abs(W.sum() - 1) < 1e-15
    d = np.eye(3)
    assert np.allclose(np.einsum("q,qa->a", W, C), 0, atol=1e-15)
    assert np.allclose(np.einsum("q,qa,qb->ab", W, C, C), d / 3, atol=1e-15)
    assert np.allclose(np.einsum("q,qa,qb,qc->abc", W, C, C, C), 0, atol=1e-15)
    iso4 = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.allclose(np.einsum("q,qa,qb,qc,qd->abcd", W, C, C, C, C), iso4, atol=1e-15)


@pytest.mark.parametrize("seed", range(4))
def test_equilibrium_moments_up_to_second_order(seed):
    rng = np.random.default_rng(seed)
    rho = 1 + 0.2 * rng.standard_normal()
    u = 0.1 * rng.standard_normal(3)
    feq = O.equilibrium3d(rho, *u)
    assert abs(feq.sum() - rho) < 1e-14
    assert np.allclose(feq @ C, rho * u, atol=1e-14)
    assert np.allclose(np.einsum("q,qa,qb->ab", feq, C, C), rho * (np.outer(u, u) + np.eye(3) / 3), atol=1e-14)
    # FP32 evaluation agrees with FP64 to float rounding
    assert np.allclose(O.equilibrium3d(rho, *u, prec=32), feq, rtol=0, atol=1e-7)


@pytest.mark.parametrize("seed", range(3))
def test_literal_equilibrium_carries_17_27_of_the_momentum(seed):
    """The reference's De Rosis form (equilibrium.cu:70-96) has linear coefficient 1 on the edge
    and corner directions: its first moment is (2/27*3*2 + 1/54*8 + 1/216*8) rho u = 17/27 rho u,
    and it differs from the second-order equilibrium by -2 w rho (c.u) on those directions only."""
    rng = np.random.default_rng(10 + seed)
    rho = 1 + 0.1 * rng.standard_normal()
    u = 0.05 * rng.standard_normal(3)
    lit = O.equilibrium3d(rho, *u, literal=True)
    ref = O.equilibrium3d(rho, *u)
    assert abs(lit.sum() - rho) < 1e-14
    assert np.allclose(lit @ C, 17 / 27 * rho * u, atol=1e-14)
    assert np.allclose(np.einsum("q,qa,qb->ab", lit, C, C), np.einsum("q,qa,qb->ab", ref, C, C), atol=1e-14)
    expect = np.where(np.arange(27) >= 7, -2 * W * rho * (C @ u), 0.0)
    assert np.allclose(lit - ref, expect, atol=1e-15)


@pytest.mark.parametrize("omega", [0.7, 1.0, 1.9])
def test_guo_source_moments(omega):
    rng = np.random.default_rng(3)
    u = 0.05 * rng.standard_normal(3)
    F = 1e-3 * rng.standard_normal(3)
    S = O.guo_source3d(u, F, omega)
    k = 1 - omega / 2
    assert abs(S.sum()) < 1e-17
    assert np.allclose(S @ C, k * F, atol=1e-17)
    assert np.allclose(np.einsum("q,qa,qb->ab", S, C, C), k * (np.outer(u, F) + np.outer(F, u)), atol=1e-17)


def test_periodic_streaming_is_a_roll_of_each_plane():
    nx, ny, nz = 7, 5, 6
    o = O.Oracle3D(cfg3(nx, ny, nz), 64)
    f = np.random.default_rng(0).random((27, nz, ny, nx))
    o.set_pops(f)
    o.stream_only()
    g = o.get_pops()
    for q in range(27):
        assert np.array_equal(g[q], np.roll(f[q], (CZ[q], CY[q], CX[q]), axis=(0, 1, 2)))


def test_non_periodic_streaming_leaves_unreached_slots_stale():
    nx, ny, nz = 5, 4, 6
    o = O.Oracle3D(cfg3(nx, ny, nz, periodic=(0, 1, 0)), 64)
    f = np.random.default_rng(1).random((27, nz, ny, nx))
    o.set_pops(f)  # both buffers hold f
    o.stream_only()
    g = o.get_pops()
    for q in range(27):
        rolled = np.roll(f[q], (CZ[q], CY[q], CX[q]), axis=(0, 1, 2))
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        xs, zs = x - CX[q], z - CZ[q]
        reached = (xs >= 0) & (xs < nx) & (zs >= 0) & (zs < nz)
        assert np.array_equal(g[q][reached], rolled[reached])
        assert np.array_equal(g[q][~reached], f[q][~reached])


def test_unforced_periodic_bgk_conserves_mass_and_momentum():
    nx, ny, nz = 8, 6, 5
    o = O.Oracle3D(cfg3(nx, ny, nz, omega=1.6), 64)
    rng = np.random.default_rng(2)
    rho = 1 + 0.05 * rng.standard_normal((nz, ny, nx))
    u = 0.05 * rng.standard_normal((nz, ny, nx, 3))
    o.init_from_macros(rho, u)
    f0 = o.get_pops()
    o.step(20)
    f1 = o.get_pops()
    assert abs(f1.sum() - f0.sum()) < 1e-11
    assert np.allclose(np.einsum("qzyx,qa->a", f1, C), np.einsum("qzyx,qa->a", f0, C), atol=1e-12)


def test_uniform_force_adds_n_times_f_of_momentum_per_step():
    nx, ny, nz = 6, 5, 4
    F = np.array([2e-5, -1e-5, 3e-5])
    o = O.Oracle3D(cfg3(nx, ny, nz, omega=1.2, force=F), 64)
    o.init_from_macros(np.ones((nz, ny, nx)), np.zeros((nz, ny, nx, 3)))
    o.step(10)
    p = np.einsum("qzyx,qa->a", o.get_pops(), C)
    F32 = F.astype(np.float32).astype(np.float64)  # lbm3d_config.force is float
    assert np.allclose(p, 10 * F32 * nx * ny * nz, rtol=1e-10, atol=0)


def test_rest_state_is_a_fixed_point_of_a_closed_bounce_back_box():
    nx, ny, nz = 5, 4, 6
    o = O.Oracle3D(cfg3(nx, ny, nz, omega=1.3, periodic=(0, 0, 0)), 64)
    fl = np.zeros((nz, ny, nx), np.uint8)
    fl[0], fl[-1], fl[:, 0], fl[:, -1], fl[:, :, 0], fl[:, :, -1] = (_abi.BOUNCE_BACK,) * 6
    o.set_flags(fl)
    o.init_from_macros(np.ones((nz, ny, nx)), np.zeros((nz, ny, nx, 3)))
    o.step(7)
    f = o.get_pops()
    assert np.abs(f - W[:, None, None, None]).max() < 1e-15


def test_corner_edge_bounce_back_copies_the_nonnegative_directions_on_edges_only():
    """EdgeCornerBounceBack (edgeCornerBounceBack.cuh:10-49) after streaming: on a node with two or
    more domain faces, the 7 directions with no negative and some positive component are copied
    into their opposites; on faces and inside, the flag changes nothing. At the (NX-1, NY-1, NZ-1)
    corner those 7 opposite slots also receive what BounceBack<3> writes there; at the other
    corners they do not (the functor ignores the normals it computes)."""
    nx, ny, nz = 5, 4, 6
    rng = np.random.default_rng(3)
    f0 = W[:, None, None, None] * (1 + 0.1 * rng.random((27, nz, ny, nx)))

    def run(fl):
        o = O.Oracle3D(cfg3(nx, ny, nz, periodic=(0, 0, 0)), 64)
        o.set_pops(f0)
        if fl is not None:
            o.set_flags(fl)
        o.stream_boundaries()
        return o.get_pops()

    ce = np.full((nz, ny, nx), _abi.CORNER_EDGE_BOUNCE_BACK, np.uint8)
    fa, fb = run(ce), run(None)
    P = [i for i in range(27) if (C[i] >= 0).all() and C[i].any()]
    assert len(P) == 7 and not set(P) & set(OPP[P])
    expected = fb.copy()
    for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
        if (x in (0, nx - 1)) + (y in (0, ny - 1)) + (z in (0, nz - 1)) >= 2:
            for i in P:
                expected[OPP[i], z, y, x] = fb[i, z, y, x]
    assert np.array_equal(fa, expected)
    assert not np.array_equal(fa, fb)
    bb = np.zeros((nz, ny, nx), np.uint8)
    bb[-1, -1, -1] = bb[0, 0, 0] = _abi.BOUNCE_BACK
    fc = run(bb)
    assert np.array_equal(fc[OPP[P], -1, -1, -1], fa[OPP[P], -1, -1, -1])
    assert not np.array_equal(fc[OPP[P], 0, 0, 0], fa[OPP[P], 0, 0, 0])


def test_zg_outflow_copies_the_entering_populations_from_the_interior_neighbour():
    """ZG_OutflowBoundary<3> (zeroGradientOutflow.cuh:8-59) after streaming: on a face node, the
    populations with c . n > 0 (n the inward normal) take the interior neighbour's post-stream
    values and the others keep the node's own; the first face in the order x, y, z decides n at
    edges and corners; inside the domain the flag changes nothing."""
    nx, ny, nz = 5, 4, 6
    rng = np.random.default_rng(4)
    f0 = W[:, None, None, None] * (1 + 0.1 * rng.random((27, nz, ny, nx)))

    def run(fl):
        o = O.Oracle3D(cfg3(nx, ny, nz, periodic=(0, 0, 0)), 64)
        o.set_pops(f0)
        if fl is not None:
            o.set_flags(fl)
        o.stream_boundaries()
        return o.get_pops()

    zg = np.full((nz, ny, nx), _abi.ZG_OUTFLOW, np.uint8)
    fa, fb = run(zg), run(None)
    expected = fb.copy()
    for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
        for axis, (p, n) in enumerate(((x, nx), (y, ny), (z, nz))):
            if p in (0, n - 1):
                nrm = np.zeros(3, int)
                nrm[axis] = 1 if p == 0 else -1
                zi, yi, xi = np.array([z, y, x]) + nrm[::-1]
                enter = C @ nrm > 0
                assert enter.sum() == 9
                expected[enter, z, y, x] = fb[enter, zi, yi, xi]
                break
    assert np.array_equal(fa, expected)
    assert not np.array_equal(fa, fb)


def test_regularized_boundary_imposes_zou_he_density_and_the_wall_velocity():
    """RegularizedBoundary (regularizedBoundary3D.cuh:10-137) after streaming: on a face node the
    populations are rebuilt so that their density is the Zou-He density of the known post-stream
    populations, rho = (sum_{c.n = 0} f + 2 sum_{c.n < 0} f) / (1 - u.n) with n the inward normal
    of the LAST matching face in the order x, y, z, and their momentum is rho * u_imposed (the
    second-order equilibrium carries rho u and the Q : Pi projection carries none); interior nodes
    are unchanged."""
    nx, ny, nz = 5, 4, 6
    u_in = float(np.float32(0.04))  # lbm3d_config.u_inlet is float
    rng = np.random.default_rng(5)
    f0 = W[:, None, None, None] * (1 + 0.1 * rng.random((27, nz, ny, nx)))

    def run(fl):
        c = cfg3(nx, ny, nz, periodic=(0, 0, 0))
        c.u_inlet = u_in
        o = O.Oracle3D(c, 64)
        o.set_pops(f0)
        if fl is not None:
            o.set_flags(fl)
        o.stream_boundaries()
        return o.get_pops()

    fl = np.full((nz, ny, nx), _abi.REGULARIZED_BOUNCE_BACK, np.uint8)
    fl[:, :, 0] = _abi.REGULARIZED_INLET_LEFT
    fa, fb = run(fl), run(None)
    checked = 0
    for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
        nrm = None
        for axis, (p, n) in enumerate(((x, nx), (y, ny), (z, nz))):
            if p in (0, n - 1):
                nrm = np.zeros(3)
                nrm[axis] = 1 if p == 0 else -1
        if nrm is None:
            assert np.array_equal(fa[:, z, y, x], fb[:, z, y, x])
            continue
        u = np.array([u_in if x == 0 else 0.0, 0.0, 0.0])
        cn = C @ nrm
        g = fb[:, z, y, x]
        rho = (g[cn == 0].sum() + 2 * g[cn < 0].sum()) / (1 - u @ nrm)
        h = fa[:, z, y, x]
        assert abs(h.sum() - rho) < 1e-13
        assert np.allclose(C.T @ h, rho * u, rtol=0, atol=1e-13)
        checked += 1
    assert checked == nx * ny * nz - (nx - 2) * (ny - 2) * (nz - 2)


def test_force_driven_channel_has_curvature_minus_f_over_nu():
    """Walls at y = 0 and y = NY-1 (BOUNCE_BACK), periodic x and z, body force along x: the
    steady profile satisfies nu u'' = -F in the bulk and is symmetric about the centre."""
    nx, ny, nz = 3, 18, 3
    omega = 1.0
    nu = (1 / omega - 0.5) / 3
    Fx = 1e-6
    o = O.Oracle3D(cfg3(nx, ny, nz, omega=omega, periodic=(1, 0, 1), force=(Fx, 0, 0)), 64)
    fl = np.zeros((nz, ny, nx), np.uint8)
    fl[:, 0], fl[:, -1] = _abi.BOUNCE_BACK, _abi.BOUNCE_BACK
    o.set_flags(fl)
    o.init_from_macros(np.ones((nz, ny, nx)), np.zeros((nz, ny, nx, 3)))
    o.step(4000)
    rho, u = o.get_macros()
    ux = u[..., 0].mean(axis=(0, 2))
    assert np.abs(u[..., 1:]).max() < 1e-12 * 1e6 * Fx
    assert np.allclose(ux, ux[::-1], rtol=1e-9, atol=0)
    d2 = ux[3:-1] - 2 * ux[2:-2] + ux[1:-3]
    assert np.allclose(d2, -Fx / nu, rtol=2e-3)
    assert ux.max() > 30 * abs(ux[0])


def test_small_taylor_green_vortex_decays_at_the_viscous_rate():
    """u = U (sin kx cos ky cos kz, -cos kx sin ky cos kz, 0) is an eigenfield of the Laplacian
    with eigenvalue -3k^2: at small U its kinetic energy decays as exp(-6 nu k^2 t)."""
    n = 24
    omega = 1.7
    nu = (1 / omega - 0.5) / 3
    k = 2 * np.pi / n
    o = O.Oracle3D(cfg3(n, n, n, omega=omega), 64)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    U = 1e-4
    u = np.stack([U * np.sin(k * x) * np.cos(k * y) * np.cos(k * z),
                  -U * np.cos(k * x) * np.sin(k * y) * np.cos(k * z),
                  np.zeros_like(x, dtype=float)], -1)
    o.init_from_macros(np.ones((n, n, n)), u)

    def energy():
        rho, uu = o.get_macros()
        return 0.5 * (rho * (uu ** 2).sum(-1)).sum()

    o.step(20)
    e0 = energy()
    o.step(60)
    e1 = energy()
    rate = -np.log(e1 / e0) / 60
    assert abs(rate / (6 * nu * k * k) - 1) < 0.02


def test_literal_equilibrium_breaks_momentum_conservation():
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(4)
    rho = np.ones((nz, ny, nx))
    u = 0.02 * rng.standard_normal((nz, ny, nx, 3))
    o = O.Oracle3D(cfg3(nx, ny, nz, omega=1.0, quirks=_abi.LBM_QUIRK_EQ3D_LINEAR), 64)
    o.init_from_macros(rho, u)
    p0 = np.einsum("qzyx,qa->a", o.get_pops(), C)
    assert np.allclose(p0, 17 / 27 * (rho[..., None] * u).sum((0, 1, 2)), atol=1e-13)
    o.step(1)
    p1 = np.einsum("qzyx,qa->a", o.get_pops(), C)
    # omega = 1: f* = f_eq(rho, u(f)) whose momentum is 17/27 of that of f
    assert np.allclose(p1, 17 / 27 * p0, atol=1e-13)


def test_reduced_storage_rounds_the_shifted_deviation():
    """16-bit storage keeps f - w rho0 to half (2^-11) or bfloat16 (2^-8) relative rounding: the
    error bound scales with the deviation from w, not with f (plus half the FP16 subnormal
    quantum 2^-25 and the FP32 rounding of f and w)."""
    nx, ny, nz = 4, 3, 3
    for st, eps in [(_abi.LBM_F16_SHIFTED, 2.0 ** -11), (_abi.LBM_BF16_SHIFTED, 2.0 ** -8)]:
        o = O.Oracle3D(cfg3(nx, ny, nz, storage=st), 32)
        rng = np.random.default_rng(5)
        f = W[:, None, None, None] * (1 + 0.05 * rng.standard_normal((27, nz, ny, nx)))
        o.set_pops(f)
        g = o.get_pops()
        dev = np.abs(f - W[:, None, None, None])
        assert np.all(np.abs(g - f) <= eps * dev + 2.0 ** -25 + 2.0 ** -22 * W[:, None, None, None])
        assert np.abs(g - f).max() > 0
        o.set_pops(g)  # idempotent
        assert np.array_equal(o.get_pops(), g)


# ---- MRT<3> ---------------------------------------------------------------------------------

def product_equilibrium(rho, u):
    """rho * prod_a phi(c_a, u_a), phi(0) = 2/3 - u^2, phi(+-1) = (1/3 + u^2 +- u)/2: the
    distribution whose 1-D moments are 1, u, 1/3 + u^2 along every axis."""
    def phi(c, ua):
        return np.where(c == 0, 2 / 3 - ua * ua, (1 / 3 + ua * ua + c * ua) / 2)
    return rho * phi(CX, u[0]) * phi(CY, u[1]) * phi(CZ, u[2])


def mrt_cfg_S(s4to8):
    S = [1.0] * 27
    S[4:9] = s4to8
    return S


def test_mrt3_basis_is_invertible_and_integer():
    M, Mi = O.mrt3_basis()
    assert np.allclose(M @ Mi, np.eye(27), atol=1e-13) and np.allclose(Mi @ M, np.eye(27), atol=1e-13)
    assert np.array_equal(M, np.round(M)) and np.all(M[0] == 1)
    assert np.array_equal(M[1:4], C.T)


@pytest.mark.parametrize("seed", range(3))
def test_mrt3_at_unit_rates_is_the_product_equilibrium(seed):
    rng = np.random.default_rng(20 + seed)
    g = W * (1 + 0.1 * rng.standard_normal(27))
    u = 0.08 * rng.standard_normal(3)
    out = O.collide_mrt3(g, u, (0, 0, 0), mrt_cfg_S([1.0] * 5))
    assert np.allclose(out, product_equilibrium(g.sum(), u), atol=1e-15)


@pytest.mark.parametrize("seed", range(3))
def test_mrt3_moments_after_collision(seed):
    rng = np.random.default_rng(30 + seed)
    g = W * (1 + 0.1 * rng.standard_normal(27))
    rho = g.sum()
    F = 1e-3 * rng.standard_normal(3)
    u = (g @ C + F / 2) / rho  # the corrected velocity the step passes in
    s = rng.uniform(0.3, 1.9, 5).astype(np.float32).astype(np.float64)  # rates cross as float
    out = O.collide_mrt3(g, u, F, mrt_cfg_S(s))
    x, y, z = C.T
    assert abs(out.sum() - rho) < 1e-15
    assert np.allclose(out @ C, g @ C + F, atol=1e-15)  # Guo: a step adds F
    pairs = [(x * y, u[0] * u[1], F[0] * u[1] + F[1] * u[0]), (x * z, u[0] * u[2], F[0] * u[2] + F[2] * u[0]),
             (y * z, u[1] * u[2], F[1] * u[2] + F[2] * u[1]),
             (x * x - y * y, u[0] ** 2 - u[1] ** 2, 2 * (F[0] * u[0] - F[1] * u[1])),
             (x * x - z * z, u[0] ** 2 - u[2] ** 2, 2 * (F[0] * u[0] - F[2] * u[2]))]
    for k, (poly, ueq, Fk) in enumerate(pairs):
        expect = (1 - s[k]) * (g @ poly) + s[k] * rho * ueq + (1 - s[k] / 2) * Fk
        assert abs(out @ poly - expect) < 1e-15
    # unforced: every moment above second order at its product-equilibrium value
    out0 = O.collide_mrt3(g, u, (0, 0, 0), mrt_cfg_S(s))
    prod = product_equilibrium(rho, u)
    for poly in (x * y * z, x * y * y, x * x * y * y, x * x * y * y * z * z, x * y * z * z, x * x + y * y + z * z):
        assert abs(out0 @ poly - prod @ poly) < 1e-15


def test_mrt3_literal_moment_19_differs_by_rho_cs2_ux2_duyz():
    """The reference's moment-19 equilibrium (2 ux^2 + 1) vs the moment of the equilibrium
    (3 ux^2 + 1): the literal collision at unit rates differs from the product equilibrium
    only in moment 19, by -rho/3 ux^2 (uy^2 - uz^2)."""
    rng = np.random.default_rng(40)
    g = W * (1 + 0.1 * rng.standard_normal(27))
    u = 0.08 * rng.standard_normal(3)
    M, _ = O.mrt3_basis()
    lit = O.collide_mrt3(g, u, (0, 0, 0), mrt_cfg_S([1.0] * 5), literal=True)
    d = M @ (lit - product_equilibrium(g.sum(), u))
    expect = np.zeros(27)
    expect[19] = -g.sum() / 3 * u[0] ** 2 * (u[1] ** 2 - u[2] ** 2)
    assert np.allclose(d, expect, atol=1e-15)


def test_mrt3_literal_quirk_adds_half_the_momentum():
    nx, ny, nz = 5, 4, 6
    F = np.array([2.0 ** -16, -(2.0 ** -17), 2.0 ** -15])
    for quirk, share in ((0, 1.0), (_abi.LBM_QUIRK_MRT3_LITERAL, 0.5)):
        c = cfg3(nx, ny, nz, force=F, quirks=quirk)
        c.collision = _abi.LBM_MRT
        for k in range(27):
            c.S[k] = 1.0
        for k in range(4, 9):
            c.S[k] = 1.3
        o = O.Oracle3D(c, 64)
        o.init_from_macros(np.ones((nz, ny, nx)), np.zeros((nz, ny, nx, 3)))
        o.step(6)
        p = np.einsum("qzyx,qa->a", o.get_pops(), C)
        assert np.allclose(p, share * 6 * F * nx * ny * nz, rtol=1e-11, atol=0)


def test_mrt3_taylor_green_decays_at_the_viscous_rate():
    n = 24
    s = 1.6
    nu = (1 / s - 0.5) / 3
    k = 2 * np.pi / n
    c = cfg3(n, n, n)
    c.collision = _abi.LBM_MRT
    for i in range(27):
        c.S[i] = 1.0
    for i in range(4, 9):
        c.S[i] = s
    o = O.Oracle3D(c, 64)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    U = 1e-4
    u = np.stack([U * np.sin(k * x) * np.cos(k * y) * np.cos(k * z),
                  -U * np.cos(k * x) * np.sin(k * y) * np.cos(k * z), np.zeros_like(x, dtype=float)], -1)
    o.init_from_macros(np.ones((n, n, n)), u)

    def energy():
        rho, uu = o.get_macros()
        return 0.5 * (rho * (uu ** 2).sum(-1)).sum()

    o.step(20)
    e0 = energy()
    o.step(60)
    e1 = energy()
    assert abs(-np.log(e1 / e0) / 60 / (6 * nu * k * k) - 1) < 0.02


# ---- CM<3, NoAdapter> ------------------------------------------------------------------------

def central_moment(f, u, a, b, c):
    return (f * (CX - u[0]) ** a * (CY - u[1]) ** b * (CZ - u[2]) ** c).sum()


@pytest.mark.parametrize("seed", range(3))
def test_cm3_at_unit_omega_is_the_product_equilibrium(seed):
    rng = np.random.default_rng(60 + seed)
    g = W * (1 + 0.1 * rng.standard_normal(27))
    u = 0.08 * rng.standard_normal(3)
    out = O.collide_cm3(g, u, (0, 0, 0), 1.0)
    assert np.allclose(out, product_equilibrium(g.sum(), u), atol=1e-14)


@pytest.mark.parametrize("seed", range(3))
def test_cm3_central_moments_after_collision(seed):
    """Conserved rho, momentum + F, the deviatoric central second moments relaxed by (1 - omega),
    the trace at rho (= 3 cs2 rho), third-order forcing moments cs2 F/2 and the sixth at rho cs2^3."""
    rng = np.random.default_rng(70 + seed)
    g = W * (1 + 0.1 * rng.standard_normal(27))
    rho = g.sum()
    F = 1e-3 * rng.standard_normal(3)
    u = (g @ C + F / 2) / rho
    om = float(np.float64(rng.uniform(0.3, 1.9)))
    out = O.collide_cm3(g, u, F, om)
    assert abs(out.sum() - rho) < 1e-14
    assert np.allclose(out @ C, g @ C + F, atol=1e-14)
    k = lambda f, a, b, c: central_moment(f, u, a, b, c)
    for idx in [(1, 1, 0), (1, 0, 1), (0, 1, 1)]:
        assert abs(k(out, *idx) - (1 - om) * k(g, *idx)) < 1e-14
    dev_in = k(g, 2, 0, 0) - k(g, 0, 2, 0)
    assert abs(k(out, 2, 0, 0) - k(out, 0, 2, 0) - (1 - om) * dev_in) < 1e-14
    assert abs(k(out, 2, 0, 0) + k(out, 0, 2, 0) + k(out, 0, 0, 2) - rho) < 1e-14
    assert np.allclose([k(out, 1, 2, 0), k(out, 2, 1, 0), k(out, 0, 2, 1)], [F[0] / 6, F[1] / 6, F[2] / 6], atol=1e-15)
    assert abs(k(out, 1, 1, 1)) < 1e-15 and abs(k(out, 2, 2, 2) - rho / 27) < 1e-14
    assert abs(k(out, 1, 2, 2) - F[0] / 18) < 1e-15


def test_cm3_taylor_green_decays_at_the_viscous_rate():
    n = 24
    om = 1.6
    nu = (1 / om - 0.5) / 3
    k = 2 * np.pi / n
    c = cfg3(n, n, n, omega=om)
    c.collision = _abi.LBM_CM
    o = O.Oracle3D(c, 64)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    U = 1e-4
    u = np.stack([U * np.sin(k * x) * np.cos(k * y) * np.cos(k * z),
                  -U * np.cos(k * x) * np.sin(k * y) * np.cos(k * z), np.zeros_like(x, dtype=float)], -1)
    o.init_from_macros(np.ones((n, n, n)), u)

    def energy():
        rho, uu = o.get_macros()
        return 0.5 * (rho * (uu ** 2).sum(-1)).sum()

    o.step(20)
    e0 = energy()
    o.step(60)
    e1 = energy()
    assert abs(-np.log(e1 / e0) / 60 / (6 * nu * k * k) - 1) < 0.02


# ---- Guo velocity inlet / pressure outlet ---------------------------------------------------

def guo_flags(nx, ny, nz, walls):
    fl = np.zeros((nz, ny, nx), np.uint8)
    if walls:
        fl[:, 0, :] = fl[:, -1, :] = _abi.BOUNCE_BACK
        fl[0], fl[-1] = _abi.BOUNCE_BACK, _abi.BOUNCE_BACK
    fl[:, :, 0] = _abi.GUO_VELOCITY_INLET
    fl[:, :, -1] = _abi.GUO_PRESSURE_OUTLET
    return fl


@pytest.mark.parametrize("quirk", [0, _abi.LBM_QUIRK_GUO3D_OWN_MOMENTS])
def test_uniform_flow_through_guo_inlet_and_outlet_is_a_fixed_point(quirk):
    nx, ny, nz = 7, 4, 5
    uin = 0.05
    c = cfg3(nx, ny, nz, omega=1.3, periodic=(0, 1, 1), quirks=quirk)
    c.u_inlet = uin
    o = O.Oracle3D(c, 64)
    o.set_flags(guo_flags(nx, ny, nz, False))
    u = np.zeros((nz, ny, nx, 3))
    u[..., 0] = np.float32(uin)
    o.init_from_macros(np.ones((nz, ny, nx)), u)
    f0 = o.get_pops()
    o.step(25)
    assert np.abs(o.get_pops() - f0).max() < 1e-14


def test_flow_from_rest_between_guo_inlet_and_outlet_converges_to_the_inlet_velocity():
    """Fluid at rest, inlet u = 0.04 at x = 0, pressure outlet (rho = 1) at x = NX-1, periodic y
    and z: the only steady state both boundaries admit is rho = 1, u = (0.04, 0, 0) everywhere,
    and the flow converges to it."""
    nx, ny, nz = 8, 3, 3
    c = cfg3(nx, ny, nz, omega=1.0, periodic=(0, 1, 1))
    c.u_inlet = 0.04
    o = O.Oracle3D(c, 64)
    o.set_flags(guo_flags(nx, ny, nz, False))
    o.init_from_macros(np.ones((nz, ny, nx)), np.zeros((nz, ny, nx, 3)))
    o.step(3000)
    rho, u = o.get_macros()
    assert np.abs(u[..., 0] - np.float32(0.04)).max() < 1e-9
    assert np.abs(rho - 1).max() < 1e-9 and np.abs(u[..., 1:]).max() < 1e-15


def test_guo_own_moments_quirk_changes_the_result():
    nx, ny, nz = 6, 4, 4
    rng = np.random.default_rng(9)
    out = []
    for quirk in (0, _abi.LBM_QUIRK_GUO3D_OWN_MOMENTS):
        c = cfg3(nx, ny, nz, omega=1.3, periodic=(0, 1, 1), quirks=quirk)
        c.u_inlet = 0.03
        o = O.Oracle3D(c, 64)
        o.set_flags(guo_flags(nx, ny, nz, False))
        rr = np.random.default_rng(9)
        o.init_from_macros(1 + 0.02 * rr.standard_normal((nz, ny, nx)), 0.02 * rr.standard_normal((nz, ny, nx, 3)))
        o.step(3)
        out.append(o.get_pops())
    assert np.abs(out[0] - out[1]).max() > 1e-6


def hermite2(P):
    """w_i / (2 cs^4) Q_i : P for a symmetric 3x3 P (the regularized non-equilibrium part), Q_i =
    c_i c_i - cs^2 I: its density and momentum vanish and its second moment is P."""
    cs2 = 1.0 / 3.0
    Qt = np.einsum("ia,ib->iab", C, C) - cs2 * np.eye(3)[None]
    return W / (2 * cs2 * cs2) * np.einsum("iab,ab->i", Qt, P)


def test_hermite_part_has_the_stated_moments():
    P = np.array([[2e-3, 5e-4, -3e-4], [5e-4, -1e-3, 7e-4], [-3e-4, 7e-4, 4e-4]])
    h = hermite2(P)
    assert abs(h.sum()) < 1e-15 and np.max(np.abs(C.T @ h)) < 1e-15
    np.testing.assert_allclose(np.einsum("i,ia,ib->ab", h, C, C), P, atol=1e-15)


@pytest.mark.parametrize("literal", [False, True])
def test_regularized_outlet_rebuilds_a_regularized_state_from_its_known_populations(literal):
    """RegularizedOutlet (regularizedOutlet.cuh:9-95) at x = NX-1 on a uniform regularized state
    f_eq(rho0, u0) + h(P): its unknown populations (c_x < 0, the slots streaming leaves stale)
    hold garbage, yet the functor returns f_eq + h exactly (non-equilibrium bounce of the knowns,
    then Pi(1) projected), with the neighbour's density and momentum. Wrong unknown sets, the
    node's own moments (which include the garbage) or a missing or mis-scaled Pi term all fail."""
    nx, ny, nz = 6, 5, 4
    rho0, u0 = 1.02, (0.03, -0.02, 0.01)
    P = np.array([[2e-3, 5e-4, -3e-4], [5e-4, -1e-3, 7e-4], [-3e-4, 7e-4, 4e-4]])
    f_unif = O.equilibrium3d(rho0, *u0, literal=literal) + hermite2(P)
    f0 = np.broadcast_to(f_unif[:, None, None, None], (27, nz, ny, nx)).copy()
    rng = np.random.default_rng(8)
    back = f0.copy()
    unknown = C[:, 0] < 0
    back[unknown, :, :, -1] = rng.uniform(0.0, 0.2, size=(int(unknown.sum()), nz, ny))
    c = cfg3(nx, ny, nz, periodic=(0, 1, 1))
    if literal:
        c.quirks = _abi.LBM_QUIRK_EQ3D_LINEAR
    fl = np.zeros((nz, ny, nx), np.uint8)
    fl[:, :, -1] = _abi.REGULARIZED_OUTLET
    fl[:, :, 0] = _abi.BOUNCE_BACK  # the x = 0 face's stale slots are not part of this check
    o = O.Oracle3D(c, 64)
    o.set_pops(f0, back)
    o.set_flags(fl)
    o.stream_boundaries()
    f = o.get_pops()
    if not literal:
        np.testing.assert_allclose(f[:, :, :, -1], f0[:, :, :, -1], atol=1e-14)
    # density and momentum are the neighbour's; the literal equilibrium carries 17/27 of the
    # momentum (its edge and corner directions' linear coefficient is 1 instead of 3)
    nb = f[:, :, :, -2]
    np.testing.assert_allclose(f[:, :, :, -1].sum(0), nb.sum(0), rtol=1e-14)
    share = 17.0 / 27.0 if literal else 1.0
    np.testing.assert_allclose(np.einsum("iab,ic->cab", f[:, :, :, -1], C),
                               share * np.einsum("iab,ic->cab", nb, C), rtol=1e-12, atol=1e-16)
    # without the functor the garbage stays
    o2 = O.Oracle3D(c, 64)
    o2.set_pops(f0, back)
    fl2 = fl.copy()
    fl2[:, :, -1] = _abi.FLUID
    o2.set_flags(fl2)
    o2.stream_boundaries()
    assert np.max(np.abs(o2.get_pops()[:, :, :, -1] - f0[:, :, :, -1])) > 1e-3


@pytest.mark.parametrize("quirk", [0, 1])
def test_extrapolated_corner_edge_is_the_inward_node_equilibrium(quirk):
    """ExtrapolatedCornerEdgeBoundary<3> (cornerBoundaries.cuh:6-36): every boundary node takes the
    second-order equilibrium of the node one step inward on each of its boundary axes, so its
    density and momentum are that node's post-stream ones (momentum z from the flat velocity
    table at 6 + i under LBM_QUIRK_CORNER3D_UZ), and its populations are f_eq of them; a flagged
    node off the faces is its own inward node and relaxes to its own equilibrium."""
    nx, ny, nz = 5, 6, 4
    rng = np.random.default_rng(9)
    f0 = W[:, None, None, None] * (1 + 0.1 * rng.random((27, nz, ny, nx)))
    c = cfg3(nx, ny, nz, periodic=(0, 0, 0))
    c.quirks = _abi.LBM_QUIRK_CORNER3D_UZ if quirk else 0

    def run(fl):
        o = O.Oracle3D(c, 64)
        o.set_pops(f0)
        if fl is not None:
            o.set_flags(fl)
        o.stream_boundaries()
        return o.get_pops()

    fe = run(np.full((nz, ny, nx), _abi.EXTRAPOLATED_CORNER_EDGE, np.uint8))
    fs = run(None)  # the post-stream state every functor reads
    flat = np.stack([C[:, 0], C[:, 1], C[:, 2]], 1).reshape(-1)  # [27][3] row-major
    for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
        xi = 1 if x == 0 else (nx - 2 if x == nx - 1 else x)
        yi = 1 if y == 0 else (ny - 2 if y == ny - 1 else y)
        zi = 1 if z == 0 else (nz - 2 if z == nz - 1 else z)
        g = fs[:, zi, yi, xi]
        r = g.sum()
        u = C.T @ g / r
        if quirk:
            u[2] = float(g @ flat[6:33]) / r
        want = O.equilibrium3d(r, *u)
        np.testing.assert_allclose(fe[:, z, y, x], want, rtol=1e-13, atol=1e-16)
        assert abs(fe[:, z, y, x].sum() - r) < 1e-13


def delta4(r):
    a = abs(r)
    if a < 1:
        return 0.125 * (3 - 2 * a + math.sqrt(1 + 4 * a - 4 * r * r))
    if a < 2:
        return 0.125 * (5 - 2 * a - math.sqrt(-7 + 12 * a - 4 * r * r))
    return 0.0


def test_peskin_kernel_square_sum_is_three_eighths():
    """The 4-point kernel's sum of squares over its support is 3/8 at every offset (one of its
    defining moment conditions), so a 4x4x4 stencil has sum phi^2 = (3/8)^3 = 27/512."""
    for r in np.linspace(0, 1, 17):
        assert abs(sum(delta4(r - j) ** 2 for j in range(-2, 3)) - 0.375) < 1e-12


def ibm3_run(u0, ds, iters, quirks=0, pos=(6.3, 5.7, 6.9), precision=64):
    n = 13
    c = cfg3(n, n, n, omega=1.2, periodic=(1, 1, 1), quirks=quirks)
    o = O.Oracle3D(c, precision)
    o.init_from_macros(np.full((n, n, n), 1.02), np.broadcast_to(np.array(u0), (n, n, n, 3)).copy())
    o.ibm_set_markers(np.array([pos]), None if ds is None else np.array([ds]), iters)
    o.step(1)
    return o.ibm_get_forces()


@pytest.mark.parametrize("ds,iters", [(None, 1), (None, 3), (0.7, 3), (2.0, 2)])
def test_single_marker_multi_direct_closed_form(ds, iters):
    """IBMManager<3>::multi_direct (IBMManager.cuh:222-252) for one marker in a uniform flow u0
    (periodic, rho0 = 1.02, no force): every interpolation is a partition of unity, so iteration k
    sees u_L = u0 (1 - a)^(k-1) with a = ds * sum phi^2 = ds * 27/512, and the Lagrangian forces sum
    to -2 rho0 u0 sum_k (1 - a)^(k-1); the force the fluid receives totals that times ds."""
    u0 = np.array([0.05, -0.03, 0.02])
    lag, tot = ibm3_run(u0, ds, iters)
    a = (1.0 if ds is None else ds) * 27.0 / 512.0
    series = sum((1 - a) ** k for k in range(iters))
    want = -2 * 1.02 * u0 * series
    np.testing.assert_allclose(lag[0], want, rtol=1e-12)
    np.testing.assert_allclose(tot, want * (1.0 if ds is None else ds), rtol=1e-12)


def test_single_marker_without_ds_and_with_the_clamp():
    """LBM_QUIRK_IBM_NO_DS spreads without the area (a = 27/512 whatever ds); the clamp quirk keeps
    only Lagrangian forces above 1e-8 and zeroes every corrected velocity component at or below
    1e-8 over the whole domain (IBM_impl.cuh:7-47): with u0 = (0.05, -0.03, 0.02) the first
    iteration keeps F_y = 0.06 rho0 only, and the corrected u_y < 0 becomes 0 everywhere, so the
    later iterations add nothing."""
    u0 = np.array([0.05, -0.03, 0.02])
    lag, tot = ibm3_run(u0, 0.7, 3, quirks=_abi.LBM_QUIRK_IBM_NO_DS)
    a = 27.0 / 512.0
    np.testing.assert_allclose(lag[0], -2 * 1.02 * u0 * (1 + (1 - a) + (1 - a) ** 2), rtol=1e-12)
    np.testing.assert_allclose(tot, lag[0], rtol=1e-12)
    lag, tot = ibm3_run(u0, None, 3, quirks=_abi.LBM_QUIRK_IBM_CLAMP)
    np.testing.assert_allclose(lag[0], [0.0, 2 * 1.02 * 0.03, 0.0], atol=1e-15)
