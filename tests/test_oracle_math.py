"""Checks that pin the oracle where the golden dumps leave room: closed forms,
invariants and brute force on tiny lattices.  Each would fail for a dropped
term, a wrong sign, or a transposed index in the function it names."""
import ctypes

import numpy as np
import pytest

from oracle.pyoracle import Oracle, hash_uniform, lib

CX = np.array([1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0, 0])
CY = np.array([0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0])
CZ = np.array([0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, 0])
G = np.array([2] * 6 + [1] * 13, dtype=float)
C3 = np.stack([CX, CY, CZ], axis=1).astype(float)


def boltz(bdist, u, du=(0.0, 0.0, 0.0)):
    out = (ctypes.c_double * 19)()
    lib().or_boltz_dist(bdist, *[float(v) for v in u], *[float(v) for v in du], out)
    return np.array(out[:])


def test_lattice_tensors():
    W = np.array([1 / 36] * 18 + [1 / 3])
    assert np.sum(G * W) == pytest.approx(1.0, abs=1e-16)
    np.testing.assert_allclose(np.einsum("i,i,ia,ib->ab", G, W, C3, C3), np.eye(3) / 3, atol=1e-16)
    # the Shan-Chen stencil sum_i g_i c_i c_i = 12 I used by the linear-field test
    np.testing.assert_allclose(np.einsum("i,ia,ib->ab", G[:18], C3[:18], C3[:18]), 12 * np.eye(3))


@pytest.mark.parametrize("bdist", [0, 2])
def test_equilibrium_moments(bdist):
    rng = np.random.default_rng(5)
    for _ in range(20):
        u = rng.uniform(-0.08, 0.08, 3)
        du = rng.uniform(-0.02, 0.02, 3)
        b = boltz(bdist, u, du)
        assert np.sum(G * b) == pytest.approx(1.0, abs=1e-15)
        np.testing.assert_allclose(np.einsum("i,i,ia->a", G, b, C3), u + du, atol=2e-16)
    b0 = boltz(bdist, (0, 0, 0))
    np.testing.assert_array_equal(b0, np.array([1 / 36] * 18 + [1 / 3]))


def test_bdist2_known_values():
    """Third-order Taylor form at u=(0.1,0,0): axis pair with s=0.3."""
    b = boltz(2, (0.1, 0.0, 0.0))
    uke = 0.5 * 0.1 * 0.3
    w = 1 / 36
    even = w * (1 - uke) + w / 2 * 0.09
    odd = w * (1 - uke) * 0.3 + w / 6 * 0.027
    assert b[0] == pytest.approx(even + odd, abs=1e-17)
    assert b[1] == pytest.approx(even - odd, abs=1e-17)
    assert b[2] == pytest.approx(w * (1 - uke), abs=1e-17)
    assert b[18] == pytest.approx((1 - uke) / 3, abs=1e-17)


def test_streaming_is_the_pull_permutation():
    """Advection alone on a 4x5x3 lattice with unique values: n_i(x) equals the
    pre-advection n_i(x - c_i), periodic wrap; the rest population stays."""
    nx, ny, nz = 4, 5, 3
    o = Oracle(nx, ny, nz, ncomp=2)
    N = o.N
    vals = np.arange(nz * ny * nx * 38, dtype=float).reshape(nz, ny, nx, 38) + 1.0
    N[1:-1, 1:-1, 1:-1, 0:19] = vals[..., :19]
    N[1:-1, 1:-1, 1:-1, 24:43] = vals[..., 19:]
    before = o.populations().copy()
    o.bc_apply(6)
    o.advection()
    after = o.populations()
    for c in range(2):
        for i in range(19):
            exp = np.roll(before[c, i], shift=(CZ[i], CY[i], CX[i]), axis=(0, 1, 2))
            np.testing.assert_array_equal(after[c, i], exp)
    # moments written by the advection: rho = sum g n
    I = o.interior()
    np.testing.assert_allclose(I[..., 19], np.einsum("i,izyx->zyx", G, after[0]), rtol=1e-15)


def test_sc_force_linear_psi_closed_form():
    """psi_b linear in space => sum_i g_i psi_b(x+c_i) c_i = 12 grad psi_b and
    F_r = -g_br psi_r(x) 12 grad psi_b (binary, code/lb3d_force_shanchen.F90:162-208)."""
    o = Oracle(8, 8, 8, ncomp=2, g_br=0.09)
    N = o.N
    z, y, x = np.meshgrid(np.arange(10), np.arange(10), np.arange(10), indexing="ij")
    grad = np.array([0.011, -0.007, 0.004])
    psi_b = 0.3 + grad[0] * x + grad[1] * y + grad[2] * z
    psi_r = 0.45 - 0.5 * grad[0] * x + 0.2 * grad[2] * z
    N[..., 43] = -np.log1p(-psi_b)
    N[..., 19] = -np.log1p(-psi_r)
    f = o.sc_force(4, 5, 3)
    pr = psi_r[3, 5, 4]
    pb = psi_b[3, 5, 4]
    np.testing.assert_allclose(f[0], -0.09 * pr * 12 * grad, rtol=1e-11, atol=1e-16)
    grad_r = np.array([-0.5 * grad[0], 0.0, 0.2 * grad[2]])
    np.testing.assert_allclose(f[1], -0.09 * pb * 12 * grad_r, rtol=1e-11, atol=1e-16)


def test_rest_state_is_a_fixed_point():
    for ncomp in (2, 3):
        o = Oracle(6, 5, 4, ncomp=ncomp, bdist=2)
        o.init_noise(seed=3, rho_mean=(0.7, 0.6, 0.5), amp=(0.0, 0.0, 0.0), dipoles=False)
        o.bc_init()
        before = o.populations().copy()
        o.run(3)
        np.testing.assert_allclose(o.populations(), before, rtol=1e-14, atol=0)


def test_mass_exact_per_component_and_determinism():
    o = Oracle(12, 10, 8, ncomp=2)
    o.init_noise(seed=7)
    o.bc_init()
    m0, p0, _ = o.sanity()
    o.run(20)
    m1, p1, c = o.sanity()
    np.testing.assert_allclose(m1, m0, rtol=1e-13)
    # total momentum of a periodic system with pair-antisymmetric forces stays ~0
    assert np.all(np.abs(p1) < 1e-12)
    o2 = Oracle(12, 10, 8, ncomp=2)
    o2.init_noise(seed=7)
    o2.bc_init()
    o2.run(20)
    np.testing.assert_array_equal(o2.populations(), o.populations())


def test_le_uniform_rest_fluid_halo_is_shifted_equilibrium():
    """A uniform fluid at rest: the x=0 halo becomes rho*feq(0,0,-2U) and the
    x=nx+1 halo rho*feq(0,0,+2U) (code/lb3d_bc_leesedwards.F90:682-825, 863-1003)."""
    U = 0.02
    o = Oracle(6, 4, 8, ncomp=2, lees_edwards=1, shear_u=U)
    o.init_noise(seed=1, rho_mean=(0.7, 0.5, 0.0), amp=(0, 0, 0), dipoles=False)
    o.bc_init()
    Ueff = U * 6 / 5
    N = o.N
    lo = N[1:-1, 1:-1, 0, :]
    hi = N[1:-1, 1:-1, -1, :]
    for c, rho in ((0, 0.7), (1, 0.5)):
        np.testing.assert_allclose(lo[..., 24 * c:24 * c + 19],
                                   np.broadcast_to(rho * boltz(2, (0, 0, -2 * Ueff)), lo[..., :19].shape),
                                   rtol=0, atol=1e-16)
        np.testing.assert_allclose(hi[..., 24 * c:24 * c + 19],
                                   np.broadcast_to(rho * boltz(2, (0, 0, 2 * Ueff)), hi[..., :19].shape),
                                   rtol=0, atol=1e-16)


def test_le_interpolation_of_a_z_linear_field():
    """Populations linear in z: the low halo equals the field shifted by
    +offset (2*shear_sum) in z, the high halo shifted by -offset, before the
    velocity correction.  Checked on the halo density with U chosen so the
    correction vanishes (shear_u=0 gives off=0, so use a direct restore)."""
    nx, ny, nz = 5, 3, 12
    o = Oracle(nx, ny, nz, ncomp=2, lees_edwards=1, shear_u=0.0, restore=1, shear_sum=1.35)
    I = o.interior()
    zz = np.arange(1, nz + 1, dtype=float)[:, None, None]
    # rest-equilibrium populations with density linear in z inside a window
    rho = 0.6 + 0.01 * zz + 0 * I[..., 0]
    W = np.array([1 / 36] * 18 + [1 / 3])
    I[..., 0:19] = rho[..., None] * W
    I[..., 24:43] = 0.5 * W
    o.bc_init()
    s, lelo, f, _ = o.le_state()
    off = 2 * s
    assert lelo == int(off) and f == pytest.approx(off - int(off))
    N = o.N
    lo_rho = np.einsum("i,zyi->zy", G, N[1:-1, 1:-1, 0, 0:19])
    hi_rho = np.einsum("i,zyi->zy", G, N[1:-1, 1:-1, -1, 0:19])
    # away from the wrap in z, interpolation is exact on a linear field
    for z in range(1, nz - 4):
        assert lo_rho[z - 1, 0] == pytest.approx(0.6 + 0.01 * (z + off), abs=1e-14)
    for z in range(5, nz + 1):
        assert hi_rho[z - 1, 0] == pytest.approx(0.6 + 0.01 * (z - off), abs=1e-14)


def test_dipole_relaxation_langevin_closed_form():
    """Uniform ternary fluid with uniform dipoles d0: after advection
    da = rho_s d0, the colour term vanishes and the surfactant sum cancels
    (sum_i g_i (da - 2 c_i c_i.da) = 0), so h = da and
    db = w_d d_0 L(beta|h|) h/|h| + (1-w_d) da/rho_s with L = coth - 1/x."""
    tau_d, beta, d_0 = 2.0, 3.0, 0.8
    o = Oracle(5, 4, 6, ncomp=3, tau_d=tau_d, beta=beta, d_0=d_0)
    o.init_noise(seed=2, rho_mean=(0.7, 0.7, 0.6), amp=(0, 0, 0), dipoles=False)
    d0 = np.array([0.3, -0.5, 0.2])
    o.interior()[..., 75:78] = d0
    o.bc_init()
    o.run(1)
    h = 0.6 * d0
    hm = beta * np.linalg.norm(h)
    L = 1 / np.tanh(hm) - 1 / hm
    wd = 1 / tau_d
    expect = wd * d_0 * L * h * beta / hm + (1 - wd) * h / 0.6
    db = o.dipole()
    np.testing.assert_allclose(db[:, 2, 1, 3], expect, rtol=1e-13)


def test_hash_uniform_range_and_determinism():
    v = [hash_uniform(1, s, i) for s in range(3) for i in range(200)]
    assert min(v) >= 0.0 and max(v) < 1.0
    assert hash_uniform(1, 0, 5) == hash_uniform(1, 0, 5)
    assert hash_uniform(1, 0, 5) != hash_uniform(2, 0, 5)


def test_stress_of_an_equilibrium_field():
    """Momentum transfer through x-planes (code/lb3d_io_stress.F90:88-173) for
    equilibrium populations rho_c(x)*feq(u) with uniform u = (ux, 0, uz):
    plane minx gets rho(minx-1)*(b9-b10) + rho(minx)*(b14-b13) per (y, z) and
    component (1-based vectors), plane maxx rho(maxx+1)*(b13-b14) +
    rho(maxx)*(b10-b9); halved.  For uniform rho both reduce to the textbook
    flux Pi_xz = rho*ux*uz."""
    nx, ny, nz = 6, 4, 5
    ux, uz = 0.03, -0.02
    b = boltz(2, (ux, 0.0, uz))
    np.testing.assert_allclose(b[8] - b[9] + b[13] - b[12], ux * uz, rtol=1e-13)
    rho_x = {0: 0.61 + 0.013 * np.arange(nx), 1: 0.52 + 0.007 * np.arange(nx) ** 2}
    o = Oracle(nx, ny, nz, ncomp=2)
    I = o.interior()
    for c in (0, 1):
        I[..., 24 * c:24 * c + 19] = rho_x[c][None, None, :, None] * b[None, None, None, :]
    o.halo_refresh()

    def rho(c, x):  # 1-based plane, periodic images at 0 and nx+1
        return rho_x[c][(x - 1) % nx]

    for minx, maxx in ((1, nx), (2, 5), (3, 3)):
        s0, s1, mean = o.stress(minx, maxx)
        e0 = sum(rho(c, minx - 1) * (b[8] - b[9]) + rho(c, minx) * (b[13] - b[12]) for c in (0, 1))
        e1 = sum(rho(c, maxx + 1) * (b[12] - b[13]) + rho(c, maxx) * (b[9] - b[8]) for c in (0, 1))
        np.testing.assert_allclose(s0, 0.5 * ny * nz * e0, rtol=1e-12)
        np.testing.assert_allclose(s1, 0.5 * ny * nz * e1, rtol=1e-12)
        np.testing.assert_allclose(mean, (s1 - s0) / (ny * nz), rtol=1e-14)
    # uniform density: the flux is rho*ux*uz through either plane (with opposite signs)
    o2 = Oracle(nx, ny, nz, ncomp=2)
    I2 = o2.interior()
    for c, r in ((0, 0.7), (1, 0.4)):
        I2[..., 24 * c:24 * c + 19] = r * b
    o2.halo_refresh()
    s0, s1, _ = o2.stress()
    np.testing.assert_allclose(s0, 0.5 * ny * nz * 1.1 * ux * uz, rtol=1e-12)
    np.testing.assert_allclose(s1, -0.5 * ny * nz * 1.1 * ux * uz, rtol=1e-12)


def test_sanity_full_uniform_flow_and_ties():
    """Uniform equilibrium flow u: every site has velocity u (amass 1), so the
    sums are ns*u, every site ties for the maximum and the reference's `>=`
    keeps the last site of its loop, (nx, ny, nz); |u| > 0.1 makes every site
    too fast.  One faster site then takes the maximum; a NaN is counted."""
    from oracle.pyoracle import Oracle
    nx, ny, nz = 5, 4, 3
    u = (0.05, -0.07, 0.06)
    b = boltz(2, u)
    o = Oracle(nx, ny, nz, ncomp=2)
    I = o.interior()
    for c, r in ((0, 0.7), (1, 0.4)):
        I[..., 24 * c:24 * c + 19] = r * b
    s = o.sanity_full()
    ns = nx * ny * nz
    assert s["fluid_sites"] == ns and s["fast_sites"] == ns and s["nan_sites"] == 0
    for c in (0, 1):
        np.testing.assert_allclose(s["vel_sum"][c], ns * np.array(u), rtol=1e-12)
        np.testing.assert_allclose(s["max_vel"][c], np.abs(u), rtol=1e-12)
        assert (s["max_pos"][c] == [nx, ny, nz]).all()
    np.testing.assert_allclose(s["mass"][:2], [0.7 * ns, 0.4 * ns], rtol=1e-13)
    np.testing.assert_allclose(s["momentum"], 1.1 * ns * np.array(u), rtol=1e-12)
    # a faster red site at (x, y, z) = (2, 3, 1), 1-based
    I[0, 2, 1, 0:19] = 0.7 * boltz(2, (0.08, -0.02, 0.0))
    s = o.sanity_full()
    assert list(s["max_pos"][0][0]) == [2, 3, 1]
    assert list(s["max_pos"][0][1]) == [nx, ny, nz] and list(s["max_pos"][1][0]) == [nx, ny, nz]
    np.testing.assert_allclose(s["max_vel"][0][0], 0.08, rtol=1e-12)
    I[2, 0, 4, 24 + 3] = np.nan
    assert o.sanity_full()["nan_sites"] == 1


def boltz_dd(bdist, u, du=(0.0, 0.0, 0.0), dd=(0.0, 0.0, 0.0)):
    out = (ctypes.c_double * 19)()
    lib().or_boltz_dist_dd(bdist, *[float(v) for v in u], *[float(v) for v in du], *[float(v) for v in dd], out)
    return np.array(out[:])


@pytest.mark.parametrize("bdist", [3, 4, 6])
def test_other_equilibria_mass_and_momentum(bdist):
    """bdist 3 (second order in u and du), 4 (Guo, second order), 6 (fourth
    order): sum g b = 1 and sum g b c = u + du, g the link multiplicities (isotropy of the D3Q19 weights up to
    the fourth moment)."""
    u, du = np.array([0.031, -0.017, 0.022]), np.array([0.004, 0.002, -0.003])
    b = boltz_dd(bdist, u, du, (0.01, -0.02, 0.005) if bdist == 4 else (0, 0, 0))
    np.testing.assert_allclose(G @ b, 1.0, rtol=0, atol=1e-15)
    np.testing.assert_allclose(C3.T @ (G * b), u + du, rtol=1e-13, atol=1e-16)


def test_guo_second_order_stress_and_fourth_order_forms():
    """bdist 4: sum b c c = T I + (u+du)(u+du) - dd dd (Guo et al., PRE 65
    046308); bdist 5 with dd = 0 is bdist 6 exactly, and with dd its first
    moment is tu + (u1ke - uke) tu1 / T with tu1 = tu - dd."""
    u, du, dd = np.array([0.02, -0.01, 0.03]), np.array([0.001, 0.0, -0.002]), np.array([0.004, -0.006, 0.002])
    tu = u + du
    b4 = boltz_dd(4, u, du, dd)
    P = (C3.T * (G * b4)) @ C3
    np.testing.assert_allclose(P, np.eye(3) / 3.0 + np.outer(tu, tu) - np.outer(dd, dd), rtol=0, atol=1e-16)
    np.testing.assert_array_equal(boltz_dd(5, u, du), boltz_dd(6, u, du))
    b5 = boltz_dd(5, u, du, dd)
    tu1 = tu - dd
    uke, u1ke = 0.5 * tu @ tu, 0.5 * tu1 @ tu1
    np.testing.assert_allclose(C3.T @ (G * b5), tu + (u1ke - uke) * tu1 * 3.0, rtol=1e-12, atol=1e-17)
    # bdist 6 conserves mass to fourth order exactly
    np.testing.assert_allclose(G @ boltz_dd(6, u, du), 1.0, rtol=0, atol=1e-15)


def test_hudong_iterated_exponential():
    """bdist 1 (code/lb3d_bdist.F90:237-298): b_i = rho_eq w_i exp((c_i.u_eq -
    |u_eq|^2/2)/T), so log(b_i / w_i) is affine in c_i.  The damped iteration
    drives the plain (unweighted, as the reference sums them) moments towards
    (1, u + du); max(2, 200|u|) iterations of step sqrt(1/n) converge further
    the faster the flow."""
    w = np.array([1 / 36.0] * 18 + [1 / 3.0])
    A = np.hstack([np.ones((19, 1)), C3])
    errs = []
    for m in (0.05, 0.1, 0.2, 0.3):
        u = np.array([0.6, -0.48, 0.64]) * m
        b = boltz_dd(1, u - 0.01, (0.01, 0.01, 0.01))
        y = np.log(b / w)
        coef = np.linalg.lstsq(A, y, rcond=None)[0]
        np.testing.assert_allclose(A @ coef, y, rtol=0, atol=1e-13)
        errs.append((abs(b.sum() - 1.0), np.abs(C3.T @ b - u).max()))
    assert all(errs[k + 1][0] < errs[k][0] for k in range(3))
    assert errs[-1][0] < 3e-4 and errs[-1][1] < 2e-4


def zou_he(n, uz, itype, low):
    buf = (ctypes.c_double * 19)(*[float(v) for v in n])
    lib().or_zou_he_site(buf, float(uz), int(itype), int(low))
    return np.array(buf[:])


@pytest.mark.parametrize("itype", [-1, 0, 1])
@pytest.mark.parametrize("low", [1, 0])
def test_zou_he_flux_rule_moments(itype, low):
    """lbe_invade_zouandheflux (code/lb3d_invasion.F90:2143-2457) on one site:
    only the populations entering the lattice change; with K, O, I the sums
    of g n over c_z = 0, leaving and entering, mass and z momentum fix
    rho = K + 2 O + m (mass flow m, itype < 0) or (K + 2 O) / (1 - uz), and
    the z momentum rho uz (= m for itype < 0), in +z on both planes.  The
    corrected diagonals (itype +-1) leave no momentum along x and y."""
    rng = np.random.default_rng(7 + itype + 3 * low)
    n = 0.02 + 0.01 * rng.random(19)
    cz = C3[:, 2]
    enter = cz > 0 if low else cz < 0
    flux = 0.013
    n2 = zou_he(n, flux, itype, low)
    np.testing.assert_array_equal(n2[~enter], n[~enter])
    F = G * n
    K, O = F[cz == 0].sum(), F[~enter & (cz != 0)].sum()
    if itype < 0:
        rho, mz = K + 2 * O + (flux if low else -flux), flux
    else:
        rho = (K + 2 * O) / (1 - flux if low else 1 + flux)
        mz = rho * flux
    F2 = G * n2
    np.testing.assert_allclose(F2.sum(), rho, rtol=1e-14)
    np.testing.assert_allclose(F2 @ cz, mz, rtol=1e-12)
    if itype != 0:
        np.testing.assert_allclose(F2 @ C3[:, :2], [0.0, 0.0], atol=1e-15)
