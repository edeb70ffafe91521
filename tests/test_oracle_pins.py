"""Pins of the oracle against facts fixed by the mathematics of the method
(the reference ships no tests or golden vectors: it has no test directory,
and its kernels need nvcc, which is not available here).

Each test is chosen so that a plausible slip in the oracle (a wrong sign or
index in a velocity table, a dropped force term, a transposed neighbour, a
missing double promotion in the EOS) fails it.
"""
import math

import numpy as np
import pytest

from oracle import oracle

EX, EY, EZ = oracle.EX, oracle.EY, oracle.EZ


def test_lattice_tables_moments():
    ex, ey, ez, opp, w32, w64 = oracle.tables()
    assert list(ex) == list(EX) and list(ey) == list(EY) and list(ez) == list(EZ)
    e = np.stack([ex, ey, ez], 1).astype(float)
    assert abs(w64.sum() - 1.0) < 1e-15
    assert np.allclose(w64 @ e, 0.0, atol=1e-16)
    second = np.einsum("i,ia,ib->ab", w64, e, e)
    assert np.allclose(second, np.eye(3) / 3.0, atol=1e-15)
    # fourth-order isotropy of D3Q19: sum w e_a^2 e_b^2 = 1/9 (a != b), 1/3 (a == b)
    fourth = np.einsum("i,ia,ia,ib,ib->ab", w64, e, e, e, e)
    assert np.allclose(fourth, np.full((3, 3), 1 / 9) + np.eye(3) * (1 / 3 - 1 / 9), atol=1e-15)
    for i in range(19):
        assert np.array_equal(e[opp[i]], -e[i])
        assert opp[opp[i]] == i
    # the reference narrows the double quotients to float (lbm_model.cuh:35-37)
    assert w32[0] == float(np.float32(1.0 / 3.0)) and w32[1] == float(np.float32(1.0 / 18.0))
    assert w32[4] == float(np.float32(1.0 / 36.0))


def test_exp_against_libm():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-30, 5, 20000), -rng.uniform(0, 3, 20000), [0.0, -1e-30, -87.0, 80.0]])
    worst = 0.0
    same_f = 0
    for x in xs:
        v = oracle.exp(float(x))
        ref = math.exp(float(x))
        worst = max(worst, abs(v - ref) / ref)
        # narrowed to float: the correctly rounded float of exp(x) in almost all cases
        if np.float32(v) == np.float32(ref):
            same_f += 1
        else:
            assert abs(float(np.float32(v)) - ref) <= float(np.spacing(np.float32(ref)))
    assert worst < 4e-16
    assert same_f >= len(xs) - 2
    assert math.isinf(oracle.exp(710.0)) and oracle.exp(-750.0) == 0.0


def _perm_state(shape, dtype, seed=0):
    rng = np.random.default_rng(seed)
    return rng.integers(1, 1000, size=shape + (19,)).astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_streaming_is_the_exact_permutation(dtype):
    shape = (5, 6, 7)
    fA = _perm_state(shape, dtype)
    P = oracle.Params(*shape)
    fB = oracle.stream(P, np.ones(shape, np.uint8), fA)
    for i in range(19):
        want = np.roll(fA[..., i], shift=(EX[i], EY[i], EZ[i]), axis=(0, 1, 2))
        assert np.array_equal(fB[..., i], want), f"direction {i}"


def test_single_fluid_node_reflects_every_population():
    shape = (5, 5, 5)
    fA = _perm_state(shape, np.float64, 1)
    mask = np.zeros(shape, np.uint8)
    mask[2, 2, 2] = 1
    fB = oracle.stream(oracle.Params(*shape), mask, fA)
    opp = [0, 10, 11, 12, 13, 14, 15, 16, 17, 18, 1, 2, 3, 4, 5, 6, 7, 8, 9]
    for i in range(19):
        assert fB[2, 2, 2, i] == fA[2, 2, 2, opp[i]]


def test_bounce_back_against_brute_force():
    shape = (6, 5, 7)
    rng = np.random.default_rng(4)
    mask = (rng.uniform(size=shape) > 0.3).astype(np.uint8)
    fA = _perm_state(shape, np.float64, 2)
    fB = oracle.stream(oracle.Params(*shape), mask, fA)
    opp = [0, 10, 11, 12, 13, 14, 15, 16, 17, 18, 1, 2, 3, 4, 5, 6, 7, 8, 9]
    for x in range(shape[0]):
        for y in range(shape[1]):
            for z in range(shape[2]):
                if not mask[x, y, z]:
                    continue
                for i in range(19):
                    ux, uy, uz = (x - EX[i]) % shape[0], (y - EY[i]) % shape[1], (z - EZ[i]) % shape[2]
                    want = fA[ux, uy, uz, i] if mask[ux, uy, uz] else fA[x, y, z, opp[i]]
                    assert fB[x, y, z, i] == want


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_uniform_psi_gives_zero_force(dtype):
    shape = (4, 5, 6)
    psi = np.full(shape, 0.37, dtype)
    F = oracle.force(oracle.Params(*shape, G=-5.0), np.ones(shape, np.uint8), psi)
    # zero up to the rounding of the ordered 10-term sums
    eps = float(np.finfo(dtype).eps)
    assert np.all(np.abs(F) <= 8 * eps * 5.0 * 0.37 * 0.37 / 3)


def test_linear_psi_gives_exact_gradient_force():
    """F = -G psi(x) c_s^2 grad(psi) exactly for linear psi (sum w e e = I/3)."""
    shape = (8, 9, 10)
    g = np.array([0.013, -0.007, 0.004])
    X, Y, Z = np.meshgrid(*[np.arange(s, dtype=float) for s in shape], indexing="ij")
    psi = 0.5 + g[0] * X + g[1] * Y + g[2] * Z
    G = -5.0
    F = oracle.force(oracle.Params(*shape, G=G), np.ones(shape, np.uint8), psi)
    inner = (slice(1, -1),) * 3
    want = -G * psi[..., None] * g / 3.0
    assert np.allclose(F[inner], want[inner], rtol=1e-12, atol=1e-15)


def test_periodic_force_sums_to_zero():
    shape = (7, 6, 8)
    rng = np.random.default_rng(5)
    psi = rng.uniform(0.1, 0.9, shape)
    F = oracle.force(oracle.Params(*shape, G=-5.0), np.ones(shape, np.uint8), psi)
    assert np.all(np.abs(F.sum(axis=(0, 1, 2))) <= 1e-12 * np.abs(F).sum())


def test_collision_moments():
    """sum f* = rho and sum f* e = m + F (velocity-shift forcing, any tau)."""
    shape = (3, 4, 5)
    rng = np.random.default_rng(6)
    w = oracle.tables()[5]
    rho0 = rng.uniform(0.5, 1.5, shape)
    fB = rho0[..., None] * w * (1 + 0.05 * rng.standard_normal(shape + (19,)))
    rho = fB.sum(-1)
    F = 1e-3 * rng.standard_normal(shape + (3,))
    e = np.stack([EX, EY, EZ], 1).astype(float)
    for tau in (1.0, 0.8, 1.7):
        P = oracle.Params(*shape, tau=tau)
        out, u = oracle.collide(P, np.ones(shape, np.uint8), fB, rho, F)
        assert np.allclose(out.sum(-1), rho, rtol=1e-14)
        m = fB @ e
        assert np.allclose(out @ e, m + F, rtol=1e-12, atol=1e-15)
        assert np.allclose(u, (m + tau * F) / rho[..., None], rtol=1e-14)


def _packing(shape, seed):
    from tests.helpers import random_spheres, sphere_mask
    return sphere_mask(shape, random_spheres(shape, 6, 2.0, 4.0, seed))


def test_fp64_mass_conservation():
    shape = (24, 20, 22)
    fluid = _packing(shape, 1)
    rng = np.random.default_rng(2)
    rho = 0.7 + rng.uniform(-0.01, 0.01, shape)
    st = oracle.State(oracle.Params(*shape, G=-5.0, rhos=0.7), fluid, rho, dtype=np.float64)
    m0 = st.fA[fluid].sum()
    st.run(60)
    assert abs(st.fA[fluid].sum() - m0) <= 1e-12 * m0
    assert st.rho[fluid].max() > 0.74  # the spinodal mixture has started to separate (initial max 0.71)


def test_fp64_momentum_conservation_periodic():
    shape = (16, 14, 12)
    rng = np.random.default_rng(3)
    rho = 0.7 + rng.uniform(-0.05, 0.05, shape)
    st = oracle.State(oracle.Params(*shape, G=-5.0), np.ones(shape, bool), rho, dtype=np.float64)
    e = np.stack([EX, EY, EZ], 1).astype(float)
    st.run(80)
    mom = (st.fA @ e).sum(axis=(0, 1, 2))
    assert np.all(np.abs(mom) <= 1e-12 * st.fA.sum())


def test_injection_adds_increment_over_tau():
    """With injection type 0 the density used by the collision gains inc, so
    the post-collision mass of each node gains inc/tau."""
    shape = (6, 10, 4)
    rho = np.full(shape, 1.0)
    P = oracle.Params(*shape, G=0.0, tau=0.8, rho_inject_period=1, rho_increment=0.01, inject_position=0.5)
    st = oracle.State(P, np.ones(shape, bool), rho, dtype=np.float64)
    st.run(1)
    mass = st.fA.sum(-1)
    ys = np.arange(shape[1])
    inj = ys < 0.5 * shape[1]
    assert np.allclose(mass[:, inj, :], 1.0 + 0.01 / 0.8, rtol=1e-14)
    assert np.allclose(mass[:, ~inj, :], 1.0, rtol=1e-14)


# --- equation of state -----------------------------------------------------

def _press(rho, G=-5.0):
    return oracle.press(oracle.Params(1, 1, 1, G=G), rho)


def _psi(rho, G=-5.0):
    return oracle.psi(oracle.Params(1, 1, 1, G=G), rho)


def test_shan_chen_eos_closed_form():
    for rho in (0.1, 0.7, 1.9):
        psi = 1 - math.exp(-rho)
        assert abs(_psi(rho) - psi) < 1e-15
        assert abs(_press(rho) - (rho / 3 + (-5.0) * psi * psi / 6)) < 1e-15


@pytest.mark.parametrize("G,spin", [(-5.0, (0.323507, 1.285931)), (-4.5, (0.405465, 1.098612)),
                                    (-5.5, (0.272969, 1.431779))])
def test_spinodal_points(G, spin):
    h = 1e-6
    for r in spin:
        dp = (_press(r + h, G) - _press(r - h, G)) / (2 * h)
        assert abs(dp) < 2e-6


@pytest.mark.parametrize("G,coex", [(-5.0, (0.156413, 1.932442)), (-4.5, (0.254125, 1.494301)),
                                    (-5.5, (0.104813, 2.307522))])
def test_coexistence_mechanical_stability(G, coex):
    """Equal pressure and int_(rho_v)^(rho_l) (p0 - p) psi'/psi^2 drho = 0."""
    from scipy.integrate import quad
    rv, rl = coex
    p0 = _press(rv, G)
    assert abs(_press(rl, G) - p0) < 2e-6
    dpsi = lambda r: math.exp(-r)  # noqa: E731
    integ, _ = quad(lambda r: (p0 - _press(r, G)) * dpsi(r) / _psi(r, G) ** 2, rv, rl, epsabs=1e-12)
    assert abs(integ) < 1e-5


def test_carnahan_starling_pressure_and_psi():
    P = oracle.Params(1, 1, 1, G=-1.0, carn_star=True, T_Tc=0.7)
    T = float(np.float32(0.7)) * float(np.float32(0.0943))
    for rho in (0.05, 0.2, 0.35):
        eta = rho
        p = rho * T * (1 + eta + eta * eta - eta ** 3) / (1 - eta) ** 3 - rho * rho
        assert abs(oracle.press(P, rho) - p) < 1e-6 * max(1.0, abs(p))
        psi = math.sqrt(2 * (oracle.press(P, rho) - rho / 3) / (-1.0 / 3))
        assert abs(oracle.psi(P, rho) - psi) < 1e-12


@pytest.mark.slow
def test_flat_interface_coexistence_g5():
    """A liquid slab in vapour relaxes to the coexistence densities; the
    velocity-shift forcing moves them slightly from the mechanical-stability
    values, so the check is at 2%."""
    shape = (64, 2, 2)
    rho = np.where((np.arange(64) >= 16) & (np.arange(64) < 48), 1.9, 0.16)[:, None, None] * np.ones(shape)
    st = oracle.State(oracle.Params(*shape, G=-5.0, tau=1.0), np.ones(shape, bool), rho, dtype=np.float64)
    st.run(4000)
    rl, rv = st.rho[32, 0, 0], st.rho[0, 0, 0]
    assert abs(rl - 1.932442) / 1.932442 < 0.02
    assert abs(rv - 0.156413) / 0.156413 < 0.02


@pytest.mark.slow
def test_laplace_law():
    """p_in - p_out = sigma / R for a droplet: the product (dp * R) is the
    same for two radii."""
    prods = []
    for R in (9.0, 12.0):  # a 6-cell droplet evaporates into the 40^3 vapour box
        L = 40
        X, Y, Z = np.meshgrid(*[np.arange(L, dtype=float)] * 3, indexing="ij")
        r = np.sqrt((X - L / 2) ** 2 + (Y - L / 2) ** 2 + (Z - L / 2) ** 2)
        rho = np.where(r < R, 1.9, 0.16)
        st = oracle.State(oracle.Params(L, L, L, G=-5.0, tau=1.0), np.ones((L,) * 3, bool), rho, dtype=np.float64)
        st.run(1000)
        liquid = st.rho > 1.0
        Rm = (liquid.sum() * 3 / (4 * math.pi)) ** (1 / 3)
        dp = st.p[L // 2, L // 2, L // 2] - st.p[0, 0, 0]
        prods.append(dp * Rm)
    assert abs(prods[0] - prods[1]) / prods[1] < 0.05


def test_poiseuille_curvature():
    """Body-force channel flow between bounce-back walls: -nu u'' = g for the
    physical velocity (m + F/2)/rho, G = 0."""
    ly = 18
    shape = (2, ly, 2)
    fluid = np.ones(shape, bool)
    fluid[:, 0, :] = False
    fluid[:, -1, :] = False
    g = 1e-6
    tau = 1.0
    st = oracle.State(oracle.Params(*shape, G=0.0, tau=tau, body_force=(g, 0.0, 0.0)), fluid,
                      np.ones(shape), dtype=np.float64)
    st.run(6000)
    e = np.stack([EX, EY, EZ], 1).astype(float)
    m = st.fA[0, 1:-1, 0] @ e
    rho = st.fA[0, 1:-1, 0].sum(-1)
    u = (m[:, 0] + 0.5 * rho * g) / rho
    nu = (tau - 0.5) / 3
    y = np.arange(1, ly - 1, dtype=float)
    c = np.polyfit(y, u, 2)
    assert abs(-2 * c[0] * nu / g - 1.0) < 0.02
    resid = u - np.polyval(c, y)
    assert np.abs(resid).max() < 1e-3 * u.max()


# --- pressure faces (anti-bounce-back, iteration_device.cuh:213-243, 282-286)

def test_pressure_faces_at_rest_keep_the_face_density():
    """Fluid at rest with density rho_w everywhere and both x faces at rho_w:
    the entering populations -A_opp + 2 t_i rho_w (1 + 4.5 (e.u)^2 - 1.5 u^2)
    are t_i rho_w again, so every population stays t_i rho_w (to rounding)
    for 200 steps.  A wrong sign of A_opp gives 3 t_i rho_w, a dropped factor
    2 gives 0, and both move the state at the first step."""
    shape = (12, 5, 4)
    rw = 1.3
    P = oracle.Params(*shape, G=0.0, tau=0.9, face_type=(1, 1, 0, 0, 0, 0), face_rho=(rw, rw, 0, 0, 0, 0))
    st = oracle.State(P, np.ones(shape, bool), np.full(shape, rw), dtype=np.float64)
    w = oracle.tables()[5]
    st.run(200)
    assert np.allclose(st.fA, w * rw, rtol=1e-13, atol=0)
    assert np.allclose(st.rho, rw, rtol=1e-13)


def test_pressure_driven_channel():
    """G = 0 channel between halfway bounce-back walls (y = 0 and ly - 1
    solid), pressure faces rho_in = 1.01 at x = 0 and rho_out = 0.99 at
    x = lx - 1, periodic z.  At steady state:
    * the mass flux runs from the high to the low density face and is the
      same through every cross-section (mass conservation);
    * the cross-section density is linear in x and extrapolates to rho_in
      and rho_out half a node outside the face nodes, where anti-bounce-back
      imposes them;
    * the velocity profile is Poiseuille's: nu rho u'' = c_s^2 drho/dx."""
    lx, ly, lz = 48, 12, 2
    shape = (lx, ly, lz)
    fluid = np.ones(shape, bool)
    fluid[:, 0, :] = fluid[:, -1, :] = False
    rin, rout, tau = 1.01, 0.99, 1.0
    P = oracle.Params(*shape, G=0.0, tau=tau, rhos=1.0, face_type=(1, 1, 0, 0, 0, 0),
                      face_rho=(rin, rout, 0, 0, 0, 0))
    st = oracle.State(P, fluid, np.ones(shape), dtype=np.float64)
    st.run(6000)
    e = np.stack([EX, EY, EZ], 1).astype(float)
    f = st.fA[:, 1:-1, 0, :]                      # fluid rows [x][y][19]
    rho = f.sum(-1)
    jx = (f @ e)[..., 0]                          # x momentum (no force, G = 0)
    flux = jx.sum(1)                              # through each cross-section
    assert np.all(flux > 0)
    assert np.ptp(flux) <= 1e-6 * flux.mean()
    rbar = rho.mean(1)
    x = np.arange(lx, dtype=float)
    inner = slice(4, lx - 4)
    c = np.polyfit(x[inner], rbar[inner], 1)
    assert np.abs(rbar[inner] - np.polyval(c, x[inner])).max() < 1e-3 * (rin - rout)
    assert abs(np.polyval(c, -0.5) - rin) < 0.02 * (rin - rout)
    assert abs(np.polyval(c, lx - 0.5) - rout) < 0.02 * (rin - rout)
    # Poiseuille curvature in the middle of the channel
    xm = lx // 2
    u = jx[xm] / rho[xm]
    y = np.arange(1, ly - 1, dtype=float)
    q = np.polyfit(y, u, 2)
    nu = (tau - 0.5) / 3
    want = (c[0] / 3.0) / (nu * rho[xm].mean())   # u'' = c_s^2 (drho/dx) / (nu rho)
    assert abs(2 * q[0] / want - 1) < 0.03


# --- injection types 1 and 2 (iteration_device.cuh:47-59) -------------------

@pytest.mark.parametrize("inject_type", [0, 1, 2])
@pytest.mark.parametrize("reverse", [1, -1])
def test_injection_mass_gain_against_brute_force(inject_type, reverse):
    """One step from rest with a random density field: the post-collision
    mass of a node is its streamed density plus inc/tau when it is eligible
    (y < inject_position * ly and, for type 1, streamed density below
    rhol_spinodal, for type 2 above), and unchanged otherwise.  The streamed
    density is formed here by brute force (a sum of rolled t_i rho)."""
    shape = (7, 10, 6)
    rng = np.random.default_rng(11 + inject_type)
    rho0 = rng.uniform(0.5, 1.5, shape)
    inc, tau, pos, thr = 0.02, 0.8, 0.45, 1.0
    P = oracle.Params(*shape, G=0.0, tau=tau, rho_inject_period=3, rho_increment=inc, inject_position=pos,
                      rhol_spinodal=thr)
    w = oracle.tables()[5]
    streamed = sum(w[i] * np.roll(rho0, shift=(EX[i], EY[i], EZ[i]), axis=(0, 1, 2)) for i in range(19))
    y = np.arange(shape[1])[None, :, None]
    elig = np.broadcast_to(y < pos * shape[1], shape)
    if inject_type == 1:
        elig = elig & (streamed < thr)
    elif inject_type == 2:
        elig = elig & (streamed > thr)
    assert 0 < elig.sum() < elig.size
    for step, fires in ((6, True), (7, False)):
        st = oracle.State(P, np.ones(shape, bool), rho0, dtype=np.float64)
        st.step = step
        st.run(1, reverse=reverse, inject_type=inject_type)
        gain = st.fA.sum(-1) - streamed
        want = np.where(elig, reverse * inc / tau, 0.0) if fires else np.zeros(shape)
        assert np.allclose(gain, want, rtol=0, atol=1e-14), (step, inject_type)
        assert np.allclose(st.rho, streamed + (np.where(elig, reverse * inc, 0.0) if fires else 0.0), atol=1e-14)


# --- solid psi (iteration_kernel.cuh:21-23, iteration_device.cuh:89-147) ----

def test_solid_neighbours_carry_psi_of_rhos():
    """Fluid at rest with density rho_f next to solids: with rhos = rho_f the
    psi field is uniform and F = 0 on every fluid node; with rhos != rho_f the
    force on each fluid node is -G psi(x) sum_i w_i psi(x + e_i) e_i with
    psi = 1 - exp(-rhos) on solid neighbours, formed here by brute force."""
    shape = (9, 8, 7)
    fluid = np.ones(shape, bool)
    fluid[3:6, 2:5, 2:4] = False
    fluid[0, 7, 6] = False                       # a solid across the periodic corner
    rf, G = 0.9, -5.0
    w = oracle.tables()[5]
    st = oracle.State(oracle.Params(*shape, G=G, rhos=rf), fluid, np.full(shape, rf), dtype=np.float64)
    st.run(1)
    assert np.abs(st.F[fluid]).max() < 1e-15
    for rhos in (0.2, 1.7):
        st = oracle.State(oracle.Params(*shape, G=G, rhos=rhos), fluid, np.full(shape, rf), dtype=np.float64)
        st.run(1)
        psi = np.where(fluid, 1 - np.exp(-st.rho), 1 - math.exp(-rhos))
        F = np.zeros(shape + (3,))
        for i in range(1, 19):
            nb = np.roll(psi, shift=(-EX[i], -EY[i], -EZ[i]), axis=(0, 1, 2))   # psi(x + e_i)
            F += (w[i] * nb)[..., None] * np.array([EX[i], EY[i], EZ[i]], float)
        F *= (-G * psi)[..., None]
        assert np.abs(F[fluid]).max() > 1e-2
        assert np.allclose(st.F[fluid], F[fluid], rtol=0, atol=1e-13)
        # solid psi enters with the sign of rhos - rho_f: the force on a
        # fluid node next to a solid points towards it when rhos > rho_f
        x, y, z = 2, 3, 2                         # fluid, solid at x + 1
        assert np.sign(st.F[x, y, z, 0]) == np.sign(rhos - rf)
