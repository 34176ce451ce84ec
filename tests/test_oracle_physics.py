"""Pins the oracle's stage RHS and time step (restatement of
ActFlow_V3-2.cu:718-821, 940-1243) against closed forms, invariants and
scalar recurrences, each chosen so that a plausible transcription mistake (a
dropped term, a wrong sign, a wrong normalisation, textbook Lawson instead of
the reference's stage inputs) fails one of them."""
import numpy as np
import pytest

from conftest import rel_l2

TWO_PI = 2 * np.pi


def cos_mode(N, n, amp):
    """Half spectrum of amp*cos(n.x) on an N x N grid, box 2pi (k = n)."""
    w = np.zeros((N, N // 2 + 1), dtype=complex)
    nx, ny = n
    if nx < 0:
        nx, ny = -nx, -ny
    w[ny % N, nx] += amp * N * N / 2
    if nx == 0:
        w[(-ny) % N, 0] += amp * N * N / 2
    return w


def spec_at(w, n):
    N = w.shape[0]
    nx, ny = n
    if nx < 0:
        nx, ny = -nx, -ny
        return np.conj(w[ny % N, nx])
    return w[ny % N, nx]


def test_energy_biot_savart_closed_form(orc):
    """omega = A cos(n.x) -> u = (n_y, -n_x)/|n|^2 A sin(n.x)... energy
    1/2<|u|^2> = A^2/(4|n|^2); with a mean M: + |M/N|^2/2 (.cu:1118-1127)."""
    N, A = 32, 0.7
    o = orc.Oracle(N, lam=0.0, beta=0.0)
    for n in [(1, 0), (0, 3), (2, 5), (4, -3)]:
        _, g, e = o.rhs(cos_mode(N, n, A))
        assert np.isclose(e, A * A / (4 * (n[0] ** 2 + n[1] ** 2)), rtol=1e-13)
    M = np.array([0.3, -0.4]) * N * N
    _, g, e = o.rhs(cos_mode(N, (2, 1), A), M)
    assert np.isclose(e, A * A / (4 * 5) + 0.5 * 0.25, rtol=1e-13)


def test_advection_two_mode_closed_form(orc):
    """omega = a cos(p.x) + b cos(q.x): -lam u.grad(omega) =
    -lam C/2 [cos((p-q).x) - cos((p+q).x)], C = ab(p_y q_x - p_x q_y)(1/|p|^2 - 1/|q|^2).
    Pins the Biot-Savart signs (.cu:1176-1180), the products (.cu:1192-1205)
    and the kx/ky assembly of nonl_subfunc_2 (.cu:1221-1223)."""
    N, lam = 32, 1.7
    o = orc.Oracle(N, lam=lam, beta=0.0, dealias=2)
    p, q, a, b = (1, 2), (3, -1), 0.8, -0.5
    w = cos_mode(N, p, a) + cos_mode(N, q, b)
    nl, g, _ = o.rhs(w)
    C = a * b * (p[1] * q[0] - p[0] * q[1]) * (1 / (p[0] ** 2 + p[1] ** 2) - 1 / (q[0] ** 2 + q[1] ** 2))
    dm = (p[0] - q[0], p[1] - q[1])
    sm = (p[0] + q[0], p[1] + q[1])
    ref = cos_mode(N, dm, -lam * C / 2) + cos_mode(N, sm, lam * C / 2)
    assert np.abs(nl - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.allclose(g, 0, atol=1e-12)


def test_shell_field_has_no_self_advection(orc):
    """A field on one |k| shell has u.grad(omega) = 0."""
    N = 32
    o = orc.Oracle(N, lam=3.0, beta=0.0)
    w = cos_mode(N, (3, 4), 0.9) + cos_mode(N, (5, 0), -0.4) + cos_mode(N, (0, 5), 0.2)
    nl, _, _ = o.rhs(w)
    assert np.abs(nl).max() < 1e-12 * np.abs(w).max()


@pytest.mark.parametrize("n,dealias,keep3", [(2, 2, True), (3, 0, True), (4, 0, False)])
def test_cubic_shear_closed_form(orc, n, dealias, keep3):
    """u = (A sin(ny), 0): |u|^2 u = A^3 (3 sin ny - sin 3ny)/4 x, so
    -beta z.curl(C) = 3 n beta A^3/4 (cos ny - cos 3ny); the 3n harmonic is kept
    or cut by the isotropic 2/3 mask (N=32: cut above |k| = 10.67)."""
    N, A, beta = 32, 0.9, 0.37
    o = orc.Oracle(N, lam=2.0, beta=beta, dealias=dealias)
    w = cos_mode(N, (0, n), -A * n)          # omega = -A n cos(ny)
    nl, g, e = o.rhs(w)
    ref = cos_mode(N, (0, n), 3 * n * beta * A ** 3 / 4)
    if keep3:
        ref = ref + cos_mode(N, (0, 3 * n), -3 * n * beta * A ** 3 / 4)
    assert np.abs(nl - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.isclose(e, A * A / 4, rtol=1e-13)
    assert np.allclose(g, 0, atol=1e-10 * N * N)


def test_cubic_shear_aliases_pseudospectrally(orc):
    """The cubic product is formed on the N-point grid (pseudo-spectral,
    .cu:1192-1205): at N=32, n=8 the 3n=24 harmonic folds onto -8, so
    |u|^2 u = A^3 sin(8y) and -beta z.curl(C) = 8 beta A^3 cos(8y)."""
    N, A, beta, n = 32, 0.9, 0.37, 8
    o = orc.Oracle(N, lam=2.0, beta=beta, dealias=0)
    nl, _, _ = o.rhs(cos_mode(N, (0, n), -A * n))
    ref = cos_mode(N, (0, n), n * beta * A ** 3)
    assert np.abs(nl - ref).max() < 1e-12 * np.abs(ref).max()


def test_mean_rhs_uniform_flow(orc):
    """omega = 0, M != 0: u = M/N uniform, g = -beta |M|^2 M / N^2
    (.cu:1086-1087 with beta' = beta/N^3, .cu:434)."""
    N, beta = 32, 0.21
    o = orc.Oracle(N, lam=1.0, beta=beta)
    M = np.array([0.6, -1.1]) * N * N
    nl, g, _ = o.rhs(np.zeros((N, N // 2 + 1), complex), M)
    assert np.abs(nl).max() < 1e-9
    assert np.allclose(g, -beta * np.dot(M, M) * M / (N * N) ** 2, rtol=1e-12)


def test_linear_exactness(orc):
    """lambda = beta = 0 -> omega_n = E^n omega_0 (N-hat = 0, only the
    integrating factor acts)."""
    N, dt = 32, 1e-2
    alpha = -0.8
    o = orc.Oracle(N, lam=0.0, beta=0.0, dt=dt, alpha=alpha)
    w0 = cos_mode(N, (1, 2), 1.0) + cos_mode(N, (0, 3), -0.5) + cos_mode(N, (4, -1), 0.25)
    o.set_state(w0)
    o.step(7)
    w, M = o.get_state()
    k2 = np.array([5.0, 9.0, 17.0])
    sig = -(1 - k2) ** 2 - alpha
    for n, s in zip([(1, 2), (0, 3), (4, -1)], sig):
        assert np.isclose(spec_at(w, n), spec_at(w0, n) * np.exp(7 * s * dt), rtol=1e-13)


def mean_recurrence(M0, beta, N, dt, e, e2, steps, lawson=False):
    """Appendix-A scalar recurrence of the mean mode (.cu:729-821); F2: no
    e*M_old term unless lawson."""
    f = lambda m: -beta * np.dot(m, m) * m / (N * N) ** 2
    M = np.array(M0, float)
    for _ in range(steps):
        Mo = M.copy()
        M = e * Mo if lawson else Mo.copy()
        g1 = f(Mo)
        M = M + dt / 6 * e * g1
        m = e2 * (Mo + dt / 2 * g1)
        g2 = f(m)
        M = M + dt / 3 * e2 * g2
        m = e2 * (Mo + dt / 2 * g2)
        g3 = f(m)
        M = M + dt / 3 * e2 * g3
        m = e * (Mo + dt * g3)
        g4 = f(m)
        M = M + dt / 6 * g4
    return M


@pytest.mark.parametrize("lawson", [False, True])
def test_mean_ode_recurrence(orc, lawson):
    N, beta, dt, alpha = 32, 0.05, 1e-2, -0.8
    c0 = -1 - alpha
    o = orc.Oracle(N, lam=2.0, beta=beta, dt=dt, alpha=alpha, mean_scheme=1 if lawson else 0)
    M0 = np.array([0.9, -0.3]) * N * N
    o.set_state(np.zeros((N, N // 2 + 1), complex), M0)
    o.step(20)
    w, M = o.get_state()
    ref = mean_recurrence(M0, beta, N, dt, np.exp(c0 * dt), np.exp(c0 * dt / 2), 20, lawson)
    assert np.allclose(M, ref, rtol=1e-13)
    assert np.abs(w).max() < 1e-9 * N * N


def test_mean_frozen_without_beta(orc):
    """F2: with beta = 0 the reference scheme keeps M constant whatever
    alpha; the lawson option gives exp(c0 t) M."""
    N, dt, alpha = 32, 1e-2, -0.8
    M0 = np.array([1.0, 2.0]) * N * N
    for lawson, expect in [(0, M0), (1, M0 * np.exp((-1 - alpha) * dt * 10))]:
        o = orc.Oracle(N, lam=1.0, beta=0.0, dt=dt, alpha=alpha, mean_scheme=lawson)
        o.set_state(np.zeros((N, N // 2 + 1), complex), M0)
        o.step(10)
        assert np.allclose(o.get_state()[1], expect, rtol=1e-12)


def _advect_errors(orc, alpha, dts, N=32, lam=1.3, T=0.2):
    U = np.array([0.8, -0.5])
    n = (2, 3)
    k2 = 13.0
    sig = -(1 - k2) ** 2 - alpha
    errs = []
    for dt in dts:
        o = orc.Oracle(N, lam=lam, beta=0.0, dt=dt, alpha=alpha)
        w0 = cos_mode(N, n, 1.0)
        o.set_state(w0, U * N * N)
        o.step(int(round(T / dt)))
        w, M = o.get_state()
        assert np.allclose(M, U * N * N, rtol=1e-14)     # beta = 0: M frozen (F2)
        ref = spec_at(w0, n) * np.exp((sig - 1j * lam * (n[0] * U[0] + n[1] * U[1])) * T)
        errs.append(abs(spec_at(w, n) - ref) / abs(ref))
    return errs


def test_uniform_advection(orc):
    """beta = 0, single-shell omega, mean flow U: omega^(t) =
    exp((sigma - i lam k.U) t) omega^(0).  With sigma(0) = 0 (alpha = -1) the
    stage inputs of the mean are exactly M and the error falls ~4x per halving
    of dt (second order, F1)."""
    e = _advect_errors(orc, -1.0, (1e-3, 5e-4))
    assert e[0] < 1e-4
    assert 3.0 < e[0] / e[1] < 5.0


def test_mean_stage_inputs_scaled_by_if(orc):
    """With sigma(0) != 0 the reference feeds the stages e2*M_o, e2*M_o, e*M_o
    (.cu:758-759, 780-781, 801-802) while M itself is not multiplied by e
    (F2), so advection by the mean is first order: the error halves with dt."""
    e = _advect_errors(orc, -0.8, (1e-3, 5e-4))
    assert 1.7 < e[0] / e[1] < 2.3


def test_temporal_order_is_two(orc):
    """The reference's stage inputs (E2(w + dt/2 N2), E(w + dt N3)) make the
    scheme second order when sigma != 0 (Lawson would be fourth)."""
    N, T = 32, 0.04
    rng = np.random.default_rng(4)
    w0 = np.zeros((N, N // 2 + 1), complex)
    for n in [(1, 1), (2, -1), (0, 2), (3, 1), (1, -3)]:
        w0 += cos_mode(N, n, rng.uniform(0.5, 1.5))
    kw = dict(lam=3.0, beta=0.5, alpha=-4.0)
    sols = {}
    for dt in (4e-3, 2e-3, 1e-3, 1.25e-4):
        o = orc.Oracle(N, dt=dt, **kw)
        o.set_state(w0)   # even omega: <|u|^2 u> = 0, M stays 0
        o.step(int(round(T / dt)))
        sols[dt] = o.get_state()[0]
    ref = sols[1.25e-4]
    e1, e2, e3 = (rel_l2(sols[d], ref) for d in (4e-3, 2e-3, 1e-3))
    p1, p2 = np.log2(e1 / e2), np.log2(e2 / e3)
    assert 1.6 < p1 < 2.5 and 1.6 < p2 < 2.5, (p1, p2)


def test_projection_restores_hermitian_column(orc):
    """reality_func (.cu:1237-1242): after a step the kx=0 column of omega is
    Hermitian for rows 1..Ny/2-1 up to the rounding of the last accumulate."""
    N = 32
    o = orc.Oracle(N, lam=3.0, beta=0.1)
    w0 = cos_mode(N, (0, 2), 1.0) + cos_mode(N, (1, 1), 0.5) + cos_mode(N, (2, -3), 0.3)
    w0[2, 0] += 0.3 * N * N            # break the symmetry of row 2 / row N-2
    o.set_state(w0)
    o.step(1)
    w, _ = o.get_state()
    col = w[:, 0]
    asym = np.abs(col[1:N // 2] - np.conj(col[N - 1:N // 2:-1])).max()
    assert asym < 1e-10 * np.abs(col).max()


def broken_column_state(nx, ny, seed=3, amp=0.5, kmax=6):
    """A state whose kx = 0 column is NOT Hermitian (rows y and Ny - y are
    independent), the rest band-limited (|k| <= kmax), the kx = Nx/2 column
    zero and the k = 0 mode zero."""
    rng = np.random.default_rng(seed)
    y = np.arange(ny)[:, None]
    x = np.arange(nx // 2 + 1)[None, :]
    yy = np.minimum(y, ny - y)
    band = (x * x + yy * yy <= kmax * kmax) & (x < nx // 2)
    band[0, 0] = False
    noise = rng.standard_normal((ny, nx // 2 + 1)) + 1j * rng.standard_normal((ny, nx // 2 + 1))
    w = np.where(band, amp * nx * ny * noise, 0.0).astype(complex)
    # kx = 0: rows y and Ny - y drawn independently above; the Nyquist row
    # (left alone by the projection) gets a value of its own
    w[ny // 2, 0] = 0.25 * nx * ny * (1.0 - 2.0j)
    return w


def projection_closed_form(w0, E):
    """One step with lambda = beta = 0 from w0: every stage's N-hat is 0, so
    the step is w1 = P(E w0), P the projection of reality_func
    (.cu:1237-1242): for y = 1..Ny/2-1, a = (x[y] + conj x[Ny-y]) / 2,
    x[y] = a, x[Ny-y] = conj a; row 0, row Ny/2 and every kx > 0 are E w0."""
    ny = w0.shape[0]
    x = E * w0
    out = x.copy()
    for y in range(1, ny // 2):
        a = 0.5 * (x[y, 0] + np.conj(x[ny - y, 0]))
        out[y, 0] = a
        out[ny - y, 0] = np.conj(a)
    return out


@pytest.mark.parametrize("nx,ny", [(32, 32), (32, 64)])
def test_projection_value_closed_form(orc, nx, ny):
    """The projection's value, not only its symmetry: from a kx = 0 column
    with broken symmetry and lambda = beta = 0, one step gives exactly
    P(E w0).  A copy instead of an average, a wrong partner row, or a
    projected Nyquist row each fail here."""
    dt, alpha = 1e-2, -0.8
    o = orc.Oracle(nx, ny, lam=0.0, beta=0.0, dt=dt, alpha=alpha, dealias=2)
    w0 = broken_column_state(nx, ny)
    o.set_state(w0, (0.0, 0.0))
    o.step(1)
    w, _ = o.get_state()
    kx = np.arange(nx // 2 + 1)[None, :]
    ky = np.fft.fftfreq(ny, 1.0 / ny)[:, None]
    k2 = kx ** 2 + ky ** 2
    E = np.exp((-(1 - k2) ** 2 - alpha) * dt)
    ref = projection_closed_form(w0, E)
    scale = np.abs(ref).max()
    assert np.abs(w - ref).max() <= 1e-14 * scale
    # the symmetry was really broken: without the projection the column differs
    assert np.abs(E * w0 - ref)[:, 0].max() > 1e-3 * scale
    assert w[ny // 2, 0] == E[ny // 2, 0] * w0[ny // 2, 0]


def _mean_closed_form_error(orc, dt, T=2.0, N=32, beta=0.5, scheme=0):
    """sigma(0) = 0 (alpha = -1 in the reference operator: c0 = -1 - alpha = 0),
    omega = 0: the mean obeys dM/dt = -beta |M|^2 M / N^4 (N^4 = (Nx Ny)^2,
    beta' = beta/N^3 of .cu:434 times the N of the k=0 DFT), whose solution is
    |M(t)|^2 = |M0|^2 / (1 + 2 beta |M0|^2 t / N^4) along M0.  Returns the
    relative error of the oracle's M at T."""
    o = orc.Oracle(N, lam=2.0, beta=beta, dt=dt, alpha=-1.0, mean_scheme=scheme)
    M0 = np.array([0.9, -0.3]) * N * N
    o.set_state(np.zeros((N, N // 2 + 1), complex), M0)
    o.step(int(round(T / dt)))
    w, M = o.get_state()
    assert np.abs(w).max() < 1e-9 * N * N
    m02 = np.dot(M0, M0)
    ref = M0 * np.sqrt(1.0 / (1.0 + 2.0 * beta * m02 * T / (N * N) ** 2))
    return np.linalg.norm(M - ref) / np.linalg.norm(ref)


@pytest.mark.parametrize("scheme", [0, 1])
def test_mean_ode_is_classical_rk4(orc, scheme):
    """With sigma(0) = 0 the reference's mean update (.cu:729-732, 756-757,
    778-779, 799-800, 820-821) is classical RK4 with the stage inputs of
    .cu:758-759, 780-781, 801-802: the error against the closed form falls
    ~16x per halving of dt (measured 14.4 and 15.3: the cubic's error
    constant reaches its asymptote slowly, so the steps are small).  A wrong
    accumulate weight (e.g. dt/3 for dt/6) leaves an O(dt) error that fails
    the bound; a wrong stage input drops the order to 1 or 2 and fails the
    ratio."""
    e = [_mean_closed_form_error(orc, dt, scheme=scheme) for dt in (0.05, 0.025, 0.0125)]
    assert e[0] < 1e-8, e
    r1, r2 = e[0] / e[1], e[1] / e[2]
    assert 13.0 < r1 < 19.0 and 13.0 < r2 < 19.0, (e, r1, r2)


def test_invariants_w0_and_mask(orc):
    """omega^[0] stays 0 and masked-out modes never gain content (T12)."""
    N = 32
    o = orc.Oracle(N, lam=3.5, beta=0.5)
    rng = np.random.default_rng(1)
    w0 = np.zeros((N, N // 2 + 1), complex)
    for n in [(1, 1), (2, -1), (0, 2), (3, 1), (4, 4)]:
        w0 += cos_mode(N, n, rng.uniform(0.5, 1.5))
    o.set_state(w0, (0.2 * N * N, 0.0))
    o.step(5)
    w, _ = o.get_state()
    assert w[0, 0] == 0
    mask = o.mask()
    assert np.abs(w[mask == 0]).max() == 0.0


def _count_iso(nx, ny):
    cnt = 0
    for y in range(ny):
        yy = min(y, ny - y)
        for x in range(nx // 2 + 1):
            if 9 * (x * x * ny * ny + yy * yy * nx * nx) <= nx * nx * ny * ny:
                cnt += 1
    return cnt


def test_masks_bruteforce(orc):
    for N in (32, 64):
        o = orc.Oracle(N, dealias=0)
        assert int(o.mask().sum()) == _count_iso(N, N)
        o = orc.Oracle(N, dealias=1)
        m = o.mask()
        assert int(m.sum()) == (N // 3 + 1) * (2 * (N // 3) + 1)
        assert int(orc.Oracle(N, dealias=2).mask().sum()) == N * (N // 2 + 1)
    # reference-literal cutoff: everything at 256^2, only k=0 from 512^2 on (F3)
    assert int(orc.Oracle(256, dealias=3).mask().sum()) == 256 * 129
    o = orc.Oracle(512, dealias=3)
    m = o.mask()
    assert int(m.sum()) == 1 and m[0, 0] == 1


def test_float32_oracle_tracks_float64(orc):
    N = 64
    rng = np.random.default_rng(2)
    w0 = np.zeros((N, N // 2 + 1), complex)
    for n in [(1, 1), (2, -1), (0, 2), (3, 1), (5, 2)]:
        w0 += cos_mode(N, n, rng.uniform(0.5, 1.5))
    res = {}
    for prec in (64, 32):
        o = orc.Oracle(N, lam=3.5, beta=0.1, dt=1e-3, precision=prec)
        o.set_state(w0)
        o.step(20)
        res[prec] = o.get_state()[0]
    assert rel_l2(res[32], res[64]) < 1e-5
