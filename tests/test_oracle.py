"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests or golden vectors for the solver (its only
runtime checks are lbm.py:420 and app.py:139-147), so each oracle function is
checked here against closed forms, invariants and known answers chosen so a
plausible slip (a dropped term, a wrong sign or index, a swapped weight)
fails at least one of them.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

C, W, OPP = O.C, O.W, O.OPP


def equilibrium(rho, u):
    """Second-order Hermite equilibrium (textbook D3Q19, c_s^2 = 1/3)."""
    cu = C @ np.asarray(u, dtype=np.float64)
    usq = float(np.dot(u, u))
    return W * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * usq)


# ---- lattice tables (lbm.py:15-19) ------------------------------------------------
def test_lattice_isotropy():
    assert C.shape == (19, 3) and abs(W.sum() - 1) < 1e-15
    assert np.allclose(W @ C, 0, atol=1e-16)
    assert np.allclose((W[:, None, None] * C[:, :, None] * C[:, None, :]).sum(0), np.eye(3) / 3, atol=1e-16)
    # fourth-order isotropy of D3Q19: sum w c_a^4 = 1/3, sum w c_a^2 c_b^2 = 1/9
    assert abs((W * C[:, 0] ** 4).sum() - 1 / 3) < 1e-15
    assert abs((W * C[:, 0] ** 2 * C[:, 1] ** 2).sum() - 1 / 9) < 1e-15
    for i in range(19):
        assert np.array_equal(C[OPP[i]], -C[i])
    assert sorted(map(tuple, C.tolist())) == sorted(
        {(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if abs(a) + abs(b) + abs(c) <= 2})


# ---- init (lbm.py:64-69) -----------------------------------------------------------
@pytest.mark.parametrize('form', [O.FP32, O.FP64])
def test_init_is_equilibrium(form):
    t = O.OracleTunnel((6, 5, 4), form)
    t.init(0.05)
    f = t.populations()
    tol = 1e-7 if form == O.FP32 else 1e-15
    u32 = float(np.float32(0.05))   # kernel parameters are float32, as Warp's are
    assert np.allclose(f, equilibrium(1.0, [u32, 0, 0])[:, None], atol=tol, rtol=0)
    assert np.allclose(t.mac, [0.05, 0, 0, 1], atol=1e-8)


# ---- fixed points ---------------------------------------------------------------------
def test_uniform_flow_is_a_fixed_point():
    """No solids, uniform inlet flow: every cell (faces included) stays at its equilibrium."""
    t = O.OracleTunnel((32, 10, 12), O.FP64)
    t.init(0.07)
    f0 = t.populations().copy()
    t.step(20, 0.07, 0.52, 0.17)
    assert np.abs(t.populations() - f0).max() < 1e-14


def test_rest_with_solids_is_a_fixed_point_and_force_free():
    """At rest, bounce-back on any body returns w_i; the momentum exchange is exactly zero."""
    dims = (20, 12, 12)
    rng = np.random.default_rng(3)
    mask = (rng.random(np.prod(dims)) < 0.15).astype(np.uint8)
    t = O.OracleTunnel(dims, O.FP64)
    t.init(0.0)
    t.set_solid(mask)
    t.step(5, 0.0, 0.55, 0.17)
    f = t.populations()
    fluid = mask == 0
    assert np.abs(f[:, fluid] - W[:, None]).max() < 1e-15
    assert np.abs(t.forces()).max() < 1e-15


# ---- collision invariants (lbm.py:125-165) ----------------------------------------------
def _perturbed_state(dims, seed, amp=0.01, u=(0.04, 0.01, -0.02)):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    eq = equilibrium(1.0, u)
    f = eq[:, None] * (1 + amp * rng.standard_normal((19, n)))
    return f


def _run_one(dims, f_in, mask=None, uin=0.05, tau=0.52, smag=0.17, form=O.FP64):
    t = O.OracleTunnel(dims, form)
    t.init(uin)
    if mask is not None:
        t.set_solid(mask)
    t.f[0][:] = f_in.reshape(-1).astype(O.ELEM[form])
    t.step(1, uin, tau, smag)
    return t


def _interior(dims, x_lo=1, x_hi=None):
    nx, ny, nz = dims
    x_hi = nx - 2 if x_hi is None else x_hi
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing='ij')
    m = (x >= x_lo) & (x <= x_hi) & (y >= 1) & (y <= ny - 2) & (z >= 1) & (z <= nz - 2)
    return m.reshape(-1)


def _streamed(f, dims):
    """Pull streaming without solids: f_i(x) <- f_i(x - c_i) (numpy, interior cells)."""
    nx, ny, nz = dims
    g = f.reshape(19, nz, ny, nx)
    out = np.empty_like(g)
    for i in range(19):
        out[i] = np.roll(g[i], shift=(C[i, 2], C[i, 1], C[i, 0]), axis=(0, 1, 2))
    return out.reshape(19, -1)


def test_collision_conserves_mass_and_momentum():
    dims = (16, 9, 9)
    f_in = _perturbed_state(dims, 1)
    t = _run_one(dims, f_in, tau=0.53, smag=0.17)
    f_out = t.populations()
    inner = _interior(dims)
    pre = _streamed(f_in, dims)[:, inner]
    rho = pre.sum(0)
    mom = C.T @ pre
    assert np.abs(f_out[:, inner].sum(0) - rho).max() < 1e-14
    assert np.abs(C.T @ f_out[:, inner] - mom).max() < 1e-14
    # macroscopic output = moments of the streamed populations (clamp inactive here)
    assert np.abs(t.mac[inner, 3] - rho).max() < 1e-14
    assert np.abs(t.mac[inner, :3] - (mom / rho).T).max() < 1e-14


def test_regularization_and_smagorinsky_closure():
    """f_out - e is the Hermite projection of keep * Pi_neq, and tau_e obeys the
    Smagorinsky closure tau_e (tau_e - tau0) = (9 sqrt2 / 2) Cs^2 Q / rho (Hou et al. 1996)."""
    dims = (40, 9, 9)   # x < nx - 24 keeps the sponge off for x <= 15
    f_in = _perturbed_state(dims, 2, amp=0.02)
    tau, cs = float(np.float32(0.52)), float(np.float32(0.17))   # float32 kernel parameters
    t = _run_one(dims, f_in, tau=tau, smag=cs)
    inner = _interior(dims, 1, 15)
    pre = _streamed(f_in, dims)[:, inner]
    rho = pre.sum(0)
    u = (C.T @ pre) / rho
    usq = (u * u).sum(0)
    cu = C @ u
    e = W[:, None] * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * usq)
    d = pre - e
    pi = np.einsum('ia,ib,in->abn', C, C, d)
    q = np.sqrt(np.einsum('abn,abn->n', pi, pi))
    dout = t.populations()[:, inner] - e
    pi_out = np.einsum('ia,ib,in->abn', C, C, dout)
    keep = (pi_out * pi).sum((0, 1)) / (pi * pi).sum((0, 1))
    assert np.abs(pi_out - keep * pi).max() < 1e-13            # stress kept along Pi only
    # Hermite form: dout_i = 4.5 w_i keep (c_i c_i - I/3) : Pi
    H = np.einsum('ia,ib->iab', C, C) - np.eye(3) / 3
    pred = 4.5 * W[:, None] * keep * np.einsum('iab,abn->in', H, pi)
    assert np.abs(dout - pred).max() < 1e-14
    tau_e = 1 / (1 - keep)
    lhs = tau_e * (tau_e - tau)
    rhs = 4.5 * math.sqrt(2) * cs * cs * q / rho
    assert np.allclose(lhs, rhs, rtol=1e-7, atol=1e-15)       # 25.4558441 ~= 18 sqrt(2) to 1e-9
    # Cs = 0 gives tau_e = tau exactly
    t0 = _run_one(dims, f_in, tau=tau, smag=0.0)
    dout0 = t0.populations()[:, inner] - e
    pi0 = np.einsum('ia,ib,in->abn', C, C, dout0)
    assert np.abs(pi0 - (1 - 1 / tau) * pi).max() < 1e-14


def test_outlet_sponge_raises_tau():
    """tau0 = tau + 0.015 (x - nx + 24) for x > nx - 24 (lbm.py:148-151)."""
    dims = (40, 9, 9)
    f_in = _perturbed_state(dims, 4, amp=0.02)
    t = _run_one(dims, f_in, tau=0.52, smag=0.0)
    nx = dims[0]
    pre = _streamed(f_in, dims)
    fo = t.populations()
    for x in (20, 30, 37):
        idx = x + nx * (4 + 9 * 4)
        p = pre[:, idx]
        rho = p.sum(); u = (C.T @ p) / rho
        e = equilibrium(rho, u)
        pi = np.einsum('ia,ib,i->ab', C, C, p - e)
        pio = np.einsum('ia,ib,i->ab', C, C, fo[:, idx] - e)
        tau0 = float(np.float32(0.52)) + max(0.0, (x - nx + 24) * 0.015)
        assert np.allclose(pio, (1 - 1 / tau0) * pi, atol=1e-14)


def test_velocity_clamp():
    """|u| > 0.35 is scaled back to 0.35 (lbm.py:130-135), and the oracle counts
    every clamped cell."""
    dims = (12, 7, 7)
    f = np.tile(equilibrium(1.0, [0.3, 0.3, 0.0])[:, None], (1, int(np.prod(dims))))
    O.clamp_hits()
    t = _run_one(dims, f, uin=0.05)
    inner = _interior(dims)
    u = t.mac[inner, :3]
    assert np.allclose(np.linalg.norm(u, axis=1), 0.35, atol=1e-14)
    assert np.allclose(u[:, 0], u[:, 1], atol=1e-15)
    assert O.clamp_hits() == (int(inner.sum()), 0)


def test_face_density_clamp_counts():
    """A face cell whose interior neighbour had rho outside [0.9, 1.1] in the previous
    step takes the clamped density (lbm.py:88); inlet and outlet faces set rho = 1
    and do not count."""
    dims = (14, 8, 9)
    nx, ny, nz = dims
    t = O.OracleTunnel(dims, O.FP64)
    t.init(0.05)
    # previous-step field: rho 1.3 on the layer y = ny - 2 and 0.8 on z = 1
    m = t.mac.reshape(nz, ny, nx, 4)
    m[:, ny - 2, :, 3] = 1.3
    m[1, :, :, 3] = 0.8
    O.clamp_hits()
    t.step(1, 0.05, 0.52, 0.17)
    vel, den = O.clamp_hits()
    mo = t.mac.reshape(nz, ny, nx, 4)
    # y = ny - 1 faces (x inside 1..nx-2) read 1.3 -> 1.1; z = 0 faces read the z = 1 layer -> 0.9
    top = mo[2:nz - 1, ny - 1, 1:nx - 1, 3]
    assert np.all(top == 1.1)
    bottom = mo[0, 1:ny - 2, 1:nx - 1, 3]
    assert np.all(bottom == 0.9)
    # the count: every face cell off the inlet and outlet whose clamped neighbour lies
    # on either modified layer
    z_, y_, x_ = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing='ij')
    face = (x_ == 0) | (x_ == nx - 1) | (y_ == 0) | (y_ == ny - 1) | (z_ == 0) | (z_ == nz - 1)
    cy, cz = np.clip(y_, 1, ny - 2), np.clip(z_, 1, nz - 2)
    prev_rho = np.where(cy == ny - 2, 1.3, np.where(cz == 1, 0.8, 1.0))
    # where the two layers cross the field holds 0.8 (written last); both values clamp
    expect = face & (x_ > 0) & (x_ < nx - 1) & ((prev_rho < 0.9) | (prev_rho > 1.1))
    assert vel == 0 and den == int(expect.sum()) > 0


def test_oracle_mass_is_the_sum_of_the_stored_populations():
    """The mass the oracle step reports (FP64, faces included, solids excluded) equals
    sum over non-solid cells of sum_i f_i of what the step stored: collision conserves
    mass and a face cell's equilibrium sums to its boundary density. Counting solid
    cells (2,109 of them at rho ~ 1) or dropping faces would be off by > 1e-3."""
    from windtunnel import geometry
    dims = (48, 20, 22)
    mask, _ = geometry.sphere(dims, (14, 10, 11), 8)
    t = O.OracleTunnel(dims, O.FP64)
    t.init(0.05)
    t.set_solid(mask)
    t.step(7, 0.05, 0.52, 0.17)
    fluid = mask == 0
    total = t.populations()[:, fluid].sum()
    assert abs(t.mass - total) <= 1e-12 * total
    assert abs(t.mass - t.mac[fluid, 3].sum()) <= 1e-12 * total


def test_oracle_mass_at_rest_is_the_fluid_cell_count():
    """At the rest fixed point (uin = 0) every non-solid cell has rho = 1."""
    from windtunnel import geometry
    dims = (30, 12, 14)
    mask, counts = geometry.sphere(dims, (10, 6, 7), 6)
    for form, tol in ((O.FP64, 1e-13), (O.FP32, 1e-6)):
        t = O.OracleTunnel(dims, form)
        t.init(0.0)
        t.set_solid(mask)
        t.step(3, 0.0, 0.52, 0.17)
        nfluid = int((mask == 0).sum())
        assert nfluid == int(np.prod(dims)) - counts['count']
        assert abs(t.mass - nfluid) <= tol * nfluid


# ---- boundaries ---------------------------------------------------------------------------
def test_moving_belt_momentum():
    """Rest fluid over the belt (solid==2): Ladd's moving-wall term 6 w_i c_ix U gives the
    first fluid layer x-momentum U/3 and no extra mass (lbm.py:121-124)."""
    dims = (20, 8, 10)
    nx, ny, nz = dims
    mask = np.zeros(dims[::-1], np.uint8)
    mask[:, 0, :] = 2
    t = O.OracleTunnel(dims, O.FP64)
    t.init(0.0)
    t.set_solid(mask.reshape(-1))
    U = float(np.float32(0.06))
    t.step(1, U, 0.55, 0.0)
    mac = t.mac.reshape(nz, ny, nx, 4)
    layer = mac[1:-1, 1, 1:-1]
    assert np.allclose(layer[..., 0], U / 3, atol=1e-15)
    assert np.allclose(layer[..., 3], 1.0, atol=1e-15)
    assert np.allclose(mac[1:-1, 2:-1, 1:-1, 0], 0.0, atol=1e-15)


def test_single_solid_cell_force_closed_form():
    """Every fluid cell at equilibrium(1, u) around one static solid: the 18 links give
    F = sum_i 2 (e_i - w_i) c_i = 2 u (lbm.py:185-193)."""
    dims = (9, 9, 9)
    u = np.array([0.05, -0.02, 0.01])
    n = 9 ** 3
    t = O.OracleTunnel(dims, O.FP64)
    t.set_solid(np.zeros(n, np.uint8))
    mask = np.zeros(n, np.uint8)
    mask[4 + 9 * (4 + 9 * 4)] = 1
    t.set_solid(mask)
    t.f[0][:] = np.repeat(equilibrium(1.0, u), n)
    t.cur = 0
    F = t.forces()
    assert np.allclose(F, 2 * u, atol=1e-15)
    # a belt cell (2) is excluded from the force
    mask[4 + 9 * (4 + 9 * 4)] = 2
    t.set_solid(mask)
    assert np.abs(t.forces()).max() == 0.0


def test_face_cells_use_previous_step_mac():
    """Face cells take equilibrium from the previous step's mac of the clamped
    interior cell (Jacobi rule), with the wall/inlet/outlet overrides of lbm.py:88-104."""
    dims = (10, 7, 8)
    nx, ny, nz = dims
    rng = np.random.default_rng(5)
    t = O.OracleTunnel(dims, O.FP64)
    t.init(0.05)
    prev = np.zeros((t.n, 4))
    prev[:, :3] = rng.uniform(-0.05, 0.08, (t.n, 3))
    prev[:, 3] = rng.uniform(0.85, 1.15, t.n)
    t.mac[:] = prev
    t.step(1, 0.05, 0.52, 0.17)
    f = t.populations()
    for (x, y, z) in [(nx - 1, 3, 3), (4, 0, 3), (4, ny - 1, 5), (5, 3, 0), (5, 3, nz - 1),
                      (nx - 1, ny - 1, nz - 1), (0, 3, 3), (0, 0, 0)]:
        idx = x + nx * (y + ny * z)
        tx, ty, tz = min(max(x, 1), nx - 2), min(max(y, 1), ny - 2), min(max(z, 1), nz - 2)
        m = prev[tx + nx * (ty + ny * tz)]
        rho = min(max(m[3], 0.9), 1.1)
        u = m[:3].copy()
        if y in (0, ny - 1):
            u[1] = 0
        if z in (0, nz - 1):
            u[2] = 0
        if x == 0:
            rho, u = 1.0, np.array([float(np.float32(0.05)), 0, 0])
        elif x == nx - 1:
            rho, u[0] = 1.0, max(u[0], 0)
        assert np.allclose(f[:, idx], equilibrium(rho, u), atol=1e-15), (x, y, z)
        assert np.allclose(t.mac[idx], [*u, rho], atol=1e-15)


def test_fp32_oracle_tracks_fp64():
    """The FP32 restatement stays within FP32 rounding of the FP64 one over 50 steps
    of flow past a sphere (a wrong FP32-only constant would show up here)."""
    from windtunnel import geometry
    dims = (48, 20, 20)
    mask, _ = geometry.sphere(dims, (14, 10, 10), 8)
    r = []
    for form in (O.FP32, O.FP64):
        t = O.OracleTunnel(dims, form)
        t.init(0.05)
        t.set_solid(mask)
        t.step(50, 0.05, 0.52, 0.17)
        r.append((t.mac.astype(np.float64), t.forces()))
    assert np.abs(r[0][0] - r[1][0]).max() < 2e-6
    assert np.allclose(r[0][1], r[1][1], rtol=1e-4, atol=5e-6)


def _bf16_round(x):
    """float64 -> bfloat16 bits, round to nearest even (numpy has no bfloat16)."""
    b = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7fff + ((b >> 16) & 1)) >> 16).astype(np.uint16)


@pytest.mark.parametrize('form', [O.FP16, O.BF16])
def test_deviation_collision_equals_reference_collision(form):
    """The FP16/BF16 forms collide on deviations g = f - w (FP32 arithmetic, no weight
    added back): one step from a state exactly representable in the storage form
    must equal the FP64 form's reference-order step (lbm.py:110-166) on f = w + g,
    to within half a storage ulp of each output plus FP32 rounding. A dropped term,
    a sign or a wrong weight class in the deviation algebra is 1e-4 or more."""
    from windtunnel import geometry
    dims = (40, 11, 12)
    rng = np.random.default_rng(21)
    n = int(np.prod(dims))
    u = np.array([0.06, -0.02, 0.015])
    f = equilibrium(1.01, u)[:, None] * (1 + 0.01 * rng.standard_normal((19, n)))
    mask, _ = geometry.sphere(dims, (14, 5, 6), 5)
    mask = mask.copy()
    mask.reshape(dims[2], dims[1], dims[0])[:, 0, :] = 2        # a moving belt under it
    if form == O.FP16:
        graw = (f - W[:, None]).astype(np.float16)
        g = graw.astype(np.float64)
        bits = graw.view(np.uint16)
    else:
        bits = _bf16_round(f - W[:, None])
        g = (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    fexact = W[:, None] + g
    lo = O.OracleTunnel(dims, form)
    lo.init(0.05)
    lo.set_solid(mask)
    lo.f[0][:] = bits.reshape(-1)
    lo.f[1][:] = bits.reshape(-1)
    lo.step(1, 0.05, 0.52, 0.17)
    hi = O.OracleTunnel(dims, O.FP64)
    hi.init(0.05)
    hi.set_solid(mask)
    hi.f[0][:] = fexact.reshape(-1)
    hi.f[1][:] = fexact.reshape(-1)
    hi.step(1, 0.05, 0.52, 0.17)
    inner = _interior(dims) & (mask == 0)
    got = lo.populations()[:, inner] - W[:, None]          # stored deviations, exactly
    want = hi.populations()[:, inner] - W[:, None]
    if form == O.FP16:
        ulp = np.spacing(np.abs(want).astype(np.float16)).astype(np.float64)
    else:
        ulp = np.abs(want) * 2.0 ** -7
    err = np.abs(got - want)
    assert (err <= 0.5 * ulp + 3e-7).all(), err.max()
    assert np.abs(lo.mac[inner, 3] - hi.mac[inner, 3]).max() < 3e-7         # rho = 1 + sum g
    assert np.abs(lo.mac[inner, :3] - hi.mac[inner, :3]).max() < 3e-8
    # the force is 2 g on each link: exactly twice the stored deviations
    assert np.allclose(lo.forces(), hi.forces(), rtol=2e-3, atol=1e-4)


@pytest.mark.parametrize('form,tol', [(O.FP16, 2e-4), (O.BF16, 2e-3)])
def test_deviation_forms_track_fp64(form, tol):
    """Over 50 steps of flow past a sphere the FP16 (BF16) deviation forms stay
    within their storage precision of the FP64 form."""
    from windtunnel import geometry
    dims = (48, 20, 20)
    mask, _ = geometry.sphere(dims, (14, 10, 10), 8)
    r = []
    for f in (form, O.FP64):
        t = O.OracleTunnel(dims, f)
        t.init(0.05)
        t.set_solid(mask)
        t.step(50, 0.05, 0.52, 0.17)
        r.append(t.mac.astype(np.float64))
    assert np.abs(r[0] - r[1]).max() < tol


# ---- link classification (lbm.py:114-124, 185-188) -----------------------------------------
def test_link_classes_brute_force():
    dims = (9, 7, 6)
    nx, ny, nz = dims
    rng = np.random.default_rng(7)
    solid = rng.choice([0, 0, 0, 1, 2, 5], size=nx * ny * nz).astype(np.uint8)
    got = O.link_classes(solid, dims)
    s3 = solid.reshape(nz, ny, nx)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                idx = x + nx * (y + ny * z)
                v = int(got[idx])
                if s3[z, y, x]:
                    assert v == 1
                    continue
                if x in (0, nx - 1) or y in (0, ny - 1) or z in (0, nz - 1):
                    assert v == 2
                    continue
                exp = 0
                for i in range(1, 19):
                    cx, cy, cz = C[i]
                    up = int(s3[z - cz, y - cy, x - cx])
                    dn = int(s3[z + cz, y + cy, x + cx])
                    exp |= int(up != 0) << (8 + i)
                    exp |= int(up == 2 and cx != 0) << (27 + i)
                    exp |= int(dn == 1) << (45 + i)
                assert v == exp, (x, y, z)


# ---- storage-form conversions ------------------------------------------------------------------
def test_half_and_bfloat_conversions_round_to_nearest_even():
    rng = np.random.default_rng(11)
    vals = np.concatenate([np.linspace(-0.05, 0.05, 2001, dtype=np.float32),
                           (rng.standard_normal(3000) * 0.01).astype(np.float32),
                           np.float32([1e-8, -3e-7, 6.1e-5, 65504, 65519, 0.0, -0.0])])
    for v in vals:
        assert O.f2h(v) == int(np.float16(v).view(np.uint16)), v     # numpy casts are IEEE RNE
    as_f = lambda b: float(np.array([b << 16], np.uint32).view(np.float32)[0])
    for v in vals:
        bits = int(np.array([v], np.float32).view(np.uint32)[0])
        lo, hi = bits >> 16, (bits >> 16) + 1          # truncation and the next bf16 away from zero
        x = float(v)
        dl, dh = abs(as_f(lo) - x), abs(as_f(hi) - x)
        want = lo if (dl < dh or (dl == dh and lo % 2 == 0)) else hi
        assert O.f2bf(v) == want, v


# ---- particles, slices, surface pressure (lbm.py:201-368) ------------------------------------------
def test_advect_uniform_field_and_kill_rules():
    dims = (40, 20, 20)
    n = 40 * 20 * 20
    mac = np.tile(np.float32([0.05, 0.01, -0.02, 1.0]), (n, 1))
    solid = np.zeros(n, np.uint8)
    pos = np.float32([[10.3, 7.2, 9.9, 5.0], [36.98, 10, 10, 1.0], [10, 10, 10, 999.5], [0.6, 10, 10, 0]])
    spd = np.ones(4, np.float32)
    p, s = O.advect(pos, spd, mac, solid, dims, 1.0, 0.05, 1000.0, 7, (2.0, 10.0, 10.0, 4.0, 4.0))
    assert np.allclose(p[0, :3], pos[0, :3] + [0.05, 0.01, -0.02], atol=1e-6) and p[0, 3] == 6.0
    assert abs(s[0] - math.sqrt(0.05 ** 2 + 0.01 ** 2 + 0.02 ** 2) / 0.05) < 1e-5
    # x > nx - 3 kills; age > maxAge kills; respawn inside the seed box with age 0 and speed 1
    for k in (1, 2):
        assert p[k, 3] == 0 and s[k] == 1.0
        assert 2.0 <= p[k, 0] <= 4.0 and 6.0 <= p[k, 1] <= 14.0 and 6.0 <= p[k, 2] <= 14.0
    assert p[3, 3] == 1.0          # moved to x ~ 0.65 >= 0.5: alive
    # a particle inside a solid cell respawns
    solid[10 + 40 * (10 + 20 * 10)] = 1
    p2, _ = O.advect(np.float32([[9.98, 10, 10, 0]]), [1.0], mac, solid, dims, 1.0, 0.05, 1000, 7,
                     (2, 10, 10, 4, 4))
    assert p2[0, 3] == 0


def test_respawn_distribution_is_uniform():
    dims = (40, 20, 20)
    n = 40 * 20 * 20
    mac = np.tile(np.float32([0.0, 0.0, 0.0, 1.0]), (n, 1))
    N = 20000
    pos = np.zeros((N, 4), np.float32)
    pos[:, :3] = [50, 10, 10]              # outside: all respawn
    p, _ = O.advect(pos, np.ones(N, np.float32), mac, np.zeros(n, np.uint8), dims, 1.0, 0.05, 1e9, 3,
                    (2.0, 10.0, 9.0, 5.0, 3.0))
    for col, lo, hi in ((0, 2, 4), (1, 5, 15), (2, 6, 12)):
        v = (p[:, col] - lo) / (hi - lo)
        assert v.min() >= 0 and v.max() < 1
        hist = np.histogram(v, bins=10, range=(0, 1))[0]
        assert np.abs(hist - N / 10).max() < 5 * math.sqrt(N / 10)
    assert abs(np.corrcoef(p[:, 0], p[:, 1])[0, 1]) < 0.03


def test_pack_particles_known_values():
    dims = (100, 50, 40)
    pos = np.float32([[0, 0, 0, 0], [100, 50, 40, 70000], [50, 25, 20, 12.7], [-5, 60, 10, 3]])
    spd = np.float32([0, 2, 1, 5])
    q16, q8 = O.pack_particles(pos, spd, dims)
    q16 = q16.reshape(-1, 4)
    assert q16[0].tolist() == [0, 0, 0, 0]
    assert q16[1].tolist() == [65535, 65535, 65535, 65535]
    assert q16[2].tolist() == [32768, 32768, 32768, 12]
    assert q16[3].tolist() == [0, 65535, 16384, 3]
    assert q8.tolist() == [0, 255, 128, 255]


def test_slice_known_fields():
    dims = (12, 8, 6)
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing='ij')
    a = 0.004
    mac = np.zeros((nz, ny, nx, 4), np.float32)
    mac[..., 0] = a * y          # shear u_x = a y: vorticity magnitude a
    mac[..., 3] = 1.0
    mac = mac.reshape(-1, 4)
    solid = np.zeros(nx * ny * nz, np.uint8)
    solid[5 + nx * (3 + ny * 2)] = 1
    uin, lref = 0.05, 40.0
    out = O.slice_(mac, solid, dims, 2, 2, 2, uin, lref, 0.0, 12.0).reshape(ny, nx, 2)
    expect = a * lref / uin / 12.0 * 255 + 0.5
    assert np.all(out[1:-1, :, 0] == int(expect))
    assert out[3, 5, 1] == 0 and out[3, 4, 1] == 255
    sp = O.slice_(mac, solid, dims, 1, 3, 0, uin, lref, 0.0, 1.6).reshape(nz, nx, 2)
    assert np.all(sp[:, :, 0] == int(min(1.0, a * 3 / uin / 1.6) * 255 + 0.5))
    w, h = O.slice_dims(dims, 0)
    assert (w, h) == (nz, ny)


def test_cp_known_density():
    dims = (20, 20, 20)
    n = 8000
    mac = np.tile(np.float32([0, 0, 0, 1.003]), (n, 1))
    solid = np.zeros(n, np.uint8)
    verts = np.float32([[10, 10, 10], [1, 1, 1]])
    norms = np.float32([[1, 0, 0], [0, 1, 0]])
    out = O.cp(mac, solid, verts, norms, dims, 0.05, 1.2)
    cpv = (1.003 - 1) / (1.5 * 0.05 ** 2)
    assert out.tolist() == [int(round(min(1, cpv / 1.2) * 127))] * 2
    # fully solid neighbourhood -> weight sum 0 -> Cp 0
    out2 = O.cp(mac, np.ones(n, np.uint8), verts, norms, dims, 0.05, 1.2)
    assert out2.tolist() == [0, 0]
