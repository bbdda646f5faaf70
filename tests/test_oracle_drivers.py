"""Pinning of the oracle's pressure-correction drivers (oracle/cap_oracle.c:
orc_calcp, orc_piso, orc_outbc, orc_correct_bc, orc_gethbya).

The reference's driver files cannot be built here: every driver links
grad_lsq_qr (GRAD_LSQ_QR.f90), which needs LAPACK's dgeqrf, and PISO / PIMPLE
call pcgsip, which needs MPI.  Their building blocks (fluxmass, grad_gauss,
fluxmc, the solvers, bpres, continuity sums) are pinned against the reference
build elsewhere; the steps between them are pinned here by what the
mathematics fixes:
  * after the flux correction the divergence of the corrected fluxes equals
    the p' equation's residual (cells off the boundary);
  * the pressure at the reference cell does not move, and the velocity
    correction takes the matching gradient component;
  * outbc's rescaled outflow equals flowin;
  * correctBoundaryConditions' copies are periodic;
  * PISO_getHbyA's closed forms (uniform old velocity, BDF without
    neighbours, buoyancy);
  * PISO's reference-cell row."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import flowcase as FC  # noqa: E402

DIMS = (16, 13, 11)


def grid(a, dims):
    ni, nj, nk = dims
    return a.reshape(nk, ni, nj)


def inner(a, dims, lo=3, shrink=0):
    ni, nj, nk = dims
    g = grid(a, dims)
    s = slice(lo - 1 + shrink, None)
    return g[lo - 1 + shrink:nk - 2 - shrink, lo - 1 + shrink:ni - 2 - shrink,
             lo - 1 + shrink:nj - 2 - shrink]


def div(f, dims):
    """-div f per cell (reference sign: su = -f1 + f1_w - f2 + f2_s - f3 + f3_b)"""
    g1, g2, g3 = (grid(f[q], dims) for q in ("f1", "f2", "f3"))
    out = np.zeros_like(g1)
    out[1:, 1:, 1:] = (-g1[1:, 1:, 1:] + g1[1:, :-1, 1:] - g2[1:, 1:, 1:] + g2[1:, 1:, :-1]
                       - g3[1:, 1:, 1:] + g3[:-1, 1:, 1:])
    return out


def residual(f, x, dims):
    """sum a_nb x_nb + su - ap x per cell"""
    g = {q: grid(f[q], dims) for q in ("ae", "aw", "an", "as", "at", "ab", "ap", "su")}
    X = grid(x, dims)
    r = np.zeros_like(X)
    c = (slice(1, -1),) * 3
    r[c] = (g["ae"][c] * X[1:-1, 2:, 1:-1] + g["aw"][c] * X[1:-1, :-2, 1:-1] +
            g["an"][c] * X[1:-1, 1:-1, 2:] + g["as"][c] * X[1:-1, 1:-1, :-2] +
            g["at"][c] * X[2:, 1:-1, 1:-1] + g["ab"][c] * X[:-2, 1:-1, 1:-1] +
            g["su"][c] - g["ap"][c] * X[c])
    return r


@pytest.mark.parametrize("solver", [4, 1, 5])
def test_calcp_corrected_fluxes_balance_the_residual(O, solver):
    """Stopped after one iteration, so that the residual is not small against the
    source: the corrected fluxes' divergence must be that residual, cell by
    cell (a sign or index slip in the assembly or the correction leaves an
    O(su) difference)."""
    f = FC.flow_case(DIMS, 5)
    p0 = f["p"].copy()
    o = FC.opts(O.DrvOpts, solver=solver, alfa=0.0 if solver == 4 else 0.92, urf=0.5, nsw=1,
                ijkpr=FC.ref_cell(DIMS, 5, 6, 4))
    reps, cont = O.calcp(DIMS, f, o)
    # pp up to a constant (ap = sum a_nb, so a constant does not change r)
    pp = (f["p"] - p0) / o.urf
    r = residual(f, pp, DIMS)
    d = div(f, DIMS)
    scale = np.max(np.abs(f["su"]))
    got, want = d[3:-3, 3:-3, 3:-3], r[3:-3, 3:-3, 3:-3]
    assert np.max(np.abs(want)) > 1e-4 * scale
    assert np.max(np.abs(got - want)) < 1e-9 * scale


def test_calcp_reference_cell(O):
    f = FC.flow_case(DIMS, 6)
    p0 = f["p"].copy()
    ipr = FC.ref_cell(DIMS, 6, 5, 7)
    o = FC.opts(O.DrvOpts, solver=4, ijkpr=ipr, ncorr=2, npcor=2)
    O.calcp(DIMS, f, o)
    assert f["p"][ipr - 1] == p0[ipr - 1]
    assert not np.array_equal(inner(f["p"], DIMS), inner(p0, DIMS))
    # pp = p on return
    np.testing.assert_array_equal(f["pp"], f["p"])


def test_calcp_velocity_correction_components(O):
    f = FC.flow_case(DIMS, 7)
    u0, v0, w0 = (f[q].copy() for q in ("u", "v", "w"))
    O.calcp(DIMS, f, FC.opts(O.DrvOpts, solver=4, ijkpr=FC.ref_cell(DIMS)))
    for m, (q, a0) in enumerate((("u", u0), ("v", v0), ("w", w0))):
        want = a0 - f["vol"] * f["ap" + q] * f["gradp"][m::3]
        np.testing.assert_allclose(inner(f[q], DIMS), inner(want, DIMS), rtol=1e-13, atol=1e-15)


def test_outbc_rescaled_outflow_equals_flowin(O):
    dims = DIMS
    f = FC.flow_case(dims, 8)
    lb = FC.boundary_types(dims)
    o = FC.opts(O.DrvOpts, lout=1, loe=1, low=1, flowin=3.7)
    for q, a in lb.items():
        setattr(o, q, a.ctypes.data)
    u_in = grid(f["u"], dims)[1:-1, -3, 1:-1].copy()
    O.outbc(dims, f, o)
    ni, nj, nk = dims
    g1 = grid(f["f1"], dims)
    outflow = g1[1:nk - 1, ni - 3, 1:nj - 1].sum()     # f1(inh), inh at i = nim - 1
    assert outflow == pytest.approx(3.7, rel=1e-12)
    # the outlet cells carry the scaled inner velocity
    fac = grid(f["u"], dims)[1:-1, -2, 1:-1] / u_in
    assert np.allclose(fac, fac.flat[0], rtol=1e-14)
    # walls elsewhere: the west flux is untouched
    assert np.all(grid(f["f1"], dims)[:, 1, :] == 0.0)


def test_correct_bc_is_periodic(O):
    dims = (14, 14, 9)
    f = FC.flow_case(dims, 9)
    f["f1"][:] = np.random.default_rng(1).normal(size=f["f1"].size)
    f["f2"][:] = np.random.default_rng(2).normal(size=f["f2"].size)
    O.correct_bc(dims, f)
    ni, nj, nk = dims
    for q in ("u", "v", "w"):
        g = grid(f[q], dims)
        np.testing.assert_array_equal(g[1:nk - 1, ni - 1, 1:nj - 1], g[1:nk - 1, 1, 1:nj - 1])
        np.testing.assert_array_equal(g[1:nk - 1, 0, 1:nj - 1], g[1:nk - 1, nj - 2, 1:nj - 1])
        np.testing.assert_array_equal(g[1:nk - 1, 1:ni - 1, nj - 1], g[1:nk - 1, 1:ni - 1, 1])
        np.testing.assert_array_equal(g[1:nk - 1, 1:ni - 1, 0], g[1:nk - 1, 1:ni - 1, nj - 2])
    g1, g2 = grid(f["f1"], dims), grid(f["f2"], dims)
    np.testing.assert_array_equal(g1[1:nk - 1, 0, 1:nj - 1], g1[1:nk - 1, nj - 2, 1:nj - 1])
    np.testing.assert_array_equal(g2[1:nk - 1, 1:ni - 1, 0], g2[1:nk - 1, 1:ni - 1, nj - 2])


def test_gethbya_closed_forms(O):
    dims = DIMS
    # Crank-Nicolson with a uniform old velocity: the old-time part is apotime * c
    f = FC.flow_case(dims, 10)
    c, U = 0.7, 1.3
    for q in ("uo", "vo", "wo"):
        f[q][:] = c
    for q in ("u", "v", "w"):
        f[q][:] = U
    o = FC.opts(O.DrvOpts, cn=1, timestep=0.02)
    H = sum(f[q] for q in ("he", "hw", "hn", "hs", "ht", "hb"))
    apot = f["den"] * f["vol"] / 0.02
    want = f["apu"] * (apot * c + U * H)
    O.gethbya(dims, f, o)
    np.testing.assert_allclose(inner(f["u"], dims), inner(want, dims), rtol=1e-12)
    # BDF2 alone, no neighbours
    f = FC.flow_case(dims, 11)
    for q in ("he", "hw", "hn", "hs", "ht", "hb"):
        f[q][:] = 0.0
    o = FC.opts(O.DrvOpts, bdf=1, btime=1.0, timestep=0.05)
    want = f["apv"] * (f["den"] * f["vol"] / 0.05) * (2.0 * f["vo"] - 0.5 * f["voo"])
    O.gethbya(dims, f, o)
    np.testing.assert_allclose(inner(f["v"], dims), inner(want, dims), rtol=1e-13)
    np.testing.assert_array_equal(inner(f["sp"], dims), 0.0)
    # Boussinesq buoyancy alone, gravity along z
    f = FC.flow_case(dims, 12)
    for q in ("he", "hw", "hn", "hs", "ht", "hb"):
        f[q][:] = 0.0
    o = FC.opts(O.DrvOpts, buoy=1, boussinesq=1, beta=2e-3, densit=1.2, tref=1.0, gravz=-9.81)
    heat = 2e-3 * 1.2 * (f["t"] - 1.0) * f["vol"]
    O.gethbya(dims, f, o)
    np.testing.assert_allclose(inner(f["sw"], dims), inner(9.81 * heat, dims), rtol=1e-13)
    np.testing.assert_array_equal(inner(f["su"], dims), 0.0)


@pytest.mark.parametrize("pimple", [False, True])
def test_piso_reference_cell_row(O, pimple):
    f = FC.flow_case(DIMS, 13)
    ipr = FC.ref_cell(DIMS, 4, 4, 4)
    o = FC.opts(O.DrvOpts, ncorr=2, npcor=2, ijkpr=ipr, nsw=200)
    reps, cont = O.piso(DIMS, f, o, pimple=pimple)
    c = ipr - 1
    assert f["ap"][c] == 1.0
    assert all(f[q][c] == 0.0 for q in ("ae", "aw", "an", "as", "at", "ab"))
    # non-final correctors solve to sor = 0.1 (a default-real literal)
    assert reps[0]["converged"] and reps[0]["resl"] < float(np.float32(0.1)) * reps[0]["resor"]
    # PISO starts from p = 0 and ends with pp = p
    np.testing.assert_array_equal(f["pp"], f["p"])
    assert len(cont) == 2 and cont[1][2] == pytest.approx(cont[0][1] + cont[1][1], rel=1e-15)


def uvw_case(O, dims, seed, **kw):
    f = FC.flow_case(dims, seed)
    rng = np.random.default_rng(seed + 100)
    for q in ("f1", "f2", "f3"):
        f[q][:] = 0.1 * rng.normal(size=f[q].size)
    f["pp"][:] = rng.normal(size=f["pp"].size)
    spv = np.zeros_like(f["u"])
    o = FC.opts(O.DrvOpts, alfa=0.92, **kw)
    q = O.UvwOpts()
    q.schemes, q.gds = 1, 1.0
    for m in range(3):
        q.urfr[m], q.urfm[m], q.sor[m], q.nsw[m] = 1 / 0.7, 0.3, 1e-6, 50
    q.nx, q.ny, q.nz = dims[0] - 2, dims[1] - 2, dims[2] - 2
    return f, spv, o, q


UVW_DIMS = (14, 12, 14)


def test_calcuvw_diagonals_and_inverse(O):
    """spv and sp receive the same additions (time term, wall shear); the
    last assembled diagonal is the w equation's, (sum a_nb + sp) urfr(iw),
    and apw = 1 / (ap + small); the w solve converged on it."""
    f, spv, o, q = uvw_case(O, UVW_DIMS, 3, bdf=1)
    reps = O.calcuvw(UVW_DIMS, f, spv, o, q)
    assert all(r["converged"] for r in reps)
    ni, nj, nk = UVW_DIMS
    g = {k: grid(f[k], UVW_DIMS)[1:nk - 1, 2:ni - 2, 2:nj - 2] for k in
         ("ae", "aw", "an", "as", "at", "ab", "ap", "sp", "apw")}
    np.testing.assert_array_equal(grid(spv, UVW_DIMS)[1:nk - 1, 2:ni - 2, 2:nj - 2], g["sp"])
    s = g["ae"] + g["aw"] + g["an"] + g["as"] + g["at"] + g["ab"] + g["sp"]
    np.testing.assert_array_equal(g["ap"], s * q.urfr[2])
    np.testing.assert_array_equal(g["apw"], 1.0 / (g["ap"] + float(np.float32(1e-30))))
    # away from the walls spv is the BDF time term alone; the wall shear adds
    # to it on the first cells above the bottom wall
    tt = grid(f["den"] * f["vol"] / o.timestep * (1.0 + 0.5 * o.btime), UVW_DIMS)
    sg = grid(spv, UVW_DIMS)
    np.testing.assert_array_equal(sg[4, 2:ni - 2, 2:nj - 2], tt[4, 2:ni - 2, 2:nj - 2])
    assert np.all(sg[2, 2:ni - 2, 2:nj - 2] > tt[2, 2:ni - 2, 2:nj - 2])


def test_calcuvw_crank_nicolson_halves_the_coefficients(O):
    fa, spa, oa, qa = uvw_case(O, UVW_DIMS, 4, bdf=1)
    fb, spb, ob, qb = uvw_case(O, UVW_DIMS, 4, bdf=1, cn=1)
    O.calcuvw(UVW_DIMS, fa, spa, oa, qa)
    O.calcuvw(UVW_DIMS, fb, spb, ob, qb)
    for k in ("ae", "aw", "an", "as", "at", "ab"):
        np.testing.assert_array_equal(fb[k], 0.5 * fa[k], err_msg=k)
    # the Crank-Nicolson time term joins the diagonal source of each equation
    assert np.all(grid(spb, UVW_DIMS)[2:-2, 2:-2, 2:-2] > grid(spa, UVW_DIMS)[2:-2, 2:-2, 2:-2])


# ------------------------------------------- calcuvw's sources, closed forms
def uniform_uvw(O, n=8, nz_int=3, **kw):
    """A uniform n^3 grid (h = 1/n, outward face areas h^2) with every field
    constant or zero: zero mass fluxes, pp = 0, rho = 1, mu = 0.02, steady.
    Each term of calcuvw's sources can then be isolated exactly.  nz_int:
    the module integer nz that the wall distance is projected on (calcuvw.f90:
    187, 277 use nx, ny, nz where a normal is meant)."""
    import configcase as CC
    import gridcase as GC
    w = GC.uniform_widths(n)
    g, dims = CC._geometry(O, w, w, w)
    f = FC.flow_case(dims, 1)
    f.update(g)
    N = dims[0] * dims[1] * dims[2]
    for q in ("u", "v", "w", "pp", "f1", "f2", "f3", "uo", "vo", "wo", "uoo", "voo", "woo",
              "su", "sv", "sw", "sp", "ae", "aw", "an", "as", "at", "ab", "ap"):
        f[q] = np.zeros(N)
    f["den"] = np.ones(N)
    f["vis"] = np.full(N, 0.02)
    f["t"] = np.ones(N)
    o = FC.opts(O.DrvOpts, alfa=0.92, const_mflux=0, **kw)
    q = O.UvwOpts()
    q.schemes, q.gds = 1, 1.0
    for m in range(3):
        q.urfr[m], q.urfm[m], q.sor[m], q.nsw[m] = 1 / 0.7, 0.3, 0.2, 1
    q.nx, q.ny, q.nz = 1, 1, nz_int
    return f, dims, o, q, 1.0 / n


def _run_uvw(O, f, dims, o, q):
    spv = np.zeros_like(f["u"])
    O.calcuvw(dims, f, spv, o, q)
    return spv


def test_calcuvw_wall_shear_closed_form(O):
    """Constant v, w on a uniform grid, no flux, no pressure gradient.  The
    wall terms of calcuvw.f90:161-337 act on the first cells above the bottom
    and below the top wall (k = 3, nkmm): sp and spv gain vsol = mu A / dpb
    with dpb = h/2 (nothing else adds to them in a steady run), and the
    sources gain vsol phi - cf phi_t with cf = mu A / dnw, dnw = (h/2) nz
    (the module integer nz, as written).  fluxuvw also acts near the walls
    here (the ghost cells' partial Gauss sums enter its explicit diffusion),
    so cf is isolated by running twice with nz = 3 and nz = 5: the tangential
    v moves by -mu A V (2/h)(1/3 - 1/5), the wall-normal w (phi_t = 0) not at
    all, and nothing else anywhere."""
    runs = []
    for nz_int in (3, 5):
        f, dims, o, q, h = uniform_uvw(O, nz_int=nz_int)
        V, W = 0.7, -0.4
        f["v"][:] = V
        f["w"][:] = W
        w_in = f["w"].copy()
        spv = _run_uvw(O, f, dims, o, q)
        runs.append((f, spv, w_in))
    ni, nj, nk = dims
    mu, A = 0.02, h * h
    vsol = mu * A / (0.5 * h)
    (fa, spva, w_in), (fb, spvb, _) = runs
    I = (slice(2, ni - 2), slice(2, nj - 2))
    for k in (2, nk - 3):           # 0-based planes k = 3 and nkmm
        np.testing.assert_allclose(grid(fa["sp"], dims)[k][I], vsol, rtol=1e-13)
        np.testing.assert_allclose(grid(spva, dims)[k][I], vsol, rtol=1e-13)
        dsv = grid(fa["sv"], dims)[k][I] - grid(fb["sv"], dims)[k][I]
        np.testing.assert_allclose(dsv, -mu * A * V * (2.0 / h) * (1.0 / 3 - 1.0 / 5), rtol=1e-9)
    for k in range(3, nk - 3):
        assert np.all(grid(fa["sp"], dims)[k][I] == 0.0)
    np.testing.assert_array_equal(fa["sw"], fb["sw"])
    wall = np.zeros((nk, ni, nj), bool)
    wall[2] = wall[nk - 3] = True
    assert np.all(grid(fa["sv"], dims)[~wall] == grid(fb["sv"], dims)[~wall])
    # away from every ghost layer, nothing acts on the constant field
    C = (slice(3, nk - 3), slice(3, ni - 3), slice(3, nj - 3))
    assert np.all(grid(fa["sv"], dims)[C] == 0.0) and np.all(grid(fa["sw"], dims)[C] == 0.0)
    # the under-relaxation source of the last (w) equation, calcuvw.f90:458:
    # su = sw + urfm ap w_old, ap already relaxed by urfr
    R = (slice(1, nk - 1), slice(2, ni - 2), slice(2, nj - 2))
    want = grid(fa["sw"], dims) + 0.3 * grid(fa["ap"], dims) * grid(w_in, dims)
    np.testing.assert_array_equal(grid(fa["su"], dims)[R], want[R])


def test_calcuvw_pressure_gradient_closed_form(O):
    """A linear pressure correction pp = a x + b y + c z, zero velocity: the
    momentum sources are -grad pp vol (calcuvw.f90:60-62), with the Gauss
    gradient exact for a linear field on the uniform grid's interior."""
    f, dims, o, q, h = uniform_uvw(O)
    a, b, c = 0.3, -1.7, 2.2
    f["pp"][:] = a * f["xc"] + b * f["yc"] + c * f["zc"]
    _run_uvw(O, f, dims, o, q)
    ni, nj, nk = dims
    I = (slice(2, nk - 2), slice(2, ni - 2), slice(2, nj - 2))
    vol = h ** 3
    np.testing.assert_allclose(grid(f["sv"], dims)[I], -b * vol, rtol=1e-12)
    np.testing.assert_allclose(grid(f["sw"], dims)[I], -c * vol, rtol=1e-12)


def test_calcuvw_time_and_buoyancy_closed_forms(O):
    """BDF2 (calcuvw.f90:86-98): s += rho vol/dt ((1 + btime) v_o - btime/2 v_oo)
    and ap += rho vol/dt (1 + btime/2); Boussinesq buoyancy (74-80):
    s -= g beta rho_ref (t - t_ref) vol.  Zero velocity, so no wall term."""
    dt, bt, beta, tref, gy, gz = 0.05, 1.0, 3e-3, 0.5, -2.0, -9.81
    f, dims, o, q, h = uniform_uvw(O, bdf=1, btime=bt, timestep=dt, buoy=1, boussinesq=1,
                                   beta=beta, tref=tref, gravx=0.0, gravy=gy, gravz=gz, densit=1.3)
    rng = np.random.default_rng(4)
    N = f["u"].size
    f["vo"][:] = rng.normal(size=N)
    f["voo"][:] = rng.normal(size=N)
    f["wo"][:] = rng.normal(size=N)
    f["t"][:] = rng.uniform(0, 1, N)
    vo, voo, wo, t = (f[k].copy() for k in ("vo", "voo", "wo", "t"))
    spv = _run_uvw(O, f, dims, o, q)
    ni, nj, nk = dims
    I = (slice(2, nk - 2), slice(2, ni - 2), slice(2, nj - 2))
    vol = h ** 3
    at = vol / dt
    heat = beta * 1.3 * (grid(t, dims) - tref) * vol
    want_v = at * ((1 + bt) * grid(vo, dims) - 0.5 * bt * grid(voo, dims)) - gy * heat
    want_w = at * ((1 + bt) * grid(wo, dims)) - gz * heat
    np.testing.assert_allclose(grid(f["sv"], dims)[I], want_v[I], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(grid(f["sw"], dims)[I], want_w[I], rtol=1e-12, atol=1e-15)
    J = (slice(3, nk - 3), slice(2, ni - 2), slice(2, nj - 2))     # no wall cells
    np.testing.assert_allclose(grid(spv, dims)[J], at * (1 + 0.5 * bt), rtol=1e-14)


def test_pimple_pressure_update_closed_form(O):
    """PIMPLE_multiple_correction.f90 solves for the pressure itself, from
    p = 0 (line 57): after one corrector p = p + urf (pp - p) = urf pp on the
    interior (275-283), so pp* = p / urf must solve the assembled system
    left in ae..ap, su (reference-cell row included: pp*(ref) = p_ref = 0) to
    the requested residual; the velocities take -apu grad(pp) vol with the
    gradient the routine leaves in gradp (276-278); and pp = p on return."""
    f = FC.flow_case(DIMS, 17)
    ipr = FC.ref_cell(DIMS, 4, 5, 4)
    urf = 0.4
    o = FC.opts(O.DrvOpts, ncorr=1, npcor=1, ijkpr=ipr, nsw=2000, sor=1e-10, urf=urf)
    pp_in, u_in, v_in = f["pp"].copy(), f["u"].copy(), f["v"].copy()
    reps, _ = O.piso(DIMS, f, o, pimple=True)
    assert reps[0]["converged"]
    ni, nj, nk = DIMS
    m = np.zeros((nk, ni, nj), bool)
    m[2:-2, 2:-2, 2:-2] = True
    m = m.ravel()
    x = pp_in.copy()                         # ghost values as the solve saw them
    x[m] = f["p"][m] / urf
    s = O.System(DIMS)
    for q in ("ae", "aw", "an", "at", "ab", "ap", "su"):
        getattr(s, q)[:] = f[q]
    s.as_[:] = f["as"]
    r = f["su"] - O.spmv(s, 3, nk - 2, x)
    r0 = f["su"] - O.spmv(s, 3, nk - 2, pp_in)
    assert np.abs(r[m]).sum() < 1e-8 * np.abs(r0[m]).sum()
    assert f["p"][ipr - 1] == 0.0
    np.testing.assert_array_equal(f["pp"], f["p"])
    gp = f["gradp"].reshape(-1, 3)
    np.testing.assert_array_equal(f["u"][m], (u_in - f["apu"] * gp[:, 0] * f["vol"])[m])
    np.testing.assert_array_equal(f["v"][m], (v_in - f["apv"] * gp[:, 1] * f["vol"])[m])
    # and the update is not p = pp, nor p += urf pp from a non-zero start:
    # the recovered pp* would not solve the system then
    bad = pp_in.copy()
    bad[m] = f["p"][m]
    rb = f["su"] - O.spmv(s, 3, nk - 2, bad)
    assert np.abs(rb[m]).sum() > 1e-3 * np.abs(r0[m]).sum()
