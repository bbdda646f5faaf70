"""The channel flow of the learned-guess and full-step configurations, held
on the device and advanced by the library's routines.

Box: nx x ny x nz cells, periodic in x (streamwise) and z (spanwise), walls
at y = -1 and 1 with the cells clustered toward them by a tanh map.  The
reference hard-codes its walls in k (calcuvw.f90:161-337,
update_values_at_ghostcells.f90:96-145, calc_vis_les.f90:90-247), so
y maps to k; x maps to j, the unit-stride index, and z to i:
(ni, nj, nk) = (nz + 4, nx + 4, ny + 4).

The grid's face areas and interpolation factors come from the library's
geometry routines (init.f90:49-144, intfac.f90).  Those loops orient the
north and top face vectors to -j and -k on a right-handed grid; the state
takes every face vector outward (the north and top vectors negated), as the
flux routines assume.

A time step follows the reference's PISO loop body for this case: the
Courant-number time step (CourantNo.h with the fixed-Courant rescale), the
periodic ghost-cell fill, the Smagorinsky viscosity (find_strain_rate,
calc_vis_les), the momentum predictor (calcuvw: CDS with full deferred
correction, BDF2, wall shear at k, the driving pressure gradient) and the
pressure correction (PISO_multiple_correction, 2 correctors, the p' solve
by PCG-SIP or its mixed-precision form).  The top wall's terms are taken on
the wall's own first cells (top_k = nkmm; the reference's lk(nimm) is that
plane only when ni = nk).
"""
from __future__ import annotations

import math
import time

import numpy as np


def _widths_uniform(n, length):
    return np.full(n + 4, length / n)


def _widths_tanh(n, factor):
    """cells of [-1, 1] clustered toward both ends by tanh(factor s)/tanh(factor),
    plus two mirrored ghost cells at each end"""
    s = np.linspace(-1.0, 1.0, n + 1)
    y = np.tanh(factor * s) / np.tanh(factor)
    h = np.diff(y)
    return np.concatenate([h[1::-1], h, h[:-3:-1]]), y[0]


def _axis(widths, start):
    """node (closing each cell) and centre coordinates of padded cells whose
    first interior cell (the third) starts at `start`"""
    nodes = start + np.cumsum(widths) - widths[0] - widths[1]
    return nodes, nodes - 0.5 * widths


class Channel:
    """Device-resident state of the channel and its time step.

    cap: the loaded package; torch: the torch module (device memory);
    nx, ny, nz: interior cells (streamwise, wall-normal, spanwise);
    solver: 4 (PCG-SIP, alfa 0) or 6 (its mixed-precision form);
    fp32_flux: the momentum face fluxes' arithmetic in FP32 (CAP_UVW_FP32)."""

    def __init__(self, cap, torch, nx, ny, nz, stretch=2.0, re_bulk=5600.0, seed=7,
                 solver=6, courant=0.5, lx=2 * math.pi, lz=math.pi, les=True, fp32_flux=False):
        self.cap, self.torch = cap, torch
        self.nx, self.ny, self.nz = nx, ny, nz
        ni, nj, nk = nz + 4, nx + 4, ny + 4
        self.dims = (ni, nj, nk)
        n = ni * nj * nk
        self.n = n
        self.nu = 1.0 / re_bulk
        self.courant = courant
        self.les = les
        dev, f64 = "cuda", torch.float64
        wz, wx = _widths_uniform(nz, lz), _widths_uniform(nx, lx)
        wy, y0 = _widths_tanh(ny, stretch)
        Zn, Zc = _axis(wz, 0.0)
        Xn, Xc = _axis(wx, 0.0)
        Yn, Yc = _axis(wy, y0)
        shape = (nk, ni, nj)

        def b(a, axis):
            t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
            return t.view([-1 if d == axis else 1 for d in range(3)]).expand(shape).contiguous().view(-1)
        f = {}
        # i <-> z (spanwise), j <-> x (streamwise), k <-> y (wall-normal)
        f["x"], f["xc"] = b(Xn, 2), b(Xc, 2)
        f["y"], f["yc"] = b(Yn, 0), b(Yc, 0)
        f["z"], f["zc"] = b(Zn, 1), b(Zc, 1)
        f["vol"] = (b(wz, 1) * b(wx, 2) * b(wy, 0)).contiguous()
        zero = lambda: torch.zeros(n, dtype=f64, device=dev)  # noqa: E731
        for q in ("fx", "fy", "fz", "ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y",
                  "ar3z"):
            f[q] = zero()
        self.ctx = cap.Context(ni, nj, nk, device_pointers=True)
        self.ctx.face_areas(f)
        self.ctx.intfac(f)
        for q in ("ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z"):
            f[q].neg_()
        # laminar profile u = 1.5 (1 - y^2) (bulk velocity 1) with 10 % noise
        g = torch.Generator(device=dev).manual_seed(seed)
        noise = lambda: 0.1 * (2 * torch.rand(n, generator=g, dtype=f64, device=dev) - 1)  # noqa: E731
        yc = f["yc"]
        inside = torch.zeros(shape, dtype=torch.bool, device=dev)
        inside[2:-2, 2:-2, 2:-2] = True
        inside = inside.view(-1)
        f["u"] = torch.where(inside, 1.5 * (1 - yc * yc).clamp(min=0) * (1 + noise()), zero())
        f["v"] = torch.where(inside, noise(), zero())
        f["w"] = torch.where(inside, noise(), zero())
        for q in ("p", "pp", "ae", "aw", "an", "as", "at", "ab", "su", "sv", "sw", "sp", "apu", "apv",
                  "apw", "res", "bb", "bs", "bw", "bp", "bn", "be", "bt", "he", "hw", "hn", "hs", "ht",
                  "hb", "t", "vart", "con"):
            f[q] = zero()
        f["ap"] = torch.ones(n, dtype=f64, device=dev)
        f["den"] = torch.ones(n, dtype=f64, device=dev)
        f["vis"] = torch.full((n,), self.nu, dtype=f64, device=dev)
        f["te"] = torch.full((n,), 1e-6, dtype=f64, device=dev)
        f["ed"] = torch.full((n,), 1e-6, dtype=f64, device=dev)
        for q in ("gradu", "gradv", "gradw", "gradp"):
            f[q] = torch.zeros(3 * n, dtype=f64, device=dev)
        self.f = f
        self._ghosts()
        for q in ("u", "v", "w"):
            f[q + "o"], f[q + "oo"] = f[q].clone(), f[q].clone()
        f["f1"], f["f2"], f["f3"] = self._face_fluxes()
        # distance to the nearer wall, for the van Driest damping
        self.walldist = (1.0 - yc.abs()).clamp(min=1e-6).contiguous()
        self.spv = zero()
        self.strain, self.vort, self.cosu = zero(), zero(), zero()
        self.time = 0.0
        self.dt = 1e-3
        o = cap.DrvOpts()
        o.ncorr, o.npcor, o.solver, o.nsw = 2, 1, solver, 4000
        o.urf, o.sor, o.alfa = 1.0, 1e-8, 0.0
        o.ijkpr = 2 * ni * nj + 2 * nj + 3                 # the cell (3, 3, 3)
        o.const_mflux, o.bdf, o.btime, o.densit = 1, 1, 1.0, 1.0
        self.opts = o
        q = cap.UvwOpts()
        # CDS, full deferred correction (with CAP_UVW_FP32 = 1024: FP32 face arithmetic)
        q.schemes, q.gds = 1 | (1024 if fp32_flux else 0), 1.0
        for m in range(3):
            q.urfr[m], q.urfm[m], q.sor[m], q.nsw[m] = 1.0, 0.0, float(np.float32(0.1)), 5
        # the laminar driving force d p / d x = -3 nu (for u = 1.5 (1 - y^2))
        q.gradpcmf = 3.0 * self.nu
        # the wall distance of calcuvw's wall terms is projected on the module
        # integers (nx, ny, nz): (0, 1, 0) makes it the distance along y
        q.nx, q.ny, q.nz = 0, 1, 0
        q.top_k = nk - 2
        self.uvw = q
        lo = cap.LesOpts()
        lo.viscos, lo.densit, lo.cappa, lo.elog, lo.urfvis = self.nu, 1.0, 0.41, 8.432, 1.0
        lo.nx, lo.ny, lo.nz, lo.top_k = 0, 1, 0, nk - 2
        self.les_opts = lo
        self.phases = {"courant": 0.0, "les": 0.0, "calcuvw": 0.0, "piso": 0.0}
        self.steps = 0
        self.last = None

    def close(self):
        if self.ctx is not None:
            self.ctx.close()
            self.ctx = None

    # ------------------------------------------------------------- helpers
    def _ghosts(self):
        f = self.f
        self.ctx.update_ghostcells_periodic([f[q] for q in ("u", "v", "w", "te", "ed", "t", "vis")],
                                            f["fz"])

    def _face_fluxes(self):
        """rho (u_f . S) on every face from the linearly interpolated velocity"""
        torch = self.torch
        ni, nj, nk = self.dims
        f = self.f
        sh = (nk, ni, nj)
        out = []
        for lam, ar, ax in (("fx", "ar1", 1), ("fy", "ar2", 2), ("fz", "ar3", 0)):
            lamv = f[lam].view(sh)
            F = torch.zeros(sh, dtype=torch.float64, device="cuda")
            P = [slice(None)] * 3
            E = [slice(None)] * 3
            P[ax] = slice(0, sh[ax] - 1)
            E[ax] = slice(1, sh[ax])
            P, E = tuple(P), tuple(E)
            face = {q: lamv[P] * f[q].view(sh)[E] + (1 - lamv[P]) * f[q].view(sh)[P]
                    for q in ("u", "v", "w")}
            F[P] = (face["u"] * f[ar + "x"].view(sh)[P] + face["v"] * f[ar + "y"].view(sh)[P] +
                    face["w"] * f[ar + "z"].view(sh)[P])
            out.append(F.view(-1).contiguous())
        return out

    # ---------------------------------------------------------- the step
    def step(self, timed=True, net=None, warm=False):
        """One time step; returns the PISO solver reports.  With net (a UNet
        on this channel's context) or warm, the last corrector's p solve
        starts from the learned guess or from the previous step's p instead
        of the previous corrector's p, when that start has the lower L1
        residual, and stops at the same absolute residual
        (Context.piso_guess); self.last then also holds `accepted` and
        `l1` (reference start, start used)."""
        torch = self.torch
        f, ctx = self.f, self.ctx
        prev = f["p"].clone() if (net is not None or warm) else None
        for a, b in (("uoo", "uo"), ("voo", "vo"), ("woo", "wo"), ("uo", "u"), ("vo", "v"),
                     ("wo", "w")):
            f[a].copy_(f[b])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        co, mean, self.dt, self.time = ctx.courant(f["f1"], f["f2"], f["f3"], f["vol"], self.cosu,
                                                   self.dt, self.time, True, self.courant)
        self.opts.timestep = self.dt
        self._ghosts()
        t1 = time.perf_counter()
        if self.les:
            ctx.find_strain_rate(f, self.strain, self.vort)
            ctx.calc_vis_les(f, self.strain, self.walldist, self.les_opts)
        t2 = time.perf_counter()
        ctx.calcuvw(f, self.spv, self.opts, self.uvw)
        self._ghosts()
        t3 = time.perf_counter()
        extra = {}
        if prev is not None:
            reps, cont, acc, l1 = ctx.piso_guess(f, self.opts, prev, net=net)
            extra = dict(accepted=acc, l1=l1)
        else:
            reps, cont = ctx.piso(f, self.opts)
        self._ghosts()
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        if timed:
            for k, v in zip(("courant", "les", "calcuvw", "piso"), (t1 - t0, t2 - t1, t3 - t2, t4 - t3)):
                self.phases[k] += v * 1e3
            self.steps += 1
        self.last = dict(reps=reps, cont=cont, courant=co, mean_courant=mean, dt=self.dt, **extra)
        return reps

    # ------------------------------------------------- the learned guess
    def snapshot(self):
        """copies of every device field and the time-stepping scalars"""
        return ({q: a.clone() for q, a in self.f.items()}, self.spv.clone(), self.strain.clone(),
                self.vort.clone(), self.cosu.clone(), self.time, self.dt)

    def restore(self, snap):
        fields, spv, strain, vort, cosu, self.time, self.dt = snap
        for q, a in fields.items():
            self.f[q].copy_(a)
        self.spv.copy_(spv)
        self.strain.copy_(strain)
        self.vort.copy_(vort)
        self.cosu.copy_(cosu)

    def pressure_system(self):
        """the p system of the last corrector of the last step (PISO leaves
        it in ae..ab, ap, su) as a Coef"""
        f = self.f
        return self.cap.Coef(ae=f["ae"], aw=f["aw"], an=f["an"], as_=f["as"], at=f["at"],
                             ab=f["ab"], ap=f["ap"], su=f["su"])

    def interior(self, a):
        """the solve range (k, i, j = 3..n?mm) of a field, in the learned
        guess's voxel order (j fastest, then i, then k)"""
        ni, nj, nk = self.dims
        return a.view(nk, ni, nj)[2:-2, 2:-2, 2:-2].reshape(-1)

    def residual(self, x):
        """r = su - M x on the solve range of the last corrector's system
        (M x = ap x - sum a_nb x_nb), in the learned guess's voxel order"""
        ni, nj, nk = self.dims
        f = self.f
        X = x.view(nk, ni, nj)
        c = lambda q: f[q].view(nk, ni, nj)[2:-2, 2:-2, 2:-2]  # noqa: E731
        r = (c("ae") * X[2:-2, 3:-1, 2:-2] + c("aw") * X[2:-2, 1:-3, 2:-2] +
             c("an") * X[2:-2, 2:-2, 3:-1] + c("as") * X[2:-2, 2:-2, 1:-3] +
             c("at") * X[3:-1, 2:-2, 2:-2] + c("ab") * X[1:-3, 2:-2, 2:-2] + c("su") -
             c("ap") * X[2:-2, 2:-2, 2:-2])
        return r.reshape(-1)

    def pressure_pair(self, prev):
        """one training pair of the learned guess from the step just taken:
        the warm start w = prev on the solve range with the ghost values the
        last corrector's solve held (those of p), inputs (r / rms r,
        ap / mean ap, w / rms w) with r = su - M w (the channels
        cap_unet_guess feeds the network in cap_piso_guess), target the
        change p - w in units of its RMS; fp32, channel-major, on the device"""
        torch = self.torch
        rms = lambda a: a.pow(2).mean().sqrt().clamp(min=1e-300)  # noqa: E731
        w = self.f["p"].clone()
        ni, nj, nk = self.dims
        w.view(nk, ni, nj)[2:-2, 2:-2, 2:-2] = prev.view(nk, ni, nj)[2:-2, 2:-2, 2:-2]
        r = self.residual(w)
        ap = self.interior(self.f["ap"])
        pv = self.interior(w)
        d = self.interior(self.f["p"]) - pv
        x = torch.stack([r / rms(r), ap / ap.mean(), pv / rms(pv) if float(pv.abs().max()) > 0 else pv])
        return x.float().reshape(-1).contiguous(), (d / rms(d)).float().contiguous()

    def phase_ms(self):
        s = max(self.steps, 1)
        return {k: v / s for k, v in self.phases.items()}

    def bulk_velocity(self):
        """volume-weighted mean of u over the interior"""
        torch = self.torch
        ni, nj, nk = self.dims
        sh = (nk, ni, nj)
        u = self.f["u"].view(sh)[2:-2, 2:-2, 2:-2]
        vol = self.f["vol"].view(sh)[2:-2, 2:-2, 2:-2]
        return float((u * vol).sum() / vol.sum())

    def continuity(self):
        """{sum of |div f|, sum of div f} over the interior (continuityErrors.h)"""
        return self.ctx.continuity(self.f["f1"], self.f["f2"], self.f["f3"])
