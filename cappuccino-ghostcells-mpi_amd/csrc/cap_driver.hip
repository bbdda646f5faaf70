// cap_driver.hip -- the pressure-correction drivers on the device: calcp
// (SIMPLE), PISO_multiple_correction with PISO_getHbyA, PIMPLE, and the
// boundary routines they call (outbc, correctBoundaryConditions).
//
// Each driver is a fixed sequence of the library's stencil kernels and
// solvers plus the element-wise steps between them, all on one stream over
// device-resident module state: nothing returns to the host inside a driver
// except the solver reports and the continuity sums.  Values that the
// reference reads back from a field in the middle of a step (ppref =
// pp(ijkpr), the reference-cell pin, the outlet rescale factor) are read by
// the kernels from device memory.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include "cap_kernels.h"
#include "cap_amd.h"
#include "cap_ctx.h"

using namespace cap;

#define HIPCHK(call)                                                                      \
    do {                                                                                  \
        hipError_t _e = (call);                                                           \
        if (_e != hipSuccess) {                                                           \
            snprintf(ctx->err, sizeof ctx->err, "%s failed: %s (%s:%d)", #call,           \
                     hipGetErrorString(_e), __FILE__, __LINE__);                          \
            return CAP_ERR_HIP;                                                           \
        }                                                                                 \
    } while (0)
#define RCHK(call)                                                                        \
    do {                                                                                  \
        int _rc = (call);                                                                 \
        if (_rc < 0) return _rc;                                                          \
    } while (0)

namespace {

constexpr double SMALL_F = (double)1e-30f;

struct Dim {
    int ni, nj, nk, nij;
};

// interior range k, i, j = lo..n?mm as (k, i, j) of a flat index q
__device__ __forceinline__ bool range_cell(const Dim &d, int lo, long q, size_t &c)
{
    const int nx = d.nj - 2 - lo + 1, ny = d.ni - 2 - lo + 1, nz = d.nk - 2 - lo + 1;
    if (q >= (long)nx * ny * nz) return false;
    const int j = lo + (int)(q % nx), i = lo + (int)((q / nx) % ny), k = lo + (int)(q / ((long)nx * ny));
    c = (size_t)(k - 1) * d.nij + (size_t)(i - 1) * d.nj + (size_t)(j - 1);
    return true;
}

__host__ __device__ long range_count(const Dim &d, int lo)
{
    return (long)(d.nj - 2 - lo + 1) * (d.ni - 2 - lo + 1) * (d.nk - 2 - lo + 1);
}

int grid_for(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

// su = -div f (calcp-multiple_correction_SIMPLE.f90:75,
// PISO_multiple_correction.f90:104-106) and / or ap = sum a_nb (76; 116) on
// k, i, j = 3..n?mm
__global__ __launch_bounds__(256) void k_su_ap(Dim d, int do_su, int do_ap, const double *f1,
                                               const double *f2, const double *f3, const double *ae,
                                               const double *aw, const double *an, const double *as,
                                               const double *at, const double *ab, double *su,
                                               double *ap)
{
    const long n = range_count(d, 3);
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        size_t c;
        range_cell(d, 3, q, c);
        if (do_su) su[c] = -f1[c] + f1[c - d.nj] - f2[c] + f2[c - 1] - f3[c] + f3[c - d.nij];
        if (do_ap) ap[c] = ae[c] + aw[c] + an[c] + as[c] + at[c] + ab[c];
    }
}

// The second-corrector flux loop of calcp (126-166) in gather form.  su is
// zeroed on the whole array (119-121), then the reference scatters, in loop
// order (k, i, j), su(P) += fmcor and su(E|N|T) -= fmcor for every cell P of
// k, i, j = 3..n?mm.  The value reaching cell C is
// ((((0 - ft_B) - fe_W) - fn_S) + fe_C) + fn_C) + ft_C, B = k-1, W = i-1,
// S = j-1, each term present when its source cell is in the loop range; the
// fluxes of the range cells lose their fmcor.
__global__ __launch_bounds__(256) void k_fmc_gather(Dim d, const double *fe, const double *fn,
                                                    const double *ft, double *su, double *f1,
                                                    double *f2, double *f3)
{
    const size_t n = (size_t)d.nij * d.nk;
    const int hi_i = d.ni - 2, hi_j = d.nj - 2, hi_k = d.nk - 2;
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        const int k = (int)(c / d.nij) + 1, rem = (int)(c % d.nij);
        const int i = rem / d.nj + 1, j = rem % d.nj + 1;
        auto in = [&](int ii, int jj, int kk) {
            return kk >= 3 && kk <= hi_k && ii >= 3 && ii <= hi_i && jj >= 3 && jj <= hi_j;
        };
        double s = 0.0;
        if (in(i, j, k - 1)) s = s - ft[c - d.nij];
        if (in(i - 1, j, k)) s = s - fe[c - d.nj];
        if (in(i, j - 1, k)) s = s - fn[c - 1];
        if (in(i, j, k)) {
            s = s + fe[c];
            s = s + fn[c];
            s = s + ft[c];
            f1[c] = f1[c] - fe[c];
            f2[c] = f2[c] - fn[c];
            f3[c] = f3[c] - ft[c];
        }
        su[c] = s;
    }
}

// mass-flux correction f -= a (x_nb - x_P) on k, i, j = lo..n?mm
// (calcp 173-191: lo = 3, x = pp; PISO 146-165: lo = 2, x = p; PIMPLE: x = pp)
__global__ __launch_bounds__(256) void k_corr_flux(Dim d, int lo, const double *x, const double *ae,
                                                   const double *an, const double *at, double *f1,
                                                   double *f2, double *f3)
{
    const long n = range_count(d, lo);
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        size_t c;
        range_cell(d, lo, q, c);
        f1[c] = f1[c] - ae[c] * (x[c + d.nj] - x[c]);
        f2[c] = f2[c] - an[c] * (x[c + 1] - x[c]);
        f3[c] = f3[c] - at[c] * (x[c + d.nij] - x[c]);
    }
}

// velocity and pressure correction on k, i, j = 3..n?mm.
//  mode 0 (calcp 196-213): u -= vol apu gradp(1), p += urf (pp - pp(ijkpr))
//  mode 1 (PISO 175-186):  u -= apu gradp(1) vol
//  mode 2 (PIMPLE 176-190): as 1, then p += urf (pp - p)
__global__ __launch_bounds__(256) void k_corr_uvwp(Dim d, int mode, const double *vol,
                                                   const double *apu, const double *apv,
                                                   const double *apw, const double *gp, double *u,
                                                   double *v, double *w, double *p,
                                                   const double *pp, size_t ipr, double urf)
{
    const long n = range_count(d, 3);
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        size_t c;
        range_cell(d, 3, q, c);
        if (mode == 0) {
            const double ppref = pp[ipr];
            u[c] = u[c] - vol[c] * apu[c] * gp[3 * c];
            v[c] = v[c] - vol[c] * apv[c] * gp[3 * c + 1];
            w[c] = w[c] - vol[c] * apw[c] * gp[3 * c + 2];
            p[c] = p[c] + urf * (pp[c] - ppref);
        } else {
            u[c] = u[c] - apu[c] * gp[3 * c] * vol[c];
            v[c] = v[c] - apv[c] * gp[3 * c + 1] * vol[c];
            w[c] = w[c] - apw[c] * gp[3 * c + 2] * vol[c];
            if (mode == 2) p[c] = p[c] + urf * (pp[c] - p[c]);
        }
    }
}

// the reference-pressure cell of PISO/PIMPLE (118-129): ap = 1, a_nb = 0,
// su = ppref = p(inpr)
__global__ void k_pin(size_t c, double *ae, double *aw, double *an, double *as, double *at,
                      double *ab, double *ap, double *su, const double *p)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    ap[c] = 1.0;
    ae[c] = 0.0;
    aw[c] = 0.0;
    an[c] = 0.0;
    as[c] = 0.0;
    at[c] = 0.0;
    ab[c] = 0.0;
    su[c] = p[c];
}

__global__ void k_copy(const double *a, double *b, size_t n)
{
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) b[c] = a[c];
}

__global__ void k_zero(double *a, size_t n)
{
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) a[c] = 0.0;
}

// ------------------------------------------------------------ PISO_getHbyA
struct HbyaIO {
    const double *den, *vol, *t, *uo, *vo, *wo, *uoo, *voo, *woo;
    const double *he, *hw, *hn, *hs, *ht, *hb, *apu, *apv, *apw;
    double *u, *v, *w, *su, *sv, *sw, *sp;
    int cn, bdf, buoy, boussinesq;
    double timestep, btime, densit, beta, tref, gravx, gravy, gravz;
};

// PISO_getHbyA.f90:41-96 (sources) and 98-198 (su, sv, sw += sum h_nb u_nb,
// with the Crank-Nicolson old-time terms), one thread per cell; u, v, w are
// only read here
__global__ __launch_bounds__(256) void k_hbya(Dim d, HbyaIO h)
{
    const long n = range_count(d, 3);
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        size_t c;
        range_cell(d, 3, q, c);
        double su = 0.0, sv = 0.0, sw = 0.0;
        h.sp[c] = 0.0;
        if (h.buoy) {
            double heat = 0.0;
            if (h.boussinesq) heat = h.beta * h.densit * (h.t[c] - h.tref) * h.vol[c];
            else heat = (h.densit - h.den[c]) * h.vol[c];
            su = su - h.gravx * heat;
            sv = sv - h.gravy * heat;
            sw = sw - h.gravz * heat;
        }
        if (h.bdf) {
            const double apotime = h.den[c] * h.vol[c] / h.timestep;
            const double sut = apotime * ((1.0 + h.btime) * h.uo[c] - 0.5 * h.btime * h.uoo[c]);
            const double svt = apotime * ((1.0 + h.btime) * h.vo[c] - 0.5 * h.btime * h.voo[c]);
            const double swt = apotime * ((1.0 + h.btime) * h.wo[c] - 0.5 * h.btime * h.woo[c]);
            su = su + sut;
            sv = sv + svt;
            sw = sw + swt;
        }
        const size_t e = c + d.nj, wv = c - d.nj, nn = c + 1, s = c - 1, t = c + d.nij, b = c - d.nij;
        double *S[3] = {&su, &sv, &sw};
        const double *X[3] = {h.u, h.v, h.w}, *XO[3] = {h.uo, h.vo, h.wo};
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            double acc = *S[m];
            if (h.cn) {
                const double *xo = XO[m];
                const double apotime = h.den[c] * h.vol[c] / h.timestep;
                acc = acc + (h.he[c] * xo[e] + h.hw[c] * xo[wv] + h.hn[c] * xo[nn] + h.hs[c] * xo[s] +
                             h.ht[c] * xo[t] + h.hb[c] * xo[b]) +
                      (apotime - h.he[c] - h.hw[c] - h.hn[c] - h.hs[c] - h.ht[c] - h.hb[c]) * xo[c];
            }
            const double *x = X[m];
            acc = acc + (h.he[c] * x[e] + h.hw[c] * x[wv] + h.hn[c] * x[nn] + h.hs[c] * x[s] +
                         h.ht[c] * x[t] + h.hb[c] * x[b]);
            *S[m] = acc;
        }
        h.su[c] = su;
        h.sv[c] = sv;
        h.sw[c] = sw;
    }
}

// PISO_getHbyA.f90:203-213: u = su * apu (v, w alike)
__global__ __launch_bounds__(256) void k_hbya_vel(Dim d, HbyaIO h)
{
    const long n = range_count(d, 3);
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        size_t c;
        range_cell(d, 3, q, c);
        h.u[c] = h.su[c] * h.apu[c];
        h.v[c] = h.sv[c] * h.apv[c];
        h.w[c] = h.sw[c] * h.apw[c];
    }
}

// ------------------------------------------------------------------ outbc
struct OutIO {
    const double *den, *ar1x, *ar1y, *ar1z;
    double *u, *v, *w, *f1, *f2, *f3;
    double *sc[9];            // t, te, ed, vis, vart, con, utt, vtt, wtt (NULL: skipped)
    const int *lb[6];         // lbw, lbe, lbs, lbn, lbb, lbt
    int on[6];                // low, loe, los, lon, lob, lot
};

// boundary cell number q of face f -> (inp, inh, the flux array, the face
// cell whose flux it is, the sign of its contribution to flow); false when
// the cell is not an outlet
struct OutCell {
    size_t inp, inh, fc;
    int arr;
    double sign;
};

__device__ bool out_cell(const Dim &d, const OutIO &o, int f, long q, OutCell &oc)
{
    const int nim = d.ni - 1, njm = d.nj - 1, nkm = d.nk - 1;
    int a, b, inbc;
    if (f < 2) {            // west / east: k = 2..nkm, j = 2..njm
        const int nkk = nkm - 1;
        a = 2 + (int)(q % nkk);     // k
        b = 2 + (int)(q / nkk);     // j
        inbc = (b - 1) * d.nk + a;
        if (o.lb[f][inbc - 1] != 2) return false;
        const int i = f == 0 ? 2 : nim;
        oc.inp = (size_t)(a - 1) * d.nij + (size_t)(i - 1) * d.nj + (size_t)(b - 1);
        oc.inh = f == 0 ? oc.inp + d.nj : oc.inp - d.nj;
        oc.fc = f == 0 ? oc.inp : oc.inh;
        oc.arr = 0;
    } else if (f < 4) {     // south / north: k = 2..nkm, i = 2..nim
        const int nkk = nkm - 1;
        a = 2 + (int)(q % nkk);     // k
        b = 2 + (int)(q / nkk);     // i
        inbc = (b - 1) * d.nk + a;
        if (o.lb[f][inbc - 1] != 2) return false;
        const int j = f == 2 ? 2 : njm;
        oc.inp = (size_t)(a - 1) * d.nij + (size_t)(b - 1) * d.nj + (size_t)(j - 1);
        oc.inh = f == 2 ? oc.inp + 1 : oc.inp - 1;
        oc.fc = f == 2 ? oc.inp : oc.inh;
        oc.arr = 1;
    } else {                // bottom / top: i = 2..nim, j = 2..njm
        const int njj = njm - 1;
        a = 2 + (int)(q % njj);     // j
        b = 2 + (int)(q / njj);     // i
        inbc = (b - 1) * d.nj + a;
        if (o.lb[f][inbc - 1] != 2) return false;
        const int k = f == 4 ? 2 : nkm;
        oc.inp = (size_t)(k - 1) * d.nij + (size_t)(b - 1) * d.nj + (size_t)(a - 1);
        oc.inh = f == 4 ? oc.inp + d.nij : oc.inp - d.nij;
        oc.fc = f == 4 ? oc.inp : oc.inh;
        oc.arr = 2;
    }
    oc.sign = (f & 1) ? 1.0 : -1.0;
    return true;
}

__host__ __device__ long face_cells(const Dim &d, int f)
{
    if (f < 2) return (long)(d.nk - 2) * (d.nj - 2);
    if (f < 4) return (long)(d.nk - 2) * (d.ni - 2);
    return (long)(d.ni - 2) * (d.nj - 2);
}

// outbc.f90:32-118: the outlet face fluxes of every face (they read only u,
// v, w, den and the areas) and, per boundary cell, its signed contribution
// to the outflow (0 when not an outlet), stored in face order
__global__ __launch_bounds__(256) void k_outbc_flux(Dim d, OutIO o, double *contrib)
{
    long base = 0;
    for (int f = 0; f < 6; ++f) {
        const long nf = face_cells(d, f);
        for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < nf; q += (long)gridDim.x * 256) {
            double cv = 0.0;
            OutCell oc;
            if (o.on[f] && out_cell(d, o, f, q, oc)) {
                double *fa = oc.arr == 0 ? o.f1 : oc.arr == 1 ? o.f2 : o.f3;
                const size_t inp = oc.inp, h = oc.inh;
                fa[oc.fc] = o.den[inp] * (o.ar1x[inp] * o.u[h] + o.ar1y[inp] * o.v[h] + o.ar1z[inp] * o.w[h]);
                cv = oc.sign * fa[oc.fc];
            }
            contrib[base + q] = cv;
        }
        base += nf;
    }
}

// flow = sum of the contributions (compensated, one block, fixed order);
// fac = flowin / (flow + small) (120)
__global__ __launch_bounds__(1024) void k_outbc_fac(long n, const double *contrib, double flowin,
                                                    double *fac)
{
    Acc s;
    for (long q = threadIdx.x; q < n; q += 1024) s += contrib[q];
    const Acc t = wave_sum(s);
    __shared__ double hh[16], ll[16];
    if ((threadIdx.x & 63) == 0) { hh[threadIdx.x >> 6] = t.hi; ll[threadIdx.x >> 6] = t.lo; }
    __syncthreads();
    if (threadIdx.x == 0) {
        Acc b;
        for (int w = 0; w < 16; ++w) { Acc x; x.hi = hh[w]; x.lo = ll[w]; b += x; }
        fac[0] = flowin / (b.value() + SMALL_F);
    }
}

// outbc.f90:127-318, one face per launch (a later face may read what an
// earlier one wrote): outlet velocities = inner values * fac, the outlet
// face flux * fac, the other fields copied
__global__ __launch_bounds__(256) void k_outbc_scale(Dim d, OutIO o, int f, const double *facp)
{
    const long nf = face_cells(d, f);
    const double fac = facp[0];
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < nf; q += (long)gridDim.x * 256) {
        OutCell oc;
        if (!out_cell(d, o, f, q, oc)) continue;
        const size_t inp = oc.inp, h = oc.inh;
        o.u[inp] = o.u[h] * fac;
        o.v[inp] = o.v[h] * fac;
        o.w[inp] = o.w[h] * fac;
        double *fa = oc.arr == 0 ? o.f1 : oc.arr == 1 ? o.f2 : o.f3;
        fa[oc.fc] = fa[oc.fc] * fac;
#pragma unroll
        for (int m = 0; m < 9; ++m)
            if (o.sc[m]) o.sc[m][inp] = o.sc[m][h];
    }
}

// correctBoundaryConditions.f90:19-84, one copy loop per launch in the
// reference's order: 0 far east (i = ni <- 2), 1 far west (i = 1 <- the
// plane li(njm), as written), 2 north (j = nj <- 2), 3 south (j = 1 <- njm);
// the west and south copies also carry f1 / f2
__global__ __launch_bounds__(256) void k_cbc(Dim d, int part, double *u, double *v, double *w,
                                             double *f1, double *f2)
{
    const int nkk = d.nk - 2;
    const long n = part < 2 ? (long)nkk * (d.nj - 2) : (long)nkk * (d.ni - 2);
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const int k = 2 + (int)(q % nkk), b = 2 + (int)(q / nkk);   // b: j or i
        const size_t lk = (size_t)(k - 1) * d.nij;
        size_t inp, inh;
        if (part == 0) { inp = lk + (size_t)(d.ni - 1) * d.nj + (b - 1); inh = lk + (size_t)1 * d.nj + (b - 1); }
        else if (part == 1) { inp = lk + (b - 1); inh = lk + (size_t)(d.nj - 2) * d.nj + (b - 1); }
        else if (part == 2) { inp = lk + (size_t)(b - 1) * d.nj + (d.nj - 1); inh = lk + (size_t)(b - 1) * d.nj + 1; }
        else { inp = lk + (size_t)(b - 1) * d.nj; inh = lk + (size_t)(b - 1) * d.nj + (d.nj - 2); }
        u[inp] = u[inh];
        v[inp] = v[inh];
        w[inp] = w[inh];
        if (part == 1) f1[inp] = f1[inh];
        if (part == 3) f2[inp] = f2[inh];
    }
}

// ------------------------------------------------------------------ host
Dim dim_of(const cap_ctx *ctx) { return Dim{ctx->ni, ctx->nj, ctx->nk, ctx->ni * ctx->nj}; }

// scratch of the drivers in the context's staging slots (device-pointer
// contexts never use those slots for staging)
int scratch(cap_ctx *ctx, int slot, size_t bytes, double **out)
{
    if (ctx->Gsz[slot] < bytes) {
        if (ctx->G[slot]) (void)hipFree(ctx->G[slot]);
        ctx->G[slot] = nullptr;
        ctx->Gsz[slot] = 0;
        HIPCHK(hipMalloc(&ctx->G[slot], bytes));
        ctx->Gsz[slot] = bytes;
    }
    *out = ctx->G[slot];
    return CAP_OK;
}

bool flow_ok(const cap_flow *f)
{
    const void *need[] = {f->vol, f->fx, f->fy, f->fz, f->ar1x, f->ar1y, f->ar1z, f->ar2x, f->ar2y,
                          f->ar2z, f->ar3x, f->ar3y, f->ar3z, f->u, f->v, f->w, f->p, f->pp, f->f1,
                          f->f2, f->f3, f->ae, f->aw, f->an, f->as, f->at, f->ab, f->ap, f->su,
                          f->gradp};
    for (const void *p : need) if (!p) return false;
    return true;
}

const double *const *geom13(const cap_flow *f, const double *buf[13])
{
    const double *g[13] = {f->fx, f->fy, f->fz, f->ar1x, f->ar1y, f->ar1z, f->ar2x, f->ar2y, f->ar2z,
                           f->ar3x, f->ar3y, f->ar3z, f->vol};
    for (int q = 0; q < 13; ++q) buf[q] = g[q];
    return buf;
}

cap_fields fields_of(const cap_flow *f)
{
    cap_fields F;
    F.x = f->x; F.y = f->y; F.z = f->z; F.xc = f->xc; F.yc = f->yc; F.zc = f->zc;
    F.vol = f->vol; F.fx = f->fx; F.fy = f->fy; F.fz = f->fz;
    F.ar1x = f->ar1x; F.ar1y = f->ar1y; F.ar1z = f->ar1z;
    F.ar2x = f->ar2x; F.ar2y = f->ar2y; F.ar2z = f->ar2z;
    F.ar3x = f->ar3x; F.ar3y = f->ar3y; F.ar3z = f->ar3z;
    F.den = f->den; F.apu = f->apu; F.apv = f->apv; F.apw = f->apw;
    F.u = f->u; F.v = f->v; F.w = f->w; F.p = f->p;
    F.gradu = f->gradu; F.gradv = f->gradv; F.gradw = f->gradw; F.gradp = f->gradp;
    return F;
}

cap_coef coef_of(const cap_flow *f)
{
    cap_coef c;
    c.ae = f->ae; c.aw = f->aw; c.an = f->an; c.as = f->as; c.at = f->at; c.ab = f->ab;
    c.ap = f->ap; c.su = f->su; c.res = f->res;
    c.bb = f->bb; c.bs = f->bs; c.bw = f->bw; c.bp = f->bp; c.bn = f->bn; c.be = f->be; c.bt = f->bt;
    return c;
}

// {ae, aw, an, as, at, ab} and {f1, f2, f3} of the flow state
double *const *a6_of(const cap_flow *f, double *a6[6])
{
    double *const a[6] = {f->ae, f->aw, f->an, f->as, f->at, f->ab};
    for (int q = 0; q < 6; ++q) a6[q] = a[q];
    return a6;
}
double *const *f3_of(const cap_flow *f, double *f3[3])
{
    f3[0] = f->f1; f3[1] = f->f2; f3[2] = f->f3;
    return f3;
}

int grad3(cap_ctx *ctx, const cap_flow *f, const double *u, double *g3)
{
    const double *gb[13];
    return cap_grad_gauss_aos(ctx, u, geom13(f, gb), g3);
}

int solve(cap_ctx *ctx, int solver, double *x, const cap_coef *c, const cap_drv_opts *o, double sor,
          cap_report *rep)
{
    switch (solver) {
    case 1: return cap_sipsol(ctx, x, c, o->alfa, sor, o->nsw, rep);
    case 2: return cap_cgstab(ctx, x, c, sor, o->nsw, rep);
    case 3: return cap_iccg(ctx, x, c, sor, o->nsw, rep);
    case 4: return cap_pcgsip(ctx, x, c, o->alfa, sor, o->nsw, rep);
    case 5: return cap_cgstab_sip(ctx, x, c, o->alfa, sor, o->nsw, rep);
    // 6 has no counterpart in the reference: pcgsip's system solved by FP64
    // refinement around FP32 PCG-SIP.  Its inner solves restart at each
    // refinement step, so they may take twice the FP64 iteration limit.
    case 6: return cap_pcgsip_mixed(ctx, x, c, o->alfa, sor, 2 * o->nsw, 3e-4, rep);
    // 7, no counterpart either: pcgsip with its preconditioner in FP32
    case 7: return cap_pcgsip_fcg(ctx, x, c, o->alfa, sor, o->nsw, rep);
    default: return CAP_ERR_ARG;
    }
}

int continuity(cap_ctx *ctx, const cap_flow *f, double *cum, double *out)
{
    double s[2];
    RCHK(cap_continuity(ctx, f->f1, f->f2, f->f3, s));
    *cum = *cum + s[0];
    out[0] = s[1];
    out[1] = s[0];
    out[2] = *cum;
    return CAP_OK;
}

bool dev_ok(const cap_ctx *ctx) { return ctx && (ctx->flags & CAP_DEVICE_POINTERS); }

}  // namespace

extern "C" int cap_outbc(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o)
{
    if (!dev_ok(ctx) || !f || !o) return CAP_ERR_ARG;
    if (!o->lout) return CAP_OK;
    const int on[6] = {o->low, o->loe, o->los, o->lon, o->lob, o->lot};
    const int *lb[6] = {o->lbw, o->lbe, o->lbs, o->lbn, o->lbb, o->lbt};
    if (!f->den || !f->ar1x || !f->ar1y || !f->ar1z || !f->u || !f->v || !f->w || !f->f1 || !f->f2 ||
        !f->f3)
        return CAP_ERR_ARG;
    for (int q = 0; q < 6; ++q) if (on[q] && !lb[q]) return CAP_ERR_ARG;
    const Dim d = dim_of(ctx);
    OutIO io;
    io.den = f->den; io.ar1x = f->ar1x; io.ar1y = f->ar1y; io.ar1z = f->ar1z;
    io.u = f->u; io.v = f->v; io.w = f->w; io.f1 = f->f1; io.f2 = f->f2; io.f3 = f->f3;
    double *sc[9] = {f->t, f->te, f->ed, f->vis, f->vart, f->con, f->utt, f->vtt, f->wtt};
    for (int m = 0; m < 9; ++m) io.sc[m] = sc[m];
    for (int q = 0; q < 6; ++q) { io.lb[q] = lb[q]; io.on[q] = on[q]; }
    long ntot = 0, nmax = 0;
    for (int q = 0; q < 6; ++q) { ntot += face_cells(d, q); nmax = std::max(nmax, face_cells(d, q)); }
    double *contrib, *fac;
    RCHK(scratch(ctx, 41, (size_t)ntot * 8, &contrib));
    RCHK(scratch(ctx, 42, 64, &fac));
    hipLaunchKernelGGL(k_outbc_flux, dim3(grid_for(nmax)), dim3(256), 0, ctx->stream, d, io, contrib);
    hipLaunchKernelGGL(k_outbc_fac, dim3(1), dim3(1024), 0, ctx->stream, ntot, (const double *)contrib,
                       o->flowin, fac);
    for (int q = 0; q < 6; ++q)
        if (on[q])
            hipLaunchKernelGGL(k_outbc_scale, dim3(grid_for(face_cells(d, q))), dim3(256), 0, ctx->stream,
                               d, io, q, (const double *)fac);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_correct_bc(cap_ctx *ctx, cap_flow *f)
{
    if (!dev_ok(ctx) || !f || !f->u || !f->v || !f->w || !f->f1 || !f->f2) return CAP_ERR_ARG;
    const Dim d = dim_of(ctx);
    // the far-west source plane li(njm) must lie inside the array
    if ((size_t)(d.nk - 2) * d.nij + (size_t)(d.nj - 2) * d.nj + (d.nj - 1) > ctx->nstd) return CAP_ERR_ARG;
    for (int part = 0; part < 4; ++part) {
        const long n = part < 2 ? (long)(d.nk - 2) * (d.nj - 2) : (long)(d.nk - 2) * (d.ni - 2);
        hipLaunchKernelGGL(k_cbc, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, part, f->u, f->v,
                           f->w, f->f1, f->f2);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_gethbya(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o)
{
    if (!dev_ok(ctx) || !f || !o || !flow_ok(f) || !f->sv || !f->sw || !f->sp || !f->apu ||
        !f->apv || !f->apw || !f->gradu || !f->gradv || !f->gradw || !f->he || !f->hw || !f->hn ||
        !f->hs || !f->ht || !f->hb)
        return CAP_ERR_ARG;
    if ((o->bdf || o->cn) && (!f->den || !f->uo || !f->vo || !f->wo)) return CAP_ERR_ARG;
    if (o->bdf && (!f->uoo || !f->voo || !f->woo)) return CAP_ERR_ARG;
    if (o->buoy && (!f->den || (o->boussinesq && !f->t))) return CAP_ERR_ARG;
    const Dim d = dim_of(ctx);
    HbyaIO h;
    h.den = f->den; h.vol = f->vol; h.t = f->t;
    h.uo = f->uo; h.vo = f->vo; h.wo = f->wo; h.uoo = f->uoo; h.voo = f->voo; h.woo = f->woo;
    h.he = f->he; h.hw = f->hw; h.hn = f->hn; h.hs = f->hs; h.ht = f->ht; h.hb = f->hb;
    h.apu = f->apu; h.apv = f->apv; h.apw = f->apw;
    h.u = f->u; h.v = f->v; h.w = f->w; h.su = f->su; h.sv = f->sv; h.sw = f->sw; h.sp = f->sp;
    h.cn = o->cn; h.bdf = o->bdf; h.buoy = o->buoy; h.boussinesq = o->boussinesq;
    h.timestep = o->timestep; h.btime = o->btime; h.densit = o->densit; h.beta = o->beta;
    h.tref = o->tref; h.gravx = o->gravx; h.gravy = o->gravy; h.gravz = o->gravz;
    const long n = range_count(d, 3);
    hipLaunchKernelGGL(k_hbya, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, h);
    hipLaunchKernelGGL(k_hbya_vel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, d, h);
    HIPCHK(hipGetLastError());
    RCHK(grad3(ctx, f, f->u, f->gradu));
    RCHK(grad3(ctx, f, f->v, f->gradv));
    RCHK(grad3(ctx, f, f->w, f->gradw));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_calcp(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o, cap_report *reps,
                         double *cont)
{
    if (!dev_ok(ctx) || !f || !o || !reps || !cont || !flow_ok(f) || !f->gradu || !f->gradv ||
        !f->gradw || !f->apu || !f->apv || !f->apw || !f->den)
        return CAP_ERR_ARG;
    if (o->ncorr < 1 || o->npcor < 1 || o->ijkpr < 1 || (size_t)o->ijkpr > ctx->nstd) return CAP_ERR_ARG;
    if (o->solver < 1 || o->solver > 7) return CAP_ERR_ARG;
    const Dim d = dim_of(ctx);
    const size_t n = ctx->nstd;
    const cap_fields F = fields_of(f);
    const cap_coef C = coef_of(f);
    const long nr = range_count(d, 3);
    double *fm = nullptr;
    if (o->npcor > 1) RCHK(scratch(ctx, 38, 3 * n * 8, &fm));
    double cum = o->cumulative;
    int worst = CAP_OK;
    for (int icorr = 1; icorr <= o->ncorr; ++icorr) {
        RCHK(grad3(ctx, f, f->u, f->gradu));
        RCHK(grad3(ctx, f, f->v, f->gradv));
        RCHK(grad3(ctx, f, f->w, f->gradw));
        // fluxmass for east, north and top (calcp-multiple_correction_SIMPLE.f90:50-60)
        double *a6[6], *f3[3];
        RCHK(cap_fluxmass3_q(ctx, &F, a6_of(f, a6), f3_of(f, f3)));
        if (!o->const_mflux) RCHK(cap_outbc(ctx, f, o));
        hipLaunchKernelGGL(k_su_ap, dim3(grid_for(nr)), dim3(256), 0, ctx->stream, d, 1, 1,
                           (const double *)f->f1, (const double *)f->f2, (const double *)f->f3,
                           (const double *)f->ae, (const double *)f->aw, (const double *)f->an,
                           (const double *)f->as, (const double *)f->at, (const double *)f->ab,
                           f->su, f->ap);
        HIPCHK(hipGetLastError());
        for (int ipcorr = 1; ipcorr <= o->npcor; ++ipcorr) {
            hipLaunchKernelGGL(k_zero, dim3(grid_for((long)n)), dim3(256), 0, ctx->stream, f->pp, n);
            cap_report *rep = &reps[(icorr - 1) * o->npcor + ipcorr - 1];
            const int rc = solve(ctx, o->solver, f->pp, &C, o, o->sor, rep);
            if (rc < 0) return rc;
            worst = std::max(worst, rc);
            RCHK(grad3(ctx, f, f->pp, f->gradp));
            if (ipcorr != o->npcor) {
                RCHK(cap_fluxmc(ctx, &F, d.nj, 1, d.nij, f->fx, fm));
                RCHK(cap_fluxmc(ctx, &F, 1, d.nij, d.nj, f->fy, fm + n));
                RCHK(cap_fluxmc(ctx, &F, d.nij, d.nj, 1, f->fz, fm + 2 * n));
                hipLaunchKernelGGL(k_fmc_gather, dim3(grid_for((long)n)), dim3(256), 0, ctx->stream, d,
                                   (const double *)fm, (const double *)(fm + n),
                                   (const double *)(fm + 2 * n), f->su, f->f1, f->f2, f->f3);
            }
            hipLaunchKernelGGL(k_corr_flux, dim3(grid_for(nr)), dim3(256), 0, ctx->stream, d, 3,
                               (const double *)f->pp, (const double *)f->ae, (const double *)f->an,
                               (const double *)f->at, f->f1, f->f2, f->f3);
            hipLaunchKernelGGL(k_corr_uvwp, dim3(grid_for(nr)), dim3(256), 0, ctx->stream, d, 0,
                               f->vol, (const double *)f->apu, (const double *)f->apv,
                               (const double *)f->apw, (const double *)f->gradp, f->u, f->v, f->w,
                               f->p, (const double *)f->pp, (size_t)o->ijkpr - 1, o->urf);
            HIPCHK(hipGetLastError());
        }
        RCHK(continuity(ctx, f, &cum, cont + 3 * (icorr - 1)));
        RCHK(cap_correct_bc(ctx, f));
    }
    RCHK(cap_bpres(ctx, f->p, f->fx, f->fy, f->fz));
    hipLaunchKernelGGL(k_copy, dim3(grid_for((long)n)), dim3(256), 0, ctx->stream, (const double *)f->p,
                       f->pp, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return worst;
}

extern "C" int cap_piso(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o, int pimple,
                        cap_report *reps, double *cont)
{
    return cap_piso_run(ctx, f, o, pimple, reps, cont, nullptr, nullptr);
}

int cap_piso_run(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o, int pimple, cap_report *reps,
                 double *cont, cap_start_hook hook, void *user)
{
    if (!dev_ok(ctx) || !f || !o || !reps || !cont || !flow_ok(f) || !f->he || !f->hw || !f->hn ||
        !f->hs || !f->ht || !f->hb || !f->be || !f->bn || !f->bt || !f->apu || !f->apv || !f->apw ||
        !f->den)
        return CAP_ERR_ARG;
    if (o->ncorr < 1 || o->npcor < 1 || o->ijkpr < 1 || (size_t)o->ijkpr > ctx->nstd) return CAP_ERR_ARG;
    if (pimple != 0 && pimple != 1) return CAP_ERR_ARG;
    const Dim d = dim_of(ctx);
    const size_t n = ctx->nstd;
    const long nr = range_count(d, 3);
    const int gn = grid_for((long)n);
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_copy, dim3(gn), dim3(256), 0, s, (const double *)f->ae, f->he, n);
    hipLaunchKernelGGL(k_copy, dim3(gn), dim3(256), 0, s, (const double *)f->aw, f->hw, n);
    hipLaunchKernelGGL(k_copy, dim3(gn), dim3(256), 0, s, (const double *)f->an, f->hn, n);
    hipLaunchKernelGGL(k_copy, dim3(gn), dim3(256), 0, s, (const double *)f->as, f->hs, n);
    hipLaunchKernelGGL(k_copy, dim3(gn), dim3(256), 0, s, (const double *)f->at, f->ht, n);
    hipLaunchKernelGGL(k_copy, dim3(gn), dim3(256), 0, s, (const double *)f->ab, f->hb, n);
    hipLaunchKernelGGL(k_zero, dim3(gn), dim3(256), 0, s, f->gradp, 3 * n);
    hipLaunchKernelGGL(k_zero, dim3(gn), dim3(256), 0, s, f->p, n);
    HIPCHK(hipGetLastError());
    double *x = pimple ? f->pp : f->p;
    const size_t inpr = (size_t)o->ijkpr - 1;
    const cap_coef C = coef_of(f);
    double cum = o->cumulative;
    int worst = CAP_OK;
    for (int icorr = 1; icorr <= o->ncorr; ++icorr) {
        if (!pimple) RCHK(cap_gethbya(ctx, f, o));
        for (int ipcorr = 1; ipcorr <= o->npcor; ++ipcorr) {
            const cap_fields F = fields_of(f);
            // piso_assemble_pressure_eq for east, north and top, bp = be, bn, bt
            double *a6[6], *f3[3];
            double *const bp3[3] = {f->be, f->bn, f->bt};
            RCHK(cap_piso_assemble3_q(ctx, &F, a6_of(f, a6), f3_of(f, f3), ipcorr, bp3, f->su));
            if (ipcorr == 1 && !o->const_mflux) RCHK(cap_outbc(ctx, f, o));
            hipLaunchKernelGGL(k_su_ap, dim3(grid_for(nr)), dim3(256), 0, s, d, ipcorr == 1 ? 1 : 0, 1,
                               (const double *)f->f1, (const double *)f->f2, (const double *)f->f3,
                               (const double *)f->ae, (const double *)f->aw, (const double *)f->an,
                               (const double *)f->as, (const double *)f->at, (const double *)f->ab,
                               f->su, f->ap);
            hipLaunchKernelGGL(k_pin, dim3(1), dim3(64), 0, s, inpr, f->ae, f->aw, f->an, f->as, f->at,
                               f->ab, f->ap, f->su, (const double *)f->p);
            HIPCHK(hipGetLastError());
            // sor(ip) = 0.1 on all but the last corrector: a default-real literal
            double sor = icorr != o->ncorr ? (double)0.1f : o->sor;
            cap_report *rep = &reps[(icorr - 1) * o->npcor + ipcorr - 1];
            if (hook && icorr == o->ncorr && ipcorr == o->npcor) {
                const int hr = hook(user, ctx, &C, x, &sor);
                if (hr < 0) return hr;
            }
            const int rc = solve(ctx, o->solver == 6 || o->solver == 7 ? o->solver : 4, x, &C, o,
                                 sor, rep);
            if (rc < 0) return rc;
            worst = std::max(worst, rc);
            RCHK(cap_bpres(ctx, x, f->fx, f->fy, f->fz));
            RCHK(grad3(ctx, f, x, f->gradp));
            if (ipcorr == o->npcor) {
                hipLaunchKernelGGL(k_corr_flux, dim3(grid_for(range_count(d, 2))), dim3(256), 0, s, d, 2,
                                   (const double *)x, (const double *)f->ae, (const double *)f->an,
                                   (const double *)f->at, f->f1, f->f2, f->f3);
                HIPCHK(hipGetLastError());
            }
        }
        RCHK(continuity(ctx, f, &cum, cont + 3 * (icorr - 1)));
        hipLaunchKernelGGL(k_corr_uvwp, dim3(grid_for(nr)), dim3(256), 0, s, d, pimple ? 2 : 1, f->vol,
                           (const double *)f->apu, (const double *)f->apv, (const double *)f->apw,
                           (const double *)f->gradp, f->u, f->v, f->w, f->p, (const double *)f->pp,
                           inpr, o->urf);
        HIPCHK(hipGetLastError());
        if (pimple) RCHK(cap_bpres(ctx, f->p, f->fx, f->fy, f->fz));
    }
    hipLaunchKernelGGL(k_copy, dim3(gn), dim3(256), 0, s, (const double *)f->p, f->pp, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return worst;
}

// ================================================================ calcuvw
namespace {

struct UvwIO {
    const double *den, *vol, *t, *uo, *vo, *wo, *uoo, *voo, *woo, *gradp, *xc, *yc, *zc, *fz;
    const double *ar3x, *ar3y, *ar3z, *vis;
    double *u, *v, *w, *su, *sv, *sw, *sp, *spv, *ap, *apu, *apv, *apw;
    double *ae, *aw, *an, *as, *at, *ab;
    int const_mflux, buoy, boussinesq, bdf, cn;
    double gradpcmf, timestep, btime, densit, beta, tref, gravx, gravy, gravz;
};

// calcuvw.f90:50-98: momentum sources on k = 2..nkm, i, j = 3..n?mm
__global__ __launch_bounds__(256) void k_uvw_src(Dim d, UvwIO q)
{
    const int nx = d.nj - 4, ny = d.ni - 4, nz = d.nk - 2;
    const long n = (long)nx * ny * nz;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int j = 3 + (int)(t % nx), i = 3 + (int)((t / nx) % ny), k = 2 + (int)(t / ((long)nx * ny));
        const size_t c = (size_t)(k - 1) * d.nij + (size_t)(i - 1) * d.nj + (size_t)(j - 1);
        double ap = 0.0, spv = 0.0, sp = 0.0;
        double su = -q.gradp[3 * c] * q.vol[c];
        double sv = -q.gradp[3 * c + 1] * q.vol[c];
        double sw = -q.gradp[3 * c + 2] * q.vol[c];
        if (q.const_mflux) su = su + q.gradpcmf * q.vol[c];
        if (q.buoy) {
            double heat = 0.0;
            if (q.boussinesq) heat = q.beta * q.densit * (q.t[c] - q.tref) * q.vol[c];
            else heat = (q.densit - q.den[c]) * q.vol[c];
            su = su - q.gravx * heat;
            sv = sv - q.gravy * heat;
            sw = sw - q.gravz * heat;
        }
        if (q.bdf) {
            const double apotime = q.den[c] * q.vol[c] / q.timestep;
            const double sut = apotime * ((1.0 + q.btime) * q.uo[c] - 0.5 * q.btime * q.uoo[c]);
            const double svt = apotime * ((1.0 + q.btime) * q.vo[c] - 0.5 * q.btime * q.voo[c]);
            const double swt = apotime * ((1.0 + q.btime) * q.wo[c] - 0.5 * q.btime * q.woo[c]);
            su = su + sut;
            sv = sv + svt;
            sw = sw + swt;
            ap = ap + apotime * (1.0 + 0.5 * q.btime);
            spv = spv + apotime * (1.0 + 0.5 * q.btime);
            sp = sp + apotime * (1.0 + 0.5 * q.btime);
        }
        q.ap[c] = ap;
        q.spv[c] = spv;
        q.sp[c] = sp;
        q.su[c] = su;
        q.sv[c] = sv;
        q.sw[c] = sw;
    }
}

// calcuvw.f90:160-310 (first-order branch): wall shear of the bottom wall
// (top = 0: cells on k = 3, ghosts below) or the top wall (top = 1: cells on
// the plane kp = nimm, as written, ghosts above), one thread per (i, j)
__global__ __launch_bounds__(256) void k_uvw_wall(Dim d, UvwIO q, int top, int kp, double dnx,
                                                  double dny, double dnz)
{
    const int nx = d.nj - 4, ny = d.ni - 4;
    const long n = (long)nx * ny;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int j = 3 + (int)(t % nx), i = 3 + (int)(t / nx);
        const size_t inbc = (size_t)(kp - 1) * d.nij + (size_t)(i - 1) * d.nj + (size_t)(j - 1);
        const size_t ingc = top ? inbc + d.nij : inbc - d.nij, iface = top ? inbc : ingc;
        const double fxp = top ? 1.0 - q.fz[inbc] : q.fz[ingc];
        const double fxe = 1.0 - fxp;
        const double gx = q.xc[inbc] * fxp + q.xc[ingc] * fxe;
        const double gy = q.yc[inbc] * fxp + q.yc[ingc] * fxe;
        const double gz = q.zc[inbc] * fxp + q.zc[ingc] * fxe;
        const double are = sqrt(q.ar3x[iface] * q.ar3x[iface] + q.ar3y[iface] * q.ar3y[iface] +
                                q.ar3z[iface] * q.ar3z[iface]);
        const double nxf = q.ar3x[iface] / are, nyf = q.ar3y[iface] / are, nzf = q.ar3z[iface] / are;
        const double dnw = fabs((q.xc[inbc] - gx) * dnx + (q.yc[inbc] - gy) * dny + (q.zc[inbc] - gz) * dnz);
        const double viss = q.vis[ingc];
        const double cf = viss * are / dnw;
        const double dxx = q.xc[inbc] - q.xc[ingc], dyy = q.yc[inbc] - q.yc[ingc],
                     dzz = q.zc[inbc] - q.zc[ingc];
        const double dpb = 0.5 * sqrt(dxx * dxx + dyy * dyy + dzz * dzz);
        const double vsol = viss * are / dpb;
        const double upb = q.u[inbc], vpb = q.v[inbc], wpb = q.w[inbc];
        const double vnp = upb * nxf + vpb * nyf + wpb * nzf;
        const double utp = upb - vnp * nxf, vtp = vpb - vnp * nyf, wtp = wpb - vnp * nzf;
        q.ap[inbc] = q.ap[inbc] + vsol;
        q.spv[inbc] = q.spv[inbc] + vsol;
        q.sp[inbc] = q.sp[inbc] + vsol;
        q.su[inbc] = q.su[inbc] + vsol * q.u[inbc] - cf * utp;
        q.sv[inbc] = q.sv[inbc] + vsol * q.v[inbc] - cf * vtp;
        q.sw[inbc] = q.sw[inbc] + vsol * q.w[inbc] - cf * wtp;
    }
}

// Crank-Nicolson: a_nb = 0.5 a_nb on the whole arrays (calcuvw.f90:313-320)
__global__ void k_half6(size_t n, double *ae, double *aw, double *an, double *as, double *at, double *ab)
{
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        ae[c] = 0.5 * ae[c];
        an[c] = 0.5 * an[c];
        at[c] = 0.5 * at[c];
        aw[c] = 0.5 * aw[c];
        as[c] = 0.5 * as[c];
        ab[c] = 0.5 * ab[c];
    }
}

// per component m (u, v, w): calcuvw.f90:324-345 / 350-370 / 376-396
__global__ __launch_bounds__(256) void k_uvw_asm(Dim d, UvwIO q, int m, double urfr, double urfm)
{
    const int nx = d.nj - 4, ny = d.ni - 4, nz = d.nk - 2;
    const long n = (long)nx * ny * nz;
    double *S = m == 0 ? q.su : m == 1 ? q.sv : q.sw;
    double *D = m == 0 ? q.ap : m == 1 ? q.spv : q.sp;
    const double *X = m == 0 ? q.u : m == 1 ? q.v : q.w;
    const double *XO = m == 0 ? q.uo : m == 1 ? q.vo : q.wo;
    double *AP = m == 0 ? q.apu : m == 1 ? q.apv : q.apw;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int j = 3 + (int)(t % nx), i = 3 + (int)((t / nx) % ny), k = 2 + (int)(t / ((long)nx * ny));
        const size_t c = (size_t)(k - 1) * d.nij + (size_t)(i - 1) * d.nj + (size_t)(j - 1);
        double s = S[c];
        if (q.cn) {
            const double apotime = q.den[c] * q.vol[c] / q.timestep;
            s = s + (q.ae[c] * XO[c + d.nj] + q.aw[c] * XO[c - d.nj] + q.an[c] * XO[c + 1] +
                     q.as[c] * XO[c - 1] + q.at[c] * XO[c + d.nij] + q.ab[c] * XO[c - d.nij]) +
                (apotime - q.ae[c] - q.aw[c] - q.an[c] - q.as[c] - q.at[c] - q.ab[c]) * XO[c];
            S[c] = s;
            D[c] = D[c] + apotime;
        }
        double ap = q.ae[c] + q.aw[c] + q.an[c] + q.as[c] + q.at[c] + q.ab[c] + D[c];
        ap = ap * urfr;
        q.ap[c] = ap;
        q.su[c] = s + urfm * ap * X[c];
        AP[c] = 1.0 / (ap + SMALL_F);
    }
}

// find_strain_rate.f90:36-77
__global__ __launch_bounds__(256) void k_strain(Dim d, const double *gu, const double *gv,
                                                const double *gw, double *strain, double *vort)
{
    const int nx = d.nj - 4, ny = d.ni - 4, nz = d.nk - 2;
    const long n = (long)nx * ny * nz;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int j = 3 + (int)(t % nx), i = 3 + (int)((t / nx) % ny), k = 2 + (int)(t / ((long)nx * ny));
        const size_t c = (size_t)(k - 1) * d.nij + (size_t)(i - 1) * d.nj + (size_t)(j - 1);
        const double dudx = gu[3 * c], dudy = gu[3 * c + 1], dudz = gu[3 * c + 2];
        const double dvdx = gv[3 * c], dvdy = gv[3 * c + 1], dvdz = gv[3 * c + 2];
        const double dwdx = gw[3 * c], dwdy = gw[3 * c + 1], dwdz = gw[3 * c + 2];
        const double s11 = dudx, s12 = 0.5 * (dudy + dvdx), s13 = 0.5 * (dudz + dwdx);
        const double s22 = dvdy, s23 = 0.5 * (dvdz + dwdy), s33 = dwdz;
        const double w12 = 0.5 * (dudy - dvdx), w13 = 0.5 * (dudz - dwdx), w23 = 0.5 * (dvdz - dwdy);
        strain[c] = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 +
                                2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
        vort[c] = sqrt(w12 * w12 + w23 * w23 + w13 * w13);
    }
}

struct LesIO {
    const double *u, *v, *w, *vol, *den, *fz, *xc, *yc, *zc, *ar3x, *ar3y, *ar3z, *strain, *wd;
    double *vis;
    double viscos, densit, cappa, elog, urf;
};

// calc_vis_les.f90:45-77: Smagorinsky with van Driest damping; default-real
// literals C_delta = 0.158, C_Les = 0.17 and the exponent 1./3.
__global__ __launch_bounds__(256) void k_les(Dim d, LesIO q)
{
    const double c_delta = (double)0.158f, c_les = (double)0.17f, third = (double)(1.0f / 3.0f);
    const int nx = d.nj - 4, ny = d.ni - 4, nz = d.nk - 2;
    const long n = (long)nx * ny * nz;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int j = 3 + (int)(t % nx), i = 3 + (int)((t / nx) % ny), k = 2 + (int)(t / ((long)nx * ny));
        const size_t c = (size_t)(k - 1) * d.nij + (size_t)(i - 1) * d.nj + (size_t)(j - 1);
        const double visold = q.vis[c];
        const double wl = q.wd[c];
        const double uref = sqrt(q.u[c] * q.u[c] + q.v[c] * q.v[c] + q.w[c] * q.w[c]);
        const double utau = sqrt(q.viscos * uref / wl);
        const double zplus = wl * utau / q.viscos;
        const double a = c_les * pow(q.vol[c], third), b = q.cappa / c_delta * wl;
        const double l = (b < a ? b : a) * (1.0 - exp(-zplus / 26.0));
        const double vs = l * l * q.strain[c] * q.den[c];
        const double vn = vs + q.viscos;
        q.vis[c] = q.urf * vn + (1.0 - q.urf) * visold;
    }
}

// calc_vis_les.f90:80-151: wall-function viscosity of the wall ghost cells
// (top = 0: bottom wall, cells on k = 3; top = 1: cells on the plane kp)
__global__ __launch_bounds__(256) void k_les_wall(Dim d, LesIO q, int top, int kp, double dnx,
                                                  double dny, double dnz)
{
    const int nx = d.nj - 4, ny = d.ni - 4;
    const long n = (long)nx * ny;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int j = 3 + (int)(t % nx), i = 3 + (int)(t / nx);
        const size_t inbc = (size_t)(kp - 1) * d.nij + (size_t)(i - 1) * d.nj + (size_t)(j - 1);
        const size_t ingc = top ? inbc + d.nij : inbc - d.nij, iface = top ? inbc : ingc;
        const double fxp = top ? 1.0 - q.fz[inbc] : q.fz[ingc];
        const double fxe = 1.0 - fxp;
        const double gx = q.xc[inbc] * fxp + q.xc[ingc] * fxe;
        const double gy = q.yc[inbc] * fxp + q.yc[ingc] * fxe;
        const double gz = q.zc[inbc] * fxp + q.zc[ingc] * fxe;
        const double are = sqrt(q.ar3x[iface] * q.ar3x[iface] + q.ar3y[iface] * q.ar3y[iface] +
                                q.ar3z[iface] * q.ar3z[iface]);
        const double nxf = q.ar3x[iface] / are, nyf = q.ar3y[iface] / are, nzf = q.ar3z[iface] / are;
        const double uu = q.u[inbc], vv = q.v[inbc], ww = q.w[inbc];
        const double vnp = uu * nxf + vv * nyf + ww * nzf;
        double xtp = uu - vnp * nxf, ytp = vv - vnp * nyf, ztp = ww - vnp * nzf;
        const double vtp = sqrt(xtp * xtp + ytp * ytp + ztp * ztp);
        xtp = xtp / vtp;
        ytp = ytp / vtp;
        ztp = ztp / vtp;
        const double ut2 = fabs(uu * xtp + vv * ytp + ww * ztp);
        const double dnw = fabs((q.xc[inbc] - gx) * dnx + (q.yc[inbc] - gy) * dny + (q.zc[inbc] - gz) * dnz);
        const double tw = q.viscos * ut2 / dnw;
        const double utau = sqrt(tw / q.densit);
        const double ypl = q.densit * utau * dnw / q.viscos;
        double viscw = 0.0;
        if (ypl > (double)11.61f) viscw = ypl * q.viscos * q.cappa / log(q.elog * ypl);
        q.vis[ingc] = viscw > q.viscos ? viscw : q.viscos;
    }
}

UvwIO uvw_io(cap_flow *f, double *spv, const cap_drv_opts *o, const cap_uvw_opts *q)
{
    UvwIO u;
    u.den = f->den; u.vol = f->vol; u.t = f->t; u.uo = f->uo; u.vo = f->vo; u.wo = f->wo;
    u.uoo = f->uoo; u.voo = f->voo; u.woo = f->woo; u.gradp = f->gradp;
    u.xc = f->xc; u.yc = f->yc; u.zc = f->zc; u.fz = f->fz;
    u.ar3x = f->ar3x; u.ar3y = f->ar3y; u.ar3z = f->ar3z; u.vis = f->vis;
    u.u = f->u; u.v = f->v; u.w = f->w; u.su = f->su; u.sv = f->sv; u.sw = f->sw; u.sp = f->sp;
    u.spv = spv; u.ap = f->ap; u.apu = f->apu; u.apv = f->apv; u.apw = f->apw;
    u.ae = f->ae; u.aw = f->aw; u.an = f->an; u.as = f->as; u.at = f->at; u.ab = f->ab;
    u.const_mflux = o->const_mflux; u.buoy = o->buoy; u.boussinesq = o->boussinesq;
    u.bdf = o->bdf; u.cn = o->cn;
    u.gradpcmf = q->gradpcmf; u.timestep = o->timestep; u.btime = o->btime; u.densit = o->densit;
    u.beta = o->beta; u.tref = o->tref; u.gravx = o->gravx; u.gravy = o->gravy; u.gravz = o->gravz;
    return u;
}

long plane_cells(const Dim &d) { return (long)(d.nj - 4) * (d.ni - 4); }
long range2_cells(const Dim &d) { return (long)(d.nj - 4) * (d.ni - 4) * (d.nk - 2); }

}  // namespace

extern "C" int cap_calcuvw(cap_ctx *ctx, cap_flow *f, double *spv, const cap_drv_opts *o,
                           const cap_uvw_opts *q, cap_report *reps)
{
    if (!dev_ok(ctx) || !f || !spv || !o || !q || !reps || !flow_ok(f) || !f->sv || !f->sw ||
        !f->sp || !f->apu || !f->apv || !f->apw || !f->gradu || !f->gradv || !f->gradw ||
        !f->vis || !f->den || !f->xc || !f->yc || !f->zc || !f->bp || !f->bt || !f->bb)
        return CAP_ERR_ARG;
    if ((o->bdf || o->cn) && (!f->uo || !f->vo || !f->wo)) return CAP_ERR_ARG;
    if (o->bdf && (!f->uoo || !f->voo || !f->woo)) return CAP_ERR_ARG;
    if (o->buoy && o->boussinesq && !f->t) return CAP_ERR_ARG;
    const Dim d = dim_of(ctx);
    // the top wall's cells: the plane lk(nimm) as written (it and its ghosts
    // must lie in the array), or the plane top_k
    if (q->top_k == 0 ? d.ni - 1 > d.nk : (q->top_k < 3 || q->top_k > d.nk - 2)) return CAP_ERR_ARG;
    const int top_plane = q->top_k ? q->top_k : d.ni - 2;
    hipStream_t s = ctx->stream;
    RCHK(grad3(ctx, f, f->u, f->gradu));
    RCHK(grad3(ctx, f, f->v, f->gradv));
    RCHK(grad3(ctx, f, f->w, f->gradw));
    RCHK(grad3(ctx, f, f->pp, f->gradp));
    const UvwIO io = uvw_io(f, spv, o, q);
    hipLaunchKernelGGL(k_uvw_src, dim3(grid_for(range2_cells(d))), dim3(256), 0, s, d, io);
    HIPCHK(hipGetLastError());
    const cap_fields F = fields_of(f);
    if (q->schemes == CAP_UVW_CDS) {   // the three directions in one pass
        double *a6[6], *f3[3];
        double *const s3[3] = {f->su, f->sv, f->sw}, *const scr3[3] = {f->bp, f->bt, f->bb};
        RCHK(cap_fluxuvw3_q(ctx, &F, f->vis, a6_of(f, a6), f3_of(f, f3), s3, scr3, q->gds));
    } else {
        RCHK(cap_fluxuvw(ctx, &F, f->vis, d.nj, 1, d.nij, f->ar1x, f->ar1y, f->ar1z, f->fx, f->ae,
                         f->aw, f->f1, f->su, f->sv, f->sw, f->bp, f->bt, f->bb, q->schemes, q->gds));
        RCHK(cap_fluxuvw(ctx, &F, f->vis, 1, d.nij, d.nj, f->ar2x, f->ar2y, f->ar2z, f->fy, f->an,
                         f->as, f->f2, f->su, f->sv, f->sw, f->bp, f->bt, f->bb, q->schemes, q->gds));
        RCHK(cap_fluxuvw(ctx, &F, f->vis, d.nij, d.nj, 1, f->ar3x, f->ar3y, f->ar3z, f->fz, f->at,
                         f->ab, f->f3, f->su, f->sv, f->sw, f->bp, f->bt, f->bb, q->schemes, q->gds));
    }
    hipLaunchKernelGGL(k_uvw_wall, dim3(grid_for(plane_cells(d))), dim3(256), 0, s, d, io, 0, 3,
                       (double)q->nx, (double)q->ny, (double)q->nz);
    hipLaunchKernelGGL(k_uvw_wall, dim3(grid_for(plane_cells(d))), dim3(256), 0, s, d, io, 1, top_plane,
                       (double)q->nx, (double)q->ny, (double)q->nz);
    if (o->cn)
        hipLaunchKernelGGL(k_half6, dim3(grid_for((long)ctx->nstd)), dim3(256), 0, s, ctx->nstd, f->ae,
                           f->aw, f->an, f->as, f->at, f->ab);
    HIPCHK(hipGetLastError());
    const cap_coef C = coef_of(f);
    double *X[3] = {f->u, f->v, f->w};
    int worst = CAP_OK;
    for (int m = 0; m < 3; ++m) {
        hipLaunchKernelGGL(k_uvw_asm, dim3(grid_for(range2_cells(d))), dim3(256), 0, s, d, io, m,
                           q->urfr[m], q->urfm[m]);
        HIPCHK(hipGetLastError());
        const int rc = cap_sipsol(ctx, X[m], &C, o->alfa, q->sor[m], q->nsw[m], &reps[m]);
        if (rc < 0) return rc;
        worst = std::max(worst, rc);
    }
    HIPCHK(hipStreamSynchronize(s));
    return worst;
}

extern "C" int cap_find_strain_rate(cap_ctx *ctx, cap_flow *f, double *strain, double *vorticity)
{
    if (!dev_ok(ctx) || !f || !strain || !vorticity || !f->u || !f->v || !f->w || !f->gradu ||
        !f->gradv || !f->gradw || !f->vol || !f->fx || !f->ar1x)
        return CAP_ERR_ARG;
    const Dim d = dim_of(ctx);
    RCHK(grad3(ctx, f, f->u, f->gradu));
    RCHK(grad3(ctx, f, f->v, f->gradv));
    RCHK(grad3(ctx, f, f->w, f->gradw));
    hipLaunchKernelGGL(k_strain, dim3(grid_for(range2_cells(d))), dim3(256), 0, ctx->stream, d,
                       (const double *)f->gradu, (const double *)f->gradv, (const double *)f->gradw,
                       strain, vorticity);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_calc_vis_les(cap_ctx *ctx, cap_flow *f, const double *strain,
                                const double *walldist, const cap_les_opts *o)
{
    if (!dev_ok(ctx) || !f || !strain || !walldist || !o || !f->u || !f->v || !f->w || !f->vol ||
        !f->den || !f->fz || !f->xc || !f->yc || !f->zc || !f->ar3x || !f->ar3y || !f->ar3z || !f->vis)
        return CAP_ERR_ARG;
    const Dim d = dim_of(ctx);
    if (o->top_k == 0 ? d.ni - 1 > d.nk : (o->top_k < 3 || o->top_k > d.nk - 2)) return CAP_ERR_ARG;
    const int top_plane = o->top_k ? o->top_k : d.ni - 2;
    LesIO q;
    q.u = f->u; q.v = f->v; q.w = f->w; q.vol = f->vol; q.den = f->den; q.fz = f->fz;
    q.xc = f->xc; q.yc = f->yc; q.zc = f->zc; q.ar3x = f->ar3x; q.ar3y = f->ar3y; q.ar3z = f->ar3z;
    q.strain = strain; q.wd = walldist; q.vis = f->vis;
    q.viscos = o->viscos; q.densit = o->densit; q.cappa = o->cappa; q.elog = o->elog; q.urf = o->urfvis;
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_les, dim3(grid_for(range2_cells(d))), dim3(256), 0, s, d, q);
    hipLaunchKernelGGL(k_les_wall, dim3(grid_for(plane_cells(d))), dim3(256), 0, s, d, q, 0, 3,
                       (double)o->nx, (double)o->ny, (double)o->nz);
    hipLaunchKernelGGL(k_les_wall, dim3(grid_for(plane_cells(d))), dim3(256), 0, s, d, q, 1, top_plane,
                       (double)o->nx, (double)o->ny, (double)o->nz);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return CAP_OK;
}
