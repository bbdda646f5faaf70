// cap_stencil.hip -- p'-path stencils on the reference layout (GRAD_GAUSS,
// bpres) and the problem generator of the measured workload.
//
// The stencils work directly on the padded j-fastest arrays: a thread per
// cell, consecutive threads along j (unit stride), so every wavefront access
// is coalesced.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <vector>
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

struct Geo {
    const double *fx, *fy, *fz, *a1x, *a1y, *a1z, *a2x, *a2y, *a2z, *a3x, *a3y, *a3z, *vol;
};

// Green-Gauss gradient in gather form (GRAD_GAUSS.f90:37-102).  The
// reference scatters face terms in loop order (k, i, j); the value that
// reaches cell P is ((((0 - t_B) - e_W) - n_S) + own faces), B = k-1,
// W = i-1, S = j-1, each term present only when that face-loop iteration is
// in the loop range k=2..nkm, i=2..nimm, j=2..njmm.  The gather below adds
// the same terms in the same order, then scales by volr = 1/vol on
// k=2..nkm, i,j=3..n?mm (88-102).
__global__ __launch_bounds__(256) void k_grad_gauss(int ni, int nj, int nk,
                                                    const double *__restrict__ u, Geo g,
                                                    double *__restrict__ dx,
                                                    double *__restrict__ dy,
                                                    double *__restrict__ dz)
{
    const int nij = ni * nj;
    const size_t n = (size_t)nij * nk;
    const int nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        const int k = (int)(c / nij) + 1;
        const int rem = (int)(c % nij);
        const int i = rem / nj + 1, j = rem % nj + 1;
        auto in_face = [&](int ii, int jj, int kk) {
            return kk >= 2 && kk <= nkm && ii >= 2 && ii <= nimm && jj >= 2 && jj <= njmm;
        };
        double sx = 0.0, sy = 0.0, sz = 0.0;
        if (in_face(i, j, k - 1)) {            // top face of B
            const size_t b = c - nij;
            const double ut = u[c] * g.fz[b] + u[b] * (1.0 - g.fz[b]);
            sx = sx - (ut * g.a3x[b]);
            sy = sy - (ut * g.a3y[b]);
            sz = sz - (ut * g.a3z[b]);
        }
        if (in_face(i - 1, j, k)) {            // east face of W
            const size_t w = c - nj;
            const double ue = u[c] * g.fx[w] + u[w] * (1.0 - g.fx[w]);
            sx = sx - (ue * g.a1x[w]);
            sy = sy - (ue * g.a1y[w]);
            sz = sz - (ue * g.a1z[w]);
        }
        if (in_face(i, j - 1, k)) {            // north face of S
            const size_t s = c - 1;
            const double un = u[c] * g.fy[s] + u[s] * (1.0 - g.fy[s]);
            sx = sx - (un * g.a2x[s]);
            sy = sy - (un * g.a2y[s]);
            sz = sz - (un * g.a2z[s]);
        }
        if (in_face(i, j, k)) {                // own east/north/top faces
            const double ue = u[c + nj] * g.fx[c] + u[c] * (1.0 - g.fx[c]);
            const double un = u[c + 1] * g.fy[c] + u[c] * (1.0 - g.fy[c]);
            const double ut = u[c + nij] * g.fz[c] + u[c] * (1.0 - g.fz[c]);
            sx = sx + (ue * g.a1x[c] + un * g.a2x[c] + ut * g.a3x[c]);
            sy = sy + (ue * g.a1y[c] + un * g.a2y[c] + ut * g.a3y[c]);
            sz = sz + (ue * g.a1z[c] + un * g.a2z[c] + ut * g.a3z[c]);
        }
        if (k >= 2 && k <= nkm && i >= 3 && i <= nimm && j >= 3 && j <= njmm) {
            const double volr = 1.0 / g.vol[c];
            sx = sx * volr;
            sy = sy * volr;
            sz = sz * volr;
        }
        dx[c] = sx;
        dy[c] = sy;
        dz[c] = sz;
    }
}

// bpres.f90:29-79 -- one thread per ghost cell of the six planes
__global__ __launch_bounds__(256) void k_bpres(int ni, int nj, int nk, double *__restrict__ p,
                                               const double *__restrict__ fx,
                                               const double *__restrict__ fy,
                                               const double *__restrict__ fz)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const int NI = nimm - 2, NJ = njmm - 2, NK = nkmm - 2;   // interior counts
    const int nb = NI * NJ, nw = NJ * NK, ns = NI * NK;
    const int tot = 2 * nb + 2 * nw + 2 * ns;
    auto IX = [&](int i, int j, int k) {
        return (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
    };
    for (int q = blockIdx.x * 256 + threadIdx.x; q < tot; q += gridDim.x * 256) {
        int r = q;
        if (r < nb) {                                   // bottom
            int i = 3 + r / NJ, j = 3 + r % NJ;
            size_t c = IX(i, j, 3);
            p[c - nij] = p[c] - (p[c + nij] - p[c]) * fz[c];
            continue;
        }
        r -= nb;
        if (r < nb) {                                   // top
            int i = 3 + r / NJ, j = 3 + r % NJ;
            size_t c = IX(i, j, nkmm);
            p[c + nij] = p[c] + (p[c] - p[c - nij]) * (1.0 - fz[c - nij]);
            continue;
        }
        r -= nb;
        if (r < nw) {                                   // west
            int j = 3 + r / NK, k = 3 + r % NK;
            size_t c = IX(3, j, k);
            p[c - nj] = p[c] - (p[c + nj] - p[c]) * fx[c];
            continue;
        }
        r -= nw;
        if (r < nw) {                                   // east
            int j = 3 + r / NK, k = 3 + r % NK;
            size_t c = IX(nimm, j, k);
            p[c + nj] = p[c] + (p[c] - p[c - nj]) * (1.0 - fx[c - nj]);
            continue;
        }
        r -= nw;
        if (r < ns) {                                   // south
            int i = 3 + r / NK, k = 3 + r % NK;
            size_t c = IX(i, 3, k);
            p[c - 1] = p[c] - (p[c + 1] - p[c]) * fy[c];
            continue;
        }
        r -= ns;
        {                                               // north
            int i = 3 + r / NK, k = 3 + r % NK;
            size_t c = IX(i, njmm, k);
            p[c + 1] = p[c] + (p[c] - p[c - 1]) * (1.0 - fy[c - 1]);
        }
    }
}

static int stage(cap_ctx *ctx, int slot, const double *p, const double **out)
{
    if (ctx->flags & CAP_DEVICE_POINTERS) { *out = p; return CAP_OK; }
    if (!ctx->G[slot]) HIPCHK(hipMalloc(&ctx->G[slot], ctx->nstd * sizeof(double)));
    if (p) HIPCHK(hipMemcpyAsync(ctx->G[slot], p, ctx->nstd * 8, hipMemcpyHostToDevice, ctx->stream));
    *out = ctx->G[slot];
    return CAP_OK;
}

extern "C" int cap_grad_gauss(cap_ctx *ctx, const double *u, const double *const *geom,
                              double *dudx, double *dudy, double *dudz)
{
    if (!ctx || !u || !geom || !dudx || !dudy || !dudz) return CAP_ERR_ARG;
    for (int q = 0; q < 13; ++q) if (!geom[q]) return CAP_ERR_ARG;
    const double *du, *gp[13];
    int rc;
    if ((rc = stage(ctx, 0, u, &du))) return rc;
    for (int q = 0; q < 13; ++q) if ((rc = stage(ctx, 1 + q, geom[q], &gp[q]))) return rc;
    Geo g = {gp[0], gp[1], gp[2], gp[3], gp[4], gp[5], gp[6], gp[7], gp[8], gp[9], gp[10], gp[11], gp[12]};
    double *ox = dudx, *oy = dudy, *zbuf = dudz;
    if (!(ctx->flags & CAP_DEVICE_POINTERS)) {
        const double *tmp;
        if ((rc = stage(ctx, 14, nullptr, &tmp))) return rc;
        if ((rc = stage(ctx, 15, nullptr, &tmp))) return rc;
        if ((rc = stage(ctx, 16, nullptr, &tmp))) return rc;
        ox = ctx->G[14];
        oy = ctx->G[15];
        zbuf = ctx->G[16];
    }
    int blocks = (int)std::min<size_t>((ctx->nstd + 255) / 256, 8192);
    hipLaunchKernelGGL(k_grad_gauss, dim3(blocks), dim3(256), 0, ctx->stream, ctx->ni, ctx->nj,
                       ctx->nk, du, g, ox, oy, zbuf);
    HIPCHK(hipGetLastError());
    if (!(ctx->flags & CAP_DEVICE_POINTERS)) {
        HIPCHK(hipMemcpyAsync(dudx, ox, ctx->nstd * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(dudy, oy, ctx->nstd * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(dudz, zbuf, ctx->nstd * 8, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_bpres(cap_ctx *ctx, double *p, const double *fx, const double *fy,
                         const double *fz)
{
    if (!ctx || !p || !fx || !fy || !fz) return CAP_ERR_ARG;
    const double *dp, *dfx, *dfy, *dfz;
    int rc;
    if ((rc = stage(ctx, 0, p, &dp))) return rc;
    if ((rc = stage(ctx, 1, fx, &dfx))) return rc;
    if ((rc = stage(ctx, 2, fy, &dfy))) return rc;
    if ((rc = stage(ctx, 3, fz, &dfz))) return rc;
    const int NI = ctx->ni - 4, NJ = ctx->nj - 4, NK = ctx->nk - 4;
    const int tot = 2 * (NI * NJ + NJ * NK + NI * NK);
    int blocks = std::max(1, std::min((tot + 255) / 256, 4096));
    hipLaunchKernelGGL(k_bpres, dim3(blocks), dim3(256), 0, ctx->stream, ctx->ni, ctx->nj, ctx->nk,
                       (double *)dp, dfx, dfy, dfz);
    HIPCHK(hipGetLastError());
    if (!(ctx->flags & CAP_DEVICE_POINTERS))
        HIPCHK(hipMemcpyAsync(p, dp, ctx->nstd * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

// ------------------------------------------------------ problem generator
// Counter-based generator (splitmix64 finaliser keyed by seed, stream and
// cell index): the same inputs on every machine, independent of how the
// work is split.
static inline uint64_t smix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline uint64_t hkey(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return smix(smix(seed ^ (stream * 0xD1B54A32D192ED03ull)) + idx);
}
static inline double gauss(uint64_t seed, uint64_t stream, uint64_t idx)
{
    double u1 = (double)((hkey(seed, stream, 2 * idx) >> 11) + 1) * 0x1.0p-53;
    double u2 = (double)(hkey(seed, stream, 2 * idx + 1) >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Log-normal Neumann p' operator on a uniform grid: interior faces
// a = h exp(sigma g), g ~ N(0,1) keyed by (seed, direction, cell); aw(E)=ae(P)
// etc.; boundary faces 0; ap = sum a_nb; su ~ N(0,1) with zero mean; ghost
// rows a* = 0, ap = 1, su = 0; reference-pressure cell (3,3,3) pinned.
extern "C" int cap_build_lognormal(cap_ctx *ctx, unsigned long long seed, double sigma,
                                   cap_coef *c)
{
    if (!ctx || !c || !c->ae || !c->aw || !c->an || !c->as || !c->at || !c->ab || !c->ap || !c->su)
        return CAP_ERR_ARG;
    const int ni = ctx->ni, nj = ctx->nj, nk = ctx->nk, nij = ni * nj;
    const size_t n = ctx->nstd;
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const double h = 1.0 / (double)(ni - 4);
    std::vector<double> A[8];
    const bool dev = ctx->flags & CAP_DEVICE_POINTERS;
    double *out[8] = {c->ae, c->aw, c->an, c->as, c->at, c->ab, c->ap, c->su};
    double *P[8];
    for (int q = 0; q < 8; ++q) {
        if (dev) { A[q].assign(n, 0.0); P[q] = A[q].data(); }
        else { P[q] = out[q]; memset(P[q], 0, n * sizeof(double)); }
    }
    double *ae = P[0], *aw = P[1], *an = P[2], *as = P[3], *at = P[4], *ab = P[5], *ap = P[6], *su = P[7];
    for (size_t q = 0; q < n; ++q) ap[q] = 1.0;
    auto IX = [&](int i, int j, int k) {
        return (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
    };
    for (int k = 3; k <= nkmm; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t q = IX(i, j, k);
                if (i < nimm) { double a = h * exp(sigma * gauss(seed, 1, q)); ae[q] = a; aw[q + nj] = a; }
                if (j < njmm) { double a = h * exp(sigma * gauss(seed, 2, q)); an[q] = a; as[q + 1] = a; }
                if (k < nkmm) { double a = h * exp(sigma * gauss(seed, 3, q)); at[q] = a; ab[q + nij] = a; }
            }
    double sum = 0.0;
    long cnt = 0;
    for (int k = 3; k <= nkmm; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t q = IX(i, j, k);
                ap[q] = ae[q] + aw[q] + an[q] + as[q] + at[q] + ab[q];
                su[q] = gauss(seed, 4, q);
                sum += su[q];
                ++cnt;
            }
    const double mean = sum / (double)cnt;
    for (int k = 3; k <= nkmm; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) su[IX(i, j, k)] -= mean;
    // reference-pressure cell (3,3,3), pinned as PISO_multiple_correction.f90:117-129
    const size_t r = IX(3, 3, 3);
    ae[r] = aw[r] = an[r] = as[r] = at[r] = ab[r] = 0.0;
    ap[r] = 1.0;
    su[r] = 0.0;
    if (dev) {
        for (int q = 0; q < 8; ++q)
            HIPCHK(hipMemcpy(out[q], P[q], n * sizeof(double), hipMemcpyHostToDevice));
    }
    return CAP_OK;
}

void cap_free_extra(cap_ctx *ctx)
{
    for (int q = 0; q < 20; ++q) if (ctx->G[q]) (void)hipFree(ctx->G[q]);
    for (int q = 0; q < F_COUNT; ++q) if (ctx->F[q]) (void)hipFree(ctx->F[q]);
}
