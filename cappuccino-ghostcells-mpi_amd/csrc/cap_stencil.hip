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
#include <mutex>
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

// Green-Gauss gradient (GRAD_GAUSS.f90:37-102) in gather form.  The
// reference scatters face terms in loop order (k, i, j); the value that
// reaches cell P is ((((0 - t_B) - e_W) - n_S) + own faces), B = k-1,
// W = i-1, S = j-1, each term present only when that face-loop iteration is
// in the loop range k=2..nkm, i=2..nimm, j=2..njmm; then it is scaled by
// volr = 1/vol on k=2..nkm, i,j=3..n?mm (88-102).
//
// A block of GG_J (j) x GG_I (i) threads, one per column, walks the planes
// of one k range.  Each cell forms the products of its own east, north and
// top faces once.  Its west and south neighbours' east and north products
// come through LDS (the block's first row and column form them again from
// memory), and its bottom face's products are the previous plane's top
// products, kept in registers; the first plane of the range forms them
// again.  Each product is the one the reference's face loop forms, and the
// terms are added in the reference's order, so the result is the same bit
// for bit, while each array is read once per plane.
constexpr int GG_J = 64, GG_I = 4;

__global__ __launch_bounds__(GG_J * GG_I) void k_grad_gauss(int ni, int nj, int nk,
                                                              const double *__restrict__ u, Geo g,
                                                              double *__restrict__ dx,
                                                              double *__restrict__ dy,
                                                              double *__restrict__ dz, int stride,
                                                              int kchunk)
{
    __shared__ double le[2][3][GG_I][GG_J], ln[2][3][GG_I][GG_J];
    const int tx = threadIdx.x % GG_J, ty = threadIdx.x / GG_J;
    const int j = 1 + (int)blockIdx.x * GG_J + tx, i = 1 + (int)blockIdx.y * GG_I + ty;
    const int k0 = 1 + (int)blockIdx.z * kchunk, k1 = min(nk, k0 + kchunk - 1);
    const bool col = i <= ni && j <= nj;
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    auto in_face = [&](int ii, int jj, int kk) {
        return kk >= 2 && kk <= nkm && ii >= 2 && ii <= nimm && jj >= 2 && jj <= njmm;
    };
    double tb[3] = {0.0, 0.0, 0.0};    // the previous plane's top products
    for (int k = k0; k <= k1; ++k) {
        const int par = k & 1;
        const size_t c = (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
        double pe[3] = {0.0, 0.0, 0.0}, pn[3] = {0.0, 0.0, 0.0}, pt[3] = {0.0, 0.0, 0.0};
        const bool own = col && in_face(i, j, k);
        if (own) {
            const double uc = u[c];
            const double ue = u[c + nj] * g.fx[c] + uc * (1.0 - g.fx[c]);
            const double un = u[c + 1] * g.fy[c] + uc * (1.0 - g.fy[c]);
            const double ut = u[c + nij] * g.fz[c] + uc * (1.0 - g.fz[c]);
            pe[0] = ue * g.a1x[c]; pe[1] = ue * g.a1y[c]; pe[2] = ue * g.a1z[c];
            pn[0] = un * g.a2x[c]; pn[1] = un * g.a2y[c]; pn[2] = un * g.a2z[c];
            pt[0] = ut * g.a3x[c]; pt[1] = ut * g.a3y[c]; pt[2] = ut * g.a3z[c];
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            le[par][q][ty][tx] = pe[q];
            ln[par][q][ty][tx] = pn[q];
        }
        __syncthreads();
        if (col) {
            double sx = 0.0, sy = 0.0, sz = 0.0;
            if (in_face(i, j, k - 1)) {            // top face of B
                if (k == k0) {
                    const size_t b = c - nij;
                    const double ut = u[c] * g.fz[b] + u[b] * (1.0 - g.fz[b]);
                    tb[0] = ut * g.a3x[b]; tb[1] = ut * g.a3y[b]; tb[2] = ut * g.a3z[b];
                }
                sx = sx - tb[0];
                sy = sy - tb[1];
                sz = sz - tb[2];
            }
            if (in_face(i - 1, j, k)) {            // east face of W
                double w3[3];
                if (ty > 0) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) w3[q] = le[par][q][ty - 1][tx];
                } else {
                    const size_t w = c - nj;
                    const double ue = u[c] * g.fx[w] + u[w] * (1.0 - g.fx[w]);
                    w3[0] = ue * g.a1x[w]; w3[1] = ue * g.a1y[w]; w3[2] = ue * g.a1z[w];
                }
                sx = sx - w3[0];
                sy = sy - w3[1];
                sz = sz - w3[2];
            }
            if (in_face(i, j - 1, k)) {            // north face of S
                double s3[3];
                if (tx > 0) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) s3[q] = ln[par][q][ty][tx - 1];
                } else {
                    const size_t sc = c - 1;
                    const double un = u[c] * g.fy[sc] + u[sc] * (1.0 - g.fy[sc]);
                    s3[0] = un * g.a2x[sc]; s3[1] = un * g.a2y[sc]; s3[2] = un * g.a2z[sc];
                }
                sx = sx - s3[0];
                sy = sy - s3[1];
                sz = sz - s3[2];
            }
            if (own) {                             // own east/north/top faces
                sx = sx + (pe[0] + pn[0] + pt[0]);
                sy = sy + (pe[1] + pn[1] + pt[1]);
                sz = sz + (pe[2] + pn[2] + pt[2]);
            }
            if (k >= 2 && k <= nkm && i >= 3 && i <= nimm && j >= 3 && j <= njmm) {
                const double volr = 1.0 / g.vol[c];
                sx = sx * volr;
                sy = sy * volr;
                sz = sz * volr;
            }
            dx[c * stride] = sx;
            dy[c * stride] = sy;
            dz[c * stride] = sz;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) tb[q] = pt[q];
    }
}

// grid of k_grad_gauss: the columns in blocks, the planes in ranges so
// that there are about 2048 blocks
static void grad_gauss_grid(int ni, int nj, int nk, dim3 &grid, int &kchunk)
{
    const int bx = (nj + GG_J - 1) / GG_J, by = (ni + GG_I - 1) / GG_I;
    const int nkc = std::max(1, std::min(nk, (2048 + bx * by - 1) / (bx * by)));
    kchunk = (nk + nkc - 1) / nkc;
    grid = dim3(bx, by, (nk + kchunk - 1) / kchunk);
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


// ------------------------------------------------------ p' assembly
// One thread per face of the reference's face loop, j (unit stride) fastest.
// Each formula keeps the reference's operand order (contraction is off).
struct Fld {
    const double *x, *y, *z, *xc, *yc, *zc, *vol, *fx, *fy, *fz;
    const double *ar1x, *ar1y, *ar1z, *ar2x, *ar2y, *ar2z, *ar3x, *ar3y, *ar3z;
    const double *den, *apu, *apv, *apw, *u, *v, *w, *p;
    const double *gu, *gv, *gw, *gp;   // 3 x n, component fastest
};
struct FaceDir {
    int idew, idns, idtb;
    const double *arvx, *arvy, *arvz, *fif;
};

// face-loop decode: k in [k0, k1], i in [2, nimm], j in [2, njmm]
__device__ __forceinline__ bool face_of(int ni, int nj, int k0, int k1, size_t q, size_t &p)
{
    const int nimm = ni - 2, njmm = nj - 2;
    const int NJ = njmm - 1, NI = nimm - 1;
    const size_t per = (size_t)NI * NJ;
    const int k = k0 + (int)(q / per);
    if (k > k1) return false;
    const int r = (int)(q % per);
    const int i = 2 + r / NJ, j = 2 + r % NJ;
    p = (size_t)(k - 1) * ni * nj + (size_t)(i - 1) * nj + (size_t)(j - 1);
    return true;
}

#define G3(g, c, q) ((g)[3 * (c) + (q)])

__global__ __launch_bounds__(256) void k_zero_st(double *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        p[i] = 0.0;
}

// fluxmass2.f90:93-188: the face east of cell p (in the direction D):
// its coefficient a (acfe(p) = acfw(e) = a) and mass flux fc
__device__ __forceinline__ void face_fluxmass(const Fld &f, const FaceDir &D, size_t p, double &a,
                                              double &fc)
{
    const size_t e = p + D.idew, s = p - D.idns, b = p - D.idtb, bs = b - D.idns;
    const double fxe = D.fif[p], fxp = 1.0 - fxe;
    const double xf = 0.25 * (f.x[p] + f.x[s] + f.x[b] + f.x[bs]);
    const double yf = 0.25 * (f.y[p] + f.y[s] + f.y[b] + f.y[bs]);
    const double zf = 0.25 * (f.z[p] + f.z[s] + f.z[b] + f.z[bs]);
    const double xi = f.xc[p] * fxp + f.xc[e] * fxe;
    const double yi = f.yc[p] * fxp + f.yc[e] * fxe;
    const double zi = f.zc[p] * fxp + f.zc[e] * fxe;
    const double xpn = f.xc[e] - f.xc[p], ypn = f.yc[e] - f.yc[p], zpn = f.zc[e] - f.zc[p];
    const double arx = D.arvx[p], ary = D.arvy[p], arz = D.arvz[p];
    const double are = sqrt(arx * arx + ary * ary + arz * arz);
    const double nxx = arx / are, nyy = ary / are, nzz = arz / are;
    const double sfdpnr = 1.0 / (arx * xpn * nxx + ary * ypn * nyy + arz * zpn * nzz);
    const double dene = f.den[p] * fxp + f.den[e] * fxe;
    const double smdpn = (arx * arx + ary * ary + arz * arz) * sfdpnr;
    a = dene * (fxp * f.vol[p] * f.apu[p] + fxe * f.vol[e] * f.apu[e]) * smdpn;
    double duxi = G3(f.gu, p, 0) * fxp + G3(f.gu, e, 0) * fxe;
    double duyi = G3(f.gu, p, 1) * fxp + G3(f.gu, e, 1) * fxe;
    double duzi = G3(f.gu, p, 2) * fxp + G3(f.gu, e, 2) * fxe;
    const double ui = f.u[p] * fxp + f.u[e] * fxe + (duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi));
    duxi = G3(f.gv, p, 0) * fxp + G3(f.gv, e, 0) * fxe;
    duyi = G3(f.gv, p, 1) * fxp + G3(f.gv, e, 1) * fxe;
    duzi = G3(f.gv, p, 2) * fxp + G3(f.gv, e, 2) * fxe;
    const double vi = f.v[p] * fxp + f.v[e] * fxe + (duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi));
    duxi = G3(f.gw, p, 0) * fxp + G3(f.gw, e, 0) * fxe;
    duyi = G3(f.gw, p, 1) * fxp + G3(f.gw, e, 1) * fxe;
    duzi = G3(f.gw, p, 2) * fxp + G3(f.gw, e, 2) * fxe;
    const double wi = f.w[p] * fxp + f.w[e] * fxe + (duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi));
    const double cu = fxe * f.vol[e] * f.apu[e] + fxp * f.vol[p] * f.apu[p];
    const double cv = fxe * f.vol[e] * f.apv[e] + fxp * f.vol[p] * f.apv[p];
    const double cw = fxe * f.vol[e] * f.apw[e] + fxp * f.vol[p] * f.apw[p];
    const double dpxi = cu * (fxe * G3(f.gp, e, 0) + fxp * G3(f.gp, p, 0)) * xpn * nxx;
    const double dpyi = cv * (fxe * G3(f.gp, e, 1) + fxp * G3(f.gp, p, 1)) * ypn * nyy;
    const double dpzi = cw * (fxe * G3(f.gp, e, 2) + fxp * G3(f.gp, p, 2)) * zpn * nzz;
    double xpp = xf - (xf - f.xc[p]) * nxx, ypp = yf - (yf - f.yc[p]) * nyy, zpp = zf - (zf - f.zc[p]) * nzz;
    double xep = xf - (xf - f.xc[e]) * nxx, yep = yf - (yf - f.yc[e]) * nyy, zep = zf - (zf - f.zc[e]) * nzz;
    xpp = xpp - f.xc[p]; ypp = ypp - f.yc[p]; zpp = zpp - f.zc[p];
    xep = xep - f.xc[e]; yep = yep - f.yc[e]; zep = zep - f.zc[e];
    // the reference's correction term, signs as written (fluxmass2.f90:171-173)
    const double dpe = (f.p[e] - f.p[p]) +
                       (G3(f.gp, e, 0) * xep + G3(f.gp, e, 1) * yep + G3(f.gp, e, 2) * zep -
                        G3(f.gp, p, 0) * xpp + G3(f.gp, p, 1) * ypp + G3(f.gp, p, 2) * zpp);
    const double dpex = cu * (arx * sfdpnr) * dpe;
    const double dpey = cv * (ary * sfdpnr) * dpe;
    const double dpez = cw * (arz * sfdpnr) * dpe;
    const double ue = ui - dpex + dpxi, ve = vi - dpey + dpyi, we = wi - dpez + dpzi;
    fc = dene * (ue * arx + ve * ary + we * arz);
}

// fluxmass2.f90:41-192
__global__ __launch_bounds__(256) void k_fluxmass(int ni, int nj, int nk, Fld f, FaceDir D,
                                                  double *__restrict__ acfe,
                                                  double *__restrict__ acfw,
                                                  double *__restrict__ fcf)
{
    const size_t nf = (size_t)(nk - 2) * (ni - 3) * (nj - 3);
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < nf; q += (size_t)gridDim.x * 256) {
        size_t p;
        if (!face_of(ni, nj, 2, nk - 1, q, p)) continue;
        double a, fc;
        face_fluxmass(f, D, p, a, fc);
        acfe[p] = a;
        acfw[p + D.idew] = a;
        fcf[p] = fc;
    }
}

// PISO_assemble_pressure_eq.f90:58-239: the face east of cell p (in the
// direction D): its coefficient a (acfe(p) = acfw(e) = a) and, when ipc > 0,
// the source suadd of the non-orthogonal correction and the mass flux fc
__device__ __forceinline__ void face_piso(const Fld &f, const FaceDir &D, int ipc, size_t p,
                                          double &a, double &suadd, double &fc)
{
    const size_t e = p + D.idew, s = p - D.idns, b = p - D.idtb, bs = b - D.idns;
    const double fxe = D.fif[p], fxp = 1.0 - fxe;
    const double xf = 0.25 * (f.x[p] + f.x[s] + f.x[b] + f.x[bs]);
    const double yf = 0.25 * (f.y[p] + f.y[s] + f.y[b] + f.y[bs]);
    const double zf = 0.25 * (f.z[p] + f.z[s] + f.z[b] + f.z[bs]);
    const double xi = f.xc[p] * fxp + f.xc[e] * fxe;
    const double yi = f.yc[p] * fxp + f.yc[e] * fxe;
    const double zi = f.zc[p] * fxp + f.zc[e] * fxe;
    const double xpn = f.xc[e] - f.xc[p], ypn = f.yc[e] - f.yc[p], zpn = f.zc[e] - f.zc[p];
    const double dpn = sqrt(xpn * xpn + ypn * ypn + zpn * zpn);
    const double ixi1 = xpn / dpn, ixi2 = ypn / dpn, ixi3 = zpn / dpn;
    const double arx = D.arvx[p], ary = D.arvy[p], arz = D.arvz[p];
    const double are = sqrt(arx * arx + ary * ary + arz * arz);
    const double vole = xpn * arx + ypn * ary + zpn * arz;
    const double nxx = arx / are, nyy = ary / are, nzz = arz / are;
    const double costn = nxx * ixi1 + nyy * ixi2 + nzz * ixi3;
    const double d2x = xpn * costn, d2y = ypn * costn, d2z = zpn * costn;
    const double dene = f.den[p] * fxp + f.den[e] * fxe;
    const double game = dene * (fxp * f.vol[p] * f.apu[p] + fxe * f.vol[e] * f.apu[e]);
    a = game * are / dpn;
    if (ipc > 0) {
        const double dfixi = G3(f.gp, p, 0) * fxp + G3(f.gp, e, 0) * fxe;
        const double dfiyi = G3(f.gp, p, 1) * fxp + G3(f.gp, e, 1) * fxe;
        const double dfizi = G3(f.gp, p, 2) * fxp + G3(f.gp, e, 2) * fxe;
        const double dfixii = dfixi * costn + arx / vole * (f.p[e] - f.p[p] - (dfixi * d2x + dfiyi * d2y + dfizi * d2z));
        const double dfiyii = dfiyi * costn + ary / vole * (f.p[e] - f.p[p] - (dfixi * d2x + dfiyi * d2y + dfizi * d2z));
        const double dfizii = dfizi * costn + arz / vole * (f.p[e] - f.p[p] - (dfixi * d2x + dfiyi * d2y + dfizi * d2z));
        const double fdfie = game * (dfixii * arx + dfiyii * ary + dfizii * arz);
        const double fdfii = game * are / dpn * (dfixi * xpn + dfiyi * ypn + dfizi * zpn);
        suadd = fdfie - fdfii;
        double duxi = G3(f.gu, p, 0) * fxp + G3(f.gu, e, 0) * fxe;
        double duyi = G3(f.gu, p, 1) * fxp + G3(f.gu, e, 1) * fxe;
        double duzi = G3(f.gu, p, 2) * fxp + G3(f.gu, e, 2) * fxe;
        const double ui = f.u[p] * fxp + f.u[e] * fxe + duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi);
        duxi = G3(f.gv, p, 0) * fxp + G3(f.gv, e, 0) * fxe;
        duyi = G3(f.gv, p, 1) * fxp + G3(f.gv, e, 1) * fxe;
        duzi = G3(f.gv, p, 2) * fxp + G3(f.gv, e, 2) * fxe;
        const double vi = f.v[p] * fxp + f.v[e] * fxe + duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi);
        duxi = G3(f.gw, p, 0) * fxp + G3(f.gw, e, 0) * fxe;
        duyi = G3(f.gw, p, 1) * fxp + G3(f.gw, e, 1) * fxe;
        duzi = G3(f.gw, p, 2) * fxp + G3(f.gw, e, 2) * fxe;
        const double wi = f.w[p] * fxp + f.w[e] * fxe + duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi);
        fc = dene * (ui * arx + vi * ary + wi * arz);
    }
}

// PISO_assemble_pressure_eq.f90:48-239 (face loop k=2..nkmm); bp zeroed before
__global__ __launch_bounds__(256) void k_piso_assemble(int ni, int nj, int nk, Fld f, FaceDir D,
                                                       int ipc, double *__restrict__ acfe,
                                                       double *__restrict__ acfw,
                                                       double *__restrict__ fcf,
                                                       double *__restrict__ bp,
                                                       double *__restrict__ su)
{
    const size_t nf = (size_t)(nk - 3) * (ni - 3) * (nj - 3);
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < nf; q += (size_t)gridDim.x * 256) {
        size_t p;
        if (!face_of(ni, nj, 2, nk - 2, q, p)) continue;
        double a, suadd = 0.0, fc = 0.0;
        face_piso(f, D, ipc, p, a, suadd, fc);
        acfe[p] = a;
        acfw[p + D.idew] = a;
        if (ipc > 0) {
            su[p] = su[p] + suadd;
            bp[p] = suadd;
            fcf[p] = fc;
        }
    }
}

// PISO_assemble_pressure_eq.f90:241-253: su(c) = su(c) - bp(c), c = cell + idew
__global__ __launch_bounds__(256) void k_piso_su(int ni, int nj, int nk, int idew,
                                                 const double *__restrict__ bp,
                                                 double *__restrict__ su)
{
    const int NI = ni - 4, NJ = nj - 4, NK = nk - 4;
    const size_t n = (size_t)NI * NJ * NK;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        const int k = 3 + (int)(q / ((size_t)NI * NJ));
        const int r = (int)(q % ((size_t)NI * NJ));
        const int i = 3 + r / NJ, j = 3 + r % NJ;
        const size_t c = (size_t)(k - 1) * ni * nj + (size_t)(i - 1) * nj + (size_t)(j - 1) + idew;
        su[c] = su[c] - bp[c];
    }
}

// assemble_pressure_eq.f90:38-118
__global__ __launch_bounds__(256) void k_assemble_p(int ni, int nj, int nk, Fld f, FaceDir D,
                                                    const double *__restrict__ su,
                                                    const double *__restrict__ sv,
                                                    const double *__restrict__ sw,
                                                    double *__restrict__ acfe,
                                                    double *__restrict__ acfw,
                                                    double *__restrict__ fcf)
{
    const size_t nf = (size_t)(nk - 3) * (ni - 3) * (nj - 3);
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < nf; q += (size_t)gridDim.x * 256) {
        size_t p;
        if (!face_of(ni, nj, 2, nk - 2, q, p)) continue;
        const size_t e = p + D.idew;
        const double fxe = D.fif[p], fxp = 1.0 - fxe;
        const double xpn = f.xc[e] - f.xc[p], ypn = f.yc[e] - f.yc[p], zpn = f.zc[e] - f.zc[p];
        const double arx = D.arvx[p], ary = D.arvy[p], arz = D.arvz[p];
        const double are = sqrt(arx * arx + ary * ary + arz * arz);
        const double nxx = arx / are, nyy = ary / are, nzz = arz / are;
        const double smdpn = (arx * arx + ary * ary + arz * arz) / (arx * xpn * nxx + ary * ypn * nyy + arz * zpn * nzz);
        const double a = (fxp * f.vol[p] + fxe * f.vol[e]) * smdpn;
        acfe[p] = a;
        acfw[e] = a;
        const double ui = su[p] * fxp + su[e] * fxe;
        const double vi = sv[p] * fxp + sv[e] * fxe;
        const double wi = sw[p] * fxp + sw[e] * fxe;
        fcf[p] = ui * arx + vi * ary + wi * arz;
    }
}

// fluxmc.f90:41-106 for every cell of k,i,j = 3..n?mm (others 0)
__global__ __launch_bounds__(256) void k_fluxmc(int ni, int nj, int nk, Fld f, FaceDir D,
                                                double *__restrict__ fmcor)
{
    const size_t n = (size_t)ni * nj * nk;
    const int nij = ni * nj;
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        const int k = (int)(c / nij) + 1, rem = (int)(c % nij);
        const int i = rem / nj + 1, j = rem % nj + 1;
        if (k < 3 || k > nk - 2 || i < 3 || i > ni - 2 || j < 3 || j > nj - 2) {
            fmcor[c] = 0.0;
            continue;
        }
        const size_t p = c, e = p + D.idew, s = p - D.idns, b = p - D.idtb, bs = b - D.idns;
        const double fxe = D.fif[p], fxp = 1. - fxe;
        double arx, ary, arz;
        if (D.idew == nj) { arx = f.ar1x[p]; ary = f.ar1y[p]; arz = f.ar1z[p]; }
        else if (D.idew == 1) { arx = f.ar2x[p]; ary = f.ar2y[p]; arz = f.ar2z[p]; }
        else { arx = f.ar3x[p]; ary = f.ar3y[p]; arz = f.ar3z[p]; }
        const double xpn = f.xc[e] - f.xc[p], ypn = f.yc[e] - f.yc[p], zpn = f.zc[e] - f.zc[p];
        const double are = sqrt(arx * arx + ary * ary + arz * arz);
        const double nxx = arx / are, nyy = ary / are, nzz = arz / are;
        const double dppnnr = 1. / ((xpn * nxx) + (ypn * nyy) + (zpn * nzz));
        const double xf = 0.25 * (f.x[p] + f.x[s] + f.x[b] + f.x[bs]);
        const double yf = 0.25 * (f.y[p] + f.y[s] + f.y[b] + f.y[bs]);
        const double zf = 0.25 * (f.z[p] + f.z[s] + f.z[b] + f.z[bs]);
        double xpp = xf - (xf - f.xc[p]) * nxx, ypp = yf - (yf - f.yc[p]) * nyy, zpp = zf - (zf - f.zc[p]) * nzz;
        double xep = xf - (xf - f.xc[e]) * nxx, yep = yf - (yf - f.yc[e]) * nyy, zep = zf - (zf - f.zc[e]) * nzz;
        xpp = xpp - f.xc[p]; ypp = ypp - f.yc[p]; zpp = zpp - f.zc[p];
        xep = xep - f.xc[e]; yep = yep - f.yc[e]; zep = zep - f.zc[e];
        const double rapr = (f.apu[p] * f.den[p] * f.vol[p] * fxp + f.apu[e] * f.den[e] * f.vol[e] * fxe);
        fmcor[c] = rapr * are * ((G3(f.gp, e, 0) * xep - G3(f.gp, p, 0) * xpp) +
                                 (G3(f.gp, e, 1) * yep - G3(f.gp, p, 1) * ypp) +
                                 (G3(f.gp, e, 2) * zep - G3(f.gp, p, 2) * zpp)) * dppnnr;
    }
}

// ------------------------------------------------------ momentum fluxes
// fluxuvw-overrelaxed-correction.f90: one thread per face of the direction's
// face range (99-118); each formula keeps the reference's operand order.
struct UvwOut {
    const double *vis, *fcf;
    double *acfe, *acfw, *su, *sv, *sw, *bp, *bt, *bb;
};

// QUICK high-order fluxes of the face east of cell p (287-341; valid where
// the reference's interior test holds)
__device__ __forceinline__ void uvw_quick(const Fld &f, const FaceDir &D, const double *fcf,
                                          size_t p, double (&h)[3])
{
    const size_t e = p + D.idew, w = p - D.idew, ee = p + 2 * (size_t)D.idew;
    const double flcf = fcf[p];
    const double ce = flcf < 0.0 ? flcf : 0.0, cp = flcf > 0.0 ? flcf : 0.0;
    const double fw_ = D.fif[w], fp_ = D.fif[p], fe_ = D.fif[e];
    const double g11 = ((2.0 - fw_) * (fp_ * fp_)) / (1.0 + fp_ - fw_);
    const double g12 = ((1.0 - fp_) * ((1.0 - fw_) * (1.0 - fw_))) / (1.0 + fp_ - fw_);
    const double g21 = ((1.0 + fe_) * ((1.0 - fp_) * (1.0 - fp_))) / (1.0 + fe_ - fp_);
    const double g22 = ((fe_ * fe_) * fp_) / (1.0 + fe_ - fp_);
    const double *vel[3] = {f.u, f.v, f.w};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double *q = vel[c];
        h[c] = ce * (g21 * q[p] + (1.0 - g21 + g22) * q[e] - g22 * q[ee]) +
               cp * (g11 * q[e] + (1.0 - g11 + g12) * q[p] - g12 * q[w]);
    }
}

__device__ __forceinline__ double fmin_f(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double fmax_f(double a, double b) { return b > a ? b : a; }

__global__ __launch_bounds__(256) void k_fluxuvw(int ni, int nj, int nk, Fld f, FaceDir D,
                                                 UvwOut o, int flags, double gam, int dir,
                                                 int kst, int ist, int jst)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const int NJ = njmm - jst + 1, NI = nimm - ist + 1, NK = nkmm - kst + 1;
    const size_t nf = (size_t)NK * NI * NJ;
    const bool lcds = flags & CAP_UVW_CDS, lcds4 = flags & CAP_UVW_CDS4;
    const bool lluds = flags & CAP_UVW_LUDS, lquds = flags & CAP_UVW_QUICK;
    const bool flim = flags & CAP_UVW_LIMITER;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < nf; q += (size_t)gridDim.x * 256) {
        const int k = kst + (int)(q / ((size_t)NI * NJ));
        const int r = (int)(q % ((size_t)NI * NJ));
        const int i = ist + r / NJ, j = jst + r % NJ;
        const size_t inp = (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
        const size_t ine = inp + D.idew, inw = inp - D.idew, ins = inp - D.idns;
        const size_t inb = inp - D.idtb, inee = inp + 2 * (size_t)D.idew, inbs = inb - D.idns;
        const double fxe = D.fif[inp], fxp = 1.0 - fxe;
        const double fxpw = D.fif[inw], fxw = 1.0 - fxpw;
        const double fxee = D.fif[ine], fxew = 1.0 - fxee;
        const double xpn = f.xc[ine] - f.xc[inp], ypn = f.yc[ine] - f.yc[inp], zpn = f.zc[ine] - f.zc[inp];
        const double dpn = sqrt(xpn * xpn + ypn * ypn + zpn * zpn);
        const double arx = D.arvx[inp], ary = D.arvy[inp], arz = D.arvz[inp];
        const double are = sqrt(arx * arx + ary * ary + arz * arz);
        const double costn = 1.0;   // orthogonal correction (178)
        const double vole = xpn * arx + ypn * ary + zpn * arz;
        const double d1x = costn, d1y = costn, d1z = costn;
        const double d2x = xpn * costn, d2y = ypn * costn, d2z = zpn * costn;
        const double game = o.vis[inp] * fxp + o.vis[ine] * fxe;
        const double xf = 0.25 * (f.x[inp] + f.x[ins] + f.x[inb] + f.x[inbs]);
        const double yf = 0.25 * (f.y[inp] + f.y[ins] + f.y[inb] + f.y[inbs]);
        const double zf = 0.25 * (f.z[inp] + f.z[ins] + f.z[inb] + f.z[inbs]);
        const double xi = f.xc[inp] * fxp + f.xc[ine] * fxe;
        const double yi = f.yc[inp] * fxp + f.yc[ine] * fxe;
        const double zi = f.zc[inp] * fxp + f.zc[ine] * fxe;
        const double *vel[3] = {f.u, f.v, f.w};
        const double *grd[3] = {f.gu, f.gv, f.gw};
        double dxi[3][3], dxii[3][3];   // [velocity][x,y,z]
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double *g = grd[c], *vq = vel[c];
            dxi[c][0] = G3(g, inp, 0) * fxp + G3(g, ine, 0) * fxe;
            dxi[c][1] = G3(g, inp, 1) * fxp + G3(g, ine, 1) * fxe;
            dxi[c][2] = G3(g, inp, 2) * fxp + G3(g, ine, 2) * fxe;
            const double dd = vq[ine] - vq[inp] - dxi[c][0] * d2x - dxi[c][1] * d2y - dxi[c][2] * d2z;
            dxii[c][0] = dxi[c][0] * d1x + arx / vole * dd;
            dxii[c][1] = dxi[c][1] * d1y + ary / vole * dd;
            dxii[c][2] = dxi[c][2] * d1z + arz / vole * dd;
        }
        // explicit (fd*e) and implicit (fd*i) diffusion (262-269)
        const double duxii = dxii[0][0], duyii = dxii[0][1], duzii = dxii[0][2];
        const double dvxii = dxii[1][0], dvyii = dxii[1][1], dvzii = dxii[1][2];
        const double dwxii = dxii[2][0], dwyii = dxii[2][1], dwzii = dxii[2][2];
        double fde[3], fdi[3];
        fde[0] = game * ((duxii + duxii) * arx + (duyii + dvxii) * ary + (duzii + dwxii) * arz);
        fde[1] = game * ((duyii + dvxii) * arx + (dvyii + dvyii) * ary + (dvzii + dwyii) * arz);
        fde[2] = game * ((duzii + dwxii) * arx + (dwyii + dvzii) * ary + (dwzii + dwzii) * arz);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            fdi[c] = game * are / dpn * (dxi[c][0] * xpn + dxi[c][1] * ypn + dxi[c][2] * zpn);
        const double de = game * are / dpn;
        const double flcf = o.fcf[inp];
        const double ce = flcf < 0.0 ? flcf : 0.0, cp = flcf > 0.0 ? flcf : 0.0;
        o.acfe[inp] = de - ce;
        o.acfw[ine] = de + cp;
        double fuds[3], fh[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < 3; ++c) fuds[c] = cp * vel[c][inp] + ce * vel[c][ine];
        bool set = false;
        if (lcds) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double ue = vel[c][inp] * fxp + vel[c][ine] * fxe +
                                  (dxi[c][0] * (xf - xi) + dxi[c][1] * (yf - yi) + dxi[c][2] * (zf - zi));
                fh[c] = flcf * ue;
            }
            set = true;
        }
        if (lcds4 && dir != 3) {
            const double c48 = 1.0 / 48.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double *vq = vel[c];
                fh[c] = flcf * (c48 * (27.0 * vq[inp] + 27.0 * vq[ine] - 3.0 * vq[inw] - 3.0 * vq[inee]));
            }
            set = true;
        }
        if (lluds) {
            const double dx[3] = {xf - f.xc[inp], yf - f.yc[inp], zf - f.zc[inp]};
            const double dxe[3] = {f.xc[ine] - xf, f.yc[ine] - yf, f.zc[ine] - zf};
            const int cc = dir - 1;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double *g = grd[c], *vq = vel[c];
                const double dfu = G3(g, inw, cc) * fxw + G3(g, inp, cc) * fxpw;
                const double dfu1 = G3(g, inee, cc) * fxee + G3(g, ine, cc) * fxew;
                fh[c] = ce * (vq[ine] + dxe[c] * dfu1) + cp * (vq[inp] + dx[c] * dfu);
            }
            set = true;
        }
        if (lquds) {
            if (k > 2 && k < nk - 1 && j > 2 && j < nj - 1 && i > 2 && i < ni - 1) {
                uvw_quick(f, D, o.fcf, inp, fh);
                set = true;
            } else if (!set) {
                // the reference keeps the previous face's values: the last face
                // of the previous row (north faces) or plane that QUICK set
                int pk = k, pi = i;
                bool any = false;
                if (dir == 2 && i - 1 >= ist) { pi = i - 1; any = true; }
                else if (dir != 3 && k - 1 >= kst) { pk = k - 1; pi = nimm; any = true; }
                if (any) {
                    const size_t pp = (size_t)(pk - 1) * nij + (size_t)(pi - 1) * nj + (size_t)(njmm - 1);
                    uvw_quick(f, D, o.fcf, pp, fh);
                }
            }
        }
        if (flim) {
            double psi[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double *g = grd[c], *vq = vel[c];
                const double rw = (2.0 * G3(g, inp, 0) * xpn + 2.0 * G3(g, inp, 1) * ypn + 2.0 * G3(g, inp, 2) * zpn) / (vq[ine] - vq[inp]) - 1.0;
                const double re = (2.0 * G3(g, ine, 0) * xpn + 2.0 * G3(g, ine, 1) * ypn + 2.0 * G3(g, ine, 2) * zpn) / (vq[inp] - vq[ine]) - 1.0;
                const double rr[2] = {rw, re};
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const double x = rr[s];
                    double ps = 0.0;
                    if (flags & CAP_UVW_SMART) ps = fmax_f(0.0, fmin_f(fmin_f(2.0 * x, 0.75 * x + 0.25), 4.0));
                    if (flags & CAP_UVW_AVL) ps = fmax_f(0.0, fmin_f(fmin_f(1.5 * x, 0.75 * x + 0.25), 2.5));
                    if (flags & CAP_UVW_MUSCL) ps = fmax_f(0.0, fmin_f(fmin_f(2.0 * x, 0.5 * x + 0.5), 2.0));
                    if (flags & CAP_UVW_UMIST) ps = fmax_f(0.0, fmin_f(fmin_f(fmin_f(2.0 * x, 0.75 * x + 0.25), 0.25 * x + 0.75), 2.0));
                    if (flags & CAP_UVW_GAMMA) ps = fmax_f(0.0, fmin_f(x, 2.0 * x / (x + 1.0)));
                    psi[3 * s + c] = ps;
                }
                fh[c] = ce * (vq[ine] + fxe * psi[3 + c] * (vq[inp] - vq[ine])) +
                        cp * (vq[inp] + fxp * psi[c] * (vq[ine] - vq[inp]));
            }
            set = true;
        }
        double *srcs[3] = {o.su, o.sv, o.sw}, *scr[3] = {o.bp, o.bt, o.bb};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            srcs[c][inp] = srcs[c][inp] + gam * (fuds[c] - fh[c]) + fde[c] - fdi[c];
            scr[c][ine] = -gam * (fuds[c] - fh[c]) - fde[c] + fdi[c];
        }
    }
}

// fluxuvw's CDS branch (the same faces and terms as k_fluxuvw with
// lcds) with the face arithmetic in FP32: the geometric differences (xpn,
// the face centre's offset from the interpolation point) are formed in FP64
// and rounded once; velocities, gradients, viscosity and the mass flux are
// rounded at the load; the face's contribution is added to su, sv, sw in FP64.
__global__ __launch_bounds__(256) void k_fluxuvw32(int ni, int nj, int nk, Fld f, FaceDir D,
                                                   UvwOut o, float gam, int kst, int ist, int jst)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const int NJ = njmm - jst + 1, NI = nimm - ist + 1, NK = nkmm - kst + 1;
    const size_t nf = (size_t)NK * NI * NJ;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < nf; q += (size_t)gridDim.x * 256) {
        const int k = kst + (int)(q / ((size_t)NI * NJ));
        const int r = (int)(q % ((size_t)NI * NJ));
        const int i = ist + r / NJ, j = jst + r % NJ;
        const size_t inp = (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
        const size_t ine = inp + D.idew, ins = inp - D.idns;
        const size_t inb = inp - D.idtb, inbs = inb - D.idns;
        const double fxe64 = D.fif[inp], fxp64 = 1.0 - fxe64;
        const float fxe = (float)fxe64, fxp = 1.0f - fxe;
        const float xpn = (float)(f.xc[ine] - f.xc[inp]), ypn = (float)(f.yc[ine] - f.yc[inp]),
                    zpn = (float)(f.zc[ine] - f.zc[inp]);
        const float dpn = sqrtf(xpn * xpn + ypn * ypn + zpn * zpn);
        const float arx = (float)D.arvx[inp], ary = (float)D.arvy[inp], arz = (float)D.arvz[inp];
        const float are = sqrtf(arx * arx + ary * ary + arz * arz);
        const float vole = xpn * arx + ypn * ary + zpn * arz;
        const float game = (float)o.vis[inp] * fxp + (float)o.vis[ine] * fxe;
        const double xf = 0.25 * (f.x[inp] + f.x[ins] + f.x[inb] + f.x[inbs]);
        const double yf = 0.25 * (f.y[inp] + f.y[ins] + f.y[inb] + f.y[inbs]);
        const double zf = 0.25 * (f.z[inp] + f.z[ins] + f.z[inb] + f.z[inbs]);
        const float dxf = (float)(xf - (f.xc[inp] * fxp64 + f.xc[ine] * fxe64));
        const float dyf = (float)(yf - (f.yc[inp] * fxp64 + f.yc[ine] * fxe64));
        const float dzf = (float)(zf - (f.zc[inp] * fxp64 + f.zc[ine] * fxe64));
        const double *vel[3] = {f.u, f.v, f.w};
        const double *grd[3] = {f.gu, f.gv, f.gw};
        float dxi[3][3], dxii[3][3], vp[3], ve[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double *g = grd[c];
            vp[c] = (float)vel[c][inp];
            ve[c] = (float)vel[c][ine];
#pragma unroll
            for (int m = 0; m < 3; ++m) dxi[c][m] = (float)G3(g, inp, m) * fxp + (float)G3(g, ine, m) * fxe;
            const float dd = ve[c] - vp[c] - dxi[c][0] * xpn - dxi[c][1] * ypn - dxi[c][2] * zpn;
            dxii[c][0] = dxi[c][0] + arx / vole * dd;
            dxii[c][1] = dxi[c][1] + ary / vole * dd;
            dxii[c][2] = dxi[c][2] + arz / vole * dd;
        }
        float fde[3], fdi[3];
        fde[0] = game * ((dxii[0][0] + dxii[0][0]) * arx + (dxii[0][1] + dxii[1][0]) * ary +
                         (dxii[0][2] + dxii[2][0]) * arz);
        fde[1] = game * ((dxii[0][1] + dxii[1][0]) * arx + (dxii[1][1] + dxii[1][1]) * ary +
                         (dxii[1][2] + dxii[2][1]) * arz);
        fde[2] = game * ((dxii[0][2] + dxii[2][0]) * arx + (dxii[2][1] + dxii[1][2]) * ary +
                         (dxii[2][2] + dxii[2][2]) * arz);
        const float de = game * are / dpn;
#pragma unroll
        for (int c = 0; c < 3; ++c) fdi[c] = de * (dxi[c][0] * xpn + dxi[c][1] * ypn + dxi[c][2] * zpn);
        const float flcf = (float)o.fcf[inp];
        const float ce = flcf < 0.0f ? flcf : 0.0f, cp = flcf > 0.0f ? flcf : 0.0f;
        o.acfe[inp] = (double)(de - ce);
        o.acfw[ine] = (double)(de + cp);
        double *srcs[3] = {o.su, o.sv, o.sw}, *scr[3] = {o.bp, o.bt, o.bb};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float fuds = cp * vp[c] + ce * ve[c];
            const float fh = flcf * (vp[c] * fxp + ve[c] * fxe +
                                     (dxi[c][0] * dxf + dxi[c][1] * dyf + dxi[c][2] * dzf));
            const float t = gam * (fuds - fh) + fde[c] - fdi[c];
            srcs[c][inp] = srcs[c][inp] + (double)t;
            scr[c][ine] = (double)(-t);
        }
    }
}

// second loop (632-644): the east cell's share, k=2..nkmm, i=2..nimm, j=2..njmm
__global__ __launch_bounds__(256) void k_fluxuvw_east(int ni, int nj, int nk, int idew, UvwOut o)
{
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const int NJ = njmm - 1, NI = nimm - 1;
    const size_t nf = (size_t)(nkmm - 1) * NI * NJ;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < nf; q += (size_t)gridDim.x * 256) {
        const int k = 2 + (int)(q / ((size_t)NI * NJ));
        const int r = (int)(q % ((size_t)NI * NJ));
        const int i = 2 + r / NJ, j = 2 + r % NJ;
        const size_t p = (size_t)(k - 1) * ni * nj + (size_t)(i - 1) * nj + (size_t)(j - 1) + idew;
        o.su[p] = o.su[p] + o.bp[p];
        o.sv[p] = o.sv[p] + o.bt[p];
        o.sw[p] = o.sw[p] + o.bb[p];
    }
}

// ------------------------------------------- three face directions in one pass
// The drivers assemble the east, north and top faces with three calls of a
// per-direction routine (each a face loop, then for PISO and fluxuvw a loop
// that moves a share to the neighbour cell).  The fused passes give every
// array the value those three calls leave in it, with the per-direction
// kernels' own face formulas (the face_* functions), so every value is
// bit-identical:
//  * fluxmass and PISO: one face per thread, the three directions in one grid
//    with the blocks of one run of faces on one XCD (k_fluxmass3x,
//    k_piso3x), then for PISO a pass over the cells for su (k_piso3_su);
//  * fluxuvw: its su/sv/sw take six terms a cell in a fixed order, so a
//    thread owns a cell c (k_uvw3), evaluates its east/north/top faces and
//    its west/south/bottom neighbours' (whose shares reach c through the
//    second loop), and applies the terms in the calls' order.  A block is 256
//    consecutive (i, j) columns, j fastest, marching F3_KSEG planes.
constexpr int F3_KSEG = 8;

struct Out3 {
    double *a[6];           // ae, aw, an, as, at, ab
    const double *fcr[3];   // the mass fluxes f1, f2, f3
    double *bp[3];          // the scratch bp, bt, bb
    double *su[3];          // su, sv, sw
    const double *vis;
};

// direction q: 0 east (i + 1), 1 north (j + 1), 2 top (k + 1), as the drivers
// pass them
__device__ __forceinline__ FaceDir dir_of(const Fld &f, int q, int nj, int nij)
{
    FaceDir D;
    if (q == 0) {
        D.idew = nj; D.idns = 1; D.idtb = nij;
        D.arvx = f.ar1x; D.arvy = f.ar1y; D.arvz = f.ar1z; D.fif = f.fx;
    } else if (q == 1) {
        D.idew = 1; D.idns = nij; D.idtb = nj;
        D.arvx = f.ar2x; D.arvy = f.ar2y; D.arvz = f.ar2z; D.fif = f.fy;
    } else {
        D.idew = nij; D.idns = nj; D.idtb = 1;
        D.arvx = f.ar3x; D.arvy = f.ar3y; D.arvz = f.ar3z; D.fif = f.fz;
    }
    return D;
}

struct Box {
    int i0, i1, j0, j1, k0, k1;
    __device__ __forceinline__ bool has(int i, int j, int k) const
    {
        return i >= i0 && i <= i1 && j >= j0 && j <= j1 && k >= k0 && k <= k1;
    }
};

// (i, j) of this thread's column, 1-based, and the first plane k0 of its
// segment; false past the plane.  Block b runs on XCD b % 8; each XCD takes
// a contiguous eighth of the column blocks, so the blocks an XCD runs one
// after the other hold neighbouring rows of columns (a face reads the row
// beside its cell), all of one plane segment.
__device__ __forceinline__ bool f3_column(int ni, int nj, int &i, int &j, int &k0)
{
    const int ncb = (int)(((long)ni * nj + 255) / 256), per = (ncb + 7) / 8;
    const int b = (int)blockIdx.x, w = b >> 3;
    const int cbl = (b & 7) * per + w % per;
    k0 = 1 + (w / per) * F3_KSEG;
    if (cbl >= ncb) return false;
    const long q = (long)cbl * 256 + threadIdx.x;
    if (q >= (long)ni * nj) return false;
    i = (int)(q / nj) + 1;
    j = (int)(q % nj) + 1;
    return true;
}

// One face per thread, the three directions in one grid: block b runs on XCD
// b % 8 (blocks are dispatched to the XCDs in turn), and the three
// directions of one run of 256 faces take blocks b, b + 8, b + 16 of the same
// XCD, so the cell values the three faces share are read from HBM once and
// from that XCD's L2 after.  Measured at 256^3 against two other mappings:
// each XCD taking a contiguous eighth of the runs (so that the row beside a
// face's cell is on the same XCD too) 2.49 against 2.27 ms for fluxmass;
// tiles of 64 x 4 faces marching 8 planes 2.97 ms.  Per block: run cb of
// the face loop, direction q.
__device__ __forceinline__ bool f3x_item(int nruns, int &cb, int &q)
{
    const int b = (int)blockIdx.x, xcd = b & 7, w = b >> 3;
    q = w % 3;
    cb = (w / 3) * 8 + xcd;
    return cb < nruns;
}
static inline int f3x_blocks(int nruns) { return (nruns + 7) / 8 * 8 * 3; }

__global__ __launch_bounds__(256) void k_fluxmass3x(int ni, int nj, int nk, Fld f,
                                                    double *__restrict__ ae, double *__restrict__ aw,
                                                    double *__restrict__ an, double *__restrict__ as,
                                                    double *__restrict__ at, double *__restrict__ ab,
                                                    double *__restrict__ f1, double *__restrict__ f2,
                                                    double *__restrict__ f3)
{
    const size_t nf = (size_t)(nk - 2) * (ni - 3) * (nj - 3);
    int cb, q;
    if (!f3x_item((int)((nf + 255) / 256), cb, q)) return;
    const size_t t = (size_t)cb * 256 + threadIdx.x;
    size_t p;
    if (t >= nf || !face_of(ni, nj, 2, nk - 1, t, p)) return;
    const FaceDir D = dir_of(f, q, nj, ni * nj);
    double a, fc;
    face_fluxmass(f, D, p, a, fc);
    double *const acfe = q == 0 ? ae : (q == 1 ? an : at);
    double *const acfw = q == 0 ? aw : (q == 1 ? as : ab);
    double *const fcf = q == 0 ? f1 : (q == 1 ? f2 : f3);
    acfe[p] = a;
    acfw[p + D.idew] = a;
    fcf[p] = fc;
}

// PISO's face loop for the three directions (bp = be, bn, bt written on the
// face range); k_piso3_su then finishes each cell
__global__ __launch_bounds__(256) void k_piso3x(int ni, int nj, int nk, Fld f, int ipc,
                                                double *__restrict__ ae, double *__restrict__ aw,
                                                double *__restrict__ an, double *__restrict__ as,
                                                double *__restrict__ at, double *__restrict__ ab,
                                                double *__restrict__ f1, double *__restrict__ f2,
                                                double *__restrict__ f3, double *__restrict__ be,
                                                double *__restrict__ bn, double *__restrict__ bt)
{
    const size_t nf = (size_t)(nk - 3) * (ni - 3) * (nj - 3);
    int cb, q;
    if (!f3x_item((int)((nf + 255) / 256), cb, q)) return;
    const size_t t = (size_t)cb * 256 + threadIdx.x;
    size_t p;
    if (t >= nf || !face_of(ni, nj, 2, nk - 2, t, p)) return;
    const FaceDir D = dir_of(f, q, nj, ni * nj);
    double a, suadd = 0.0, fc = 0.0;
    face_piso(f, D, ipc, p, a, suadd, fc);
    double *const acfe = q == 0 ? ae : (q == 1 ? an : at);
    double *const acfw = q == 0 ? aw : (q == 1 ? as : ab);
    acfe[p] = a;
    acfw[p + D.idew] = a;
    (q == 0 ? be : (q == 1 ? bn : bt))[p] = suadd;
    if (ipc > 0) (q == 0 ? f1 : (q == 1 ? f2 : f3))[p] = fc;
}

// each cell after k_piso3x: bp zero outside the face range; on it, su takes
// the three directions' face source and (241-253) loses it again where
// c - idew lies in k,i,j = 3..nkmm, 3..nimm, 3..njmm, in the calls' order
__global__ __launch_bounds__(256) void k_piso3_su(int ni, int nj, int nk, int ipc,
                                                  double *__restrict__ be, double *__restrict__ bn,
                                                  double *__restrict__ bt, double *__restrict__ su)
{
    const size_t n = (size_t)ni * nj * nk;
    const Box F = {2, ni - 2, 2, nj - 2, 2, nk - 2};
    const Box C = {3, ni - 2, 3, nj - 2, 3, nk - 2};
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        const int k = (int)(c / ((size_t)ni * nj)) + 1;
        const int r = (int)(c % ((size_t)ni * nj));
        const int i = r / nj + 1, j = r % nj + 1;
        if (!F.has(i, j, k)) {
            be[c] = 0.0; bn[c] = 0.0; bt[c] = 0.0;
            continue;
        }
        if (ipc <= 0) continue;
        const double b3[3] = {be[c], bn[c], bt[c]};
        double s = su[c];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            s = s + b3[q];
            if (C.has(i - (q == 0), j - (q == 1), k - (q == 2))) s = s - b3[q];
        }
        su[c] = s;
    }
}

// fluxuvw-overrelaxed-correction.f90 with the CDS scheme alone: the face
// east of cell p (direction D), as k_fluxuvw evaluates it with lcds: the
// coefficient terms de, ce, cp and, per velocity component, the deferred
// correction g = gam*(fuds - fh) and the explicit / implicit diffusion
// fde, fdi (su(p) += g + fde - fdi; the east cell's share -g - fde + fdi)
__device__ __forceinline__ void face_uvw_cds(const Fld &f, const FaceDir &D, const double *vis,
                                             const double *fcf, size_t inp, double gam,
                                             double &de, double &ce, double &cp, double (&g)[3],
                                             double (&fde)[3], double (&fdi)[3])
{
    const size_t ine = inp + D.idew, ins = inp - D.idns;
    const size_t inb = inp - D.idtb, inbs = inb - D.idns;
    const double fxe = D.fif[inp], fxp = 1.0 - fxe;
    const double xpn = f.xc[ine] - f.xc[inp], ypn = f.yc[ine] - f.yc[inp], zpn = f.zc[ine] - f.zc[inp];
    const double dpn = sqrt(xpn * xpn + ypn * ypn + zpn * zpn);
    const double arx = D.arvx[inp], ary = D.arvy[inp], arz = D.arvz[inp];
    const double are = sqrt(arx * arx + ary * ary + arz * arz);
    const double costn = 1.0;
    const double vole = xpn * arx + ypn * ary + zpn * arz;
    const double d1x = costn, d1y = costn, d1z = costn;
    const double d2x = xpn * costn, d2y = ypn * costn, d2z = zpn * costn;
    const double game = vis[inp] * fxp + vis[ine] * fxe;
    const double xf = 0.25 * (f.x[inp] + f.x[ins] + f.x[inb] + f.x[inbs]);
    const double yf = 0.25 * (f.y[inp] + f.y[ins] + f.y[inb] + f.y[inbs]);
    const double zf = 0.25 * (f.z[inp] + f.z[ins] + f.z[inb] + f.z[inbs]);
    const double xi = f.xc[inp] * fxp + f.xc[ine] * fxe;
    const double yi = f.yc[inp] * fxp + f.yc[ine] * fxe;
    const double zi = f.zc[inp] * fxp + f.zc[ine] * fxe;
    const double *vel[3] = {f.u, f.v, f.w};
    const double *grd[3] = {f.gu, f.gv, f.gw};
    double dxi[3][3], dxii[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double *gr = grd[c], *vq = vel[c];
        dxi[c][0] = G3(gr, inp, 0) * fxp + G3(gr, ine, 0) * fxe;
        dxi[c][1] = G3(gr, inp, 1) * fxp + G3(gr, ine, 1) * fxe;
        dxi[c][2] = G3(gr, inp, 2) * fxp + G3(gr, ine, 2) * fxe;
        const double dd = vq[ine] - vq[inp] - dxi[c][0] * d2x - dxi[c][1] * d2y - dxi[c][2] * d2z;
        dxii[c][0] = dxi[c][0] * d1x + arx / vole * dd;
        dxii[c][1] = dxi[c][1] * d1y + ary / vole * dd;
        dxii[c][2] = dxi[c][2] * d1z + arz / vole * dd;
    }
    const double duxii = dxii[0][0], duyii = dxii[0][1], duzii = dxii[0][2];
    const double dvxii = dxii[1][0], dvyii = dxii[1][1], dvzii = dxii[1][2];
    const double dwxii = dxii[2][0], dwyii = dxii[2][1], dwzii = dxii[2][2];
    fde[0] = game * ((duxii + duxii) * arx + (duyii + dvxii) * ary + (duzii + dwxii) * arz);
    fde[1] = game * ((duyii + dvxii) * arx + (dvyii + dvyii) * ary + (dvzii + dwyii) * arz);
    fde[2] = game * ((duzii + dwxii) * arx + (dwyii + dvzii) * ary + (dwzii + dwzii) * arz);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        fdi[c] = game * are / dpn * (dxi[c][0] * xpn + dxi[c][1] * ypn + dxi[c][2] * zpn);
    de = game * are / dpn;
    const double flcf = fcf[inp];
    ce = flcf < 0.0 ? flcf : 0.0;
    cp = flcf > 0.0 ? flcf : 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double fuds = cp * vel[c][inp] + ce * vel[c][ine];
        const double ue = vel[c][inp] * fxp + vel[c][ine] * fxe +
                          (dxi[c][0] * (xf - xi) + dxi[c][1] * (yf - yi) + dxi[c][2] * (zf - zi));
        const double fh = flcf * ue;
        g[c] = gam * (fuds - fh);
    }
}

// fluxuvw (CDS) for the three directions.  Direction q's face loop covers
// k=kst..nkmm, i=ist..nimm, j=jst..njmm with the starts of 99-118; the east
// loop (632-644) adds the scratch bp/bt/bb of c to su/sv/sw(c) for c - idew in
// k,i,j = 2..nkmm, 2..nimm, 2..njmm, and the call ends by zeroing the
// scratch.  So the east loop adds the share the west/south/bottom face wrote
// at c, or, where that face is outside the face loop, the scratch's value:
// what the caller left in it for the east call, zero for the later calls.
__global__ __launch_bounds__(256) void k_uvw3(int ni, int nj, int nk, Fld f, Out3 o, double gam)
{
    int i, j, k0;
    if (!f3_column(ni, nj, i, j, k0)) return;
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const Box G = {2, nimm, 2, njmm, 2, nkmm};
    const Box F[3] = {{2, nimm, 3, njmm, 3, nkmm}, {3, nimm, 2, njmm, 3, nkmm},
                      {3, nimm, 3, njmm, 2, nkmm}};
    const int k1 = min(k0 + F3_KSEG - 1, nk);
    for (int k = k0; k <= k1; ++k) {
        const size_t c = (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
        double scr0[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) scr0[m] = o.bp[m][c];
        bool touched = false;
#pragma unroll
        for (int q = 0; q < 3; ++q)
            touched |= F[q].has(i, j, k) || G.has(i - (q == 0), j - (q == 1), k - (q == 2));
        if (!touched) {
#pragma unroll
            for (int m = 0; m < 3; ++m) o.bp[m][c] = 0.0;
            continue;
        }
        // every face first (the own face's coefficient and source terms,
        // the west/south/bottom face's coefficient and share), the stores
        // after them
        double s[3] = {o.su[0][c], o.su[1][c], o.su[2][c]};
        bool own[3], wst[3], shr[3];
        double aown[3], awst[3], tg[3][3], tfe[3][3], tfi[3][3], sh[3][3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const FaceDir D = dir_of(f, q, nj, nij);
            double de, ce, cp, g[3], fde[3], fdi[3];
            own[q] = F[q].has(i, j, k);
            if (own[q]) {
                face_uvw_cds(f, D, o.vis, o.fcr[q], c, gam, de, ce, cp, g, fde, fdi);
                aown[q] = de - ce;
#pragma unroll
                for (int m = 0; m < 3; ++m) { tg[q][m] = g[m]; tfe[q][m] = fde[m]; tfi[q][m] = fdi[m]; }
            }
            const int iw = i - (q == 0), jw = j - (q == 1), kw = k - (q == 2);
            shr[q] = G.has(iw, jw, kw);
            wst[q] = F[q].has(iw, jw, kw);
            if (wst[q]) {
                face_uvw_cds(f, D, o.vis, o.fcr[q], c - D.idew, gam, de, ce, cp, g, fde, fdi);
                awst[q] = de + cp;
#pragma unroll
                for (int m = 0; m < 3; ++m) sh[q][m] = -g[m] - fde[m] + fdi[m];
            } else {
#pragma unroll
                for (int m = 0; m < 3; ++m) sh[q][m] = q == 0 ? scr0[m] : 0.0;
            }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (own[q]) {
                o.a[2 * q][c] = aown[q];
#pragma unroll
                for (int m = 0; m < 3; ++m) s[m] = s[m] + tg[q][m] + tfe[q][m] - tfi[q][m];
            }
            if (wst[q]) o.a[2 * q + 1][c] = awst[q];
            if (shr[q]) {
#pragma unroll
                for (int m = 0; m < 3; ++m) s[m] = s[m] + sh[q][m];
            }
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            o.su[m][c] = s[m];
            o.bp[m][c] = 0.0;
        }
    }
}

// ------------------------------------------------------ ghost-cell values
// update_values_at_ghostcells_nonperiodic.f90: one thread per boundary cell
// of the six faces (west, east, south, north, bottom, top); the six loops of
// the reference write disjoint ghost cells and read only inner cells and the
// ghost cell itself, so they run as one launch.
struct GhostIO {
    double *var[7];          // u, v, w, te, ed, t, vis
    const double *geo[12];   // fx, fy, fz, ar1x..ar3z
    const int *lb[6];        // lbw, lbe, lbs, lbn, lbb, lbt
    double inl[6];           // uin, vin, win, tein, edin, tin
};

__device__ __forceinline__ void ghost_cell(const GhostIO &g, size_t inp, size_t ingc, size_t nb,
                                           double fxp, double fxe, const double *ext_f,
                                           size_t ext_c, int ext_sign, const double *ax,
                                           const double *ay, const double *az, size_t iface,
                                           int lb, bool side_wall_vis)
{
    double *u = g.var[0], *v = g.var[1], *w = g.var[2], *te = g.var[3], *ed = g.var[4],
           *t = g.var[5], *vis = g.var[6];
    const double fxer = 1.0 / fxe;
    if (lb == 1) {   // inlet
        u[ingc] = (g.inl[0] - u[inp] * fxp) * fxer;
        v[ingc] = (g.inl[1] - v[inp] * fxp) * fxer;
        w[ingc] = (g.inl[2] - w[inp] * fxp) * fxer;
        te[ingc] = (g.inl[3] - te[inp] * fxp) * fxer;
        ed[ingc] = (g.inl[4] - ed[inp] * fxp) * fxer;
        t[ingc] = (g.inl[5] - t[inp] * fxp) * fxer;
        vis[ingc] = 0.09 * (te[ingc] * te[ingc]) / ed[ingc];
    } else if (lb == 2) {   // outlet: linear extrapolation from two inner cells
#pragma unroll
        for (int c = 0; c < 7; ++c) {
            double *q = g.var[c];
            q[ingc] = ext_sign < 0 ? q[inp] - (q[inp + nb] - q[inp]) * ext_f[ext_c]
                                   : q[inp] + (q[inp] - q[inp - nb]) * (1.0 - ext_f[ext_c]);
        }
    } else if (lb == 3) {   // symmetry
        const double un = u[inp] * ax[iface] + v[inp] * ay[iface] + w[inp] * az[iface];
        u[ingc] = u[inp] - un * ax[iface];
        v[ingc] = v[inp] - un * ay[iface];
        w[ingc] = w[inp] - un * az[iface];
        te[ingc] = te[inp];
        ed[ingc] = ed[inp];
        t[ingc] = t[inp];
        vis[ingc] = vis[inp];
    } else if (lb == 4) {   // wall
        u[ingc] = (-u[inp] * fxp) * fxer;
        v[ingc] = (-v[inp] * fxp) * fxer;
        w[ingc] = (-w[inp] * fxp) * fxer;
        te[ingc] = (-te[inp] * fxp) * fxer;
        ed[ingc] = ed[inp];
        t[ingc] = (g.inl[5] - t[inp] * fxp) * fxer;
        vis[ingc] = side_wall_vis ? (vis[ingc] - vis[inp] * fxp) * fxer : vis[inp];
    }
}

__global__ __launch_bounds__(256) void k_ghostcells(int ni, int nj, int nk, GhostIO g)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const size_t nJK = (size_t)(nj - 4) * (nk - 4), nIK = (size_t)(ni - 4) * (nk - 4),
                 nIJ = (size_t)(ni - 4) * (nj - 4);
    const size_t total = 2 * (nJK + nIK + nIJ);
    const double *fx = g.geo[0], *fy = g.geo[1], *fz = g.geo[2];
    auto IXd = [&](int i, int j, int k) {
        return (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
    };
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < total; q += (size_t)gridDim.x * 256) {
        size_t r = q;
        if (r < 2 * nJK) {   // west / east: (k, j)
            const bool east = r >= nJK;
            if (east) r -= nJK;
            const int k = 3 + (int)(r / (nj - 4)), j = 3 + (int)(r % (nj - 4));
            const int lb = (east ? g.lb[1] : g.lb[0])[(j - 1) * nk + k - 1];
            if (!east) {
                const size_t inp = IXd(3, j, k);
                const double fxp = fx[inp - nj];
                ghost_cell(g, inp, inp - nj, nj, fxp, 1.0 - fxp, fx, inp, -1, g.geo[3], g.geo[4],
                           g.geo[5], inp - nj, lb, true);
            } else {
                const size_t inp = IXd(nimm, j, k);
                const double fxp = 1.0 - fx[inp];
                ghost_cell(g, inp, inp + nj, nj, fxp, 1.0 - fxp, fx, inp - nj, 1, g.geo[3],
                           g.geo[4], g.geo[5], inp, lb, true);
            }
            continue;
        }
        r -= 2 * nJK;
        if (r < 2 * nIK) {   // south / north: (k, i)
            const bool north = r >= nIK;
            if (north) r -= nIK;
            const int k = 3 + (int)(r / (ni - 4)), i = 3 + (int)(r % (ni - 4));
            const int lb = (north ? g.lb[3] : g.lb[2])[(i - 1) * nk + k - 1];
            if (!north) {
                const size_t inp = IXd(i, 3, k);
                const double fxp = 1.0 - fy[inp];
                ghost_cell(g, inp, inp - 1, 1, fxp, 1.0 - fxp, fy, inp, -1, g.geo[6], g.geo[7],
                           g.geo[8], inp - 1, lb, true);
            } else {
                const size_t inp = IXd(i, njmm, k);
                const double fxp = 1.0 - fy[inp];
                ghost_cell(g, inp, inp + 1, 1, fxp, 1.0 - fxp, fy, inp - 1, 1, g.geo[6], g.geo[7],
                           g.geo[8], inp, lb, true);
            }
            continue;
        }
        r -= 2 * nIK;
        {   // bottom / top: (i, j)
            const bool top = r >= nIJ;
            if (top) r -= nIJ;
            const int i = 3 + (int)(r / (nj - 4)), j = 3 + (int)(r % (nj - 4));
            const int lb = (top ? g.lb[5] : g.lb[4])[(i - 1) * nj + j - 1];
            if (!top) {
                const size_t inp = IXd(i, j, 3), ingc = inp - nij;
                const double fxp = fz[ingc];
                ghost_cell(g, inp, ingc, nij, fxp, 1.0 - fxp, fz, inp, -1, g.geo[9], g.geo[10],
                           g.geo[11], ingc, lb, false);
            } else {
                const size_t inp = IXd(i, j, nkmm);
                const double fxp = 1.0 - fz[inp];
                ghost_cell(g, inp, inp + nij, nij, fxp, 1.0 - fxp, fz, inp - nij, 1, g.geo[9],
                           g.geo[10], g.geo[11], inp, lb, false);
            }
        }
    }
}

// ------------------------------------------------------ Courant number
// CourantNo.h:30-51: per-cell su and per-block partials {sum su (hi, lo),
// sum vol (hi, lo), max su/vol}; k_courant_fin adds the partials in block
// order (compensated sums: the result does not depend on the block count).
__global__ __launch_bounds__(256) void k_courant(int ni, int nj, int nk, const double *__restrict__ f1,
                                                 const double *__restrict__ f2,
                                                 const double *__restrict__ f3,
                                                 const double *__restrict__ vol,
                                                 double *__restrict__ su, double *part)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2;
    const size_t n = (size_t)nij * nk;
    Acc ssu, svol;
    double mx = 0.0;
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        const int k = (int)(c / nij) + 1, r = (int)(c % nij);
        const int i = r / nj + 1, j = r % nj + 1;
        double s = 0.0;
        if (k >= 2 && k <= nk - 1 && i >= 3 && i <= nimm && j >= 3 && j <= njmm) {
            s = fabs(f1[c]) + fabs(f1[c - nj]) + fabs(f2[c]) + fabs(f2[c - 1]) + fabs(f3[c]) +
                fabs(f3[c - nij]);
            const double q = s / vol[c];
            mx = q > mx ? q : mx;
            ssu += s;
            svol += vol[c];
        }
        su[c] = s;
    }
    __shared__ double sh[5][4];
    Acc a = wave_sum(ssu), b = wave_sum(svol);
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = a.hi; sh[1][w] = a.lo; sh[2][w] = b.hi; sh[3][w] = b.lo; sh[4][w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Acc ta, tb;
        double m = 0.0;
        for (int q = 0; q < 4; ++q) {
            Acc x, y;
            x.hi = sh[0][q]; x.lo = sh[1][q]; y.hi = sh[2][q]; y.lo = sh[3][q];
            ta += x;
            tb += y;
            m = sh[4][q] > m ? sh[4][q] : m;
        }
        double *pp = part + (size_t)blockIdx.x * 5;
        pp[0] = ta.hi; pp[1] = ta.lo; pp[2] = tb.hi; pp[3] = tb.lo; pp[4] = m;
    }
}

// CourantNo.h:52-64; out = {CoNum, meanCoNum, timestep, time}
__global__ void k_courant_fin(const double *part, int nblocks, double timestep, double time,
                              int conumfix, double conumfixvalue, double *out)
{
    if (threadIdx.x != 0) return;
    Acc ssu, svol;
    double conum = 0.0;
    for (int b = 0; b < nblocks; ++b) {
        Acc x, y;
        x.hi = part[5 * b]; x.lo = part[5 * b + 1];
        y.hi = part[5 * b + 2]; y.lo = part[5 * b + 3];
        ssu += x;
        svol += y;
        conum = part[5 * b + 4] > conum ? part[5 * b + 4] : conum;
    }
    conum = 0.5 * conum * timestep;
    double mean = 0.5 * ssu.value() / svol.value() * timestep;
    if (conumfix) {
        const double dt = timestep;
        timestep = conumfixvalue / conum * timestep;
        conum = conum * timestep / dt;
        mean = mean * timestep / dt;
    }
    out[0] = conum;
    out[1] = mean;
    out[2] = timestep;
    out[3] = time + timestep;
}

// continuityErrors.h:63-77: global and sum|local| over k=2..nkm, i,j=3..n?mm
__global__ __launch_bounds__(256) void k_continuity(int ni, int nj, int nk,
                                                    const double *__restrict__ f1,
                                                    const double *__restrict__ f2,
                                                    const double *__restrict__ f3, DevState *st,
                                                    double *partials, unsigned *ticket)
{
    const int NI = ni - 4, NJ = nj - 4, NK = nk - 2;
    const size_t n = (size_t)NI * NJ * NK;
    const int nij = ni * nj;
    Acc acc[2];
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        const int k = 2 + (int)(q / ((size_t)NI * NJ));
        const int r = (int)(q % ((size_t)NI * NJ));
        const int i = 3 + r / NJ, j = 3 + r % NJ;
        const size_t c = (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
        const double loc = -f1[c] + f1[c - nj] - f2[c] + f2[c - 1] - f3[c] + f3[c - nij];
        acc[0] += loc;
        acc[1] += fabs(loc);
    }
    grid_reduce<2, F_RAW>(acc, blockIdx.x, gridDim.x, partials, ticket, st);
}

// device copy of a caller array in host-pointer mode (slot-cached buffers of
// `mult` x nxyz doubles); the pointer itself in device-pointer mode
static int stage_n(cap_ctx *ctx, int slot, const double *p, int mult, const double **out)
{
    if (ctx->flags & CAP_DEVICE_POINTERS) { *out = p; return CAP_OK; }
    if (slot < 0 || slot >= CAP_NSTAGE) return CAP_ERR_ARG;
    const size_t bytes = ctx->nstd * sizeof(double) * (size_t)mult;
    if (ctx->Gsz[slot] < bytes) {
        if (ctx->G[slot]) (void)hipFree(ctx->G[slot]);
        ctx->G[slot] = nullptr;
        HIPCHK(hipMalloc(&ctx->G[slot], bytes));
        ctx->Gsz[slot] = bytes;
    }
    if (p) HIPCHK(hipMemcpyAsync(ctx->G[slot], p, bytes, hipMemcpyHostToDevice, ctx->stream));
    *out = ctx->G[slot];
    return CAP_OK;
}

double *cap_slot_buffer(cap_ctx *ctx, int slot)
{
    if (slot < 0 || slot >= CAP_NSTAGE) return nullptr;
    const size_t bytes = ctx->nstd * sizeof(double);
    if (ctx->Gsz[slot] < bytes) {
        if (ctx->G[slot]) (void)hipFree(ctx->G[slot]);
        ctx->G[slot] = nullptr;
        ctx->Gsz[slot] = 0;
        const hipError_t e = hipMalloc(&ctx->G[slot], bytes);
        if (e != hipSuccess) {
            ctx->G[slot] = nullptr;
            snprintf(ctx->err, sizeof ctx->err, "hipMalloc of a staging array failed: %s",
                     hipGetErrorString(e));
            return nullptr;
        }
        ctx->Gsz[slot] = bytes;
    }
    return ctx->G[slot];
}

static int stage(cap_ctx *ctx, int slot, const double *p, const double **out)
{
    return stage_n(ctx, slot, p, 1, out);
}

static int unstage(cap_ctx *ctx, double *host, const double *dev, int mult = 1)
{
    if (!host || (ctx->flags & CAP_DEVICE_POINTERS)) return CAP_OK;
    HIPCHK(hipMemcpyAsync(host, dev, ctx->nstd * 8 * (size_t)mult, hipMemcpyDeviceToHost, ctx->stream));
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
    dim3 grid;
    int kchunk;
    grad_gauss_grid(ctx->ni, ctx->nj, ctx->nk, grid, kchunk);
    hipLaunchKernelGGL(k_grad_gauss, grid, dim3(GG_J * GG_I), 0, ctx->stream, ctx->ni, ctx->nj,
                       ctx->nk, du, g, ox, oy, zbuf, 1, kchunk);
    HIPCHK(hipGetLastError());
    if (!(ctx->flags & CAP_DEVICE_POINTERS)) {
        HIPCHK(hipMemcpyAsync(dudx, ox, ctx->nstd * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(dudy, oy, ctx->nstd * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(dudz, zbuf, ctx->nstd * 8, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

// grad_gauss into the reference's gradu(3, nxyz) (component fastest), device
// pointers, no synchronisation
int cap_grad_gauss_aos(cap_ctx *ctx, const double *u, const double *const *geom, double *g3)
{
    Geo g = {geom[0], geom[1], geom[2], geom[3], geom[4], geom[5], geom[6], geom[7], geom[8],
             geom[9], geom[10], geom[11], geom[12]};
    dim3 grid;
    int kchunk;
    grad_gauss_grid(ctx->ni, ctx->nj, ctx->nk, grid, kchunk);
    hipLaunchKernelGGL(k_grad_gauss, grid, dim3(GG_J * GG_I), 0, ctx->stream, ctx->ni, ctx->nj,
                       ctx->nk, u, g, g3, g3 + 1, g3 + 2, 3, kchunk);
    HIPCHK(hipGetLastError());
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


// ------------------------------------------------ assembly entry points
static int stage_fields(cap_ctx *ctx, const cap_fields *F, Fld &f)
{
    const double *src[31] = {F->x, F->y, F->z, F->xc, F->yc, F->zc, F->vol, F->fx, F->fy, F->fz,
                             F->ar1x, F->ar1y, F->ar1z, F->ar2x, F->ar2y, F->ar2z, F->ar3x,
                             F->ar3y, F->ar3z, F->den, F->apu, F->apv, F->apw, F->u, F->v,
                             F->w, F->p, F->gradu, F->gradv, F->gradw, F->gradp};
    const double *dst[31];
    for (int q = 0; q < 31; ++q) {
        if (!src[q]) { dst[q] = nullptr; continue; }
        int rc = stage_n(ctx, q, src[q], q >= 27 ? 3 : 1, &dst[q]);
        if (rc) return rc;
    }
    f = Fld{dst[0], dst[1], dst[2], dst[3], dst[4], dst[5], dst[6], dst[7], dst[8], dst[9],
            dst[10], dst[11], dst[12], dst[13], dst[14], dst[15], dst[16], dst[17], dst[18],
            dst[19], dst[20], dst[21], dst[22], dst[23], dst[24], dst[25], dst[26], dst[27],
            dst[28], dst[29], dst[30]};
    return CAP_OK;
}

static bool dir_ok(const cap_ctx *ctx, int idew, int idns, int idtb)
{
    const int nj = ctx->nj, nij = ctx->ni * ctx->nj;
    return (idew == nj && idns == 1 && idtb == nij) || (idew == 1 && idns == nij && idtb == nj) ||
           (idew == nij && idns == nj && idtb == 1);
}

static int stage_dir(cap_ctx *ctx, int idew, int idns, int idtb, const double *arvx,
                     const double *arvy, const double *arvz, const double *fif, FaceDir &D)
{
    int rc;
    D.idew = idew; D.idns = idns; D.idtb = idtb;
    if ((rc = stage(ctx, 31, arvx, &D.arvx))) return rc;
    if ((rc = stage(ctx, 32, arvy, &D.arvy))) return rc;
    if ((rc = stage(ctx, 33, arvz, &D.arvz))) return rc;
    return stage(ctx, 34, fif, &D.fif);
}

static int out3(cap_ctx *ctx, double *a, double *b, double *c, double **da, double **db, double **dc)
{
    const double *t;
    int rc;
    if ((rc = stage(ctx, 35, a, &t))) return rc; *da = (double *)t;
    if ((rc = stage(ctx, 36, b, &t))) return rc; *db = (double *)t;
    if ((rc = stage(ctx, 37, c, &t))) return rc; *dc = (double *)t;
    return CAP_OK;
}

static int blocks_for(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }

// ------------------------------------ the three directions fused (drivers)
// Device pointers only; queued on the context's stream, not synchronised.
static dim3 f3_grid(const cap_ctx *ctx)
{
    const int ncb = (int)(((long)ctx->ni * ctx->nj + 255) / 256), per = (ncb + 7) / 8;
    return dim3((unsigned)(8 * per * ((ctx->nk + F3_KSEG - 1) / F3_KSEG)));
}

int cap_fluxmass3_q(cap_ctx *ctx, const cap_fields *F, double *const *a6, double *const *f3)
{
    if (!(ctx->flags & CAP_DEVICE_POINTERS)) return CAP_ERR_ARG;
    Fld f;
    int rc;
    if ((rc = stage_fields(ctx, F, f))) return rc;
    const size_t nf = (size_t)(ctx->nk - 2) * (ctx->ni - 3) * (ctx->nj - 3);
    hipLaunchKernelGGL(k_fluxmass3x, dim3(f3x_blocks((int)((nf + 255) / 256))), dim3(256), 0,
                       ctx->stream, ctx->ni, ctx->nj, ctx->nk, f, a6[0], a6[1], a6[2], a6[3], a6[4],
                       a6[5], f3[0], f3[1], f3[2]);
    HIPCHK(hipGetLastError());
    return CAP_OK;
}

int cap_piso_assemble3_q(cap_ctx *ctx, const cap_fields *F, double *const *a6, double *const *f3,
                         int ipc, double *const *bp3, double *su)
{
    if (!(ctx->flags & CAP_DEVICE_POINTERS)) return CAP_ERR_ARG;
    Fld f;
    int rc;
    if ((rc = stage_fields(ctx, F, f))) return rc;
    const size_t nf = (size_t)(ctx->nk - 3) * (ctx->ni - 3) * (ctx->nj - 3);
    hipLaunchKernelGGL(k_piso3x, dim3(f3x_blocks((int)((nf + 255) / 256))), dim3(256), 0,
                       ctx->stream, ctx->ni, ctx->nj, ctx->nk, f, ipc, a6[0], a6[1], a6[2], a6[3],
                       a6[4], a6[5], f3[0], f3[1], f3[2], bp3[0], bp3[1], bp3[2]);
    hipLaunchKernelGGL(k_piso3_su, dim3(blocks_for(ctx->nstd)), dim3(256), 0, ctx->stream, ctx->ni,
                       ctx->nj, ctx->nk, ipc, bp3[0], bp3[1], bp3[2], su);
    HIPCHK(hipGetLastError());
    return CAP_OK;
}

int cap_fluxuvw3_q(cap_ctx *ctx, const cap_fields *F, const double *vis, double *const *a6,
                   const double *const *f3, double *const *s3, double *const *scr3, double gds)
{
    if (!(ctx->flags & CAP_DEVICE_POINTERS) || ctx->ni < 6 || ctx->nj < 6 || ctx->nk < 6)
        return CAP_ERR_ARG;
    Fld f;
    int rc;
    if ((rc = stage_fields(ctx, F, f))) return rc;
    Out3 o = {};
    for (int q = 0; q < 6; ++q) o.a[q] = a6[q];
    for (int q = 0; q < 3; ++q) { o.fcr[q] = f3[q]; o.su[q] = s3[q]; o.bp[q] = scr3[q]; }
    o.vis = vis;
    hipLaunchKernelGGL(k_uvw3, f3_grid(ctx), dim3(256), 0, ctx->stream, ctx->ni, ctx->nj, ctx->nk,
                       f, o, gds);
    HIPCHK(hipGetLastError());
    return CAP_OK;
}

extern "C" int cap_fluxmass3(cap_ctx *ctx, const cap_fields *F, double *ae, double *aw,
                             double *an, double *as, double *at, double *ab, double *f1,
                             double *f2, double *f3)
{
    if (!ctx || !F || !ae || !aw || !an || !as || !at || !ab || !f1 || !f2 || !f3)
        return CAP_ERR_ARG;
    double *const a6[6] = {ae, aw, an, as, at, ab}, *const fc[3] = {f1, f2, f3};
    int rc;
    if ((rc = cap_fluxmass3_q(ctx, F, a6, fc))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_piso_assemble_pressure_eq3(cap_ctx *ctx, const cap_fields *F, double *ae,
                                              double *aw, double *an, double *as, double *at,
                                              double *ab, double *f1, double *f2, double *f3,
                                              int ipc, double *be, double *bn, double *bt,
                                              double *su)
{
    if (!ctx || !F || !ae || !aw || !an || !as || !at || !ab || !f1 || !f2 || !f3 || !be ||
        !bn || !bt || !su)
        return CAP_ERR_ARG;
    double *const a6[6] = {ae, aw, an, as, at, ab}, *const fc[3] = {f1, f2, f3};
    double *const bp3[3] = {be, bn, bt};
    int rc;
    if ((rc = cap_piso_assemble3_q(ctx, F, a6, fc, ipc, bp3, su))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_fluxuvw3(cap_ctx *ctx, const cap_fields *F, const double *vis, double *ae,
                            double *aw, double *an, double *as, double *at, double *ab,
                            const double *f1, const double *f2, const double *f3, double *su,
                            double *sv, double *sw, double *bp, double *bt, double *bb,
                            int schemes, double gds)
{
    if (!ctx || !F || !vis || !ae || !aw || !an || !as || !at || !ab || !f1 || !f2 || !f3 ||
        !su || !sv || !sw || !bp || !bt || !bb || schemes != CAP_UVW_CDS)
        return CAP_ERR_ARG;
    double *const a6[6] = {ae, aw, an, as, at, ab};
    const double *const fc[3] = {f1, f2, f3};
    double *const s3[3] = {su, sv, sw}, *const scr3[3] = {bp, bt, bb};
    int rc;
    if ((rc = cap_fluxuvw3_q(ctx, F, vis, a6, fc, s3, scr3, gds))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_fluxmass(cap_ctx *ctx, const cap_fields *F, int idew, int idns, int idtb,
                            const double *arvx, const double *arvy, const double *arvz,
                            const double *fif, double *acfe, double *acfw, double *fcf)
{
    if (!ctx || !F || !arvx || !arvy || !arvz || !fif || !acfe || !acfw || !fcf ||
        !dir_ok(ctx, idew, idns, idtb))
        return CAP_ERR_ARG;
    Fld f;
    FaceDir D;
    double *a, *b, *c;
    int rc;
    if ((rc = stage_fields(ctx, F, f)) || (rc = stage_dir(ctx, idew, idns, idtb, arvx, arvy, arvz, fif, D)) ||
        (rc = out3(ctx, acfe, acfw, fcf, &a, &b, &c)))
        return rc;
    const size_t nf = (size_t)(ctx->nk - 2) * (ctx->ni - 3) * (ctx->nj - 3);
    hipLaunchKernelGGL(k_fluxmass, dim3(blocks_for(nf)), dim3(256), 0, ctx->stream, ctx->ni,
                       ctx->nj, ctx->nk, f, D, a, b, c);
    HIPCHK(hipGetLastError());
    if ((rc = unstage(ctx, acfe, a)) || (rc = unstage(ctx, acfw, b)) || (rc = unstage(ctx, fcf, c)))
        return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_piso_assemble_pressure_eq(cap_ctx *ctx, const cap_fields *F, int idew,
                                             int idns, int idtb, const double *arvx,
                                             const double *arvy, const double *arvz,
                                             const double *fif, double *acfe, double *acfw,
                                             double *fcf, int ipc, double *bp, double *su)
{
    if (!ctx || !F || !arvx || !arvy || !arvz || !fif || !acfe || !acfw || !fcf || !bp || !su ||
        !dir_ok(ctx, idew, idns, idtb))
        return CAP_ERR_ARG;
    Fld f;
    FaceDir D;
    double *a, *b, *c;
    const double *tb, *ts;
    int rc;
    if ((rc = stage_fields(ctx, F, f)) || (rc = stage_dir(ctx, idew, idns, idtb, arvx, arvy, arvz, fif, D)) ||
        (rc = out3(ctx, acfe, acfw, fcf, &a, &b, &c)) || (rc = stage(ctx, 38, nullptr, &tb)) ||
        (rc = stage(ctx, 39, su, &ts)))
        return rc;
    double *dbp = (ctx->flags & CAP_DEVICE_POINTERS) ? bp : (double *)tb;
    double *dsu = (double *)ts;
    hipLaunchKernelGGL(k_zero_st, dim3(blocks_for(ctx->nstd)), dim3(256), 0, ctx->stream, dbp, ctx->nstd);
    const size_t nf = (size_t)(ctx->nk - 3) * (ctx->ni - 3) * (ctx->nj - 3);
    hipLaunchKernelGGL(k_piso_assemble, dim3(blocks_for(nf)), dim3(256), 0, ctx->stream, ctx->ni,
                       ctx->nj, ctx->nk, f, D, ipc, a, b, c, dbp, dsu);
    if (ipc > 0) {
        const size_t nc = (size_t)(ctx->ni - 4) * (ctx->nj - 4) * (ctx->nk - 4);
        hipLaunchKernelGGL(k_piso_su, dim3(blocks_for(nc)), dim3(256), 0, ctx->stream, ctx->ni,
                           ctx->nj, ctx->nk, idew, (const double *)dbp, dsu);
    }
    HIPCHK(hipGetLastError());
    if ((rc = unstage(ctx, acfe, a)) || (rc = unstage(ctx, acfw, b)) || (rc = unstage(ctx, fcf, c)) ||
        (rc = unstage(ctx, bp, dbp)) || (rc = unstage(ctx, su, dsu)))
        return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_assemble_pressure_eq(cap_ctx *ctx, const cap_fields *F, int idew, int idns,
                                        int idtb, const double *arvx, const double *arvy,
                                        const double *arvz, const double *fif, double *acfe,
                                        double *acfw, double *fcf, const double *su,
                                        const double *sv, const double *sw)
{
    if (!ctx || !F || !arvx || !arvy || !arvz || !fif || !acfe || !acfw || !fcf || !su || !sv ||
        !sw || !dir_ok(ctx, idew, idns, idtb))
        return CAP_ERR_ARG;
    Fld f;
    FaceDir D;
    double *a, *b, *c;
    const double *dsu, *dsv, *dsw;
    int rc;
    if ((rc = stage_fields(ctx, F, f)) || (rc = stage_dir(ctx, idew, idns, idtb, arvx, arvy, arvz, fif, D)) ||
        (rc = out3(ctx, acfe, acfw, fcf, &a, &b, &c)) || (rc = stage(ctx, 39, su, &dsu)) ||
        (rc = stage(ctx, 40, sv, &dsv)) || (rc = stage(ctx, 41, sw, &dsw)))
        return rc;
    const size_t nf = (size_t)(ctx->nk - 3) * (ctx->ni - 3) * (ctx->nj - 3);
    hipLaunchKernelGGL(k_assemble_p, dim3(blocks_for(nf)), dim3(256), 0, ctx->stream, ctx->ni,
                       ctx->nj, ctx->nk, f, D, dsu, dsv, dsw, a, b, c);
    HIPCHK(hipGetLastError());
    if ((rc = unstage(ctx, acfe, a)) || (rc = unstage(ctx, acfw, b)) || (rc = unstage(ctx, fcf, c)))
        return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_fluxmc(cap_ctx *ctx, const cap_fields *F, int idew, int idns, int idtb,
                          const double *fif, double *fmcor)
{
    if (!ctx || !F || !fif || !fmcor || !dir_ok(ctx, idew, idns, idtb)) return CAP_ERR_ARG;
    Fld f;
    FaceDir D = {idew, idns, idtb, nullptr, nullptr, nullptr, nullptr};
    const double *t;
    int rc;
    if ((rc = stage_fields(ctx, F, f)) || (rc = stage(ctx, 34, fif, &D.fif)) ||
        (rc = stage(ctx, 35, fmcor, &t)))
        return rc;
    hipLaunchKernelGGL(k_fluxmc, dim3(blocks_for(ctx->nstd)), dim3(256), 0, ctx->stream, ctx->ni,
                       ctx->nj, ctx->nk, f, D, (double *)t);
    HIPCHK(hipGetLastError());
    if ((rc = unstage(ctx, fmcor, t))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_fluxuvw(cap_ctx *ctx, const cap_fields *F, const double *vis, int idew,
                           int idns, int idtb, const double *arvx, const double *arvy,
                           const double *arvz, const double *fif, double *acfe, double *acfw,
                           const double *fcf, double *su, double *sv, double *sw, double *bp,
                           double *bt, double *bb, int schemes, double gds)
{
    if (!ctx || !F || !vis || !arvx || !arvy || !arvz || !fif || !acfe || !acfw || !fcf || !su ||
        !sv || !sw || !bp || !bt || !bb || !dir_ok(ctx, idew, idns, idtb) ||
        (schemes & ~(CAP_UVW_ALL | CAP_UVW_FP32)) ||
        ((schemes & CAP_UVW_FP32) && schemes != (CAP_UVW_FP32 | CAP_UVW_CDS)) || ctx->ni < 6 ||
        ctx->nj < 6 || ctx->nk < 6)
        return CAP_ERR_ARG;
    Fld f;
    FaceDir D;
    UvwOut o;
    int rc;
    double *a, *b, *unused;
    const double *t[8];
    if ((rc = stage_fields(ctx, F, f)) || (rc = stage_dir(ctx, idew, idns, idtb, arvx, arvy, arvz, fif, D)) ||
        (rc = out3(ctx, acfe, acfw, nullptr, &a, &b, &unused)))
        return rc;
    const double *src[8] = {vis, fcf, su, sv, sw, bp, bt, bb};
    for (int q = 0; q < 8; ++q)
        if ((rc = stage(ctx, 40 + q, src[q], &t[q]))) return rc;
    o.vis = t[0]; o.fcf = t[1]; o.acfe = a; o.acfw = b;
    o.su = (double *)t[2]; o.sv = (double *)t[3]; o.sw = (double *)t[4];
    o.bp = (double *)t[5]; o.bt = (double *)t[6]; o.bb = (double *)t[7];
    const int dir = idew == ctx->nj ? 1 : (idew == 1 ? 2 : 3);
    const int kst = dir == 3 ? 2 : 3, ist = dir == 1 ? 2 : 3, jst = dir == 2 ? 2 : 3;
    const size_t nf = (size_t)(ctx->nk - 2 - kst + 1) * (ctx->ni - 2 - ist + 1) * (ctx->nj - 2 - jst + 1);
    if (schemes & CAP_UVW_FP32)
        hipLaunchKernelGGL(k_fluxuvw32, dim3(blocks_for(nf)), dim3(256), 0, ctx->stream, ctx->ni,
                           ctx->nj, ctx->nk, f, D, o, (float)gds, kst, ist, jst);
    else
        hipLaunchKernelGGL(k_fluxuvw, dim3(blocks_for(nf)), dim3(256), 0, ctx->stream, ctx->ni,
                           ctx->nj, ctx->nk, f, D, o, schemes, gds, dir, kst, ist, jst);
    const size_t ne = (size_t)(ctx->nk - 3) * (ctx->ni - 3) * (ctx->nj - 3);
    hipLaunchKernelGGL(k_fluxuvw_east, dim3(blocks_for(ne)), dim3(256), 0, ctx->stream, ctx->ni,
                       ctx->nj, ctx->nk, idew, o);
    for (int q = 5; q < 8; ++q)
        hipLaunchKernelGGL(k_zero_st, dim3(blocks_for(ctx->nstd)), dim3(256), 0, ctx->stream,
                           (double *)t[q], ctx->nstd);
    HIPCHK(hipGetLastError());
    if ((rc = unstage(ctx, acfe, a)) || (rc = unstage(ctx, acfw, b)) ||
        (rc = unstage(ctx, su, t[2])) || (rc = unstage(ctx, sv, t[3])) || (rc = unstage(ctx, sw, t[4])) ||
        (rc = unstage(ctx, bp, t[5])) || (rc = unstage(ctx, bt, t[6])) || (rc = unstage(ctx, bb, t[7])))
        return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

// device copy of `bytes` of a caller array (host-pointer mode), or the
// pointer itself (device-pointer mode)
static int stage_bytes(cap_ctx *ctx, int slot, const void *p, size_t bytes, const void **out)
{
    if (ctx->flags & CAP_DEVICE_POINTERS) { *out = p; return CAP_OK; }
    if (slot < 0 || slot >= CAP_NSTAGE) return CAP_ERR_ARG;
    const size_t cap = ctx->nstd * sizeof(double);
    if (bytes > cap) return CAP_ERR_ARG;
    if (ctx->Gsz[slot] < cap) {
        if (ctx->G[slot]) (void)hipFree(ctx->G[slot]);
        ctx->G[slot] = nullptr;
        HIPCHK(hipMalloc(&ctx->G[slot], cap));
        ctx->Gsz[slot] = cap;
    }
    HIPCHK(hipMemcpyAsync(ctx->G[slot], p, bytes, hipMemcpyHostToDevice, ctx->stream));
    *out = ctx->G[slot];
    return CAP_OK;
}

extern "C" int cap_update_ghostcells(cap_ctx *ctx, double *const *var, const double *const *geom,
                                     const cap_bc *bc)
{
    if (!ctx || !var || !geom || !bc || ctx->ni < 6 || ctx->nj < 6 || ctx->nk < 6)
        return CAP_ERR_ARG;
    for (int q = 0; q < 7; ++q) if (!var[q]) return CAP_ERR_ARG;
    for (int q = 0; q < 12; ++q) if (!geom[q]) return CAP_ERR_ARG;
    const int *lbs[6] = {bc->lbw, bc->lbe, bc->lbs, bc->lbn, bc->lbb, bc->lbt};
    for (int q = 0; q < 6; ++q) if (!lbs[q]) return CAP_ERR_ARG;
    const int ni = ctx->ni, nj = ctx->nj, nk = ctx->nk;
    const size_t lbn[6] = {(size_t)nj * nk, (size_t)nj * nk, (size_t)ni * nk, (size_t)ni * nk,
                           (size_t)ni * nj, (size_t)ni * nj};
    GhostIO g;
    int rc;
    for (int q = 0; q < 7; ++q) {
        const double *t;
        if ((rc = stage(ctx, q, var[q], &t))) return rc;
        g.var[q] = (double *)t;
    }
    for (int q = 0; q < 12; ++q)
        if ((rc = stage(ctx, 7 + q, geom[q], &g.geo[q]))) return rc;
    for (int q = 0; q < 6; ++q) {
        const void *t;
        if ((rc = stage_bytes(ctx, 19 + q, lbs[q], lbn[q] * sizeof(int), &t))) return rc;
        g.lb[q] = (const int *)t;
    }
    g.inl[0] = bc->uin; g.inl[1] = bc->vin; g.inl[2] = bc->win;
    g.inl[3] = bc->tein; g.inl[4] = bc->edin; g.inl[5] = bc->tin;
    const size_t total = 2 * ((size_t)(nj - 4) * (nk - 4) + (size_t)(ni - 4) * (nk - 4) +
                              (size_t)(ni - 4) * (nj - 4));
    hipLaunchKernelGGL(k_ghostcells, dim3(blocks_for(total)), dim3(256), 0, ctx->stream, ni, nj, nk, g);
    HIPCHK(hipGetLastError());
    for (int q = 0; q < 7; ++q)
        if ((rc = unstage(ctx, var[q], g.var[q]))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

// update_values_at_ghostcells.f90:3-148, the periodic variant: one thread
// per ghost cell of the i planes (2, nim), the j planes (2, njm) on
// k = 3..nkmm, and of the two wall planes k = 2 and nkm.  Every ghost value
// is read from interior cells (or, vis at the walls, from the ghost cell
// itself), so the cells are independent.
struct PerIO {
    double *v[6];          // u, v, w, te, ed, vis
    const double *fz;
};

__global__ __launch_bounds__(256) void k_ghost_periodic(int ni, int nj, int nk, PerIO g)
{
    const int nij = ni * nj, nis = ni - 4, njs = nj - 4, nks = nk - 4;
    const size_t ni_cnt = 2 * (size_t)njs * nks, nj_cnt = 2 * (size_t)nis * nks;
    const size_t nk_cnt = 2 * (size_t)nis * njs, total = ni_cnt + nj_cnt + nk_cnt;
    auto ix = [&](int i, int j, int k) {
        return (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + (size_t)(j - 1);
    };
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (size_t)gridDim.x * 256) {
        if (t < ni_cnt) {                       // i planes (39-66)
            const int side = (int)(t / ((size_t)njs * nks));
            const size_t r = t % ((size_t)njs * nks);
            const int j = 3 + (int)(r % njs), k = 3 + (int)(r / njs);
            const size_t dst = side ? ix(ni - 1, j, k) : ix(2, j, k);
            const size_t src = side ? ix(3, j, k) : ix(ni - 2, j, k);
            for (int q = 0; q < 6; ++q) g.v[q][dst] = g.v[q][src];
        } else if (t < ni_cnt + nj_cnt) {       // j planes (68-94)
            const size_t u = t - ni_cnt;
            const int side = (int)(u / ((size_t)nis * nks));
            const size_t r = u % ((size_t)nis * nks);
            const int i = 3 + (int)(r % nis), k = 3 + (int)(r / nis);
            const size_t dst = side ? ix(i, nj - 1, k) : ix(i, 2, k);
            const size_t src = side ? ix(i, 3, k) : ix(i, nj - 2, k);
            for (int q = 0; q < 6; ++q) g.v[q][dst] = g.v[q][src];
        } else {                                // walls at k (96-145)
            const size_t u = t - ni_cnt - nj_cnt;
            const int top = (int)(u / ((size_t)nis * njs));
            const size_t r = u % ((size_t)nis * njs);
            const int j = 3 + (int)(r % njs), i = 3 + (int)(r / njs);
            const size_t gc = top ? ix(i, j, nk - 1) : ix(i, j, 2);
            const size_t p = top ? ix(i, j, nk - 2) : ix(i, j, 3);
            double fxp, fxe;
            if (top) { fxe = g.fz[p]; fxp = 1.0 - fxe; }
            else { fxp = g.fz[gc]; fxe = 1.0 - fxp; }
            for (int q = 0; q < 4; ++q) g.v[q][gc] = -g.v[q][p] * fxp / fxe;
            g.v[4][gc] = g.v[4][p];
            g.v[5][gc] = (g.v[5][gc] - g.v[5][p] * fxp) / fxe;
        }
    }
}

extern "C" int cap_update_ghostcells_periodic(cap_ctx *ctx, double *const *var, const double *fz)
{
    if (!ctx || !var || !fz || ctx->ni < 7 || ctx->nj < 7 || ctx->nk < 7) return CAP_ERR_ARG;
    const int qs[6] = {0, 1, 2, 3, 4, 6};       // u, v, w, te, ed, vis of var[7]; t untouched
    for (int q = 0; q < 6; ++q) if (!var[qs[q]]) return CAP_ERR_ARG;
    const int ni = ctx->ni, nj = ctx->nj, nk = ctx->nk;
    PerIO g;
    int rc;
    for (int q = 0; q < 6; ++q) {
        const double *t;
        if ((rc = stage(ctx, q, var[qs[q]], &t))) return rc;
        g.v[q] = (double *)t;
    }
    if ((rc = stage(ctx, 6, fz, &g.fz))) return rc;
    const size_t total = 2 * ((size_t)(nj - 4) * (nk - 4) + (size_t)(ni - 4) * (nk - 4) +
                              (size_t)(ni - 4) * (nj - 4));
    hipLaunchKernelGGL(k_ghost_periodic, dim3(blocks_for(total)), dim3(256), 0, ctx->stream, ni, nj,
                       nk, g);
    HIPCHK(hipGetLastError());
    for (int q = 0; q < 6; ++q)
        if ((rc = unstage(ctx, var[qs[q]], g.v[q]))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_courant(cap_ctx *ctx, const double *f1, const double *f2, const double *f3,
                           const double *vol, double *su, double *timestep, double *time,
                           int conumfix, double conumfixvalue, double *conum)
{
    if (!ctx || !f1 || !f2 || !f3 || !vol || !su || !timestep || !time || !conum)
        return CAP_ERR_ARG;
    const double *d1, *d2, *d3, *dv, *ds;
    int rc;
    if ((rc = stage(ctx, 0, f1, &d1)) || (rc = stage(ctx, 1, f2, &d2)) || (rc = stage(ctx, 2, f3, &d3)) ||
        (rc = stage(ctx, 3, vol, &dv)) || (rc = stage(ctx, 4, nullptr, &ds)))
        return rc;
    double *dsu = (ctx->flags & CAP_DEVICE_POINTERS) ? su : (double *)ds;
    const int nb = blocks_for(ctx->nstd);
    const size_t need = ((size_t)nb * 5 + 8) * sizeof(double);
    if (ctx->scr_bytes < need) {
        if (ctx->scr) (void)hipFree(ctx->scr);
        ctx->scr = nullptr;
        HIPCHK(hipMalloc(&ctx->scr, need));
        ctx->scr_bytes = need;
    }
    double *out_d = ctx->scr + (size_t)nb * 5;
    hipLaunchKernelGGL(k_courant, dim3(nb), dim3(256), 0, ctx->stream, ctx->ni, ctx->nj, ctx->nk,
                       d1, d2, d3, dv, dsu, ctx->scr);
    hipLaunchKernelGGL(k_courant_fin, dim3(1), dim3(64), 0, ctx->stream, (const double *)ctx->scr,
                       nb, *timestep, *time, conumfix, conumfixvalue, out_d);
    HIPCHK(hipGetLastError());
    double out[4];
    HIPCHK(hipMemcpyAsync(out, out_d, sizeof out, hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = unstage(ctx, su, dsu))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    conum[0] = out[0];
    conum[1] = out[1];
    *timestep = out[2];
    *time = out[3];
    return CAP_OK;
}

// ------------------------------------------- allocation-free Courant number
// Scratch of cap_courant_noalloc: k_courant's per-block partials (5 doubles
// for each of at most 8192 blocks, the cap of blocks_for) followed by
// k_courant_fin's 4 results.  Each GPU holds its own copy of this array for
// as long as the process runs.  In host code the name `scratch` is only a
// host-side stand-in: kernels get this device's address of the array, looked
// up once per device by cap_scratch_setup (at cap_init) and kept in
// scratch_ptr.
constexpr size_t CAP_SCRATCH_DOUBLES = (size_t)8192 * 5 + 8;
static __device__ __align__(256) double scratch[CAP_SCRATCH_DOUBLES];

constexpr int CAP_MAX_DEVICES = 64;
static double *scratch_ptr[CAP_MAX_DEVICES];
static std::mutex scratch_mu;

int cap_scratch_setup(int device)
{
    if (device < 0 || device >= CAP_MAX_DEVICES) return CAP_ERR_NODEV;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_ptr[device]) return CAP_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        (void)hipGetLastError();
        return CAP_ERR_HIP;
    }
    scratch_ptr[device] = (double *)p;
    return CAP_OK;
}

extern "C" int cap_courant_noalloc(cap_ctx *ctx, const double *f1, const double *f2,
                                   const double *f3, const double *vol, double *su,
                                   double *timestep, double *time, int conumfix,
                                   double conumfixvalue, double *conum)
{
    if (!ctx || !f1 || !f2 || !f3 || !vol || !su || !timestep || !time || !conum)
        return CAP_ERR_ARG;
    if (!(ctx->flags & CAP_DEVICE_POINTERS)) {
        snprintf(ctx->err, sizeof ctx->err,
                 "cap_courant_noalloc needs a CAP_DEVICE_POINTERS context (host arrays would "
                 "be staged through allocated device buffers)");
        return CAP_ERR_ARG;
    }
    double *scr = (ctx->device >= 0 && ctx->device < CAP_MAX_DEVICES) ? scratch_ptr[ctx->device]
                                                                      : nullptr;
    if (!scr) {
        snprintf(ctx->err, sizeof ctx->err, "no scratch address for device %d", ctx->device);
        return CAP_ERR_HIP;
    }
    const int nb = blocks_for(ctx->nstd);
    const size_t need = (size_t)nb * 5 + 4;
    if (need > CAP_SCRATCH_DOUBLES) {
        snprintf(ctx->err, sizeof ctx->err,
                 "cap_courant_noalloc: %d blocks need %zu scratch doubles, more than its %zu",
                 nb, need, CAP_SCRATCH_DOUBLES);
        return CAP_ERR_ARG;
    }
    // every partial k_courant_fin reads is written by k_courant before it on
    // the same stream, so nothing needs clearing between calls
    double *out_d = scr + (size_t)nb * 5;
    hipLaunchKernelGGL(k_courant, dim3(nb), dim3(256), 0, ctx->stream, ctx->ni, ctx->nj, ctx->nk,
                       f1, f2, f3, vol, su, scr);
    hipLaunchKernelGGL(k_courant_fin, dim3(1), dim3(64), 0, ctx->stream, (const double *)scr, nb,
                       *timestep, *time, conumfix, conumfixvalue, out_d);
    HIPCHK(hipGetLastError());
    double out[4];
    HIPCHK(hipMemcpyAsync(out, out_d, sizeof out, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    conum[0] = out[0];
    conum[1] = out[1];
    *timestep = out[2];
    *time = out[3];
    return CAP_OK;
}

extern "C" int cap_continuity(cap_ctx *ctx, const double *f1, const double *f2, const double *f3,
                              double *sums)
{
    if (!ctx || !f1 || !f2 || !f3 || !sums) return CAP_ERR_ARG;
    const double *d1, *d2, *d3;
    int rc;
    if ((rc = stage(ctx, 35, f1, &d1)) || (rc = stage(ctx, 36, f2, &d2)) || (rc = stage(ctx, 37, f3, &d3)))
        return rc;
    HIPCHK(hipMemsetAsync(ctx->ticket, 0, 64, ctx->stream));
    const size_t n = (size_t)(ctx->ni - 4) * (ctx->nj - 4) * (ctx->nk - 2);
    hipLaunchKernelGGL(k_continuity, dim3(std::min(blocks_for(n), 1024)), dim3(256), 0,
                       ctx->stream, ctx->ni, ctx->nj, ctx->nk, d1, d2, d3, ctx->st,
                       ctx->partials, ctx->ticket);
    HIPCHK(hipGetLastError());
    DevState h;
    HIPCHK(hipMemcpyAsync(ctx->st_host, ctx->st, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    memcpy(&h, ctx->st_host, sizeof h);
    sums[0] = h.red[0];
    sums[1] = h.red[1];
    return CAP_OK;
}

// ------------------------------------------------------ problem generator
// Counter-based generator (splitmix64 finaliser keyed by seed, stream and
// cell index): the same inputs on every machine, independent of how the
// work is split.  Its log, cos(2 pi u) and exp are its own: range reduction
// and fixed-degree polynomials in +, -, *, / (contraction is off for the
// whole library), so any IEEE binary64 machine gives the same bits.
__device__ __forceinline__ uint64_t gen_smix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ uint64_t gen_key(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return gen_smix(gen_smix(seed ^ (stream * 0xD1B54A32D192ED03ull)) + idx);
}

constexpr double GEN_LN2_HI = 6.93147180369123816490e-01;   // low 32 bits zero: e * hi exact
constexpr double GEN_LN2_LO = 1.9082149292705877e-10;

// ln x, x > 0 normal: x = m 2^e, m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s)
// with s = (m - 1) / (m + 1), |s| <= 0.1716, odd series through s^21
__device__ double gen_log(double x)
{
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff) - 1022;
    double m = __longlong_as_double((long long)((b & 0x000FFFFFFFFFFFFFull) | 0x3FE0000000000000ull));
    if (m < 0.70710678118654757) {
        m = m * 2.0;
        e = e - 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    // 1/21, 1/19, ..., 1/3 (Horner in s^2)
    const double cf[10] = {0.047619047619047616, 0.05263157894736842, 0.058823529411764705,
                           0.06666666666666667, 0.07692307692307693, 0.09090909090909091,
                           0.1111111111111111, 0.14285714285714285, 0.2, 0.3333333333333333};
    double t = cf[0];
    for (int q = 1; q < 10; ++q) t = cf[q] + z * t;
    const double lm = 2.0 * s + 2.0 * s * (z * t);
    const double de = (double)e;
    return de * GEN_LN2_HI + (de * GEN_LN2_LO + lm);
}

// cos(2 pi u), u in [0, 1): v = |u - rint(u)|, quadrant q = rint(4 v),
// r = v - q / 4 in [-1/8, 1/8] (all exact), theta = 2 pi r; Taylor series
// of cos through theta^18 and of sin through theta^19
__device__ double gen_cos2pi(double u)
{
    double v = u - rint(u);
    if (v < 0.0) v = -v;
    const double q = rint(4.0 * v);
    const double r = v - 0.25 * q;
    const double th = 6.283185307179586 * r;
    const double z = th * th;
    // (-1)^n / (2n)!, n = 9..1, and (-1)^n / (2n+1)!, n = 9..1
    const double cc[9] = {-1.5619206968586225e-16, 4.779477332387385e-14, -1.1470745597729725e-11,
                          2.08767569878681e-09, -2.755731922398589e-07, 2.48015873015873e-05,
                          -0.001388888888888889, 0.041666666666666664, -0.5};
    const double cs[9] = {-8.22063524662433e-18, 2.8114572543455206e-15, -7.647163731819816e-13,
                          1.6059043836821613e-10, -2.505210838544172e-08, 2.7557319223985893e-06,
                          -0.0001984126984126984, 0.008333333333333333, -0.16666666666666666};
    double c = cc[0], sn = cs[0];
    for (int k = 1; k < 9; ++k) {
        c = cc[k] + z * c;
        sn = cs[k] + z * sn;
    }
    c = 1.0 + z * c;
    sn = th + th * (z * sn);
    if (q == 0.0) return c;
    if (q == 1.0) return -sn;
    return -c;
}

// e^x, |x| < 700: k = rint(x / ln 2), r = x - k ln 2 (two-part ln 2),
// |r| <= 0.347, Taylor series through r^13, times 2^k built from its bits
__device__ double gen_exp(double x)
{
    const double k = rint(x * 1.4426950408889634);
    const double r = (x - k * GEN_LN2_HI) - k * GEN_LN2_LO;
    // 1/13!, 1/12!, ..., 1/2!, 1/1!, 1/0!
    const double cf[14] = {1.6059043836821613e-10, 2.08767569878681e-09, 2.505210838544172e-08,
                           2.755731922398589e-07, 2.7557319223985893e-06, 2.48015873015873e-05,
                           0.0001984126984126984, 0.001388888888888889, 0.008333333333333333,
                           0.041666666666666664, 0.16666666666666666, 0.5, 1.0, 1.0};
    double p = cf[0];
    for (int q = 1; q < 14; ++q) p = cf[q] + r * p;
    return p * __longlong_as_double(((long long)k + 1023) << 52);
}

// Box-Muller on two counter-keyed uniforms, u1 in (0, 1], u2 in [0, 1)
__device__ double gen_normal(uint64_t seed, uint64_t stream, uint64_t idx)
{
    const double u1 = (double)((gen_key(seed, stream, 2 * idx) >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(gen_key(seed, stream, 2 * idx + 1) >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * gen_log(u1)) * gen_cos2pi(u2);
}

// One thread per cell: ghost rows a* = 0, ap = 1, su = 0; an interior cell's
// faces to interior neighbours a = h exp(sigma g), g keyed by (seed,
// direction, index of the face's west / south / bottom cell), so that
// aw(E) = ae(P) etc.; ap = ae + aw + an + as + at + ab; su = g (stream 4),
// its mean removed by k_gen_center.
__global__ __launch_bounds__(256) void k_gen_lognormal(int ni, int nj, int nk, uint64_t seed,
                                                       double sigma, double h,
                                                       double *__restrict__ ae, double *__restrict__ aw,
                                                       double *__restrict__ an, double *__restrict__ as,
                                                       double *__restrict__ at, double *__restrict__ ab,
                                                       double *__restrict__ ap, double *__restrict__ su)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const size_t n = (size_t)nij * nk;
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        const int k = (int)(c / nij) + 1, rr = (int)(c % nij);
        const int i = rr / nj + 1, j = rr % nj + 1;
        double e = 0.0, w = 0.0, no = 0.0, so = 0.0, t = 0.0, bo = 0.0, p = 1.0, s = 0.0;
        if (k >= 3 && k <= nkmm && i >= 3 && i <= nimm && j >= 3 && j <= njmm) {
            if (i < nimm) e = h * gen_exp(sigma * gen_normal(seed, 1, c));
            if (i > 3) w = h * gen_exp(sigma * gen_normal(seed, 1, c - nj));
            if (j < njmm) no = h * gen_exp(sigma * gen_normal(seed, 2, c));
            if (j > 3) so = h * gen_exp(sigma * gen_normal(seed, 2, c - 1));
            if (k < nkmm) t = h * gen_exp(sigma * gen_normal(seed, 3, c));
            if (k > 3) bo = h * gen_exp(sigma * gen_normal(seed, 3, c - nij));
            p = e + w + no + so + t + bo;
            s = gen_normal(seed, 4, c);
        }
        ae[c] = e; aw[c] = w; an[c] = no; as[c] = so; at[c] = t; ab[c] = bo; ap[c] = p; su[c] = s;
    }
}

// Sum of su over each interior (k, i) row, j in order (one thread a row).
__global__ __launch_bounds__(256) void k_gen_rows(int ni, int nj, int nk, const double *__restrict__ su,
                                                  double *__restrict__ rows)
{
    const int nis = ni - 4, njs = nj - 4, nks = nk - 4, nij = ni * nj;
    const int nrow = nis * nks;
    for (int r = blockIdx.x * 256 + threadIdx.x; r < nrow; r += gridDim.x * 256) {
        const int k = r / nis + 3, i = r % nis + 3;
        const size_t c0 = (size_t)(k - 1) * nij + (size_t)(i - 1) * nj + 2;   // j = 3
        double row = 0.0;
        for (int j = 0; j < njs; ++j) row += su[c0 + j];
        rows[r] = row;
    }
}

// The rows of each plane in i order (one thread a plane), then the planes in
// k order: mean = sum / count, stored in rows[nrow + nks].
__global__ __launch_bounds__(256) void k_gen_mean(int ni, int nj, int nk, double *rows)
{
    const int nis = ni - 4, njs = nj - 4, nks = nk - 4;
    const int nrow = nis * nks;
    double *planes = rows + nrow;
    for (int k = threadIdx.x; k < nks; k += blockDim.x) {
        double plane = 0.0;
        for (int i = 0; i < nis; ++i) plane += rows[k * nis + i];
        planes[k] = plane;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sum = 0.0;
        for (int k = 0; k < nks; ++k) sum += planes[k];
        planes[nks] = sum / (double)((long)nis * njs * nks);
    }
}

// su -= mean on the interior; then the reference-pressure cell (3,3,3)
// pinned as PISO_multiple_correction.f90:117-129 does (a_nb = 0, ap = 1,
// su = 0; its neighbours keep their faces to it).
__global__ __launch_bounds__(256) void k_gen_center(int ni, int nj, int nk, const double *mean_p,
                                                    double *__restrict__ ae, double *__restrict__ aw,
                                                    double *__restrict__ an, double *__restrict__ as,
                                                    double *__restrict__ at, double *__restrict__ ab,
                                                    double *__restrict__ ap, double *__restrict__ su)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const size_t n = (size_t)nij * nk;
    const size_t ref = (size_t)2 * nij + (size_t)2 * nj + 2;
    const double mean = *mean_p;
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
        const int k = (int)(c / nij) + 1, rr = (int)(c % nij);
        const int i = rr / nj + 1, j = rr % nj + 1;
        if (!(k >= 3 && k <= nkmm && i >= 3 && i <= nimm && j >= 3 && j <= njmm)) continue;
        if (c == ref) {
            ae[c] = aw[c] = an[c] = as[c] = at[c] = ab[c] = 0.0;
            ap[c] = 1.0;
            su[c] = 0.0;
        } else {
            su[c] = su[c] - mean;
        }
    }
}

// Log-normal Neumann p' operator on a uniform grid: interior faces
// a = h exp(sigma g), g ~ N(0,1) keyed by (seed, direction, cell); aw(E)=ae(P)
// etc.; boundary faces 0; ap = sum a_nb; su ~ N(0,1) with zero mean; ghost
// rows a* = 0, ap = 1, su = 0; reference-pressure cell (3,3,3) pinned.
// Everything runs on the device; the mean is summed in a fixed order (rows,
// planes, total), so the result is the same bits on every run.
extern "C" int cap_build_lognormal(cap_ctx *ctx, unsigned long long seed, double sigma,
                                   cap_coef *c)
{
    if (!ctx || !c || !c->ae || !c->aw || !c->an || !c->as || !c->at || !c->ab || !c->ap || !c->su)
        return CAP_ERR_ARG;
    const int ni = ctx->ni, nj = ctx->nj, nk = ctx->nk;
    if (ni < 7 || nj < 7 || nk < 7) return CAP_ERR_ARG;   // at least 3 interior cells each way
    const double h = 1.0 / (double)(ni - 4);
    double *out[8] = {c->ae, c->aw, c->an, c->as, c->at, c->ab, c->ap, c->su};
    double *d[8];
    int rc;
    for (int q = 0; q < 8; ++q) {
        const double *p;
        if ((rc = stage(ctx, q, nullptr, &p))) return rc;   // device pointers: the caller's array
        d[q] = (ctx->flags & CAP_DEVICE_POINTERS) ? out[q] : (double *)p;
    }
    const size_t nrow = (size_t)(ni - 4) * (nk - 4);
    double *rows;
    {   // row sums, plane sums and the mean, in staging slot 47
        const size_t bytes = (nrow + (nk - 4) + 1) * sizeof(double);
        if (ctx->Gsz[47] < bytes) {
            if (ctx->G[47]) (void)hipFree(ctx->G[47]);
            ctx->G[47] = nullptr;
            ctx->Gsz[47] = 0;
            HIPCHK(hipMalloc(&ctx->G[47], bytes));
            ctx->Gsz[47] = bytes;
        }
        rows = ctx->G[47];
    }
    hipLaunchKernelGGL(k_gen_lognormal, dim3(blocks_for(ctx->nstd)), dim3(256), 0, ctx->stream, ni,
                       nj, nk, (uint64_t)seed, sigma, h, d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7]);
    hipLaunchKernelGGL(k_gen_rows, dim3(blocks_for(nrow)), dim3(256), 0, ctx->stream, ni, nj, nk,
                       (const double *)d[7], rows);
    hipLaunchKernelGGL(k_gen_mean, dim3(1), dim3(256), 0, ctx->stream, ni, nj, nk, rows);
    hipLaunchKernelGGL(k_gen_center, dim3(blocks_for(ctx->nstd)), dim3(256), 0, ctx->stream, ni, nj,
                       nk, (const double *)(rows + nrow + (nk - 4)), d[0], d[1], d[2], d[3], d[4], d[5],
                       d[6], d[7]);
    HIPCHK(hipGetLastError());
    for (int q = 0; q < 8; ++q)
        if ((rc = unstage(ctx, out[q], d[q]))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

void cap_free_extra(cap_ctx *ctx)
{
    if (ctx->scr) (void)hipFree(ctx->scr);
    for (int q = 0; q < CAP_NSTAGE; ++q) if (ctx->G[q]) (void)hipFree(ctx->G[q]);
    for (int q = 0; q < F_COUNT; ++q) if (ctx->F[q]) (void)hipFree(ctx->F[q]);
    if (ctx->pkd) (void)hipFree(ctx->pkd);
    if (ctx->pkb) (void)hipFree(ctx->pkb);
    if (ctx->pkf) (void)hipFree(ctx->pkf);
}
