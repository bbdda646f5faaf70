// cap_geometry.hip -- the grid's derived geometry on the device, and the
// reader of the reference's unformatted grid file.
//
//   cap_face_areas   face area vectors from the nodes     (init.f90:49-144)
//   cap_intfac       interpolation factors fx, fy, fz      (intfac.f90:3-51,
//                    find_intersection_point.f90:3-76, as init.f90:147-149
//                    calls them)
//   cap_grid_dims /  the grid file of read_grid.f90:126-161 (host code)
//   cap_read_grid
//
// One thread per cell of each loop range; every value is formed with the
// reference's operations in its order (contraction is off for the library),
// so the results are bit-identical to the reference routines.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <vector>
#include <algorithm>
#include "cap_amd.h"
#include "cap_ctx.h"

#define HIPCHK(call)                                                                      \
    do {                                                                                  \
        hipError_t _e = (call);                                                           \
        if (_e != hipSuccess) {                                                           \
            snprintf(ctx->err, sizeof ctx->err, "%s failed: %s (%s:%d)", #call,           \
                     hipGetErrorString(_e), __FILE__, __LINE__);                          \
            return CAP_ERR_HIP;                                                           \
        }                                                                                 \
    } while (0)

namespace {

struct Box {
    int k0, i0, j0, nk, ni, nj;   // first cell (1-based) and counts of a loop range
};

__device__ __forceinline__ bool cell_of(const Box &b, int ni, int nj, size_t t, size_t *p)
{
    const size_t n = (size_t)b.nk * b.ni * b.nj;
    if (t >= n) return false;
    const int j = b.j0 + (int)(t % b.nj);
    const int i = b.i0 + (int)((t / b.nj) % b.ni);
    const int k = b.k0 + (int)(t / ((size_t)b.nj * b.ni));
    *p = (size_t)(k - 1) * ni * nj + (size_t)(i - 1) * nj + (size_t)(j - 1);
    return true;
}

// The face of cell P spanned by its corner P, the corner one step back along
// idns and one back along idtb: the cross product of its two mean edge
// vectors (init.f90:59-77 and the two loops after it).
__global__ __launch_bounds__(256) void k_face_areas(int ni, int nj, Box b, int idns, int idtb,
                                                    const double *__restrict__ x,
                                                    const double *__restrict__ y,
                                                    const double *__restrict__ z,
                                                    double *__restrict__ arx,
                                                    double *__restrict__ ary,
                                                    double *__restrict__ arz)
{
    const size_t n = (size_t)b.nk * b.ni * b.nj;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256) {
        size_t p;
        cell_of(b, ni, nj, t, &p);
        const size_t s = p - idns, bb = p - idtb, bs = bb - idns;
        const double xe = 0.5 * (x[p] - x[s] + x[bb] - x[bs]);
        const double ye = 0.5 * (y[p] - y[s] + y[bb] - y[bs]);
        const double ze = 0.5 * (z[p] - z[s] + z[bb] - z[bs]);
        const double xd = 0.5 * (x[p] - x[bb] + x[s] - x[bs]);
        const double yd = 0.5 * (y[p] - y[bb] + y[s] - y[bs]);
        const double zd = 0.5 * (z[p] - z[bb] + z[s] - z[bs]);
        arx[p] = ye * zd - yd * ze;
        ary[p] = xd * ze - xe * zd;
        arz[p] = xe * yd - ye * xd;
    }
}

// Where the line of centres P-E crosses the plane of the face's corners
// (find_intersection_point.f90:60-69: both determinants expanded along their
// first row, summed in the reference's order), then the interpolation factor
// |crossing - P| / (|E - P| + small) (intfac.f90:42-44, small in the z term
// of the denominator as written).
__global__ __launch_bounds__(256) void k_intfac(int ni, int nj, Box b, int idew, int idns,
                                                int idtb, const double *__restrict__ x,
                                                const double *__restrict__ y,
                                                const double *__restrict__ z,
                                                const double *__restrict__ xc,
                                                const double *__restrict__ yc,
                                                const double *__restrict__ zc,
                                                double *__restrict__ lambda)
{
    const double small = (double)1e-30f;
    const size_t n = (size_t)b.nk * b.ni * b.nj;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256) {
        size_t p;
        cell_of(b, ni, nj, t, &p);
        const size_t e = p + idew;
        const double x1 = x[p], y1 = y[p], z1 = z[p];
        const double x2 = x[p - idns], y2 = y[p - idns], z2 = z[p - idns];
        const double x3 = x[p - idtb], y3 = y[p - idtb], z3 = z[p - idtb];
        const double x4 = xc[p], y4 = yc[p], z4 = zc[p];
        const double dx = xc[e] - x4, dy = yc[e] - y4, dz = zc[e] - z4;
        const double m34 = y3 * z4 - y4 * z3, m24 = y2 * z4 - y4 * z2, m14 = y1 * z4 - y4 * z1;
        const double m23 = y2 * z3 - y3 * z2, m13 = y1 * z3 - y3 * z1, m12 = y1 * z2 - y2 * z1;
        const double num = x2 * m34 - x1 * m34 - x3 * m24 + x1 * m24 + x3 * m14 - x2 * m14 +
                           x4 * m23 - x1 * m23 - x4 * m13 + x2 * m13 + x4 * m12 - x3 * m12;
        const double d3 = y3 * dz - dy * z3, d2 = y2 * dz - dy * z2, d1 = y1 * dz - dy * z1;
        const double den = x2 * d3 - x1 * d3 - x3 * d2 + x1 * d2 + x3 * d1 - x2 * d1 + dx * m23 -
                           dx * m13 + dx * m12;
        const double tt = -num / den;
        const double fxp = x4 + dx * tt, fyp = y4 + dy * tt, fzp = z4 + dz * tt;
        const double ax = fxp - x4, ay = fyp - y4, az = fzp - z4;
        const double bx = xc[e] - x4, by = yc[e] - y4, bz = zc[e] - z4 + small;
        lambda[p] = sqrt(ax * ax + ay * ay + az * az) / (sqrt(bx * bx + by * by + bz * bz) + small);
    }
}

int grid_blocks(const Box &b)
{
    const size_t n = (size_t)b.nk * b.ni * b.nj;
    return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192));
}

// device copy of a caller array (host-pointer contexts), or the array itself
struct Dev {
    cap_ctx *ctx;
    std::vector<double *> own;
    ~Dev() { for (double *p : own) (void)hipFree(p); }
    int in(const double *h, const double **out)
    {
        if (ctx->flags & CAP_DEVICE_POINTERS) { *out = h; return CAP_OK; }
        double *d = nullptr;
        HIPCHK(hipMalloc(&d, ctx->nstd * sizeof(double)));
        own.push_back(d);
        HIPCHK(hipMemcpyAsync(d, h, ctx->nstd * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        *out = d;
        return CAP_OK;
    }
    // in/out array: uploaded (cells outside the loop ranges keep the caller's values)
    int io(double *h, double **out)
    {
        const double *p;
        int rc = in(h, &p);
        *out = (double *)p;
        return rc;
    }
    int back(double *h, const double *d)
    {
        if (ctx->flags & CAP_DEVICE_POINTERS) return CAP_OK;
        HIPCHK(hipMemcpyAsync(h, d, ctx->nstd * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        return CAP_OK;
    }
};

}  // namespace

extern "C" int cap_face_areas(cap_ctx *ctx, const double *x, const double *y, const double *z,
                              double *const *ar)
{
    if (!ctx || !x || !y || !z || !ar) return CAP_ERR_ARG;
    for (int q = 0; q < 9; ++q) if (!ar[q]) return CAP_ERR_ARG;
    const int ni = ctx->ni, nj = ctx->nj, nk = ctx->nk, nij = ni * nj;
    Dev D{ctx, {}};
    const double *dx, *dy, *dz;
    double *da[9];
    int rc;
    if ((rc = D.in(x, &dx)) || (rc = D.in(y, &dy)) || (rc = D.in(z, &dz))) return rc;
    for (int q = 0; q < 9; ++q) if ((rc = D.io(ar[q], &da[q]))) return rc;
    // east k = 2..nkm, i = 1..nim, j = 2..njm; north k = 2..nkm, i = 2..nim,
    // j = 1..njm; top k = 1..nkm, i = 2..nim, j = 2..njm
    const Box be{2, 1, 2, nk - 2, ni - 1, nj - 2};
    const Box bn{2, 2, 1, nk - 2, ni - 2, nj - 1};
    const Box bt{1, 2, 2, nk - 1, ni - 2, nj - 2};
    hipLaunchKernelGGL(k_face_areas, dim3(grid_blocks(be)), dim3(256), 0, ctx->stream, ni, nj, be,
                       1, nij, dx, dy, dz, da[0], da[1], da[2]);
    hipLaunchKernelGGL(k_face_areas, dim3(grid_blocks(bn)), dim3(256), 0, ctx->stream, ni, nj, bn,
                       nj, nij, dx, dy, dz, da[3], da[4], da[5]);
    hipLaunchKernelGGL(k_face_areas, dim3(grid_blocks(bt)), dim3(256), 0, ctx->stream, ni, nj, bt,
                       1, nj, dx, dy, dz, da[6], da[7], da[8]);
    HIPCHK(hipGetLastError());
    for (int q = 0; q < 9; ++q) if ((rc = D.back(ar[q], da[q]))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

extern "C" int cap_intfac(cap_ctx *ctx, const double *const *nodes_centres, double *fx,
                          double *fy, double *fz)
{
    if (!ctx || !nodes_centres || !fx || !fy || !fz) return CAP_ERR_ARG;
    for (int q = 0; q < 6; ++q) if (!nodes_centres[q]) return CAP_ERR_ARG;
    const int ni = ctx->ni, nj = ctx->nj, nk = ctx->nk, nij = ni * nj;
    Dev D{ctx, {}};
    const double *g[6];
    double *f[3];
    double *host[3] = {fx, fy, fz};
    int rc;
    for (int q = 0; q < 6; ++q) if ((rc = D.in(nodes_centres[q], &g[q]))) return rc;
    for (int q = 0; q < 3; ++q) if ((rc = D.io(host[q], &f[q]))) return rc;
    const Box b{2, 2, 2, nk - 2, ni - 2, nj - 2};      // k = 2..nkm, i = 2..nim, j = 2..njm
    // (idew, idns, idtb) of init.f90:147-149: (nj, 1, nij), (1, nj, nij), (nij, 1, nj)
    const int dirs[3][3] = {{nj, 1, nij}, {1, nj, nij}, {nij, 1, nj}};
    for (int d = 0; d < 3; ++d)
        hipLaunchKernelGGL(k_intfac, dim3(grid_blocks(b)), dim3(256), 0, ctx->stream, ni, nj, b,
                           dirs[d][0], dirs[d][1], dirs[d][2], g[0], g[1], g[2], g[3], g[4], g[5],
                           f[d]);
    HIPCHK(hipGetLastError());
    for (int q = 0; q < 3; ++q) if ((rc = D.back(host[q], f[q]))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CAP_OK;
}

// ------------------------------------------------------------ grid file
// Fortran unformatted sequential records: each record is a 4-byte length
// marker, the data, and the same marker again.  A record longer than
// 2^31 - 1 bytes is split into subrecords whose leading marker is negated
// while more subrecords follow.
namespace {

struct RecReader {
    FILE *f = nullptr;
    char err[160] = {0};
    // append the next record's bytes to buf; false at an error
    bool next(std::vector<unsigned char> &buf)
    {
        buf.clear();
        for (;;) {
            int32_t head, tail;
            if (fread(&head, 4, 1, f) != 1) { snprintf(err, sizeof err, "unexpected end of file"); return false; }
            const bool more = head < 0;
            const size_t len = (size_t)(more ? -(int64_t)head : head);
            const size_t at = buf.size();
            buf.resize(at + len);
            if (len && fread(buf.data() + at, 1, len, f) != len) {
                snprintf(err, sizeof err, "record shorter than its marker (%zu bytes)", len);
                return false;
            }
            if (fread(&tail, 4, 1, f) != 1 || (size_t)(tail < 0 ? -(int64_t)tail : tail) != len) {
                snprintf(err, sizeof err, "record markers do not match");
                return false;
            }
            if (!more) return true;
        }
    }
};

}  // namespace

extern "C" int cap_grid_dims(const char *path, int *ni, int *nj, int *nk)
{
    if (!path || !ni || !nj || !nk) return CAP_ERR_ARG;
    RecReader r;
    r.f = fopen(path, "rb");
    if (!r.f) return CAP_ERR_ARG;
    std::vector<unsigned char> b;
    const bool ok = r.next(b) && b.size() == 12;
    fclose(r.f);
    if (!ok) return CAP_ERR_ARG;
    int32_t v[3];
    memcpy(v, b.data(), 12);
    *ni = v[0]; *nj = v[1]; *nk = v[2];
    return (v[0] > 0 && v[1] > 0 && v[2] > 0) ? CAP_OK : CAP_ERR_ARG;
}

extern "C" int cap_read_grid(const char *path, int ni, int nj, int nk, double *const *arrays,
                             int *const *lb, int *flags)
{
    if (!path || !arrays || ni < 1 || nj < 1 || nk < 1) return CAP_ERR_ARG;
    RecReader r;
    r.f = fopen(path, "rb");
    if (!r.f) return CAP_ERR_ARG;
    std::vector<unsigned char> b;
    int rc = CAP_OK;
    const size_t nij = (size_t)ni * nj, nik = (size_t)ni * nk, njk = (size_t)nj * nk;
    const size_t nijk = nij * nk;
    do {
        int32_t d[3];
        if (!r.next(b) || b.size() != 12) { rc = CAP_ERR_ARG; break; }
        memcpy(d, b.data(), 12);
        if (d[0] != ni || d[1] != nj || d[2] != nk) { rc = CAP_ERR_ARG; break; }
        // nij, nik, njk, nijk, ijs, iks, jks, ijks; lbw, lbe (njk), lbs, lbn
        // (nik), lbb, lbt (nij); low, loe, los, lon, lob, lot, lout,
        // lthermbc, lconbc (default logicals, 4 bytes)
        const size_t nints = 8 + 2 * njk + 2 * nik + 2 * nij + 9;
        if (!r.next(b) || b.size() != nints * 4) { rc = CAP_ERR_ARG; break; }
        std::vector<int32_t> iv(nints);
        memcpy(iv.data(), b.data(), nints * 4);
        if ((size_t)iv[0] != nij || (size_t)iv[1] != nik || (size_t)iv[2] != njk ||
            (size_t)iv[3] != nijk) { rc = CAP_ERR_ARG; break; }
        size_t at = 8;
        const size_t lens[6] = {njk, njk, nik, nik, nij, nij};
        for (int q = 0; q < 6; ++q) {
            if (lb && lb[q]) memcpy(lb[q], iv.data() + at, lens[q] * 4);
            at += lens[q];
        }
        if (flags)
            for (int q = 0; q < 9; ++q) flags[q] = iv[at + q] != 0;
        // x, y, z, xc, yc, zc, vol, fx, fy, fz, ar1x .. ar3z
        for (int q = 0; q < 19; ++q) {
            if (!r.next(b) || b.size() != nijk * 8) { rc = CAP_ERR_ARG; break; }
            if (arrays[q]) memcpy(arrays[q], b.data(), nijk * 8);
        }
    } while (0);
    fclose(r.f);
    return rc;
}
