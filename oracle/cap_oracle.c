/*
 * cap_oracle.c -- CPU oracle (test infrastructure) for the p' hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product library never
 * links or calls it.
 *
 * Each routine restates one reference routine with the reference's loop
 * ranges, loop order (k outer, i middle, j inner = unit stride), operand
 * order and single-precision literals promoted to double:
 *   small = 1e-30   (modules_allocatable.f90:39)  -> (double)1e-30f
 *   s0    = 1.e20   (pcg-sip.f90:89, iccg.f90:83)  -> (double)1e20f
 * Compiled with -O2 -ffp-contract=off so no multiply-add is fused, which is
 * what the reference's -O2 x86-64 build does (no FMA in the base ISA).
 *
 * Parity pinning (tests/test_oracle.py):
 *  - orc_sipsol, orc_cgstab_sip, orc_grad_gauss, orc_bpres: against golden
 *    vectors written by the reference Fortran compiled with amdflang
 *    (oracle/Makefile target `ref`, fixtures in tests/golden/).
 *  - orc_pcgsip, orc_iccg: the reference files call global_sum / exchange,
 *    whose only definitions need mpif.h (absent), so they are not built.
 *    Pinned instead by: (1) the SIP factors they share with sipsol/cgstab_sip
 *    (golden); (2) forward/backward sweeps == dense solve with the assembled
 *    L and U (numpy); (3) the PCG scalar recurrence against an independent
 *    numpy PCG using dense L U as the preconditioner (iteration history to
 *    1e-12); (4) the IC(0) diagonal/sweeps against a dense IC(0) of the same
 *    7-diagonal matrix.
 */
#include "cap_oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>
#include <time.h>

#define SMALL ((double)1e-30f)

/* Summation of the solvers' scalar products and L1 norms.  0 (default): the
 * reference's single running sum in loop order.  1: one running sum per
 * k-plane, planes added in order.  2: a compensated running sum (every
 * addition is Knuth's TwoSum, the rounding errors are carried in a second
 * word), read as hi + lo: the nearly exact sum, whatever the order -- what
 * the GPU's reductions compute.  Modes 1 and 2 are used by the tests: mode 2
 * reproduces the GPU's arithmetic, and the distance between modes measures
 * how far a solve moves when only the summation changes. */
static int g_plane_sums = 0;
void orc_set_plane_sums(int on) { g_plane_sums = on; }

typedef struct { double tot, part; int k; double hi, lo; } psum;
#define PS0 {0.0, 0.0, -1, 0.0, 0.0}
/* adds v at plane k; returns the sum so far (bit-identical to `s = s + v`
 * from 0.0 in mode 0) */
static inline double ps_add(psum *ps, double v, int k)
{
    if (!g_plane_sums) { ps->part = ps->part + v; return ps->part; }
    if (g_plane_sums == 2) {
        const double s = ps->hi + v;
        const double bp = s - ps->hi;
        ps->lo += (ps->hi - (s - bp)) + (v - bp);
        ps->hi = s;
        return ps->hi + ps->lo;
    }
    if (k != ps->k) {
        if (ps->k != -1) ps->tot = ps->tot + ps->part;
        ps->part = 0.0;
        ps->k = k;
    }
    ps->part = ps->part + v;
    return ps->tot + ps->part;
}
#define S0INIT ((double)1e20f)

/* Fortran MIN/MAX of two ordered values (no NaN inputs) */
static inline double orc_min2(double a, double b) { return b < a ? b : a; }
static inline double orc_max2(double a, double b) { return b > a ? b : a; }

/* 1-based (i,j,k) -> C offset; li/lk of setind.f90:34-40 */
#define IX(i, j, k) ((size_t)((k)-1) * (size_t)nij + (size_t)((i)-1) * (size_t)nj + (size_t)((j)-1))

/* ------------------------------------------------------------------ RNG -- */
static inline uint64_t smix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint64_t orc_hash(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return smix(smix(seed ^ (stream * 0xD1B54A32D192ED03ull)) + idx);
}

double orc_u01(uint64_t seed, uint64_t stream, uint64_t idx)
{
    return (double)(orc_hash(seed, stream, idx) >> 11) * 0x1.0p-53;
}

/*
 * The generator's own log, cos(2 pi u) and exp: range reduction plus
 * fixed-degree polynomials in +, -, *, / only (no libm), so that any machine
 * with IEEE binary64 arithmetic and no contraction -- this oracle and the
 * device kernel of cap_build_lognormal -- produces the same bits.  Accuracy
 * is a few ulp, which is all a synthetic input needs.
 */
static inline double gen_bits(uint64_t b)
{
    double x;
    memcpy(&x, &b, sizeof x);
    return x;
}

static inline uint64_t gen_u64(double x)
{
    uint64_t b;
    memcpy(&b, &x, sizeof b);
    return b;
}

#define GEN_LN2_HI 6.93147180369123816490e-01   /* 0x3FE62E42FEE00000: e * hi exact */
#define GEN_LN2_LO 1.9082149292705877e-10

/* ln x for a normal x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716, series to s^21 */
double orc_gen_log(double x)
{
    const uint64_t b = gen_u64(x);
    int e = (int)((b >> 52) & 0x7ff) - 1022;
    double m = gen_bits((b & 0x000FFFFFFFFFFFFFull) | 0x3FE0000000000000ull);   /* [0.5, 1) */
    if (m < 0.70710678118654757) {
        m = m * 2.0;
        e = e - 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double t = 0.047619047619047616;                 /* 1/21 */
    t = 0.05263157894736842 + z * t;                 /* 1/19 */
    t = 0.058823529411764705 + z * t;                /* 1/17 */
    t = 0.06666666666666667 + z * t;                 /* 1/15 */
    t = 0.07692307692307693 + z * t;                 /* 1/13 */
    t = 0.09090909090909091 + z * t;                 /* 1/11 */
    t = 0.1111111111111111 + z * t;                  /* 1/9 */
    t = 0.14285714285714285 + z * t;                 /* 1/7 */
    t = 0.2 + z * t;                                 /* 1/5 */
    t = 0.3333333333333333 + z * t;                  /* 1/3 */
    const double lm = 2.0 * s + 2.0 * s * (z * t);
    const double de = (double)e;
    return de * GEN_LN2_HI + (de * GEN_LN2_LO + lm);
}

/* cos(2 pi u) for u in [0, 1): u - rint(u) exactly, folded to |v| <= 1/2,
 * quadrant q = rint(4 v), r = v - q / 4 in [-1/8, 1/8] exactly, theta =
 * 2 pi r, Taylor series of cos to theta^18 and of sin to theta^19 */
double orc_gen_cos2pi(double u)
{
    double v = u - rint(u);
    if (v < 0.0) v = -v;
    const double q = rint(4.0 * v);
    const double r = v - 0.25 * q;
    const double th = 6.283185307179586 * r;
    const double z = th * th;
    double c = -1.5619206968586225e-16;              /* -1/18! */
    c = 4.779477332387385e-14 + z * c;               /* 1/16! */
    c = -1.1470745597729725e-11 + z * c;             /* -1/14! */
    c = 2.08767569878681e-09 + z * c;                /* 1/12! */
    c = -2.755731922398589e-07 + z * c;              /* -1/10! */
    c = 2.48015873015873e-05 + z * c;                /* 1/8! */
    c = -0.001388888888888889 + z * c;               /* -1/6! */
    c = 0.041666666666666664 + z * c;                /* 1/4! */
    c = -0.5 + z * c;                                /* -1/2! */
    c = 1.0 + z * c;
    double sn = -8.22063524662433e-18;               /* -1/19! */
    sn = 2.8114572543455206e-15 + z * sn;            /* 1/17! */
    sn = -7.647163731819816e-13 + z * sn;            /* -1/15! */
    sn = 1.6059043836821613e-10 + z * sn;            /* 1/13! */
    sn = -2.505210838544172e-08 + z * sn;            /* -1/11! */
    sn = 2.7557319223985893e-06 + z * sn;            /* 1/9! */
    sn = -0.0001984126984126984 + z * sn;            /* -1/7! */
    sn = 0.008333333333333333 + z * sn;              /* 1/5! */
    sn = -0.16666666666666666 + z * sn;              /* -1/3! */
    sn = th + th * (z * sn);
    if (q == 0.0) return c;
    if (q == 1.0) return -sn;
    return -c;                                       /* q == 2 */
}

/* e^x for |x| < 700: k = rint(x / ln 2), r = x - k ln 2 (two-part ln 2),
 * |r| <= 0.347, Taylor series to r^13, times 2^k built from its bits */
double orc_gen_exp(double x)
{
    const double k = rint(x * 1.4426950408889634);
    const double r = (x - k * GEN_LN2_HI) - k * GEN_LN2_LO;
    double p = 1.6059043836821613e-10;               /* 1/13! */
    p = 2.08767569878681e-09 + r * p;                /* 1/12! */
    p = 2.505210838544172e-08 + r * p;               /* 1/11! */
    p = 2.755731922398589e-07 + r * p;               /* 1/10! */
    p = 2.7557319223985893e-06 + r * p;              /* 1/9! */
    p = 2.48015873015873e-05 + r * p;                /* 1/8! */
    p = 0.0001984126984126984 + r * p;               /* 1/7! */
    p = 0.001388888888888889 + r * p;                /* 1/6! */
    p = 0.008333333333333333 + r * p;                /* 1/5! */
    p = 0.041666666666666664 + r * p;                /* 1/4! */
    p = 0.16666666666666666 + r * p;                 /* 1/3! */
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    return p * gen_bits((uint64_t)((long long)k + 1023) << 52);
}

double orc_normal(uint64_t seed, uint64_t stream, uint64_t idx)
{
    /* Box-Muller on two counter-keyed uniforms, u1 in (0,1], u2 in [0,1) */
    const double u1 = (double)((orc_hash(seed, stream, 2 * idx) >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(orc_hash(seed, stream, 2 * idx + 1) >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * orc_gen_log(u1)) * orc_gen_cos2pi(u2);
}

/* ------------------------------------------------------------ builders -- */
static void zero_sys(size_t n, orc_sys *s)
{
    double *arrs[] = {s->ae, s->aw, s->an, s->as, s->at, s->ab, s->su};
    for (int a = 0; a < 7; ++a) memset(arrs[a], 0, n * sizeof(double));
    for (size_t c = 0; c < n; ++c) s->ap[c] = 1.0;
}

/*
 * Neumann p' operator with log-normal face coefficients on a uniform grid:
 * interior face (east/north/top of an interior cell whose neighbour is also
 * interior) a_f = h*exp(sigma*g_f), g_f ~ N(0,1) keyed by (seed, dir, ijk);
 * aw(E)=ae(P), as(N)=an(P), ab(T)=at(P); boundary faces 0; ap = sum a_nb;
 * su ~ N(0,1) projected to zero mean; ghost rows a*=0, ap=1, su=0; the
 * reference-pressure cell (3,3,3) pinned.
 */
void orc_build_lognormal(int ni, int nj, int nk, uint64_t seed, double sigma,
                         orc_sys *s)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const size_t n = (size_t)nij * nk;
    const double h = 1.0 / (double)(ni - 4);
    zero_sys(n, s);
    for (int k = 3; k <= nkmm; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t c = IX(i, j, k);
                if (i < nimm) {
                    double a = h * orc_gen_exp(sigma * orc_normal(seed, 1, c));
                    s->ae[c] = a;
                    s->aw[c + nj] = a;
                }
                if (j < njmm) {
                    double a = h * orc_gen_exp(sigma * orc_normal(seed, 2, c));
                    s->an[c] = a;
                    s->as[c + 1] = a;
                }
                if (k < nkmm) {
                    double a = h * orc_gen_exp(sigma * orc_normal(seed, 3, c));
                    s->at[c] = a;
                    s->ab[c + nij] = a;
                }
            }
    /* the mean of su: a sum per (k, i) row over j, per plane over the rows
     * in i order, then over the planes in k order (an order a device kernel
     * reproduces exactly) */
    double sum = 0.0;
    long cnt = 0;
    for (int k = 3; k <= nkmm; ++k) {
        double plane = 0.0;
        for (int i = 3; i <= nimm; ++i) {
            double row = 0.0;
            for (int j = 3; j <= njmm; ++j) {
                size_t c = IX(i, j, k);
                s->ap[c] = s->ae[c] + s->aw[c] + s->an[c] + s->as[c] + s->at[c] + s->ab[c];
                s->su[c] = orc_normal(seed, 4, c);
                row += s->su[c];
                ++cnt;
            }
            plane += row;
        }
        sum += plane;
    }
    double mean = sum / (double)cnt;
    for (int k = 3; k <= nkmm; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) s->su[IX(i, j, k)] -= mean;
    /* reference-pressure cell (3,3,3) pinned as PISO_multiple_correction.f90:
     * 117-129 does: ap = 1, a_nb = 0, su = p_ref (= 0, the starting value) */
    size_t r = IX(3, 3, 3);
    s->ae[r] = s->aw[r] = s->an[r] = s->as[r] = s->at[r] = s->ab[r] = 0.0;
    s->ap[r] = 1.0;
    s->su[r] = 0.0;
}

/*
 * General non-symmetric, diagonally dominant 7-diagonal system whose rows
 * couple to the ghost cells (Dirichlet data) and whose ghost planes k=2 and
 * k=nkm carry rows of their own (the k-range of iccg / cgstab_sip).  fi gets
 * random values everywhere (initial guess and ghost data).
 */
void orc_build_general(int ni, int nj, int nk, uint64_t seed, orc_sys *s,
                       double *fi)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    const size_t n = (size_t)nij * nk;
    zero_sys(n, s);
    for (size_t c = 0; c < n; ++c) fi[c] = 2.0 * orc_u01(seed, 20, c) - 1.0;
    for (int k = 2; k <= nkm; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t c = IX(i, j, k);
                s->ae[c] = 0.5 + orc_u01(seed, 11, c);
                s->aw[c] = 0.5 + orc_u01(seed, 12, c);
                s->an[c] = 0.5 + orc_u01(seed, 13, c);
                s->as[c] = 0.5 + orc_u01(seed, 14, c);
                s->at[c] = 0.5 + orc_u01(seed, 15, c);
                s->ab[c] = 0.5 + orc_u01(seed, 16, c);
                s->ap[c] = (s->ae[c] + s->aw[c] + s->an[c] + s->as[c] + s->at[c] + s->ab[c]) *
                           (1.0 + 0.05 * orc_u01(seed, 17, c));
                s->su[c] = 2.0 * orc_u01(seed, 18, c) - 1.0;
            }
}

/* ------------------------------------------------------- SIP building -- */
/* SIP (Stone) factorisation: sipsol.f90:28-50 == pcg-sip.f90:68-87 ==
 * cgstab_sip.f90:60-79; k range is the caller's. */
void orc_sip_factor(int ni, int nj, int nk, int k_lo, int k_hi, orc_sys *s,
                    double alfa)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2;
    (void)nk;
    double *bb = s->bb, *bw = s->bw, *bs = s->bs, *bp = s->bp, *bn = s->bn,
           *be = s->be, *bt = s->bt;
    for (int k = k_lo; k <= k_hi; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t c = IX(i, j, k);
                bb[c] = -s->ab[c] / (1. + alfa * (bn[c - nij] + be[c - nij]));
                bw[c] = -s->aw[c] / (1. + alfa * (bn[c - nj] + bt[c - nj]));
                bs[c] = -s->as[c] / (1. + alfa * (be[c - 1] + bt[c - 1]));
                double p1 = alfa * (bb[c] * bn[c - nij] + bw[c] * bn[c - nj]);
                double p2 = alfa * (bb[c] * be[c - nij] + bs[c] * be[c - 1]);
                double p3 = alfa * (bw[c] * bt[c - nj] + bs[c] * bt[c - 1]);
                bp[c] = 1. / (s->ap[c] + p1 + p2 + p3 - bb[c] * bt[c - nij] -
                              bw[c] * be[c - nj] - bs[c] * bn[c - 1] + SMALL);
                bn[c] = (-s->an[c] - p1) * bp[c];
                be[c] = (-s->ae[c] - p2) * bp[c];
                bt[c] = (-s->at[c] - p3) * bp[c];
            }
}

/* z = (LU)^-1 rhs: forward pcg-sip.f90:98-106, backward pcg-sip.f90:112-121.
 * z must be zero outside the range (as the reference's zeroed zk). */
void orc_sip_apply(int ni, int nj, int nk, int k_lo, int k_hi,
                   const orc_sys *s, const double *rhs, double *z)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2;
    (void)nk;
    for (int k = k_lo; k <= k_hi; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t c = IX(i, j, k);
                z[c] = (rhs[c] - s->bb[c] * z[c - nij] - s->bw[c] * z[c - nj] -
                        s->bs[c] * z[c - 1]) * s->bp[c];
            }
    for (int k = k_hi; k >= k_lo; --k)
        for (int i = nimm; i >= 3; --i)
            for (int j = njmm; j >= 3; --j) {
                size_t c = IX(i, j, k);
                z[c] = z[c] - s->bn[c] * z[c + 1] - s->be[c] * z[c + nj] -
                       s->bt[c] * z[c + nij];
            }
}

/* z = A p with the operand order of pcg-sip.f90:147-149 */
void orc_spmv(int ni, int nj, int nk, int k_lo, int k_hi, const orc_sys *s,
              const double *p, double *z)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2;
    (void)nk;
    for (int k = k_lo; k <= k_hi; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t c = IX(i, j, k);
                z[c] = s->ap[c] * p[c] - s->ae[c] * p[c + nj] - s->aw[c] * p[c - nj] -
                       s->an[c] * p[c + 1] - s->as[c] * p[c - 1] - s->at[c] * p[c + nij] -
                       s->ab[c] * p[c - nij];
            }
}

/* residual r = sum a_nb fi_nb + su - ap fi: pcg-sip.f90:53-55 (same operand
 * order in sipsol.f90:66-68, iccg.f90:57-59, cgstab_sip.f90:49-51) */
static inline double resid(const orc_sys *s, const double *fi, size_t c, int nj, int nij)
{
    return s->ae[c] * fi[c + nj] + s->aw[c] * fi[c - nj] + s->an[c] * fi[c + 1] +
           s->as[c] * fi[c - 1] + s->at[c] * fi[c + nij] + s->ab[c] * fi[c - nij] +
           s->su[c] - s->ap[c] * fi[c];
}

/* --------------------------------------------------------------- SIPSOL -- */
/* sipsol.f90:3-106 */
void orc_sipsol(int ni, int nj, int nk, double *fi, orc_sys *s, double alfa,
                double sor, int nsw, orc_report *rep)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    double *res = s->res;
    orc_sip_factor(ni, nj, nk, 3, nkmm, s, alfa);            /* 28-50 */
    double resor = 0.0, res1 = 0.0;
    int n, conv = 0;
    for (n = 1; n <= nsw; ++n) {
        res1 = 0.0;
        psum res1_ps = PS0;
        for (int k = 3; k <= nkmm; ++k)                       /* 60-75 */
            for (int i = 3; i <= nimm; ++i)
                for (int j = 3; j <= njmm; ++j) {
                    size_t c = IX(i, j, k);
                    res[c] = resid(s, fi, c, nj, nij);
                    res1 = ps_add(&res1_ps, fabs(res[c]), k);
                    res[c] = (res[c] - s->bb[c] * res[c - nij] - s->bw[c] * res[c - nj] -
                              s->bs[c] * res[c - 1]) * s->bp[c];
                }
        if (n == 1) resor = res1;                             /* 77 */
        double rsm = res1 / (resor + SMALL);                  /* 78 */
        for (int k = nkmm; k >= 3; --k)                       /* 81-92 */
            for (int i = nimm; i >= 3; --i)
                for (int j = njmm; j >= 3; --j) {
                    size_t c = IX(i, j, k);
                    res[c] = res[c] - s->bn[c] * res[c + 1] - s->be[c] * res[c + nj] -
                             s->bt[c] * res[c + nij];
                    fi[c] = fi[c] + res[c];
                }
        if (rsm < sor) { conv = 1; break; }                   /* 96 */
    }
    rep->resor = resor;
    rep->resl = res1;
    rep->iters = n;   /* N after EXIT, or NS+1 when the DO loop ran out */
    rep->converged = conv;
}

/* --------------------------------------------------------------- PCGSIP -- */
/* pcg-sip.f90:3-200 (global_sum is the identity on one domain) */
static void pcgsip_core(int ni, int nj, int nk, double *fi, orc_sys *s,
                        double alfa, double sor, int nsw, int fixed_iters,
                        orc_report *rep)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const size_t n = (size_t)nij * nk;
    double *pk = calloc(n, sizeof(double)), *zk = calloc(n, sizeof(double));
    double *res = s->res;
    memset(res, 0, n * sizeof(double));                        /* 42-44 */
    double res0 = 0.0;
    psum res0_ps = PS0;
    for (int k = 3; k <= nkmm; ++k)                            /* 49-59 */
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t c = IX(i, j, k);
                res[c] = resid(s, fi, c, nj, nij);
                res0 = ps_add(&res0_ps, fabs(res[c]), k);
            }
    orc_sip_factor(ni, nj, nk, 3, nkmm, s, alfa);              /* 68-87 */
    double s0 = S0INIT, resl = 0.0, resor = 0.0;
    int l, conv = 0;
    for (l = 1; l <= nsw; ++l) {
        for (int k = 3; k <= nkmm; ++k)                        /* 98-106 */
            for (int i = 3; i <= nimm; ++i)
                for (int j = 3; j <= njmm; ++j) {
                    size_t c = IX(i, j, k);
                    zk[c] = (res[c] - s->bb[c] * zk[c - nij] - s->bw[c] * zk[c - nj] -
                             s->bs[c] * zk[c - 1]) * s->bp[c];
                }
        double sk = 0.0;
        psum sk_ps = PS0;
        for (int k = nkmm; k >= 3; --k)                        /* 111-121 */
            for (int i = nimm; i >= 3; --i)
                for (int j = njmm; j >= 3; --j) {
                    size_t c = IX(i, j, k);
                    zk[c] = zk[c] - s->bn[c] * zk[c + 1] - s->be[c] * zk[c + nj] -
                            s->bt[c] * zk[c + nij];
                    sk = ps_add(&sk_ps, res[c] * zk[c], k);
                }
        double bet = sk / s0;                                  /* 127 */
        for (int k = 3; k <= nkmm; ++k)                        /* 131-138 */
            for (int i = 3; i <= nimm; ++i)
                for (int j = 3; j <= njmm; ++j) {
                    size_t c = IX(i, j, k);
                    pk[c] = zk[c] + bet * pk[c];
                }
        double pkapk = 0.0;
        psum pkapk_ps = PS0;
        for (int k = 3; k <= nkmm; ++k)                        /* 142-153 */
            for (int i = 3; i <= nimm; ++i)
                for (int j = 3; j <= njmm; ++j) {
                    size_t c = IX(i, j, k);
                    zk[c] = s->ap[c] * pk[c] - s->ae[c] * pk[c + nj] - s->aw[c] * pk[c - nj] -
                            s->an[c] * pk[c + 1] - s->as[c] * pk[c - 1] -
                            s->at[c] * pk[c + nij] - s->ab[c] * pk[c - nij];
                    pkapk = ps_add(&pkapk_ps, pk[c] * zk[c], k);
                }
        double alf = sk / pkapk;                               /* 157 */
        resl = 0.0;
        psum resl_ps = PS0;
        for (int k = 3; k <= nkmm; ++k)                        /* 161-175 */
            for (int i = 3; i <= nimm; ++i)
                for (int j = 3; j <= njmm; ++j) {
                    size_t c = IX(i, j, k);
                    fi[c] = fi[c] + alf * pk[c];
                    res[c] = res[c] - alf * zk[c];
                    resl = ps_add(&resl_ps, fabs(res[c]), k);
                }
        s0 = sk;                                               /* 181 */
        if (l == 1) resor = res0;                              /* 185 */
        double rsm = resl / (resor + SMALL);                   /* 186 */
        if (!fixed_iters && rsm < sor) { conv = 1; break; }    /* 188 */
    }
    rep->resor = res0;
    rep->resl = resl;
    rep->iters = l;
    rep->converged = conv;
    free(pk);
    free(zk);
}

void orc_pcgsip(int ni, int nj, int nk, double *fi, orc_sys *s, double alfa,
                double sor, int nsw, orc_report *rep)
{
    pcgsip_core(ni, nj, nk, fi, s, alfa, sor, nsw, 0, rep);
}

double orc_pcgsip_time_iters(int ni, int nj, int nk, double *fi, orc_sys *s,
                             double alfa, int n_iters)
{
    struct timespec t0, t1;
    orc_report rep;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    pcgsip_core(ni, nj, nk, fi, s, alfa, 0.0, n_iters, 1, &rep);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ----------------------------------------------------------------- ICCG -- */
/* iccg.f90:3-197; k range 2..nkm */
void orc_iccg(int ni, int nj, int nk, double *fi, orc_sys *s, double sor,
              int nsw, orc_report *rep)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    const size_t n = (size_t)nij * nk;
    double *pk = calloc(n, sizeof(double)), *zk = calloc(n, sizeof(double)),
           *d = calloc(n, sizeof(double));
    double *res = s->res;
    memset(res, 0, n * sizeof(double));                        /* 45-48 */
    double res0 = 0.0;
    psum res0_ps = PS0;
    for (int k = 2; k <= nkm; ++k)                             /* 52-63 */
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t c = IX(i, j, k);
                res[c] = resid(s, fi, c, nj, nij);
                res0 = ps_add(&res0_ps, fabs(res[c]), k);
            }
    for (int k = 2; k <= nkm; ++k)                             /* 73-81 */
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t c = IX(i, j, k);
                d[c] = 1. / (s->ap[c] - s->aw[c] * s->aw[c] * d[c - nj] -
                             s->as[c] * s->as[c] * d[c - 1] -
                             s->ab[c] * s->ab[c] * d[c - nij]);
            }
    double s0 = S0INIT, resl = 0.0, resor = 0.0;
    int l, conv = 0;
    for (l = 1; l <= nsw; ++l) {
        for (int k = 2; k <= nkm; ++k)                         /* 92-100 */
            for (int i = 3; i <= nimm; ++i)
                for (int j = 3; j <= njmm; ++j) {
                    size_t c = IX(i, j, k);
                    zk[c] = (res[c] + s->aw[c] * zk[c - nj] + s->as[c] * zk[c - 1] +
                             s->ab[c] * zk[c - nij]) * d[c];
                }
        for (int k = 2; k <= nkm; ++k)                         /* 102-109 */
            for (int i = 3; i <= nimm; ++i)
                for (int j = 3; j <= njmm; ++j) {
                    size_t c = IX(i, j, k);
                    zk[c] = zk[c] / (d[c] + SMALL);
                }
        double sk = 0.0;
        psum sk_ps = PS0;
        for (int k = nkm; k >= 2; --k)                         /* 113-123 */
            for (int i = nimm; i >= 3; --i)
                for (int j = njmm; j >= 3; --j) {
                    size_t c = IX(i, j, k);
                    zk[c] = (zk[c] + s->ae[c] * zk[c + nj] + s->an[c] * zk[c + 1] +
                             s->at[c] * zk[c + nij]) * d[c];
                    sk = ps_add(&sk_ps, res[c] * zk[c], k);
                }
        double bet = sk / s0;                                  /* 129 */
        for (int k = 2; k <= nkm; ++k)                         /* 133-140 */
            for (int i = 3; i <= nimm; ++i)
                for (int j = 3; j <= njmm; ++j) {
                    size_t c = IX(i, j, k);
                    pk[c] = zk[c] + bet * pk[c];
                }
        double pkapk = 0.0;
        psum pkapk_ps = PS0;
        for (int k = 2; k <= nkm; ++k)                         /* 144-155 */
            for (int i = 3; i <= nimm; ++i)
                for (int j = 3; j <= njmm; ++j) {
                    size_t c = IX(i, j, k);
                    zk[c] = s->ap[c] * pk[c] - s->ae[c] * pk[c + nj] - s->aw[c] * pk[c - nj] -
                            s->an[c] * pk[c + 1] - s->as[c] * pk[c - 1] -
                            s->at[c] * pk[c + nij] - s->ab[c] * pk[c - nij];
                    pkapk = ps_add(&pkapk_ps, pk[c] * zk[c], k);
                }
        double alf = sk / pkapk;                               /* 160 */
        resl = 0.0;
        psum resl_ps = PS0;
        for (int k = 2; k <= nkm; ++k)                         /* 164-174 */
            for (int i = 3; i <= nimm; ++i)
                for (int j = 3; j <= njmm; ++j) {
                    size_t c = IX(i, j, k);
                    fi[c] = fi[c] + alf * pk[c];
                    res[c] = res[c] - alf * zk[c];
                    resl = ps_add(&resl_ps, fabs(res[c]), k);
                }
        s0 = sk;                                               /* 178 */
        if (l == 1) resor = res0;                              /* 182 */
        double rsm = resl / (resor + SMALL);                   /* 183 */
        if (rsm < sor) { conv = 1; break; }                    /* 185 */
    }
    rep->resor = res0;
    rep->resl = resl;
    rep->iters = l;
    rep->converged = conv;
    free(pk);
    free(zk);
    free(d);
}

/* ----------------------------------------------------------- CGSTAB_SIP -- */
/* cgstab_sip.f90:1-278; k range 2..nkm; its local work arrays are zero
 * outside the solve range (they are uninitialised in the reference). */
void orc_cgstab_sip(int ni, int nj, int nk, double *fi, orc_sys *s,
                    double alfa, double sor, int nsw, orc_report *rep)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    const size_t n = (size_t)nij * nk;
    double *reso = calloc(n, sizeof(double)), *pk = calloc(n, sizeof(double)),
           *uk = calloc(n, sizeof(double)), *zk = calloc(n, sizeof(double)),
           *vk = calloc(n, sizeof(double));
    double *res = s->res;
#define LOOP_F for (int k = 2; k <= nkm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j)
#define LOOP_B for (int k = nkm; k >= 2; --k) for (int i = nimm; i >= 3; --i) for (int j = njmm; j >= 3; --j)
    double res0 = 0.0;
    psum res0_ps = PS0;
    LOOP_F {                                                   /* 44-55 */
        size_t c = IX(i, j, k);
        res[c] = resid(s, fi, c, nj, nij);
        res0 = ps_add(&res0_ps, fabs(res[c]), k);
    }
    orc_sip_factor(ni, nj, nk, 2, nkm, s, alfa);               /* 60-79 */
    LOOP_F {                                                   /* 83-94 */
        size_t c = IX(i, j, k);
        reso[c] = res[c];
        pk[c] = 0.0; uk[c] = 0.0; zk[c] = 0.0; vk[c] = 0.0;
    }
    double alf = 1.0, beto = 1.0, gam = 1.0, resl = 0.0, resor = 0.0;
    int l, conv = 0;
    for (l = 1; l <= nsw; ++l) {
        double bet = 0.0;
        psum bet_ps = PS0;                                      /* 106-116 */
        LOOP_F { size_t c = IX(i, j, k); bet = ps_add(&bet_ps, res[c] * reso[c], k); }
        double om = bet * gam / (alf * beto + SMALL);
        beto = bet;
        LOOP_F {                                               /* 120-127 */
            size_t c = IX(i, j, k);
            pk[c] = res[c] + om * (pk[c] - alf * uk[c]);
        }
        LOOP_F {                                               /* 131-139 */
            size_t c = IX(i, j, k);
            zk[c] = (pk[c] - s->bb[c] * zk[c - nij] - s->bw[c] * zk[c - nj] -
                     s->bs[c] * zk[c - 1]) * s->bp[c];
        }
        LOOP_B {                                               /* 144-152 */
            size_t c = IX(i, j, k);
            zk[c] = zk[c] - s->bn[c] * zk[c + 1] - s->be[c] * zk[c + nj] - s->bt[c] * zk[c + nij];
        }
        LOOP_F {                                               /* 156-166 */
            size_t c = IX(i, j, k);
            uk[c] = s->ap[c] * zk[c] - s->ae[c] * zk[c + nj] - s->aw[c] * zk[c - nj] -
                    s->an[c] * zk[c + 1] - s->as[c] * zk[c - 1] - s->at[c] * zk[c + nij] -
                    s->ab[c] * zk[c - nij];
        }
        double ukreso = 0.0;
        psum ukreso_ps = PS0;                                   /* 170-179 */
        LOOP_F { size_t c = IX(i, j, k); ukreso = ps_add(&ukreso_ps, uk[c] * reso[c], k); }
        gam = bet / ukreso;
        LOOP_F {                                               /* 183-191 */
            size_t c = IX(i, j, k);
            fi[c] = fi[c] + gam * zk[c];
            res[c] = res[c] - gam * uk[c];
        }
        LOOP_F {                                               /* 195-203 */
            size_t c = IX(i, j, k);
            zk[c] = (res[c] - s->bb[c] * zk[c - nij] - s->bw[c] * zk[c - nj] -
                     s->bs[c] * zk[c - 1]) * s->bp[c];
        }
        LOOP_B {                                               /* 207-215 */
            size_t c = IX(i, j, k);
            zk[c] = zk[c] - s->bn[c] * zk[c + 1] - s->be[c] * zk[c + nj] - s->bt[c] * zk[c + nij];
        }
        LOOP_F {                                               /* 219-229 */
            size_t c = IX(i, j, k);
            vk[c] = s->ap[c] * zk[c] - s->ae[c] * zk[c + nj] - s->aw[c] * zk[c - nj] -
                    s->an[c] * zk[c + 1] - s->as[c] * zk[c - 1] - s->at[c] * zk[c + nij] -
                    s->ab[c] * zk[c - nij];
        }
        double vres = 0.0, vv = 0.0;
        psum vres_ps = PS0, vv_ps = PS0;                           /* 233-245 */
        LOOP_F {
            size_t c = IX(i, j, k);
            vres = ps_add(&vres_ps, vk[c] * res[c], k);
            vv = ps_add(&vv_ps, vk[c] * vk[c], k);
        }
        alf = vres / (vv + SMALL);
        resl = 0.0;
        psum resl_ps = PS0;                                            /* 249-259 */
        LOOP_F {
            size_t c = IX(i, j, k);
            fi[c] = fi[c] + alf * zk[c];
            res[c] = res[c] - alf * vk[c];
            resl = ps_add(&resl_ps, fabs(res[c]), k);
        }
        if (l == 1) resor = res0;                              /* 263 */
        double rsm = resl / (resor + SMALL);
        if (rsm < sor) { conv = 1; break; }                    /* 266 */
    }
#undef LOOP_F
#undef LOOP_B
    rep->resor = res0;
    rep->resl = resl;
    rep->iters = l;
    rep->converged = conv;
    free(reso); free(pk); free(uk); free(zk); free(vk);
}

/* ----------------------------------------------------------- stencils -- */
/* GRAD_GAUSS.f90:1-105: zero, face scatter over k=2..nkm, i=2..nimm,
 * j=2..njmm, then scale by volr=1/vol on k=2..nkm, i,j=3..n?mm. */
void orc_grad_gauss(int ni, int nj, int nk, const double *u, const double *fx,
                    const double *fy, const double *fz, const double *ar1x,
                    const double *ar1y, const double *ar1z, const double *ar2x,
                    const double *ar2y, const double *ar2z, const double *ar3x,
                    const double *ar3y, const double *ar3z, const double *vol,
                    double *dudx, double *dudy, double *dudz)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    const size_t n = (size_t)nij * nk;
    memset(dudx, 0, n * sizeof(double));
    memset(dudy, 0, n * sizeof(double));
    memset(dudz, 0, n * sizeof(double));
    for (int k = 2; k <= nkm; ++k)
        for (int i = 2; i <= nimm; ++i)
            for (int j = 2; j <= njmm; ++j) {
                size_t p = IX(i, j, k), e = p + nj, nn = p + 1, t = p + nij;
                double ue = u[e] * fx[p] + u[p] * (1.0 - fx[p]);
                double un = u[nn] * fy[p] + u[p] * (1.0 - fy[p]);
                double ut = u[t] * fz[p] + u[p] * (1.0 - fz[p]);
                dudx[p] = dudx[p] + (ue * ar1x[p] + un * ar2x[p] + ut * ar3x[p]);
                dudx[e] = dudx[e] - (ue * ar1x[p]);
                dudx[nn] = dudx[nn] - (un * ar2x[p]);
                dudx[t] = dudx[t] - (ut * ar3x[p]);
                dudy[p] = dudy[p] + (ue * ar1y[p] + un * ar2y[p] + ut * ar3y[p]);
                dudy[e] = dudy[e] - (ue * ar1y[p]);
                dudy[nn] = dudy[nn] - (un * ar2y[p]);
                dudy[t] = dudy[t] - (ut * ar3y[p]);
                dudz[p] = dudz[p] + (ue * ar1z[p] + un * ar2z[p] + ut * ar3z[p]);
                dudz[e] = dudz[e] - (ue * ar1z[p]);
                dudz[nn] = dudz[nn] - (un * ar2z[p]);
                dudz[t] = dudz[t] - (ut * ar3z[p]);
            }
    for (int k = 2; k <= nkm; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t p = IX(i, j, k);
                double volr = 1.0 / vol[p];
                dudx[p] = dudx[p] * volr;
                dudy[p] = dudy[p] * volr;
                dudz[p] = dudz[p] * volr;
            }
}

/* bpres.f90:29-79: linear extrapolation of p into the six ghost planes */
void orc_bpres(int ni, int nj, int nk, double *p, const double *fx,
               const double *fy, const double *fz)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    (void)nk;
    for (int i = 3; i <= nimm; ++i)
        for (int j = 3; j <= njmm; ++j) {
            size_t c = IX(i, j, 3);
            p[c - nij] = p[c] - (p[c + nij] - p[c]) * fz[c];
        }
    for (int i = 3; i <= nimm; ++i)
        for (int j = 3; j <= njmm; ++j) {
            size_t c = IX(i, j, nkmm);
            p[c + nij] = p[c] + (p[c] - p[c - nij]) * (1.0 - fz[c - nij]);
        }
    for (int j = 3; j <= njmm; ++j)
        for (int k = 3; k <= nkmm; ++k) {
            size_t c = IX(3, j, k);
            p[c - nj] = p[c] - (p[c + nj] - p[c]) * fx[c];
        }
    for (int j = 3; j <= njmm; ++j)
        for (int k = 3; k <= nkmm; ++k) {
            size_t c = IX(nimm, j, k);
            p[c + nj] = p[c] + (p[c] - p[c - nj]) * (1.0 - fx[c - nj]);
        }
    for (int i = 3; i <= nimm; ++i)
        for (int k = 3; k <= nkmm; ++k) {
            size_t c = IX(i, 3, k);
            p[c - 1] = p[c] - (p[c + 1] - p[c]) * fy[c];
        }
    for (int i = 3; i <= nimm; ++i)
        for (int k = 3; k <= nkmm; ++k) {
            size_t c = IX(i, njmm, k);
            p[c + 1] = p[c] + (p[c] - p[c - 1]) * (1.0 - fy[c - 1]);
        }
}

/* ------------------------------------------------------ p' assembly -- */
#define G3(g, c, q) ((g)[3 * (c) + (q)])

/* fluxmass2.f90:41-192 -- face loop k=2..nkm, i=2..nimm, j=2..njmm */
void orc_fluxmass(int ni, int nj, int nk, const orc_fields *f, int idew, int idns, int idtb,
                  const double *arvx, const double *arvy, const double *arvz,
                  const double *fif, double *acfe, double *acfw, double *fcf)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    for (int k = 2; k <= nkm; ++k)
        for (int i = 2; i <= nimm; ++i)
            for (int j = 2; j <= njmm; ++j) {
                size_t p = IX(i, j, k), e = p + idew, s = p - idns, b = p - idtb, bs = b - idns;
                double fxe = fif[p], fxp = 1.0 - fxe;
                double xf = 0.25 * (f->x[p] + f->x[s] + f->x[b] + f->x[bs]);
                double yf = 0.25 * (f->y[p] + f->y[s] + f->y[b] + f->y[bs]);
                double zf = 0.25 * (f->z[p] + f->z[s] + f->z[b] + f->z[bs]);
                double xi = f->xc[p] * fxp + f->xc[e] * fxe;
                double yi = f->yc[p] * fxp + f->yc[e] * fxe;
                double zi = f->zc[p] * fxp + f->zc[e] * fxe;
                double xpn = f->xc[e] - f->xc[p], ypn = f->yc[e] - f->yc[p], zpn = f->zc[e] - f->zc[p];
                double arx = arvx[p], ary = arvy[p], arz = arvz[p];
                double are = sqrt(arx * arx + ary * ary + arz * arz);
                double nxx = arx / are, nyy = ary / are, nzz = arz / are;
                double sfdpnr = 1.0 / (arx * xpn * nxx + ary * ypn * nyy + arz * zpn * nzz);
                double dene = f->den[p] * fxp + f->den[e] * fxe;
                double smdpn = (arx * arx + ary * ary + arz * arz) * sfdpnr;
                acfe[p] = dene * (fxp * f->vol[p] * f->apu[p] + fxe * f->vol[e] * f->apu[e]) * smdpn;
                acfw[e] = acfe[p];
                double duxi = G3(f->gradu, p, 0) * fxp + G3(f->gradu, e, 0) * fxe;
                double duyi = G3(f->gradu, p, 1) * fxp + G3(f->gradu, e, 1) * fxe;
                double duzi = G3(f->gradu, p, 2) * fxp + G3(f->gradu, e, 2) * fxe;
                double ui = f->u[p] * fxp + f->u[e] * fxe + (duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi));
                duxi = G3(f->gradv, p, 0) * fxp + G3(f->gradv, e, 0) * fxe;
                duyi = G3(f->gradv, p, 1) * fxp + G3(f->gradv, e, 1) * fxe;
                duzi = G3(f->gradv, p, 2) * fxp + G3(f->gradv, e, 2) * fxe;
                double vi = f->v[p] * fxp + f->v[e] * fxe + (duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi));
                duxi = G3(f->gradw, p, 0) * fxp + G3(f->gradw, e, 0) * fxe;
                duyi = G3(f->gradw, p, 1) * fxp + G3(f->gradw, e, 1) * fxe;
                duzi = G3(f->gradw, p, 2) * fxp + G3(f->gradw, e, 2) * fxe;
                double wi = f->w[p] * fxp + f->w[e] * fxe + (duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi));
                double dpxi = (fxe * f->vol[e] * f->apu[e] + fxp * f->vol[p] * f->apu[p]) *
                              (fxe * G3(f->gradp, e, 0) + fxp * G3(f->gradp, p, 0)) * xpn * nxx;
                double dpyi = (fxe * f->vol[e] * f->apv[e] + fxp * f->vol[p] * f->apv[p]) *
                              (fxe * G3(f->gradp, e, 1) + fxp * G3(f->gradp, p, 1)) * ypn * nyy;
                double dpzi = (fxe * f->vol[e] * f->apw[e] + fxp * f->vol[p] * f->apw[p]) *
                              (fxe * G3(f->gradp, e, 2) + fxp * G3(f->gradp, p, 2)) * zpn * nzz;
                double xpp = xf - (xf - f->xc[p]) * nxx, ypp = yf - (yf - f->yc[p]) * nyy,
                       zpp = zf - (zf - f->zc[p]) * nzz;
                double xep = xf - (xf - f->xc[e]) * nxx, yep = yf - (yf - f->yc[e]) * nyy,
                       zep = zf - (zf - f->zc[e]) * nzz;
                xpp = xpp - f->xc[p]; ypp = ypp - f->yc[p]; zpp = zpp - f->zc[p];
                xep = xep - f->xc[e]; yep = yep - f->yc[e]; zep = zep - f->zc[e];
                double dpe = (f->p[e] - f->p[p]) +
                             (G3(f->gradp, e, 0) * xep + G3(f->gradp, e, 1) * yep + G3(f->gradp, e, 2) * zep -
                              G3(f->gradp, p, 0) * xpp + G3(f->gradp, p, 1) * ypp + G3(f->gradp, p, 2) * zpp);
                double dpex = (fxe * f->vol[e] * f->apu[e] + fxp * f->vol[p] * f->apu[p]) * (arx * sfdpnr) * dpe;
                double dpey = (fxe * f->vol[e] * f->apv[e] + fxp * f->vol[p] * f->apv[p]) * (ary * sfdpnr) * dpe;
                double dpez = (fxe * f->vol[e] * f->apw[e] + fxp * f->vol[p] * f->apw[p]) * (arz * sfdpnr) * dpe;
                double ue = ui - dpex + dpxi, ve = vi - dpey + dpyi, we = wi - dpez + dpzi;
                fcf[p] = dene * (ue * arx + ve * ary + we * arz);
            }
}

/* PISO_assemble_pressure_eq.f90:44-257 (bp is the scratch array the caller
 * passes, zeroed first; su is module `coef` su) */
void orc_piso_assemble(int ni, int nj, int nk, const orc_fields *f, int idew, int idns,
                       int idtb, const double *arvx, const double *arvy, const double *arvz,
                       const double *fif, double *acfe, double *acfw, double *fcf, int ipc,
                       double *bp, double *su)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    memset(bp, 0, (size_t)nij * nk * sizeof(double));
    for (int k = 2; k <= nkmm; ++k)
        for (int i = 2; i <= nimm; ++i)
            for (int j = 2; j <= njmm; ++j) {
                size_t p = IX(i, j, k), e = p + idew, s = p - idns, b = p - idtb, bs = b - idns;
                double fxe = fif[p], fxp = 1.0 - fxe;
                double xf = 0.25 * (f->x[p] + f->x[s] + f->x[b] + f->x[bs]);
                double yf = 0.25 * (f->y[p] + f->y[s] + f->y[b] + f->y[bs]);
                double zf = 0.25 * (f->z[p] + f->z[s] + f->z[b] + f->z[bs]);
                double xi = f->xc[p] * fxp + f->xc[e] * fxe;
                double yi = f->yc[p] * fxp + f->yc[e] * fxe;
                double zi = f->zc[p] * fxp + f->zc[e] * fxe;
                double xpn = f->xc[e] - f->xc[p], ypn = f->yc[e] - f->yc[p], zpn = f->zc[e] - f->zc[p];
                double dpn = sqrt(xpn * xpn + ypn * ypn + zpn * zpn);
                double ixi1 = xpn / dpn, ixi2 = ypn / dpn, ixi3 = zpn / dpn;
                double arx = arvx[p], ary = arvy[p], arz = arvz[p];
                double are = sqrt(arx * arx + ary * ary + arz * arz);
                double vole = xpn * arx + ypn * ary + zpn * arz;
                double nxx = arx / are, nyy = ary / are, nzz = arz / are;
                double costn = nxx * ixi1 + nyy * ixi2 + nzz * ixi3;
                double d1x = costn, d1y = costn, d1z = costn;
                double d2x = xpn * costn, d2y = ypn * costn, d2z = zpn * costn;
                double dene = f->den[p] * fxp + f->den[e] * fxe;
                double game = dene * (fxp * f->vol[p] * f->apu[p] + fxe * f->vol[e] * f->apu[e]);
                acfe[p] = game * are / dpn;
                acfw[e] = acfe[p];
                if (ipc > 0) {
                    double dfixi = G3(f->gradp, p, 0) * fxp + G3(f->gradp, e, 0) * fxe;
                    double dfiyi = G3(f->gradp, p, 1) * fxp + G3(f->gradp, e, 1) * fxe;
                    double dfizi = G3(f->gradp, p, 2) * fxp + G3(f->gradp, e, 2) * fxe;
                    double dfixii = dfixi * d1x + arx / vole * (f->p[e] - f->p[p] - (dfixi * d2x + dfiyi * d2y + dfizi * d2z));
                    double dfiyii = dfiyi * d1y + ary / vole * (f->p[e] - f->p[p] - (dfixi * d2x + dfiyi * d2y + dfizi * d2z));
                    double dfizii = dfizi * d1z + arz / vole * (f->p[e] - f->p[p] - (dfixi * d2x + dfiyi * d2y + dfizi * d2z));
                    double fdfie = game * (dfixii * arx + dfiyii * ary + dfizii * arz);
                    double fdfii = game * are / dpn * (dfixi * xpn + dfiyi * ypn + dfizi * zpn);
                    double suadd = fdfie - fdfii;
                    su[p] = su[p] + suadd;
                    bp[p] = suadd;
                    double duxi = G3(f->gradu, p, 0) * fxp + G3(f->gradu, e, 0) * fxe;
                    double duyi = G3(f->gradu, p, 1) * fxp + G3(f->gradu, e, 1) * fxe;
                    double duzi = G3(f->gradu, p, 2) * fxp + G3(f->gradu, e, 2) * fxe;
                    double ui = f->u[p] * fxp + f->u[e] * fxe + duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi);
                    duxi = G3(f->gradv, p, 0) * fxp + G3(f->gradv, e, 0) * fxe;
                    duyi = G3(f->gradv, p, 1) * fxp + G3(f->gradv, e, 1) * fxe;
                    duzi = G3(f->gradv, p, 2) * fxp + G3(f->gradv, e, 2) * fxe;
                    double vi = f->v[p] * fxp + f->v[e] * fxe + duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi);
                    duxi = G3(f->gradw, p, 0) * fxp + G3(f->gradw, e, 0) * fxe;
                    duyi = G3(f->gradw, p, 1) * fxp + G3(f->gradw, e, 1) * fxe;
                    duzi = G3(f->gradw, p, 2) * fxp + G3(f->gradw, e, 2) * fxe;
                    double wi = f->w[p] * fxp + f->w[e] * fxe + duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi);
                    fcf[p] = dene * (ui * arx + vi * ary + wi * arz);
                }
            }
    if (ipc > 0) {
        /* PISO_assemble_pressure_eq.f90:241-253: su(c) -= bp(c) at c = cell + idew */
        for (int k = 3; k <= nkmm; ++k)
            for (int i = 3; i <= nimm; ++i)
                for (int j = 3; j <= njmm; ++j) {
                    size_t c = IX(i, j, k) + idew;
                    su[c] = su[c] - bp[c];
                }
    }
}

/* assemble_pressure_eq.f90:38-118 (projection form: no density or apu in the
 * coefficient; face velocity from su/sv/sw = HbyA, gradients unused) */
void orc_assemble_p(int ni, int nj, int nk, const orc_fields *f, int idew, int idns, int idtb,
                    const double *arvx, const double *arvy, const double *arvz,
                    const double *fif, double *acfe, double *acfw, double *fcf,
                    const double *su, const double *sv, const double *sw)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    for (int k = 2; k <= nkmm; ++k)
        for (int i = 2; i <= nimm; ++i)
            for (int j = 2; j <= njmm; ++j) {
                size_t p = IX(i, j, k), e = p + idew;
                double fxe = fif[p], fxp = 1.0 - fxe;
                double xpn = f->xc[e] - f->xc[p], ypn = f->yc[e] - f->yc[p], zpn = f->zc[e] - f->zc[p];
                double arx = arvx[p], ary = arvy[p], arz = arvz[p];
                double are = sqrt(arx * arx + ary * ary + arz * arz);
                double nxx = arx / are, nyy = ary / are, nzz = arz / are;
                double smdpn = (arx * arx + ary * ary + arz * arz) / (arx * xpn * nxx + ary * ypn * nyy + arz * zpn * nzz);
                acfe[p] = (fxp * f->vol[p] + fxe * f->vol[e]) * smdpn;
                acfw[e] = acfe[p];
                double ui = su[p] * fxp + su[e] * fxe;
                double vi = sv[p] * fxp + sv[e] * fxe;
                double wi = sw[p] * fxp + sw[e] * fxe;
                fcf[p] = ui * arx + vi * ary + wi * arz;
            }
    (void)idns; (void)idtb;
}

/* fluxmc.f90:41-106 for every cell of the corrector loop range of
 * calcp-multiple_correction_SIMPLE.f90:120-162 (k,i,j = 3..n?mm) */
void orc_fluxmc_all(int ni, int nj, int nk, const orc_fields *f, int idew, int idns, int idtb,
                    const double *fif, double *fmcor)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    memset(fmcor, 0, (size_t)nij * nk * sizeof(double));
    for (int k = 3; k <= nkmm; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t p = IX(i, j, k), e = p + idew, s = p - idns, b = p - idtb, bs = b - idns;
                double fxe = fif[p], fxp = 1. - fxe;
                double arx, ary, arz;
                if (idew == nj) { arx = f->ar1x[p]; ary = f->ar1y[p]; arz = f->ar1z[p]; }
                else if (idew == 1) { arx = f->ar2x[p]; ary = f->ar2y[p]; arz = f->ar2z[p]; }
                else { arx = f->ar3x[p]; ary = f->ar3y[p]; arz = f->ar3z[p]; }
                double xpn = f->xc[e] - f->xc[p], ypn = f->yc[e] - f->yc[p], zpn = f->zc[e] - f->zc[p];
                double are = sqrt(arx * arx + ary * ary + arz * arz);
                double nxx = arx / are, nyy = ary / are, nzz = arz / are;
                double dppnnr = 1. / ((xpn * nxx) + (ypn * nyy) + (zpn * nzz));
                double xf = 0.25 * (f->x[p] + f->x[s] + f->x[b] + f->x[bs]);
                double yf = 0.25 * (f->y[p] + f->y[s] + f->y[b] + f->y[bs]);
                double zf = 0.25 * (f->z[p] + f->z[s] + f->z[b] + f->z[bs]);
                double xpp = xf - (xf - f->xc[p]) * nxx, ypp = yf - (yf - f->yc[p]) * nyy, zpp = zf - (zf - f->zc[p]) * nzz;
                double xep = xf - (xf - f->xc[e]) * nxx, yep = yf - (yf - f->yc[e]) * nyy, zep = zf - (zf - f->zc[e]) * nzz;
                xpp = xpp - f->xc[p]; ypp = ypp - f->yc[p]; zpp = zpp - f->zc[p];
                xep = xep - f->xc[e]; yep = yep - f->yc[e]; zep = zep - f->zc[e];
                double rapr = (f->apu[p] * f->den[p] * f->vol[p] * fxp + f->apu[e] * f->den[e] * f->vol[e] * fxe);
                fmcor[p] = rapr * are * ((G3(f->gradp, e, 0) * xep - G3(f->gradp, p, 0) * xpp) +
                                         (G3(f->gradp, e, 1) * yep - G3(f->gradp, p, 1) * ypp) +
                                         (G3(f->gradp, e, 2) * zep - G3(f->gradp, p, 2) * zpp)) * dppnnr;
            }
}

/* continuityErrors.h:63-81: sums over k=2..nkm, i,j=3..n?mm of
 * LocalContErr = -f1+f1(w) - f2+f2(s) - f3+f3(b); sums = {global, sum|local|} */
void orc_continuity(int ni, int nj, int nk, const double *f1, const double *f2,
                    const double *f3, double *sums)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    double g = 0.0, sl = 0.0;
    for (int k = 2; k <= nkm; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                size_t c = IX(i, j, k);
                double loc = -f1[c] + f1[c - nj] - f2[c] + f2[c - 1] - f3[c] + f3[c - nij];
                g = g + loc;
                sl = sl + fabs(loc);
            }
    sums[0] = g;
    sums[1] = sl;
}

/* fluxuvw-overrelaxed-correction.f90:3-659: momentum face fluxes of one
 * direction.  Face loop over the direction's range (99-118), UDS implicit
 * coefficients acfe(P)/acfw(E) (276-288), explicit minus implicit diffusion
 * with the orthogonal correction costn = 1 (178-269), deferred correction
 * gam*(F_uds - F_high) with the schemes selected by `flags` (297-619); the P
 * side goes into su/sv/sw (625-627), the E side through the scratch
 * bp/bt/bb (628-630) added in a second loop (632-644), after which the
 * scratch is zeroed (652-656).  As in the reference, fuhigh/fvhigh/fwhigh
 * keep their value from the previous face when no scheme sets them (QUICK
 * skips faces next to the first cell of the range). */
void orc_fluxuvw(int ni, int nj, int nk, const orc_fields *f, const double *vis, int idew,
                 int idns, int idtb, const double *arvx, const double *arvy, const double *arvz,
                 const double *fif, double *acfe, double *acfw, const double *fcf, double *su,
                 double *sv, double *sw, double *bp, double *bt, double *bb, int flags,
                 double gds)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const int lcds = flags & ORC_UVW_CDS, lcds4 = flags & ORC_UVW_CDS4, lluds = flags & ORC_UVW_LUDS;
    const int lquds = flags & ORC_UVW_QUICK, flim = flags & ORC_UVW_LIMITER;
    const double gam = gds;
    double psie1 = 0.0, psie2 = 0.0, psie3 = 0.0, psiw1 = 0.0, psiw2 = 0.0, psiw3 = 0.0;
    double fuhigh = 0.0, fvhigh = 0.0, fwhigh = 0.0;
    int dir = 0;
    if (idew == nj) dir = 1;
    if (idew == 1) dir = 2;
    if (idew == nij) dir = 3;
    int kst = 2, kend = nkmm, ist = 2, iend = nimm, jst = 2, jend = njmm;
    if (dir == 1) { jst = 3; kst = 3; }
    else if (dir == 2) { ist = 3; kst = 3; }
    else if (dir == 3) { ist = 3; jst = 3; }
    const double *gu = f->gradu, *gv = f->gradv, *gw = f->gradw;
    const double *u = f->u, *v = f->v, *w = f->w;
    for (int k = kst; k <= kend; ++k)
        for (int i = ist; i <= iend; ++i)
            for (int j = jst; j <= jend; ++j) {
                const size_t inp = IX(i, j, k);
                const size_t ine = inp + idew, inw = inp - idew, ins = inp - idns;
                const size_t inb = inp - idtb, inee = inp + 2 * (size_t)idew, inbs = inb - idns;
                const double fxe = fif[inp], fxp = 1.0 - fxe;
                const double fxpw = fif[inw], fxw = 1.0 - fxpw;
                const double fxee = fif[ine], fxew = 1.0 - fxee;
                const double xpn = f->xc[ine] - f->xc[inp], ypn = f->yc[ine] - f->yc[inp],
                             zpn = f->zc[ine] - f->zc[inp];
                const double dpn = sqrt(xpn * xpn + ypn * ypn + zpn * zpn);
                const double arx = arvx[inp], ary = arvy[inp], arz = arvz[inp];
                const double are = sqrt(arx * arx + ary * ary + arz * arz);
                const double costn = 1.0;
                const double vole = xpn * arx + ypn * ary + zpn * arz;
                const double d1x = costn, d1y = costn, d1z = costn;
                const double d2x = xpn * costn, d2y = ypn * costn, d2z = zpn * costn;
                const double game = vis[inp] * fxp + vis[ine] * fxe;
                const double xf = 0.25 * (f->x[inp] + f->x[ins] + f->x[inb] + f->x[inbs]);
                const double yf = 0.25 * (f->y[inp] + f->y[ins] + f->y[inb] + f->y[inbs]);
                const double zf = 0.25 * (f->z[inp] + f->z[ins] + f->z[inb] + f->z[inbs]);
                const double xi = f->xc[inp] * fxp + f->xc[ine] * fxe;
                const double yi = f->yc[inp] * fxp + f->yc[ine] * fxe;
                const double zi = f->zc[inp] * fxp + f->zc[ine] * fxe;
                const double duxi = G3(gu, inp, 0) * fxp + G3(gu, ine, 0) * fxe;
                const double duyi = G3(gu, inp, 1) * fxp + G3(gu, ine, 1) * fxe;
                const double duzi = G3(gu, inp, 2) * fxp + G3(gu, ine, 2) * fxe;
                const double du = u[ine] - u[inp] - duxi * d2x - duyi * d2y - duzi * d2z;
                const double duxii = duxi * d1x + arx / vole * du;
                const double duyii = duyi * d1y + ary / vole * du;
                const double duzii = duzi * d1z + arz / vole * du;
                const double dvxi = G3(gv, inp, 0) * fxp + G3(gv, ine, 0) * fxe;
                const double dvyi = G3(gv, inp, 1) * fxp + G3(gv, ine, 1) * fxe;
                const double dvzi = G3(gv, inp, 2) * fxp + G3(gv, ine, 2) * fxe;
                const double dv = v[ine] - v[inp] - dvxi * d2x - dvyi * d2y - dvzi * d2z;
                const double dvxii = dvxi * d1x + arx / vole * dv;
                const double dvyii = dvyi * d1y + ary / vole * dv;
                const double dvzii = dvzi * d1z + arz / vole * dv;
                const double dwxi = G3(gw, inp, 0) * fxp + G3(gw, ine, 0) * fxe;
                const double dwyi = G3(gw, inp, 1) * fxp + G3(gw, ine, 1) * fxe;
                const double dwzi = G3(gw, inp, 2) * fxp + G3(gw, ine, 2) * fxe;
                const double dw = w[ine] - w[inp] - dwxi * d2x - dwyi * d2y - dwzi * d2z;
                const double dwxii = dwxi * d1x + arx / vole * dw;
                const double dwyii = dwyi * d1y + ary / vole * dw;
                const double dwzii = dwzi * d1z + arz / vole * dw;
                const double fdue = game * ((duxii + duxii) * arx + (duyii + dvxii) * ary + (duzii + dwxii) * arz);
                const double fdve = game * ((duyii + dvxii) * arx + (dvyii + dvyii) * ary + (dvzii + dwyii) * arz);
                const double fdwe = game * ((duzii + dwxii) * arx + (dwyii + dvzii) * ary + (dwzii + dwzii) * arz);
                const double fdui = game * are / dpn * (duxi * xpn + duyi * ypn + duzi * zpn);
                const double fdvi = game * are / dpn * (dvxi * xpn + dvyi * ypn + dvzi * zpn);
                const double fdwi = game * are / dpn * (dwxi * xpn + dwyi * ypn + dwzi * zpn);
                const double de = game * are / dpn;
                const double flcf = fcf[inp];
                const double ce = flcf < 0.0 ? flcf : 0.0;
                const double cp = flcf > 0.0 ? flcf : 0.0;
                acfe[inp] = de - ce;
                acfw[ine] = de + cp;
                const double fuuds = cp * u[inp] + ce * u[ine];
                const double fvuds = cp * v[inp] + ce * v[ine];
                const double fwuds = cp * w[inp] + ce * w[ine];
                if (lcds) {
                    const double ue = u[inp] * fxp + u[ine] * fxe + (duxi * (xf - xi) + duyi * (yf - yi) + duzi * (zf - zi));
                    const double ve = v[inp] * fxp + v[ine] * fxe + (dvxi * (xf - xi) + dvyi * (yf - yi) + dvzi * (zf - zi));
                    const double we = w[inp] * fxp + w[ine] * fxe + (dwxi * (xf - xi) + dwyi * (yf - yi) + dwzi * (zf - zi));
                    fuhigh = flcf * ue;
                    fvhigh = flcf * ve;
                    fwhigh = flcf * we;
                }
                if (lcds4 && (idew == nj || idew == 1)) {
                    const double c48 = 1.0 / 48.0;
                    const double ue = c48 * (27.0 * u[inp] + 27.0 * u[ine] - 3.0 * u[inw] - 3.0 * u[inee]);
                    const double ve = c48 * (27.0 * v[inp] + 27.0 * v[ine] - 3.0 * v[inw] - 3.0 * v[inee]);
                    const double we = c48 * (27.0 * w[inp] + 27.0 * w[ine] - 3.0 * w[inw] - 3.0 * w[inee]);
                    fuhigh = flcf * ue;
                    fvhigh = flcf * ve;
                    fwhigh = flcf * we;
                }
                if (lluds) {
                    const double dxc = xf - f->xc[inp], dyc = yf - f->yc[inp], dzc = zf - f->zc[inp];
                    const double dxe = f->xc[ine] - xf, dye = f->yc[ine] - yf, dze = f->zc[ine] - zf;
                    const int c = dir - 1;
                    const double dfidx_u = G3(gu, inw, c) * fxw + G3(gu, inp, c) * fxpw;
                    const double dfidy_u = G3(gv, inw, c) * fxw + G3(gv, inp, c) * fxpw;
                    const double dfidz_u = G3(gw, inw, c) * fxw + G3(gw, inp, c) * fxpw;
                    const double dfidx_u1 = G3(gu, inee, c) * fxee + G3(gu, ine, c) * fxew;
                    const double dfidy_u1 = G3(gv, inee, c) * fxee + G3(gv, ine, c) * fxew;
                    const double dfidz_u1 = G3(gw, inee, c) * fxee + G3(gw, ine, c) * fxew;
                    fuhigh = ce * (u[ine] + dxe * dfidx_u1) + cp * (u[inp] + dxc * dfidx_u);
                    fvhigh = ce * (v[ine] + dye * dfidy_u1) + cp * (v[inp] + dyc * dfidy_u);
                    fwhigh = ce * (w[ine] + dze * dfidz_u1) + cp * (w[inp] + dzc * dfidz_u);
                }
                if (lquds && k > 2 && k < nk - 1 && j > 2 && j < nj - 1 && i > 2 && i < ni - 1) {
                    const double fw_ = fif[inp - idew], fp_ = fif[inp], fe_ = fif[inp + idew];
                    const double g11 = ((2.0 - fw_) * (fp_ * fp_)) / (1.0 + fp_ - fw_);
                    const double g12 = ((1.0 - fp_) * ((1.0 - fw_) * (1.0 - fw_))) / (1.0 + fp_ - fw_);
                    const double g21 = ((1.0 + fe_) * ((1.0 - fp_) * (1.0 - fp_))) / (1.0 + fe_ - fp_);
                    const double g22 = ((fe_ * fe_) * fp_) / (1.0 + fe_ - fp_);
                    fuhigh = ce * (g21 * u[inp] + (1.0 - g21 + g22) * u[ine] - g22 * u[inee]) +
                             cp * (g11 * u[ine] + (1.0 - g11 + g12) * u[inp] - g12 * u[inw]);
                    fvhigh = ce * (g21 * v[inp] + (1.0 - g21 + g22) * v[ine] - g22 * v[inee]) +
                             cp * (g11 * v[ine] + (1.0 - g11 + g12) * v[inp] - g12 * v[inw]);
                    fwhigh = ce * (g21 * w[inp] + (1.0 - g21 + g22) * w[ine] - g22 * w[inee]) +
                             cp * (g11 * w[ine] + (1.0 - g11 + g12) * w[inp] - g12 * w[inw]);
                }
                if (flim) {
                    const double r1 = (2.0 * G3(gu, inp, 0) * xpn + 2.0 * G3(gu, inp, 1) * ypn + 2.0 * G3(gu, inp, 2) * zpn) / (u[ine] - u[inp]) - 1.0;
                    const double r2 = (2.0 * G3(gv, inp, 0) * xpn + 2.0 * G3(gv, inp, 1) * ypn + 2.0 * G3(gv, inp, 2) * zpn) / (v[ine] - v[inp]) - 1.0;
                    const double r3 = (2.0 * G3(gw, inp, 0) * xpn + 2.0 * G3(gw, inp, 1) * ypn + 2.0 * G3(gw, inp, 2) * zpn) / (w[ine] - w[inp]) - 1.0;
                    const double r4 = (2.0 * G3(gu, ine, 0) * xpn + 2.0 * G3(gu, ine, 1) * ypn + 2.0 * G3(gu, ine, 2) * zpn) / (u[inp] - u[ine]) - 1.0;
                    const double r5 = (2.0 * G3(gv, ine, 0) * xpn + 2.0 * G3(gv, ine, 1) * ypn + 2.0 * G3(gv, ine, 2) * zpn) / (v[inp] - v[ine]) - 1.0;
                    const double r6 = (2.0 * G3(gw, ine, 0) * xpn + 2.0 * G3(gw, ine, 1) * ypn + 2.0 * G3(gw, ine, 2) * zpn) / (w[inp] - w[ine]) - 1.0;
                    const double r[6] = {r1, r2, r3, r4, r5, r6};
                    double psi[6] = {psiw1, psiw2, psiw3, psie1, psie2, psie3};
                    for (int q = 0; q < 6; ++q) {
                        const double x = r[q];
                        if (flags & ORC_UVW_SMART) psi[q] = orc_max2(0.0, orc_min2(orc_min2(2.0 * x, 0.75 * x + 0.25), 4.0));
                        if (flags & ORC_UVW_AVL) psi[q] = orc_max2(0.0, orc_min2(orc_min2(1.5 * x, 0.75 * x + 0.25), 2.5));
                        if (flags & ORC_UVW_MUSCL) psi[q] = orc_max2(0.0, orc_min2(orc_min2(2.0 * x, 0.5 * x + 0.5), 2.0));
                        if (flags & ORC_UVW_UMIST) psi[q] = orc_max2(0.0, orc_min2(orc_min2(orc_min2(2.0 * x, 0.75 * x + 0.25), 0.25 * x + 0.75), 2.0));
                        if (flags & ORC_UVW_GAMMA) psi[q] = orc_max2(0.0, orc_min2(x, 2.0 * x / (x + 1.0)));
                    }
                    psiw1 = psi[0]; psiw2 = psi[1]; psiw3 = psi[2];
                    psie1 = psi[3]; psie2 = psi[4]; psie3 = psi[5];
                    fuhigh = ce * (u[ine] + fxe * psie1 * (u[inp] - u[ine])) + cp * (u[inp] + fxp * psiw1 * (u[ine] - u[inp]));
                    fvhigh = ce * (v[ine] + fxe * psie2 * (v[inp] - v[ine])) + cp * (v[inp] + fxp * psiw2 * (v[ine] - v[inp]));
                    fwhigh = ce * (w[ine] + fxe * psie3 * (w[inp] - w[ine])) + cp * (w[inp] + fxp * psiw3 * (w[ine] - w[inp]));
                }
                su[inp] = su[inp] + gam * (fuuds - fuhigh) + fdue - fdui;
                sv[inp] = sv[inp] + gam * (fvuds - fvhigh) + fdve - fdvi;
                sw[inp] = sw[inp] + gam * (fwuds - fwhigh) + fdwe - fdwi;
                bp[ine] = -gam * (fuuds - fuhigh) - fdue + fdui;
                bt[ine] = -gam * (fvuds - fvhigh) - fdve + fdvi;
                bb[ine] = -gam * (fwuds - fwhigh) - fdwe + fdwi;
            }
    for (int k = 2; k <= nkmm; ++k)
        for (int i = 2; i <= nimm; ++i)
            for (int j = 2; j <= njmm; ++j) {
                const size_t inp = IX(i, j, k) + idew;
                su[inp] = su[inp] + bp[inp];
                sv[inp] = sv[inp] + bt[inp];
                sw[inp] = sw[inp] + bb[inp];
            }
    const size_t n = (size_t)nij * nk;
    for (size_t c = 0; c < n; ++c) { bp[c] = 0.0; bt[c] = 0.0; bb[c] = 0.0; }
}

/* update_values_at_ghostcells_nonperiodic.f90:3-642: values of u, v, w, te,
 * ed, t, vis in the ghost cells next to the six boundary faces, by the
 * boundary type of each boundary cell (lbw/lbe indexed (j-1)*nk+k, lbs/lbn
 * (i-1)*nk+k, lbb/lbt (i-1)*nj+j): 1 inlet (Dirichlet uin..tin), 2 outlet
 * (linear extrapolation from two inner cells), 3 symmetry, 4 wall.  Edges
 * and corners are not filled.  var = {u, v, w, te, ed, t, vis}; geo = {fx,
 * fy, fz, ar1x, ar1y, ar1z, ar2x, ar2y, ar2z, ar3x, ar3y, ar3z}; inl = {uin,
 * vin, win, tein, edin, tin}. */
static void orc_ghost_cell(double *const *var, size_t inp, size_t ingc, long nb, double fxp,
                           double fxe, const double *ext_f, size_t ext_c, int ext_sign,
                           const double *ax, const double *ay, const double *az, size_t iface,
                           int lb, const double *inl, int side_wall_vis)
{
    double *u = var[0], *v = var[1], *w = var[2], *te = var[3], *ed = var[4], *t = var[5],
           *vis = var[6];
    const double fxer = 1.0 / fxe;
    if (lb == 1) {
        u[ingc] = (inl[0] - u[inp] * fxp) * fxer;
        v[ingc] = (inl[1] - v[inp] * fxp) * fxer;
        w[ingc] = (inl[2] - w[inp] * fxp) * fxer;
        te[ingc] = (inl[3] - te[inp] * fxp) * fxer;
        ed[ingc] = (inl[4] - ed[inp] * fxp) * fxer;
        t[ingc] = (inl[5] - t[inp] * fxp) * fxer;
        vis[ingc] = 0.09 * (te[ingc] * te[ingc]) / ed[ingc];
    }
    if (lb == 2) {
        /* ext_sign < 0: q(inp) - (q(inp+nb) - q(inp)) * f(ext_c)   (west, south, bottom)
         * ext_sign > 0: q(inp) + (q(inp) - q(inp-nb)) * (1 - f(ext_c))  (east, north, top) */
        for (int c = 0; c < 7; ++c) {
            double *q = var[c];
            if (ext_sign < 0) q[ingc] = q[inp] - (q[inp + nb] - q[inp]) * ext_f[ext_c];
            else q[ingc] = q[inp] + (q[inp] - q[inp - nb]) * (1.0 - ext_f[ext_c]);
        }
    }
    if (lb == 3) {
        const double un = u[inp] * ax[iface] + v[inp] * ay[iface] + w[inp] * az[iface];
        u[ingc] = u[inp] - un * ax[iface];
        v[ingc] = v[inp] - un * ay[iface];
        w[ingc] = w[inp] - un * az[iface];
        te[ingc] = te[inp];
        ed[ingc] = ed[inp];
        t[ingc] = t[inp];
        vis[ingc] = vis[inp];
    }
    if (lb == 4) {
        u[ingc] = (-u[inp] * fxp) * fxer;
        v[ingc] = (-v[inp] * fxp) * fxer;
        w[ingc] = (-w[inp] * fxp) * fxer;
        te[ingc] = (-te[inp] * fxp) * fxer;
        ed[ingc] = ed[inp];
        t[ingc] = (inl[5] - t[inp] * fxp) * fxer;
        vis[ingc] = side_wall_vis ? (vis[ingc] - vis[inp] * fxp) * fxer : vis[inp];
    }
}

void orc_ghostcells(int ni, int nj, int nk, double *const *var, const double *const *geo,
                    const int *lbw, const int *lbe, const int *lbs, const int *lbn,
                    const int *lbb, const int *lbt, const double *inl)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const double *fx = geo[0], *fy = geo[1], *fz = geo[2];
    for (int k = 3; k <= nkmm; ++k)          /* west (45-139) */
        for (int j = 3; j <= njmm; ++j) {
            const size_t inp = IX(3, j, k), ingc = inp - nj, iface = inp - nj;
            const double fxp = fx[inp - nj], fxe = 1.0 - fxp;
            orc_ghost_cell(var, inp, ingc, nj, fxp, fxe, fx, inp, -1, geo[3], geo[4], geo[5],
                           iface, lbw[(j - 1) * nk + k - 1], inl, 1);
        }
    for (int k = 3; k <= nkmm; ++k)          /* east (143-239) */
        for (int j = 3; j <= njmm; ++j) {
            const size_t inp = IX(nimm, j, k), ingc = inp + nj, iface = inp;
            const double fxp = 1.0 - fx[inp], fxe = 1.0 - fxp;
            orc_ghost_cell(var, inp, ingc, nj, fxp, fxe, fx, inp - nj, 1, geo[3], geo[4],
                           geo[5], iface, lbe[(j - 1) * nk + k - 1], inl, 1);
        }
    for (int k = 3; k <= nkmm; ++k)          /* south (243-337) */
        for (int i = 3; i <= nimm; ++i) {
            const size_t inp = IX(i, 3, k), ingc = inp - 1, iface = inp - 1;
            const double fxp = 1.0 - fy[inp], fxe = 1.0 - fxp;
            orc_ghost_cell(var, inp, ingc, 1, fxp, fxe, fy, inp, -1, geo[6], geo[7], geo[8],
                           iface, lbs[(i - 1) * nk + k - 1], inl, 1);
        }
    for (int k = 3; k <= nkmm; ++k)          /* north (341-437) */
        for (int i = 3; i <= nimm; ++i) {
            const size_t inp = IX(i, njmm, k), ingc = inp + 1, iface = inp;
            const double fxp = 1.0 - fy[inp], fxe = 1.0 - fxp;
            orc_ghost_cell(var, inp, ingc, 1, fxp, fxe, fy, inp - 1, 1, geo[6], geo[7], geo[8],
                           iface, lbn[(i - 1) * nk + k - 1], inl, 1);
        }
    for (int i = 3; i <= nimm; ++i)          /* bottom (441-537) */
        for (int j = 3; j <= njmm; ++j) {
            const size_t inp = IX(i, j, 3), ingc = inp - nij, iface = inp - nij;
            const double fxp = fz[ingc], fxe = 1.0 - fxp;
            orc_ghost_cell(var, inp, ingc, nij, fxp, fxe, fz, inp, -1, geo[9], geo[10], geo[11],
                           iface, lbb[(i - 1) * nj + j - 1], inl, 0);
        }
    for (int i = 3; i <= nimm; ++i)          /* top (541-638) */
        for (int j = 3; j <= njmm; ++j) {
            const size_t inp = IX(i, j, nkmm), ingc = inp + nij, iface = inp;
            const double fxp = 1.0 - fz[inp], fxe = 1.0 - fxp;
            orc_ghost_cell(var, inp, ingc, nij, fxp, fxe, fz, inp - nij, 1, geo[9], geo[10],
                           geo[11], iface, lbt[(i - 1) * nj + j - 1], inl, 0);
        }
}

/* CourantNo.h:10-81 (ltransient branch): su = |f1|+|f1(w)|+|f2|+|f2(s)|+
 * |f3|+|f3(b)| per cell on k = 2..nkm, i, j = 3..n?mm (su = 0 elsewhere),
 * CoNum = 0.5 * max(su/vol) * dt, meanCoNum = 0.5 * sum(su)/sum(vol) * dt;
 * with CoNumFix the time step is rescaled to CoNumFixValue (58-64), and time
 * advances by the time step.  out = {CoNum, meanCoNum}.  (global_max is the
 * identity on one rank.) */
void orc_courant(int ni, int nj, int nk, const double *f1, const double *f2, const double *f3,
                 const double *vol, double *su, double *timestep, double *time, int conumfix,
                 double conumfixvalue, double *out)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    double conum = 0.0, meanconum = 0.0, suma = 0.0;
    psum mean_ps = PS0, suma_ps = PS0;
    memset(su, 0, (size_t)nij * nk * sizeof(double));
    for (int k = 2; k <= nkm; ++k)
        for (int i = 3; i <= nimm; ++i)
            for (int j = 3; j <= njmm; ++j) {
                const size_t inp = IX(i, j, k);
                su[inp] = fabs(f1[inp]) + fabs(f1[inp - nj]) + fabs(f2[inp]) + fabs(f2[inp - 1]) +
                          fabs(f3[inp]) + fabs(f3[inp - nij]);
                conum = orc_max2(conum, su[inp] / vol[inp]);
                suma = ps_add(&suma_ps, vol[inp], k);
                meanconum = ps_add(&mean_ps, su[inp], k);
            }
    conum = 0.5 * conum * *timestep;
    meanconum = 0.5 * meanconum / suma * *timestep;
    if (conumfix) {
        const double dt = *timestep;
        *timestep = conumfixvalue / conum * *timestep;
        conum = conum * *timestep / dt;
        meanconum = meanconum * *timestep / dt;
    }
    *time = *time + *timestep;
    out[0] = conum;
    out[1] = meanconum;
}

/* ================================================ pressure-correction drivers */

static orc_fields drv_fields(const orc_flow *F)
{
    orc_fields f = {F->x, F->y, F->z, F->xc, F->yc, F->zc, F->vol, F->fx, F->fy, F->fz,
                    F->ar1x, F->ar1y, F->ar1z, F->ar2x, F->ar2y, F->ar2z, F->ar3x, F->ar3y,
                    F->ar3z, F->den, F->apu, F->apv, F->apw, F->u, F->v, F->w, F->p,
                    F->gradu, F->gradv, F->gradw, F->gradp};
    return f;
}

/* `call grad_gauss(u, gradu(1,:), gradu(2,:), gradu(3,:))`: the strided
 * sections are copied in, zeroed and filled by the callee, copied out */
static void drv_grad(int ni, int nj, int nk, const orc_flow *F, const double *u, double *g3)
{
    const size_t n = (size_t)ni * nj * nk;
    double *d = (double *)malloc(3 * n * sizeof(double));
    orc_grad_gauss(ni, nj, nk, u, F->fx, F->fy, F->fz, F->ar1x, F->ar1y, F->ar1z, F->ar2x,
                   F->ar2y, F->ar2z, F->ar3x, F->ar3y, F->ar3z, F->vol, d, d + n, d + 2 * n);
    for (size_t c = 0; c < n; ++c) {
        g3[3 * c] = d[c];
        g3[3 * c + 1] = d[n + c];
        g3[3 * c + 2] = d[2 * n + c];
    }
    free(d);
}

static orc_sys drv_sys(orc_flow *F)
{
    orc_sys s = {F->ae, F->aw, F->an, F->as, F->at, F->ab, F->ap, F->su, F->res,
                 F->bb, F->bs, F->bw, F->bp, F->bn, F->be, F->bt};
    return s;
}

/* continuityErrors.h:55-87 */
static void drv_continuity(int ni, int nj, int nk, const orc_flow *F, double *cum, double *out)
{
    double s[2];
    orc_continuity(ni, nj, nk, F->f1, F->f2, F->f3, s);
    *cum = *cum + s[0];
    out[0] = s[1];
    out[1] = s[0];
    out[2] = *cum;
}

/* outbc.f90:3-318.  The outflow face fluxes use ar1x/ar1y/ar1z on every
 * face, as the reference does; flow is the running sum in face order. */
void orc_outbc(int ni, int nj, int nk, orc_flow *F, const orc_drv_opts *o)
{
    if (!o->lout) return;
    const int nij = ni * nj, nim = ni - 1, njm = nj - 1, nkm = nk - 1;
    psum fl = PS0;
    double flow = 0.0;
#define OUTF(fa, cface, cin)                                                                    \
    fa[cface] = F->den[inp] * (F->ar1x[inp] * F->u[cin] + F->ar1y[inp] * F->v[cin] +              \
                               F->ar1z[inp] * F->w[cin])
    /* flux pass: west, east, south, north, bottom, top (40-118) */
    if (o->low)
        for (int k = 2; k <= nkm; ++k) for (int j = 2; j <= njm; ++j) {
            size_t inp = IX(2, j, k); int inbc = (j - 1) * nk + k;
            if (o->lbw[inbc - 1] == 2) { size_t inh = inp + nj; OUTF(F->f1, inp, inh); flow = ps_add(&fl, -F->f1[inp], 0); }
        }
    if (o->loe)
        for (int k = 2; k <= nkm; ++k) for (int j = 2; j <= njm; ++j) {
            size_t inp = IX(nim, j, k); int inbc = (j - 1) * nk + k;
            if (o->lbe[inbc - 1] == 2) { size_t inh = inp - nj; OUTF(F->f1, inh, inh); flow = ps_add(&fl, F->f1[inh], 0); }
        }
    if (o->los)
        for (int k = 2; k <= nkm; ++k) for (int i = 2; i <= nim; ++i) {
            size_t inp = IX(i, 2, k); int inbc = (i - 1) * nk + k;
            if (o->lbs[inbc - 1] == 2) { size_t inh = inp + 1; OUTF(F->f2, inp, inh); flow = ps_add(&fl, -F->f2[inp], 0); }
        }
    if (o->lon)
        for (int k = 2; k <= nkm; ++k) for (int i = 2; i <= nim; ++i) {
            size_t inp = IX(i, njm, k); int inbc = (i - 1) * nk + k;
            if (o->lbn[inbc - 1] == 2) { size_t inh = inp - 1; OUTF(F->f2, inh, inh); flow = ps_add(&fl, F->f2[inh], 0); }
        }
    if (o->lob)
        for (int i = 2; i <= nim; ++i) for (int j = 2; j <= njm; ++j) {
            size_t inp = IX(i, j, 2); int inbc = (i - 1) * nj + j;
            if (o->lbb[inbc - 1] == 2) { size_t inh = inp + nij; OUTF(F->f3, inp, inh); flow = ps_add(&fl, -F->f3[inp], 0); }
        }
    if (o->lot)
        for (int i = 2; i <= nim; ++i) for (int j = 2; j <= njm; ++j) {
            size_t inp = IX(i, j, nkm); int inbc = (i - 1) * nj + j;
            if (o->lbt[inbc - 1] == 2) { size_t inh = inp - nij; OUTF(F->f3, inh, inh); flow = ps_add(&fl, F->f3[inh], 0); }
        }
#undef OUTF
    /* the running sum starts from flow = 0. (a REAL*4 zero, exact) */
    const double fac = o->flowin / (flow + SMALL);
#define COPYS(inp, inh)                                                                         \
    do {                                                                                        \
        if (F->te) F->te[inp] = F->te[inh];                                                     \
        if (F->ed) F->ed[inp] = F->ed[inh];                                                     \
        if (F->vis) F->vis[inp] = F->vis[inh];                                                  \
        if (F->t) F->t[inp] = F->t[inh];                                                        \
        if (F->vart) F->vart[inp] = F->vart[inh];                                               \
        if (F->con) F->con[inp] = F->con[inh];                                                  \
        if (F->utt) F->utt[inp] = F->utt[inh];                                                  \
        if (F->vtt) F->vtt[inp] = F->vtt[inh];                                                  \
        if (F->wtt) F->wtt[inp] = F->wtt[inh];                                                  \
    } while (0)
#define SCALE(inp, inh)                                                                         \
    do {                                                                                        \
        F->u[inp] = F->u[inh] * fac;                                                            \
        F->v[inp] = F->v[inh] * fac;                                                            \
        F->w[inp] = F->w[inh] * fac;                                                            \
    } while (0)
    /* scaling pass (127-318) */
    if (o->low)
        for (int k = 2; k <= nkm; ++k) for (int j = 2; j <= njm; ++j) {
            size_t inp = IX(2, j, k); int inbc = (j - 1) * nk + k;
            if (o->lbw[inbc - 1] == 2) { size_t inh = inp + nj; SCALE(inp, inh); F->f1[inp] = F->f1[inp] * fac; COPYS(inp, inh); }
        }
    if (o->loe)
        for (int k = 2; k <= nkm; ++k) for (int j = 2; j <= njm; ++j) {
            size_t inp = IX(nim, j, k); int inbc = (j - 1) * nk + k;
            if (o->lbe[inbc - 1] == 2) { size_t inh = inp - nj; SCALE(inp, inh); F->f1[inh] = F->f1[inh] * fac; COPYS(inp, inh); }
        }
    if (o->los)
        for (int k = 2; k <= nkm; ++k) for (int i = 2; i <= nim; ++i) {
            size_t inp = IX(i, 2, k); int inbc = (i - 1) * nk + k;
            if (o->lbs[inbc - 1] == 2) { size_t inh = inp + 1; SCALE(inp, inh); F->f2[inp] = F->f2[inp] * fac; COPYS(inp, inh); }
        }
    if (o->lon)
        for (int k = 2; k <= nkm; ++k) for (int i = 2; i <= nim; ++i) {
            size_t inp = IX(i, njm, k); int inbc = (i - 1) * nk + k;
            if (o->lbn[inbc - 1] == 2) { size_t inh = inp - 1; SCALE(inp, inh); F->f2[inh] = F->f2[inh] * fac; COPYS(inp, inh); }
        }
    if (o->lob)
        for (int i = 2; i <= nim; ++i) for (int j = 2; j <= njm; ++j) {
            size_t inp = IX(i, j, 2); int inbc = (i - 1) * nj + j;
            if (o->lbb[inbc - 1] == 2) { size_t inh = inp + nij; SCALE(inp, inh); F->f3[inp] = F->f3[inp] * fac; COPYS(inp, inh); }
        }
    if (o->lot)
        for (int i = 2; i <= nim; ++i) for (int j = 2; j <= njm; ++j) {
            size_t inp = IX(i, j, nkm); int inbc = (i - 1) * nj + j;
            if (o->lbt[inbc - 1] == 2) { size_t inh = inp - nij; SCALE(inp, inh); F->f3[inh] = F->f3[inh] * fac; COPYS(inp, inh); }
        }
#undef COPYS
#undef SCALE
}

/* correctBoundaryConditions.f90:1-86: periodic copies into the padding
 * planes, in the reference's order (far east, far west -- whose source plane
 * is li(njm), as written -- north, south) */
void orc_correct_bc(int ni, int nj, int nk, orc_flow *F)
{
    const int nij = ni * nj, nim = ni - 1, njm = nj - 1, nkm = nk - 1;
    (void)nij;
#define CP3(inp, inh) do { F->u[inp] = F->u[inh]; F->v[inp] = F->v[inh]; F->w[inp] = F->w[inh]; } while (0)
    for (int k = 2; k <= nkm; ++k) for (int j = 2; j <= njm; ++j) {
        size_t inp = IX(ni, j, k), inh = IX(2, j, k);
        CP3(inp, inh);
    }
    for (int k = 2; k <= nkm; ++k) for (int j = 2; j <= njm; ++j) {
        size_t inp = IX(1, j, k), inh = IX(njm, j, k);
        CP3(inp, inh);
        F->f1[inp] = F->f1[inh];
    }
    for (int k = 2; k <= nkm; ++k) for (int i = 2; i <= nim; ++i) {
        size_t inp = IX(i, nj, k), inh = IX(i, 2, k);
        CP3(inp, inh);
    }
    for (int k = 2; k <= nkm; ++k) for (int i = 2; i <= nim; ++i) {
        size_t inp = IX(i, 1, k), inh = IX(i, njm, k);
        CP3(inp, inh);
        F->f2[inp] = F->f2[inh];
    }
#undef CP3
}

/* PISO_getHbyA.f90:3-233: su, sv, sw = b + sum h_nb u_nb on k, i, j =
 * 3..n?mm (buoyancy 54-73, BDF 76-91, Crank-Nicolson 109-123 / 144-158 /
 * 178-192), then u = su * apu (and v, w) (203-213), then the velocity
 * gradients (Gauss branch, 222-226) */
void orc_gethbya(int ni, int nj, int nk, orc_flow *F, const orc_drv_opts *o)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    for (int k = 3; k <= nkmm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
        size_t inp = IX(i, j, k);
        F->sp[inp] = 0.0;
        F->su[inp] = 0.0;
        F->sv[inp] = 0.0;
        F->sw[inp] = 0.0;
        if (o->buoy) {
            double heat = 0.0;
            if (o->boussinesq) heat = o->beta * o->densit * (F->t[inp] - o->tref) * F->vol[inp];
            else heat = (o->densit - F->den[inp]) * F->vol[inp];
            F->su[inp] = F->su[inp] - o->gravx * heat;
            F->sv[inp] = F->sv[inp] - o->gravy * heat;
            F->sw[inp] = F->sw[inp] - o->gravz * heat;
        }
        if (o->bdf) {
            const double apotime = F->den[inp] * F->vol[inp] / o->timestep;
            const double sut = apotime * ((1.0 + o->btime) * F->uo[inp] - 0.5 * o->btime * F->uoo[inp]);
            const double svt = apotime * ((1.0 + o->btime) * F->vo[inp] - 0.5 * o->btime * F->voo[inp]);
            const double swt = apotime * ((1.0 + o->btime) * F->wo[inp] - 0.5 * o->btime * F->woo[inp]);
            F->su[inp] = F->su[inp] + sut;
            F->sv[inp] = F->sv[inp] + svt;
            F->sw[inp] = F->sw[inp] + swt;
        }
    }
    double *S[3] = {F->su, F->sv, F->sw};
    const double *X[3] = {F->u, F->v, F->w}, *XO[3] = {F->uo, F->vo, F->wo};
    for (int q = 0; q < 3; ++q)
        for (int k = 3; k <= nkmm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
            size_t inp = IX(i, j, k), ine = inp + nj, inw = inp - nj, inn = inp + 1, ins = inp - 1,
                   intp = inp + nij, inb = inp - nij;
            const double *x = X[q];
            if (o->cn) {
                const double *xo = XO[q];
                const double apotime = F->den[inp] * F->vol[inp] / o->timestep;
                S[q][inp] = S[q][inp] + (F->he[inp] * xo[ine] + F->hw[inp] * xo[inw] +
                                         F->hn[inp] * xo[inn] + F->hs[inp] * xo[ins] +
                                         F->ht[inp] * xo[intp] + F->hb[inp] * xo[inb]) +
                            (apotime - F->he[inp] - F->hw[inp] - F->hn[inp] - F->hs[inp] -
                             F->ht[inp] - F->hb[inp]) * xo[inp];
            }
            S[q][inp] = S[q][inp] + (F->he[inp] * x[ine] + F->hw[inp] * x[inw] +
                                     F->hn[inp] * x[inn] + F->hs[inp] * x[ins] +
                                     F->ht[inp] * x[intp] + F->hb[inp] * x[inb]);
        }
    for (int k = 3; k <= nkmm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
        size_t inp = IX(i, j, k);
        F->u[inp] = F->su[inp] * F->apu[inp];
        F->v[inp] = F->sv[inp] * F->apv[inp];
        F->w[inp] = F->sw[inp] * F->apw[inp];
    }
    drv_grad(ni, nj, nk, F, F->u, F->gradu);
    drv_grad(ni, nj, nk, F, F->v, F->gradv);
    drv_grad(ni, nj, nk, F, F->w, F->gradw);
}

static void drv_solve(int ni, int nj, int nk, orc_flow *F, double *x, int solver, double alfa,
                      double sor, int nsw, orc_report *rep)
{
    orc_sys s = drv_sys(F);
    switch (solver) {
    case 1: orc_sipsol(ni, nj, nk, x, &s, alfa, sor, nsw, rep); break;
    case 2: orc_cgstab(ni, nj, nk, x, &s, sor, nsw, rep); break;
    case 3: orc_iccg(ni, nj, nk, x, &s, sor, nsw, rep); break;
    case 4: orc_pcgsip(ni, nj, nk, x, &s, alfa, sor, nsw, rep); break;
    default: orc_cgstab_sip(ni, nj, nk, x, &s, alfa, sor, nsw, rep); break;
    }
}

/* calcp-multiple_correction_SIMPLE.f90:3-238 with the Gauss gradients
 * (gauss branch of 38-46 and 112-116) */
void orc_calcp(int ni, int nj, int nk, orc_flow *F, const orc_drv_opts *o, orc_report *reps,
               double *cont)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const size_t n = (size_t)nij * nk;
    double cum = o->cumulative;
    double *fm = (double *)malloc(3 * n * sizeof(double));
    for (int icorr = 1; icorr <= o->ncorr; ++icorr) {
        drv_grad(ni, nj, nk, F, F->u, F->gradu);
        drv_grad(ni, nj, nk, F, F->v, F->gradv);
        drv_grad(ni, nj, nk, F, F->w, F->gradw);
        orc_fields f = drv_fields(F);
        orc_fluxmass(ni, nj, nk, &f, nj, 1, nij, F->ar1x, F->ar1y, F->ar1z, F->fx, F->ae, F->aw, F->f1);
        orc_fluxmass(ni, nj, nk, &f, 1, nij, nj, F->ar2x, F->ar2y, F->ar2z, F->fy, F->an, F->as, F->f2);
        orc_fluxmass(ni, nj, nk, &f, nij, nj, 1, F->ar3x, F->ar3y, F->ar3z, F->fz, F->at, F->ab, F->f3);
        if (!o->const_mflux) orc_outbc(ni, nj, nk, F, o);
        for (int k = 3; k <= nkmm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
            size_t inp = IX(i, j, k);
            F->su[inp] = -F->f1[inp] + F->f1[inp - nj] - F->f2[inp] + F->f2[inp - 1] - F->f3[inp] +
                         F->f3[inp - nij];
            F->ap[inp] = F->ae[inp] + F->aw[inp] + F->an[inp] + F->as[inp] + F->at[inp] + F->ab[inp];
        }
        for (int ipcorr = 1; ipcorr <= o->npcor; ++ipcorr) {
            memset(F->pp, 0, n * sizeof(double));
            drv_solve(ni, nj, nk, F, F->pp, o->solver, o->alfa, o->sor, o->nsw,
                      &reps[(icorr - 1) * o->npcor + ipcorr - 1]);
            const double ppref = F->pp[o->ijkpr - 1];
            drv_grad(ni, nj, nk, F, F->pp, F->gradp);
            if (ipcorr != o->npcor) {
                memset(F->su, 0, n * sizeof(double));
                f = drv_fields(F);
                orc_fluxmc_all(ni, nj, nk, &f, nj, 1, nij, F->fx, fm);
                orc_fluxmc_all(ni, nj, nk, &f, 1, nij, nj, F->fy, fm + n);
                orc_fluxmc_all(ni, nj, nk, &f, nij, nj, 1, F->fz, fm + 2 * n);
                for (int k = 3; k <= nkmm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
                    size_t inp = IX(i, j, k);
                    double fmcor = fm[inp];
                    F->f1[inp] = F->f1[inp] - fmcor;
                    F->su[inp] = F->su[inp] + fmcor;
                    F->su[inp + nj] = F->su[inp + nj] - fmcor;
                    fmcor = fm[n + inp];
                    F->f2[inp] = F->f2[inp] - fmcor;
                    F->su[inp] = F->su[inp] + fmcor;
                    F->su[inp + 1] = F->su[inp + 1] - fmcor;
                    fmcor = fm[2 * n + inp];
                    F->f3[inp] = F->f3[inp] - fmcor;
                    F->su[inp] = F->su[inp] + fmcor;
                    F->su[inp + nij] = F->su[inp + nij] - fmcor;
                }
            }
            for (int k = 3; k <= nkmm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
                size_t inp = IX(i, j, k);
                F->f1[inp] = F->f1[inp] - F->ae[inp] * (F->pp[inp + nj] - F->pp[inp]);
                F->f2[inp] = F->f2[inp] - F->an[inp] * (F->pp[inp + 1] - F->pp[inp]);
                F->f3[inp] = F->f3[inp] - F->at[inp] * (F->pp[inp + nij] - F->pp[inp]);
            }
            for (int k = 3; k <= nkmm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
                size_t inp = IX(i, j, k);
                F->u[inp] = F->u[inp] - F->vol[inp] * F->apu[inp] * F->gradp[3 * inp];
                F->v[inp] = F->v[inp] - F->vol[inp] * F->apv[inp] * F->gradp[3 * inp + 1];
                F->w[inp] = F->w[inp] - F->vol[inp] * F->apw[inp] * F->gradp[3 * inp + 2];
                F->p[inp] = F->p[inp] + o->urf * (F->pp[inp] - ppref);
            }
        }
        drv_continuity(ni, nj, nk, F, &cum, cont + 3 * (icorr - 1));
        orc_correct_bc(ni, nj, nk, F);
    }
    orc_bpres(ni, nj, nk, F->p, F->fx, F->fy, F->fz);
    memcpy(F->pp, F->p, n * sizeof(double));
    free(fm);
}

/* PISO_multiple_correction.f90:3-297 (pimple = 0) and
 * PIMPLE_multiple_correction.f90:3-302 (pimple = 1), Gauss gradients */
void orc_piso(int ni, int nj, int nk, orc_flow *F, const orc_drv_opts *o, int pimple,
              orc_report *reps, double *cont)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const size_t n = (size_t)nij * nk;
    double cum = o->cumulative;
    memcpy(F->he, F->ae, n * sizeof(double));
    memcpy(F->hw, F->aw, n * sizeof(double));
    memcpy(F->hn, F->an, n * sizeof(double));
    memcpy(F->hs, F->as, n * sizeof(double));
    memcpy(F->ht, F->at, n * sizeof(double));
    memcpy(F->hb, F->ab, n * sizeof(double));
    memset(F->gradp, 0, 3 * n * sizeof(double));
    memset(F->p, 0, n * sizeof(double));
    double *x = pimple ? F->pp : F->p;     /* the variable solved for */
    const size_t inpr = (size_t)o->ijkpr - 1;
    for (int icorr = 1; icorr <= o->ncorr; ++icorr) {
        if (!pimple) orc_gethbya(ni, nj, nk, F, o);
        for (int ipcorr = 1; ipcorr <= o->npcor; ++ipcorr) {
            orc_fields f = drv_fields(F);
            orc_piso_assemble(ni, nj, nk, &f, nj, 1, nij, F->ar1x, F->ar1y, F->ar1z, F->fx, F->ae,
                              F->aw, F->f1, ipcorr, F->be, F->su);
            orc_piso_assemble(ni, nj, nk, &f, 1, nij, nj, F->ar2x, F->ar2y, F->ar2z, F->fy, F->an,
                              F->as, F->f2, ipcorr, F->bn, F->su);
            orc_piso_assemble(ni, nj, nk, &f, nij, nj, 1, F->ar3x, F->ar3y, F->ar3z, F->fz, F->at,
                              F->ab, F->f3, ipcorr, F->bt, F->su);
            if (ipcorr == 1) {
                if (!o->const_mflux) orc_outbc(ni, nj, nk, F, o);
                for (int k = 3; k <= nkmm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
                    size_t inp = IX(i, j, k);
                    F->su[inp] = -F->f1[inp] + F->f1[inp - nj] - F->f2[inp] + F->f2[inp - 1] -
                                 F->f3[inp] + F->f3[inp - nij];
                }
            }
            for (int k = 3; k <= nkmm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
                size_t inp = IX(i, j, k);
                F->ap[inp] = F->ae[inp] + F->aw[inp] + F->an[inp] + F->as[inp] + F->at[inp] + F->ab[inp];
            }
            F->ap[inpr] = 1.0;
            F->ae[inpr] = 0.0;
            F->aw[inpr] = 0.0;
            F->an[inpr] = 0.0;
            F->as[inpr] = 0.0;
            F->at[inpr] = 0.0;
            F->ab[inpr] = 0.0;
            const double ppref = F->p[inpr];
            F->su[inpr] = ppref;
            /* sor(ip) = 0.1 is a default-real literal */
            const double sor = icorr != o->ncorr ? (double)0.1f : o->sor;
            drv_solve(ni, nj, nk, F, x, 4, o->alfa, sor, o->nsw,
                      &reps[(icorr - 1) * o->npcor + ipcorr - 1]);
            orc_bpres(ni, nj, nk, x, F->fx, F->fy, F->fz);
            drv_grad(ni, nj, nk, F, x, F->gradp);
            if (ipcorr == o->npcor)
                for (int k = 2; k <= nkmm; ++k) for (int i = 2; i <= nimm; ++i) for (int j = 2; j <= njmm; ++j) {
                    size_t inp = IX(i, j, k);
                    F->f1[inp] = F->f1[inp] - F->ae[inp] * (x[inp + nj] - x[inp]);
                    F->f2[inp] = F->f2[inp] - F->an[inp] * (x[inp + 1] - x[inp]);
                    F->f3[inp] = F->f3[inp] - F->at[inp] * (x[inp + nij] - x[inp]);
                }
        }
        drv_continuity(ni, nj, nk, F, &cum, cont + 3 * (icorr - 1));
        for (int k = 3; k <= nkmm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
            size_t inp = IX(i, j, k);
            F->u[inp] = F->u[inp] - F->apu[inp] * F->gradp[3 * inp] * F->vol[inp];
            F->v[inp] = F->v[inp] - F->apv[inp] * F->gradp[3 * inp + 1] * F->vol[inp];
            F->w[inp] = F->w[inp] - F->apw[inp] * F->gradp[3 * inp + 2] * F->vol[inp];
            if (pimple) F->p[inp] = F->p[inp] + o->urf * (F->pp[inp] - F->p[inp]);
        }
        if (pimple) orc_bpres(ni, nj, nk, F->p, F->fx, F->fy, F->fz);
    }
    memcpy(F->pp, F->p, n * sizeof(double));
}

/* ======================================================= LES and momentum */

void orc_strain_rate(int ni, int nj, int nk, const double *gradu, const double *gradv,
                     const double *gradw, double *strain, double *vorticity)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    (void)nij;
    for (int k = 2; k <= nkm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
        const size_t c = IX(i, j, k);
        const double dudx = gradu[3 * c], dudy = gradu[3 * c + 1], dudz = gradu[3 * c + 2];
        const double dvdx = gradv[3 * c], dvdy = gradv[3 * c + 1], dvdz = gradv[3 * c + 2];
        const double dwdx = gradw[3 * c], dwdy = gradw[3 * c + 1], dwdz = gradw[3 * c + 2];
        const double s11 = dudx, s12 = 0.5 * (dudy + dvdx), s13 = 0.5 * (dudz + dwdx);
        const double s22 = dvdy, s23 = 0.5 * (dvdz + dwdy), s33 = dwdz;
        const double w12 = 0.5 * (dudy - dvdx), w13 = 0.5 * (dudz - dwdx), w23 = 0.5 * (dvdz - dwdy);
        strain[c] = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 +
                                2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
        vorticity[c] = sqrt(w12 * w12 + w23 * w23 + w13 * w13);
    }
}

/* the wall-function viscosity of one wall cell pair (calc_vis_les.f90:84-115) */
static void les_wall(size_t inbc, size_t ingc, size_t iface, double fxp, const double *u,
                     const double *v, const double *w, const double *xc, const double *yc,
                     const double *zc, const double *ar3x, const double *ar3y, const double *ar3z,
                     double *vis, double viscos, double densit, double cappa, double elog,
                     double dnx, double dny, double dnz)
{
    const double fxe = 1.0 - fxp;
    const double gx = xc[inbc] * fxp + xc[ingc] * fxe;
    const double gy = yc[inbc] * fxp + yc[ingc] * fxe;
    const double gz = zc[inbc] * fxp + zc[ingc] * fxe;
    const double are = sqrt(ar3x[iface] * ar3x[iface] + ar3y[iface] * ar3y[iface] +
                            ar3z[iface] * ar3z[iface]);
    const double nxf = ar3x[iface] / are, nyf = ar3y[iface] / are, nzf = ar3z[iface] / are;
    const double vnp = u[inbc] * nxf + v[inbc] * nyf + w[inbc] * nzf;
    double xtp = u[inbc] - vnp * nxf, ytp = v[inbc] - vnp * nyf, ztp = w[inbc] - vnp * nzf;
    const double vtp = sqrt(xtp * xtp + ytp * ytp + ztp * ztp);
    xtp = xtp / vtp;
    ytp = ytp / vtp;
    ztp = ztp / vtp;
    const double ut2 = fabs(u[inbc] * xtp + v[inbc] * ytp + w[inbc] * ztp);
    const double dnw = fabs((xc[inbc] - gx) * dnx + (yc[inbc] - gy) * dny + (zc[inbc] - gz) * dnz);
    const double tw = viscos * ut2 / dnw;
    const double utau = sqrt(tw / densit);
    const double ypl = densit * utau * dnw / viscos;
    double viscw = 0.0;
    if (ypl > (double)11.61f) viscw = ypl * viscos * cappa / log(elog * ypl);
    vis[ingc] = orc_max2(viscos, viscw);
}

void orc_calc_vis_les(int ni, int nj, int nk, const double *u, const double *v, const double *w,
                      const double *vol, const double *den, const double *fz, const double *xc,
                      const double *yc, const double *zc, const double *ar3x, const double *ar3y,
                      const double *ar3z, const double *strain, const double *walldist,
                      double *vis, double viscos, double densit, double cappa, double elog,
                      double urfvis, int nx, int ny, int nz, int top_k)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    /* default-real literals: C_delta = 0.158, C_Les = 0.17, 1./3. */
    const double c_delta = (double)0.158f, c_les = (double)0.17f, third = (double)(1.0f / 3.0f);
    for (int k = 2; k <= nkm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
        const size_t c = IX(i, j, k);
        const double visold = vis[c];
        const double wl = walldist[c];
        const double uref = sqrt(u[c] * u[c] + v[c] * v[c] + w[c] * w[c]);
        const double utau = sqrt(viscos * uref / wl);
        const double zplus = wl * utau / viscos;
        const double l = orc_min2(c_les * pow(vol[c], third), cappa / c_delta * wl) *
                         (1.0 - exp(-zplus / 26.0));
        const double vs = l * l * strain[c] * den[c];
        vis[c] = vs + viscos;
        vis[c] = urfvis * vis[c] + (1.0 - urfvis) * visold;
    }
    const double dnx = nx, dny = ny, dnz = nz;
    for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
        const size_t inbc = IX(i, j, 3), ingc = inbc - nij;
        les_wall(inbc, ingc, ingc, fz[ingc], u, v, w, xc, yc, zc, ar3x, ar3y, ar3z, vis, viscos,
                 densit, cappa, elog, dnx, dny, dnz);
    }
    /* the top wall's cells are taken on the plane lk(nimm), as written (or
     * on top_k when set) */
    const int ktop_les = top_k ? top_k : nimm;
    for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
        const size_t inbc = IX(i, j, ktop_les), ingc = inbc + nij;
        les_wall(inbc, ingc, inbc, 1.0 - fz[inbc], u, v, w, xc, yc, zc, ar3x, ar3y, ar3z, vis,
                 viscos, densit, cappa, elog, dnx, dny, dnz);
    }
}

/* the wall shear of one wall cell pair in calcuvw (calcuvw.f90:160-240,
 * first-order branch) */
static void uvw_wall(orc_flow *F, double *spv, size_t inbc, size_t ingc, size_t iface, double fxp,
                     double dnx, double dny, double dnz)
{
    const double fxe = 1.0 - fxp;
    const double gx = F->xc[inbc] * fxp + F->xc[ingc] * fxe;
    const double gy = F->yc[inbc] * fxp + F->yc[ingc] * fxe;
    const double gz = F->zc[inbc] * fxp + F->zc[ingc] * fxe;
    const double are = sqrt(F->ar3x[iface] * F->ar3x[iface] + F->ar3y[iface] * F->ar3y[iface] +
                            F->ar3z[iface] * F->ar3z[iface]);
    const double nxf = F->ar3x[iface] / are, nyf = F->ar3y[iface] / are, nzf = F->ar3z[iface] / are;
    const double dnw = fabs((F->xc[inbc] - gx) * dnx + (F->yc[inbc] - gy) * dny +
                            (F->zc[inbc] - gz) * dnz);
    const double viss = F->vis[ingc];
    const double cf = viss * are / dnw;
    const double dxx = F->xc[inbc] - F->xc[ingc], dyy = F->yc[inbc] - F->yc[ingc],
                 dzz = F->zc[inbc] - F->zc[ingc];
    const double dpb = 0.5 * sqrt(dxx * dxx + dyy * dyy + dzz * dzz);
    const double vsol = viss * are / dpb;
    const double upb = F->u[inbc], vpb = F->v[inbc], wpb = F->w[inbc];
    const double vnp = upb * nxf + vpb * nyf + wpb * nzf;
    const double utp = upb - vnp * nxf, vtp = vpb - vnp * nyf, wtp = wpb - vnp * nzf;
    F->ap[inbc] = F->ap[inbc] + vsol;
    spv[inbc] = spv[inbc] + vsol;
    F->sp[inbc] = F->sp[inbc] + vsol;
    F->su[inbc] = F->su[inbc] + vsol * F->u[inbc] - cf * utp;
    F->sv[inbc] = F->sv[inbc] + vsol * F->v[inbc] - cf * vtp;
    F->sw[inbc] = F->sw[inbc] + vsol * F->w[inbc] - cf * wtp;
}

void orc_calcuvw(int ni, int nj, int nk, orc_flow *F, double *spv, const orc_drv_opts *o,
                 const orc_uvw_opts *q, orc_report *reps)
{
    const int nij = ni * nj, nimm = ni - 2, njmm = nj - 2, nkm = nk - 1;
    const size_t n = (size_t)nij * nk;
    drv_grad(ni, nj, nk, F, F->u, F->gradu);
    drv_grad(ni, nj, nk, F, F->v, F->gradv);
    drv_grad(ni, nj, nk, F, F->w, F->gradw);
    drv_grad(ni, nj, nk, F, F->pp, F->gradp);
    for (int k = 2; k <= nkm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
        const size_t c = IX(i, j, k);
        F->ap[c] = 0.0;
        spv[c] = 0.0;
        F->sp[c] = 0.0;
        F->su[c] = -F->gradp[3 * c] * F->vol[c];
        F->sv[c] = -F->gradp[3 * c + 1] * F->vol[c];
        F->sw[c] = -F->gradp[3 * c + 2] * F->vol[c];
        if (o->const_mflux) F->su[c] = F->su[c] + q->gradpcmf * F->vol[c];
        if (o->buoy) {
            double heat = 0.0;
            if (o->boussinesq) heat = o->beta * o->densit * (F->t[c] - o->tref) * F->vol[c];
            else heat = (o->densit - F->den[c]) * F->vol[c];
            F->su[c] = F->su[c] - o->gravx * heat;
            F->sv[c] = F->sv[c] - o->gravy * heat;
            F->sw[c] = F->sw[c] - o->gravz * heat;
        }
        if (o->bdf) {
            const double apotime = F->den[c] * F->vol[c] / o->timestep;
            const double sut = apotime * ((1.0 + o->btime) * F->uo[c] - 0.5 * o->btime * F->uoo[c]);
            const double svt = apotime * ((1.0 + o->btime) * F->vo[c] - 0.5 * o->btime * F->voo[c]);
            const double swt = apotime * ((1.0 + o->btime) * F->wo[c] - 0.5 * o->btime * F->woo[c]);
            F->su[c] = F->su[c] + sut;
            F->sv[c] = F->sv[c] + svt;
            F->sw[c] = F->sw[c] + swt;
            F->ap[c] = F->ap[c] + apotime * (1.0 + 0.5 * o->btime);
            spv[c] = spv[c] + apotime * (1.0 + 0.5 * o->btime);
            F->sp[c] = F->sp[c] + apotime * (1.0 + 0.5 * o->btime);
        }
    }
    orc_fields f = drv_fields(F);
    orc_fluxuvw(ni, nj, nk, &f, F->vis, nj, 1, nij, F->ar1x, F->ar1y, F->ar1z, F->fx, F->ae, F->aw,
                F->f1, F->su, F->sv, F->sw, F->bp, F->bt, F->bb, q->schemes, q->gds);
    orc_fluxuvw(ni, nj, nk, &f, F->vis, 1, nij, nj, F->ar2x, F->ar2y, F->ar2z, F->fy, F->an, F->as,
                F->f2, F->su, F->sv, F->sw, F->bp, F->bt, F->bb, q->schemes, q->gds);
    orc_fluxuvw(ni, nj, nk, &f, F->vis, nij, nj, 1, F->ar3x, F->ar3y, F->ar3z, F->fz, F->at, F->ab,
                F->f3, F->su, F->sv, F->sw, F->bp, F->bt, F->bb, q->schemes, q->gds);
    const double dnx = q->nx, dny = q->ny, dnz = q->nz;
    for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
        const size_t inbc = IX(i, j, 3), ingc = inbc - nij;
        uvw_wall(F, spv, inbc, ingc, ingc, F->fz[ingc], dnx, dny, dnz);
    }
    /* the top wall's cells are taken on the plane lk(nimm), as written (or
     * on q->top_k when set) */
    const int ktop = q->top_k ? q->top_k : nimm;
    for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
        const size_t inbc = IX(i, j, ktop), ingc = inbc + nij;
        uvw_wall(F, spv, inbc, ingc, inbc, 1.0 - F->fz[inbc], dnx, dny, dnz);
    }
    if (o->cn)
        for (size_t c = 0; c < n; ++c) {
            F->ae[c] = 0.5 * F->ae[c];
            F->an[c] = 0.5 * F->an[c];
            F->at[c] = 0.5 * F->at[c];
            F->aw[c] = 0.5 * F->aw[c];
            F->as[c] = 0.5 * F->as[c];
            F->ab[c] = 0.5 * F->ab[c];
        }
    double *X[3] = {F->u, F->v, F->w}, *S[3] = {F->su, F->sv, F->sw}, *D[3] = {F->ap, spv, F->sp};
    const double *XO[3] = {F->uo, F->vo, F->wo};
    double *AP[3] = {F->apu, F->apv, F->apw};
    orc_sys s = drv_sys(F);
    for (int m = 0; m < 3; ++m) {
        for (int k = 2; k <= nkm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j) {
            const size_t c = IX(i, j, k);
            if (o->cn) {
                const double *xo = XO[m];
                const double apotime = F->den[c] * F->vol[c] / o->timestep;
                S[m][c] = S[m][c] + (F->ae[c] * xo[c + nj] + F->aw[c] * xo[c - nj] +
                                     F->an[c] * xo[c + 1] + F->as[c] * xo[c - 1] +
                                     F->at[c] * xo[c + nij] + F->ab[c] * xo[c - nij]) +
                          (apotime - F->ae[c] - F->aw[c] - F->an[c] - F->as[c] - F->at[c] -
                           F->ab[c]) * xo[c];
                D[m][c] = D[m][c] + apotime;
            }
            F->ap[c] = F->ae[c] + F->aw[c] + F->an[c] + F->as[c] + F->at[c] + F->ab[c] + D[m][c];
            F->ap[c] = F->ap[c] * q->urfr[m];
            F->su[c] = S[m][c] + q->urfm[m] * F->ap[c] * X[m][c];
            AP[m][c] = 1.0 / (F->ap[c] + SMALL);
        }
        orc_sipsol(ni, nj, nk, X[m], &s, o->alfa, q->sor[m], q->nsw[m], &reps[m]);
    }
}

/* ------------------------------------------------------------ geometry -- */
/*
 * Face area vectors from the node coordinates (init.f90:49-144).  Node
 * ijk is the (i, j, k) corner that closes cell ijk.  Per direction the face
 * of cell P is spanned by its corner P, the corner one step back along idns
 * ("south") and one back along idtb ("bottom"):
 *   east  (idns 1,  idtb nij): k = 2..nkm, i = 1..nim, j = 2..njm  -> ar1
 *   north (idns nj, idtb nij): k = 2..nkm, i = 2..nim, j = 1..njm  -> ar2
 *   top   (idns 1,  idtb nj ): k = 1..nkm, i = 2..nim, j = 2..njm  -> ar3
 * The vector is the cross product of the face's two mean edge vectors (the
 * "et" and "zd" directions).  Cells outside these ranges keep their values.
 */
static void orc_face_dir(int ni, int nj, int nk, const double *x, const double *y,
                         const double *z, int idns, int idtb, int klo, int ilo, int jlo,
                         double *arx, double *ary, double *arz)
{
    const int nij = ni * nj, nim = ni - 1, njm = nj - 1, nkm = nk - 1;
    for (int k = klo; k <= nkm; ++k)
        for (int i = ilo; i <= nim; ++i)
            for (int j = jlo; j <= njm; ++j) {
                const size_t p = IX(i, j, k), s = p - idns, b = p - idtb, bs = b - idns;
                const double xe = 0.5 * (x[p] - x[s] + x[b] - x[bs]);
                const double ye = 0.5 * (y[p] - y[s] + y[b] - y[bs]);
                const double ze = 0.5 * (z[p] - z[s] + z[b] - z[bs]);
                const double xd = 0.5 * (x[p] - x[b] + x[s] - x[bs]);
                const double yd = 0.5 * (y[p] - y[b] + y[s] - y[bs]);
                const double zd = 0.5 * (z[p] - z[b] + z[s] - z[bs]);
                arx[p] = ye * zd - yd * ze;
                ary[p] = xd * ze - xe * zd;
                arz[p] = xe * yd - ye * xd;
            }
}

void orc_face_areas(int ni, int nj, int nk, const double *x, const double *y, const double *z,
                    double *const *ar)
{
    const int nij = ni * nj;
    orc_face_dir(ni, nj, nk, x, y, z, 1, nij, 2, 1, 2, ar[0], ar[1], ar[2]);
    orc_face_dir(ni, nj, nk, x, y, z, nj, nij, 2, 2, 1, ar[3], ar[4], ar[5]);
    orc_face_dir(ni, nj, nk, x, y, z, 1, nj, 1, 2, 2, ar[6], ar[7], ar[8]);
}

/*
 * Intersection of the plane through p1, p2, p3 with the line through p4,
 * p5 (find_intersection_point.f90:3-76): p4 + (p5 - p4) t with
 * t = -det[1 1 1 1; p1 p2 p3 p4] / det[1 1 1 0; p1 p2 p3 p5-p4], both
 * determinants expanded along the first row into the reference's sum order.
 */
void orc_find_intersection_point(const double *p1, const double *p2, const double *p3,
                                  const double *p4, const double *p5, double *pj)
{
    const double x1 = p1[0], y1 = p1[1], z1 = p1[2], x2 = p2[0], y2 = p2[1], z2 = p2[2];
    const double x3 = p3[0], y3 = p3[1], z3 = p3[2], x4 = p4[0], y4 = p4[1], z4 = p4[2];
    const double dx = p5[0] - x4, dy = p5[1] - y4, dz = p5[2] - z4;
    /* 2x2 minors of the y, z rows */
    const double m34 = y3 * z4 - y4 * z3, m24 = y2 * z4 - y4 * z2, m14 = y1 * z4 - y4 * z1;
    const double m23 = y2 * z3 - y3 * z2, m13 = y1 * z3 - y3 * z1, m12 = y1 * z2 - y2 * z1;
    const double num = x2 * m34 - x1 * m34 - x3 * m24 + x1 * m24 + x3 * m14 - x2 * m14 + x4 * m23 -
                       x1 * m23 - x4 * m13 + x2 * m13 + x4 * m12 - x3 * m12;
    const double d3 = y3 * dz - dy * z3, d2 = y2 * dz - dy * z2, d1 = y1 * dz - dy * z1;
    const double den = x2 * d3 - x1 * d3 - x3 * d2 + x1 * d2 + x3 * d1 - x2 * d1 + dx * m23 -
                       dx * m13 + dx * m12;
    const double t = -num / den;
    pj[0] = x4 + dx * t;
    pj[1] = y4 + dy * t;
    pj[2] = z4 + dz * t;
}

/*
 * Interpolation factor of the face between P and E = P + idew (intfac.f90:
 * 3-51) on k = 2..nkm, i = 2..nim, j = 2..njm: the distance from P's centre
 * to where the line of centres P-E crosses the face plane (corner P, P - idns,
 * P - idtb), over the P-E distance.  `small` enters the z term of the
 * denominator's square and the denominator itself, as written.
 */
void orc_intfac(int ni, int nj, int nk, int idew, int idns, int idtb, const double *x,
                const double *y, const double *z, const double *xc, const double *yc,
                const double *zc, double *lambda)
{
    const int nij = ni * nj, nim = ni - 1, njm = nj - 1, nkm = nk - 1;
    const double small = (double)1e-30f;
    for (int k = 2; k <= nkm; ++k)
        for (int i = 2; i <= nim; ++i)
            for (int j = 2; j <= njm; ++j) {
                const size_t p = IX(i, j, k), e = p + idew;
                const double c1[3] = {x[p], y[p], z[p]};
                const double c2[3] = {x[p - idns], y[p - idns], z[p - idns]};
                const double c3[3] = {x[p - idtb], y[p - idtb], z[p - idtb]};
                const double c4[3] = {xc[p], yc[p], zc[p]};
                const double c5[3] = {xc[e], yc[e], zc[e]};
                double f[3];
                orc_find_intersection_point(c1, c2, c3, c4, c5, f);
                const double ax = f[0] - xc[p], ay = f[1] - yc[p], az = f[2] - zc[p];
                const double bx = xc[e] - xc[p], by = yc[e] - yc[p], bz = zc[e] - zc[p] + small;
                lambda[p] = sqrt(ax * ax + ay * ay + az * az) / (sqrt(bx * bx + by * by + bz * bz) + small);
            }
}

/* -------------------------------------------------- cgstab, pcg-jacobi -- */
/*
 * BiCGStab with the diagonal ILU(0) preconditioner (cgstab.f90:1-284):
 * d = 1/(ap - aw d(i-1) ae(i-1) - as d(j-1) an(j-1) - ab d(k-1) at(k-1)) on
 * k = 2..nkm (78-86); each application z = (r + aw z(i-1) + as z(j-1) +
 * ab z(k-1)) d, z /= (d + small), z = (z + ae z(i+1) + an z(j+1) +
 * at z(k+1)) d.  The second application's backward sweep covers only
 * k = 3..nkmm (212), so the planes k = 2 and nkm keep the scaled forward
 * values there.  The local arrays are zero outside the range (the reference
 * leaves them uninitialised); res keeps the caller's values outside it.
 */
void orc_cgstab(int ni, int nj, int nk, double *fi, orc_sys *s, double sor, int nsw,
                orc_report *rep)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkm = nk - 1, nkmm = nk - 2;
    const size_t n = (size_t)nij * nk;
    double *reso = calloc(n, sizeof(double)), *pk = calloc(n, sizeof(double)),
           *uk = calloc(n, sizeof(double)), *zk = calloc(n, sizeof(double)),
           *vk = calloc(n, sizeof(double)), *d = calloc(n, sizeof(double));
    double *res = s->res;
#define LOOP_F for (int k = 2; k <= nkm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j)
    double res0 = 0.0;
    psum res0_ps = PS0;
    LOOP_F {                                                   /* 42-53 */
        size_t c = IX(i, j, k);
        res[c] = resid(s, fi, c, nj, nij);
        res0 = ps_add(&res0_ps, fabs(res[c]), k);
    }
    LOOP_F {                                                   /* 58-66 */
        size_t c = IX(i, j, k);
        d[c] = 1. / (s->ap[c] - s->aw[c] * d[c - nj] * s->ae[c - nj] -
                     s->as[c] * d[c - 1] * s->an[c - 1] -
                     s->ab[c] * d[c - nij] * s->at[c - nij]);
    }
    LOOP_F {                                                   /* 71-83 */
        size_t c = IX(i, j, k);
        reso[c] = res[c];
    }
    double alf = 1.0, beto = 1.0, gam = 1.0, resl = 0.0, resor = 0.0;
    int l, conv = 0;
    /* z = M^-1 r with the backward sweep on k = kb_hi..kb_lo (downwards) */
#define ILU_APPLY(rhs, kb_hi, kb_lo)                                                         \
    do {                                                                                     \
        LOOP_F {                                                                             \
            size_t c = IX(i, j, k);                                                          \
            zk[c] = (rhs[c] + s->aw[c] * zk[c - nj] + s->as[c] * zk[c - 1] +                 \
                     s->ab[c] * zk[c - nij]) * d[c];                                         \
        }                                                                                    \
        LOOP_F {                                                                             \
            size_t c = IX(i, j, k);                                                          \
            zk[c] = zk[c] / (d[c] + SMALL);                                                  \
        }                                                                                    \
        for (int k = (kb_hi); k >= (kb_lo); --k)                                             \
            for (int i = nimm; i >= 3; --i)                                                  \
                for (int j = njmm; j >= 3; --j) {                                            \
                    size_t c = IX(i, j, k);                                                  \
                    zk[c] = (zk[c] + s->ae[c] * zk[c + nj] + s->an[c] * zk[c + 1] +          \
                             s->at[c] * zk[c + nij]) * d[c];                                 \
                }                                                                            \
    } while (0)
#define APPLY_A(out)                                                                         \
    LOOP_F {                                                                                 \
        size_t c = IX(i, j, k);                                                              \
        out[c] = s->ap[c] * zk[c] - s->ae[c] * zk[c + nj] - s->aw[c] * zk[c - nj] -          \
                 s->an[c] * zk[c + 1] - s->as[c] * zk[c - 1] - s->at[c] * zk[c + nij] -      \
                 s->ab[c] * zk[c - nij];                                                     \
    }
    for (l = 1; l <= nsw; ++l) {
        double bet = 0.0;
        psum bet_ps = PS0;                                      /* 94-104 */
        LOOP_F { size_t c = IX(i, j, k); bet = ps_add(&bet_ps, res[c] * reso[c], k); }
        double om = bet * gam / (alf * beto + SMALL);
        beto = bet;
        LOOP_F {                                               /* 108-115 */
            size_t c = IX(i, j, k);
            pk[c] = res[c] + om * (pk[c] - alf * uk[c]);
        }
        ILU_APPLY(pk, nkm, 2);                                 /* 119-148 */
        APPLY_A(uk);                                           /* 152-162 */
        double ukreso = 0.0;
        psum ukreso_ps = PS0;                                   /* 166-175 */
        LOOP_F { size_t c = IX(i, j, k); ukreso = ps_add(&ukreso_ps, uk[c] * reso[c], k); }
        gam = bet / ukreso;
        LOOP_F {                                               /* 179-187 */
            size_t c = IX(i, j, k);
            fi[c] = fi[c] + gam * zk[c];
            res[c] = res[c] - gam * uk[c];
        }
        ILU_APPLY(res, nkmm, 3);                               /* 191-220: the backward sweep of 212 is k = nkmm..3 */
        APPLY_A(vk);                                           /* 224-234 */
        double vres = 0.0, vv = 0.0;
        psum vres_ps = PS0, vv_ps = PS0;                       /* 238-250 */
        LOOP_F {
            size_t c = IX(i, j, k);
            vres = ps_add(&vres_ps, vk[c] * res[c], k);
            vv = ps_add(&vv_ps, vk[c] * vk[c], k);
        }
        alf = vres / (vv + SMALL);
        resl = 0.0;
        psum resl_ps = PS0;                                    /* 254-264 */
        LOOP_F {
            size_t c = IX(i, j, k);
            fi[c] = fi[c] + alf * zk[c];
            res[c] = res[c] - alf * vk[c];
            resl = ps_add(&resl_ps, fabs(res[c]), k);
        }
        if (l == 1) resor = res0;                              /* 268 */
        double rsm = resl / (resor + SMALL);
        if (rsm < sor) { conv = 1; break; }                    /* 271 */
    }
#undef ILU_APPLY
#undef APPLY_A
#undef LOOP_F
    rep->resor = res0;
    rep->resl = resl;
    rep->iters = l;
    rep->converged = conv;
    free(reso); free(pk); free(uk); free(zk); free(vk); free(d);
}

/*
 * CG with the Jacobi (diagonal) preconditioner (pcg-jacobi.f90:3-155,
 * subroutine Diagonally_Preconditioned_CG): range k, i, j = 3..n?mm;
 * pk, zk and the whole of res start at zero (33-37); z = r / (ap + small)
 * (divided, 66); s0 = 1e20 as a REAL*4 literal.
 */
void orc_pcg_jacobi(int ni, int nj, int nk, double *fi, orc_sys *s, double sor, int nsw,
                    orc_report *rep)
{
    const int nij = ni * nj;
    const int nimm = ni - 2, njmm = nj - 2, nkmm = nk - 2;
    const size_t n = (size_t)nij * nk;
    double *pk = calloc(n, sizeof(double)), *zk = calloc(n, sizeof(double));
    double *res = s->res;
    memset(res, 0, n * sizeof(double));                        /* 41-48 */
#define LOOP_F for (int k = 3; k <= nkmm; ++k) for (int i = 3; i <= nimm; ++i) for (int j = 3; j <= njmm; ++j)
    double res0 = 0.0;
    psum res0_ps = PS0;
    LOOP_F {                                                   /* 52-64 */
        size_t c = IX(i, j, k);
        res[c] = resid(s, fi, c, nj, nij);
        res0 = ps_add(&res0_ps, fabs(res[c]), k);
    }
    double s0 = S0INIT, resl = 0.0, resor = 0.0;
    int l, conv = 0;
    for (l = 1; l <= nsw; ++l) {
        double sk = 0.0;
        psum sk_ps = PS0;
        LOOP_F {                                               /* 77-88 */
            size_t c = IX(i, j, k);
            zk[c] = res[c] / (s->ap[c] + SMALL);
            sk = ps_add(&sk_ps, res[c] * zk[c], k);
        }
        double bet = sk / s0;                                  /* 92 */
        LOOP_F {                                               /* 96-103 */
            size_t c = IX(i, j, k);
            pk[c] = zk[c] + bet * pk[c];
        }
        double pkapk = 0.0;
        psum pkapk_ps = PS0;
        LOOP_F {                                               /* 107-118 */
            size_t c = IX(i, j, k);
            zk[c] = s->ap[c] * pk[c] - s->ae[c] * pk[c + nj] - s->aw[c] * pk[c - nj] -
                    s->an[c] * pk[c + 1] - s->as[c] * pk[c - 1] - s->at[c] * pk[c + nij] -
                    s->ab[c] * pk[c - nij];
            pkapk = ps_add(&pkapk_ps, pk[c] * zk[c], k);
        }
        double alf = sk / pkapk;                               /* 120 */
        resl = 0.0;
        psum resl_ps = PS0;
        LOOP_F {                                               /* 124-134 */
            size_t c = IX(i, j, k);
            fi[c] = fi[c] + alf * pk[c];
            res[c] = res[c] - alf * zk[c];
            resl = ps_add(&resl_ps, fabs(res[c]), k);
        }
        s0 = sk;                                               /* 136 */
        if (l == 1) resor = res0;                              /* 140 */
        double rsm = resl / (resor + SMALL);
        if (rsm < sor) { conv = 1; break; }                    /* 143 */
    }
#undef LOOP_F
    rep->resor = res0;
    rep->resl = resl;
    rep->iters = l;
    rep->converged = conv;
    free(pk);
    free(zk);
}

/* ------------------------------------------------ periodic ghost cells -- */
/*
 * update_values_at_ghostcells.f90:3-148 (the channel's periodic variant):
 * ghost planes i = 2 / nim take the values of i = nimm / 3 (39-66), j = 2 /
 * njm those of j = njmm / 3 (68-94), both on k = 3..nkmm, for u, v, w, te,
 * ed, vis; walls at k = 2 and nkm (96-145): u, v, w, te = -phi_P f_P / f_E,
 * ed copied, vis = (vis_g - vis_P f_P) / f_E with the face's interpolation
 * factor fz.  var = {u, v, w, te, ed, t, vis}; t is not touched.
 */
void orc_ghostcells_periodic(int ni, int nj, int nk, double *const *var, const double *fz)
{
    const int nij = ni * nj, nim = ni - 1, njm = nj - 1, nimm = ni - 2, njmm = nj - 2;
    const int nkm = nk - 1, nkmm = nk - 2;
    double *u = var[0], *v = var[1], *w = var[2], *te = var[3], *ed = var[4], *vis = var[6];
    double *cp[6] = {u, v, w, te, ed, vis};
    for (int k = 3; k <= nkmm; ++k)
        for (int j = 3; j <= njmm; ++j) {
            const size_t gw = IX(2, j, k), p = IX(3, j, k), e = IX(nimm, j, k), ge = IX(nim, j, k);
            for (int q = 0; q < 6; ++q) cp[q][gw] = cp[q][e];
            for (int q = 0; q < 6; ++q) cp[q][ge] = cp[q][p];
        }
    for (int k = 3; k <= nkmm; ++k)
        for (int i = 3; i <= nimm; ++i) {
            const size_t gw = IX(i, 2, k), p = IX(i, 3, k), e = IX(i, njmm, k), ge = IX(i, njm, k);
            for (int q = 0; q < 6; ++q) cp[q][gw] = cp[q][e];
            for (int q = 0; q < 6; ++q) cp[q][ge] = cp[q][p];
        }
    for (int i = 3; i <= nimm; ++i)
        for (int j = 3; j <= njmm; ++j) {
            const size_t g = IX(i, j, 2), p = IX(i, j, 3);
            const double fxp = fz[g], fxe = 1.0 - fxp;
            u[g] = -u[p] * fxp / fxe;
            v[g] = -v[p] * fxp / fxe;
            w[g] = -w[p] * fxp / fxe;
            te[g] = -te[p] * fxp / fxe;
            ed[g] = ed[p];
            vis[g] = (vis[g] - vis[p] * fxp) / fxe;
        }
    for (int i = 3; i <= nimm; ++i)
        for (int j = 3; j <= njmm; ++j) {
            const size_t g = IX(i, j, nkm), p = IX(i, j, nkmm);
            const double fxe = fz[p], fxp = 1.0 - fxe;
            u[g] = -u[p] * fxp / fxe;
            v[g] = -v[p] * fxp / fxe;
            w[g] = -w[p] * fxp / fxe;
            te[g] = -te[p] * fxp / fxe;
            ed[g] = ed[p];
            vis[g] = (vis[g] - vis[p] * fxp) / fxe;
        }
}
