/*
 * cap_oracle.h -- CPU oracle for the pressure-correction (p') hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links this; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and
 * only as the checker (or the timed CPU baseline), never as the thing shipped.
 *
 * A single-threaded C restatement of the reference Fortran routines, with the
 * reference's loop ranges, loop order, operand order and single-precision
 * literals, on the reference's padded j-fastest layout:
 *   ijk = lk(k) + li(i) + j, li(i) = (i-1)*nj, lk(k) = (k-1)*ni*nj   (1-based)
 *   C offset = ijk - 1.
 * Each function names the reference file:line it follows.
 */
#ifndef CAP_ORACLE_H
#define CAP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Module `coef` + `coefb` arrays the reference solvers read and write
 * (modules_allocatable.f90:161-180).  Every pointer is a full padded array of
 * ni*nj*nk doubles. */
typedef struct {
    double *ae, *aw, *an, *as, *at, *ab, *ap, *su, *res;
    double *bb, *bs, *bw, *bp, *bn, *be, *bt;
} orc_sys;

typedef struct {
    double resor;   /* initial L1 residual (reference `resor(ifi)`)          */
    double resl;    /* last L1 residual the reference prints as "Final"      */
    int iters;      /* the loop index the reference prints ("No Iterations") */
    int converged;  /* 1 if the loop exited on the relative-residual test    */
} orc_report;

/* counter-based generator shared bit-for-bit with the product's generator */
uint64_t orc_hash(uint64_t seed, uint64_t stream, uint64_t idx);
double orc_u01(uint64_t seed, uint64_t stream, uint64_t idx);
double orc_normal(uint64_t seed, uint64_t stream, uint64_t idx);
/* the generator's libm-free log, cos(2 pi u) and exp (bit-reproducible) */
double orc_gen_log(double x);
double orc_gen_cos2pi(double u);
double orc_gen_exp(double x);

/* problem builders (synthetic inputs of the measured workloads) */
void orc_build_lognormal(int ni, int nj, int nk, uint64_t seed, double sigma,
                         orc_sys *s);
void orc_build_general(int ni, int nj, int nk, uint64_t seed, orc_sys *s,
                       double *fi);

/* solvers */
void orc_sipsol(int ni, int nj, int nk, double *fi, orc_sys *s, double alfa,
                double sor, int nsw, orc_report *rep);
void orc_pcgsip(int ni, int nj, int nk, double *fi, orc_sys *s, double alfa,
                double sor, int nsw, orc_report *rep);
void orc_iccg(int ni, int nj, int nk, double *fi, orc_sys *s, double sor,
              int nsw, orc_report *rep);
void orc_cgstab_sip(int ni, int nj, int nk, double *fi, orc_sys *s,
                    double alfa, double sor, int nsw, orc_report *rep);

/* pieces used by the parity tests to pin single steps */
void orc_sip_factor(int ni, int nj, int nk, int k_lo, int k_hi, orc_sys *s,
                    double alfa);
void orc_sip_apply(int ni, int nj, int nk, int k_lo, int k_hi,
                   const orc_sys *s, const double *rhs, double *z);
void orc_spmv(int ni, int nj, int nk, int k_lo, int k_hi, const orc_sys *s,
              const double *p, double *z);

/* stencils */
void orc_grad_gauss(int ni, int nj, int nk, const double *u, const double *fx,
                    const double *fy, const double *fz, const double *ar1x,
                    const double *ar1y, const double *ar1z, const double *ar2x,
                    const double *ar2y, const double *ar2z, const double *ar3x,
                    const double *ar3y, const double *ar3z, const double *vol,
                    double *dudx, double *dudy, double *dudz);
void orc_bpres(int ni, int nj, int nk, double *p, const double *fx,
               const double *fy, const double *fz);

/* 1: scalar products / norms summed per k-plane (tests measure the solvers'
 * sensitivity to summation order with it); 0: the reference's order */
void orc_set_plane_sums(int on);

/* timing helper for the CPU baseline: n PCG-SIP iterations from fi */
double orc_pcgsip_time_iters(int ni, int nj, int nk, double *fi, orc_sys *s,
                             double alfa, int n_iters);

#ifdef __cplusplus
}
#endif
#endif

#ifdef __cplusplus
extern "C" {
#endif
/* Fields and geometry the p' assembly reads (modules geometry, variables,
 * coef, gradients of modules_allocatable.f90).  Gradients are the reference's
 * array-of-structures gradu(3, nxyz): component fastest. */
typedef struct {
    const double *x, *y, *z, *xc, *yc, *zc, *vol;
    const double *fx, *fy, *fz;
    const double *ar1x, *ar1y, *ar1z, *ar2x, *ar2y, *ar2z, *ar3x, *ar3y, *ar3z;
    const double *den, *apu, *apv, *apw, *u, *v, *w, *p;
    const double *gradu, *gradv, *gradw, *gradp;
} orc_fields;

void orc_fluxmass(int ni, int nj, int nk, const orc_fields *f, int idew, int idns, int idtb,
                  const double *arvx, const double *arvy, const double *arvz,
                  const double *fif, double *acfe, double *acfw, double *fcf);
void orc_piso_assemble(int ni, int nj, int nk, const orc_fields *f, int idew, int idns,
                       int idtb, const double *arvx, const double *arvy, const double *arvz,
                       const double *fif, double *acfe, double *acfw, double *fcf, int ipc,
                       double *bp, double *su);
void orc_assemble_p(int ni, int nj, int nk, const orc_fields *f, int idew, int idns, int idtb,
                    const double *arvx, const double *arvy, const double *arvz,
                    const double *fif, double *acfe, double *acfw, double *fcf,
                    const double *su, const double *sv, const double *sw);
void orc_fluxmc_all(int ni, int nj, int nk, const orc_fields *f, int idew, int idns, int idtb,
                    const double *fif, double *fmcor);
void orc_continuity(int ni, int nj, int nk, const double *f1, const double *f2,
                    const double *f3, double *sums);

/* scheme flags of orc_fluxuvw: the reference's logicals lcds, lcds4, lluds,
 * lquds, flux_limiter and the limiter choice lsmart, lavl, lmuscl, lumist,
 * lgamma (module parameters of modules_allocatable.f90) */
#define ORC_UVW_CDS     1
#define ORC_UVW_CDS4    2
#define ORC_UVW_LUDS    4
#define ORC_UVW_QUICK   8
#define ORC_UVW_LIMITER 16
#define ORC_UVW_SMART   32
#define ORC_UVW_AVL     64
#define ORC_UVW_MUSCL   128
#define ORC_UVW_UMIST   256
#define ORC_UVW_GAMMA   512
void orc_fluxuvw(int ni, int nj, int nk, const orc_fields *f, const double *vis, int idew,
                 int idns, int idtb, const double *arvx, const double *arvy, const double *arvz,
                 const double *fif, double *acfe, double *acfw, const double *fcf, double *su,
                 double *sv, double *sw, double *bp, double *bt, double *bb, int flags,
                 double gds);
void orc_courant(int ni, int nj, int nk, const double *f1, const double *f2, const double *f3,
                 const double *vol, double *su, double *timestep, double *time, int conumfix,
                 double conumfixvalue, double *out);
void orc_ghostcells(int ni, int nj, int nk, double *const *var, const double *const *geo,
                    const int *lbw, const int *lbe, const int *lbs, const int *lbn,
                    const int *lbb, const int *lbt, const double *inl);
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
/* Module state of the pressure-correction drivers (modules geometry,
 * variables, coef, coefb, hcoef, gradients of modules_allocatable.f90).
 * Gradients are gradu(3, nxyz), component fastest.  uo..woo and t are read
 * only by PISO_getHbyA's optional terms; t..wtt are copied by outbc (NULL:
 * not copied). */
typedef struct {
    const double *x, *y, *z, *xc, *yc, *zc, *vol, *fx, *fy, *fz;
    const double *ar1x, *ar1y, *ar1z, *ar2x, *ar2y, *ar2z, *ar3x, *ar3y, *ar3z;
    const double *den;
    double *u, *v, *w, *p, *pp, *f1, *f2, *f3;
    double *ae, *aw, *an, *as, *at, *ab, *ap, *su, *sv, *sw, *sp, *apu, *apv, *apw;
    double *res, *bb, *bs, *bw, *bp, *bn, *be, *bt;
    double *he, *hw, *hn, *hs, *ht, *hb;
    double *gradu, *gradv, *gradw, *gradp;
    const double *uo, *vo, *wo, *uoo, *voo, *woo;
    double *t, *te, *ed, *vis, *vart, *con, *utt, *vtt, *wtt;
} orc_flow;

/* Settings the drivers read from module parameters / boundc / time_mod.
 * solver: calcp's isol (1 sipsol, 3 iccg, 4 pcgsip, 5 cgstab_sip). */
typedef struct {
    int ncorr, npcor, solver, nsw;
    double urf, sor, alfa;
    int ijkpr;                       /* 1-based reference-pressure cell */
    int const_mflux;                 /* outbc skipped when set */
    int lout, low, loe, los, lon, lob, lot;
    double flowin;
    const int *lbw, *lbe, *lbs, *lbn, *lbb, *lbt;
    int cn, bdf, buoy, boussinesq;
    double timestep, btime, densit, beta, tref, gravx, gravy, gravz;
    double cumulative;               /* cumulativeContErr on entry */
} orc_drv_opts;

void orc_outbc(int ni, int nj, int nk, orc_flow *F, const orc_drv_opts *o);
void orc_correct_bc(int ni, int nj, int nk, orc_flow *F);
void orc_gethbya(int ni, int nj, int nk, orc_flow *F, const orc_drv_opts *o);
/* reps: ncorr * npcor solver reports; cont: per corrector {sumLocalContErr,
 * globalContErr, cumulativeContErr} */
void orc_calcp(int ni, int nj, int nk, orc_flow *F, const orc_drv_opts *o, orc_report *reps,
               double *cont);
/* pimple = 0: PISO_multiple_correction, 1: PIMPLE_multiple_correction */
void orc_piso(int ni, int nj, int nk, orc_flow *F, const orc_drv_opts *o, int pimple,
              orc_report *reps, double *cont);
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
/* find_strain_rate.f90:21-79 after its gradients: strain and vorticity
 * magnitudes on k = 2..nkm, i, j = 3..n?mm from gradu/gradv/gradw(3, nxyz) */
void orc_strain_rate(int ni, int nj, int nk, const double *gradu, const double *gradv,
                     const double *gradw, double *strain, double *vorticity);
/* calc_vis_les.f90:45-151 (Smagorinsky with van Driest damping, then the
 * wall-function viscosity of the bottom and top wall ghost planes).  nx, ny,
 * nz are the module integers the wall distance is (as written) projected on. */
void orc_calc_vis_les(int ni, int nj, int nk, const double *u, const double *v, const double *w,
                      const double *vol, const double *den, const double *fz, const double *xc,
                      const double *yc, const double *zc, const double *ar3x, const double *ar3y,
                      const double *ar3z, const double *strain, const double *walldist,
                      double *vis, double viscos, double densit, double cappa, double elog,
                      double urfvis, int nx, int ny, int nz, int top_k);

/* calcuvw settings beyond the driver settings */
typedef struct {
    int schemes;                 /* fluxuvw scheme flags (ORC_UVW_*) */
    double gds;                  /* gds(iu) */
    double urfr[3], urfm[3];     /* urfr(iu..iw), urfm(iu..iw) */
    double sor[3];
    int nsw[3];
    double gradpcmf;
    int nx, ny, nz;              /* module integers of the wall loops */
    int top_k;                   /* top wall's plane: 0 = lk(nimm) as written */
} orc_uvw_opts;

/* calcuvw.f90:3-464 (Gauss gradients, isol = 1): momentum sources, fluxuvw
 * for the three directions, the bottom / top wall terms, Crank-Nicolson
 * halving, then per component under-relaxed assembly, apu/apv/apw and
 * sipsol.  spv is the module array of the v equation's diagonal source.
 * calcstress (Reynolds stresses of the EVM / ASM models, read by neither
 * fluxuvw nor the assembly) is not part of it. */
void orc_calcuvw(int ni, int nj, int nk, orc_flow *F, double *spv, const orc_drv_opts *o,
                 const orc_uvw_opts *q, orc_report *reps);
/* geometry: face areas (init.f90:49-144) into ar[9] = {ar1x, ar1y, ar1z,
 * ar2x, .., ar3z}; the plane/line intersection of find_intersection_point.f90
 * and the interpolation factors of intfac.f90 */
void orc_face_areas(int ni, int nj, int nk, const double *x, const double *y, const double *z,
                    double *const *ar);
void orc_find_intersection_point(const double *p1, const double *p2, const double *p3,
                                  const double *p4, const double *p5, double *pj);
void orc_intfac(int ni, int nj, int nk, int idew, int idns, int idtb, const double *x,
                const double *y, const double *z, const double *xc, const double *yc,
                const double *zc, double *lambda);
/* cgstab.f90 (BiCGStab, diagonal ILU(0); k = 2..nkm) and pcg-jacobi.f90
 * (Diagonally_Preconditioned_CG; k = 3..nkmm) */
void orc_cgstab(int ni, int nj, int nk, double *fi, orc_sys *s, double sor, int nsw,
                orc_report *rep);
void orc_pcg_jacobi(int ni, int nj, int nk, double *fi, orc_sys *s, double sor, int nsw,
                    orc_report *rep);
/* update_values_at_ghostcells.f90 (periodic i and j, walls at k); var =
 * {u, v, w, te, ed, t, vis}, t untouched */
void orc_ghostcells_periodic(int ni, int nj, int nk, double *const *var, const double *fz);
#ifdef __cplusplus
}
#endif
