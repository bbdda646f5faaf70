/*
 * cap_amd.h -- C ABI of the MI355X pressure-correction (p') library.
 *
 * Drop-in boundary for the reference's linear solvers and p' stencils.  The
 * reference calls Fortran external procedures with implicit interfaces
 * (`call pcgsip(fi, ifi)` etc.) that read and write module arrays.  Here the
 * module arrays are passed explicitly as plain pointers in the reference's
 * padded j-fastest layout (ijk = (k-1)*ni*nj + (i-1)*nj + j, 1-based;
 * C offset = ijk-1; ni,nj,nk = interior cells + 4).  The Fortran side binds
 * these with ISO_C_BINDING (cappuccino-ghostcells-mpi_amd/fortran/cap_hip.f90).
 *
 * Pointers are host pointers unless the context was created with
 * CAP_DEVICE_POINTERS, in which case they are device (HBM) pointers and no
 * host<->device copy is made.  Every entry point returns a status code; no
 * exception crosses the ABI.  One context drives one HIP stream; calls are
 * synchronous on return.
 */
#ifndef CAP_AMD_H
#define CAP_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CAP_OK              0
#define CAP_MAXITER         1   /* nsw iterations done without reaching sor */
#define CAP_NAN             2   /* a residual norm became NaN or Inf       */
#define CAP_GUESS_REJECTED  3   /* learned guess raised the residual, unused */
#define CAP_ERR_ARG        -1
#define CAP_ERR_HIP        -2
#define CAP_ERR_TIMEOUT    -3   /* a wavefront dependency wait expired; the
                                   context's reduction and claim state is
                                   reset before the call returns, so the
                                   context stays usable                     */
#define CAP_ERR_NOMEM      -4
#define CAP_ERR_NODEV      -5

/* context flags */
#define CAP_DEVICE_POINTERS 1

typedef struct cap_ctx cap_ctx;

/*
 * Module `coef` and `coefb` arrays (modules_allocatable.f90:161-180).
 * ae..ap and su are read.  res and bb..bt are the reference's module work
 * arrays: may be NULL (their halo values are then taken as zero and nothing
 * is written back); when given, their values outside the solve range enter
 * the recursions exactly as in the reference (e.g. sipsol.f90:70-71 reads
 * res on the boundary planes, pcg-sip.f90:72-81 reads bn/be/bt there) and on
 * return they hold what the reference leaves in them.
 */
typedef struct {
    double *ae, *aw, *an, *as, *at, *ab, *ap, *su, *res;
    double *bb, *bs, *bw, *bp, *bn, *be, *bt;
} cap_coef;

typedef struct {
    double resor;   /* initial L1 residual, the reference's resor(ifi)      */
    double resl;    /* final L1 residual the reference reports              */
    int iters;      /* the reference's reported iteration number: the index
                       of the converging iteration, or nsw+1 when it ran out */
    int status;     /* CAP_OK / CAP_MAXITER / CAP_NAN                        */
    double ms;      /* device time of the solve, milliseconds               */
} cap_report;

/* Context for arrays of padded extents ni x nj x nk. */
int cap_init(cap_ctx **ctx, int ni, int nj, int nk, int flags);
int cap_finalize(cap_ctx *ctx);
const char *cap_last_error(const cap_ctx *ctx);

/* Replaces `call sipsol(fi, ifi)` (sipsol.f90:3-106): Stone's SIP,
 * k,i,j = 3..n?mm, stop when RES1/RESOR < sor. */
int cap_sipsol(cap_ctx *ctx, double *fi, const cap_coef *c, double alfa,
               double sor, int nsw, cap_report *rep);

/* Replaces `call pcgsip(fi, ifi)` (pcg-sip.f90:3-200): CG with the SIP
 * preconditioner, k,i,j = 3..n?mm, stop when resl/res0 < sor. */
int cap_pcgsip(cap_ctx *ctx, double *fi, const cap_coef *c, double alfa,
               double sor, int nsw, cap_report *rep);

/* Replaces `call iccg(fi, ifi)` (iccg.f90:3-197): CG with IC(0),
 * k = 2..nkm, i,j = 3..n?mm. */
int cap_iccg(cap_ctx *ctx, double *fi, const cap_coef *c, double sor, int nsw,
             cap_report *rep);

/* Replaces `call cgstab_sip(fi, ifi)` (cgstab_sip.f90:1-278): BiCGStab with
 * the SIP preconditioner, k = 2..nkm, i,j = 3..n?mm. */
int cap_cgstab_sip(cap_ctx *ctx, double *fi, const cap_coef *c, double alfa,
                   double sor, int nsw, cap_report *rep);

/* Replaces `call cgstab(fi, ifi)` (cgstab.f90:1-284): BiCGStab with the
 * diagonal ILU(0) preconditioner (d of cgstab.f90:58-66), k = 2..nkm,
 * i,j = 3..n?mm; its second backward sweep covers k = 3..nkmm only, as
 * written (cgstab.f90:212).  Reached from calcp with isol = 2
 * (calcp-multiple_correction_SIMPLE.f90:99).  bb..bt are not used. */
int cap_cgstab(cap_ctx *ctx, double *fi, const cap_coef *c, double sor, int nsw,
               cap_report *rep);

/* Replaces `call Diagonally_Preconditioned_CG(fi, ifi)` (pcg-jacobi.f90:
 * 3-155): CG with the Jacobi preconditioner z = r / (ap + small),
 * k,i,j = 3..n?mm; res is zero outside the range on return. */
int cap_pcg_jacobi(cap_ctx *ctx, double *fi, const cap_coef *c, double sor, int nsw,
                   cap_report *rep);

/* No counterpart in the reference (its solvers are FP64 only).  Solves the
 * pcgsip system (same arrays, same range, same L1 stop test against the FP64
 * residual of the starting fi) by FP64 iterative refinement around an FP32
 * PCG with an FP32 SIP preconditioner and FP32 SpMV; the true residual, the
 * dot products and the norms are FP64.  inner_rtol is the L1 reduction the
 * FP32 inner solve is asked for per refinement step (half the reduction the
 * stop test still needs, when that is less).  A system too ill-conditioned
 * for FP32 (a step that leaves more than 3/4 of the residual it started
 * from, or whose inner solve runs out of its 4 (ni + nj + nk) iterations) is
 * finished in FP64 by cap_pcgsip from the current fi, to the same absolute
 * residual.  rep->iters counts FP32 inner iterations plus those FP64
 * ones; nsw bounds their sum. */
int cap_pcgsip_mixed(cap_ctx *ctx, double *fi, const cap_coef *c, double alfa,
                     double sor, int nsw, double inner_rtol, cap_report *rep);

/* No counterpart in the reference: PCG-SIP with the preconditioner in FP32.
 * The SIP factors (alfa as in cap_pcgsip) are formed in FP64 and rounded to
 * FP32; the forward and backward sweeps run on an FP32 copy of the residual;
 * the SpMV, residual, solution, scalar products and the stop test (L1
 * residual relative to the start, as cap_pcgsip) stay FP64.  beta is the
 * flexible (Polak-Ribiere) form, which the varying rounding of an FP32
 * preconditioner needs.  The Krylov space is never restarted, so the
 * solution's error stays at the FP64 solve's.  After 4 (nis + njs + nks)
 * iterations without convergence (a system too ill-conditioned for the FP32
 * preconditioner) the solve goes on as cap_pcgsip from the current fi, to
 * the same absolute residual; the report counts both parts. */
int cap_pcgsip_fcg(cap_ctx *ctx, double *fi, const cap_coef *c, double alfa, double sor,
                   int nsw, cap_report *rep);

/* Replaces `call grad_gauss(u, dudx, dudy, dudz)` (GRAD_GAUSS.f90:1-105).
 * geom holds fx, fy, fz, ar1x, ar1y, ar1z, ar2x, ar2y, ar2z, ar3x, ar3y,
 * ar3z, vol in this order (module `geometry`). */
int cap_grad_gauss(cap_ctx *ctx, const double *u, const double *const *geom,
                   double *dudx, double *dudy, double *dudz);

/* Fields and geometry read by the p' assembly: modules geometry (x, y, z
 * nodes; xc, yc, zc centres; vol; fx, fy, fz; ar1*-ar3* face areas),
 * variables (den, u, v, w, p), coef (apu, apv, apw) and gradients (gradu,
 * gradv, gradw, gradp: the reference's gradu(3, nxyz), component fastest,
 * 3*nxyz doubles).  A pointer a routine does not read may be NULL. */
typedef struct {
    const double *x, *y, *z, *xc, *yc, *zc, *vol, *fx, *fy, *fz;
    const double *ar1x, *ar1y, *ar1z, *ar2x, *ar2y, *ar2z, *ar3x, *ar3y, *ar3z;
    const double *den, *apu, *apv, *apw, *u, *v, *w, *p;
    const double *gradu, *gradv, *gradw, *gradp;
} cap_fields;

/* Replaces `call fluxmass(idew,idns,idtb, arvx,arvy,arvz, fif, acfe,acfw,fcf)`
 * (fluxmass2.f90:3-195): p' face coefficients and Rhie-Chow mass flux on the
 * face loop k=2..nkm, i=2..nimm, j=2..njmm.  (idew,idns,idtb) is one of
 * (nj,1,nij), (1,nij,nj), (nij,nj,1) as at calcp-multiple_correction_SIMPLE.f90:50-60. */
int cap_fluxmass(cap_ctx *ctx, const cap_fields *f, int idew, int idns, int idtb,
                 const double *arvx, const double *arvy, const double *arvz,
                 const double *fif, double *acfe, double *acfw, double *fcf);

/* Replaces `call piso_assemble_pressure_eq(idew,idns,idtb, arvx,arvy,arvz,
 * fif, acfe,acfw,fcf, ipc, bp)` (PISO_assemble_pressure_eq.f90:3-260); su is
 * module coef's su, updated in place like the reference. */
int cap_piso_assemble_pressure_eq(cap_ctx *ctx, const cap_fields *f, int idew, int idns,
                                  int idtb, const double *arvx, const double *arvy,
                                  const double *arvz, const double *fif, double *acfe,
                                  double *acfw, double *fcf, int ipc, double *bp, double *su);

/* Replaces `call assemble_pressure_eq(...)` (assemble_pressure_eq.f90:3-124);
 * su, sv, sw are the face-velocity sources (HbyA) it reads from module coef. */
int cap_assemble_pressure_eq(cap_ctx *ctx, const cap_fields *f, int idew, int idns, int idtb,
                             const double *arvx, const double *arvy, const double *arvz,
                             const double *fif, double *acfe, double *acfw, double *fcf,
                             const double *su, const double *sv, const double *sw);

/* `call fluxmc(inp,idew,idns,idtb,fif,fmcor)` (fluxmc.f90:3-109) for every
 * cell k,i,j = 3..n?mm at once: fmcor(inp) for those cells, 0 elsewhere. */
int cap_fluxmc(cap_ctx *ctx, const cap_fields *f, int idew, int idns, int idtb,
               const double *fif, double *fmcor);

/* The drivers' three calls in one pass (no single counterpart in the
 * reference, which calls these routines once per direction: east with
 * (idew,idns,idtb) = (nj,1,nij) and fx, ar1*; north (1,nij,nj), fy, ar2*; top
 * (nij,nj,1), fz, ar3*).  Each leaves every array bit for bit as the three
 * calls do: cap_fluxmass3 = cap_fluxmass for the three directions
 * (calcp-multiple_correction_SIMPLE.f90:50-60), cap_piso_assemble_pressure_eq3
 * = cap_piso_assemble_pressure_eq for them with bp = be, bn, bt
 * (PISO_multiple_correction.f90).  Device pointers only (CAP_DEVICE_POINTERS
 * contexts); other contexts get CAP_ERR_ARG. */
int cap_fluxmass3(cap_ctx *ctx, const cap_fields *f, double *ae, double *aw, double *an,
                  double *as, double *at, double *ab, double *f1, double *f2, double *f3);
int cap_piso_assemble_pressure_eq3(cap_ctx *ctx, const cap_fields *f, double *ae, double *aw,
                                   double *an, double *as, double *at, double *ab, double *f1,
                                   double *f2, double *f3, int ipc, double *be, double *bn,
                                   double *bt, double *su);

/* ---- Pressure-correction drivers ---------------------------------------
 * Module state of calcp / PISO / PIMPLE (modules geometry, variables, coef,
 * coefb, hcoef, gradients of modules_allocatable.f90), all device (HBM)
 * pointers of a CAP_DEVICE_POINTERS context, reference layout.  Gradients are
 * gradu(3, nxyz), component fastest.  uo..woo are read only by
 * PISO_getHbyA's BDF / Crank-Nicolson terms and t by its buoyancy term; outbc
 * also copies t, te, ed, vis, vart, con, utt, vtt, wtt into the outlet cells
 * (NULL: not copied). */
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
} cap_flow;

/* Settings the drivers read from modules parameters, boundc and time_mod.
 * solver is calcp's isol: 1 sipsol, 2 cgstab, 3 iccg, 4 pcgsip, 5 cgstab_sip
 * (the reference's setting); PISO and PIMPLE use pcgsip.  solver 6 (no
 * counterpart in the reference) selects cap_pcgsip_mixed (inner_rtol 3e-4,
 * at most 2 nsw FP32 inner iterations) in all three drivers, solver 7 (no
 * counterpart either) cap_pcgsip_fcg.  lbw..lbt are
 * device arrays indexed as in cap_bc (outbc reads them when lout is set). */
typedef struct {
    int ncorr, npcor, solver, nsw;
    double urf, sor, alfa;
    int ijkpr;                       /* 1-based reference-pressure cell */
    int const_mflux;                 /* outbc is skipped when set */
    int lout, low, loe, los, lon, lob, lot;
    double flowin;
    const int *lbw, *lbe, *lbs, *lbn, *lbb, *lbt;
    int cn, bdf, buoy, boussinesq;
    double timestep, btime, densit, beta, tref, gravx, gravy, gravz;
    double cumulative;               /* cumulativeContErr on entry */
} cap_drv_opts;

/* `call calcp` (calcp-multiple_correction_SIMPLE.f90:3-238, Gauss
 * gradients): every step on the device.  reps: ncorr * npcor solver reports;
 * cont: per corrector {sumLocalContErr, globalContErr, cumulativeContErr}. */
int cap_calcp(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o, cap_report *reps, double *cont);
/* `call PISO_multiple_correction` (PISO_multiple_correction.f90:3-297, with
 * PISO_getHbyA.f90:3-233) for pimple = 0, `call PIMPLE_multiple_correction`
 * (PIMPLE_multiple_correction.f90:3-302) for pimple = 1.  Reports as cap_calcp. */
int cap_piso(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o, int pimple, cap_report *reps,
             double *cont);
/* `call outbc` (outbc.f90:3-318): outlet face fluxes, then velocities and
 * fluxes of the outlet cells scaled by flowin / flow.  Returns at once
 * unless o->lout. */
int cap_outbc(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o);
/* `call correctBoundaryConditions` (correctBoundaryConditions.f90:1-86):
 * periodic copies of u, v, w (and f1, f2 on the west / south planes). */
int cap_correct_bc(cap_ctx *ctx, cap_flow *f);
/* `call PISO_getHbyA` (PISO_getHbyA.f90:3-233). */
int cap_gethbya(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o);

/* calcuvw settings beyond cap_drv_opts (of which calcuvw reads const_mflux,
 * buoy, boussinesq, bdf, cn, timestep, btime, densit, beta, tref, grav*,
 * alfa): fluxuvw's scheme flags (CAP_UVW_*) and gds(iu); urfr, urfm, sor, nsw
 * of u, v, w; gradpcmf; the module integers nx, ny, nz of the wall loops.
 * top_k (no counterpart in the reference): 0 takes the top wall's cells on
 * the plane lk(nimm), as calcuvw.f90:254 writes it (the wall's first cells
 * only when ni = nk; then ni <= nk + 1 is required); 3..nk-2 takes them on
 * that plane instead (nk - 2: the wall's own first cells for any ni). */
typedef struct {
    int schemes;
    double gds;
    double urfr[3], urfm[3], sor[3];
    int nsw[3];
    double gradpcmf;
    int nx, ny, nz;
    int top_k;
} cap_uvw_opts;

/* `call calcuvw` (calcuvw.f90:3-464, Gauss gradients, isol = 1): gradients of
 * u, v, w, pp; momentum sources; fluxuvw for the three directions; bottom and
 * top wall shear (the top wall's cells on the plane lk(nimm), as written:
 * needs ni <= nk + 1); Crank-Nicolson halving; per component the
 * under-relaxed assembly, apu / apv / apw and sipsol.  spv is the module
 * array of the v equation's diagonal source (device).  reps[3]: the three
 * sipsol reports.  calcstress (Reynolds stresses of the EVM / ASM models,
 * read neither by fluxuvw nor by the assembly) is not part of it. */
int cap_calcuvw(cap_ctx *ctx, cap_flow *f, double *spv, const cap_drv_opts *o,
                const cap_uvw_opts *q, cap_report *reps);

/* `call find_strain_rate` (find_strain_rate.f90:1-79, Gauss gradients):
 * gradu, gradv, gradw of f, then strain and vorticity magnitudes on
 * k = 2..nkm, i, j = 3..n?mm (device arrays). */
int cap_find_strain_rate(cap_ctx *ctx, cap_flow *f, double *strain, double *vorticity);

typedef struct {
    double viscos, densit, cappa, elog, urfvis;   /* urfvis = urf(ivis) */
    int nx, ny, nz;                               /* module integers of the wall loops */
    int top_k;                                    /* as cap_uvw_opts.top_k */
} cap_les_opts;

/* `call calc_vis_les` (calc_vis_les.f90:3-151): Smagorinsky viscosity with
 * van Driest damping (C_Les = 0.17, C_delta = 0.158) into f->vis on
 * k = 2..nkm, i, j = 3..n?mm, then the wall-function viscosity of the bottom
 * and top wall ghost cells (top wall on the plane lk(nimm), as written:
 * needs ni <= nk + 1; or on o->top_k when that is set).  strain and walldist are device arrays.  pow, exp
 * and log are the device's: within a few ulp of the host's. */
int cap_calc_vis_les(cap_ctx *ctx, cap_flow *f, const double *strain, const double *walldist,
                     const cap_les_opts *o);

/* ---- Learned initial guess (no counterpart in the reference) ----------
 * A 3-D U-Net (4 levels, 16/32/48/64 channels, 3x3x3 convolutions, BF16
 * weights and activations, FP32 accumulation on the matrix cores) that maps
 * (su, ap, previous p') on the nx x ny x nz interior cells (x = the j index,
 * fastest) to the direction of the correction of p'.  Inputs of
 * cap_unet_forward / _gradients / _train_step are three fp32 channels,
 * channel-major, normalised by the caller; pointers follow the context's
 * CAP_DEVICE_POINTERS flag. */
typedef struct cap_unet cap_unet;
int cap_unet_create(cap_ctx *ctx, int nx, int ny, int nz, unsigned long long seed,
                    cap_unet **net);
int cap_unet_destroy(cap_unet *net);
int cap_unet_num_params(const cap_unet *net, long *count);
int cap_unet_get_params(cap_unet *net, float *params);
int cap_unet_set_params(cap_unet *net, const float *params);
int cap_unet_forward(cap_unet *net, const float *in, float *out);
/* loss = mean((out - target)^2); grads (may be NULL): d loss / d params */
int cap_unet_gradients(cap_unet *net, const float *in, const float *target, float *loss,
                       float *grads);
/* one Adam step (beta 0.9 / 0.999) on the same loss */
int cap_unet_train_step(cap_unet *net, const float *in, const float *target, float lr,
                        float *loss);
/* fi (device, reference layout): in the start (0 in calcp, p in PISO),
 * out the guess fi + beta (prev - fi) + alpha g on the solve range.  g is
 * the network's output for the channels (r / s, ap / a, prev / ps): prev the
 * previous-step p' (NULL: zero, and beta = 0), r = su - M prev its residual
 * (M x = ap x - sum a_nb x_nb; su - M fi when prev is NULL), s the RMS of r,
 * a the mean of ap and ps the RMS of prev (1 when zero).  (beta, alpha) is the least-squares fit of
 * the residual's 2-norm, so the guess never raises that norm above the
 * start's or prev's.  scales = {s, a, ps, alpha, beta} (may be NULL).  Needs
 * a CAP_DEVICE_POINTERS context whose interior is nx x ny x nz. */
int cap_unet_guess(cap_ctx *ctx, cap_unet *net, double *fi, const double *prev,
                   const cap_coef *c, double *scales);
/* PCG-SIP (as cap_pcgsip) from the network's guess when that guess lowers the
 * L1 residual of the start fi, otherwise from fi (then the status is
 * CAP_GUESS_REJECTED when the solve converged).  *accepted tells which.  The
 * solver stops on its L1 residual relative to its own start, so from an
 * accepted guess sor is rescaled by r0 / r_guess: the solve stops at the
 * absolute residual that the solve from fi stops at, in fewer iterations when
 * the guess is good. */
int cap_pcgsip_guess(cap_ctx *ctx, cap_unet *net, double *fi, const double *prev,
                     const cap_coef *c, double alfa, double sor, int nsw, cap_report *rep,
                     int *accepted);

/* No counterpart in the reference: cap_piso (PISO_multiple_correction.f90)
 * with the last corrector's p solve (PISO_multiple_correction.f90:141, where
 * the learned guess's training pairs arise) started from a guess instead of
 * the previous corrector's p.  The warm start w is prev (the previous time
 * step's p) on the solve range with the ghost values of the previous
 * corrector's p, which the solve holds fixed.  With net NULL the start is w;
 * with net it is cap_unet_guess's guess for the previous corrector's p and
 * w.  It is kept (*accepted = 1) only when it lowers the L1 residual of
 * the reference's start, and sor is then scaled by the ratio of the two, so
 * the solve stops at the absolute residual the reference's start stops at.
 * l1 (may be NULL) = {L1 residual of the reference's start, of the start
 * used}.  Device pointers only. */
int cap_piso_guess(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o, int pimple, cap_unet *net,
                   const double *prev, cap_report *reps, double *cont, int *accepted, double *l1);

/* The Courant-number block included in the time loop (CourantNo.h:10-81,
 * ltransient branch): su = sum of |face mass fluxes| f1..f3 of each cell on
 * k = 2..nkm, i, j = 3..n?mm (0 elsewhere), conum = {CoNum, meanCoNum}; with
 * conumfix the time step is rescaled so that CoNum = conumfixvalue.  The time
 * advances by the (new) time step.  Sums are compensated on the device. */
int cap_courant(cap_ctx *ctx, const double *f1, const double *f2, const double *f3,
                const double *vol, double *su, double *timestep, double *time, int conumfix,
                double conumfixvalue, double *conum);

/* No counterpart in the reference.  cap_courant without any device memory
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight time-step loops, harnesses that forbid hipMalloc): the same
 * kernels and results as cap_courant, with the per-block partials in a
 * static device array of the library (8192 x 5 + 8 doubles, 320 KiB per GPU,
 * held until the process ends) instead of a buffer the context allocates.
 * Needs a CAP_DEVICE_POINTERS context (CAP_ERR_ARG otherwise: host arrays
 * would be staged through allocated buffers).  All contexts on one GPU share
 * that GPU's array: calls that overlap on one GPU (from several host threads
 * or streams) race and must be serialised by the caller. */
int cap_courant_noalloc(cap_ctx *ctx, const double *f1, const double *f2, const double *f3,
                        const double *vol, double *su, double *timestep, double *time,
                        int conumfix, double conumfixvalue, double *conum);

/* Boundary types of the ghost-cell update, per boundary cell: lbw / lbe
 * indexed (j-1)*nk + k-1, lbs / lbn (i-1)*nk + k-1, lbb / lbt (i-1)*nj + j-1
 * (module bc of modules_allocatable.f90:217-219); 1 inlet, 2 outlet,
 * 3 symmetry, 4 wall.  Inlet values: module inlet (uin, vin, win, tein,
 * edin, tin). */
typedef struct {
    const int *lbw, *lbe, *lbs, *lbn, *lbb, *lbt;
    double uin, vin, win, tein, edin, tin;
} cap_bc;

/* Replaces `call update_values_at_ghostcells`
 * (update_values_at_ghostcells_nonperiodic.f90:3-642): ghost-cell values of
 * var = {u, v, w, te, ed, t, vis} next to the six boundary faces (edges and
 * corners untouched), geom = {fx, fy, fz, ar1x, ar1y, ar1z, ar2x, ar2y, ar2z,
 * ar3x, ar3y, ar3z}.  The on-device ghost fill of the path (one GPU: no
 * exchange between ranks). */
int cap_update_ghostcells(cap_ctx *ctx, double *const *var, const double *const *geom,
                          const cap_bc *bc);

/* Replaces the periodic `call update_values_at_ghostcells`
 * (update_values_at_ghostcells.f90:3-148, the channel's variant): the ghost
 * planes i = 2, nim and j = 2, njm take u, v, w, te, ed, vis of the opposite
 * interior planes (k = 3..nkmm); at the walls k = 2 and nkm, u, v, w, te =
 * -phi_P f_P / f_E, ed copied, vis = (vis_g - vis_P f_P) / f_E, with the
 * wall face's interpolation factor fz.  var = {u, v, w, te, ed, t, vis} as
 * cap_update_ghostcells takes it; t is not touched. */
int cap_update_ghostcells_periodic(cap_ctx *ctx, double *const *var, const double *fz);

/* Convection schemes of cap_fluxuvw: the reference's module logicals
 * lcds, lcds4, lluds, lquds, flux_limiter and the limiter choice lsmart,
 * lavl, lmuscl, lumist, lgamma (modules_allocatable.f90:58, 90), applied in
 * the reference's order (a later one overrides an earlier one). */
#define CAP_UVW_CDS      1
#define CAP_UVW_CDS4     2
#define CAP_UVW_LUDS     4
#define CAP_UVW_QUICK    8
#define CAP_UVW_LIMITER  16
#define CAP_UVW_SMART    32
#define CAP_UVW_AVL      64
#define CAP_UVW_MUSCL    128
#define CAP_UVW_UMIST    256
#define CAP_UVW_GAMMA    512
#define CAP_UVW_ALL      1023
/* No counterpart in the reference: with CAP_UVW_CDS alone, the face
 * arithmetic of fluxuvw in FP32 (geometry differences formed in FP64 and
 * rounded once; velocities, gradients, viscosity and mass flux rounded at
 * the load) and each face's contribution added to su, sv, sw in FP64. */
#define CAP_UVW_FP32     1024

/* Replaces `call fluxuvw(idew, idns, idtb, arvx, arvy, arvz, fif, acfe, acfw,
 * fcf)` (fluxuvw-overrelaxed-correction.f90:3-659, called for the three
 * directions at calcuvw.f90:143-149): UDS coefficients acfe(P) / acfw(E),
 * explicit minus implicit diffusion and the deferred correction
 * gds * (F_uds - F_high) added to su / sv / sw (module coef), the east cell's
 * share through the scratch bp / bt / bb, which are zero on return.  vis is
 * the module `variables` viscosity; gds the reference's gds(iu). */
int cap_fluxuvw(cap_ctx *ctx, const cap_fields *f, const double *vis, int idew, int idns,
                int idtb, const double *arvx, const double *arvy, const double *arvz,
                const double *fif, double *acfe, double *acfw, const double *fcf, double *su,
                double *sv, double *sw, double *bp, double *bt, double *bb, int schemes,
                double gds);

/* calcuvw.f90:100-140's three cap_fluxuvw calls (east, north, top, with f1,
 * f2, f3 and ae/aw, an/as, at/ab) in one pass, for the CDS scheme alone
 * (schemes must be CAP_UVW_CDS); every array ends bit for bit as after the
 * three calls, bp, bt, bb zeroed.  No single counterpart in the reference.
 * Device pointers only. */
int cap_fluxuvw3(cap_ctx *ctx, const cap_fields *f, const double *vis, double *ae, double *aw,
                 double *an, double *as, double *at, double *ab, const double *f1,
                 const double *f2, const double *f3, double *su, double *sv, double *sw,
                 double *bp, double *bt, double *bb, int schemes, double gds);

/* The continuity-error sums of continuityErrors.h:55-81 over k=2..nkm,
 * i,j=3..n?mm: sums[0] = global (signed), sums[1] = sum of |local|.  FP64,
 * deterministic reduction order. */
int cap_continuity(cap_ctx *ctx, const double *f1, const double *f2, const double *f3,
                   double *sums);

/* Replaces `call bpres(p)` (bpres.f90:3-81). */
int cap_bpres(cap_ctx *ctx, double *p, const double *fx, const double *fy,
              const double *fz);

/* ---- Grid geometry ------------------------------------------------------
 * Replaces the face-area loops of init.f90:49-144: from the nodes x, y, z
 * (node ijk closes cell ijk) the face area vectors ar[9] = {ar1x, ar1y, ar1z,
 * ar2x, .., ar3z} on the reference's loop ranges (east k = 2..nkm,
 * i = 1..nim, j = 2..njm; north k = 2..nkm, i = 2..nim, j = 1..njm; top
 * k = 1..nkm, i = 2..nim, j = 2..njm); other cells keep their values.  As
 * the reference's loops are written, on a right-handed grid the east vector
 * points to +i and the north and top vectors to -j and -k. */
int cap_face_areas(cap_ctx *ctx, const double *x, const double *y, const double *z,
                   double *const *ar);

/* Replaces `call intfac(nj,1,nij,fx); call intfac(1,nj,nij,fy);
 * call intfac(nij,1,nj,fz)` (init.f90:147-149; intfac.f90:3-51 with
 * find_intersection_point.f90:3-76): the interpolation factor of each face
 * on k = 2..nkm, i = 2..nim, j = 2..njm from nodes and centres
 * nodes_centres[6] = {x, y, z, xc, yc, zc}; other cells keep their values. */
int cap_intfac(cap_ctx *ctx, const double *const *nodes_centres, double *fx, double *fy,
               double *fz);

/* The grid file of read_grid.f90:126-161 (Fortran unformatted sequential
 * records: ni, nj, nk; nij, nik, njk, nijk, ijs, iks, jks, ijks, lbw, lbe, lbs,
 * lbn, lbb, lbt, low, loe, los, lon, lob, lot, lout, lthermbc, lconbc; then
 * x, y, z, xc, yc, zc, vol, fx, fy, fz, ar1x .. ar3z).  Host code: plain
 * host pointers, no context.  cap_grid_dims reads the extents; cap_read_grid
 * fills arrays[19] (in the file's order, ni*nj*nk doubles each; a NULL entry
 * is skipped), lb[6] = {lbw, lbe, lbs, lbn, lbb, lbt} (int32, may be NULL)
 * and flags[9] = {low, loe, los, lon, lob, lot, lout, lthermbc, lconbc}
 * (0 / 1, may be NULL).  CAP_ERR_ARG for a missing or malformed file or
 * extents that differ from the file's. */
int cap_grid_dims(const char *path, int *ni, int *nj, int *nk);
int cap_read_grid(const char *path, int ni, int nj, int nk, double *const *arrays, int *const *lb,
                  int *flags);

/* Problem generator of the measured workload (not a reference routine):
 * the log-normal Neumann p' operator on a uniform grid with the reference-
 * pressure cell (3,3,3) pinned as PISO_multiple_correction.f90:117-129 does,
 * written into the given arrays (host or device per the context).  Computed
 * on the device from a counter-based generator with its own log, cos and
 * exp (IEEE basic operations only), so the arrays are the same bits on every
 * machine and equal to the oracle's orc_build_lognormal. */
int cap_build_lognormal(cap_ctx *ctx, unsigned long long seed, double sigma,
                        cap_coef *c);

/* Measurement hook, no counterpart in the reference: average device time
 * (ms, HIP events on the context's stream) of each kernel of one PCG-SIP
 * iteration re-launched `reps` times on the vectors of the last pcgsip
 * (fp32 = 0) or pcgsip_mixed (fp32 = 1) solve.  ms[0..5]: SIP forward sweep,
 * SIP backward sweep + dot, pk update, SpMV + dot, fi/res update + L1, and
 * the forward sweep that first updates the residual (0 for fp32 = 1).  An FP64
 * pcgsip iteration after the first runs ms[5], ms[1], ms[2], ms[3]; its pk
 * update (ms[2] for fp32 = 0) also does the previous iteration's fi update. */
int cap_time_kernels(cap_ctx *ctx, int fp32, int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif
