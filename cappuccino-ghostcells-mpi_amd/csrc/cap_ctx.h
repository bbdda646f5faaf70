// cap_ctx.h -- the library context behind the opaque cap_ctx handle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include "cap_amd.h"

namespace cap { struct DevState; }

// solver work vectors in the tile-sheared layout
enum {
    T_FI, T_RES, T_ZK, T_PK, T_W, T_RESO, T_UK, T_VK, T_D,
    T_AE, T_AW, T_AN, T_AS, T_AT, T_AB, T_AP, T_SU,
    T_BB, T_BW, T_BS, T_BP, T_BN, T_BE, T_BT,
    T_COUNT
};

// staging copies of caller arrays (reference layout) in host-pointer mode;
// S_AE..S_SU follow the order ae, aw, an, as, at, ab, ap, su
enum {
    S_FI, S_AE, S_AW, S_AN, S_AS, S_AT, S_AB, S_AP, S_SU, S_RES,
    S_BB, S_BS, S_BW, S_BP, S_BN, S_BE, S_BT,
    S_COUNT
};

// FP32 vectors of the mixed-precision inner solver (tile-sheared layout)
enum {
    F_R, F_Z, F_P, F_W, F_X,
    F_AE, F_AW, F_AN, F_AS, F_AT, F_AB, F_AP,
    F_BB, F_BW, F_BS, F_BP, F_BN, F_BE, F_BT,
    F_COUNT
};

#define CAP_NSTAGE 48

struct cap_ctx {
    int ni, nj, nk;
    size_t nstd;
    int flags;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    double *T[T_COUNT];
    void *tpool;                  // one allocation holding every T[] vector
    size_t tcap;
    double *S[S_COUNT];
    float *F[F_COUNT];
    size_t fcap;
    // the forward sweep's SIP factors {bb, bw, bs, bp} packed four to an
    // element (FP64 and the FP32 copy of the mixed-precision solver)
    double *pkd;
    float *pkf;
    // iccg's backward sweep: {ae, an, at, d} packed four to an element
    // (allocated by the first ICCG solve; pbcap elements)
    double *pkb;
    size_t pbcap;
    int bad_launch;               // a sweep was queued without its packed factors
    double *G[CAP_NSTAGE];      // stencil staging (reference layout)
    size_t Gsz[CAP_NSTAGE];
    void *msg;                  // wavefront message slots
    size_t msg_bytes;
    double *scr;                  // small reduction scratch of the stencil entry points
    size_t scr_bytes;
    unsigned long long *claim;
    unsigned long long claim_base;
    int wave_lag;                 // steps a wavefront tile keeps behind its producer
    cap::DevState *st;
    cap::DevState *st2;            // inner solve of the mixed-precision solver
    cap::DevState *st_host;
    double *partials;
    int *asym;                    // per SpMV block: 1 = reads aw, as, ab (FP64 grid, then FP32 grid)
    unsigned *ticket;
    int device;                   // the HIP device current at cap_init
    char err[256];
};

void cap_free_extra(cap_ctx *ctx);
// a reference-layout array of nstd doubles kept in staging slot `slot`
// (allocated on first use; null when that fails, with ctx->err set)
double *cap_slot_buffer(cap_ctx *ctx, int slot);
// The corrector loop of cap_piso.  Before the p solve of the last corrector
// it calls hook (when not null), which may replace the solve's start x on the
// solve range and scale the relative tolerance *sor; a negative return ends
// the loop with that status.
typedef int (*cap_start_hook)(void *user, cap_ctx *ctx, const cap_coef *c, double *x,
                              double *sor);
int cap_piso_run(cap_ctx *ctx, cap_flow *f, const cap_drv_opts *o, int pimple, cap_report *reps,
                 double *cont, cap_start_hook hook, void *user);
// looks up this device's copy of the static scratch array of the
// allocation-free entry points (once per device; the device must be current)
int cap_scratch_setup(int device);
// grad_gauss into gradu(3, nxyz) layout (device pointers; queued, not synchronised)
int cap_grad_gauss_aos(cap_ctx *ctx, const double *u, const double *const *geom, double *g3);
// The east, north and top directions of fluxmass, PISO_assemble_pressure_eq
// and fluxuvw (CDS scheme alone) in one pass, leaving every array as the
// three per-direction calls do (device pointers; queued, not synchronised).
// a6 = {ae, aw, an, as, at, ab}; f3 = {f1, f2, f3}; bp3 = {be, bn, bt};
// s3 = {su, sv, sw}; scr3 = fluxuvw's scratch {bp, bt, bb}.
int cap_fluxmass3_q(cap_ctx *ctx, const cap_fields *F, double *const *a6, double *const *f3);
int cap_piso_assemble3_q(cap_ctx *ctx, const cap_fields *F, double *const *a6, double *const *f3,
                         int ipc, double *const *bp3, double *su);
int cap_fluxuvw3_q(cap_ctx *ctx, const cap_fields *F, const double *vis, double *const *a6,
                   const double *const *f3, double *const *s3, double *const *scr3, double gds);
