// cap_wave.h -- hyperplane-wavefront schedule for the SIP / IC(0) recursions
// over the tile-sheared layout (cap_layout.h), plus the per-cell recursions.
//
// One workgroup of 256 lanes owns one tile (16x16 columns, all k).  Lane
// (a, b) has chain coordinates (pa, pb) = (a, b) for a forward recursion and
// (15-a, 15-b) for a backward one.  At step t it handles chain index
// rel = t - pa - pb, i.e. k-plane kk = rel (forward) or S-1-rel (backward).
// Its three predecessors in the recursion are
//   k-side: the lane itself at step t-1,
//   j-side: lane (pa-1, pb) at step t-1, or -- at the tile edge -- the
//           neighbouring tile's edge lane, or the ring column at the box edge,
//   i-side: lane (pa, pb-1) at step t-1, or the neighbouring tile / ring,
// which is the (k-1, i-1, j-1) dependency of e.g. pcg-sip.f90:98-106 and the
// (k+1, i+1, j+1) dependency of pcg-sip.f90:112-121.
//
// Values cross tile edges as messages: the edge lane of the producing tile
// writes its value for step t into a message slot; the consuming tile reads
// it at its own step t-15 (same cell).  Every slot holds a NaN sentinel until
// it is written, so a message is its own ready flag.
//
// Cells are classified per step: COMPUTE (solve cell, 1<=kk<=nks), LOAD
// (k-halo plane: the value is data, read from the dependency array), IDLE.
//
// The recursion formulas (Op*::compute) keep the reference's operand order so
// that, with contraction disabled, every value is bit-identical to the
// sequential loop.
#pragma once
#include "cap_layout.h"

namespace cap {

enum { K_IDLE = 0, K_LOAD = 1, K_COMPUTE = 2 };
enum { SRC_LDS = 0, SRC_RING = 1, SRC_MSG = 2 };

// bit pattern of an unwritten message slot (a NaN no arithmetic produces)
constexpr uint64_t MSG_EMPTY = 0x7FF4DEADBEEFCAFEull;
constexpr uint32_t MSG_EMPTY_F32 = 0x7FA5CAFEu;

struct LaneInfo {
    int a, b, pa, pb;       // tile-local column and chain coordinates
    int ci, cj;             // solve column
    bool valid;
    int src_a, src_b;       // where the j-side / i-side predecessor comes from
    int nb_a, nb_b;         // its lane (SRC_LDS)
    int ptile_a, ptile_b;   // its tile (SRC_MSG)
    bool prod_a, prod_b;    // this lane's value goes to the next tile
};

template <bool FWD>
CAP_HD void lane_setup(const Lay &L, int ti, int tj, int lane, LaneInfo &li)
{
    li.a = lane % TW;
    li.b = lane / TW;
    li.pa = FWD ? li.a : TW - 1 - li.a;
    li.pb = FWD ? li.b : TW - 1 - li.b;
    li.ci = ti * TW + li.b;
    li.cj = tj * TW + li.a;
    li.valid = col_valid(L, li.ci, li.cj);
    if (FWD) {
        li.src_a = li.cj == 0 ? SRC_RING : (li.a > 0 ? SRC_LDS : SRC_MSG);
        li.src_b = li.ci == 0 ? SRC_RING : (li.b > 0 ? SRC_LDS : SRC_MSG);
    } else {
        li.src_a = li.cj + 1 == L.njs ? SRC_RING : (li.a < TW - 1 ? SRC_LDS : SRC_MSG);
        li.src_b = li.ci + 1 == L.nis ? SRC_RING : (li.b < TW - 1 ? SRC_LDS : SRC_MSG);
    }
    li.nb_a = li.src_a == SRC_LDS ? (FWD ? lane - 1 : lane + 1) : lane;
    li.nb_b = li.src_b == SRC_LDS ? (FWD ? lane - TW : lane + TW) : lane;
    li.ptile_a = ti * L.nTJ + (FWD ? tj - 1 : tj + 1);
    li.ptile_b = (FWD ? ti - 1 : ti + 1) * L.nTJ + tj;
    li.prod_a = li.valid && (FWD ? (li.a == TW - 1 && tj + 1 < L.nTJ) : (li.a == 0 && tj > 0));
    li.prod_b = li.valid && (FWD ? (li.b == TW - 1 && ti + 1 < L.nTI) : (li.b == 0 && ti > 0));
}

CAP_HD int nsteps(const Lay &L) { return L.S + 2 * (TW - 1); }

// slot every lane of the tile touches at step t
template <bool FWD>
CAP_HD int step_slot(const Lay &L, int t)
{
    return FWD ? posmod(t, L.S) : posmod(L.S - 1 + 2 * (TW - 1) - t, L.S);
}

template <bool FWD>
CAP_HD int step_kind(const Lay &L, const LaneInfo &li, int t, int &kk)
{
    int rel = t - li.pa - li.pb;
    if (!li.valid || rel < 0 || rel >= L.S) { kk = -1; return K_IDLE; }
    kk = FWD ? rel : L.S - 1 - rel;
    if (kk >= 1 && kk <= L.nks) return K_COMPUTE;
    return K_LOAD;
}

// Message slots: [side a/b][component q < nd][tile][step][16], one double each.
CAP_HD size_t msg_index(const Lay &L, int nd, int side, int q, int tile, int t, int s)
{
    return ((((size_t)side * nd + q) * L.ntiles + tile) * (size_t)nsteps(L) + t) * TW + s;
}
CAP_HD size_t msg_count(const Lay &L, int nd) { return msg_index(L, nd, 2, 0, 0, 0, 0); }

// After the messages: per tile, MSG_DUMMY slots of 256 lanes that take the
// stores of lanes with no cell at a step (one slot per chained value and per
// extra output, at most 3 + 4).
constexpr int MSG_DUMMY = 8;
CAP_HD size_t msg_dummy(const Lay &L, int nd, int tile, int lane)
{
    return msg_count(L, nd) + (size_t)tile * MSG_DUMMY * TCOL + lane;
}
// message buffer length (elements) for sweeps with up to nd chained values
CAP_HD size_t msg_alloc(const Lay &L, int nd) { return msg_dummy(L, nd, L.ntiles, 0); }

// the producer of the value consumed at step t wrote it at its step t + 15
CAP_HD size_t msg_in_a(const Lay &L, int nd, int q, const LaneInfo &li, int t)
{
    return msg_index(L, nd, 0, q, li.ptile_a, t + TW - 1, li.b);
}
CAP_HD size_t msg_in_b(const Lay &L, int nd, int q, const LaneInfo &li, int t)
{
    return msg_index(L, nd, 1, q, li.ptile_b, t + TW - 1, li.a);
}
CAP_HD size_t msg_out_a(const Lay &L, int nd, int q, const LaneInfo &li, int tile, int t)
{
    return msg_index(L, nd, 0, q, tile, t, li.b);
}
CAP_HD size_t msg_out_b(const Lay &L, int nd, int q, const LaneInfo &li, int tile, int t)
{
    return msg_index(L, nd, 1, q, tile, t, li.a);
}

// ring elements holding the j-side / i-side predecessor at plane kk
template <bool FWD>
CAP_HD size_t ring_in_a(const Lay &L, const LaneInfo &li, int kk)
{
    return ring_elem(L, FWD ? 2 : 3, li.ci, kk);
}
template <bool FWD>
CAP_HD size_t ring_in_b(const Lay &L, const LaneInfo &li, int kk)
{
    return ring_elem(L, FWD ? 0 : 1, li.cj, kk);
}

// tile processed by the c-th claim (a topological order of the tile DAG)
template <bool FWD>
CAP_HD void claim_tile(const Lay &L, int c, int &ti, int &tj)
{
    int t = FWD ? c : L.ntiles - 1 - c;
    ti = t / L.nTJ;
    tj = t % L.nTJ;
}

// ---------------------------------------------------------------- recursions
// Each op: NIN per-cell inputs x[], ND chained values (dk/di/dj are the k-,
// i-, j-side predecessors in the sweep's direction), NX extra outputs, NR
// reduction terms.

// SIP forward substitution, pcg-sip.f90:102-103 / cgstab_sip.f90:135-136:
//   z = (rhs - bb*z(k-1) - bw*z(i-1) - bs*z(j-1)) * bp
// x = {rhs, bb, bw, bs, bp}
template <class V>
struct OpSipFwdT {
    using T = V;
    static constexpr int NIN = 5, ND = 1, NX = 0, NR = 0;
    static constexpr int PKN = 4;   // x[1..4] come packed: {bb, bw, bs, bp}
    CAP_HD static void compute(const V *x, const V *dk, const V *di, const V *dj, V *d, V *,
                               double *)
    {
        d[0] = (x[0] - x[1] * dk[0] - x[2] * di[0] - x[3] * dj[0]) * x[4];
    }
};
using OpSipFwd = OpSipFwdT<double>;

// The previous PCG-SIP iteration's residual update (pcg-sip.f90:169-171)
// done by the next iteration's forward sweep (pcg-sip.f90:102-103), which
// reads the new residual right after producing it:
//   r = res - alf*zk;  sum |r|;  z = (r - bb*z(k-1) - bw*z(i-1) - bs*z(j-1)) * bp
// x = {res, bb, bw, bs, bp, A pk}; extra output r; alf is the previous
// iteration's step length, read from the solver state when the sweep starts.
// (That iteration's fi update, pcg-sip.f90:168, goes with the pk update.)
template <class V>
struct OpSipFwdResT {
    using T = V;
    static constexpr int NIN = 6, ND = 1, NX = 1, NR = 1;
    static constexpr int PKN = 4;   // x[1..4] come packed: {bb, bw, bs, bp}
    static constexpr int LOOKAHEAD = 12;
    static constexpr bool TAKES_ALF = true;
    V alf;
    CAP_HD void compute(const V *x, const V *dk, const V *di, const V *dj, V *d, V *xo,
                        double *r) const
    {
        const V res = x[0] - alf * x[5];
        xo[0] = res;
        r[0] = __builtin_fabs((double)res);
        d[0] = (res - x[1] * dk[0] - x[2] * di[0] - x[3] * dj[0]) * x[4];
    }
};
using OpSipFwdRes = OpSipFwdResT<double>;

// SIP backward substitution + scalar product, pcg-sip.f90:116-118:
//   z = z - bn*z(j+1) - be*z(i+1) - bt*z(k+1);  sk += res*z
// x = {z, bn, be, bt, res}
// (the FP32 form accumulates the product in FP64)
template <class V>
struct OpSipBwdDotT {
    using T = V;
    static constexpr int NIN = 5, ND = 1, NX = 0, NR = 1;
    CAP_HD static void compute(const V *x, const V *dk, const V *di, const V *dj, V *d, V *,
                               double *r)
    {
        d[0] = x[0] - x[1] * dj[0] - x[2] * di[0] - x[3] * dk[0];
        r[0] = (double)x[4] * (double)d[0];
    }
};
using OpSipBwdDot = OpSipBwdDotT<double>;

// The backward substitution of OpSipBwdDotT with a second product against
// the previous preconditioned residual z_old, for the flexible PCG's
// (Polak-Ribiere) beta.  x = {z, bn, be, bt, res, z_old}
template <class V>
struct OpSipBwdDot2T {
    using T = V;
    static constexpr int NIN = 6, ND = 1, NX = 0, NR = 2;
    CAP_HD static void compute(const V *x, const V *dk, const V *di, const V *dj, V *d, V *,
                               double *r)
    {
        d[0] = x[0] - x[1] * dj[0] - x[2] * di[0] - x[3] * dk[0];
        r[0] = (double)x[4] * (double)d[0];
        r[1] = (double)x[4] * (double)x[5];
    }
};

// SIP backward substitution without a product, cgstab_sip.f90:148-149
// x = {z, bn, be, bt}
struct OpSipBwd {
    using T = double;
    static constexpr int NIN = 4, ND = 1, NX = 0, NR = 0;
    CAP_HD static void compute(const double *x, const double *dk, const double *di,
                               const double *dj, double *d, double *, double *)
    {
        d[0] = x[0] - x[1] * dj[0] - x[2] * di[0] - x[3] * dk[0];
    }
};

// sipsol backward sweep with the update, sipsol.f90:86-88:
//   res = res - bn*res(j+1) - be*res(i+1) - bt*res(k+1);  fi = fi + res
// x = {res, bn, be, bt, fi}; extra output: fi
struct OpSipBwdUpdate {
    using T = double;
    static constexpr int NIN = 5, ND = 1, NX = 1, NR = 0;
    CAP_HD static void compute(const double *x, const double *dk, const double *di,
                               const double *dj, double *d, double *xo, double *)
    {
        d[0] = x[0] - x[1] * dj[0] - x[2] * di[0] - x[3] * dk[0];
        xo[0] = x[4] + d[0];
    }
};

// SIP (Stone) factorisation, pcg-sip.f90:72-84 (== sipsol.f90:34-46):
// x = {ab, aw, as, ap, an, ae, at}; chained {bn, be, bt}; extra {bb, bw, bs, bp}
struct OpSipFactor {
    using T = double;
    static constexpr int NIN = 7, ND = 3, NX = 4, NR = 0;
    double alfa, small;
    CAP_HD void compute(const double *x, const double *dk, const double *di,
                        const double *dj, double *d, double *xo, double *) const
    {
        // dk = {bn,be,bt}(k-1), di = (i-1), dj = (j-1)
        double bb = -x[0] / (1. + alfa * (dk[0] + dk[1]));
        double bw = -x[1] / (1. + alfa * (di[0] + di[2]));
        double bs = -x[2] / (1. + alfa * (dj[1] + dj[2]));
        double p1 = alfa * (bb * dk[0] + bw * di[0]);
        double p2 = alfa * (bb * dk[1] + bs * dj[1]);
        double p3 = alfa * (bw * di[2] + bs * dj[2]);
        double bp = 1. / (x[3] + p1 + p2 + p3 - bb * dk[2] - bw * di[1] - bs * dj[0] + small);
        d[0] = (-x[4] - p1) * bp;
        d[1] = (-x[5] - p2) * bp;
        d[2] = (-x[6] - p3) * bp;
        xo[0] = bb; xo[1] = bw; xo[2] = bs; xo[3] = bp;
    }
};

// IC(0) diagonal, iccg.f90:77-78:
//   d = 1/(ap - aw^2 d(i-1) - as^2 d(j-1) - ab^2 d(k-1));  x = {ap, aw, as, ab}
struct OpIcDiag {
    using T = double;
    static constexpr int NIN = 4, ND = 1, NX = 0, NR = 0;
    CAP_HD static void compute(const double *x, const double *dk, const double *di,
                               const double *dj, double *d, double *, double *)
    {
        d[0] = 1. / (x[0] - x[1] * x[1] * di[0] - x[2] * x[2] * dj[0] - x[3] * x[3] * dk[0]);
    }
};

// IC(0) forward, iccg.f90:96-97:
//   z = (res + aw*z(i-1) + as*z(j-1) + ab*z(k-1)) * d;  x = {res, aw, as, ab, d}
struct OpIcFwd {
    using T = double;
    static constexpr int NIN = 5, ND = 1, NX = 0, NR = 0;
    static constexpr int PKN = 4;   // x[1..4] come packed: {aw, as, ab, d}
    CAP_HD static void compute(const double *x, const double *dk, const double *di,
                               const double *dj, double *d, double *, double *)
    {
        d[0] = (x[0] + x[1] * di[0] + x[2] * dj[0] + x[3] * dk[0]) * x[4];
    }
};

// The previous ICCG iteration's residual update (iccg.f90, as pcg-sip.f90:
// 169-171) done by the next iteration's IC(0) forward sweep, as OpSipFwdResT
// does for PCG-SIP:  r = res - alf*zk;  sum |r|;  z = (r + aw*z(i-1) + ...)*d
// x = {res, aw, as, ab, d, A pk}; extra output r
struct OpIcFwdRes {
    using T = double;
    static constexpr int NIN = 6, ND = 1, NX = 1, NR = 1;
    static constexpr int PKN = 4;   // x[1..4] come packed: {aw, as, ab, d}
    static constexpr int LOOKAHEAD = 12;
    static constexpr bool TAKES_ALF = true;
    double alf;
    CAP_HD void compute(const double *x, const double *dk, const double *di, const double *dj,
                        double *d, double *xo, double *r) const
    {
        const double res = x[0] - alf * x[5];
        xo[0] = res;
        r[0] = __builtin_fabs(res);
        d[0] = (res + x[1] * di[0] + x[2] * dj[0] + x[3] * dk[0]) * x[4];
    }
};

// IC(0) scaling + backward + product, iccg.f90:106 then 118-120:
//   z = z/(d+small);  z = (z + ae*z(i+1) + an*z(j+1) + at*z(k+1)) * d;
//   sk += res*z.   x = {z, ae, an, at, d, res}
struct OpIcBwdDot {
    using T = double;
    static constexpr int NIN = 6, ND = 1, NX = 0, NR = 1;
    // x[1..4] = {ae, an, at, d} from one packed element; four loads a step
    static constexpr int PKN = 4, LOOKAHEAD = 12;
    double small;
    CAP_HD void compute(const double *x, const double *dk, const double *di,
                        const double *dj, double *d, double *, double *r) const
    {
        double z = x[0] / (x[4] + small);
        d[0] = (z + x[1] * di[0] + x[2] * dj[0] + x[3] * dk[0]) * x[4];
        r[0] = x[5] * d[0];
    }
};

// Diagonal ILU(0) of cgstab.f90:58-66:
//   d = 1/(ap - aw d(i-1) ae(i-1) - as d(j-1) an(j-1) - ab d(k-1) at(k-1))
// The west, south and bottom neighbours' own ae, an, at enter the product
// after their d, so they travel with d as chained values: chained
// {d, ae, an, at}, x = {ap, aw, as, ab, ae, an, at}.
struct OpIluDiag {
    using T = double;
    static constexpr int NIN = 7, ND = 4, NX = 0, NR = 0;
    CAP_HD static void compute(const double *x, const double *dk, const double *di,
                               const double *dj, double *d, double *, double *)
    {
        d[0] = 1. / (x[0] - x[1] * di[0] * di[1] - x[2] * dj[0] * dj[2] - x[3] * dk[0] * dk[3]);
        d[1] = x[4];
        d[2] = x[5];
        d[3] = x[6];
    }
};

// ILU(0) scaling + backward substitution, cgstab.f90:129-148:
//   z = z/(d+small);  z = (z + ae*z(i+1) + an*z(j+1) + at*z(k+1)) * d
// x = {z, ae, an, at, d}, x[1..4] packed
struct OpIcBwd {
    using T = double;
    static constexpr int NIN = 5, ND = 1, NX = 0, NR = 0;
    static constexpr int PKN = 4;
    double small;
    CAP_HD void compute(const double *x, const double *dk, const double *di, const double *dj,
                        double *d, double *, double *) const
    {
        const double z = x[0] / (x[4] + small);
        d[0] = (z + x[1] * di[0] + x[2] * dj[0] + x[3] * dk[0]) * x[4];
    }
};

// cgstab.f90's second preconditioner application: the scaling covers
// k = 2..nkm (201-208) but the backward sweep only k = 3..nkmm (212-220), so
// the cells of planes k = 2 and nkm keep the scaled value.  Those cells carry
// a NaN in place of ae in their packed element (no coefficient is NaN).
struct OpIcBwdMid {
    using T = double;
    static constexpr int NIN = 5, ND = 1, NX = 0, NR = 0;
    static constexpr int PKN = 4;
    double small;
    CAP_HD void compute(const double *x, const double *dk, const double *di, const double *dj,
                        double *d, double *, double *) const
    {
        const double z = x[0] / (x[4] + small);
        const bool end_plane = x[1] != x[1];
        d[0] = end_plane ? z : (z + x[1] * di[0] + x[2] * dj[0] + x[3] * dk[0]) * x[4];
    }
};

}  // namespace cap
