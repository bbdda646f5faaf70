// cap_solver.hip -- gfx950 kernels and host drivers of the p' linear solvers
// (sipsol, pcgsip, iccg, cgstab_sip of the reference) behind the C ABI of
// include/cap_amd.h.
//
// Data flow of one solve:
//   reference arrays (padded, j-fastest)  --gather-->  tile-sheared vectors
//   --> Krylov iterations entirely on the device (wavefront sweeps, SpMV,
//       vector updates, deterministic FP64 reductions; the scalars live in a
//       DevState in device memory, and every kernel is a no-op once the
//       stop test has fired)  --scatter-->  reference arrays.
// The host enqueues iterations in batches and reads the 100-byte DevState
// back once per batch to learn whether the solve has stopped.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <algorithm>
#include <type_traits>
#include <vector>
#include "cap_kernels.h"
#include "cap_amd.h"
#include "cap_ctx.h"

using namespace cap;

#define SPIN_LIMIT (1 << 20)

// ============================================================ layout moves
// One block per MOVE_SLOTS slots of a tile, staged through LDS (the schedule
// is in cap_layout.h: move_grid, move_item, move_step).

// reference -> tiled layout: every element (0 for tile columns beyond the
// solve box), the ring columns included
__global__ __launch_bounds__(BLOCK) void k_gather(Lay L, const double *__restrict__ src,
                                                  double *__restrict__ dst)
{
    __shared__ double buf[MOVE_SLOTS * TCOL];
    int tile, s0;
    if (move_item(L, blockIdx.x, tile, s0)) {
        const int ti = tile / L.nTJ, tj = tile % L.nTJ;
        const int lane = threadIdx.x, b = lane / TW, a = lane % TW;
        const int ci = ti * TW + b, cj = tj * TW + a;
        const bool valid = col_valid(L, ci, cj);
        const int ns = min(MOVE_SLOTS, L.S - s0);
        // all row steps' loads first (a lane without an element at a step
        // reads element 0), then the LDS stores, so that the loads are in
        // flight together
        constexpr int NJ = MOVE_SLOTS + TW - 1;
        static_assert(NJ <= 64, "one bit per row step in okm");
        double v[NJ];
        unsigned long long okm = 0ull;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            int r, kk;
            const bool ok = move_step(L, s0, ns, b, a, j, r, kk) && valid;
            okm |= (unsigned long long)ok << j;
            v[j] = src[ok ? std_index(L, ci, cj, kk) : 0];
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            int r, kk;
            if (move_step(L, s0, ns, b, a, j, r, kk)) buf[r * TCOL + lane] = (okm >> j) & 1ull ? v[j] : 0.0;
        }
        __syncthreads();
        const size_t tb = tile_base(L, ti, tj);
        for (int r = 0; r < ns; ++r) dst[elem_at(tb, s0 + r, lane)] = buf[r * TCOL + lane];
    }
    for (size_t e = L.nmain + (size_t)blockIdx.x * BLOCK + threadIdx.x; e < L.n;
         e += (size_t)gridDim.x * BLOCK) {
        int ci, cj, kk;
        double v = 0.0;
        if (decode(L, e, ci, cj, kk)) v = src[std_index(L, ci, cj, kk)];
        dst[e] = v;
    }
}

// tiled -> reference layout, solve-range cells only
__global__ __launch_bounds__(BLOCK) void k_scatter(Lay L, const double *__restrict__ src,
                                                   double *__restrict__ dst)
{
    __shared__ double buf[MOVE_SLOTS * TCOL];
    int tile, s0;
    if (!move_item(L, blockIdx.x, tile, s0)) return;
    const int ti = tile / L.nTJ, tj = tile % L.nTJ;
    const int lane = threadIdx.x, b = lane / TW, a = lane % TW;
    const int ci = ti * TW + b, cj = tj * TW + a;
    const int ns = min(MOVE_SLOTS, L.S - s0);
    const size_t tb = tile_base(L, ti, tj);
    double v[MOVE_SLOTS];   // the loads first (slot s0 again past ns), then the LDS stores
#pragma unroll
    for (int r = 0; r < MOVE_SLOTS; ++r) v[r] = src[elem_at(tb, s0 + (r < ns ? r : 0), lane)];
#pragma unroll
    for (int r = 0; r < MOVE_SLOTS; ++r)
        if (r < ns) buf[r * TCOL + lane] = v[r];
    __syncthreads();
    if (!col_valid(L, ci, cj)) return;
    for (int j = 0; j < MOVE_SLOTS + TW - 1; ++j) {
        int r, kk;
        if (move_step(L, s0, ns, b, a, j, r, kk) && kk >= 1 && kk <= L.nks)
            dst[std_index(L, ci, cj, kk)] = buf[r * TCOL + lane];
    }
}

__global__ __launch_bounds__(BLOCK) void k_fill(double *p, size_t n, double v)
{
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK)
        p[i] = v;
}

// ========================================================== vector kernels
template <class V>
struct Coef7T { const V *ae, *aw, *an, *as, *at, *ab, *ap; };

using Coef7 = Coef7T<double>;

// r = ae fi_e + aw fi_w + an fi_n + as fi_s + at fi_t + ab fi_b + su - ap fi
// (pcg-sip.f90:53-55); sum|r| -> FIN
template <int FIN>
__global__ __launch_bounds__(BLOCK) void k_resid(Lay L, Coef7 A, const double *__restrict__ su,
                                                 const double *__restrict__ fi,
                                                 double *__restrict__ r, DevState *st,
                                                 double *partials, unsigned *ticket)
{
    if (st->done) return;
    Acc acc[1];
    for_cells_skewed(L, [&](const Cell &c) {
        Nb nb = neighbours_at(L, c.ti, c.tj, c.e, c.slot, c.b, c.a, c.kk);
        const size_t e = c.e;
        double v = A.ae[e] * fi[nb.e] + A.aw[e] * fi[nb.w] + A.an[e] * fi[nb.n] +
                   A.as[e] * fi[nb.s] + A.at[e] * fi[nb.t] + A.ab[e] * fi[nb.bt] + su[e] -
                   A.ap[e] * fi[e];
        r[e] = v;
        acc[0] += fabs(v);
    });
    grid_reduce<1, FIN>(acc, blockIdx.x, gridDim.x, partials, ticket, st);
}

// Flags the blocks of the flat-kernel grid whose solve cells are not coupled
// symmetrically entry for entry: aw, as, ab of a cell differ (bit for bit)
// from ae, an, at of its west, south and bottom neighbours.  The p' matrix
// of the reference is assembled face by face (the west coefficient of a
// cell's east neighbour is its east coefficient, fluxmass2.f90:95-96), so
// only a few cells differ, e.g. the neighbours of a pinned reference cell;
// the SpMV reads 4 coefficient arrays instead of 7 in every other block.
__global__ __launch_bounds__(BLOCK) void k_symcheck(Lay L, Coef7 A, int *asym)
{
    auto bits = [](double x) { return __double_as_longlong(x); };
    bool differ = false;
    for_cells_skewed(L, [&](const Cell &c) {
        Nb nb = neighbours_at(L, c.ti, c.tj, c.e, c.slot, c.b, c.a, c.kk);
        const size_t e = c.e;
        differ |= bits(A.aw[e]) != bits(A.ae[nb.w]) || bits(A.as[e]) != bits(A.an[nb.s]) ||
                  bits(A.ab[e]) != bits(A.at[nb.bt]);
    });
    const int any = __syncthreads_or(differ ? 1 : 0);
    if (threadIdx.x == 0) asym[blockIdx.x] = any;
}

// z = ap p - ae p_e - aw p_w - an p_n - as p_s - at p_t - ab p_b
// (pcg-sip.f90:147-149).  MODE 0: sum p.z   MODE 1: sum z.y   MODE 2: sum z.y, sum z.z
// In blocks that k_symcheck found symmetric (asym[block] == 0), aw, as, ab
// are read as ae, an, at of the neighbour: the same values, bit for bit.
template <class V, int MODE, int FIN>
__global__ __launch_bounds__(BLOCK) void k_spmv(Lay L, Coef7T<V> A, const V *__restrict__ p,
                                                V *__restrict__ z, const V *__restrict__ y,
                                                DevState *st, double *partials,
                                                unsigned *ticket, const int *asym)
{
    if (st->done) return;
    Acc acc[MODE == 2 ? 2 : 1];
    // a cell's operands: p of the cell and its six neighbours, the seven
    // coefficients, y (MODE 1 and 2)
    struct Ops { V pc, pe, pw, pn, ps, pt, pb, ap, ae, aw, an, as, at, ab, y; };
    auto run = [&](auto symc) {
        constexpr bool SYM = decltype(symc)::value;
        for_cells_pipelined<Ops>(L, [&](const Cell &c, bool valid) {
            const size_t e = c.e;
            const Nb nb = valid ? neighbours_at(L, c.ti, c.tj, c.e, c.slot, c.b, c.a, c.kk)
                                : Nb{e, e, e, e, e, e};
            Ops o;
            o.pc = p[e]; o.pe = p[nb.e]; o.pw = p[nb.w]; o.pn = p[nb.n]; o.ps = p[nb.s];
            o.pt = p[nb.t]; o.pb = p[nb.bt];
            o.ap = A.ap[e]; o.ae = A.ae[e]; o.an = A.an[e]; o.at = A.at[e];
            o.aw = SYM ? A.ae[nb.w] : A.aw[e];
            o.as = SYM ? A.an[nb.s] : A.as[e];
            o.ab = SYM ? A.at[nb.bt] : A.ab[e];
            o.y = MODE == 0 ? V(0) : y[e];
            return o;
        }, [&](const Cell &c, const Ops &o) {
            V v = o.ap * o.pc - o.ae * o.pe - o.aw * o.pw - o.an * o.pn - o.as * o.ps -
                  o.at * o.pt - o.ab * o.pb;
            z[c.e] = v;
            if constexpr (MODE == 0) acc[0] += (double)o.pc * (double)v;
            if constexpr (MODE == 1) acc[0] += (double)v * (double)o.y;
            if constexpr (MODE == 2) { acc[0] += (double)v * (double)o.y; acc[1] += (double)v * (double)v; }
        });
    };
    if (asym && asym[blockIdx.x] == 0)
        run(std::true_type());
    else
        run(std::false_type());
    grid_reduce<MODE == 2 ? 2 : 1, FIN>(acc, blockIdx.x, gridDim.x, partials, ticket, st);
}

// pk = zk + bet pk   (pcg-sip.f90:135)
template <class V>
__global__ __launch_bounds__(BLOCK) void k_pupdate(Lay L, const V *__restrict__ z,
                                                   V *__restrict__ p, DevState *st)
{
    if (st->done) return;
    const V bet = (V)st->bet;
    for_cells(L, [&](const Cell &c) { p[c.e] = z[c.e] + bet * p[c.e]; });
}

// The previous iteration's fi = fi + alf pk (pcg-sip.f90:168), while its pk
// and alf are still current, then pk = zk + bet pk (pcg-sip.f90:135)
template <class V>
__global__ __launch_bounds__(BLOCK) void k_pupdate_fi(Lay L, const V *__restrict__ z,
                                                      V *__restrict__ p, V *__restrict__ fi,
                                                      DevState *st)
{
    if (st->done) return;
    const V bet = (V)st->bet, alf = (V)st->alf;
    for_cells(L, [&](const Cell &c) {
        const V po = p[c.e];
        fi[c.e] = fi[c.e] + alf * po;
        p[c.e] = z[c.e] + bet * po;
    });
}

// fi = fi + alf pk after a stop in the forward sweep, whose residual update
// already ran; the state's done flag is set by then (pcg-sip.f90:168)
template <class V>
__global__ __launch_bounds__(BLOCK) void k_fi_last(Lay L, V *__restrict__ fi, const V *__restrict__ p,
                                                   const DevState *st)
{
    const V alf = (V)st->alf;
    for_cells(L, [&](const Cell &c) { fi[c.e] = fi[c.e] + alf * p[c.e]; });
}

// pk = res + om (pk - alf uk)   (cgstab_sip.f90:124)
__global__ __launch_bounds__(BLOCK) void k_bic_p(Lay L, const double *__restrict__ res,
                                                 const double *__restrict__ uk,
                                                 double *__restrict__ p, DevState *st)
{
    if (st->done) return;
    const double om = st->om, alf = st->alf;
    for_cells(L, [&](const Cell &c) { p[c.e] = res[c.e] + om * (p[c.e] - alf * uk[c.e]); });
}

// fi += a p; res -= a z; with a = alf (which 0) or gam (which 1);
// L1 of res -> FIN when FIN != F_NONE   (pcg-sip.f90:167-171, cgstab_sip.f90:187-188, 254-256)
template <class V, int WHICH, int FIN>
__global__ __launch_bounds__(BLOCK) void k_update(Lay L, V *__restrict__ fi, V *__restrict__ res,
                                                  const V *__restrict__ p,
                                                  const V *__restrict__ z, DevState *st,
                                                  double *partials, unsigned *ticket)
{
    if (st->done) return;
    const V a = (V)(WHICH == 0 ? st->alf : st->gam);
    Acc acc[1];
    for_cells(L, [&](const Cell &c) {
        fi[c.e] = fi[c.e] + a * p[c.e];
        V r = res[c.e] - a * z[c.e];
        res[c.e] = r;
        acc[0] += fabs((double)r);
    });
    if constexpr (FIN != F_NONE)
        grid_reduce<1, FIN>(acc, blockIdx.x, gridDim.x, partials, ticket, st);
}

// k_update<double, 0, F_BIC_RESL> (fi += alf zk, res -= alf vk, L1 of res and
// the stop test, cgstab_sip.f90:254-266) with the next iteration's
// bet = res.reso (cgstab_sip.f90:106-116) summed from the same new residual
__global__ __launch_bounds__(BLOCK) void k_bic_update_bet(Lay L, double *__restrict__ fi,
                                                          double *__restrict__ res,
                                                          const double *__restrict__ p,
                                                          const double *__restrict__ z,
                                                          const double *__restrict__ reso,
                                                          DevState *st, double *partials,
                                                          unsigned *ticket)
{
    if (st->done) return;
    const double a = st->alf;
    Acc acc[2];
    for_cells(L, [&](const Cell &c) {
        fi[c.e] = fi[c.e] + a * p[c.e];
        const double r = res[c.e] - a * z[c.e];
        res[c.e] = r;
        acc[0] += fabs(r);
        acc[1] += r * reso[c.e];
    });
    grid_reduce<2, F_BIC_RESL_BET>(acc, blockIdx.x, gridDim.x, partials, ticket, st);
}

// sum x.y -> FIN; optionally y := x first (COPY: reso = res, cgstab_sip.f90:87)
template <int FIN, bool COPY>
__global__ __launch_bounds__(BLOCK) void k_dot(Lay L, const double *__restrict__ x,
                                               double *__restrict__ y, DevState *st,
                                               double *partials, unsigned *ticket)
{
    if (st->done) return;
    Acc acc[1];
    for_cells(L, [&](const Cell &c) {
        if constexpr (COPY) y[c.e] = x[c.e];
        acc[0] += x[c.e] * y[c.e];
    });
    grid_reduce<1, FIN>(acc, blockIdx.x, gridDim.x, partials, ticket, st);
}

// ======================================================= wavefront sweeps
template <class Op>
struct WaveIO {
    using V = typename Op::T;
    const V *in[Op::NIN];
    V *dep[Op::ND];
    V *xo[Op::NX > 0 ? Op::NX : 1];
    V *msg;                         // message slots (filled with the sentinel)
    const V *pk = nullptr;          // inputs 1..PKN packed four to an element
};

// how many inputs (from x[1]) an op reads packed four to an element
template <class Op, class = void>
struct PkN {
    static constexpr int v = 0;
};
template <class Op>
struct PkN<Op, std::void_t<decltype(Op::PKN)>> {
    static constexpr int v = Op::PKN;
};

// ops that take the iteration's step length (alf) from the solver state
template <class Op, class = void>
struct TakesAlf : std::false_type {};
template <class Op>
struct TakesAlf<Op, std::void_t<decltype(Op::TAKES_ALF)>> : std::true_type {};

// depth U of a compute lane's input ring (U-1 steps of inputs in flight): 16
// for one chained value and up to five inputs, else 8 (register budget),
// unless the op says
template <class Op, class = void>
struct LookAhead {
    static constexpr int v = Op::ND == 1 && Op::NIN <= 5 ? 16 : 8;
};
template <class Op>
struct LookAhead<Op, std::void_t<decltype(Op::LOOKAHEAD)>> {
    static constexpr int v = Op::LOOKAHEAD;
};

// one element of four packed values: two 16-byte loads (FP64) or one (FP32)
__device__ __forceinline__ void load4(const double *p, double (&v)[4])
{
    const double2 a = *reinterpret_cast<const double2 *>(p);
    const double2 b = *reinterpret_cast<const double2 *>(p + 2);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void load4(const float *p, float (&v)[4])
{
    const float4 a = *reinterpret_cast<const float4 *>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}

// pack four tiled vectors (d may be null: zeros) into out[4 e + q]
template <class V>
__global__ void k_pack4(size_t n, const V *a, const V *b, const V *c, const V *d, V *out)
{
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        out[4 * e] = a[e];
        out[4 * e + 1] = b[e];
        out[4 * e + 2] = c[e];
        out[4 * e + 3] = d ? d[e] : V(0);
    }
}

struct WaveSync {
    unsigned long long *claim;      // tile claim counter
    unsigned long long claim_base;
    DevState *st;
    double *partials;
    unsigned *ticket;
    int lag;                        // steps a consumer tile keeps behind its producer
};

template <class V>
__device__ __forceinline__ V coh_load(const V *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class V>
__device__ __forceinline__ void coh_store(V *p, V v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool msg_empty(double v) { return (uint64_t)__double_as_longlong(v) == MSG_EMPTY; }
__device__ __forceinline__ bool msg_empty(float v) { return (uint32_t)__float_as_uint(v) == MSG_EMPTY_F32; }

template <class V> struct MsgEmpty;
template <> struct MsgEmpty<double> {
    static __device__ __forceinline__ double get() { return __longlong_as_double((long long)MSG_EMPTY); }
};
template <> struct MsgEmpty<float> {
    static __device__ __forceinline__ float get() { return __uint_as_float(MSG_EMPTY_F32); }
};

template <class V>
__global__ __launch_bounds__(BLOCK) void k_msg_fill(V *p, size_t n)
{
    const V s = MsgEmpty<V>::get();
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK)
        p[i] = s;
}

// spin until a message slot has been written (bounded: on expiry the solve is
// flagged and ends with an error, every wave still runs to completion; once
// a solve is flagged, waits give up at once)
template <class V>
__device__ __forceinline__ V msg_wait(const V *p, V v, DevState *st)
{
    for (int it = 0; it < SPIN_LIMIT && msg_empty(v); ++it) {
        if (__hip_atomic_load(&st->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return v;
        __builtin_amdgcn_s_sleep(1);
        v = coh_load(p);
    }
    if (msg_empty(v)) {
        st->err = CAP_ERR_TIMEOUT;
        st->done = 1;
    }
    return v;
}

// One workgroup = one tile (16x16 columns, all planes) marching through the
// tile's steps, in five waves:
//  * waves 0-3 (256 lanes, one per column) compute.  Their inputs for step
//    t+U-1 are loaded at the start of step t (a register ring of depth U), so
//    loads stay in flight across steps; values cross lanes through LDS, one
//    barrier a step.
//  * wave 4 carries the values that come from outside the tile: it reads the
//    neighbouring tiles' messages (or the ring columns at the box edge) DM
//    steps ahead and hands them to the compute waves through LDS, so a late
//    message holds up only this wave's memory counter.  When a message is
//    late it also waits until the producer is sy.lag steps further on, so
//    that later messages are found written.
// Tiles are claimed from a counter in a topological order, so a workgroup only
// ever waits for tiles held by running workgroups.
constexpr int WAVE_BLOCK = BLOCK + 64;

template <class Op, bool FWD, int FIN>
__global__ __launch_bounds__(WAVE_BLOCK) void k_wave(Lay L, Op op, WaveIO<Op> io, WaveSync sy)
{
    using V = typename Op::T;
    constexpr int PKN = PkN<Op>::v;
    constexpr int ND = Op::ND, NIN = Op::NIN, U = LookAhead<Op>::v;
    // messages are read DM steps ahead of use: further ahead, more of them are
    // found unwritten and the consumer tile falls further behind its producer
    constexpr int DM = 2;
    constexpr int NXS = Op::NX > 0 ? Op::NX : 1, NRS = Op::NR > 0 ? Op::NR : 1;
    __shared__ V lds[2][ND][TCOL];
    __shared__ V mbuf[2][ND][2][TW];   // [step parity][q][j-side by b / i-side by a]
    __shared__ int sh_tile, sh_flag;
    if (threadIdx.x == 0) {
        const unsigned long long c = atomicAdd(sy.claim, 1ull) - sy.claim_base;
        sh_flag = sy.st->done;
        if (c >= (unsigned long long)L.ntiles) {   // launch bookkeeping broken
            sh_flag = 1;
            sy.st->err = CAP_ERR_ARG;
            sy.st->done = 1;
        }
        sh_tile = (int)c;
    }
    __syncthreads();
    if (sh_flag) return;
    int ti, tj;
    claim_tile<FWD>(L, sh_tile, ti, tj);
    ti = __builtin_amdgcn_readfirstlane(ti);
    tj = __builtin_amdgcn_readfirstlane(tj);
    const int tile = ti * L.nTJ + tj;
    const int NST = nsteps(L);
    const int last_t = NST - 1;
    // steps padded to a multiple of both loop strides (padded steps are idle)
    constexpr int UM = U > DM ? U : DM;
    static_assert(UM % U == 0 && UM % DM == 0, "strides must divide each other");
    const int NSTP = (NST + UM - 1) / UM * UM;
    V *const msgp = io.msg;
    Acc red[NRS];

    if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) == BLOCK / 64) {
        // ------------------------------------------ wave 4: the tile's inputs
        // lanes  0-15: j-side input of row b = idx (chain-start lane of the row)
        //       16-31: i-side input of column a = idx
        const int c = threadIdx.x - BLOCK;
        const int grp = c >> 4, idx = c & (TW - 1), side = grp & 1;
        const bool in = grp < 2;
        const int a_last = min(L.njs - 1 - tj * TW, TW - 1);
        const int b_last = min(L.nis - 1 - ti * TW, TW - 1);
        const int sa = FWD ? 0 : (tj == L.nTJ - 1 ? a_last : TW - 1);
        const int sb = FWD ? 0 : (ti == L.nTI - 1 ? b_last : TW - 1);
        const int cl = side == 0 ? idx * TW + sa : sb * TW + idx;
        LaneInfo li;
        lane_setup<FWD>(L, ti, tj, cl, li);
        const bool ring = side == 0 ? (FWD ? tj == 0 : tj == L.nTJ - 1)
                                    : (FWD ? ti == 0 : ti == L.nTI - 1);
        const int ptile = side == 0 ? li.ptile_a : li.ptile_b;
        const int msl = side == 0 ? li.b : li.a;
        V *depp[ND];
#pragma unroll
        for (int q = 0; q < ND; ++q) depp[q] = io.dep[q];
        // where the value for step t comes from (any valid address for lanes
        // 32-63, whose loads are not used)
        auto src = [&](int q, int t) -> const V * {
            if (ring) {   // ring column at the plane of the chain-start lane
                int kk = t - li.pa - li.pb;
                if (!FWD) kk = L.S - 1 - kk;
                kk = kk < 0 ? 0 : (kk > L.S - 1 ? L.S - 1 : kk);
                return depp[q] + (side == 0 ? ring_elem(L, FWD ? 2 : 3, min(li.ci, L.nis - 1), kk)
                                            : ring_elem(L, FWD ? 0 : 1, min(li.cj, L.njs - 1), kk));
            }
            const int tm = min(t + TW - 1, last_t);
            return msgp + msg_index(L, ND, side, q, ptile < 0 ? 0 : ptile, tm, msl);
        };
        auto computes = [&](int t) {
            int kk;
            return step_kind<FWD>(L, li, t, kk) == K_COMPUTE;
        };
        auto needed = [&](int t) { return in && !ring && computes(t); };
        auto fetch = [&](int t, V (&y)[ND]) {
#pragma unroll
            for (int q = 0; q < ND; ++q) y[q] = coh_load(src(q, t < last_t ? t : last_t));
        };
        V m[DM][ND];   // m[j] holds step t with t % DM == j
        // Hand the values of step t (in m[t % DM]) to the cell waves.  When a
        // message is late, the lanes wait for their values of step t and of
        // step t + sy.lag (the producer is then sy.lag steps ahead) and the
        // whole wave reads its messages in flight again.  Returns whether the
        // wave was late.
        auto deliver = [&](int t, V (&y)[ND]) -> bool {
            bool late = false;
            if (needed(t)) {
#pragma unroll
                for (int q = 0; q < ND; ++q) late |= msg_empty(y[q]);
            }
            const bool any_late = __any(late);
            if (any_late) {
                if (late) {
#pragma unroll
                    for (int q = 0; q < ND; ++q) y[q] = msg_wait(src(q, t), y[q], sy.st);
                }
                if (needed(t + sy.lag))
                    (void)msg_wait(src(0, t + sy.lag), MsgEmpty<V>::get(), sy.st);
            }
            if (in) {
#pragma unroll
                for (int q = 0; q < ND; ++q) mbuf[t & 1][q][side][idx] = y[q];
            }
            return any_late;
        };
#pragma unroll
        for (int d = 0; d < DM; ++d) fetch(d, m[d]);
        (void)deliver(0, m[0]);
        fetch(DM, m[0]);
        __syncthreads();
        for (int t0 = 0; t0 < NSTP; t0 += DM) {
#pragma unroll
            for (int d = 0; d < DM; ++d) {
                // during step t: hand over step t+1, read step t+1+DM
                const int t = t0 + d;
                if (deliver(t + 1, m[(d + 1) % DM])) {
#pragma unroll
                    for (int j = 2; j <= DM; ++j) fetch(t + j, m[(d + j) % DM]);
                }
                fetch(t + 1 + DM, m[(d + 1) % DM]);
                __syncthreads();
            }
        }
    } else {
        // --------------------------------------------------- waves 0-3: cells
        if constexpr (TakesAlf<Op>::value) op.alf = (V)sy.st->alf;
        const int lane = threadIdx.x;
        LaneInfo li;
        lane_setup<FWD>(L, ti, tj, lane, li);
        const size_t tb = tile_base(L, ti, tj);
        V inb[U][NIN];
        V prev[ND];
        for (int q = 0; q < ND; ++q) prev[q] = V(0);
        // array pointers in registers (a per-lane choice between two of them
        // must not turn into a load of the pointer itself)
        const V *inp[NIN];
        V *depp[ND];
        V *xop[NXS];
#pragma unroll
        for (int q = 0; q < NIN; ++q) inp[q] = io.in[q];
#pragma unroll
        for (int q = 0; q < ND; ++q) depp[q] = io.dep[q];
#pragma unroll
        for (int q = 0; q < NXS; ++q) xop[q] = io.xo[q];
        // stores of lanes with no cell at a step go to private dummy slots, so
        // every step issues the same memory operations
        V *const dummy = msgp + msg_dummy(L, ND, tile, lane);
        const bool sa_ext = li.src_a != SRC_LDS, sb_ext = li.src_b != SRC_LDS;
        // the lane's two halo planes of the chained values (kk = 0 and nks+1)
        V hlo[ND], hhi[ND];
        {
            const size_t e0 = elem_at(tb, (li.a + li.b) % L.S, lane);
            const size_t e1 = elem_at(tb, (L.nks + 1 + li.a + li.b) % L.S, lane);
#pragma unroll
            for (int q = 0; q < ND; ++q) {
                hlo[q] = depp[q][e0];
                hhi[q] = depp[q][e1];
            }
        }
        const V *const pkp = io.pk;
        // Addresses are a uniform row pointer (the tile's slot, in scalar
        // registers) plus the lane's 32-bit offset, so that the loads and
        // stores take the scalar-base form and no 64-bit address is formed
        // per lane.
        const unsigned ul = (unsigned)lane;
        auto fetch_row = [&](size_t ro, V (&x)[NIN]) {
            if constexpr (PKN > 0) {
                x[0] = (inp[0] + ro)[ul];
                V p4[4];
                load4(pkp + 4 * ro + 4 * ul, p4);
#pragma unroll
                for (int q = 1; q <= PKN; ++q) x[q] = p4[q - 1];
#pragma unroll
                for (int q = PKN + 1; q < NIN; ++q) x[q] = (inp[q] + ro)[ul];
            } else {
#pragma unroll
                for (int q = 0; q < NIN; ++q) x[q] = (inp[q] + ro)[ul];
            }
        };
        // the slot of the next step (one further in the sweep's direction)
        auto advance = [&](int sl) {
            return FWD ? (sl + 1 == L.S ? 0 : sl + 1) : (sl == 0 ? L.S - 1 : sl - 1);
        };
        // Steps at which every lane of a full tile has a solve cell: there
        // the step kinds, halo values and dummy stores drop out.
        const bool full = L.nis - ti * TW >= TW && L.njs - tj * TW >= TW;
        const int t_lo = 2 * (TW - 1) + 1, t_hi = L.nks;
        // Inputs are loaded U-1 steps ahead, at the start of a step (right
        // after its LDS reads, so the loads issue while those are in flight),
        // into the ring slot the previous step has just consumed.
#pragma unroll
        for (int d = 0; d < U - 1; ++d) fetch_row(tb + (size_t)step_slot<FWD>(L, d) * TCOL, inb[d]);
        // slots of step t and of step t + U - 1
        int sl_t = step_slot<FWD>(L, 0), sl_u = step_slot<FWD>(L, U - 1);
        __syncthreads();   // step-0 inputs from outside the tile are in mbuf
        for (int t0 = 0; t0 < NSTP; t0 += U) {
#pragma unroll
            for (int d = 0; d < U; ++d) {
                const int t = t0 + d;
                const size_t ro = tb + (size_t)sl_t * TCOL;
                const size_t e = ro + ul;
                const int rp = (t + 1) & 1, wp = t & 1;
                V dk[ND], di[ND], dj[ND], out[ND];
#pragma unroll
                for (int q = 0; q < ND; ++q) {
                    dk[q] = prev[q];
                    dj[q] = sa_ext ? mbuf[wp][q][0][li.b] : lds[rp][q][li.nb_a];
                    di[q] = sb_ext ? mbuf[wp][q][1][li.a] : lds[rp][q][li.nb_b];
                }
                __builtin_amdgcn_sched_barrier(0);
                fetch_row(tb + (size_t)sl_u * TCOL, inb[(d + U - 1) % U]);
                V xv[NXS];
                double r[NRS];
                op.compute(inb[d], dk, di, dj, out, xv, r);
                if (full && t >= t_lo && t <= t_hi) {
#pragma unroll
                    for (int q = 0; q < ND; ++q) lds[wp][q][lane] = out[q];
#pragma unroll
                    for (int q = 0; q < ND; ++q) {
                        // streaming stores: the next sweep reads them from HBM
                        __builtin_nontemporal_store(out[q], &(depp[q] + ro)[ul]);
                        if (li.prod_a) coh_store(msgp + msg_index(L, ND, 0, q, tile, t, 0) + li.b, out[q]);
                        if (li.prod_b) coh_store(msgp + msg_index(L, ND, 1, q, tile, t, 0) + li.a, out[q]);
                    }
#pragma unroll
                    for (int q = 0; q < Op::NX; ++q) __builtin_nontemporal_store(xv[q], &(xop[q] + ro)[ul]);
#pragma unroll
                    for (int q = 0; q < Op::NR; ++q) red[q] += r[q];
                } else {
                    int kk;
                    const int kind = step_kind<FWD>(L, li, t, kk);
                    const bool comp = kind == K_COMPUTE, idle = kind == K_IDLE;
#pragma unroll
                    for (int q = 0; q < ND; ++q) {
                        out[q] = comp ? out[q] : (kind == K_LOAD ? (kk == 0 ? hlo[q] : hhi[q]) : V(0));
                        lds[wp][q][lane] = out[q];
                        *(idle ? dummy + q * TCOL : depp[q] + e) = out[q];
                    }
#pragma unroll
                    for (int q = 0; q < Op::NX; ++q) *(comp ? xop[q] + e : dummy + (ND + q) * TCOL) = xv[q];
                    if (comp) {
#pragma unroll
                        for (int q = 0; q < ND; ++q) {
                            if (li.prod_a) coh_store(msgp + msg_index(L, ND, 0, q, tile, t, li.b), out[q]);
                            if (li.prod_b) coh_store(msgp + msg_index(L, ND, 1, q, tile, t, li.a), out[q]);
                        }
                    }
#pragma unroll
                    for (int q = 0; q < Op::NR; ++q) red[q] += comp ? r[q] : 0.0;
                }
#pragma unroll
                for (int q = 0; q < ND; ++q) prev[q] = out[q];
                sl_t = advance(sl_t);
                sl_u = advance(sl_u);
                __syncthreads();
            }
        }
    }
    if constexpr (FIN != F_NONE) {
        grid_reduce<NRS, FIN>(red, tile, L.ntiles, sy.partials, sy.ticket, sy.st);
    }
}

// ============================================================ host helpers
// grid of the flat kernels: eight blocks a tile (one per XCD, each an eighth of
// the tile's slots), at most 2048 blocks (tiles beyond 256 are strided)
static int grid_for(const Lay &L)
{
    return 8 * std::min(L.ntiles, 256);
}

// grid of the SpMV (and of its symmetry check, whose per-block flags it
// reads): fewer, longer blocks, so that the p and coefficient lines of a
// block's previous planes, which its stencil reads again, are still in the
// XCD's L2 (measured at 256^3 with the skewed walk: 0.215 ms with 512
// blocks, 0.217 with 1024, 0.222 with 2048)
static int grid_spmv(const Lay &L)
{
    return 8 * std::min(L.ntiles, 64);
}

static int fail(cap_ctx *c, int code, const char *what)
{
    snprintf(c->err, sizeof c->err, "%s", what);
    return code;
}

#define HIPCHK(call)                                                                      \
    do {                                                                                  \
        hipError_t _e = (call);                                                           \
        if (_e != hipSuccess) {                                                           \
            snprintf(ctx->err, sizeof ctx->err, "%s failed: %s (%s:%d)", #call,           \
                     hipGetErrorString(_e), __FILE__, __LINE__);                          \
            return CAP_ERR_HIP;                                                           \
        }                                                                                 \
    } while (0)

static int ensure_tiled(cap_ctx *ctx, size_t n)
{
    if (ctx->tcap >= n) return CAP_OK;
    if (ctx->tpool) { (void)hipFree(ctx->tpool); ctx->tpool = nullptr; }
    for (int v = 0; v < T_COUNT; ++v) ctx->T[v] = nullptr;
    // One pool; consecutive vectors start 4352 bytes (17 x 256) apart modulo
    // 64 KiB, so the same element of different vectors does not fall on the
    // same HBM channel when a kernel streams several vectors at once.
    const size_t stride = ((n * sizeof(double) + 65535) / 65536) * 65536 + 4352;
    HIPCHK(hipMalloc(&ctx->tpool, stride * T_COUNT));
    for (int v = 0; v < T_COUNT; ++v) ctx->T[v] = (double *)((char *)ctx->tpool + stride * v);
    if (ctx->pkd) { (void)hipFree(ctx->pkd); ctx->pkd = nullptr; }
    HIPCHK(hipMalloc(&ctx->pkd, 4 * n * sizeof(double)));

    ctx->tcap = n;
    return CAP_OK;
}

// the backward-sweep packs {ae, an, at, d} of iccg (one) and cgstab (two,
// the second at pkb + 4 n): only those solvers read them, so they are
// allocated on the first such solve of a size, not with the other vectors
static int ensure_pkb(cap_ctx *ctx, size_t n, int packs = 1)
{
    const size_t need = n * (size_t)packs;
    if (ctx->pbcap >= need) return CAP_OK;
    if (ctx->pkb) { (void)hipFree(ctx->pkb); ctx->pkb = nullptr; }
    ctx->pbcap = 0;
    HIPCHK(hipMalloc(&ctx->pkb, 4 * need * sizeof(double)));
    ctx->pbcap = need;
    return CAP_OK;
}

static int ensure_std(cap_ctx *ctx)
{
    if (ctx->S[0]) return CAP_OK;
    for (int v = 0; v < S_COUNT; ++v) HIPCHK(hipMalloc(&ctx->S[v], ctx->nstd * sizeof(double)));
    return CAP_OK;
}

// device pointer for a caller array (uploading it in host-pointer mode)
static int dev_in(cap_ctx *ctx, int slot, const double *p, double **out)
{
    if (!p) { *out = nullptr; return CAP_OK; }
    if (ctx->flags & CAP_DEVICE_POINTERS) { *out = (double *)p; return CAP_OK; }
    int rc = ensure_std(ctx);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->S[slot], p, ctx->nstd * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    *out = ctx->S[slot];
    return CAP_OK;
}

static int dev_out(cap_ctx *ctx, double *host, const double *dev)
{
    if (!host || (ctx->flags & CAP_DEVICE_POINTERS)) return CAP_OK;
    HIPCHK(hipMemcpyAsync(host, dev, ctx->nstd * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    return CAP_OK;
}

template <class Op, bool FWD, int FIN>
static void launch_wave(cap_ctx *ctx, const Lay &L, const Op &op, WaveIO<Op> io,
                        DevState *st = nullptr)
{
    using V = typename Op::T;
    if (PkN<Op>::v > 0 && !io.pk) {     // never queue a sweep without its packed factors
        ctx->bad_launch = 1;
        return;
    }
    WaveSync sy;
    sy.claim = ctx->claim;
    sy.claim_base = ctx->claim_base;
    sy.st = st ? st : ctx->st;
    sy.partials = ctx->partials;
    sy.ticket = ctx->ticket;
    sy.lag = ctx->wave_lag;
    ctx->claim_base += (unsigned long long)L.ntiles;
    const size_t nmsg = msg_count(L, Op::ND);
    io.msg = (V *)ctx->msg;
    hipLaunchKernelGGL(k_msg_fill<V>, dim3((int)std::min<size_t>((nmsg + BLOCK - 1) / BLOCK, 2048)),
                       dim3(BLOCK), 0, ctx->stream, io.msg, nmsg);
    hipLaunchKernelGGL((k_wave<Op, FWD, FIN>), dim3(L.ntiles), dim3(WAVE_BLOCK), 0, ctx->stream, L, op,
                       io, sy);
}

static Coef7 coef7(cap_ctx *ctx)
{
    Coef7 A;
    A.ae = ctx->T[T_AE]; A.aw = ctx->T[T_AW]; A.an = ctx->T[T_AN]; A.as = ctx->T[T_AS];
    A.at = ctx->T[T_AT]; A.ab = ctx->T[T_AB]; A.ap = ctx->T[T_AP];
    return A;
}

// flags the SpMV blocks of this solve's layout that must read aw, as, ab
// (asym, one flag a block of the SpMV grid).  The FP32 SpMV runs on the same
// grid and reads the same flags: two doubles that are equal bit for bit
// round to equal floats.
static void sym_check(cap_ctx *ctx, const Lay &L)
{
    hipLaunchKernelGGL(k_symcheck, dim3(grid_spmv(L)), dim3(BLOCK), 0, ctx->stream, L, coef7(ctx),
                       ctx->asym);
}

// Everything a solve needs in device memory, in both layouts.
struct SolveIn {
    double *fi, *ae, *aw, *an, *as, *at, *ab, *ap, *su, *res;
    double *bb, *bs, *bw, *bp, *bn, *be, *bt;
};

static int prepare(cap_ctx *ctx, const Lay &L, double *fi, const cap_coef *c, SolveIn &d,
                   bool need_res_halo, bool need_b_halo)
{
    int rc;
    if (!fi || !c || !c->ae || !c->aw || !c->an || !c->as || !c->at || !c->ab || !c->ap || !c->su)
        return fail(ctx, CAP_ERR_ARG, "null array argument");
    if ((rc = ensure_tiled(ctx, L.n))) return rc;
    {   // message slots for the largest sweep (four chained FP64 values)
        const size_t bytes = msg_alloc(L, 4) * sizeof(double);
        if (ctx->msg_bytes < bytes) {
            if (ctx->msg) (void)hipFree(ctx->msg);
            ctx->msg = nullptr;
            HIPCHK(hipMalloc(&ctx->msg, bytes));
            ctx->msg_bytes = bytes;
        }
    }
    if ((rc = dev_in(ctx, S_FI, fi, &d.fi))) return rc;
    const double *src[9] = {c->ae, c->aw, c->an, c->as, c->at, c->ab, c->ap, c->su, nullptr};
    double **dst[9] = {&d.ae, &d.aw, &d.an, &d.as, &d.at, &d.ab, &d.ap, &d.su, nullptr};
    for (int v = 0; v < 8; ++v)
        if ((rc = dev_in(ctx, S_AE + v, src[v], dst[v]))) return rc;
    d.res = d.bb = d.bs = d.bw = d.bp = d.bn = d.be = d.bt = nullptr;
    if (c->res && (rc = dev_in(ctx, S_RES, c->res, &d.res))) return rc;
    if (c->bn) {
        if ((rc = dev_in(ctx, S_BN, c->bn, &d.bn))) return rc;
        if ((rc = dev_in(ctx, S_BE, c->be, &d.be))) return rc;
        if ((rc = dev_in(ctx, S_BT, c->bt, &d.bt))) return rc;
    }
    d.bb = c->bb ? ((ctx->flags & CAP_DEVICE_POINTERS) ? c->bb : ctx->S[S_BB]) : nullptr;
    d.bs = c->bs ? ((ctx->flags & CAP_DEVICE_POINTERS) ? c->bs : ctx->S[S_BS]) : nullptr;
    d.bw = c->bw ? ((ctx->flags & CAP_DEVICE_POINTERS) ? c->bw : ctx->S[S_BW]) : nullptr;
    d.bp = c->bp ? ((ctx->flags & CAP_DEVICE_POINTERS) ? c->bp : ctx->S[S_BP]) : nullptr;
    if (!(ctx->flags & CAP_DEVICE_POINTERS) && (c->bb || c->bs || c->bw || c->bp)) {
        if ((rc = ensure_std(ctx))) return rc;
        // the reference leaves bb..bp untouched outside the range: keep them
        if (c->bb) HIPCHK(hipMemcpyAsync(ctx->S[S_BB], c->bb, ctx->nstd * 8, hipMemcpyHostToDevice, ctx->stream));
        if (c->bs) HIPCHK(hipMemcpyAsync(ctx->S[S_BS], c->bs, ctx->nstd * 8, hipMemcpyHostToDevice, ctx->stream));
        if (c->bw) HIPCHK(hipMemcpyAsync(ctx->S[S_BW], c->bw, ctx->nstd * 8, hipMemcpyHostToDevice, ctx->stream));
        if (c->bp) HIPCHK(hipMemcpyAsync(ctx->S[S_BP], c->bp, ctx->nstd * 8, hipMemcpyHostToDevice, ctx->stream));
    }
    const int gb = grid_for(L), gm = move_grid(L);
    hipStream_t s = ctx->stream;
    double *const *T = ctx->T;
    HIPCHK(hipEventRecord(ctx->ev0, s));
    HIPCHK(hipMemsetAsync(ctx->ticket, 0, 64, s));
    hipLaunchKernelGGL(k_gather, dim3(gm), dim3(BLOCK), 0, s, L, (const double *)d.fi, T[T_FI]);
    const double *cs[8] = {d.ae, d.aw, d.an, d.as, d.at, d.ab, d.ap, d.su};
    const int ct[8] = {T_AE, T_AW, T_AN, T_AS, T_AT, T_AB, T_AP, T_SU};
    for (int v = 0; v < 8; ++v)
        hipLaunchKernelGGL(k_gather, dim3(gm), dim3(BLOCK), 0, s, L, cs[v], T[ct[v]]);
    // halo values of res and bn/be/bt are data for the recursions
    if (need_res_halo && d.res)
        hipLaunchKernelGGL(k_gather, dim3(gm), dim3(BLOCK), 0, s, L, (const double *)d.res, T[T_RES]);
    else
        hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_RES], L.n, 0.0);
    if (need_b_halo && d.bn) {
        hipLaunchKernelGGL(k_gather, dim3(gm), dim3(BLOCK), 0, s, L, (const double *)d.bn, T[T_BN]);
        hipLaunchKernelGGL(k_gather, dim3(gm), dim3(BLOCK), 0, s, L, (const double *)d.be, T[T_BE]);
        hipLaunchKernelGGL(k_gather, dim3(gm), dim3(BLOCK), 0, s, L, (const double *)d.bt, T[T_BT]);
    } else {
        hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_BN], L.n, 0.0);
        hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_BE], L.n, 0.0);
        hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_BT], L.n, 0.0);
    }
    HIPCHK(hipGetLastError());
    return CAP_OK;
}

static int init_state(cap_ctx *ctx, double sor, int nsw)
{
    DevState h;
    memset(&h, 0, sizeof h);
    h.sor = sor;
    h.nsw = nsw;
    h.small = (double)1e-30f;
    h.s0 = (double)1e20f;
    h.alf = 1.0; h.beto = 1.0; h.gam = 1.0;
    if (nsw <= 0) h.done = 1;
    memcpy(ctx->st_host, &h, sizeof h);
    HIPCHK(hipMemcpyAsync(ctx->st, ctx->st_host, sizeof h, hipMemcpyHostToDevice, ctx->stream));
    return CAP_OK;
}

static int read_state(cap_ctx *ctx, DevState &h)
{
    HIPCHK(hipMemcpyAsync(ctx->st_host, ctx->st, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    memcpy(&h, ctx->st_host, sizeof h);
    return CAP_OK;
}

// SIP factorisation into T_BB..T_BT for the solve range (pcg-sip.f90:68-87)
static void sip_factor(cap_ctx *ctx, const Lay &L, double alfa)
{
    double *const *T = ctx->T;
    OpSipFactor op;
    op.alfa = alfa;
    op.small = (double)1e-30f;
    WaveIO<OpSipFactor> io;
    const int in[7] = {T_AB, T_AW, T_AS, T_AP, T_AN, T_AE, T_AT};
    for (int q = 0; q < 7; ++q) io.in[q] = T[in[q]];
    io.dep[0] = T[T_BN]; io.dep[1] = T[T_BE]; io.dep[2] = T[T_BT];
    io.xo[0] = T[T_BB]; io.xo[1] = T[T_BW]; io.xo[2] = T[T_BS]; io.xo[3] = T[T_BP];
    launch_wave<OpSipFactor, true, F_NONE>(ctx, L, op, io);
    // the forward sweep reads its four factors packed: {bb, bw, bs, bp}
    const int gp = (int)std::min<size_t>((L.n + BLOCK - 1) / BLOCK, 8192);
    hipLaunchKernelGGL(k_pack4<double>, dim3(gp), dim3(BLOCK), 0, ctx->stream, L.n,
                       (const double *)T[T_BB], (const double *)T[T_BW], (const double *)T[T_BS],
                       (const double *)T[T_BP], ctx->pkd);
}

static void sip_fwd(cap_ctx *ctx, const Lay &L, const double *rhs, double *z)
{
    double *const *T = ctx->T;
    WaveIO<OpSipFwd> io;
    io.in[0] = rhs; io.in[1] = T[T_BB]; io.in[2] = T[T_BW]; io.in[3] = T[T_BS]; io.in[4] = T[T_BP];
    io.dep[0] = z;
    io.xo[0] = nullptr;
    io.pk = ctx->pkd;
    launch_wave<OpSipFwd, true, F_NONE>(ctx, L, OpSipFwd(), io);
}

// the previous PCG-SIP iteration's residual update and L1 stop test, fused
// into this iteration's forward sweep (z := forward sweep of the new res)
static void sip_fwd_res(cap_ctx *ctx, const Lay &L)
{
    double *const *T = ctx->T;
    WaveIO<OpSipFwdRes> io;
    io.in[0] = T[T_RES]; io.in[1] = T[T_BB]; io.in[2] = T[T_BW]; io.in[3] = T[T_BS];
    io.in[4] = T[T_BP]; io.in[5] = T[T_W];
    io.dep[0] = T[T_ZK];
    io.xo[0] = T[T_RES];
    io.pk = ctx->pkd;
    launch_wave<OpSipFwdRes, true, F_PCG_RESL>(ctx, L, OpSipFwdRes(), io);
}

static void sip_bwd(cap_ctx *ctx, const Lay &L, double *z)
{
    double *const *T = ctx->T;
    WaveIO<OpSipBwd> io;
    io.in[0] = z; io.in[1] = T[T_BN]; io.in[2] = T[T_BE]; io.in[3] = T[T_BT];
    io.dep[0] = z;
    io.xo[0] = nullptr;
    launch_wave<OpSipBwd, false, F_NONE>(ctx, L, OpSipBwd(), io);
}

// write back what the reference leaves in its module arrays
static int writeback(cap_ctx *ctx, const Lay &L, double *fi_host, const cap_coef *c,
                     SolveIn &d, bool res_zero_first, bool factors)
{
    const int gm = move_grid(L);
    hipStream_t s = ctx->stream;
    double *const *T = ctx->T;
    double *dsts[7] = {d.bb, d.bs, d.bw, d.bp, d.bn, d.be, d.bt};
    double *hs[7] = {c->bb, c->bs, c->bw, c->bp, c->bn, c->be, c->bt};
    const int ts[7] = {T_BB, T_BS, T_BW, T_BP, T_BN, T_BE, T_BT};
    hipLaunchKernelGGL(k_scatter, dim3(gm), dim3(BLOCK), 0, s, L, (const double *)T[T_FI], d.fi);
    if (d.res) {
        if (res_zero_first) {
            int gs = (int)std::min<size_t>((ctx->nstd + BLOCK - 1) / BLOCK, 4096);
            hipLaunchKernelGGL(k_fill, dim3(gs), dim3(BLOCK), 0, s, d.res, ctx->nstd, 0.0);
        }
        hipLaunchKernelGGL(k_scatter, dim3(gm), dim3(BLOCK), 0, s, L, (const double *)T[T_RES], d.res);
    }
    if (factors)
        for (int v = 0; v < 7; ++v)
            if (dsts[v])
                hipLaunchKernelGGL(k_scatter, dim3(gm), dim3(BLOCK), 0, s, L, (const double *)T[ts[v]], dsts[v]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev1, s));
    int rc = dev_out(ctx, fi_host, d.fi);
    if (rc) return rc;
    if (d.res && (rc = dev_out(ctx, c->res, d.res))) return rc;
    if (factors)
        for (int v = 0; v < 7; ++v)
            if (dsts[v] && (rc = dev_out(ctx, hs[v], dsts[v]))) return rc;
    HIPCHK(hipStreamSynchronize(s));
    return CAP_OK;
}

// After a failed solve (an expired wait, a tile claim out of range, a sweep
// that was not queued) the context's shared reduction and claim state may be
// out of step: a block that lost the wait may still store its partial later,
// and the tile claim counter may not match claim_base.  Wait for every queued
// kernel, then put the empty marker back into every partial slot and zero
// the ticket and the claim counter, so that the next solve on the context
// starts from a clean state.  Returns `code` (the failure being reported).
static int recover(cap_ctx *ctx, int code, const char *what)
{
    (void)hipStreamSynchronize(ctx->stream);
    double empty;
    const uint64_t bits = PARTIAL_EMPTY;
    memcpy(&empty, &bits, sizeof empty);
    hipLaunchKernelGGL(k_fill, dim3(64), dim3(BLOCK), 0, ctx->stream, ctx->partials,
                       (size_t)MAX_RED_BLOCKS * 64, empty);
    (void)hipMemsetAsync(ctx->ticket, 0, 64, ctx->stream);
    (void)hipMemsetAsync(ctx->claim, 0, 64, ctx->stream);
    ctx->claim_base = 0;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess || hipGetLastError() != hipSuccess)
        return fail(ctx, CAP_ERR_HIP, "context reset after a failed solve did not complete");
    return fail(ctx, code, what);
}

static int finish_report(cap_ctx *ctx, const DevState &h, cap_report *rep, float ms)
{
    if (ctx->bad_launch) {
        ctx->bad_launch = 0;
        return recover(ctx, CAP_ERR_ARG, "a SIP sweep was queued without its packed factors");
    }
    if (h.err == CAP_ERR_TIMEOUT)
        return recover(ctx, CAP_ERR_TIMEOUT, "wavefront dependency wait expired");
    if (h.err == CAP_ERR_ARG) return recover(ctx, CAP_ERR_ARG, "wavefront tile claim out of range");
    if (rep) {
        rep->resor = h.l >= 1 ? h.resor : h.res0;
        rep->resl = h.resl;
        rep->iters = h.conv ? h.l : h.nsw + 1;
        rep->status = h.err == 2 ? CAP_NAN : (h.conv ? CAP_OK : CAP_MAXITER);
        rep->ms = ms;
    }
    return h.err == 2 ? CAP_NAN : (h.conv ? CAP_OK : CAP_MAXITER);
}

// Run `one_iter` in batches until the device says stop.
template <class F>
static int iterate(cap_ctx *ctx, int nsw, F one_iter, DevState &h)
{
    int rc;
    int batch = 4;
    for (int l = 0; l < nsw;) {
        int nb = std::min(batch, nsw - l);
        for (int b = 0; b < nb; ++b) one_iter();
        HIPCHK(hipGetLastError());
        l += nb;
        if ((rc = read_state(ctx, h))) return rc;
        if (h.done || h.err) break;
        batch = std::min(batch * 2, 32);
    }
    return CAP_OK;
}

static Lay lay_for(cap_ctx *ctx, int klo, int khi)
{
    return make_lay(ctx->ni, ctx->nj, ctx->nk, 3, ctx->ni - 2, 3, ctx->nj - 2, klo, khi);
}

// ======================================================== the four solvers
extern "C" int cap_pcgsip(cap_ctx *ctx, double *fi, const cap_coef *c, double alfa, double sor,
                          int nsw, cap_report *rep)
{
    if (!ctx) return CAP_ERR_ARG;
    const Lay L = lay_for(ctx, 3, ctx->nk - 2);
    SolveIn d;
    int rc = prepare(ctx, L, fi, c, d, false, true);
    if (rc) return rc;
    if ((rc = init_state(ctx, sor, nsw))) return rc;
    sym_check(ctx, L);
    double *const *T = ctx->T;
    const int gb = grid_for(L);
    hipStream_t s = ctx->stream;
    const Coef7 A = coef7(ctx);
    hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_ZK], L.n, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_PK], L.n, 0.0);
    hipLaunchKernelGGL((k_resid<F_RES0>), dim3(gb), dim3(BLOCK), 0, s, L, A,
                       (const double *)T[T_SU], (const double *)T[T_FI], T[T_RES], ctx->st,
                       ctx->partials, ctx->ticket);
    sip_factor(ctx, L, alfa);
    // Iteration l's update (pcg-sip.f90:167-171) is split over iteration
    // l+1: the residual update, its L1 sum and the stop test (172-188) run at
    // the start of the forward sweep, the fi update in the pk update.  When
    // the forward sweep stops the solve, the last fi update follows the loop;
    // when the loop runs out, the last iteration's whole update does.
    int queued = 0;
    auto one = [&]() {
        const bool first = queued++ == 0;
        if (first)
            sip_fwd(ctx, L, T[T_RES], T[T_ZK]);
        else
            sip_fwd_res(ctx, L);
        WaveIO<OpSipBwdDot> io;
        io.in[0] = T[T_ZK]; io.in[1] = T[T_BN]; io.in[2] = T[T_BE]; io.in[3] = T[T_BT];
        io.in[4] = T[T_RES];
        io.dep[0] = T[T_ZK];
        io.xo[0] = nullptr;
        launch_wave<OpSipBwdDot, false, F_PCG_SK>(ctx, L, OpSipBwdDot(), io);
        if (first)
            hipLaunchKernelGGL(k_pupdate<double>, dim3(gb), dim3(BLOCK), 0, s, L,
                               (const double *)T[T_ZK], T[T_PK], ctx->st);
        else
            hipLaunchKernelGGL(k_pupdate_fi<double>, dim3(gb), dim3(BLOCK), 0, s, L, (const double *)T[T_ZK],
                               T[T_PK], T[T_FI], ctx->st);
        hipLaunchKernelGGL((k_spmv<double, 0, F_PCG_PKAPK>), dim3(grid_spmv(L)), dim3(BLOCK), 0, s, L, A,
                           (const double *)T[T_PK], T[T_W], (const double *)nullptr, ctx->st,
                           ctx->partials, ctx->ticket, ctx->asym);
    };
    DevState h;
    if ((rc = iterate(ctx, nsw, one, h))) return rc;
    if (queued > 0 && !h.err) {
        if (h.done)   // stopped by a forward sweep: its iteration's fi update
            hipLaunchKernelGGL(k_fi_last<double>, dim3(gb), dim3(BLOCK), 0, s, L, T[T_FI],
                               (const double *)T[T_PK], (const DevState *)ctx->st);
        else          // ran out of iterations: the last one's whole update
            hipLaunchKernelGGL((k_update<double, 0, F_PCG_RESL>), dim3(gb), dim3(BLOCK), 0, s, L,
                               T[T_FI], T[T_RES], (const double *)T[T_PK], (const double *)T[T_W],
                               ctx->st, ctx->partials, ctx->ticket);
        HIPCHK(hipGetLastError());
    }
    if ((rc = read_state(ctx, h))) return rc;
    if ((rc = writeback(ctx, L, fi, c, d, true, true))) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    return finish_report(ctx, h, rep, ms);
}

extern "C" int cap_sipsol(cap_ctx *ctx, double *fi, const cap_coef *c, double alfa, double sor,
                          int nsw, cap_report *rep)
{
    if (!ctx) return CAP_ERR_ARG;
    const Lay L = lay_for(ctx, 3, ctx->nk - 2);
    SolveIn d;
    int rc = prepare(ctx, L, fi, c, d, true, true);
    if (rc) return rc;
    if ((rc = init_state(ctx, sor, nsw))) return rc;
    double *const *T = ctx->T;
    const int gb = grid_for(L);
    hipStream_t s = ctx->stream;
    const Coef7 A = coef7(ctx);
    sip_factor(ctx, L, alfa);
    auto one = [&]() {
        // residual of the current fi (sipsol.f90:66-69) into T_W, then the
        // forward sweep over res (70-71), backward sweep and update (86-88)
        hipLaunchKernelGGL((k_resid<F_SIP_RES1>), dim3(gb), dim3(BLOCK), 0, s, L, A,
                           (const double *)T[T_SU], (const double *)T[T_FI], T[T_W], ctx->st,
                           ctx->partials, ctx->ticket);
        sip_fwd(ctx, L, T[T_W], T[T_RES]);
        WaveIO<OpSipBwdUpdate> io;
        io.in[0] = T[T_RES]; io.in[1] = T[T_BN]; io.in[2] = T[T_BE]; io.in[3] = T[T_BT];
        io.in[4] = T[T_FI];
        io.dep[0] = T[T_RES];
        io.xo[0] = T[T_FI];
        launch_wave<OpSipBwdUpdate, false, F_SIP_STEP>(ctx, L, OpSipBwdUpdate(), io);
    };
    DevState h;
    if ((rc = iterate(ctx, nsw, one, h))) return rc;
    if ((rc = writeback(ctx, L, fi, c, d, false, true))) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    return finish_report(ctx, h, rep, ms);
}

extern "C" int cap_iccg(cap_ctx *ctx, double *fi, const cap_coef *c, double sor, int nsw,
                        cap_report *rep)
{
    if (!ctx) return CAP_ERR_ARG;
    const Lay L = lay_for(ctx, 2, ctx->nk - 1);
    SolveIn d;
    int rc = prepare(ctx, L, fi, c, d, false, false);
    if (rc) return rc;
    if ((rc = ensure_pkb(ctx, L.n))) return rc;
    if ((rc = init_state(ctx, sor, nsw))) return rc;
    sym_check(ctx, L);
    double *const *T = ctx->T;
    const int gb = grid_for(L);
    hipStream_t s = ctx->stream;
    const Coef7 A = coef7(ctx);
    hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_ZK], L.n, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_PK], L.n, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_D], L.n, 0.0);
    hipLaunchKernelGGL((k_resid<F_RES0>), dim3(gb), dim3(BLOCK), 0, s, L, A,
                       (const double *)T[T_SU], (const double *)T[T_FI], T[T_RES], ctx->st,
                       ctx->partials, ctx->ticket);
    {   // d = 1/(ap - aw^2 d_w - as^2 d_s - ab^2 d_b)  (iccg.f90:73-81)
        WaveIO<OpIcDiag> io;
        io.in[0] = T[T_AP]; io.in[1] = T[T_AW]; io.in[2] = T[T_AS]; io.in[3] = T[T_AB];
        io.dep[0] = T[T_D];
        io.xo[0] = nullptr;
        launch_wave<OpIcDiag, true, F_NONE>(ctx, L, OpIcDiag(), io);
        // the forward sweep reads {aw, as, ab, d} packed in one 32-byte
        // element, the backward sweep {ae, an, at, d}
        const int gp = (int)std::min<size_t>((L.n + BLOCK - 1) / BLOCK, 8192);
        hipLaunchKernelGGL(k_pack4<double>, dim3(gp), dim3(BLOCK), 0, s, L.n, (const double *)T[T_AW],
                           (const double *)T[T_AS], (const double *)T[T_AB], (const double *)T[T_D],
                           ctx->pkd);
        hipLaunchKernelGGL(k_pack4<double>, dim3(gp), dim3(BLOCK), 0, s, L.n, (const double *)T[T_AE],
                           (const double *)T[T_AN], (const double *)T[T_AT], (const double *)T[T_D],
                           ctx->pkb);
    }
    // As in cap_pcgsip: iteration l's residual update and stop test run in
    // iteration l+1's forward sweep, its fi update in the p update.
    int queued = 0;
    auto one = [&]() {
        const bool first = queued++ == 0;
        if (first) {
            WaveIO<OpIcFwd> f;
            f.in[0] = T[T_RES]; f.in[1] = T[T_AW]; f.in[2] = T[T_AS]; f.in[3] = T[T_AB];
            f.in[4] = T[T_D];
            f.dep[0] = T[T_ZK];
            f.xo[0] = nullptr;
            f.pk = ctx->pkd;
            launch_wave<OpIcFwd, true, F_NONE>(ctx, L, OpIcFwd(), f);
        } else {
            WaveIO<OpIcFwdRes> f;
            f.in[0] = T[T_RES]; f.in[1] = T[T_AW]; f.in[2] = T[T_AS]; f.in[3] = T[T_AB];
            f.in[4] = T[T_D]; f.in[5] = T[T_W];
            f.dep[0] = T[T_ZK];
            f.xo[0] = T[T_RES];
            f.pk = ctx->pkd;
            launch_wave<OpIcFwdRes, true, F_PCG_RESL>(ctx, L, OpIcFwdRes(), f);
        }
        OpIcBwdDot ob;
        ob.small = (double)1e-30f;
        WaveIO<OpIcBwdDot> b;
        b.in[0] = T[T_ZK]; b.in[1] = T[T_AE]; b.in[2] = T[T_AN]; b.in[3] = T[T_AT];
        b.in[4] = T[T_D]; b.in[5] = T[T_RES];
        b.dep[0] = T[T_ZK];
        b.xo[0] = nullptr;
        b.pk = ctx->pkb;
        launch_wave<OpIcBwdDot, false, F_PCG_SK>(ctx, L, ob, b);
        if (first)
            hipLaunchKernelGGL(k_pupdate<double>, dim3(gb), dim3(BLOCK), 0, s, L,
                               (const double *)T[T_ZK], T[T_PK], ctx->st);
        else
            hipLaunchKernelGGL(k_pupdate_fi<double>, dim3(gb), dim3(BLOCK), 0, s, L, (const double *)T[T_ZK],
                               T[T_PK], T[T_FI], ctx->st);
        hipLaunchKernelGGL((k_spmv<double, 0, F_PCG_PKAPK>), dim3(grid_spmv(L)), dim3(BLOCK), 0, s, L, A,
                           (const double *)T[T_PK], T[T_W], (const double *)nullptr, ctx->st,
                           ctx->partials, ctx->ticket, ctx->asym);
    };
    DevState h;
    if ((rc = iterate(ctx, nsw, one, h))) return rc;
    if (queued > 0 && !h.err) {
        if (h.done)   // stopped by a forward sweep: its iteration's fi update
            hipLaunchKernelGGL(k_fi_last<double>, dim3(gb), dim3(BLOCK), 0, s, L, T[T_FI],
                               (const double *)T[T_PK], (const DevState *)ctx->st);
        else          // ran out of iterations: the last one's whole update
            hipLaunchKernelGGL((k_update<double, 0, F_PCG_RESL>), dim3(gb), dim3(BLOCK), 0, s, L,
                               T[T_FI], T[T_RES], (const double *)T[T_PK], (const double *)T[T_W],
                               ctx->st, ctx->partials, ctx->ticket);
        HIPCHK(hipGetLastError());
    }
    if ((rc = read_state(ctx, h))) return rc;
    if ((rc = writeback(ctx, L, fi, c, d, true, false))) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    return finish_report(ctx, h, rep, ms);
}

extern "C" int cap_cgstab_sip(cap_ctx *ctx, double *fi, const cap_coef *c, double alfa,
                              double sor, int nsw, cap_report *rep)
{
    if (!ctx) return CAP_ERR_ARG;
    const Lay L = lay_for(ctx, 2, ctx->nk - 1);
    SolveIn d;
    int rc = prepare(ctx, L, fi, c, d, false, true);
    if (rc) return rc;
    if ((rc = init_state(ctx, sor, nsw))) return rc;
    sym_check(ctx, L);
    double *const *T = ctx->T;
    const int gb = grid_for(L);
    hipStream_t s = ctx->stream;
    const Coef7 A = coef7(ctx);
    // the solver's own work arrays are zero on and around the range
    const int zero[5] = {T_ZK, T_PK, T_UK, T_VK, T_RESO};
    for (int v = 0; v < 5; ++v)
        hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[zero[v]], L.n, 0.0);
    hipLaunchKernelGGL((k_resid<F_RES0>), dim3(gb), dim3(BLOCK), 0, s, L, A,
                       (const double *)T[T_SU], (const double *)T[T_FI], T[T_RES], ctx->st,
                       ctx->partials, ctx->ticket);
    sip_factor(ctx, L, alfa);
    int queued = 0;
    auto one = [&]() {
        // bet = res.reso; om; beto (cgstab_sip.f90:106-116).  After the first
        // iteration the previous iteration's last update has summed it.
        if (queued++ == 0)
            hipLaunchKernelGGL((k_dot<F_BIC_BET, false>), dim3(gb), dim3(BLOCK), 0, s, L,
                               (const double *)T[T_RES], T[T_RESO], ctx->st, ctx->partials, ctx->ticket);
        hipLaunchKernelGGL(k_bic_p, dim3(gb), dim3(BLOCK), 0, s, L, (const double *)T[T_RES],
                           (const double *)T[T_UK], T[T_PK], ctx->st);
        sip_fwd(ctx, L, T[T_PK], T[T_ZK]);
        sip_bwd(ctx, L, T[T_ZK]);
        hipLaunchKernelGGL((k_spmv<double, 1, F_BIC_GAM>), dim3(grid_spmv(L)), dim3(BLOCK), 0, s, L, A,
                           (const double *)T[T_ZK], T[T_UK], (const double *)T[T_RESO], ctx->st,
                           ctx->partials, ctx->ticket, ctx->asym);
        hipLaunchKernelGGL((k_update<double, 1, F_NONE>), dim3(gb), dim3(BLOCK), 0, s, L, T[T_FI],
                           T[T_RES], (const double *)T[T_ZK], (const double *)T[T_UK], ctx->st,
                           ctx->partials, ctx->ticket);
        sip_fwd(ctx, L, T[T_RES], T[T_ZK]);
        sip_bwd(ctx, L, T[T_ZK]);
        hipLaunchKernelGGL((k_spmv<double, 2, F_BIC_ALF>), dim3(grid_spmv(L)), dim3(BLOCK), 0, s, L, A,
                           (const double *)T[T_ZK], T[T_VK], (const double *)T[T_RES], ctx->st,
                           ctx->partials, ctx->ticket, ctx->asym);
        hipLaunchKernelGGL(k_bic_update_bet, dim3(gb), dim3(BLOCK), 0, s, L, T[T_FI], T[T_RES],
                           (const double *)T[T_ZK], (const double *)T[T_VK], (const double *)T[T_RESO],
                           ctx->st, ctx->partials, ctx->ticket);
    };
    // reso = res before the first iteration (cgstab_sip.f90:87)
    hipLaunchKernelGGL((k_dot<F_RAW, true>), dim3(gb), dim3(BLOCK), 0, s, L,
                       (const double *)T[T_RES], T[T_RESO], ctx->st, ctx->partials, ctx->ticket);
    DevState h;
    if ((rc = iterate(ctx, nsw, one, h))) return rc;
    if ((rc = writeback(ctx, L, fi, c, d, false, true))) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    return finish_report(ctx, h, rep, ms);
}


// ------------------------------------------------------- cgstab (ILU(0))
// Marks the planes k = 2 and nkm (kk = 1 and nks) of cgstab's second backward
// sweep, which does not reach them (cgstab.f90:212): a NaN in place of ae in
// their packed element.
__global__ __launch_bounds__(BLOCK) void k_mark_end_planes(Lay L, double *__restrict__ pk)
{
    const double nan = __longlong_as_double(0x7FF8000000000001ll);
    for_cells(L, [&](const Cell &c) {
        if (c.kk == 1 || c.kk == L.nks) pk[4 * c.e] = nan;
    });
}

// Replaces `call cgstab(fi, ifi)` (cgstab.f90:1-284): BiCGStab with the
// diagonal ILU(0) preconditioner on k = 2..nkm, i, j = 3..n?mm.  The
// preconditioner is the IC(0) sweep pair of iccg with the diagonal of
// cgstab.f90:58-66; the BiCGStab kernels are cgstab_sip's.
extern "C" int cap_cgstab(cap_ctx *ctx, double *fi, const cap_coef *c, double sor, int nsw,
                          cap_report *rep)
{
    if (!ctx) return CAP_ERR_ARG;
    const Lay L = lay_for(ctx, 2, ctx->nk - 1);
    SolveIn d;
    int rc = prepare(ctx, L, fi, c, d, false, false);
    if (rc) return rc;
    if ((rc = ensure_pkb(ctx, L.n, 2))) return rc;
    if ((rc = init_state(ctx, sor, nsw))) return rc;
    sym_check(ctx, L);
    double *const *T = ctx->T;
    const int gb = grid_for(L);
    hipStream_t s = ctx->stream;
    const Coef7 A = coef7(ctx);
    // the solver's automatic arrays are zero on and around the range; BB, BW,
    // BS carry the diagonal's chained ae, an, at
    const int zero[9] = {T_ZK, T_PK, T_UK, T_VK, T_RESO, T_D, T_BB, T_BW, T_BS};
    for (int v = 0; v < 9; ++v)
        hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[zero[v]], L.n, 0.0);
    hipLaunchKernelGGL((k_resid<F_RES0>), dim3(gb), dim3(BLOCK), 0, s, L, A,
                       (const double *)T[T_SU], (const double *)T[T_FI], T[T_RES], ctx->st,
                       ctx->partials, ctx->ticket);
    {
        WaveIO<OpIluDiag> io;
        io.in[0] = T[T_AP]; io.in[1] = T[T_AW]; io.in[2] = T[T_AS]; io.in[3] = T[T_AB];
        io.in[4] = T[T_AE]; io.in[5] = T[T_AN]; io.in[6] = T[T_AT];
        io.dep[0] = T[T_D]; io.dep[1] = T[T_BB]; io.dep[2] = T[T_BW]; io.dep[3] = T[T_BS];
        io.xo[0] = nullptr;
        launch_wave<OpIluDiag, true, F_NONE>(ctx, L, OpIluDiag(), io);
        const int gp = (int)std::min<size_t>((L.n + BLOCK - 1) / BLOCK, 8192);
        double *pkb2 = ctx->pkb + 4 * L.n;
        hipLaunchKernelGGL(k_pack4<double>, dim3(gp), dim3(BLOCK), 0, s, L.n, (const double *)T[T_AW],
                           (const double *)T[T_AS], (const double *)T[T_AB], (const double *)T[T_D],
                           ctx->pkd);
        hipLaunchKernelGGL(k_pack4<double>, dim3(gp), dim3(BLOCK), 0, s, L.n, (const double *)T[T_AE],
                           (const double *)T[T_AN], (const double *)T[T_AT], (const double *)T[T_D],
                           ctx->pkb);
        hipLaunchKernelGGL(k_pack4<double>, dim3(gp), dim3(BLOCK), 0, s, L.n, (const double *)T[T_AE],
                           (const double *)T[T_AN], (const double *)T[T_AT], (const double *)T[T_D],
                           pkb2);
        hipLaunchKernelGGL(k_mark_end_planes, dim3(gb), dim3(BLOCK), 0, s, L, pkb2);
    }
    // z = M^-1 rhs: forward sweep, then the scaled backward sweep over all
    // planes (mid = false) or over k = 3..nkmm only (mid = true)
    auto apply = [&](const double *rhs, bool mid) {
        WaveIO<OpIcFwd> f;
        f.in[0] = rhs; f.in[1] = T[T_AW]; f.in[2] = T[T_AS]; f.in[3] = T[T_AB]; f.in[4] = T[T_D];
        f.dep[0] = T[T_ZK];
        f.xo[0] = nullptr;
        f.pk = ctx->pkd;
        launch_wave<OpIcFwd, true, F_NONE>(ctx, L, OpIcFwd(), f);
        if (!mid) {
            OpIcBwd ob;
            ob.small = (double)1e-30f;
            WaveIO<OpIcBwd> b;
            b.in[0] = T[T_ZK]; b.in[1] = T[T_AE]; b.in[2] = T[T_AN]; b.in[3] = T[T_AT]; b.in[4] = T[T_D];
            b.dep[0] = T[T_ZK];
            b.xo[0] = nullptr;
            b.pk = ctx->pkb;
            launch_wave<OpIcBwd, false, F_NONE>(ctx, L, ob, b);
        } else {
            OpIcBwdMid ob;
            ob.small = (double)1e-30f;
            WaveIO<OpIcBwdMid> b;
            b.in[0] = T[T_ZK]; b.in[1] = T[T_AE]; b.in[2] = T[T_AN]; b.in[3] = T[T_AT]; b.in[4] = T[T_D];
            b.dep[0] = T[T_ZK];
            b.xo[0] = nullptr;
            b.pk = ctx->pkb + 4 * L.n;
            launch_wave<OpIcBwdMid, false, F_NONE>(ctx, L, ob, b);
        }
    };
    int queued = 0;
    auto one = [&]() {
        // bet = res.reso, om, beto (cgstab.f90:94-105); after the first
        // iteration the previous iteration's last update has summed it
        if (queued++ == 0)
            hipLaunchKernelGGL((k_dot<F_BIC_BET, false>), dim3(gb), dim3(BLOCK), 0, s, L,
                               (const double *)T[T_RES], T[T_RESO], ctx->st, ctx->partials, ctx->ticket);
        hipLaunchKernelGGL(k_bic_p, dim3(gb), dim3(BLOCK), 0, s, L, (const double *)T[T_RES],
                           (const double *)T[T_UK], T[T_PK], ctx->st);
        apply(T[T_PK], false);
        hipLaunchKernelGGL((k_spmv<double, 1, F_BIC_GAM>), dim3(grid_spmv(L)), dim3(BLOCK), 0, s, L, A,
                           (const double *)T[T_ZK], T[T_UK], (const double *)T[T_RESO], ctx->st,
                           ctx->partials, ctx->ticket, ctx->asym);
        hipLaunchKernelGGL((k_update<double, 1, F_NONE>), dim3(gb), dim3(BLOCK), 0, s, L, T[T_FI],
                           T[T_RES], (const double *)T[T_ZK], (const double *)T[T_UK], ctx->st,
                           ctx->partials, ctx->ticket);
        apply(T[T_RES], true);
        hipLaunchKernelGGL((k_spmv<double, 2, F_BIC_ALF>), dim3(grid_spmv(L)), dim3(BLOCK), 0, s, L, A,
                           (const double *)T[T_ZK], T[T_VK], (const double *)T[T_RES], ctx->st,
                           ctx->partials, ctx->ticket, ctx->asym);
        hipLaunchKernelGGL(k_bic_update_bet, dim3(gb), dim3(BLOCK), 0, s, L, T[T_FI], T[T_RES],
                           (const double *)T[T_ZK], (const double *)T[T_VK], (const double *)T[T_RESO],
                           ctx->st, ctx->partials, ctx->ticket);
    };
    // reso = res before the first iteration (cgstab.f90:75)
    hipLaunchKernelGGL((k_dot<F_RAW, true>), dim3(gb), dim3(BLOCK), 0, s, L,
                       (const double *)T[T_RES], T[T_RESO], ctx->st, ctx->partials, ctx->ticket);
    DevState h;
    if ((rc = iterate(ctx, nsw, one, h))) return rc;
    if ((rc = writeback(ctx, L, fi, c, d, false, false))) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    return finish_report(ctx, h, rep, ms);
}

// --------------------------------------------------------- Jacobi PCG
// z = r/(ap + small) and sk += r z (pcg-jacobi.f90:77-92), alone (FIN =
// F_PCG_SK, the first iteration) or after the previous iteration's update
// fi += alf p, r -= alf z, sum |r| and the stop test (pcg-jacobi.f90:124-143):
// then FIN = F_JAC_RESL_SK, whose finaliser runs the stop test before it
// takes sk and bet for the next iteration.
template <bool UPDATE, int FIN>
__global__ __launch_bounds__(BLOCK) void k_jac_z(Lay L, double *__restrict__ fi,
                                                 double *__restrict__ res,
                                                 const double *__restrict__ p,
                                                 double *__restrict__ z,
                                                 const double *__restrict__ ap, DevState *st,
                                                 double *partials, unsigned *ticket)
{
    if (st->done) return;
    const double a = st->alf, small = st->small;
    Acc acc[2];
    for_cells(L, [&](const Cell &c) {
        double r = res[c.e];
        if constexpr (UPDATE) {
            fi[c.e] = fi[c.e] + a * p[c.e];
            r = r - a * z[c.e];
            res[c.e] = r;
            acc[0] += fabs(r);
        }
        const double zz = r / (ap[c.e] + small);
        z[c.e] = zz;
        acc[1] += r * zz;
    });
    if constexpr (UPDATE) {
        grid_reduce<2, FIN>(acc, blockIdx.x, gridDim.x, partials, ticket, st);
    } else {
        Acc one[1] = {acc[1]};
        grid_reduce<1, FIN>(one, blockIdx.x, gridDim.x, partials, ticket, st);
    }
}

// Replaces `call Diagonally_Preconditioned_CG(fi, ifi)` (pcg-jacobi.f90:
// 3-155): CG with the Jacobi preconditioner on k, i, j = 3..n?mm.  Each
// iteration is three passes: p = z + bet p; z = A p with (p, z); then the
// update with the next iteration's z and (r, z).
extern "C" int cap_pcg_jacobi(cap_ctx *ctx, double *fi, const cap_coef *c, double sor, int nsw,
                              cap_report *rep)
{
    if (!ctx) return CAP_ERR_ARG;
    const Lay L = lay_for(ctx, 3, ctx->nk - 2);
    SolveIn d;
    int rc = prepare(ctx, L, fi, c, d, false, false);
    if (rc) return rc;
    if ((rc = init_state(ctx, sor, nsw))) return rc;
    sym_check(ctx, L);
    double *const *T = ctx->T;
    const int gb = grid_for(L);
    hipStream_t s = ctx->stream;
    const Coef7 A = coef7(ctx);
    hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_ZK], L.n, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_PK], L.n, 0.0);
    hipLaunchKernelGGL((k_resid<F_RES0>), dim3(gb), dim3(BLOCK), 0, s, L, A,
                       (const double *)T[T_SU], (const double *)T[T_FI], T[T_RES], ctx->st,
                       ctx->partials, ctx->ticket);
    int queued = 0;
    auto one = [&]() {
        if (queued++ == 0)
            hipLaunchKernelGGL((k_jac_z<false, F_PCG_SK>), dim3(gb), dim3(BLOCK), 0, s, L, T[T_FI],
                               T[T_RES], (const double *)T[T_PK], T[T_ZK], (const double *)T[T_AP],
                               ctx->st, ctx->partials, ctx->ticket);
        hipLaunchKernelGGL(k_pupdate<double>, dim3(gb), dim3(BLOCK), 0, s, L, (const double *)T[T_ZK],
                           T[T_PK], ctx->st);
        hipLaunchKernelGGL((k_spmv<double, 0, F_PCG_PKAPK>), dim3(grid_spmv(L)), dim3(BLOCK), 0, s, L,
                           A, (const double *)T[T_PK], T[T_ZK], (const double *)nullptr, ctx->st,
                           ctx->partials, ctx->ticket, ctx->asym);
        hipLaunchKernelGGL((k_jac_z<true, F_JAC_RESL_SK>), dim3(gb), dim3(BLOCK), 0, s, L, T[T_FI],
                           T[T_RES], (const double *)T[T_PK], T[T_ZK], (const double *)T[T_AP],
                           ctx->st, ctx->partials, ctx->ticket);
    };
    DevState h;
    if ((rc = iterate(ctx, nsw, one, h))) return rc;
    if ((rc = writeback(ctx, L, fi, c, d, true, false))) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    return finish_report(ctx, h, rep, ms);
}


// ============================================ mixed precision (FP32 inner)
__global__ __launch_bounds__(BLOCK) void k_to_f32(const double *__restrict__ a,
                                                  float *__restrict__ b, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK)
        b[i] = (float)a[i];
}

__global__ __launch_bounds__(BLOCK) void k_fillf(float *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK)
        p[i] = 0.0f;
}

// x += e on the solve range (FP64 accumulation of the FP32 correction)
__global__ __launch_bounds__(BLOCK) void k_add_corr(Lay L, double *__restrict__ x,
                                                    const float *__restrict__ e)
{
    for_cells(L, [&](const Cell &c) { x[c.e] = x[c.e] + (double)e[c.e]; });
}

static int ensure_f32(cap_ctx *ctx, size_t n)
{
    if (ctx->fcap >= n) return CAP_OK;
    for (int v = 0; v < F_COUNT; ++v)
        if (ctx->F[v]) { (void)hipFree(ctx->F[v]); ctx->F[v] = nullptr; }
    for (int v = 0; v < F_COUNT; ++v) HIPCHK(hipMalloc(&ctx->F[v], n * sizeof(float)));
    if (ctx->pkf) { (void)hipFree(ctx->pkf); ctx->pkf = nullptr; }
    HIPCHK(hipMalloc(&ctx->pkf, 4 * n * sizeof(float)));
    ctx->fcap = n;
    return CAP_OK;
}

// FP64 iterative refinement: r = b - A x in FP64 on the tile layout, then
// an FP32 PCG-SIP solve A e = r (FP32 coefficients, FP32 SIP factors taken
// from the FP64 factorisation, FP32 SpMV, FP64 dot products and norms) to
// max(inner_rtol, half the remaining reduction) of its start, x += e in FP64,
// until sum|r| / sum|r0| < sor with r the FP64 residual.  When a step leaves
// more than 3/4 of the residual it started from, or its inner solve runs
// out of its 4 (nis + njs + nks) iterations, the solve ends in FP64
// (cap_pcgsip from the current fi, to the same absolute residual).
extern "C" int cap_pcgsip_mixed(cap_ctx *ctx, double *fi, const cap_coef *c, double alfa,
                                double sor, int nsw, double inner_rtol, cap_report *rep)
{
    if (!ctx || !(inner_rtol > 0.0) || inner_rtol >= 1.0) return CAP_ERR_ARG;
    const Lay L = lay_for(ctx, 3, ctx->nk - 2);
    SolveIn d;
    int rc = prepare(ctx, L, fi, c, d, false, true);
    if (rc) return rc;
    if ((rc = ensure_f32(ctx, L.n))) return rc;
    if ((rc = init_state(ctx, sor, nsw))) return rc;
    sym_check(ctx, L);
    double *const *T = ctx->T;
    float *const *F = ctx->F;
    const int gb = grid_for(L);
    hipStream_t s = ctx->stream;
    const Coef7 A = coef7(ctx);
    hipLaunchKernelGGL((k_resid<F_RES0>), dim3(gb), dim3(BLOCK), 0, s, L, A,
                       (const double *)T[T_SU], (const double *)T[T_FI], T[T_RES], ctx->st,
                       ctx->partials, ctx->ticket);
    sip_factor(ctx, L, alfa);
    const int c64[14] = {T_AE, T_AW, T_AN, T_AS, T_AT, T_AB, T_AP,
                         T_BB, T_BW, T_BS, T_BP, T_BN, T_BE, T_BT};
    const int c32[14] = {F_AE, F_AW, F_AN, F_AS, F_AT, F_AB, F_AP,
                         F_BB, F_BW, F_BS, F_BP, F_BN, F_BE, F_BT};
    const int gl = (int)std::min<size_t>((L.n + BLOCK - 1) / BLOCK, 8192);
    for (int v = 0; v < 14; ++v)
        hipLaunchKernelGGL(k_to_f32, dim3(gl), dim3(BLOCK), 0, s, (const double *)T[c64[v]], F[c32[v]], L.n);
    hipLaunchKernelGGL(k_pack4<float>, dim3(gl), dim3(BLOCK), 0, s, L.n, (const float *)F[F_BB],
                       (const float *)F[F_BW], (const float *)F[F_BS], (const float *)F[F_BP], ctx->pkf);

    Coef7T<float> A32;
    A32.ae = F[F_AE]; A32.aw = F[F_AW]; A32.an = F[F_AN]; A32.as = F[F_AS];
    A32.at = F[F_AT]; A32.ab = F[F_AB]; A32.ap = F[F_AP];
    DevState h, hi;
    if ((rc = read_state(ctx, h))) return rc;
    const double res0 = h.res0;
    double resl = res0;
    int inner_total = 0, conv = 0, nan = 0, stalled = 0;
    const int cycle_cap = 4 * (L.nis + L.njs + L.nks);
    DevState *st2 = ctx->st2;
    while (true) {
        if (!(resl == resl)) { nan = 1; break; }
        if (resl / (res0 + (double)1e-30f) < sor) { conv = 1; break; }
        if (inner_total >= nsw) break;
        const double resl_start = resl;
        // inner solve from e = 0 with rhs r (FP32)
        hipLaunchKernelGGL(k_to_f32, dim3(gl), dim3(BLOCK), 0, s, (const double *)T[T_RES], F[F_R], L.n);
        hipLaunchKernelGGL(k_fillf, dim3(gl), dim3(BLOCK), 0, s, F[F_X], L.n);
        hipLaunchKernelGGL(k_fillf, dim3(gl), dim3(BLOCK), 0, s, F[F_Z], L.n);
        hipLaunchKernelGGL(k_fillf, dim3(gl), dim3(BLOCK), 0, s, F[F_P], L.n);
        memset(&hi, 0, sizeof hi);
        // an inner solve reduces its residual by inner_rtol, or by half of
        // what the outer test still needs when that is less (the last cycle
        // then stops near the requested residual instead of far below it)
        hi.sor = std::max(inner_rtol, 0.5 * sor * (res0 + (double)1e-30f) / resl);
        // an inner solve gets at most 4 (nis + njs + nks) iterations: CG on
        // these grids reduces a residual by inner_rtol in a fraction of that
        hi.nsw = std::min(nsw - inner_total, cycle_cap);
        hi.small = (double)1e-30f;
        hi.s0 = (double)1e20f;
        hi.res0 = resl;
        memcpy(ctx->st_host, &hi, sizeof hi);
        HIPCHK(hipMemcpyAsync(st2, ctx->st_host, sizeof hi, hipMemcpyHostToDevice, s));
        // the inner PCG-SIP iteration, its update split as in cap_pcgsip
        int queued = 0;
        auto one = [&]() {
            const bool first = queued++ == 0;
            if (first) {
                WaveIO<OpSipFwdT<float>> f;
                f.in[0] = F[F_R]; f.in[1] = F[F_BB]; f.in[2] = F[F_BW]; f.in[3] = F[F_BS];
                f.in[4] = F[F_BP];
                f.dep[0] = F[F_Z];
                f.xo[0] = nullptr;
                f.pk = ctx->pkf;
                launch_wave<OpSipFwdT<float>, true, F_NONE>(ctx, L, OpSipFwdT<float>(), f, st2);
            } else {
                WaveIO<OpSipFwdResT<float>> f;
                f.in[0] = F[F_R]; f.in[1] = F[F_BB]; f.in[2] = F[F_BW]; f.in[3] = F[F_BS];
                f.in[4] = F[F_BP]; f.in[5] = F[F_W];
                f.dep[0] = F[F_Z];
                f.xo[0] = F[F_R];
                f.pk = ctx->pkf;
                launch_wave<OpSipFwdResT<float>, true, F_PCG_RESL>(ctx, L, OpSipFwdResT<float>(), f, st2);
            }
            WaveIO<OpSipBwdDotT<float>> b;
            b.in[0] = F[F_Z]; b.in[1] = F[F_BN]; b.in[2] = F[F_BE]; b.in[3] = F[F_BT];
            b.in[4] = F[F_R];
            b.dep[0] = F[F_Z];
            b.xo[0] = nullptr;
            launch_wave<OpSipBwdDotT<float>, false, F_PCG_SK>(ctx, L, OpSipBwdDotT<float>(), b, st2);
            if (first)
                hipLaunchKernelGGL(k_pupdate<float>, dim3(gb), dim3(BLOCK), 0, s, L,
                                   (const float *)F[F_Z], F[F_P], st2);
            else
                hipLaunchKernelGGL(k_pupdate_fi<float>, dim3(gb), dim3(BLOCK), 0, s, L,
                                   (const float *)F[F_Z], F[F_P], F[F_X], st2);
            hipLaunchKernelGGL((k_spmv<float, 0, F_PCG_PKAPK>), dim3(grid_spmv(L)), dim3(BLOCK), 0, s, L,
                               A32, (const float *)F[F_P], F[F_W], (const float *)nullptr, st2,
                               ctx->partials, ctx->ticket, ctx->asym);
        };
        int batch = 4;
        for (int l = 0; l < hi.nsw;) {
            int nb = std::min(batch, hi.nsw - l);
            for (int q = 0; q < nb; ++q) one();
            HIPCHK(hipGetLastError());
            l += nb;
            HIPCHK(hipMemcpyAsync(ctx->st_host, st2, sizeof hi, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            memcpy(&hi, ctx->st_host, sizeof hi);
            if (hi.done || hi.err) break;
            batch = std::min(batch * 2, 32);
        }
        if (queued > 0 && !hi.err) {
            if (hi.done)
                hipLaunchKernelGGL(k_fi_last<float>, dim3(gb), dim3(BLOCK), 0, s, L, F[F_X],
                                   (const float *)F[F_P], (const DevState *)st2);
            else
                hipLaunchKernelGGL((k_update<float, 0, F_PCG_RESL>), dim3(gb), dim3(BLOCK), 0, s, L,
                                   F[F_X], F[F_R], (const float *)F[F_P], (const float *)F[F_W], st2,
                                   ctx->partials, ctx->ticket);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(ctx->st_host, st2, sizeof hi, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            memcpy(&hi, ctx->st_host, sizeof hi);
        }
        if (hi.err == CAP_ERR_TIMEOUT)
            return recover(ctx, CAP_ERR_TIMEOUT, "wavefront dependency wait expired");
        if (hi.err == CAP_ERR_ARG) return recover(ctx, CAP_ERR_ARG, "wavefront tile claim out of range");
        inner_total += hi.l;
        hipLaunchKernelGGL(k_add_corr, dim3(gb), dim3(BLOCK), 0, s, L, T[T_FI], (const float *)F[F_X]);
        hipLaunchKernelGGL((k_resid<F_RAW>), dim3(gb), dim3(BLOCK), 0, s, L, A,
                           (const double *)T[T_SU], (const double *)T[T_FI], T[T_RES], ctx->st,
                           ctx->partials, ctx->ticket);
        if ((rc = read_state(ctx, h))) return rc;
        resl = h.red[0];
        if (hi.l == 0) break;
        // every step asks for at most half of the remaining reduction; one
        // that leaves more than 3/4 of the FP64 residual, or an inner solve
        // that used its iterations without reaching its target, means the
        // FP32 solve cannot resolve this system (too ill-conditioned for FP32)
        if (!(resl / (res0 + (double)1e-30f) < sor) &&
            (resl > 0.75 * resl_start || (!hi.conv && hi.l >= hi.nsw))) {
            stalled = 1;
            break;
        }
    }
    if ((rc = writeback(ctx, L, fi, c, d, true, true))) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    int st_code = nan ? CAP_NAN : (conv ? CAP_OK : CAP_MAXITER);
    double resl_out = resl;
    if (stalled && inner_total < nsw && resl == resl) {
        // the rest of the solve in FP64: PCG-SIP from the current fi to the
        // same absolute L1 residual (its stop test is relative to its start)
        cap_report r64;
        const double sor64 = sor * (res0 + (double)1e-30f) / resl;
        rc = cap_pcgsip(ctx, fi, c, alfa, sor64, nsw - inner_total, &r64);
        if (rc < 0) return rc;
        st_code = r64.status;
        resl_out = r64.resl;
        inner_total += r64.iters;
        ms += (float)r64.ms;
    }
    if (rep) {
        rep->resor = res0;
        rep->resl = resl_out;
        rep->iters = inner_total;
        rep->status = st_code;
        rep->ms = ms;
    }
    return st_code;
}


// ============================================ flexible PCG, FP32 preconditioner
// The previous iteration's update of the flexible PCG (FP64): fi += alf p,
// r -= alf w, sum|r| and the stop test (pcg-sip.f90:167-188), and the FP32
// copy of r that the next preconditioner application reads.
__global__ __launch_bounds__(BLOCK) void k_fcg_res(Lay L, double *__restrict__ fi,
                                                   double *__restrict__ res,
                                                   const double *__restrict__ p,
                                                   const double *__restrict__ w,
                                                   float *__restrict__ r32, DevState *st,
                                                   double *partials, unsigned *ticket)
{
    if (st->done) return;
    const double a = st->alf;
    Acc acc[1];
    for_cells(L, [&](const Cell &c) {
        fi[c.e] = fi[c.e] + a * p[c.e];
        const double r = res[c.e] - a * w[c.e];
        res[c.e] = r;
        r32[c.e] = (float)r;
        acc[0] += fabs(r);
    });
    grid_reduce<1, F_PCG_RESL>(acc, blockIdx.x, gridDim.x, partials, ticket, st);
}

// p = z + bet p with z the FP32 preconditioned residual
__global__ __launch_bounds__(BLOCK) void k_pupdate_z32(Lay L, const float *__restrict__ z,
                                                       double *__restrict__ p, DevState *st)
{
    if (st->done) return;
    const double bet = st->bet;
    for_cells(L, [&](const Cell &c) { p[c.e] = (double)z[c.e] + bet * p[c.e]; });
}

// PCG-SIP (pcg-sip.f90) with the preconditioner applied in FP32: the SIP
// factors come from the FP64 factorisation rounded to FP32, the forward and
// backward sweeps run on the FP32 copy of the residual, and everything else
// -- SpMV, residual, solution, scalar products, the stop test -- stays FP64.
// The FP32 preconditioner is not exactly the same linear operator from one
// iteration to the next, so beta is the flexible (Polak-Ribiere) form
// ((r, z) - (r, z_old)) / (r_old, z_old), with z_old kept in a second FP32
// vector; with an exact preconditioner (r, z_old) vanishes and this is the
// reference's beta.  Unlike the refinement form (cap_pcgsip_mixed) the Krylov
// space is never restarted, so the solution's error stays at the FP64 solve's.
// A system that needs more than 4 (nis + njs + nks) iterations is too
// ill-conditioned for the FP32 preconditioner (its rounding then stalls the
// iteration): the solve goes on in FP64, cap_pcgsip from the current fi to
// the same absolute residual.
extern "C" int cap_pcgsip_fcg(cap_ctx *ctx, double *fi, const cap_coef *c, double alfa,
                              double sor, int nsw, cap_report *rep)
{
    if (!ctx) return CAP_ERR_ARG;
    const Lay L = lay_for(ctx, 3, ctx->nk - 2);
    SolveIn d;
    int rc = prepare(ctx, L, fi, c, d, false, true);
    if (rc) return rc;
    if ((rc = ensure_f32(ctx, L.n))) return rc;
    const int cap32 = 4 * (L.nis + L.njs + L.nks);
    const int nsw32 = std::min(nsw, cap32);
    if ((rc = init_state(ctx, sor, nsw32))) return rc;
    sym_check(ctx, L);
    double *const *T = ctx->T;
    float *const *F = ctx->F;
    const int gb = grid_for(L);
    const int gl = (int)std::min<size_t>((L.n + BLOCK - 1) / BLOCK, 8192);
    hipStream_t s = ctx->stream;
    const Coef7 A = coef7(ctx);
    hipLaunchKernelGGL(k_fill, dim3(gb), dim3(BLOCK), 0, s, T[T_PK], L.n, 0.0);
    hipLaunchKernelGGL((k_resid<F_RES0>), dim3(gb), dim3(BLOCK), 0, s, L, A,
                       (const double *)T[T_SU], (const double *)T[T_FI], T[T_RES], ctx->st,
                       ctx->partials, ctx->ticket);
    sip_factor(ctx, L, alfa);
    const int c64[7] = {T_BB, T_BW, T_BS, T_BP, T_BN, T_BE, T_BT};
    const int c32[7] = {F_BB, F_BW, F_BS, F_BP, F_BN, F_BE, F_BT};
    for (int v = 0; v < 7; ++v)
        hipLaunchKernelGGL(k_to_f32, dim3(gl), dim3(BLOCK), 0, s, (const double *)T[c64[v]], F[c32[v]], L.n);
    hipLaunchKernelGGL(k_pack4<float>, dim3(gl), dim3(BLOCK), 0, s, L.n, (const float *)F[F_BB],
                       (const float *)F[F_BW], (const float *)F[F_BS], (const float *)F[F_BP], ctx->pkf);
    hipLaunchKernelGGL(k_to_f32, dim3(gl), dim3(BLOCK), 0, s, (const double *)T[T_RES], F[F_R], L.n);
    hipLaunchKernelGGL(k_fillf, dim3(gl), dim3(BLOCK), 0, s, F[F_Z], L.n);
    hipLaunchKernelGGL(k_fillf, dim3(gl), dim3(BLOCK), 0, s, F[F_X], L.n);
    float *zn = F[F_Z], *zo = F[F_X];
    int queued = 0;
    auto one = [&]() {
        if (queued++ > 0)
            hipLaunchKernelGGL(k_fcg_res, dim3(gb), dim3(BLOCK), 0, s, L, T[T_FI], T[T_RES],
                               (const double *)T[T_PK], (const double *)T[T_W], F[F_R], ctx->st,
                               ctx->partials, ctx->ticket);
        WaveIO<OpSipFwdT<float>> fw;
        fw.in[0] = F[F_R]; fw.in[1] = F[F_BB]; fw.in[2] = F[F_BW]; fw.in[3] = F[F_BS];
        fw.in[4] = F[F_BP];
        fw.dep[0] = zn;
        fw.xo[0] = nullptr;
        fw.pk = ctx->pkf;
        launch_wave<OpSipFwdT<float>, true, F_NONE>(ctx, L, OpSipFwdT<float>(), fw);
        WaveIO<OpSipBwdDot2T<float>> bw;
        bw.in[0] = zn; bw.in[1] = F[F_BN]; bw.in[2] = F[F_BE]; bw.in[3] = F[F_BT];
        bw.in[4] = F[F_R]; bw.in[5] = zo;
        bw.dep[0] = zn;
        bw.xo[0] = nullptr;
        launch_wave<OpSipBwdDot2T<float>, false, F_FCG_SK>(ctx, L, OpSipBwdDot2T<float>(), bw);
        hipLaunchKernelGGL(k_pupdate_z32, dim3(gb), dim3(BLOCK), 0, s, L, (const float *)zn,
                           T[T_PK], ctx->st);
        hipLaunchKernelGGL((k_spmv<double, 0, F_PCG_PKAPK>), dim3(grid_spmv(L)), dim3(BLOCK), 0, s, L, A,
                           (const double *)T[T_PK], T[T_W], (const double *)nullptr, ctx->st,
                           ctx->partials, ctx->ticket, ctx->asym);
        std::swap(zn, zo);
    };
    DevState h;
    if ((rc = iterate(ctx, nsw32, one, h))) return rc;
    if (queued > 0 && !h.err) {
        // the loop ran out: the last iteration's update (a no-op when a
        // residual update has stopped the solve)
        hipLaunchKernelGGL(k_fcg_res, dim3(gb), dim3(BLOCK), 0, s, L, T[T_FI], T[T_RES],
                           (const double *)T[T_PK], (const double *)T[T_W], F[F_R], ctx->st,
                           ctx->partials, ctx->ticket);
        HIPCHK(hipGetLastError());
    }
    if ((rc = read_state(ctx, h))) return rc;
    if ((rc = writeback(ctx, L, fi, c, d, true, true))) return rc;
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    cap_report r32;
    rc = finish_report(ctx, h, &r32, ms);
    if (rc == CAP_MAXITER && nsw32 < nsw && r32.resl == r32.resl && r32.resl > 0.0) {
        // too ill-conditioned for the FP32 preconditioner: FP64 from here
        cap_report r64;
        rc = cap_pcgsip(ctx, fi, c, alfa, sor * (r32.resor + (double)1e-30f) / r32.resl,
                        nsw - h.l, &r64);
        if (rc < 0) return rc;
        r32.resl = r64.resl;
        r32.iters = h.l + r64.iters;
        r32.status = r64.status;
        r32.ms += r64.ms;
    }
    if (rep) *rep = r32;
    return r32.status;
}

// ================================================================ timing
// Measurement hook (no counterpart in the reference).  Re-launches each
// kernel of one PCG-SIP iteration `reps` times on the context's stream, on
// the tile vectors left by the last pcgsip (fp32 == 0) or pcgsip_mixed
// (fp32 != 0) solve, bracketed by HIP events; ms[q] receives the average
// duration of: 0 SIP forward sweep, 1 SIP backward sweep + dot, 2 pk update,
// 3 SpMV + dot, 4 fi/res update + L1.
extern "C" int cap_time_kernels(cap_ctx *ctx, int fp32, int reps, double *ms)
{
    if (!ctx || !ms || reps < 1 || ctx->tcap == 0 || (fp32 && ctx->fcap == 0)) return CAP_ERR_ARG;
    const Lay L = lay_for(ctx, 3, ctx->nk - 2);
    int rc;
    if ((rc = init_state(ctx, 0.0, 1 << 30))) return rc;
    sym_check(ctx, L);
    HIPCHK(hipMemsetAsync(ctx->ticket, 0, 64, ctx->stream));
    double *const *T = ctx->T;
    float *const *F = ctx->F;
    const int gb = grid_for(L);
    hipStream_t s = ctx->stream;
    const Coef7 A = coef7(ctx);
    Coef7T<float> A32;
    A32.ae = F[F_AE]; A32.aw = F[F_AW]; A32.an = F[F_AN]; A32.as = F[F_AS];
    A32.at = F[F_AT]; A32.ab = F[F_AB]; A32.ap = F[F_AP];
    for (int q = 0; q < 6; ++q) {
        if (fp32 && q == 5) {   // the fused sweep has no FP32 form
            ms[q] = 0.0;
            continue;
        }
        HIPCHK(hipEventRecord(ctx->ev0, s));
        for (int r = 0; r < reps; ++r) {
            if (!fp32) {
                if (q == 0) sip_fwd(ctx, L, T[T_RES], T[T_ZK]);
                if (q == 5) sip_fwd_res(ctx, L);
                if (q == 1) {
                    WaveIO<OpSipBwdDot> io;
                    io.in[0] = T[T_ZK]; io.in[1] = T[T_BN]; io.in[2] = T[T_BE]; io.in[3] = T[T_BT];
                    io.in[4] = T[T_RES];
                    io.dep[0] = T[T_ZK];
                    io.xo[0] = nullptr;
                                launch_wave<OpSipBwdDot, false, F_PCG_SK>(ctx, L, OpSipBwdDot(), io);
                }
                if (q == 2)
                    hipLaunchKernelGGL(k_pupdate_fi<double>, dim3(gb), dim3(BLOCK), 0, s, L,
                                       (const double *)T[T_ZK], T[T_PK], T[T_FI], ctx->st);
                if (q == 3)
                    hipLaunchKernelGGL((k_spmv<double, 0, F_PCG_PKAPK>), dim3(grid_spmv(L)), dim3(BLOCK), 0, s,
                                       L, A, (const double *)T[T_PK], T[T_W],
                                       (const double *)nullptr, ctx->st, ctx->partials, ctx->ticket, ctx->asym);
                if (q == 4)
                    hipLaunchKernelGGL((k_update<double, 0, F_PCG_RESL>), dim3(gb), dim3(BLOCK), 0,
                                       s, L, T[T_FI], T[T_RES], (const double *)T[T_PK],
                                       (const double *)T[T_W], ctx->st, ctx->partials, ctx->ticket);
            } else {
                if (q == 0) {
                    WaveIO<OpSipFwdT<float>> f;
                    f.in[0] = F[F_R]; f.in[1] = F[F_BB]; f.in[2] = F[F_BW]; f.in[3] = F[F_BS];
                    f.in[4] = F[F_BP];
                    f.dep[0] = F[F_Z];
                    f.xo[0] = nullptr;
                    f.pk = ctx->pkf;
                    launch_wave<OpSipFwdT<float>, true, F_NONE>(ctx, L, OpSipFwdT<float>(), f);
                }
                if (q == 1) {
                    WaveIO<OpSipBwdDotT<float>> b;
                    b.in[0] = F[F_Z]; b.in[1] = F[F_BN]; b.in[2] = F[F_BE]; b.in[3] = F[F_BT];
                    b.in[4] = F[F_R];
                    b.dep[0] = F[F_Z];
                    b.xo[0] = nullptr;
                            launch_wave<OpSipBwdDotT<float>, false, F_PCG_SK>(ctx, L, OpSipBwdDotT<float>(), b);
                }
                if (q == 2)
                    hipLaunchKernelGGL(k_pupdate<float>, dim3(gb), dim3(BLOCK), 0, s, L,
                                       (const float *)F[F_Z], F[F_P], ctx->st);
                if (q == 3)
                    hipLaunchKernelGGL((k_spmv<float, 0, F_PCG_PKAPK>), dim3(grid_spmv(L)), dim3(BLOCK), 0, s,
                                       L, A32, (const float *)F[F_P], F[F_W],
                                       (const float *)nullptr, ctx->st, ctx->partials, ctx->ticket, ctx->asym);
                if (q == 4)
                    hipLaunchKernelGGL((k_update<float, 0, F_PCG_RESL>), dim3(gb), dim3(BLOCK), 0,
                                       s, L, F[F_X], F[F_R], (const float *)F[F_P],
                                       (const float *)F[F_W], ctx->st, ctx->partials, ctx->ticket);
            }
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ctx->ev1, s));
        HIPCHK(hipEventSynchronize(ctx->ev1));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
        ms[q] = (double)t / reps;
    }
    DevState h;
    if ((rc = read_state(ctx, h))) return rc;
    if (h.done || h.err) return fail(ctx, CAP_ERR_ARG, "kernel timing hit a stop condition");
    return CAP_OK;
}

// ================================================================ context
extern "C" int cap_init(cap_ctx **out, int ni, int nj, int nk, int flags)
{
    if (!out || ni < 5 || nj < 5 || nk < 5) return CAP_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CAP_ERR_NODEV;
    cap_ctx *ctx = (cap_ctx *)calloc(1, sizeof(cap_ctx));
    if (!ctx) return CAP_ERR_NOMEM;
    ctx->ni = ni; ctx->nj = nj; ctx->nk = nk;
    ctx->nstd = (size_t)ni * nj * nk;
    ctx->flags = flags;
    HIPCHK(hipGetDevice(&ctx->device));
    if (cap_scratch_setup(ctx->device) != CAP_OK)
        return fail(ctx, CAP_ERR_HIP, "device address of the static scratch array not found");
    HIPCHK(hipStreamCreate(&ctx->stream));
    HIPCHK(hipEventCreate(&ctx->ev0));
    HIPCHK(hipEventCreate(&ctx->ev1));
    HIPCHK(hipMalloc(&ctx->st, sizeof(DevState)));
    HIPCHK(hipMalloc(&ctx->st2, sizeof(DevState)));
    HIPCHK(hipHostMalloc(&ctx->st_host, sizeof(DevState), hipHostMallocDefault));
    HIPCHK(hipMalloc(&ctx->partials, (size_t)MAX_RED_BLOCKS * 64 * sizeof(double)));
    // no block is known to be symmetric until a solve's k_symcheck says so
    HIPCHK(hipMalloc(&ctx->asym, 2 * (size_t)MAX_RED_BLOCKS * sizeof(int)));
    HIPCHK(hipMemset(ctx->asym, 1, 2 * (size_t)MAX_RED_BLOCKS * sizeof(int)));
    {   // every reduction slot starts empty (grid_reduce puts the marker back)
        double empty;
        const uint64_t bits = PARTIAL_EMPTY;
        memcpy(&empty, &bits, sizeof empty);
        hipLaunchKernelGGL(k_fill, dim3(64), dim3(BLOCK), 0, ctx->stream, ctx->partials,
                           (size_t)MAX_RED_BLOCKS * 64, empty);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    HIPCHK(hipMalloc(&ctx->ticket, 64));
    HIPCHK(hipMemset(ctx->ticket, 0, 64));
    HIPCHK(hipMalloc(&ctx->claim, 64));
    HIPCHK(hipMemset(ctx->claim, 0, 64));
    ctx->claim_base = 0;
    // steps a consumer tile keeps behind its producer in the wavefront
    // sweeps (CAP_WAVE_LAG in the environment overrides it, 0..64)
    ctx->wave_lag = 1;
    if (const char *lag = getenv("CAP_WAVE_LAG")) {
        const int v = atoi(lag);
        if (v >= 0 && v <= 64) ctx->wave_lag = v;
    }
    *out = ctx;
    return CAP_OK;
}

extern "C" int cap_finalize(cap_ctx *ctx)
{
    if (!ctx) return CAP_OK;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->tpool) (void)hipFree(ctx->tpool);
    for (int v = 0; v < S_COUNT; ++v) if (ctx->S[v]) (void)hipFree(ctx->S[v]);
    cap_free_extra(ctx);
    if (ctx->msg) (void)hipFree(ctx->msg);
    (void)hipFree(ctx->st);
    (void)hipFree(ctx->st2);
    (void)hipHostFree(ctx->st_host);
    (void)hipFree(ctx->partials);
    (void)hipFree(ctx->asym);
    (void)hipFree(ctx->ticket);
    (void)hipFree(ctx->claim);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamDestroy(ctx->stream);
    free(ctx);
    return CAP_OK;
}

extern "C" const char *cap_last_error(const cap_ctx *ctx)
{
    return ctx ? ctx->err : "no context";
}
