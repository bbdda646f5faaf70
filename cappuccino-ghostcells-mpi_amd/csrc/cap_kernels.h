// cap_kernels.h -- device-side building blocks shared by the solver kernels:
// device scalar state, deterministic grid reductions and the finalisers that
// turn reduced sums into the Krylov scalars of the reference solvers.
#pragma once
#include <hip/hip_runtime.h>
#include "cap_wave.h"
#include "cap_amd.h"

namespace cap {

constexpr int BLOCK = 256;
constexpr int MAX_RED_BLOCKS = 4096;

// Scalars of one solve, resident on the device.  The reference keeps them as
// locals of pcgsip / iccg / cgstab_sip / sipsol; the names follow it.
struct DevState {
    double res0, resor, resl, rsm;
    double s0, sk, bet, pkapk, alf;          // pcg-sip.f90:34-35, iccg.f90:37-38
    double beto, gam, om, ukreso, vres, vv;  // cgstab_sip.f90:36-37
    double sor, small;
    double red[4];                           // raw sums (tests / mixed solver)
    int l;          // iterations completed
    int nsw;
    int done;       // 1: later kernels of this solve do nothing
    int conv;       // 1: stopped on the relative-residual test
    int stop;       // sipsol: stop after this sweep's update
    int err;        // CAP_ERR_TIMEOUT from a wavefront wait, CAP_NAN
    int pad[2];
};

enum Fin {
    F_NONE = 0,
    F_RES0,        // res0 = sum|r|                          pcg-sip.f90:56
    F_PCG_SK,      // sk, bet = sk/s0                        pcg-sip.f90:118,127
    F_PCG_PKAPK,   // pkapk, alf = sk/pkapk                  pcg-sip.f90:150,157
    F_PCG_RESL,    // resl, s0 = sk, stop test               pcg-sip.f90:171-188
    F_SIP_RES1,    // RES1, RESOR, RSM                       sipsol.f90:69,77-78
    F_SIP_STEP,    // end of a sipsol sweep                  sipsol.f90:96
    F_BIC_BET,     // bet, om, beto                          cgstab_sip.f90:111-116
    F_BIC_GAM,     // ukreso, gam                            cgstab_sip.f90:175-179
    F_BIC_ALF,     // vres, vv, alf                          cgstab_sip.f90:239-245
    F_BIC_RESL,    // resl, stop test                        cgstab_sip.f90:256-266
    F_RAW,         // keep the sums in red[]
    F_BIC_RESL_BET,  // F_BIC_RESL on v[0]; unless that stops, F_BIC_BET on v[1]
    F_JAC_RESL_SK,   // F_PCG_RESL on v[0]; unless that stops, F_PCG_SK on v[1]
    F_FCG_SK,        // flexible PCG: sk = (r, z), bet = ((r, z) - (r, z_old)) / s0
};

__device__ __forceinline__ void fin_resl(DevState *st, double resl)
{
    st->resl = resl;
    st->l += 1;
    if (st->l == 1) st->resor = st->res0;
    st->rsm = resl / (st->resor + st->small);
    if (!(resl == resl) || resl > 1.0e300) { st->err = 2; st->done = 1; return; }
    if (st->rsm < st->sor) { st->conv = 1; st->done = 1; }
    else if (st->l >= st->nsw) st->done = 1;
}

template <int FIN>
__device__ __forceinline__ void finalize(DevState *st, const double *v)
{
    if constexpr (FIN == F_RES0) {
        st->res0 = v[0];
    } else if constexpr (FIN == F_PCG_SK) {
        st->sk = v[0];
        st->bet = v[0] / st->s0;
    } else if constexpr (FIN == F_PCG_PKAPK) {
        st->pkapk = v[0];
        st->alf = st->sk / v[0];
    } else if constexpr (FIN == F_PCG_RESL) {
        st->s0 = st->sk;
        fin_resl(st, v[0]);
    } else if constexpr (FIN == F_SIP_RES1) {
        st->resl = v[0];
        if (st->l == 0) st->resor = v[0];
        st->rsm = v[0] / (st->resor + st->small);
        st->stop = st->rsm < st->sor;
        if (!(v[0] == v[0]) || v[0] > 1.0e300) st->err = 2;
    } else if constexpr (FIN == F_SIP_STEP) {
        st->l += 1;
        if (st->err) st->done = 1;
        else if (st->stop) { st->conv = 1; st->done = 1; }
        else if (st->l >= st->nsw) st->done = 1;
    } else if constexpr (FIN == F_BIC_BET) {
        st->om = v[0] * st->gam / (st->alf * st->beto + st->small);
        st->beto = v[0];
    } else if constexpr (FIN == F_BIC_GAM) {
        st->ukreso = v[0];
        st->gam = st->beto / v[0];
    } else if constexpr (FIN == F_BIC_ALF) {
        st->vres = v[0];
        st->vv = v[1];
        st->alf = v[0] / (v[1] + st->small);
    } else if constexpr (FIN == F_BIC_RESL) {
        fin_resl(st, v[0]);
    } else if constexpr (FIN == F_RAW) {
        for (int q = 0; q < 4; ++q) st->red[q] = v[q];
    } else if constexpr (FIN == F_JAC_RESL_SK) {
        // the stop test (pcg-jacobi.f90:136-143), then the next iteration's
        // sk and bet (77-92), summed in the same pass from the new residual
        st->s0 = st->sk;
        fin_resl(st, v[0]);
        if (!st->done) {
            st->sk = v[1];
            st->bet = v[1] / st->s0;
        }
    } else if constexpr (FIN == F_FCG_SK) {
        st->sk = v[0];
        st->bet = (v[0] - v[1]) / st->s0;
    } else if constexpr (FIN == F_BIC_RESL_BET) {
        // the stop test, then (when the solve goes on) the next iteration's
        // bet, om, beto, which its k_dot would compute from the same residual
        fin_resl(st, v[0]);
        if (!st->done) {
            st->om = v[1] * st->gam / (st->alf * st->beto + st->small);
            st->beto = v[1];
        }
    }
}

__device__ __forceinline__ double wave_sum(double x)
{
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Sum as an unevaluated pair hi + lo: every addition is Knuth's TwoSum (its
// rounding error is carried in lo), so a reduction of ~1e7 terms keeps about
// 30 more bits than a plain running sum and rounds to (nearly) the exact sum
// whatever the reduction order.  The solvers' scalar products therefore do not
// depend on the order in which the GPU adds them.  (Code is compiled without
// FMA contraction, which TwoSum needs.)
struct Acc {
    double hi = 0.0, lo = 0.0;
    __device__ __forceinline__ Acc &operator+=(double x)
    {
        const double s = hi + x;
        const double bp = s - hi;
        lo += (hi - (s - bp)) + (x - bp);
        hi = s;
        return *this;
    }
    __device__ __forceinline__ Acc &operator+=(const Acc &o)
    {
        *this += o.hi;
        lo += o.lo;
        return *this;
    }
    __device__ __forceinline__ double value() const { return hi + lo; }
};

__device__ __forceinline__ Acc wave_sum(Acc x)
{
    for (int off = 32; off > 0; off >>= 1) {
        Acc o;
        o.hi = __shfl_xor(x.hi, off, 64);
        o.lo = __shfl_xor(x.lo, off, 64);
        x += o;
    }
    return x;
}

// Deterministic two-level reduction of NV per-thread sums over the grid.
// Each block stores its sums (hi, lo) in partials[(slot*NV + q)*2 + {0,1}]
// (slot fixed per block) with device-coherent stores; the last block to take
// a ticket adds the partials in slot order and runs the finaliser on the
// rounded totals.  No fence is used (on this GPU a device-scope release writes
// back the whole L2): every partial slot holds an empty marker (PARTIAL_EMPTY)
// until its block's store arrives, the last block waits on each slot until it
// is written and puts the marker back.
constexpr uint64_t PARTIAL_EMPTY = 0x7FF4C0FFEEC0FFEEull;
#define CAP_RED_SPIN (1 << 20)

template <int NV, int FIN>
__device__ void grid_reduce(const Acc (&v)[NV], int slot, int nslots, double *partials,
                            unsigned *ticket, DevState *st)
{
    __shared__ double shh[NV][BLOCK / 64], shl[NV][BLOCK / 64];
    __shared__ int last;
    auto sh_put = [&](int q, int k, const Acc &x) { shh[q][k] = x.hi; shl[q][k] = x.lo; };
    auto sh_sum = [&](int q) {
        Acc b;
        b.hi = shh[q][0];
        b.lo = shl[q][0];
        for (int k = 1; k < BLOCK / 64; ++k) {
            Acc o;
            o.hi = shh[q][k];
            o.lo = shl[q][k];
            b += o;
        }
        return b;
    };
    // only the first BLOCK threads take part (a larger block may call this
    // with every thread; the others only join the barriers)
    const int w = threadIdx.x >> 6;
    const bool part = threadIdx.x < BLOCK;
    const double empty = __longlong_as_double((long long)PARTIAL_EMPTY);
    for (int q = 0; q < NV; ++q) {
        Acc x = wave_sum(v[q]);
        if (part && (threadIdx.x & 63) == 0) sh_put(q, w, x);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < NV; ++q) {
            const Acc b = sh_sum(q);
            double *pp = &partials[((size_t)slot * NV + q) * 2];
            __hip_atomic_store(pp, b.hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(pp + 1, b.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    Acc acc[NV];
    bool lost = false;
    for (int s = threadIdx.x; part && s < nslots; s += BLOCK)
        for (int q = 0; q < NV; ++q) {
            double *pp = &partials[((size_t)s * NV + q) * 2];
            double x[2];
            for (int h = 0; h < 2; ++h) {
                x[h] = __hip_atomic_load(pp + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                for (int it = 0; it < CAP_RED_SPIN &&
                                 (uint64_t)__double_as_longlong(x[h]) == PARTIAL_EMPTY; ++it) {
                    __builtin_amdgcn_s_sleep(1);
                    x[h] = __hip_atomic_load(pp + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                lost |= (uint64_t)__double_as_longlong(x[h]) == PARTIAL_EMPTY;
                __hip_atomic_store(pp + h, empty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            Acc o;
            o.hi = x[0];
            o.lo = x[1];
            acc[q] += o;
        }
    if (lost) {
        st->err = CAP_ERR_TIMEOUT;
        st->done = 1;
    }
    __syncthreads();
    for (int q = 0; q < NV; ++q) {
        Acc x = wave_sum(acc[q]);
        if (part && (threadIdx.x & 63) == 0) sh_put(q, w, x);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot[NV > 4 ? NV : 4] = {0.0, 0.0, 0.0, 0.0};
        for (int q = 0; q < NV; ++q) {
            tot[q] = sh_sum(q).value();
        }
        finalize<FIN>(st, tot);
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// element / cell decode for kernels that sweep the tiled vectors
struct Cell {
    size_t e;
    int ti, tj, slot, b, a, kk;
    bool in;   // interior cell (solve range)
};

__device__ __forceinline__ Cell cell_of(const Lay &L, size_t ts, int lane)
{
    Cell c;
    const int t = (int)(ts / (size_t)L.S);
    c.slot = (int)(ts % (size_t)L.S);
    c.ti = t / L.nTJ;
    c.tj = t % L.nTJ;
    c.b = lane / TW;
    c.a = lane % TW;
    c.kk = posmod(c.slot - c.a - c.b, L.S);
    c.e = elem_ts(ts, c.slot, lane);
    const int ci = c.ti * TW + c.b, cj = c.tj * TW + c.a;
    c.in = col_valid(L, ci, cj) && c.kk >= 1 && c.kk <= L.nks;
    return c;
}

// Blocks of the flat kernels come in groups of eight (the launch grid is a
// multiple of 8): block b takes the (b % 8)-th eighth of the slots of tiles
// b/8, b/8 + G, ... (G = gridDim / 8).  Blocks are spread over the eight XCDs
// by b % 8, so every tile's same eighth of the planes is handled on the same
// XCD: a stencil's reads of a neighbouring tile's edge columns (eight-byte
// pieces of separate cache lines for the j-edges) find the lines that the
// neighbour's block has pulled into that XCD's L2.
template <class F>
__device__ __forceinline__ void for_cells(const Lay &L, F &&f)
{
    const int c = blockIdx.x & 7, G = (int)(gridDim.x >> 3);
    const int s0 = c * L.S / 8, s1 = (c + 1) * L.S / 8;
    if (G == 0 || s0 >= s1) return;
    const int lane = threadIdx.x;
    for (int t = (int)(blockIdx.x >> 3); t < L.ntiles; t += G) {
        const size_t lo = (size_t)t * L.S + s0;
        SlotWalk w;
        w.start(L, lo, lane);
        for (int s = s0; s < s1; ++s) {
            if (w.cell(L)) {
                Cell cl;
                cl.e = elem_ts(lo + (size_t)(s - s0), w.slot, lane);
                cl.ti = w.ti; cl.tj = w.tj; cl.slot = w.slot; cl.b = w.b; cl.a = w.a; cl.kk = w.kk;
                cl.in = true;
                f(cl);
            }
            w.next(L);
        }
    }
}

// The SpMV's walk: blocks as in for_cells (block b, eighth b % 8 of the
// slots of tiles b/8, b/8 + G, ...), each tile's range rotated by
// skew_slot0 (TileWalk), so that the edge rows a stencil reads from the
// neighbouring tiles are the ones those tiles' blocks work on at the same
// step and are still in the XCD's L2.  A block visits the same cells as in
// for_cells, in another order, so kernels that share per-block results (the
// symmetry flags of k_symcheck) must both walk this way.
template <class F>
__device__ __forceinline__ void for_cells_skewed(const Lay &L, F &&f)
{
    const int c = blockIdx.x & 7, G = (int)(gridDim.x >> 3);
    const int s0 = c * L.S / 8, s1 = (c + 1) * L.S / 8;
    if (G == 0 || s0 >= s1) return;
    const int lane = threadIdx.x;
    for (int t = (int)(blockIdx.x >> 3); t < L.ntiles; t += G) {
        TileWalk w;
        w.start(L, t, skew_slot0(L, t, s0), lane);
        for (int s = s0; s < s1; ++s) {
            if (w.cell(L)) {
                Cell cl;
                cl.e = w.elem(lane);
                cl.ti = w.ti; cl.tj = w.tj; cl.slot = w.slot; cl.b = w.b; cl.a = w.a; cl.kk = w.kk;
                cl.in = true;
                f(cl);
            }
            w.next(L);
        }
    }
}

// for_cells_skewed with the memory access of each cell split from its
// arithmetic: load(cell, valid) reads what the cell needs into a value of
// type T (a lane with no cell at that slot, valid false, reads only its own
// element, which is always in the vector), compute(cell, values) does the
// rest.  The walk issues the loads of the lane's next slot before it
// computes the current one, alternating two value buffers (no copy between
// them, so no wait for a load before its cell is computed), and each wave
// keeps two slots' loads in flight.  The cells are computed in
// for_cells_skewed's order.
template <class T, class LD, class CP>
__device__ __forceinline__ void for_cells_pipelined(const Lay &L, LD &&load, CP &&compute)
{
    const int c = blockIdx.x & 7, G = (int)(gridDim.x >> 3);
    const int s0 = c * L.S / 8, s1 = (c + 1) * L.S / 8;
    if (G == 0 || s0 >= s1) return;
    const int lane = threadIdx.x;
    for (int t = (int)(blockIdx.x >> 3); t < L.ntiles; t += G) {
        TileWalk w;
        w.start(L, t, skew_slot0(L, t, s0), lane);
        // the cell at the walk's current slot (past the range it is the
        // tile's next slot, still an element of the vector)
        auto cell_here = [&]() {
            Cell cl;
            cl.e = w.elem(lane);
            cl.ti = w.ti; cl.tj = w.tj; cl.slot = w.slot; cl.b = w.b; cl.a = w.a; cl.kk = w.kk;
            cl.in = true;
            return cl;
        };
        Cell ca = cell_here();
        bool ha = w.cell(L);
        T va = load(ca, ha);
        w.next(L);
        for (int s = s0 + 1;; s += 2) {
            const bool in_b = s < s1;
            const Cell cb = cell_here();
            const bool hb = in_b && w.cell(L);
            const T vb = load(cb, hb);
            w.next(L);
            if (ha) compute(ca, va);
            if (!in_b) break;
            const bool in_a = s + 1 < s1;
            ca = cell_here();
            ha = in_a && w.cell(L);
            va = load(ca, ha);
            w.next(L);
            if (hb) compute(cb, vb);
            if (!in_a) break;
        }
    }
}

}  // namespace cap
