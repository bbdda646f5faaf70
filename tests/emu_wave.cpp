// emu_wave.cpp -- host emulation of the GPU wavefront schedule (test helper).
//
// Runs the tile-sheared layout and the per-step lane schedule of
// cap_wave.h sequentially: tiles in claim order, steps in order, the 256
// lanes of a step one after another, with the previous step's lane values
// standing in for the workgroup's LDS exchange.  Used by tests to check the
// schedule and the layout against the oracle bit for bit without a GPU.
#include "../cappuccino-ghostcells-mpi_amd/csrc/cap_wave.h"
#include <algorithm>
#include <vector>
#include <cstring>
#include <stdexcept>
#include <cmath>

using namespace cap;

template <class Op, bool FWD>
static double run(const Lay &L, const Op &op, const double *const *in, double *const *dep,
                  double *const *xo)
{
    constexpr int ND = Op::ND, NIN = Op::NIN;
    constexpr int NX = Op::NX > 0 ? Op::NX : 1, NR = Op::NR > 0 ? Op::NR : 1;
    double red = 0.0;
    const int NST = nsteps(L);
    // message slots start empty; a consumer must only ever read written ones
    std::vector<uint64_t> msg(msg_count(L, ND), MSG_EMPTY);
    auto rd = [&](size_t i) {
        if (msg[i] == MSG_EMPTY) throw std::runtime_error("message read before it was written");
        double v;
        std::memcpy(&v, &msg[i], 8);
        msg[i] = MSG_EMPTY;   // each slot is consumed once
        return v;
    };
    auto wr = [&](size_t i, double v) {
        if (msg[i] != MSG_EMPTY) throw std::runtime_error("message slot written twice");
        std::memcpy(&msg[i], &v, 8);
    };
    for (int c = 0; c < L.ntiles; ++c) {
        int ti, tj;
        claim_tile<FWD>(L, c, ti, tj);
        const int tile = ti * L.nTJ + tj;
        std::vector<LaneInfo> li(TCOL);
        for (int lane = 0; lane < TCOL; ++lane) lane_setup<FWD>(L, ti, tj, lane, li[lane]);
        std::vector<double> prev(TCOL * ND, 0.0), cur(TCOL * ND, 0.0);
        const size_t tb = tile_base(L, ti, tj);
        for (int t = 0; t < NST; ++t) {
            const int slot = step_slot<FWD>(L, t);
            for (int lane = 0; lane < TCOL; ++lane) {
                const LaneInfo &l = li[lane];
                int kk;
                int kind = step_kind<FWD>(L, l, t, kk);
                size_t e = elem_at(tb, slot, lane);
                double out[ND];
                if (kind == K_COMPUTE) {
                    double x[NIN], dk[ND], di[ND], dj[ND], xv[NX], r[NR];
                    for (int q = 0; q < NIN; ++q) x[q] = in[q][e];
                    for (int q = 0; q < ND; ++q) {
                        dk[q] = prev[lane * ND + q];
                        dj[q] = l.src_a == SRC_LDS ? prev[l.nb_a * ND + q]
                              : l.src_a == SRC_RING ? dep[q][ring_in_a<FWD>(L, l, kk)]
                                                    : rd(msg_in_a(L, ND, q, l, t));
                        di[q] = l.src_b == SRC_LDS ? prev[l.nb_b * ND + q]
                              : l.src_b == SRC_RING ? dep[q][ring_in_b<FWD>(L, l, kk)]
                                                    : rd(msg_in_b(L, ND, q, l, t));
                    }
                    op.compute(x, dk, di, dj, out, xv, r);
                    for (int q = 0; q < ND; ++q) dep[q][e] = out[q];
                    for (int q = 0; q < Op::NX; ++q) xo[q][e] = xv[q];
                    if (Op::NR) red += r[0];
                    for (int q = 0; q < ND; ++q) {
                        if (l.prod_a) wr(msg_out_a(L, ND, q, l, tile, t), out[q]);
                        if (l.prod_b) wr(msg_out_b(L, ND, q, l, tile, t), out[q]);
                    }
                } else if (kind == K_LOAD) {
                    for (int q = 0; q < ND; ++q) out[q] = dep[q][e];
                } else {
                    // the GPU kernel sends this lane's stores to its dummy slots
                    for (int q = 0; q < ND + Op::NX; ++q)
                        if (q >= MSG_DUMMY ||
                            msg_dummy(L, ND, tile, lane) + (size_t)q * TCOL >= msg_alloc(L, ND) ||
                            msg_dummy(L, ND, tile, lane) < msg_count(L, ND))
                            throw std::runtime_error("dummy slot out of range");
                    for (int q = 0; q < ND; ++q) out[q] = 0.0;
                }
                for (int q = 0; q < ND; ++q) cur[lane * ND + q] = out[q];
            }
            prev.swap(cur);
        }
    }
    for (size_t i = 0; i < msg.size(); ++i)
        if (msg[i] != MSG_EMPTY) throw std::runtime_error("message left unconsumed");
    return red;
}

extern "C" {

void emu_lay(int ni, int nj, int nk, int klo, int khi, int *out)
{
    Lay L = make_lay(ni, nj, nk, 3, ni - 2, 3, nj - 2, klo, khi);
    out[0] = L.nTI; out[1] = L.nTJ; out[2] = L.S; out[3] = L.ntiles;
    out[4] = (int)L.n;
}

void emu_gather(int ni, int nj, int nk, int klo, int khi, const double *std_, double *t)
{
    Lay L = make_lay(ni, nj, nk, 3, ni - 2, 3, nj - 2, klo, khi);
    for (size_t e = 0; e < L.n; ++e) {
        int ci, cj, kk;
        t[e] = decode(L, e, ci, cj, kk) ? std_[std_index(L, ci, cj, kk)] : 0.0;
    }
}

void emu_scatter(int ni, int nj, int nk, int klo, int khi, const double *t, double *std_)
{
    Lay L = make_lay(ni, nj, nk, 3, ni - 2, 3, nj - 2, klo, khi);
    for (size_t e = 0; e < L.n; ++e) {
        int ci, cj, kk;
        if (e < L.nmain && decode(L, e, ci, cj, kk) && kk >= 1 && kk <= L.nks)
            std_[std_index(L, ci, cj, kk)] = t[e];
    }
}

// k_gather / k_scatter block by block as the kernels run them (the LDS buffer
// of each block, every block's rows and slots); returns the number of
// accesses outside the arrays and of tiled elements not written exactly once
long emu_gather_blocks(int ni, int nj, int nk, int klo, int khi, const double *std_, double *t)
{
    Lay L = make_lay(ni, nj, nk, 3, ni - 2, 3, nj - 2, klo, khi);
    const size_t nstd = (size_t)ni * nj * nk;
    std::vector<int> writes(L.n, 0);
    std::vector<double> buf(MOVE_SLOTS * TCOL);
    long bad = 0;
    const int G = move_grid(L);
    for (int bid = 0; bid < G; ++bid) {
        int tile, s0;
        if (move_item(L, bid, tile, s0)) {
            const int ti = tile / L.nTJ, tj = tile % L.nTJ, ns = std::min(MOVE_SLOTS, L.S - s0);
            for (int lane = 0; lane < TCOL; ++lane) {
                const int b = lane / TW, a = lane % TW, ci = ti * TW + b, cj = tj * TW + a;
                const bool valid = col_valid(L, ci, cj);
                for (int j = 0; j < MOVE_SLOTS + TW - 1; ++j) {
                    int r, kk;
                    if (!move_step(L, s0, ns, b, a, j, r, kk)) continue;
                    const size_t q = std_index(L, ci, cj, kk);
                    if (valid && q >= nstd) { ++bad; continue; }
                    buf[r * TCOL + lane] = valid ? std_[q] : 0.0;
                }
            }
            const size_t tb = tile_base(L, ti, tj);
            for (int lane = 0; lane < TCOL; ++lane)
                for (int r = 0; r < ns; ++r) {
                    const size_t e = elem_at(tb, s0 + r, lane);
                    if (e >= L.nmain) { ++bad; continue; }
                    t[e] = buf[r * TCOL + lane];
                    ++writes[e];
                }
        }
        for (size_t e = L.nmain + (size_t)bid * 256; e < L.n; e += (size_t)G * 256)
            for (size_t q = e; q < e + 256 && q < L.n; ++q) {
                int ci, cj, kk;
                t[q] = decode(L, q, ci, cj, kk) ? std_[std_index(L, ci, cj, kk)] : 0.0;
                ++writes[q];
            }
    }
    for (size_t e = 0; e < L.n; ++e) bad += writes[e] != 1;
    return bad;
}

long emu_scatter_blocks(int ni, int nj, int nk, int klo, int khi, const double *t, double *std_)
{
    Lay L = make_lay(ni, nj, nk, 3, ni - 2, 3, nj - 2, klo, khi);
    const size_t nstd = (size_t)ni * nj * nk;
    std::vector<int> writes(nstd, 0);
    std::vector<double> buf(MOVE_SLOTS * TCOL);
    long bad = 0;
    const int G = move_grid(L);
    for (int bid = 0; bid < G; ++bid) {
        int tile, s0;
        if (!move_item(L, bid, tile, s0)) continue;
        const int ti = tile / L.nTJ, tj = tile % L.nTJ, ns = std::min(MOVE_SLOTS, L.S - s0);
        const size_t tb = tile_base(L, ti, tj);
        for (int lane = 0; lane < TCOL; ++lane)
            for (int r = 0; r < ns; ++r) buf[r * TCOL + lane] = t[elem_at(tb, s0 + r, lane)];
        for (int lane = 0; lane < TCOL; ++lane) {
            const int b = lane / TW, a = lane % TW, ci = ti * TW + b, cj = tj * TW + a;
            if (!col_valid(L, ci, cj)) continue;
            for (int j = 0; j < MOVE_SLOTS + TW - 1; ++j) {
                int r, kk;
                if (!move_step(L, s0, ns, b, a, j, r, kk) || kk < 1 || kk > L.nks) continue;
                const size_t q = std_index(L, ci, cj, kk);
                if (q >= nstd) { ++bad; continue; }
                std_[q] = buf[r * TCOL + lane];
                ++writes[q];
            }
        }
    }
    for (size_t q = 0; q < nstd; ++q) bad += writes[q] > 1;
    return bad;
}

// op: 0 sip factor, 1 sip fwd, 2 sip bwd+dot, 3 ic diag, 4 ic fwd, 5 ic bwd+dot,
//     6 sip bwd (no dot), 7 sipsol bwd+update
double emu_wave(int op, int ni, int nj, int nk, int klo, int khi, double alfa,
                const double *const *in, double *const *dep, double *const *xo)
{
    Lay L = make_lay(ni, nj, nk, 3, ni - 2, 3, nj - 2, klo, khi);
    const double small = (double)1e-30f;
    try {
    switch (op) {
    case 0: { OpSipFactor o; o.alfa = alfa; o.small = small; return run<OpSipFactor, true>(L, o, in, dep, xo); }
    case 1: return run<OpSipFwd, true>(L, OpSipFwd(), in, dep, xo);
    case 2: return run<OpSipBwdDot, false>(L, OpSipBwdDot(), in, dep, xo);
    case 3: return run<OpIcDiag, true>(L, OpIcDiag(), in, dep, xo);
    case 4: return run<OpIcFwd, true>(L, OpIcFwd(), in, dep, xo);
    case 5: { OpIcBwdDot o; o.small = small; return run<OpIcBwdDot, false>(L, o, in, dep, xo); }
    case 6: return run<OpSipBwd, false>(L, OpSipBwd(), in, dep, xo);
    case 7: return run<OpSipBwdUpdate, false>(L, OpSipBwdUpdate(), in, dep, xo);
    }
    } catch (const std::exception &) {
        return std::nan("");
    }
    return 0.0;
}

// The flat kernels' walk (SlotWalk over per-block slot ranges, neighbours_at)
// checked against decode() / neighbours(): every solve cell visited once per
// lane, with the same plane and neighbours.  Returns the number of mismatches.
int emu_check_walk(int ni, int nj, int nk, int klo, int khi, int nblocks)
{
    Lay L = make_lay(ni, nj, nk, 3, ni - 2, 3, nj - 2, klo, khi);
    const size_t nts = (size_t)L.ntiles * L.S;
    std::vector<int> seen(L.nmain, 0);
    int bad = 0;
    const size_t per = (nts + nblocks - 1) / nblocks;
    for (int blk = 0; blk < nblocks; ++blk) {
        size_t lo = (size_t)blk * per, hi = lo + per < nts ? lo + per : nts;
        if (lo > hi) lo = hi;
        if (lo >= hi) continue;
        for (int lane = 0; lane < TCOL; ++lane) {
            SlotWalk w;
            w.start(L, lo, lane);
            for (size_t ts = lo; ts < hi; ++ts) {
                const size_t e = elem_ts(ts, w.slot, lane);
                int ci, cj, kk;
                const bool valid = decode(L, e, ci, cj, kk);
                const bool in = valid && kk >= 1 && kk <= L.nks;
                if (w.cell(L) != in) ++bad;
                if (w.cell(L)) {
                    seen[e]++;
                    if (w.kk != kk || w.ti * TW + w.b != ci || w.tj * TW + w.a != cj ||
                        w.slot != (int)(ts % (size_t)L.S)) ++bad;
                    Nb r = neighbours(L, w.ti, w.tj, w.slot, w.b, w.a, w.kk);
                    Nb f = neighbours_at(L, w.ti, w.tj, e, w.slot, w.b, w.a, w.kk);
                    if (r.e != f.e || r.w != f.w || r.n != f.n || r.s != f.s || r.t != f.t ||
                        r.bt != f.bt) ++bad;
                }
                w.next(L);
            }
        }
    }
    for (size_t e = 0; e < L.nmain; ++e) {
        int ci, cj, kk;
        const bool in = decode(L, e, ci, cj, kk) && kk >= 1 && kk <= L.nks;
        if (seen[e] != (in ? 1 : 0)) ++bad;
    }
    return bad;
}

// The SpMV's skewed walk (for_cells_skewed: block b, eighth b % 8 of the
// slots of tiles b/8, b/8 + G, ..., each tile's range rotated by
// skew_slot0): every solve cell visited once, with the plane and neighbours
// of decode() / neighbours(), and every neighbour in another tile lying in
// the slot that tile's block of the same eighth visits at the same step.
// Returns the number of mismatches.
int emu_check_skewed_walk(int ni, int nj, int nk, int klo, int khi, int nblocks)
{
    Lay L = make_lay(ni, nj, nk, 3, ni - 2, 3, nj - 2, klo, khi);
    std::vector<int> seen(L.nmain, 0);
    int bad = 0;
    const int G = nblocks >> 3;
    for (int blk = 0; blk < nblocks; ++blk) {
        const int c = blk & 7;
        const int s0 = c * L.S / 8, s1 = (c + 1) * L.S / 8;
        if (G == 0 || s0 >= s1) continue;
        for (int t = blk >> 3; t < L.ntiles; t += G) {
            for (int lane = 0; lane < TCOL; ++lane) {
                TileWalk w;
                w.start(L, t, skew_slot0(L, t, s0), lane);
                for (int s = s0; s < s1; ++s) {
                    const size_t e = w.elem(lane);
                    int ci, cj, kk;
                    const bool valid = decode(L, e, ci, cj, kk);
                    const bool in = valid && kk >= 1 && kk <= L.nks;
                    if (w.cell(L) != in) ++bad;
                    if (w.cell(L)) {
                        seen[e]++;
                        if (w.kk != kk || w.ti * TW + w.b != ci || w.tj * TW + w.a != cj) ++bad;
                        Nb r = neighbours(L, w.ti, w.tj, w.slot, w.b, w.a, w.kk);
                        Nb f = neighbours_at(L, w.ti, w.tj, e, w.slot, w.b, w.a, w.kk);
                        if (r.e != f.e || r.w != f.w || r.n != f.n || r.s != f.s || r.t != f.t ||
                            r.bt != f.bt) ++bad;
                        for (size_t q : {r.e, r.w, r.n, r.s}) {
                            if (q >= L.nmain) continue;   // ring column
                            const int tq = (int)(q / ((size_t)L.S * TCOL));
                            if (tq == t) continue;
                            const int sq = (int)((q / TCOL) % (size_t)L.S);
                            if (posmod(sq - skew_slot0(L, tq, s0), L.S) != s - s0) ++bad;
                        }
                    }
                    w.next(L);
                }
            }
        }
    }
    for (size_t e = 0; e < L.nmain; ++e) {
        int ci, cj, kk;
        const bool in = decode(L, e, ci, cj, kk) && kk >= 1 && kk <= L.nks;
        if (seen[e] != (in ? 1 : 0)) ++bad;
    }
    return bad;
}

// z = A p on interior cells with the operand order of pcg-sip.f90:147-149
void emu_spmv(int ni, int nj, int nk, int klo, int khi, const double *const *A,
              const double *p, double *z)
{
    Lay L = make_lay(ni, nj, nk, 3, ni - 2, 3, nj - 2, klo, khi);
    // A = {ae, aw, an, as, at, ab, ap}
    for (size_t e = 0; e < L.n; ++e) {
        int ci, cj, kk;
        if (e >= L.nmain || !decode(L, e, ci, cj, kk) || kk < 1 || kk > L.nks) continue;
        int ti = ci / TW, b = ci % TW, tj = cj / TW, a = cj % TW;
        int slot = posmod(kk + a + b, L.S);
        Nb nb = neighbours(L, ti, tj, slot, b, a, kk);
        z[e] = A[6][e] * p[e] - A[0][e] * p[nb.e] - A[1][e] * p[nb.w] - A[2][e] * p[nb.n] -
               A[3][e] * p[nb.s] - A[4][e] * p[nb.t] - A[5][e] * p[nb.bt];
    }
}

}  // extern "C"
