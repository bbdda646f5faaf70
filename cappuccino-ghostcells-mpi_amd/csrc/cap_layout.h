// cap_layout.h -- tile-sheared layout of the solver's work vectors.
//
// The reference keeps every field in the padded j-fastest layout
//   ijk = (k-1)*ni*nj + (i-1)*nj + j         (setind.f90:34-40)
// and its SIP / IC(0) sweeps are recursions over (i-1, j-1, k-1) (forward) or
// (i+1, j+1, k+1) (backward), e.g. pcg-sip.f90:98-121.  On the GPU the sweeps
// run as a hyperplane wavefront.  In a hyperplane the cells of one cache line
// belong to different planes, so the solver keeps its work vectors (fi, res,
// zk, pk, ...) and its copies of the coefficients in a layout built for the
// wavefront:
//
//  * the columns (i,j) of the solve box are grouped into tiles of 16 (j) x
//    16 (i) columns;
//  * inside a tile, element (slot s, b, a) holds column (i0+b, j0+a) at the
//    k-plane kk = (s - a - b) mod S, where S = nks + 2 k-planes (the solve
//    range plus one halo plane at each end);
//  * element address = ((ti*nTJ + tj)*S + s)*256 + b*16 + a;
//  * the four halo columns-rows around the solve box (i = ia-1, i = ib+1,
//    j = ja-1, j = jb+1; all S planes) live in four "ring" arrays after the
//    tiles, so a 256^3 box is exactly 16 x 16 tiles.
//
// A wavefront lane (a,b) handles its column; at step t it works on
// kk = t - a - b, so every lane of the tile touches slot t mod S at step t:
// 4 waves x 64 lanes read 256 contiguous doubles.
//
// Everything here is plain index arithmetic shared by the GPU kernels and the
// host-side schedule check.
#pragma once
#include <stdint.h>
#include <stddef.h>

#if defined(__HIPCC__)
#define CAP_HD __host__ __device__ __forceinline__
#else
#define CAP_HD inline
#endif

namespace cap {

constexpr int TW = 16;          // tile width in a (j) and b (i)
constexpr int TCOL = TW * TW;   // columns per tile = lanes per workgroup

struct Lay {
    int ni, nj, nk, nij;      // padded extents of the reference arrays
    int ia, ja, ka;           // first solve cell (1-based)
    int nis, njs, nks;        // solve-box extents
    int nTI, nTJ;             // tiles along i and j
    int S;                    // slots per tile = nks + 2
    int ntiles;
    size_t nmain;             // tile elements
    size_t ring[4];           // ring arrays: i=ia-1, i=ib+1 (len njs), j=ja-1, j=jb+1 (len nis)
    size_t n;                 // elements per tiled vector
};

CAP_HD Lay make_lay(int ni, int nj, int nk, int ia, int ib, int ja, int jb, int ka, int kb)
{
    Lay L;
    L.ni = ni; L.nj = nj; L.nk = nk; L.nij = ni * nj;
    L.ia = ia; L.ja = ja; L.ka = ka;
    L.nis = ib - ia + 1; L.njs = jb - ja + 1; L.nks = kb - ka + 1;
    L.nTI = (L.nis + TW - 1) / TW;
    L.nTJ = (L.njs + TW - 1) / TW;
    L.S = L.nks + 2;
    L.ntiles = L.nTI * L.nTJ;
    L.nmain = (size_t)L.ntiles * (size_t)L.S * TCOL;
    L.ring[0] = L.nmain;
    L.ring[1] = L.ring[0] + (size_t)L.S * L.njs;
    L.ring[2] = L.ring[1] + (size_t)L.S * L.njs;
    L.ring[3] = L.ring[2] + (size_t)L.S * L.nis;
    L.n = L.ring[3] + (size_t)L.S * L.nis;
    return L;
}

CAP_HD int posmod(int x, int m) { int r = x % m; return r < 0 ? r + m : r; }

CAP_HD size_t tile_base(const Lay &L, int ti, int tj)
{
    return ((size_t)(ti * L.nTJ + tj) * (size_t)L.S) * TCOL;
}

// element of lane (b*16 + a) at slot s of the tile whose first element is tb
CAP_HD size_t elem_at(size_t tb, int slot, int lane)
{
    return tb + (size_t)slot * TCOL + lane;
}

CAP_HD size_t elem(const Lay &L, int ti, int tj, int slot, int b, int a)
{
    return elem_at(tile_base(L, ti, tj), slot, b * TW + a);
}

// element of the lane at tile-slot ts = tile*S + slot
CAP_HD size_t elem_ts(size_t ts, int slot, int lane)
{
    return elem_at((ts - (size_t)slot) * TCOL, slot, lane);
}

// ring element: side 0 (i=ia-1) / 1 (i=ib+1) indexed by cj, side 2 (j=ja-1) /
// 3 (j=jb+1) indexed by ci
CAP_HD size_t ring_elem(const Lay &L, int side, int idx, int kk)
{
    const int len = side < 2 ? L.njs : L.nis;
    return L.ring[side] + (size_t)kk * len + idx;
}

// element of solve column (ci, cj) (0-based) at plane kk; ring columns
// ci = -1 / nis or cj = -1 / njs map to the ring arrays
CAP_HD size_t elem_of(const Lay &L, int ci, int cj, int kk)
{
    if (ci < 0) return ring_elem(L, 0, cj, kk);
    if (ci >= L.nis) return ring_elem(L, 1, cj, kk);
    if (cj < 0) return ring_elem(L, 2, ci, kk);
    if (cj >= L.njs) return ring_elem(L, 3, ci, kk);
    int ti = ci / TW, b = ci % TW, tj = cj / TW, a = cj % TW;
    return elem(L, ti, tj, posmod(kk + a + b, L.S), b, a);
}

// offset into the reference array of column (ci, cj), plane kk
CAP_HD size_t std_index(const Lay &L, int ci, int cj, int kk)
{
    int i = L.ia + ci, j = L.ja + cj, k = L.ka - 1 + kk;
    return (size_t)(k - 1) * L.nij + (size_t)(i - 1) * L.nj + (size_t)(j - 1);
}

CAP_HD bool col_valid(const Lay &L, int ci, int cj) { return ci < L.nis && cj < L.njs; }

// Neighbour elements of solve cell (ti,tj,slot,b,a) at plane kk.  The
// reference's names: east = i+1 (b+1), west = i-1, north = j+1 (a+1),
// south = j-1, top = k+1, bottom = k-1.
struct Nb { size_t e, w, n, s, t, bt; };

CAP_HD Nb neighbours(const Lay &L, int ti, int tj, int slot, int b, int a, int kk)
{
    const int S = L.S;
    const int sp = slot + 1 == S ? 0 : slot + 1;
    const int sm = slot == 0 ? S - 1 : slot - 1;
    const int ci = ti * TW + b, cj = tj * TW + a;
    Nb r;
    r.t = elem(L, ti, tj, sp, b, a);
    r.bt = elem(L, ti, tj, sm, b, a);
    r.n = cj + 1 == L.njs ? ring_elem(L, 3, ci, kk)
          : a < TW - 1    ? elem(L, ti, tj, sp, b, a + 1)
                          : elem(L, ti, tj + 1, posmod(kk + b, S), b, 0);
    r.s = cj == 0      ? ring_elem(L, 2, ci, kk)
          : a > 0      ? elem(L, ti, tj, sm, b, a - 1)
                       : elem(L, ti, tj - 1, posmod(kk + TW - 1 + b, S), b, TW - 1);
    r.e = ci + 1 == L.nis ? ring_elem(L, 1, cj, kk)
          : b < TW - 1    ? elem(L, ti, tj, sp, b + 1, a)
                          : elem(L, ti + 1, tj, posmod(kk + a, S), 0, a);
    r.w = ci == 0      ? ring_elem(L, 0, cj, kk)
          : b > 0      ? elem(L, ti, tj, sm, b - 1, a)
                       : elem(L, ti - 1, tj, posmod(kk + a + TW - 1, S), TW - 1, a);
    return r;
}

// x mod S for 0 <= x (a few multiples of S at most)
CAP_HD int wrap_slot(int x, int S)
{
    while (x >= S) x -= S;
    return x;
}

// neighbours() for a solve cell whose element e, slot and plane kk are known:
// the same lane one slot up / down is e +- 256 (wrapping at S), so only the
// neighbours in other tiles need slot arithmetic.
CAP_HD Nb neighbours_at(const Lay &L, int ti, int tj, size_t e, int slot, int b, int a, int kk)
{
    const int S = L.S;
    const int lane = b * TW + a;
    const size_t tb = e - elem_at(0, slot, lane);
    const size_t up = elem_at(tb, slot + 1 == S ? 0 : slot + 1, lane);
    const size_t dn = elem_at(tb, slot == 0 ? S - 1 : slot - 1, lane);
    const int ci = ti * TW + b, cj = tj * TW + a;
    Nb r;
    r.t = up;
    r.bt = dn;
    r.n = cj + 1 == L.njs ? ring_elem(L, 3, ci, kk)
          : a < TW - 1    ? up + 1
                          : elem(L, ti, tj + 1, wrap_slot(kk + b, S), b, 0);
    r.s = cj == 0      ? ring_elem(L, 2, ci, kk)
          : a > 0      ? dn - 1
                       : elem(L, ti, tj - 1, wrap_slot(kk + TW - 1 + b, S), b, TW - 1);
    r.e = ci + 1 == L.nis ? ring_elem(L, 1, cj, kk)
          : b < TW - 1    ? up + TW
                          : elem(L, ti + 1, tj, wrap_slot(kk + a, S), 0, a);
    r.w = ci == 0      ? ring_elem(L, 0, cj, kk)
          : b > 0      ? dn - TW
                       : elem(L, ti - 1, tj, wrap_slot(kk + a + TW - 1, S), TW - 1, a);
    return r;
}

// One lane's walk over consecutive tile slots ts = lo, lo+1, ... (ts = tile*S
// + slot): tile, slot and plane advance by wrapping, with no division per
// step.  cell() tells whether the lane's element at the current slot is a
// solve cell.
struct SlotWalk {
    int a, b, slot, ti, tj, kk;
    bool colv;
    CAP_HD void start(const Lay &L, size_t lo, int lane)
    {
        a = lane % TW;
        b = lane / TW;
        slot = (int)(lo % (size_t)L.S);
        const int tile = (int)(lo / (size_t)L.S);
        ti = tile / L.nTJ;
        tj = tile % L.nTJ;
        colv = col_valid(L, ti * TW + b, tj * TW + a);
        kk = slot - a - b;
        while (kk < 0) kk += L.S;
    }
    CAP_HD bool cell(const Lay &L) const { return colv && kk >= 1 && kk <= L.nks; }
    CAP_HD void next(const Lay &L)
    {
        if (++kk == L.S) kk = 0;
        if (++slot == L.S) {
            slot = 0;
            if (++tj == L.nTJ) { tj = 0; ++ti; }
            colv = col_valid(L, ti * TW + b, tj * TW + a);
        }
    }
};

// Walk over a range of one tile's slots that wraps from slot S-1 to slot 0
// of the same tile (the SpMV's skewed walk).  Cell (b, a) of slot s lies at
// plane (s - a - b) mod S, so the j-side neighbour of an a = 15 cell is the
// next tile's a = 0 cell in slot s - 15 and the i-side neighbour of a b = 15
// cell is in slot s - 15 of the tile below; the i-1 / j-1 sides read slot
// s + 15.  Starting tile (ti, tj) at slot s0 - 15 (ti + tj), every block of a
// stencil reads its neighbour tiles' edge rows in the slots that their own
// blocks, started at the same time, work on at the same step.
struct TileWalk {
    int a, b, slot, ti, tj, kk;
    bool colv;
    size_t base;
    CAP_HD void start(const Lay &L, int tile, int slot0, int lane)
    {
        a = lane % TW;
        b = lane / TW;
        ti = tile / L.nTJ;
        tj = tile % L.nTJ;
        slot = slot0;
        base = (size_t)tile * (size_t)L.S * TCOL;
        colv = col_valid(L, ti * TW + b, tj * TW + a);
        kk = posmod(slot - a - b, L.S);
    }
    CAP_HD bool cell(const Lay &L) const { return colv && kk >= 1 && kk <= L.nks; }
    CAP_HD size_t elem(int lane) const { return base + (size_t)slot * TCOL + lane; }
    CAP_HD void next(const Lay &L)
    {
        if (++kk == L.S) kk = 0;
        if (++slot == L.S) slot = 0;
    }
};

// first slot of the skewed walk of tile `tile` over its range starting at s0
CAP_HD int skew_slot0(const Lay &L, int tile, int s0)
{
    const int ti = tile / L.nTJ, tj = tile % L.nTJ;
    return posmod(s0 - (TW - 1) * (ti + tj), L.S);
}

// Layout moves (k_gather, k_scatter).  Block bid moves MOVE_SLOTS
// consecutive slots [s0, s0 + ns) of one tile between the layouts through
// LDS.  On the reference side a row of a tile (its 16 columns of one i, at
// one plane) is 16 consecutive doubles, and its elements lie in 16
// consecutive slots (slot = plane + a + b mod S); the block reads or writes
// whole rows there (row step j: plane s0 - b - 15 + j) and whole slots on
// the tiled side, so both sides are coalesced.  A row whose slots straddle
// the ends of the block's range is shared with the neighbouring range.  The
// hardware hands block b to XCD b mod 8, and XCD x takes the items
// [x per, (x + 1) per) in order, so neighbouring ranges of a tile run on one
// XCD one after the other and the shared rows' lines are still in its L2.
constexpr int MOVE_SLOTS = 32;

CAP_HD int move_grid(const Lay &L)
{
    const long W = (long)L.ntiles * ((L.S + MOVE_SLOTS - 1) / MOVE_SLOTS);
    return (int)(8 * ((W + 7) / 8));
}

// the (tile, first slot) of block bid of the move_grid grid; false for the
// blocks past the last item
CAP_HD bool move_item(const Lay &L, int bid, int &tile, int &s0)
{
    const int nch = (L.S + MOVE_SLOTS - 1) / MOVE_SLOTS;
    const long W = (long)L.ntiles * nch;
    const long per = (W + 7) / 8;
    const long r = bid >> 3;
    const long w = (long)(bid & 7) * per + r;
    if (r >= per || w >= W) return false;
    tile = (int)(w / nch);
    s0 = (int)(w % nch) * MOVE_SLOTS;
    return true;
}

// row step j of lane (b, a): its element's slot s0 + r and plane kk; false
// when the lane has no element of the range at that step
CAP_HD bool move_step(const Lay &L, int s0, int ns, int b, int a, int j, int &r, int &kk)
{
    r = j - (TW - 1) + a;
    if (r < 0 || r >= ns) return false;
    kk = posmod(s0 + r - a - b, L.S);
    return true;
}

// Decompose an element index into (ci, cj, kk) (ring columns included);
// returns false for slots of tile columns beyond the solve box.
CAP_HD bool decode(const Lay &L, size_t e, int &ci, int &cj, int &kk)
{
    if (e >= L.nmain) {
        int side = e >= L.ring[3] ? 3 : e >= L.ring[2] ? 2 : e >= L.ring[1] ? 1 : 0;
        const int len = side < 2 ? L.njs : L.nis;
        const size_t off = e - L.ring[side];
        kk = (int)(off / len);
        const int idx = (int)(off % len);
        if (side == 0) { ci = -1; cj = idx; }
        else if (side == 1) { ci = L.nis; cj = idx; }
        else if (side == 2) { ci = idx; cj = -1; }
        else { ci = idx; cj = L.njs; }
        return kk <= L.nks + 1;
    }
    const size_t tslot = e / TCOL;
    const int lane = (int)(e % TCOL);
    int b = lane / TW, a = lane % TW;
    int slot = (int)(tslot % (size_t)L.S);
    int t = (int)(tslot / (size_t)L.S);
    int ti = t / L.nTJ, tj = t % L.nTJ;
    ci = ti * TW + b;
    cj = tj * TW + a;
    kk = posmod(slot - a - b, L.S);
    return col_valid(L, ci, cj) && kk <= L.nks + 1;
}

}  // namespace cap
