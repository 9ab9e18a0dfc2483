// Physics mode: fused pull collide+stream, one launch per time step.
//
// The step has the arithmetic of the reference's col.F90:127-229 and the inlet/outlet formulas of
// bcond.bc.F90, with the y/z periodicity applied by index wrap (ghosts always fresh, so there is
// no stale seam and no frozen edge ghost), bounce-back taken from the site's own opposite
// population (bcond.obs.F90:56-73 with fresh neighbours), and solid sites left untouched.
//
// Layout in HBM: two buffers (A current, B next), tile-major.  The lattice is cut into tiles of 64
// consecutive sites along x; a tile stores its 19 populations as 19 contiguous runs of 64 values
// (PopIndex), one guard tile before and after the lattice.  One wavefront updates one tile, so
// every load and store of a population is a contiguous 64-lane access (256 B in FP32, 128 B in
// FP16) whose address is a wave-uniform row pointer (scalar registers), one of three lane offsets
// (x shift -1, 0, +1) and the constant population offset Q * 64 values.  Per tile a 4-byte code
// says whether the tile is pure fluid (no solid site in or next to it: no mask traffic at all),
// all solid (skipped: no loads, no stores), or mixed (the 64 per-site link masks are read: bit q
// set = the source of population q is solid, bit 31 = the site itself is solid).  The codes are
// built once per mask with a wavefront ballot (k_classify).
//
// Arithmetic: FP64 storage collides full values in the reference's evaluation order (collide,
// bit-identical to the oracle); FP32, FP16 and BF16 storage collide the stored values themselves
// (deviations from a constant reference per population, zero for FP32) in single precision
// (collide_dev).
//
// Boundaries in x: the inlet ghost (x = 0) is the constant equilibrium at (u_in, 0, 0); the outlet
// ghost plane (x = l+1, 5 populations) is written by the sites at x = l while they collide and
// read one step later, so no extra pass over the lattice is needed.
#include <stdlib.h>

#include <type_traits>

#include "eup_internal.h"

namespace eup {

namespace {

template <typename S> struct XArgs {
    using R = typename Store<S>::compute;
    const S *A;  // tile 0 of the current state (tile-major layout)
    S *B;        // tile 0 of the next state
    const int32_t *tile_code;
    const uint32_t *link;
    const R *out_cur; // 5 x (m*n)
    R *out_next;
    R inl[5];    // inlet ghost values, in the kernel's representation (deviations for FP32 compute)
    DevRef ref;  // deviation references of the storage (conversions to and from full values)
    BgkF bgk;    // constants of the FP32-compute collision (unused by FP64)
    R omega, fgrad;
    int l, m, n, tpr;
    UDiv dtpr, dm; // t / tpr and row / m (t < tiles < 2^31)
    uint32_t tiles;
    uint32_t chunk; // logical blocks per XCD (order 1)
    uint32_t grp;   // logical blocks of one tile row, rounded up (order 2)
    UDiv dgrp;      // n / grp
    uint32_t order; // block order: 0 dispatch, 1 XCD-contiguous ranges, 2 XCD-owned tile rows
};

// FP32-compute storages (FP32, FP16 and BF16 deviations) run the step on the stored values
// themselves: deviations d_q = f_q - r_q, with r_q = 0 for full-value FP32 (collide_dev).  FP64
// computes on full values in the reference's evaluation order (collide).
template <typename S> constexpr bool kOnStored = std::is_same_v<typename Store<S>::compute, float>;

// e3m12 loads leave the NaN test of the stored codes to one test per site (`top`, nan_if_top)
template <typename S> constexpr bool kTopTest = std::is_same_v<S, e3m12_t>;

// a stored deviation as the kernel computes on it (FP32-compute storages)
template <typename S> __device__ __forceinline__ float step_raw(S v, int q, float &top)
{
    if constexpr (kTopTest<S>) return Store<S>::step_load(v, q, top);
    else return Store<S>::raw_load(v, q);
}

template <typename S>
__device__ __forceinline__ typename Store<S>::compute ld(const XArgs<S> &a, const S *p, int q, float &top)
{
    if constexpr (kOnStored<S>) return step_raw<S>(*p, q, top);
    else return Store<S>::load(*p, a.ref.r[q], q);
}

template <int... Qs> struct QSeq {};
using Q18 = QSeq<0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17>;
using Q19 = QSeq<0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18>;

// Addressing of one wavefront's tile: the tile of each of the nine source rows (wave-uniform
// byte pointers, indexed [dy+1][dz+1] for the row j+dy, k+dz with periodic wrap) and the lane byte
// offsets of the three x shifts.  A population pulled along c with c_x = +1 comes from lane - 1,
// which for lane 0 is lane 63 of the previous tile: it is addressed from the row pointer minus one
// tile so that every lane offset is non-negative.  For c_x = -1 lane 63 reads lane 0 of the next
// tile.
struct Wave {
    const char *row[3][3];
    uint32_t o0, op, om;
    int jj[3], kk[3];
    int ti, i;
};

template <typename S>
__device__ __forceinline__ void wave_setup(const XArgs<S> &a, Wave &w, int ti, int j, int k, int lane)
{
    constexpr int64_t TB = (int64_t)kTileElems * sizeof(S);
    const int m = a.m, n = a.n;
    w.ti = ti;
    w.i = ti * kTile + lane;
    w.jj[0] = (j == 0) ? m - 1 : j - 1;
    w.jj[1] = j;
    w.jj[2] = (j == m - 1) ? 0 : j + 1;
    w.kk[0] = (k == 0) ? n - 1 : k - 1;
    w.kk[1] = k;
    w.kk[2] = (k == n - 1) ? 0 : k + 1;
    // row tile indices fit in 32 bits (tiles < 2^32): one widening multiply per row pointer
    const char *A = reinterpret_cast<const char *>(a.A);
    const uint32_t tpr = (uint32_t)a.tpr, tpm = tpr * (uint32_t)m;
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dz = 0; dz < 3; ++dz) {
            const uint32_t rt = (uint32_t)ti + tpr * (uint32_t)w.jj[dy] + tpm * (uint32_t)w.kk[dz];
            w.row[dy][dz] = A + (uint64_t)rt * (uint64_t)TB;
        }
    // The first tile of a row never uses what lane 0 pulls along c_x = +1 (x = 0 is the inlet, or
    // x_edges reloads the far end of the row), nor the last tile what lane 63 pulls along c_x = -1:
    // those lanes are pointed at their own tile's run, a line the other lanes fetch anyway, instead
    // of the neighbouring row's last (first) line, which another XCD owns (FP16 fetches 6.6 % less).
    w.o0 = (uint32_t)lane * (uint32_t)sizeof(S);
    const uint32_t op0 = (ti == 0) ? (uint32_t)TB : (uint32_t)(kTile - 1) * (uint32_t)sizeof(S);
    const uint32_t om63 = (ti == a.tpr - 1) ? (uint32_t)(kTile - 1) * (uint32_t)sizeof(S) : (uint32_t)TB;
    w.op = (lane == 0) ? op0 : (uint32_t)TB + (uint32_t)(lane - 1) * (uint32_t)sizeof(S);
    w.om = (lane == kTile - 1) ? om63 : (uint32_t)(lane + 1) * (uint32_t)sizeof(S);
}

template <typename S>
__device__ __forceinline__ const S *at(const char *p, int64_t bytes, uint32_t off)
{
    return reinterpret_cast<const S *>(p + bytes + (uint64_t)off);
}

// Pull of population Q, first pass: the load from the interior source x - c_Q, or (MIXED, source
// solid) the site's own opposite population.  All 19 loads are issued back to back; the sites
// whose source lies outside the lattice in x are fixed afterwards (x_edges).
template <typename S, bool MIXED, int Q>
__device__ __forceinline__ void pull(const XArgs<S> &a, const Wave &w, uint32_t lm, typename Store<S>::compute *x,
                                     float &top)
{
    constexpr int cx = Dir<Q>::cx, cy = Dir<Q>::cy, cz = Dir<Q>::cz, op = Dir<Q>::opp;
    constexpr int64_t TB = (int64_t)kTileElems * sizeof(S);
    constexpr int64_t QB = (int64_t)Q * kTile * sizeof(S);
    const char *rb = w.row[1 - cy][1 - cz]; // source row j - c_y, k - c_z
    const S *p;
    if constexpr (cx == 1) p = at<S>(rb, QB - TB, w.op);
    else if constexpr (cx == 0) p = at<S>(rb, QB, w.o0);
    else p = at<S>(rb, QB, w.om);
    if constexpr (kOnStored<S>) {
        float corr = 0.0f;
        if constexpr (MIXED) {
            // the source is solid: bounce back the site's own opposite population (as a deviation
            // from r_Q it carries r_opp - r_Q on top of the stored deviation)
            if ((lm >> Q) & 1u) {
                p = at<S>(w.row[1][1], (int64_t)op * kTile * sizeof(S), w.o0);
                corr = a.bgk.bb[Q];
            }
        }
        // (the bounce-back path is taken by FP64 only: the other storages use the at-once form)
        if constexpr (MIXED && Store<S>::deviation) x[Q] = step_raw<S>(*p, Q, top) + corr;
        else x[Q] = step_raw<S>(*p, Q, top);
    } else {
        float ref = a.ref.r[Q];
        if constexpr (MIXED) {
            // the source is solid: bounce back the site's own opposite population
            if ((lm >> Q) & 1u) {
                p = at<S>(w.row[1][1], (int64_t)op * kTile * sizeof(S), w.o0);
                ref = a.ref.r[op];
            }
        }
        x[Q] = Store<S>::load(*p, ref, Q);
    }
}

template <typename S, bool MIXED, int... Qs>
__device__ __forceinline__ void pull_all(QSeq<Qs...>, const XArgs<S> &a, const Wave &w, uint32_t lm,
                                         typename Store<S>::compute *x, float &top)
{
    (pull<S, MIXED, Qs>(a, w, lm, x, top), ...);
}

// Second pass, populations with c_x = +1 (Q = 0..4) at the site x = 0 and c_x = -1 (Q = 9..13) at
// x = l - 1: open ends take the inlet equilibrium and the outlet ghost plane, a periodic x takes
// the far end of the row (unless the bounce-back of the first pass applies).  Only the first and
// the last tile of a row reach this code (wave-uniform tests).
template <typename S, bool XOPEN>
__device__ __forceinline__ void x_edges(const XArgs<S> &a, const Wave &w, uint32_t lm, typename Store<S>::compute *x,
                                        float &top)
{
    using R = typename Store<S>::compute;
    constexpr int64_t TB = (int64_t)kTileElems * sizeof(S);
    constexpr int cxs[19] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
    constexpr int cys[19] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1, 0};
    constexpr int czs[19] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1, 0};
    if (w.ti == 0) {
        const bool edge = w.i == 0;
#pragma unroll
        for (int Q = 0; Q < 5; ++Q) {
            if constexpr (XOPEN) {
                x[Q] = edge ? a.inl[Q] : x[Q];
            } else {
                if (edge && !((lm >> Q) & 1u)) {
                    const char *rb = w.row[1 - cys[Q]][1 - czs[Q]];
                    x[Q] = ld<S>(a, at<S>(rb, (int64_t)Q * kTile * sizeof(S) + (int64_t)(a.tpr - 1) * TB,
                                          (uint32_t)((a.l - 1) & 63) * (uint32_t)sizeof(S)),
                                 Q, top);
                }
            }
        }
    }
    if (w.ti == a.tpr - 1) {
        const bool edge = w.i == a.l - 1;
#pragma unroll
        for (int Q = 9; Q < 14; ++Q) {
            if constexpr (XOPEN) {
                // wave-uniform address into a plane the launch only reads: a scalar load
                typedef const __attribute__((address_space(4))) R ConstR;
                const R o = *(ConstR *)(a.out_cur + ((Q - 9) * (size_t)(a.m * a.n) + w.jj[1 - cys[Q]] +
                                                     (size_t)a.m * w.kk[1 - czs[Q]]));
                x[Q] = edge ? o : x[Q];
            } else {
                if (edge && !((lm >> Q) & 1u)) {
                    const char *rb = w.row[1 - cys[Q]][1 - czs[Q]];
                    x[Q] = ld<S>(a, at<S>(rb, (int64_t)Q * kTile * sizeof(S) - (int64_t)w.ti * TB, 0u), Q, top);
                }
            }
        }
    }
    (void)cxs;
}

#ifndef EUP_NT_STORE
#define EUP_NT_STORE 1
#endif

template <typename S> __device__ __forceinline__ void store_val(S *p, S v)
{
#if EUP_NT_STORE
    if constexpr (std::is_same_v<S, bf16_t> || std::is_same_v<S, e3m12_t>)
        __builtin_nontemporal_store(v.bits, reinterpret_cast<uint16_t *>(p));
    else
        __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// the stored value of population Q from the kernel's representation
template <typename S, int Q>
__device__ __forceinline__ void conv(const XArgs<S> &a, const typename Store<S>::compute *x, S *sv)
{
    if constexpr (kOnStored<S>) sv[Q] = Store<S>::raw_store(x[Q], Q);
    else sv[Q] = Store<S>::store(x[Q], a.ref.r[Q], Q);
}

// convert population Q, keep the stored value in sv and store it
template <typename S, int Q>
__device__ __forceinline__ void put(const XArgs<S> &a, char *tb, uint32_t o0, const typename Store<S>::compute *x,
                                    S *sv)
{
    conv<S, Q>(a, x, sv);
    store_val<S>(reinterpret_cast<S *>(tb + (int64_t)Q * kTile * sizeof(S) + (uint64_t)o0), sv[Q]);
}

template <typename S, int... Qs>
__device__ __forceinline__ void put_all(QSeq<Qs...>, const XArgs<S> &a, char *tb, uint32_t o0,
                                        const typename Store<S>::compute *x, S *sv)
{
    (put<S, Qs>(a, tb, o0, x, sv), ...);
}

// The x = l sites write the next step's outlet ghost values (bcond.bc.F90:127-164) from their
// post-collision populations as stored (sv), in the step kernel's representation: the plane is a
// function of the stored state, as the reference's ghost fill reads its stored arrays, and
// k_outlet_refresh computes the same plane from a state the host sets.
template <typename S>
__device__ __forceinline__ void outlet_next(const XArgs<S> &a, const Wave &w, int j, int k, const S *sv)
{
    using R = typename Store<S>::compute;
    if (w.i != a.l - 1) return;
    R o[5], xf[19];
    // ref + stored deviation, as Store<S>::load forms it (bgk.r holds the same references as ref.r;
    // the collision already keeps it in scalar registers); the identity for full-value storages
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        if constexpr (kTopTest<S>) xf[q] = a.bgk.r[q] + Store<S>::load_magnitude(sv[q], q);
        else if constexpr (kOnStored<S> && Store<S>::deviation) xf[q] = a.bgk.r[q] + Store<S>::raw_load(sv[q], q);
        else xf[q] = Store<S>::raw_load(sv[q], q);
    }
    outlet_eq<R>(xf, o);
    if constexpr (kOnStored<S> && Store<S>::deviation) {
        // the ghost plane holds deviations
#pragma unroll
        for (int v = 0; v < 5; ++v) o[v] -= a.bgk.r[9 + v];
    }
    const size_t mn = (size_t)a.m * a.n;
#pragma unroll
    for (int v = 0; v < 5; ++v) a.out_next[v * mn + j + (size_t)a.m * k] = o[v];
}

// Store the lane's post-collision populations x of tile t, converted once: the outlet (the x = l
// sites) takes the values as stored.
template <typename S, bool XOPEN>
__device__ __forceinline__ void finish(const XArgs<S> &a, uint32_t t, const Wave &w, int j, int k,
                                       const typename Store<S>::compute *x)
{
    S sv[19];
    put_all<S>(Q19{}, a, reinterpret_cast<char *>(a.B) + (int64_t)t * ((int64_t)kTileElems * sizeof(S)), w.o0, x, sv);
    if constexpr (XOPEN) outlet_next<S>(a, w, j, k, sv);
}

// Update tile t (addressing w, row j, k), one site per lane.  MIXED applies the link mask lm of the
// lane's site: bounce-back from solid sources, solid sites frozen.
template <typename S, bool XOPEN, bool MIXED>
__device__ __forceinline__ void update_tile(const XArgs<S> &a, uint32_t t, const Wave &w, int j, int k,
                                            uint32_t lm)
{
    using R = typename Store<S>::compute;
    if constexpr (MIXED) {
        if (lm & 0x80000000u) return; // solid site: frozen
    }
    R x[19];
    float top = 0.0f;
    pull_all<S, MIXED>(Q18{}, a, w, lm, x, top);
    x[18] = ld<S>(a, at<S>(w.row[1][1], 18 * kTile * (int64_t)sizeof(S), w.o0), 18, top);
    x_edges<S, XOPEN>(a, w, lm, x, top);
    if (w.i >= a.l) return; // padding lanes of a partial last tile
    // a NaN code among the loads makes the site's density, and so its whole update, NaN
    if constexpr (kTopTest<S>) x[18] = Store<S>::nan_if_top(top, x[18]);
    if constexpr (kOnStored<S>) collide_dev<Store<S>::deviation>(x, a.bgk);
    else collide<R>(x, x, a.omega, a.fgrad);
    finish<S, XOPEN>(a, t, w, j, k, x);
}

// Mixed tile with every load issued at once: the link masks, the 19 pulls from the fluid
// addresses and the site's own 18 moving populations (the bounce-back sources), then a per-lane
// choice in registers.  update_tile<MIXED> instead waits for the link masks before it can address
// its pulls (two dependent memory round trips); this form costs 18 more loads per mixed tile
// (lines that the neighbouring tiles read as well) and as many registers.  Measured on MI355X
// (512 x 256 x 256 sponge, 8 % mixed tiles): FP32 +3.3 %, FP16 +2.4 %, FP64 -0.6 % (whose 100
// registers per lane leave no room), so FP64 keeps update_tile<MIXED>.
template <typename S, int Q>
__device__ __forceinline__ void own_val(const XArgs<S> &a, const Wave &w, typename Store<S>::compute *o, float &top)
{
    o[Q] = ld<S>(a, at<S>(w.row[1][1], (int64_t)Q * kTile * sizeof(S), w.o0), Q, top);
}

template <typename S, int... Qs>
__device__ __forceinline__ void own_all(QSeq<Qs...>, const XArgs<S> &a, const Wave &w, typename Store<S>::compute *o,
                                        float &top)
{
    (own_val<S, Qs>(a, w, o, top), ...);
}

template <typename S, int Q>
__device__ __forceinline__ void choose(const XArgs<S> &a, uint32_t lm, const typename Store<S>::compute *own,
                                       typename Store<S>::compute *x)
{
    constexpr int op = Dir<Q>::opp;
    typename Store<S>::compute b = own[op];
    if constexpr (kOnStored<S> && Store<S>::deviation) b += a.bgk.bb[Q];
    x[Q] = ((lm >> Q) & 1u) ? b : x[Q];
}

template <typename S, int... Qs>
__device__ __forceinline__ void choose_all(QSeq<Qs...>, const XArgs<S> &a, uint32_t lm,
                                           const typename Store<S>::compute *own, typename Store<S>::compute *x)
{
    (choose<S, Qs>(a, lm, own, x), ...);
}

template <typename S, bool XOPEN>
__device__ __forceinline__ void update_mixed_at_once(const XArgs<S> &a, uint32_t t, const Wave &w, int j, int k,
                                                     const uint32_t *lmp)
{
    using R = typename Store<S>::compute;
    const uint32_t lm = *lmp;
    R x[19], own[18];
    float top = 0.0f;
    pull_all<S, false>(Q18{}, a, w, 0u, x, top);
    x[18] = ld<S>(a, at<S>(w.row[1][1], 18 * kTile * (int64_t)sizeof(S), w.o0), 18, top);
    own_all<S>(Q18{}, a, w, own, top);
    if (lm & 0x80000000u) return; // solid site: frozen
    choose_all<S>(Q18{}, a, lm, own, x);
    x_edges<S, XOPEN>(a, w, lm, x, top);
    if (w.i >= a.l) return;
    if constexpr (kTopTest<S>) x[18] = Store<S>::nan_if_top(top, x[18]);
    if constexpr (kOnStored<S>) collide_dev<Store<S>::deviation>(x, a.bgk);
    else collide<R>(x, x, a.omega, a.fgrad);
    finish<S, XOPEN>(a, t, w, j, k, x);
}

template <typename S> constexpr bool kMixedAtOnce = sizeof(S) <= 4;

// One step of one tile per wavefront (the body of k_fused).
template <typename S, bool XOPEN>
__device__ __forceinline__ void fused_step(const XArgs<S> &a)
{
    // Block order (chosen by fused_block_order).  Hardware block b runs on XCD b % 8.  Order 0:
    // the eight XCDs sweep the lattice together in dispatch order.  Order 1: each XCD gets a
    // contiguous range of tiles.  Order 2: the XCDs still sweep together, but every group of
    // 8 * grp hardware blocks is dealt out so that each XCD takes grp consecutive logical blocks
    // (a whole tile row), so the x-neighbouring tiles that share cache lines through their
    // x-shifted pulls run on one XCD and share its L2; a last partial group keeps dispatch order
    // (the mapping stays a bijection).
    const uint32_t b = blockIdx.x;
    uint32_t lb = b;
    if (a.order == 1) {
        lb = (b & 7u) * a.chunk + (b >> 3);
    } else if (a.order == 2) {
        const uint32_t r8 = b >> 3, gq = a.dgrp.div(r8), g0 = gq * 8u * a.grp;
        if (g0 + 8u * a.grp <= gridDim.x) lb = g0 + (b & 7u) * a.grp + (r8 - gq * a.grp);
    }
    const uint32_t t = __builtin_amdgcn_readfirstlane(lb * (kBlock / kTile) + (threadIdx.x >> 6));
    if (t >= a.tiles) return;
    // the codes are read-only during the launch: a scalar load through the constant address space
    // (consecutive wavefronts share its scalar-cache line) instead of a vector load + readfirstlane;
    // the addressing below does not depend on the code and overlaps the load
    typedef const __attribute__((address_space(4))) int32_t ConstI32;
    const int32_t code = *(ConstI32 *)(a.tile_code + t);
    const uint32_t row = a.dtpr.div(t);
    const int ti = (int)(t - row * (uint32_t)a.tpr);
    const uint32_t kk = a.dm.div(row);
    const int j = (int)(row - kk * (uint32_t)a.m), k = (int)kk;
    Wave w;
    wave_setup<S>(a, w, ti, j, k, (int)(threadIdx.x & 63));
    if (code == kTileSolid) return;
    if (code == kTileFluid) {
        update_tile<S, XOPEN, false>(a, t, w, j, k, 0u);
    } else if constexpr (kMixedAtOnce<S>) {
        update_mixed_at_once<S, XOPEN>(a, t, w, j, k, a.link + ((size_t)code * kTile + (threadIdx.x & 63)));
    } else {
        const uint32_t lm = a.link[(size_t)code * kTile + (threadIdx.x & 63)];
        update_tile<S, XOPEN, true>(a, t, w, j, k, lm);
    }
}

// No register target beyond the block size: asking the compiler for 8 wavefronts per SIMD (64
// registers) makes it interleave the 19 loads with their uses and costs the 2-byte storages
// 7-13 % on MI355X, although it removes only one register.
template <typename S, bool XOPEN>
__global__ __launch_bounds__(kBlock) void k_fused(XArgs<S> a)
{
    fused_step<S, XOPEN>(a);
}

// e3m12 with open x faces, with a budget of 96 scalar registers (101 without): still seven
// wavefronts per SIMD (its 70 vector registers decide), and the schedule the compiler builds under
// the budget ran the 512 x 256 x 256 sponge step in 0.431 / 0.427 ms instead of 0.442 / 0.438 ms
// on two MI355X boxes, level with FP16 (0.425 ms).  The same budget costs FP16 and BF16 3 % and
// 7 %: there it lets an eighth wavefront in, whose extra 19 loads in flight per lane disorder the
// memory request stream.
template <>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_num_sgpr(96))) void k_fused<e3m12_t, true>(XArgs<e3m12_t> a)
{
    fused_step<e3m12_t, true>(a);
}

// FP64 with open x faces, compiled for the three wavefronts per SIMD that its block cap allows:
// 112 registers instead of 94 (so at most four wavefronts where the cap is lifted), and 1.624
// instead of 1.637 ms per 512 x 256 x 256 sponge step on MI355X (the parity stream-collide under
// the same hint: no change).
template <>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(1, 3))) void k_fused<double, true>(XArgs<double> a)
{
    fused_step<double, true>(a);
}

// ---------------------------------------------------------------------------------------------
// mask preparation
// ---------------------------------------------------------------------------------------------

// per site: link mask (bit q = source x - c_q solid, bit 31 = self solid); per tile: class via
// ballot.  cls: 0 fluid, 1 mixed, 2 all solid
__global__ __launch_bounds__(kBlock) void k_classify(const uint8_t *__restrict__ solid, int l, int m,
                                                     int n, int tpr, uint32_t tiles, int xopen,
                                                     int noskip, uint8_t *__restrict__ cls,
                                                     uint32_t *__restrict__ link_all)
{
    const uint32_t t = blockIdx.x * (kBlock / kTile) + (threadIdx.x >> 6);
    if (t >= tiles) return; // whole wavefront leaves together
    const uint32_t row = t / (uint32_t)tpr;
    const int i = (int)(t - row * (uint32_t)tpr) * kTile + (int)(threadIdx.x & 63);
    const int j = (int)(row % (uint32_t)m), k = (int)(row / (uint32_t)m);
    const bool active = i < l;
    uint32_t lm = 0;
    if (active) {
        auto S = [&](int ii, int jj, int kk) -> bool {
            if (ii < 0 || ii >= l) {
                if (xopen) return false;
                ii = (ii + l) % l;
            }
            jj = (jj + m) % m;
            kk = (kk + n) % n;
            return solid[(size_t)ii + (size_t)l * ((size_t)jj + (size_t)m * kk)] != 0;
        };
        const int cxs[18] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0};
        const int cys[18] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1};
        const int czs[18] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1};
#pragma unroll
        for (int q = 0; q < 18; ++q)
            if (S(i - cxs[q], j - cys[q], k - czs[q])) lm |= 1u << q;
        if (S(i, j, k)) lm |= 0x80000000u;
    }
    const uint64_t any_link = __ballot(active && lm != 0);
    const uint64_t not_solid = __ballot(active && !(lm & 0x80000000u));
    uint8_t c;
    if (noskip) c = 1;
    else if (not_solid == 0) c = 2;
    else if (any_link == 0) c = 0;
    else c = 1;
    if ((threadIdx.x & 63) == 0) cls[t] = c;
    link_all[(size_t)t * kTile + (threadIdx.x & 63)] = lm;
}

__global__ void k_compact_links(const int32_t *__restrict__ code, const uint32_t *__restrict__ link_all,
                                uint32_t tiles, uint32_t *__restrict__ link)
{
    const uint32_t t = blockIdx.x * (kBlock / kTile) + (threadIdx.x >> 6);
    if (t >= tiles) return;
    const int32_t c = code[t];
    if (c < 0) return;
    link[(size_t)c * kTile + (threadIdx.x & 63)] = link_all[(size_t)t * kTile + (threadIdx.x & 63)];
}

// ---------------------------------------------------------------------------------------------
// state conversion and outlet refresh
// ---------------------------------------------------------------------------------------------
template <typename S>
__global__ void k_from_staging(const double *__restrict__ stg, S *__restrict__ d0, S *__restrict__ d1,
                               PopIndex P, int l, int m, int n, int q, float ref)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t N = (int64_t)l * m * n;
    if (t >= N) return;
    const int i = (int)(t % l);
    const int64_t r = t / l;
    const int j = (int)(r % m), k = (int)(r / m);
    const double v = stg[(int64_t)(i + 1) + (int64_t)(l + 2) * ((int64_t)(j + 1) + (int64_t)(m + 2) * (k + 1))];
    const S sv = Store<S>::store64(v, ref, q);
    const int64_t e = P(q, i, j, k);
    d0[e] = sv;
    d1[e] = sv;
}

template <typename S>
__global__ void k_to_staging(const S *__restrict__ src, double *__restrict__ stg, PopIndex P, int l, int m,
                             int n, int q, float ref)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t N = (int64_t)l * m * n;
    if (t >= N) return;
    const int i = (int)(t % l);
    const int64_t r = t / l;
    const int j = (int)(r % m), k = (int)(r / m);
    stg[(int64_t)(i + 1) + (int64_t)(l + 2) * ((int64_t)(j + 1) + (int64_t)(m + 2) * (k + 1))] =
        Store<S>::load64(src[P(q, i, j, k)], ref, q);
}

// counter-based uniform in [-1, 1) (splitmix64 finaliser)
__device__ __forceinline__ double hash_uniform(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// equilibrium start over nsites sites; site t is (t % l, (t / l) % m, t / (l m)) of the layout P
// (the parity layout passes its ghosted extents l + 2, m + 2)
template <typename S>
__global__ void k_init_eq(S *base, PopIndex P, int l, int m, int64_t nsites, double ux, double uy0, double uz,
                          double noise, uint64_t seed, DevRef ref)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsites) return;
    const int i = (int)(t % l);
    const int64_t r = t / l;
    const int j = (int)(r % m), k = (int)(r / m);
    const double uy = uy0 + noise * hash_uniform(seed, (uint64_t)t);
    const double usq = ux * ux + uy * uy + uz * uz;
    const int cx[19] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
    const int cy[19] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1, 0};
    const int cz[19] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1, 0};
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const double cu = cx[q] * ux + cy[q] * uy + cz[q] * uz;
        const double f = weight<double>(q) * (1.0 + (3.0 * cu + 1.5 * (3.0 * cu * cu - usq)));
        base[P(q, i, j, k)] = Store<S>::store64(f, ref.r[q], q);
    }
}

template <typename S>
__global__ void k_outlet_refresh(XArgs<S> a, typename Store<S>::compute *o0,
                                 typename Store<S>::compute *o1)
{
    using R = typename Store<S>::compute;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t mn = (int64_t)a.m * a.n;
    if (t >= mn) return;
    const int j = (int)(t % a.m), k = (int)(t / a.m);
    PopIndex P;
    P.stride = 0;
    P.off = 0;
    P.pitch = a.l;
    P.m = a.m;
    P.tpr = a.tpr;
    P.tiled = 1;
    R x[19], o[5];
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        // as outlet_next takes the stored values
        if constexpr (kTopTest<S>) x[q] = a.ref.r[q] + Store<S>::load_magnitude(a.A[P(q, a.l - 1, j, k)], q);
        else x[q] = Store<S>::load(a.A[P(q, a.l - 1, j, k)], a.ref.r[q], q);
    }
    outlet_eq<R>(x, o);
#pragma unroll
    for (int u = 0; u < 5; ++u) {
        // in the step kernel's representation (r = 0 for full values)
        const R v = kOnStored<S> ? o[u] - (R)a.bgk.r[9 + u] : o[u];
        o0[u * mn + t] = v;
        o1[u * mn + t] = v;
    }
}

template <typename S> XArgs<S> make_args(Handle *h)
{
    using R = typename Store<S>::compute;
    XArgs<S> a;
    a.A = static_cast<const S *>(h->buf[h->cur]);
    a.B = static_cast<S *>(h->buf[h->cur ^ 1]);
    a.tile_code = h->tile_code;
    a.link = h->link_mask;
    a.out_cur = static_cast<const R *>(h->outlet[h->cur]);
    a.out_next = static_cast<R *>(h->outlet[h->cur ^ 1]);
    a.ref = dev_ref(h);
    a.omega = (R)h->cfg.omega;
    a.fgrad = (R)h->cfg.fgrad;
    // r_q of the FP32-compute step: the storage's deviation references, or 0 for full values
    static const int opp[19] = {11, 12, 9, 10, 13, 14, 15, 16, 17, 2, 3, 0, 1, 4, 5, 6, 7, 8, 18};
    static const int cx[19] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
    static const int cy[19] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1, 0};
    static const int cz[19] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1, 0};
    double r[19];
    for (int q = 0; q < 19; ++q) r[q] = Store<S>::deviation ? (double)a.ref.r[q] : 0.0;
    const double om = h->cfg.omega;
    BgkF &c = a.bgk;
    double rsum = 0.0, jr[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < 19; ++q) {
        c.r[q] = (float)r[q];
        c.wr[q] = (float)(om * r[q]);
        c.bb[q] = (float)(r[opp[q]] - r[q]);
        rsum += r[q];
        jr[0] += cx[q] * r[q];
        jr[1] += cy[q] * r[q];
        jr[2] += cz[q] * r[q];
    }
    c.rsum = (float)rsum;
    for (int d = 0; d < 3; ++d) c.jr[d] = (float)jr[d];
    c.omega1 = (float)(1.0 - om);
    c.ow1 = (float)(om / 18.0);
    c.ow2 = (float)(om / 36.0);
    c.fgrad = (float)h->cfg.fgrad;
    if constexpr (kOnStored<S>) {
        double inl[5];
        inlet_eq<double>(h->cfg.u_inflow, inl);
        for (int u = 0; u < 5; ++u) a.inl[u] = (R)(inl[u] - r[u]);
    } else {
        R inl[5];
        inlet_eq<R>((R)h->cfg.u_inflow, inl);
        for (int u = 0; u < 5; ++u) a.inl[u] = inl[u];
    }
    a.l = h->g.l;
    a.m = h->g.m;
    a.n = h->g.n;
    a.tpr = h->g.tiles_per_row;
    a.dtpr = UDiv::make((uint32_t)a.tpr);
    a.dm = UDiv::make((uint32_t)a.m);
    a.tiles = (uint32_t)h->g.tiles;
    const uint32_t per_block = kBlock / kTile;
    const uint32_t nb = (uint32_t)((h->g.tiles + per_block - 1) / per_block);
    a.chunk = (nb + 7) / 8;
    a.grp = (uint32_t)((h->g.tiles_per_row + per_block - 1) / per_block);
    a.dgrp = UDiv::make(a.grp);
    a.order = h->fused_remap > 0 ? (uint32_t)h->fused_remap : 0u; // set by the step launch (fused_block_order)
    return a;
}

inline bool x_open(const Handle *h) { return h->cfg.front != EUP_FACE_PERIODIC; }

// Occupancy cap of the fused kernel in blocks (4 wavefronts each) per CU.  The kernel uses no LDS;
// a dynamic LDS allocation of LDS-per-CU / cap bytes per block is what limits the blocks that
// share a CU.  Every lane keeps the 19 loads of its site in flight; fewer resident wavefronts give
// the HBM controllers a more orderly request stream when each load moves 256 B or more per
// wavefront.  Measured on MI355X (512 x 256 x 256 sponge): FP32 runs 2-8 % faster at 5 blocks per
// CU than uncapped (8, set by its registers), FP64 about 3 % faster at 3; the 2-byte storages,
// whose loads move 128 B, are fastest uncapped (0).  EUP_FUSED_BLOCKS_PER_CU overrides the choice.
static int fused_cap(const Handle *h)
{
    int cap = h->elem == 2 ? 0 : (h->elem == 4 ? 5 : 3);
    // a grid that fits the GPU in one round of 8 blocks per CU runs uncapped: with the cap it
    // would take two rounds (128 x 64 x 64 cylinder, FP32: +2.7 % uncapped)
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->cfg.device) == hipSuccess && cus > 0 &&
        (h->g.tiles + 3) / 4 <= (int64_t)cus * 8)
        cap = 0;
    if (const char *e = getenv("EUP_FUSED_BLOCKS_PER_CU")) cap = atoi(e);
    return cap;
}

// Block order of the fused kernel.  In dispatch order the resident wavefronts of the whole GPU
// form one window that moves through the lattice; when that window covers at least a plane of
// tiles, the rows a tile pulls from are fetched while their other readers are resident too, and
// the XCDs sweeping together is faster, best with each XCD owning whole tile rows (order 2: the
// lines that x-neighbouring tiles share through their x-shifted pulls stay in one L2).  Otherwise
// each XCD is given a contiguous range of tiles (order 1).  Measured on MI355X (window / plane of
// tiles; MLUPS in orders 0 / 1 / 2): 512 x 256 x 256 sponge FP32 (2.5) 39 100-39 400 / - /
// 40 000-40 200, FP16 (4.0) 75 100-75 300 / - / 78 000, BF16 (4.0) - / 71 300 / 78 500, FP64 (1.5)
// - / 18 970 / 20 210; 1536 x 768 x 768 sponge FP32 (0.28) - / 39 280 / 36 360, FP16 (0.44)
// - / 77 900 / 74 700; its 192 x 768 x 768 slab FP32 (2.2) 39 810 / - / 40 290.
// EUP_FUSED_XCD_REMAP=0/1/2 forces the order.
static int fused_block_order(const Handle *h, int cap)
{
    if (const char *e = getenv("EUP_FUSED_XCD_REMAP")) {
        const int v = atoi(e);
        return v < 0 ? 0 : (v > 2 ? 2 : v);
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->cfg.device) != hipSuccess || cus <= 0)
        cus = 256;
    const double window = (double)cus * (cap > 0 ? cap : 8) * (kBlock / kTile);
    const double plane = (double)h->g.tiles_per_row * h->g.m;
    return window >= plane ? 2 : 1;
}

static size_t fused_lds_bytes(const Handle *h) { return occupancy_lds_bytes(h, fused_cap(h)); }

} // namespace

// dynamic LDS per block that limits a kernel without LDS to `cap` blocks per CU (0 = no limit)
size_t occupancy_lds_bytes(const Handle *h, int cap)
{
    if (cap <= 0) return 0;
    int per_cu = 0;
    if (hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, h->cfg.device) !=
            hipSuccess || per_cu <= 0)
        return 0;
    int per_block = 0;
    if (hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, h->cfg.device) != hipSuccess)
        return 0;
    size_t b = ((size_t)per_cu / (size_t)cap) & ~(size_t)1023;
    if (b > (size_t)per_block) b = (size_t)per_block & ~(size_t)1023;
    return b;
}

namespace {

template <typename S> int step_t(Handle *h)
{
    if (h->fused_remap < 0) h->fused_remap = fused_block_order(h, fused_cap(h));
    XArgs<S> a = make_args<S>(h);
    const unsigned grid = 8u * a.chunk; // multiple of 8 so the XCD remap is a bijection
    if (h->fused_lds == (size_t)-1) h->fused_lds = fused_lds_bytes(h);
    const size_t lds = h->fused_lds;
    if (x_open(h))
        hipLaunchKernelGGL((k_fused<S, true>), dim3(grid), dim3(kBlock), lds, h->stream, a);
    else
        hipLaunchKernelGGL((k_fused<S, false>), dim3(grid), dim3(kBlock), lds, h->stream, a);
    if (hip_check(h, hipGetLastError(), "fused step launch")) return EUP_EHIP;
    h->cur ^= 1;
    return 0;
}

template <typename S> int refresh_t(Handle *h)
{
    using R = typename Store<S>::compute;
    XArgs<S> a = make_args<S>(h);
    const int64_t mn = (int64_t)h->g.m * h->g.n;
    hipLaunchKernelGGL((k_outlet_refresh<S>), dim3((unsigned)((mn + 255) / 256)), dim3(256), 0,
                       h->stream, a, static_cast<R *>(h->outlet[0]), static_cast<R *>(h->outlet[1]));
    return hip_check(h, hipGetLastError(), "outlet refresh") ? EUP_EHIP : 0;
}

template <typename S> int set_t(Handle *h, int q)
{
    S *d0 = static_cast<S *>(h->buf[0]);
    S *d1 = static_cast<S *>(h->buf[1]);
    hipLaunchKernelGGL((k_from_staging<S>), dim3((unsigned)((h->g.sites + 255) / 256)), dim3(256), 0,
                       h->stream, h->staging, d0, d1, pop_index(h), h->g.l, h->g.m, h->g.n, q - 1,
                       dev_ref(h).r[q - 1]);
    return hip_check(h, hipGetLastError(), "set pops") ? EUP_EHIP : 0;
}

template <typename S> int get_t(Handle *h, int q)
{
    const S *s = static_cast<const S *>(h->buf[h->cur]);
    if (hip_check(h, hipMemsetAsync(h->staging, 0, sizeof(double) * h->g.ghosted, h->stream), "memset"))
        return EUP_EHIP;
    hipLaunchKernelGGL((k_to_staging<S>), dim3((unsigned)((h->g.sites + 255) / 256)), dim3(256), 0,
                       h->stream, s, h->staging, pop_index(h), h->g.l, h->g.m, h->g.n, q - 1,
                       dev_ref(h).r[q - 1]);
    return hip_check(h, hipGetLastError(), "get pops") ? EUP_EHIP : 0;
}

template <typename F> int dispatch(Handle *h, F &&f)
{
    switch (h->cfg.storage) {
    case EUP_STORE_F64: return f((double *)nullptr);
    case EUP_STORE_F32: return f((float *)nullptr);
    case EUP_STORE_F16: return f((_Float16 *)nullptr);
    case EUP_STORE_BF16: return f((bf16_t *)nullptr);
    case EUP_STORE_E3M12: return f((e3m12_t *)nullptr);
    }
    return fail(h, EUP_EINVAL, "unknown storage");
}

} // namespace

int physics_step(Handle *h)
{
    if (h->outlet_dirty) {
        int rc = physics_refresh_outlet(h);
        if (rc) return rc;
    }
    return dispatch(h, [&](auto *p) { return step_t<std::remove_pointer_t<decltype(p)>>(h); });
}

int physics_refresh_outlet(Handle *h)
{
    h->outlet_dirty = false;
    return dispatch(h, [&](auto *p) { return refresh_t<std::remove_pointer_t<decltype(p)>>(h); });
}

int physics_set_pop(Handle *h, int q)
{
    h->outlet_dirty = true;
    return dispatch(h, [&](auto *p) { return set_t<std::remove_pointer_t<decltype(p)>>(h, q); });
}

int physics_get_pop(Handle *h, int q)
{
    return dispatch(h, [&](auto *p) { return get_t<std::remove_pointer_t<decltype(p)>>(h, q); });
}

int init_equilibrium(Handle *h, double ux, double uy, double uz, double noise, uint64_t seed)
{
    const bool parity = h->cfg.mode == EUP_MODE_PARITY;
    const int64_t ns = parity ? h->g.ghosted : h->g.sites;
    const unsigned grid = (unsigned)((ns + 255) / 256);
    const PopIndex P = pop_index(h);
    const int il = parity ? h->g.l + 2 : h->g.l, im = parity ? h->g.m + 2 : h->g.m;
    for (int b = 0; b < 2; ++b) {
        int rc = dispatch(h, [&](auto *p) {
            using S = std::remove_pointer_t<decltype(p)>;
            hipLaunchKernelGGL((k_init_eq<S>), dim3(grid), dim3(256), 0, h->stream, static_cast<S *>(h->buf[b]),
                               P, il, im, ns, ux, uy, uz, noise, seed, dev_ref(h));
            return hip_check(h, hipGetLastError(), "init equilibrium") ? EUP_EHIP : 0;
        });
        if (rc) return rc;
    }
    h->outlet_dirty = true;
    return hip_check(h, hipStreamSynchronize(h->stream), "sync") ? EUP_EHIP : 0;
}

int physics_prepare_mask(Handle *h)
{
    const Geometry &g = h->g;
    const bool noskip = (h->cfg.flags & EUP_FLAG_NO_TILE_SKIP) != 0;
    uint8_t *cls = nullptr;
    uint32_t *link_all = nullptr;
    int rc = 0;
    if (hip_check(h, hipMalloc((void **)&cls, g.tiles), "hipMalloc") ||
        hip_check(h, hipMalloc((void **)&link_all, sizeof(uint32_t) * kTile * g.tiles), "hipMalloc")) {
        hipFree(cls);
        return EUP_EHIP;
    }
    const unsigned grid = (unsigned)((g.tiles + 3) / 4);
    hipLaunchKernelGGL(k_classify, dim3(grid), dim3(kBlock), 0, h->stream, h->solid, g.l, g.m, g.n,
                       g.tiles_per_row, (uint32_t)g.tiles, (int)x_open(h), (int)noskip, cls, link_all);
    if (hip_check(h, hipGetLastError(), "classify")) rc = EUP_EHIP;
    uint8_t *cls_h = nullptr;
    int32_t *code_h = nullptr;
    if (!rc && (hip_check(h, hipHostMalloc((void **)&cls_h, g.tiles), "hipHostMalloc") ||
                hip_check(h, hipHostMalloc((void **)&code_h, sizeof(int32_t) * g.tiles), "hipHostMalloc")))
        rc = EUP_EHIP;
    if (!rc && hip_check(h, hipMemcpyAsync(cls_h, cls, g.tiles, hipMemcpyDeviceToHost, h->stream), "copy") )
        rc = EUP_EHIP;
    if (!rc && hip_check(h, hipStreamSynchronize(h->stream), "sync")) rc = EUP_EHIP;
    int64_t nf = 0, nm = 0, ns = 0;
    if (!rc) {
        for (int64_t t = 0; t < g.tiles; ++t) {
            if (cls_h[t] == 0) { code_h[t] = kTileFluid; ++nf; }
            else if (cls_h[t] == 2) { code_h[t] = kTileSolid; ++ns; }
            else code_h[t] = (int32_t)(nm++);
        }
        h->tiles_fluid = nf;
        h->tiles_mixed = nm;
        h->tiles_solid = ns;
        hipFree(h->tile_code);
        hipFree(h->link_mask);
        h->tile_code = nullptr;
        h->link_mask = nullptr;
        if (hip_check(h, hipMalloc((void **)&h->tile_code, sizeof(int32_t) * g.tiles), "hipMalloc") ||
            hip_check(h, hipMalloc((void **)&h->link_mask, sizeof(uint32_t) * kTile * (nm > 0 ? nm : 1)),
                      "hipMalloc") ||
            hip_check(h, hipMemcpyAsync(h->tile_code, code_h, sizeof(int32_t) * g.tiles,
                                        hipMemcpyHostToDevice, h->stream), "copy"))
            rc = EUP_EHIP;
    }
    if (!rc && nm > 0) {
        hipLaunchKernelGGL(k_compact_links, dim3(grid), dim3(kBlock), 0, h->stream, h->tile_code,
                           link_all, (uint32_t)g.tiles, h->link_mask);
        if (hip_check(h, hipGetLastError(), "compact links")) rc = EUP_EHIP;
    }
    if (!rc && nm > 0) {
        // mixed tile list for the momentum-exchange reduction
        int32_t *list_h = nullptr;
        if (hip_check(h, hipHostMalloc((void **)&list_h, sizeof(int32_t) * nm), "hipHostMalloc")) rc = EUP_EHIP;
        if (!rc) {
            for (int64_t t = 0; t < g.tiles; ++t)
                if (code_h[t] >= 0) list_h[code_h[t]] = (int32_t)t;
            hipFree(h->mixed_list);
            h->mixed_list = nullptr;
            if (hip_check(h, hipMalloc((void **)&h->mixed_list, sizeof(int32_t) * nm), "hipMalloc") ||
                hip_check(h, hipMemcpyAsync(h->mixed_list, list_h, sizeof(int32_t) * nm,
                                            hipMemcpyHostToDevice, h->stream), "copy"))
                rc = EUP_EHIP;
            if (!rc && hip_check(h, hipStreamSynchronize(h->stream), "sync")) rc = EUP_EHIP;
            hipHostFree(list_h);
        }
    }
    if (!rc && hip_check(h, hipStreamSynchronize(h->stream), "sync")) rc = EUP_EHIP;
    hipFree(cls);
    hipFree(link_all);
    if (cls_h) hipHostFree(cls_h);
    if (code_h) hipHostFree(code_h);
    h->outlet_dirty = true;
    return rc;
}

} // namespace eup
