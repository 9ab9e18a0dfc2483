// Parity mode: the reference's serial FP64 time step on the GPU, bit for bit.
//
// One step of CODE/SRC/bgk3d.F90:128-138 under -DSERIAL -DORIGINAL -DOBSTACLES -DCYLINDER_INF is
// bcond_obs -> bcond_bc -> movef -> col on ghosted arrays a01..a19.  Here it is four kernels on
// two ghosted FP64 buffers A (current) and B (next):
//   1. k_bb_gather / k_bb_scatter: the in-place bounce-back of bcond.obs.F90:47-77.  The serial
//      loop reads a neighbour's opposite population, which is already overwritten when that
//      neighbour is solid and comes earlier in (k, j, i) order; in that case the value read is
//      the site's own old population (the neighbour wrote it from us).  Gathering every new value
//      from old values first and scattering afterwards reproduces the serial result exactly,
//      solid sites included, with no race.
//   2. k_ghost: the ghost fills of bcond.bc.F90 for CYLINDER_INF (outflow + border fix :127-204,
//      inflow :260-321, periodic z :385-409/:472-496, periodic y with the SERIAL corner copies
//      :535-560/:600-625).  Every ghost cell is written with the value of its last writer in the
//      reference's statement order; cells the reference never writes keep their value.  Ghosts
//      are written into both A and B, because the next step's bounce-back reads the ghosts left
//      by this step's fill (the one-step-stale seam of the serial path).
//   3. k_stream_collide: movef.F90:43-68 fused with col.F90:96-235 (pull from A, collide, write
//      the interior of B); the FP64 collision is the reference's expression order.  One wavefront
//      per 64-site row segment, so the nine source rows are wave-uniform pointers.
#include <stdlib.h>

#include "eup_internal.h"

namespace eup {

namespace {

struct PArgs {
    const double *A[19];
    double *B[19];
    const double *A0; // population 0 of A; population q at A0 + q * stride
    double *B0;
    int64_t stride;
    int pitch;        // row pitch (elements) of the ghosted layout
    int l, m, n;
    double omega, fgrad, u_in;
    double *outp;     // outlet equilibria of the x = l sites, 5 x m x n (written by the stream-collide)
    const uint8_t *solid;
    UDiv dpitch, dmp; // / pitch and / (m + 2): ghosted offset -> (i, j, k)
};

__device__ __forceinline__ int64_t gidx(int lp, int mp, int i, int j, int k)
{
    return (int64_t)i + (int64_t)lp * ((int64_t)j + (int64_t)mp * (int64_t)k);
}

template <int Q> __device__ __forceinline__ constexpr bool precedes()
{
    // neighbour x + c_Q comes before x in the serial (k, j, i) loop order
    return Dir<Q>::cz < 0 || (Dir<Q>::cz == 0 && (Dir<Q>::cy < 0 || (Dir<Q>::cy == 0 && Dir<Q>::cx < 0)));
}

template <int Q>
__device__ __forceinline__ void bb_one(const PArgs &a, const uint8_t *solid, int i, int j, int k,
                                       int64_t s, double *out)
{
    constexpr int cx = Dir<Q>::cx, cy = Dir<Q>::cy, cz = Dir<Q>::cz, op = Dir<Q>::opp;
    const int lp = a.pitch, mp = a.m + 2;
    const int yi = i + cx, yj = j + cy, yk = k + cz;
    bool ysolid = false;
    if (yi >= 1 && yi <= a.l && yj >= 1 && yj <= a.m && yk >= 1 && yk <= a.n)
        ysolid = solid[(int64_t)(yi - 1) + (int64_t)a.l * ((int64_t)(yj - 1) + (int64_t)a.m * (yk - 1))] != 0;
    if (precedes<Q>() && ysolid)
        out[Q] = a.A[Q][s];
    else
        out[Q] = a.A[op][gidx(lp, mp, yi, yj, yk)];
}

template <int... Qs> struct QSeq {};

template <int... Qs>
__device__ __forceinline__ void bb_all(QSeq<Qs...>, const PArgs &a, const uint8_t *solid, int i,
                                       int j, int k, int64_t s, double *out)
{
    (bb_one<Qs>(a, solid, i, j, k, s, out), ...);
}

using Q18 = QSeq<0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17>;

__global__ void k_bb_gather(PArgs a, const int32_t *list, int64_t nsolid, const uint8_t *solid,
                            double *tmp)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsolid) return;
    const int64_t s = list[t];
    // offsets of the padded ghosted layout stay below 2^31
    const uint32_t rowi = a.dpitch.div((uint32_t)s);
    const uint32_t kk = a.dmp.div(rowi);
    const int i = (int)((uint32_t)s - rowi * (uint32_t)a.pitch);
    const int j = (int)(rowi - kk * (uint32_t)(a.m + 2));
    const int k = (int)kk;
    double out[18];
    bb_all(Q18{}, a, solid, i, j, k, s, out);
    // population-major, so that consecutive solid sites write consecutive values
#pragma unroll
    for (int q = 0; q < 18; ++q) tmp[q * nsolid + t] = out[q];
}

__global__ void k_bb_scatter(PArgs a, const int32_t *list, int64_t nsolid, const double *tmp)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsolid) return;
    const int64_t s = list[t];
#pragma unroll
    for (int q = 0; q < 18; ++q) const_cast<double *>(a.A[q])[s] = tmp[q * nsolid + t];
}

__device__ __forceinline__ void put2(const PArgs &a, int q, int64_t s, double v)
{
    const_cast<double *>(a.A[q])[s] = v;
    a.B[q][s] = v;
}

// flattened work: [0, P) x = l+1 plane, [P, 2P) x = 0 plane, then l*m z-face columns, then
// l*n y-face rows; P = (m+2)(n+2).  The inflow values (x = 0, populations 1..5) are constants that
// nothing else writes except the corner copy a03(0,0,k) of the y fill, so they are written only
// when `inflow` is set (the first step after the host set the state); the corner copy is done
// every step, and so every cell ends each step with the value the reference leaves in it.
__global__ void k_ghost(PArgs a, int inflow)
{
    // inflow = 1 also means that the outlet plane was not written by a stream-collide since the
    // host last set the state: the outflow values are then computed here from A
    const int l = a.l, m = a.m, n = a.n, lp = a.pitch, mp = m + 2;
    const int64_t P = (int64_t)(m + 2) * (n + 2);
    const int64_t Z = (int64_t)l * m, Y = (int64_t)l * n;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < P) { // outflow plane x = l+1
        const int j = (int)(t % (m + 2)), k = (int)(t / (m + 2));
        const bool jin = j >= 1 && j <= m, kin = k >= 1 && k <= n;
        if (!jin && !kin) return; // corners of the plane: never written
        const int sj = jin ? j : (j == 0 ? 1 : m);     // border fix sources (:171-204)
        const int sk = kin ? k : (k == 0 ? 1 : n);
        double o[5];
        // a solid site at x = l had its populations replaced by this step's bounce-back after the
        // stream-collide computed the plane: recompute those from A
        const int64_t mn = (int64_t)m * n, pl = (int64_t)(sj - 1) + (int64_t)m * (sk - 1);
        if (!inflow && a.solid[(int64_t)(l - 1) + (int64_t)l * pl] == 0) {
#pragma unroll
            for (int u = 0; u < 5; ++u) o[u] = a.outp[u * mn + pl];
        } else {
            const int64_t src = gidx(lp, mp, l, sj, sk);
            double x[19];
#pragma unroll
            for (int q = 0; q < 19; ++q) x[q] = a.A[q][src];
            outlet_eq<double>(x, o);
        }
        const int64_t d = gidx(lp, mp, l + 1, j, k);
#pragma unroll
        for (int u = 0; u < 5; ++u) put2(a, 9 + u, d, o[u]);
        if (j == m + 1 && kin) put2(a, 9, d, a.A[9][gidx(lp, mp, 1, 1, k)]); // :612
        return;
    }
    t -= P;
    if (t < P) { // inflow plane x = 0
        const int j = (int)(t % (m + 2)), k = (int)(t / (m + 2));
        const int64_t d = gidx(lp, mp, 0, j, k);
        if (inflow) {
            double o[5];
            inlet_eq<double>(a.u_in, o);
#pragma unroll
            for (int u = 0; u < 5; ++u) put2(a, u, d, o[u]);
        }
        if (j == 0 && k >= 1 && k <= n) put2(a, 2, d, a.A[2][gidx(lp, mp, l, m, k)]); // :547
        return;
    }
    t -= P;
    if (t < Z) { // z = n+1 and z = 0
        const int i = 1 + (int)(t % l), j = 1 + (int)(t / l);
        const int64_t up = gidx(lp, mp, i, j, n + 1), s1 = gidx(lp, mp, i, j, 1);
        const int64_t dn = gidx(lp, mp, i, j, 0), sn = gidx(lp, mp, i, j, n);
        const int qu[5] = {1, 8, 10, 14, 15}; // a02 a09 a11 a15 a16
        const int qd[5] = {3, 5, 6, 12, 17};  // a04 a06 a07 a13 a18
#pragma unroll
        for (int u = 0; u < 5; ++u) put2(a, qu[u], up, a.A[qu[u]][s1]);
#pragma unroll
        for (int u = 0; u < 5; ++u) put2(a, qd[u], dn, a.A[qd[u]][sn]);
        return;
    }
    t -= Z;
    if (t < Y) { // y = 0 and y = m+1
        const int i = 1 + (int)(t % l), k = 1 + (int)(t / l);
        const int64_t lf = gidx(lp, mp, i, 0, k), sm = gidx(lp, mp, i, m, k);
        const int64_t rt = gidx(lp, mp, i, m + 1, k), s1 = gidx(lp, mp, i, 1, k);
        const int ql[5] = {2, 6, 7, 8, 11};   // a03 a07 a08 a09 a12
        const int qr[5] = {0, 9, 15, 16, 17}; // a01 a10 a16 a17 a18
#pragma unroll
        for (int u = 0; u < 5; ++u) put2(a, ql[u], lf, a.A[ql[u]][sm]);
#pragma unroll
        for (int u = 0; u < 5; ++u) put2(a, qr[u], rt, a.A[qr[u]][s1]);
    }
}

// pull of population Q into x: the source row (j - c_y, k - c_z) is wave-uniform, so its pointer
// is formed in scalar registers (base pointer + population and row offsets); the lane adds the
// zero-extended x offset i - c_x >= 0 (ghost column 0 included)
template <int Q>
__device__ __forceinline__ void pull_one(const PArgs &a, int64_t row, int64_t lp, int64_t lpmp, int i,
                                         double *x)
{
    constexpr int cx = Dir<Q>::cx, cy = Dir<Q>::cy, cz = Dir<Q>::cz;
    const double *p = a.A0 + ((int64_t)Q * a.stride + row - cy * lp - cz * lpmp);
    x[Q] = p[(uint64_t)(uint32_t)(i - cx)];
}

template <int... Qs>
__device__ __forceinline__ void pull_all(QSeq<Qs...>, const PArgs &a, int64_t row, int64_t lp,
                                         int64_t lpmp, int i, double *x)
{
    (pull_one<Qs>(a, row, lp, lpmp, i, x), ...);
}

// one wavefront per 64-site segment of an interior row (tpr segments per row): j and k are
// wave-uniform, i = 1 + 64 * segment + lane
struct SegMap {
    int tpr;
    uint32_t segments;
    UDiv dtpr, dm;
    uint32_t chunk; // blocks per XCD (order 1)
    uint32_t grp;   // blocks of one row, rounded up (order 2)
    UDiv dgrp;      // n / grp
    uint32_t order; // 0 dispatch, 1 XCD-contiguous ranges of segments, 2 XCD-owned rows
};

__global__ __launch_bounds__(256) void k_stream_collide(PArgs a, SegMap s)
{
    // block order as in the physics kernel: hardware block b runs on XCD b % 8; order 2 deals
    // each group of 8 * grp blocks out so that every XCD takes a whole row (the lines that
    // x-neighbouring segments share through their x-shifted pulls stay in one L2), a last
    // partial group keeps dispatch order
    const uint32_t b = blockIdx.x;
    uint32_t lb = b;
    if (s.order == 1) {
        lb = (b & 7u) * s.chunk + (b >> 3);
    } else if (s.order == 2) {
        const uint32_t r8 = b >> 3, gq = s.dgrp.div(r8), g0 = gq * 8u * s.grp;
        if (g0 + 8u * s.grp <= gridDim.x) lb = g0 + (b & 7u) * s.grp + (r8 - gq * s.grp);
    }
    const uint32_t w = __builtin_amdgcn_readfirstlane(lb * 4u + (threadIdx.x >> 6));
    if (w >= s.segments) return;
    const uint32_t rowi = s.dtpr.div(w);
    const int seg = (int)(w - rowi * (uint32_t)s.tpr);
    const uint32_t kk = s.dm.div(rowi);
    const int j = 1 + (int)(rowi - kk * (uint32_t)a.m);
    const int k = 1 + (int)kk;
    const int i = 1 + seg * 64 + (int)(threadIdx.x & 63);
    if (i > a.l) return;
    const int64_t lp = a.pitch, lpmp = lp * (a.m + 2);
    const int64_t row = lp * j + lpmp * k;
    double x[19];
    pull_all(Q18{}, a, row, lp, lpmp, i, x);
    x[18] = (a.A0 + (18 * a.stride + row))[(uint64_t)(uint32_t)i];
    collide<double>(x, x, a.omega, a.fgrad);
    // non-temporal: the next reads of these values come a step later
#pragma unroll
    for (int q = 0; q < 19; ++q)
        __builtin_nontemporal_store(x[q], &(a.B0 + ((int64_t)q * a.stride + row))[(uint64_t)(uint32_t)i]);
    if (i == a.l) {
        // the next step's outflow ghost values of this row (bcond.bc.F90:127-164), from the values
        // just stored; the ghost fill uses them unless bounce-back changes this site first
        double o[5];
        outlet_eq<double>(x, o);
        const int64_t mn = (int64_t)a.m * a.n, pl = (int64_t)(j - 1) + (int64_t)a.m * (k - 1);
#pragma unroll
        for (int u = 0; u < 5; ++u) a.outp[u * mn + pl] = o[u];
    }
}

// one population between the reference's ghosted array (row pitch l + 2; the staging buffer) and
// the padded layout (row pitch `pitch`), ghosted site t = i + (l+2) (j + (m+2) k)
__global__ void k_parity_put(const double *__restrict__ stg, double *__restrict__ d0, double *__restrict__ d1,
                             int l, int m, int n, int pitch)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)(l + 2) * (m + 2) * (n + 2)) return;
    const int i = (int)(t % (l + 2));
    const int64_t r = t / (l + 2);
    const int64_t e = i + (int64_t)pitch * r;
    d0[e] = stg[t];
    d1[e] = stg[t];
}

__global__ void k_parity_take(const double *__restrict__ src, double *__restrict__ stg, int l, int m, int n,
                              int pitch)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)(l + 2) * (m + 2) * (n + 2)) return;
    const int i = (int)(t % (l + 2));
    const int64_t r = t / (l + 2);
    stg[t] = src[i + (int64_t)pitch * r];
}

PArgs make_args(Handle *h, int from)
{
    PArgs a;
    double *src = static_cast<double *>(h->buf[from]);
    double *dst = static_cast<double *>(h->buf[from ^ 1]);
    // every population pointer sits on its offset base (ghosted site (0, 0, 0))
    for (int q = 0; q < 19; ++q) {
        a.A[q] = src + (int64_t)q * h->stride + h->poff;
        a.B[q] = dst + (int64_t)q * h->stride + h->poff;
    }
    a.A0 = src + h->poff;
    a.B0 = dst + h->poff;
    a.stride = h->stride;
    a.pitch = h->ppitch;
    a.l = h->g.l;
    a.m = h->g.m;
    a.n = h->g.n;
    a.omega = h->cfg.omega;
    a.fgrad = h->cfg.fgrad;
    a.u_in = h->cfg.u_inflow;
    a.outp = static_cast<double *>(h->outlet[0]);
    a.solid = h->solid;
    a.dpitch = UDiv::make((uint32_t)h->ppitch);
    a.dmp = UDiv::make((uint32_t)(h->g.m + 2));
    return a;
}

inline unsigned nblocks(int64_t work, int bs) { return (unsigned)((work + bs - 1) / bs); }

} // namespace

int parity_prepare_mask(Handle *h)
{
    // solid list in the reference's bbox loop order (k, j, i), as offsets in the padded layout
    const Geometry &g = h->g;
    uint8_t *sol_h = nullptr;
    if (hip_check(h, hipHostMalloc((void **)&sol_h, g.sites), "hipHostMalloc")) return EUP_EHIP;
    if (hip_check(h, hipMemcpy(sol_h, h->solid, g.sites, hipMemcpyDeviceToHost), "hipMemcpy solid")) {
        hipHostFree(sol_h);
        return EUP_EHIP;
    }
    int64_t cnt = 0;
    for (int64_t s = 0; s < g.sites; ++s) cnt += sol_h[s] != 0;
    int32_t *list_h = nullptr;
    if (cnt > 0 && hip_check(h, hipHostMalloc((void **)&list_h, sizeof(int32_t) * cnt), "hipHostMalloc")) {
        hipHostFree(sol_h);
        return EUP_EHIP;
    }
    int64_t c = 0;
    for (int k = 1; k <= g.n; ++k)
        for (int j = 1; j <= g.m; ++j)
            for (int i = 1; i <= g.l; ++i) {
                int64_t o = (int64_t)(i - 1) + (int64_t)g.l * ((int64_t)(j - 1) + (int64_t)g.m * (k - 1));
                if (sol_h[o] != 0) list_h[c++] = (int32_t)parity_index(h, i, j, k);
            }
    hipHostFree(sol_h);
    hipFree(h->solid_list);
    hipFree(h->bb_tmp);
    h->solid_list = nullptr;
    h->bb_tmp = nullptr;
    int rc = 0;
    if (cnt > 0) {
        if (hip_check(h, hipMalloc((void **)&h->solid_list, sizeof(int32_t) * cnt), "hipMalloc") ||
            hip_check(h, hipMalloc((void **)&h->bb_tmp, sizeof(double) * 18 * cnt), "hipMalloc") ||
            hip_check(h, hipMemcpy(h->solid_list, list_h, sizeof(int32_t) * cnt, hipMemcpyHostToDevice),
                      "hipMemcpy list"))
            rc = EUP_EHIP;
    }
    if (list_h) hipHostFree(list_h);
    h->nsolid = cnt;
    return rc;
}

int parity_step(Handle *h)
{
    const Geometry &g = h->g;
    PArgs a = make_args(h, h->cur);
    if (h->nsolid > 0) {
        hipLaunchKernelGGL(k_bb_gather, dim3(nblocks(h->nsolid, 256)), dim3(256), 0, h->stream, a,
                           h->solid_list, h->nsolid, h->solid, h->bb_tmp);
        hipLaunchKernelGGL(k_bb_scatter, dim3(nblocks(h->nsolid, 256)), dim3(256), 0, h->stream, a,
                           h->solid_list, h->nsolid, h->bb_tmp);
    }
    const int64_t gwork = 2 * (int64_t)(g.m + 2) * (g.n + 2) + (int64_t)g.l * g.m + (int64_t)g.l * g.n;
    // the host's set / init calls mark the ghosts dirty (outlet_dirty): fill the inflow plane
    hipLaunchKernelGGL(k_ghost, dim3(nblocks(gwork, 256)), dim3(256), 0, h->stream, a, h->outlet_dirty ? 1 : 0);
    h->outlet_dirty = false;
    SegMap sm;
    sm.tpr = (g.l + 63) / 64;
    sm.segments = (uint32_t)((int64_t)sm.tpr * g.m * g.n);
    sm.dtpr = UDiv::make((uint32_t)sm.tpr);
    sm.dm = UDiv::make((uint32_t)g.m);
    const uint32_t nb = (sm.segments + 3) / 4;
    sm.chunk = (nb + 7) / 8;
    sm.grp = (uint32_t)((sm.tpr + 3) / 4);
    sm.dgrp = UDiv::make(sm.grp);
    // XCD-owned rows (order 2).  On the 512 x 256 x 256 sponge (MI355X) order 2 ran 17 150-17 180
    // MLUPS against 17 100 in dispatch order (0), which had measured 5 % faster than contiguous
    // ranges (1).  EUP_PARITY_XCD_REMAP=0/1/2 forces the order.
    sm.order = 2;
    if (const char *e = getenv("EUP_PARITY_XCD_REMAP")) {
        const int v = atoi(e);
        sm.order = v < 0 ? 0u : (v > 2 ? 2u : (uint32_t)v);
    }
    if (h->parity_lds == (size_t)-1) {
        // occupancy cap of the stream-collide (as for the fused FP64 kernel): 3 blocks per CU run
        // about 4 % faster on MI355X than the register-limited occupancy;
        // EUP_PARITY_BLOCKS_PER_CU overrides (0 = no cap)
        int cap = 3;
        if (const char *e = getenv("EUP_PARITY_BLOCKS_PER_CU")) cap = atoi(e);
        h->parity_lds = occupancy_lds_bytes(h, cap);
    }
    hipLaunchKernelGGL(k_stream_collide, dim3(8u * sm.chunk), dim3(256), h->parity_lds, h->stream, a, sm);
    if (hip_check(h, hipGetLastError(), "parity step launch")) return EUP_EHIP;
    h->cur ^= 1;
    return 0;
}

int parity_set_pop(Handle *h, int q)
{
    const Geometry &g = h->g;
    double *d0 = static_cast<double *>(h->buf[0]) + (int64_t)(q - 1) * h->stride + h->poff;
    double *d1 = static_cast<double *>(h->buf[1]) + (int64_t)(q - 1) * h->stride + h->poff;
    hipLaunchKernelGGL(k_parity_put, dim3(nblocks(g.ghosted, 256)), dim3(256), 0, h->stream, h->staging, d0, d1,
                       g.l, g.m, g.n, h->ppitch);
    return hip_check(h, hipGetLastError(), "set pops") ? EUP_EHIP : 0;
}

int parity_get_pop(Handle *h, int q)
{
    const Geometry &g = h->g;
    const double *src = static_cast<const double *>(h->buf[h->cur]) + (int64_t)(q - 1) * h->stride + h->poff;
    hipLaunchKernelGGL(k_parity_take, dim3(nblocks(g.ghosted, 256)), dim3(256), 0, h->stream, src, h->staging,
                       g.l, g.m, g.n, h->ppitch);
    return hip_check(h, hipGetLastError(), "get pops") ? EUP_EHIP : 0;
}

} // namespace eup
