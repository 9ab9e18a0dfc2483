// Diagnostics in FP64: probe (probe_global.F90:41-77), control-volume stress drag (flux_X.F90,
// flux_Y.F90), momentum-exchange drag (no reference counterpart), global means (diagno.F90) and the
// velocity / vorticity fields of the binary VTK writers (vtk_*.binary.F90).
//
// Populations are converted to double on load whatever the storage.  The control-volume drag
// computes one stress per face site on the GPU with the reference's expression order and returns
// them in the reference's loop order, so the host sums them exactly as flux_X/flux_Y do.  The
// momentum-exchange force and the global means are wavefront-reduced on the GPU in a fixed order
// (deterministic run to run).
#include <mutex>
#include <type_traits>
#include <vector>

#include "eup_internal.h"

// Device scratch of eup_drag_cv_noalloc: the per-site stresses of the two faces of one axis,
// at most kNoAllocFace sites per face (16 MiB per GPU, held for the life of the process).  Host
// code never passes this name to a kernel or a copy: it is a host-side stand-in; the device
// address comes from hipGetSymbolAddress, once per device (noalloc_scratch).
static constexpr int64_t kNoAllocFace = int64_t(1) << 20;
static __device__ __align__(256) double scratch[2 * kNoAllocFace];

namespace eup {

namespace {

constexpr int kRedBlocks = 2048;

// population q of interior site (i, j, k) (1-based) as double
struct ParityAcc {
    const double *p[19]; // population q at its offset base
    int l, m, n, pitch;
    __device__ __forceinline__ double operator()(int q, int i, int j, int k) const
    {
        return p[q][(int64_t)i + (int64_t)pitch * ((int64_t)j + (int64_t)(m + 2) * k)];
    }
};

template <typename S> struct PhysAcc {
    using value_type = S;
    const S *base;
    PopIndex ix;
    DevRef ref;
    int l, m, n;
    __device__ __forceinline__ double operator()(int q, int i, int j, int k) const
    {
        return Store<S>::load64(base[ix(q, i - 1, j - 1, k - 1)], ref.r[q], q);
    }
};

template <typename Acc>
__global__ void k_site19(Acc acc, int i, int j, int k, double *out)
{
    const int q = threadIdx.x;
    if (q < 19) out[q] = acc(q, i, j, k);
}

// flux_X.F90:173-242 / flux_Y.F90:168-242 for one site; axis 0 -> sigma_xx, 1 -> sigma_xy
__device__ double site_sigma(const double *x, double omega, int axis)
{
    const double P1 = 1.0 / 18.0, P2 = 1.0 / 36.0, RF = 3.0, QF = 1.5;
    double rho = +x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9] + x[10]
                 + x[11] + x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18];
    double rhoinv = 1.0 / rho;
    double xj = +(x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rhoinv;
    double yj = +(x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rhoinv;
    double zj = +(x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rhoinv;
    double cvsq = xj * xj + yj * yj + zj * zj;
    double fe[18];
    fe[0] = rho * P2 * (1.0 + RF * (xj - yj) + QF * (3.0 * (xj - yj) * (xj - yj) - cvsq));
    fe[1] = rho * P2 * (1.0 + RF * (xj - zj) + QF * (3.0 * (xj - zj) * (xj - zj) - cvsq));
    fe[2] = rho * P2 * (1.0 + RF * (xj + yj) + QF * (3.0 * (xj + yj) * (xj + yj) - cvsq));
    fe[3] = rho * P2 * (1.0 + RF * (xj + zj) + QF * (3.0 * (xj + zj) * (xj + zj) - cvsq));
    fe[4] = rho * P1 * (1.0 + RF * (xj) + QF * (3.0 * (xj) * (xj)-cvsq));
    fe[5] = rho * P1 * (1.0 + RF * (zj) + QF * (3.0 * (zj) * (zj)-cvsq));
    fe[6] = rho * P2 * (1.0 + RF * (yj + zj) + QF * (3.0 * (yj + zj) * (yj + zj) - cvsq));
    fe[7] = rho * P1 * (1.0 + RF * (yj) + QF * (3.0 * (yj) * (yj)-cvsq));
    fe[8] = rho * P2 * (1.0 + RF * (yj - zj) + QF * (3.0 * (yj - zj) * (yj - zj) - cvsq));
    fe[9] = rho * P2 * (1.0 + RF * (-xj - yj) + QF * (3.0 * (xj + yj) * (xj + yj) - cvsq));
    fe[10] = rho * P2 * (1.0 + RF * (-xj - zj) + QF * (3.0 * (xj + zj) * (xj + zj) - cvsq));
    fe[11] = rho * P2 * (1.0 + RF * (-xj + yj) + QF * (3.0 * (xj - yj) * (xj - yj) - cvsq));
    fe[12] = rho * P2 * (1.0 + RF * (-xj + zj) + QF * (3.0 * (xj - zj) * (xj - zj) - cvsq));
    fe[13] = rho * P1 * (1.0 + RF * (-xj) + QF * (3.0 * (xj) * (xj)-cvsq));
    fe[14] = rho * P1 * (1.0 + RF * (-zj) + QF * (3.0 * (zj) * (zj)-cvsq));
    fe[15] = rho * P2 * (1.0 + RF * (-yj - zj) + QF * (3.0 * (yj + zj) * (yj + zj) - cvsq));
    fe[16] = rho * P1 * (1.0 + RF * (-yj) + QF * (3.0 * (yj) * (yj)-cvsq));
    fe[17] = rho * P2 * (1.0 + RF * (-yj + zj) + QF * (3.0 * (yj - zj) * (yj - zj) - cvsq));
    const double cxd[18] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0};
    const double cyd[18] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1};
    const double c = (1 - 0.5 * omega);
    double sg;
    if (axis == 0) {
        sg = -rho / RF;
        for (int q = 0; q < 18; ++q) sg = sg - c * cxd[q] * cxd[q] * (x[q] - fe[q]);
    } else {
        sg = 0.0;
        for (int q = 0; q < 18; ++q) sg = sg - c * cxd[q] * cyd[q] * (x[q] - fe[q]);
    }
    return sg;
}

// face sites of one drag face in the reference's loop order (k outer, then j or i)
template <typename Acc>
__global__ void k_sigma(Acc acc, int axis, int plane, int lo, int hi, int k1, int k2, double omega,
                        double *out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nin = hi - lo + 1;
    const int64_t total = (int64_t)nin * (k2 - k1 + 1);
    if (t >= total) return;
    const int u = lo + (int)(t % nin);
    const int k = k1 + (int)(t / nin);
    const int i = axis == 0 ? plane : u;
    const int j = axis == 0 ? u : plane;
    double x[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) x[q] = acc(q, i, j, k);
    out[t] = site_sigma(x, omega, axis);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// block partial sums of NV values; lane 0 of the block writes partial[blockIdx.x * NV + v]
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double *partial)
{
    __shared__ double sh[NV][kBlock / 64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double s = wave_sum(v[c]);
        if (lane == 0) sh[c][w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            double s = 0.0;
            for (int ww = 0; ww < kBlock / 64; ++ww) s += sh[c][ww];
            partial[(int64_t)blockIdx.x * NV + c] = s;
        }
    }
}

template <int NV>
__global__ void k_final(const double *partial, int nblocks, double *out)
{
    double v[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) v[c] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock)
#pragma unroll
        for (int c = 0; c < NV; ++c) v[c] += partial[(int64_t)b * NV + c];
    block_reduce<NV>(v, out);
}

// momentum exchange over all interior sites (parity layout), neighbour test on the solid flags
template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_mem_sites(Acc acc, const uint8_t *solid, int xopen,
                                                      double *partial)
{
    const int l = acc.l, m = acc.m, n = acc.n;
    const int64_t N = (int64_t)l * m * n;
    double v[3] = {0.0, 0.0, 0.0};
    const int cxs[18] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0};
    const int cys[18] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1};
    const int czs[18] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1};
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < N; s += (int64_t)gridDim.x * blockDim.x) {
        if (solid[s]) continue;
        const int i = (int)(s % l);
        const int64_t r = s / l;
        const int j = (int)(r % m), k = (int)(r / m);
        for (int p = 0; p < 18; ++p) {
            int ti = i + cxs[p], tj = (j + cys[p] + m) % m, tk = (k + czs[p] + n) % n;
            if (ti < 0 || ti >= l) {
                if (xopen) continue;
                ti = (ti + l) % l;
            }
            if (!solid[(int64_t)ti + (int64_t)l * ((int64_t)tj + (int64_t)m * tk)]) continue;
            const double f2 = 2.0 * acc(p, i + 1, j + 1, k + 1);
            v[0] += cxs[p] * f2;
            v[1] += cys[p] * f2;
            v[2] += czs[p] * f2;
        }
    }
    block_reduce<3>(v, partial);
}

// momentum exchange over the mixed tiles only (physics layout): bit opp(p) of the link mask of a
// fluid site says that x + c_p is solid
template <typename S>
__global__ __launch_bounds__(kBlock) void k_mem_tiles(PhysAcc<S> acc, const int32_t *mixed_list,
                                                      int64_t nmixed, const uint32_t *link, int tpr,
                                                      double *partial)
{
    const int cxs[18] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0};
    const int cys[18] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1};
    const int czs[18] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1};
    const int opp[18] = {11, 12, 9, 10, 13, 14, 15, 16, 17, 2, 3, 0, 1, 4, 5, 6, 7, 8};
    double v[3] = {0.0, 0.0, 0.0};
    const int lane = threadIdx.x & 63;
    const int64_t wpb = kBlock / 64;
    for (int64_t c = (int64_t)blockIdx.x * wpb + (threadIdx.x >> 6); c < nmixed; c += (int64_t)gridDim.x * wpb) {
        const uint32_t lm = link[c * kTile + lane];
        if (lm == 0 || (lm & 0x80000000u)) continue;
        const int64_t t = mixed_list[c];
        const int64_t row = t / tpr;
        const int i = (int)(t - row * tpr) * kTile + lane;
        if (i >= acc.l) continue;
        const int j = (int)(row % acc.m), k = (int)(row / acc.m);
        for (int p = 0; p < 18; ++p) {
            if (!((lm >> opp[p]) & 1u)) continue;
            const double f2 = 2.0 * acc(p, i + 1, j + 1, k + 1);
            v[0] += cxs[p] * f2;
            v[1] += cys[p] * f2;
            v[2] += czs[p] * f2;
        }
    }
    block_reduce<3>(v, partial);
}

// velocity of one site as the reference's binary VTK writers compute it (vtk_xy.binary.F90:93-115):
// momenta and density summed in FP64 in the written order, each rounded to float, divided in
// float (IEEE division; the translation unit is built without contraction)
template <typename Acc>
__device__ __forceinline__ void vtk_site(const Acc &acc, int i, int j, int k, float *o)
{
    double x[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) x[q] = acc(q, i, j, k);
    const float u = (float)(x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]);
    const float w = (float)(x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]);
    const float v = (float)(x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]);
    const float den = (float)(x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9]
                              + x[10] + x[11] + x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18]);
    o[0] = u / den;
    o[1] = w / den;
    o[2] = v / den;
}

// one plane of velocities in the writers' loop order: plane 0 = yz at x = index (k, j), 1 = xz at
// y = index (k, i), 2 = xy at z = index (j, i)
template <typename Acc>
__global__ void k_vtk_velocity(Acc acc, int plane, int index, float *out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = acc.l, m = acc.m, n = acc.n;
    int i, j, k;
    if (plane == 0) {
        if (t >= (int64_t)m * n) return;
        i = index; j = 1 + (int)(t % m); k = 1 + (int)(t / m);
    } else if (plane == 1) {
        if (t >= (int64_t)l * n) return;
        i = 1 + (int)(t % l); j = index; k = 1 + (int)(t / l);
    } else {
        if (t >= (int64_t)l * m) return;
        i = 1 + (int)(t % l); j = 1 + (int)(t / l); k = index;
    }
    vtk_site(acc, i, j, k, out + 3 * t);
}

// vorticity of the xy plane z = k0 as vtk_om.binary.F90:88-131 computes it (i = 2..l-1 fastest,
// j = 2..m-1): densities rounded to float; in each momentum only the last population is divided
// by the density, as the writer's expression reads; the difference is formed in float
template <typename Acc>
__global__ void k_vtk_vorticity(Acc acc, int k0, float *out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = acc.l, m = acc.m;
    if (t >= (int64_t)(l - 2) * (m - 2)) return;
    const int i = 2 + (int)(t % (l - 2)), j = 2 + (int)(t / (l - 2));
    const int di[4] = {0, 0, -1, 1}, dj[4] = {1, -1, 0, 0}; // up, down, left, right
    float mom[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        double x[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) x[q] = acc(q, i + di[s], j + dj[s], k0);
        const float d = (float)(x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9]
                                + x[10] + x[11] + x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18]);
        if (s < 2)
            mom[s] = (float)(x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13] / (double)d);
        else
            mom[s] = (float)(x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17] / (double)d);
    }
    out[t] = 0.5f * (mom[0] - mom[1]) - 0.5f * (mom[3] - mom[2]);
}

template <typename Acc>
__global__ __launch_bounds__(kBlock) void k_diagno(Acc acc, const uint8_t *solid, double *partial)
{
    const int l = acc.l, m = acc.m, n = acc.n;
    const int64_t N = (int64_t)l * m * n;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < N; s += (int64_t)gridDim.x * blockDim.x) {
        double rho, xj, yj, zj;
        if (solid[s]) {
            rho = 1.0; xj = 0.0; yj = 0.0; zj = 0.0;
        } else {
            const int i = 1 + (int)(s % l);
            const int64_t r = s / l;
            const int j = 1 + (int)(r % m), k = 1 + (int)(r / m);
            double x[19];
#pragma unroll
            for (int q = 0; q < 19; ++q) x[q] = acc(q, i, j, k);
            rho = +x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9] + x[10]
                  + x[11] + x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18];
            const double rhoinv = 1.0 / rho;
            xj = +(x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rhoinv;
            yj = +(x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rhoinv;
            zj = +(x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rhoinv;
        }
        v[0] += rho;
        v[1] += xj;
        v[2] += yj;
        v[3] += zj;
        v[4] += xj * xj + yj * yj + zj * zj;
    }
    block_reduce<5>(v, partial);
}

ParityAcc parity_acc(Handle *h)
{
    ParityAcc a;
    const double *base = static_cast<const double *>(h->buf[h->cur]);
    for (int q = 0; q < 19; ++q) a.p[q] = base + (int64_t)q * h->stride + h->poff;
    a.l = h->g.l;
    a.m = h->g.m;
    a.n = h->g.n;
    a.pitch = h->ppitch;
    return a;
}

template <typename S> PhysAcc<S> phys_acc(Handle *h)
{
    PhysAcc<S> a;
    a.base = static_cast<const S *>(h->buf[h->cur]);
    a.ix = pop_index(h);
    a.ref = dev_ref(h);
    a.l = h->g.l;
    a.m = h->g.m;
    a.n = h->g.n;
    return a;
}

// run f(acc) with the accessor of the handle's layout
template <typename F> int with_acc(Handle *h, F &&f)
{
    if (h->cfg.mode == EUP_MODE_PARITY) return f(parity_acc(h));
    switch (h->cfg.storage) {
    case EUP_STORE_F64: return f(phys_acc<double>(h));
    case EUP_STORE_F32: return f(phys_acc<float>(h));
    case EUP_STORE_F16: return f(phys_acc<_Float16>(h));
    case EUP_STORE_BF16: return f(phys_acc<bf16_t>(h));
    case EUP_STORE_E3M12: return f(phys_acc<e3m12_t>(h));
    }
    return fail(h, EUP_EINVAL, "unknown storage");
}

int ensure_diag(Handle *h, int64_t len)
{
    if (h->diag_len >= len) return 0;
    hipFree(h->diag);
    h->diag = nullptr;
    h->diag_len = 0;
    if (hip_check(h, hipMalloc((void **)&h->diag, sizeof(double) * len), "hipMalloc diag")) return EUP_EHIP;
    h->diag_len = len;
    return 0;
}

} // namespace

int diag_site(Handle *h, int i, int j, int k, double out[19])
{
    if (!h->site19 && hip_check(h, hipMalloc((void **)&h->site19, sizeof(double) * 32), "hipMalloc"))
        return EUP_EHIP;
    int rc = with_acc(h, [&](auto acc) {
        hipLaunchKernelGGL((k_site19<decltype(acc)>), dim3(1), dim3(64), 0, h->stream, acc, i, j, k,
                           h->site19);
        return hip_check(h, hipGetLastError(), "site gather") ? EUP_EHIP : 0;
    });
    if (rc) return rc;
    if (hip_check(h, hipMemcpyAsync(out, h->site19, sizeof(double) * 19, hipMemcpyDeviceToHost, h->stream),
                  "copy site") ||
        hip_check(h, hipStreamSynchronize(h->stream), "sync"))
        return EUP_EHIP;
    return 0;
}

// index clipping of flux_X.F90:76-172 / flux_Y.F90:76-165 for a single task (offsets 0)
static bool clip_faces(const Geometry &g, const int32_t box[6], int axis, int *f1, int *f2,
                       double *n1, double *n2, int *lo, int *hi, int *k1, int *k2)
{
    const int L = axis == 0 ? g.l : g.m, M = axis == 0 ? g.m : g.l, n = g.n;
    const int s = axis == 0 ? box[0] : box[2], e = axis == 0 ? box[1] : box[3];
    const int os = axis == 0 ? box[2] : box[0], oe = axis == 0 ? box[3] : box[1];
    *n1 = 1.0;
    *n2 = 1.0;
    *f1 = *f2 = 0;
    if (s > 0 && e <= L) { *f1 = s; *f2 = e; }
    if (s > 0 && e > L) { *f1 = s; *f2 = L; *n2 = 0.0; }
    if (s <= 0 && e <= L) { *f1 = 1; *n1 = 0.0; *f2 = e; }
    if (s < 0 && e > L) return false;
    if (s > L || e < 0) return false;
    *lo = *hi = 0;
    if (os > 0 && oe <= M) { *lo = os; *hi = oe; }
    if (os > 0 && oe > M) { *lo = os; *hi = M; }
    if (os <= 0 && oe <= M) { *lo = 1; *hi = oe; }
    if (os < 0 && oe > M) { *lo = 1; *hi = M; }
    if (os > M || oe < 0) return false;
    *k1 = *k2 = 0;
    const int ks = box[4], ke = box[5];
    if (ks > 0 && ke <= n) { *k1 = ks; *k2 = ke; }
    if (ks > 0 && ke > n) { *k1 = ks; *k2 = n; }
    if (ks <= 0 && ke <= n) { *k1 = 1; *k2 = ke; }
    if (ks < 0 && ke > n) { *k1 = 1; *k2 = n; }
    if (ks > n || ke < 0) return false;
    return true;
}

// fixed != nullptr: the per-site stresses go to that device array (room for 2 * fixed_cap values),
// and are summed on the host through a stack buffer, so nothing is allocated
static int drag_axis(Handle *h, const int32_t box[6], int axis, double *res, double *fixed = nullptr,
                     int64_t fixed_cap = 0)
{
    *res = 0.0;
    int f[2], lo, hi, k1, k2;
    double n1, n2;
    if (!clip_faces(h->g, box, axis, &f[0], &f[1], &n1, &n2, &lo, &hi, &k1, &k2)) return 0;
    const int lim_in = axis == 0 ? h->g.m : h->g.l;
    const int lim_f = axis == 0 ? h->g.l : h->g.m;
    // a box ending at 0 puts its second face on plane 0: the ghost plane of the reference's arrays,
    // which only parity mode holds; the other layouts have no plane 0
    const int fmin = h->cfg.mode == EUP_MODE_PARITY ? 0 : 1;
    if (lo < 1 || hi > lim_in || k1 < 1 || k2 > h->g.n || f[0] < 1 || f[1] < fmin || f[1] > lim_f)
        return fail(h, EUP_ERANGE, "drag box outside the lattice after clipping");
    const int64_t per = (int64_t)(hi - lo + 1) * (k2 - k1 + 1);
    if (per <= 0) return 0;
    int rc = 0;
    double *buf = fixed;
    if (fixed) {
        if (per > fixed_cap)
            return fail(h, EUP_EINVAL, "eup_drag_cv_noalloc: a face of " + std::to_string(per) +
                                           " sites exceeds the no-allocation scratch (at most " +
                                           std::to_string(fixed_cap) + " sites per face)");
    } else {
        rc = ensure_diag(h, 2 * per);
        if (rc) return rc;
        buf = h->diag;
    }
    for (int ii = 0; ii < 2; ++ii) {
        rc = with_acc(h, [&](auto acc) {
            hipLaunchKernelGGL((k_sigma<decltype(acc)>), dim3((unsigned)((per + 255) / 256)), dim3(256), 0,
                               h->stream, acc, axis, f[ii], lo, hi, k1, k2, h->cfg.omega, buf + ii * per);
            return hip_check(h, hipGetLastError(), "sigma") ? EUP_EHIP : 0;
        });
        if (rc) return rc;
    }
    if (fixed) {
        // the same left-to-right sums as below, fed through a fixed host buffer
        double part[4096];
        double fl[2] = {0.0, 0.0};
        for (int ii = 0; ii < 2; ++ii)
            for (int64_t off = 0; off < per; off += 4096) {
                const int64_t cnt = per - off < 4096 ? per - off : 4096;
                if (hip_check(h, hipMemcpyAsync(part, buf + ii * per + off, sizeof(double) * cnt,
                                                hipMemcpyDeviceToHost, h->stream), "copy sigma") ||
                    hip_check(h, hipStreamSynchronize(h->stream), "sync"))
                    return EUP_EHIP;
                for (int64_t t = 0; t < cnt; ++t) fl[ii] = fl[ii] + part[t];
            }
        *res = axis == 0 ? -fl[0] * n1 + fl[1] * n2 : +fl[0] * n1 + fl[1] * n2;
        return 0;
    }
    std::vector<double> sg(2 * per);
    if (hip_check(h, hipMemcpyAsync(sg.data(), h->diag, sizeof(double) * 2 * per, hipMemcpyDeviceToHost,
                                    h->stream), "copy sigma") ||
        hip_check(h, hipStreamSynchronize(h->stream), "sync"))
        return EUP_EHIP;
    double fl[2] = {0.0, 0.0};
    for (int ii = 0; ii < 2; ++ii)
        for (int64_t t = 0; t < per; ++t) fl[ii] = fl[ii] + sg[ii * per + t];
    *res = axis == 0 ? -fl[0] * n1 + fl[1] * n2 : +fl[0] * n1 + fl[1] * n2;
    return 0;
}

int diag_drag_cv(Handle *h, const int32_t box[6], double *fx, double *fy)
{
    int rc = drag_axis(h, box, 0, fx);
    if (rc) return rc;
    return drag_axis(h, box, 1, fy);
}

// Device address of the no-allocation scratch on the handle's device (current when called):
// looked up once per device and kept.
static int noalloc_scratch(Handle *h, double **out)
{
    static std::mutex mu;
    static double *ptr[64] = {};
    const int dev = h->cfg.device;
    if (dev < 0 || dev >= 64) return fail(h, EUP_EINVAL, "eup_drag_cv_noalloc: device ordinal above 63");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        if (hip_check(h, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
            return EUP_EHIP;
        ptr[dev] = static_cast<double *>(p);
    }
    *out = ptr[dev];
    return 0;
}

int diag_drag_cv_noalloc(Handle *h, const int32_t box[6], double *fx, double *fy)
{
    double *s = nullptr;
    int rc = noalloc_scratch(h, &s);
    if (rc) return rc;
    rc = drag_axis(h, box, 0, fx, s, kNoAllocFace);
    if (rc) return rc;
    return drag_axis(h, box, 1, fy, s, kNoAllocFace);
}

int diag_drag_mem(Handle *h, double F[3])
{
    int rc = ensure_diag(h, 3 * kRedBlocks + 8);
    if (rc) return rc;
    const int xopen = h->cfg.front != EUP_FACE_PERIODIC;
    if (h->cfg.mode == EUP_MODE_PARITY) {
        hipLaunchKernelGGL((k_mem_sites<ParityAcc>), dim3(kRedBlocks), dim3(kBlock), 0, h->stream,
                           parity_acc(h), h->solid, xopen, h->diag);
    } else if (h->tiles_mixed > 0) {
        rc = with_acc(h, [&](auto acc) -> int {
            using A = decltype(acc);
            if constexpr (std::is_same_v<A, ParityAcc>) {
                return 0;
            } else {
                using S = typename A::value_type;
                hipLaunchKernelGGL((k_mem_tiles<S>), dim3(kRedBlocks), dim3(kBlock), 0, h->stream, acc,
                                   h->mixed_list, h->tiles_mixed, h->link_mask, h->g.tiles_per_row,
                                   h->diag);
                return 0;
            }
        });
        if (rc) return rc;
    } else {
        F[0] = F[1] = F[2] = 0.0;
        return 0;
    }
    double *res = h->diag + 3 * kRedBlocks;
    hipLaunchKernelGGL((k_final<3>), dim3(1), dim3(kBlock), 0, h->stream, h->diag, kRedBlocks, res);
    if (hip_check(h, hipGetLastError(), "mem drag") ||
        hip_check(h, hipMemcpyAsync(F, res, sizeof(double) * 3, hipMemcpyDeviceToHost, h->stream), "copy") ||
        hip_check(h, hipStreamSynchronize(h->stream), "sync"))
        return EUP_EHIP;
    return 0;
}

// one velocity plane into out (host), through a plane-sized device buffer
static int vtk_plane(Handle *h, int plane, int index, float *out)
{
    const Geometry &g = h->g;
    const int64_t count = plane == 0 ? (int64_t)g.m * g.n : (plane == 1 ? (int64_t)g.l * g.n : (int64_t)g.l * g.m);
    int rc = ensure_diag(h, (3 * count + 1) / 2 + 1);
    if (rc) return rc;
    float *dev = reinterpret_cast<float *>(h->diag);
    rc = with_acc(h, [&](auto acc) {
        hipLaunchKernelGGL((k_vtk_velocity<decltype(acc)>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                           h->stream, acc, plane, index, dev);
        return hip_check(h, hipGetLastError(), "vtk velocity") ? EUP_EHIP : 0;
    });
    if (rc) return rc;
    if (hip_check(h, hipMemcpyAsync(out, dev, sizeof(float) * 3 * count, hipMemcpyDeviceToHost, h->stream),
                  "copy vtk") ||
        hip_check(h, hipStreamSynchronize(h->stream), "sync"))
        return EUP_EHIP;
    return 0;
}

int diag_vtk_velocity(Handle *h, int plane, int index, float *out)
{
    if (plane != 3) return vtk_plane(h, plane, index, out);
    // the volume in the writer's order (k, j, i) = the xy planes one after the other
    const int64_t per = (int64_t)h->g.l * h->g.m * 3;
    for (int k = 1; k <= h->g.n; ++k) {
        int rc = vtk_plane(h, 2, k, out + (k - 1) * per);
        if (rc) return rc;
    }
    return 0;
}

int diag_vtk_vorticity(Handle *h, int k0, float *out)
{
    const int64_t count = (int64_t)(h->g.l - 2) * (h->g.m - 2);
    int rc = ensure_diag(h, count / 2 + 1);
    if (rc) return rc;
    float *dev = reinterpret_cast<float *>(h->diag);
    rc = with_acc(h, [&](auto acc) {
        hipLaunchKernelGGL((k_vtk_vorticity<decltype(acc)>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                           h->stream, acc, k0, dev);
        return hip_check(h, hipGetLastError(), "vtk vorticity") ? EUP_EHIP : 0;
    });
    if (rc) return rc;
    if (hip_check(h, hipMemcpyAsync(out, dev, sizeof(float) * count, hipMemcpyDeviceToHost, h->stream),
                  "copy vtk") ||
        hip_check(h, hipStreamSynchronize(h->stream), "sync"))
        return EUP_EHIP;
    return 0;
}

int diag_diagno(Handle *h, double out[5])
{
    int rc = ensure_diag(h, 5 * kRedBlocks + 8);
    if (rc) return rc;
    rc = with_acc(h, [&](auto acc) {
        hipLaunchKernelGGL((k_diagno<decltype(acc)>), dim3(kRedBlocks), dim3(kBlock), 0, h->stream, acc,
                           h->solid, h->diag);
        return 0;
    });
    if (rc) return rc;
    double *res = h->diag + 5 * kRedBlocks;
    hipLaunchKernelGGL((k_final<5>), dim3(1), dim3(kBlock), 0, h->stream, h->diag, kRedBlocks, res);
    double tot[5];
    if (hip_check(h, hipGetLastError(), "diagno") ||
        hip_check(h, hipMemcpyAsync(tot, res, sizeof(double) * 5, hipMemcpyDeviceToHost, h->stream), "copy") ||
        hip_check(h, hipStreamSynchronize(h->stream), "sync"))
        return EUP_EHIP;
    const double rdv = 1. / ((double)h->g.l * (double)h->g.m * (double)h->g.n);
    for (int c = 0; c < 5; ++c) out[c] = tot[c] * rdv;
    return 0;
}

} // namespace eup
