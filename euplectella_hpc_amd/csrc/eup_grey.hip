// Grey-lattice mode: partial bounce-back on a solid-fraction field s in [0, 1], the drag
// objective averaged over a step window, and its discrete adjoint with checkpoint-recompute.
// (No reference counterpart; the binary-mask limit is the physics step's bounce-back.)
//
// Link model: the link between x and its source y = x - c_q has the fraction
//   sigma = 1 - (1 - s(x))(1 - s(y))          (s = 0 on the inlet / outlet ghosts)
// and the pulled population is f_q(x) = (1 - sigma) src_q + sigma a_qbar(x), with src_q = a_q(y),
// the inlet equilibrium or the outlet equilibrium of bcond.bc.F90.  sigma is symmetric in the
// link, so every link conserves mass; s in {0, 1} gives the halfway bounce-back exactly.
// Force on the grey body: F = sum over interior links (x, q) of sigma c_q (a_q(y) - a_qbar(x)).
// Objective: J = (1 / ((t1 - t0 + 1) norm)) sum_{t = t0..t1} F_x(a^t).
//
// Adjoint: lam^t = (d a^{t+1} / d a^t)^T lam^{t+1} + [t >= t0] w dF_x/da(a^t), and
// dJ/ds accumulates (d a^{t+1} / ds)^T lam^{t+1} + [t >= t0] w dF_x/ds(a^t).  Per step two kernels:
// k_adj_collide recomputes the pulled populations f(x) of a^t and applies the transposed BGK
// Jacobian to lam^{t+1}, and adds the gradient terms of the links the site pulls through;
// k_adj_stream gathers the transposed pull (all gathers, no atomics):
//   lam_p(z) = (1 - sigma_l) lin_p(z + c_p) + sigma_l lin_pbar(z)   (sigma_l of link z -- z + c_p)
// plus the outlet-equilibrium transpose on the last x column and the objective terms, and adds the
// gradient terms of the links of which the site is the source (FP64 accumulation).
// Inside a recomputed segment k_adj_fused (the FP64 default) runs the stream half of step t and the
// collide half of step t - 1 in one pass, so lam^t is never stored.
// States are checkpointed every K steps; each segment is recomputed forward and stored before its
// backward sweep, so memory is (t1/K + K) states instead of t1.
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "eup_internal.h"

namespace eup {

namespace {

// direction table (0-based population index; bgk3d.F90:31-49), folded into the fully unrolled
// loops below
constexpr int kCX[19] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
constexpr int kCY[19] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1, 0};
constexpr int kCZ[19] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1, 0};
constexpr int kOPP[19] = {11, 12, 9, 10, 13, 14, 15, 16, 17, 2, 3, 0, 1, 4, 5, 6, 7, 8, 18};

// a state: 19 SoA arrays of l*m*n sites, population q at A + q * stride
template <typename R> struct GArgs {
    const R *A; // the state read
    R *B;       // the state written, or null
    int64_t stride;
    const R *s;
    R inl[5];
    R omega;
    int l, m, n;
    int xopen;
    __device__ __forceinline__ R a(int q, int64_t x) const { return A[q * stride + x]; }
};

struct Nbr {
    int l, m, n, xopen;
    // site index of (i + dx, j + dy, k + dz), wrapping y and z (and x when periodic); -1 = inlet
    // ghost, -2 = outlet ghost
    __device__ __forceinline__ int64_t at(int i, int j, int k, int dx, int dy, int dz) const
    {
        int ii = i + dx, jj = j + dy, kk = k + dz;
        jj = jj < 0 ? jj + m : (jj >= m ? jj - m : jj);
        kk = kk < 0 ? kk + n : (kk >= n ? kk - n : kk);
        if (ii < 0) {
            if (xopen) return -1;
            ii += l;
        } else if (ii >= l) {
            if (xopen) return -2;
            ii -= l;
        }
        return (int64_t)ii + (int64_t)l * ((int64_t)jj + (int64_t)m * kk);
    }
};

template <typename R> __device__ __forceinline__ R wq(int q)
{
    return q == 18 ? R(1) / R(3) : ((kCX[q] * kCX[q] + kCY[q] * kCY[q] + kCZ[q] * kCZ[q]) == 1 ? R(1) / R(18) : R(1) / R(36));
}

// link q of site (i,j,k) (0-based): sv the source value, syv the source's solid fraction (0 on a
// ghost), kd 0 interior, 1 inlet, 2 outlet.  Called inside fully unrolled loops, so q is a
// compile-time constant.
template <typename R>
__device__ __forceinline__ void grey_link(const GArgs<R> &a, const Nbr &nb, int i, int j, int k, int q, R &sv,
                                          R &syv, int &kd)
{
    const int64_t y = nb.at(i, j, k, -kCX[q], -kCY[q], -kCZ[q]);
    if (kCX[q] == 1 && y == -1) { // only c_x = +1 pulls from the inlet ghost
        sv = a.inl[kCX[q] == 1 ? q : 0];
        syv = R(0);
        kd = 1;
    } else if (kCX[q] == -1 && y == -2) { // only c_x = -1 pulls from the outlet ghost
        const int64_t z = nb.at(a.l - 1, j, k, 0, -kCY[q], -kCZ[q]);
        R yv[19], o[5];
#pragma unroll
        for (int p = 0; p < 19; ++p) yv[p] = a.a(p, z);
        outlet_eq<R>(yv, o);
        sv = o[kCX[q] == -1 ? q - 9 : 0];
        syv = R(0);
        kd = 2;
    } else {
        sv = a.a(q, y);
        syv = a.s[y];
        kd = 0;
    }
}

// pulled populations f[19] of site (i,j,k) (0-based): per link sigma = 1 - (1 - s(x))(1 - s(y))
// and f_q = (1 - sigma) src_q + sigma a_qbar(x)
template <typename R>
__device__ __forceinline__ void grey_pull(const GArgs<R> &a, const Nbr &nb, int i, int j, int k, int64_t x, R *f)
{
    const R sx = a.s[x];
#pragma unroll
    for (int q = 0; q < 18; ++q) {
        R sv, syv;
        int kd;
        grey_link<R>(a, nb, i, j, k, q, sv, syv, kd);
        const R sg = R(1) - (R(1) - sx) * (R(1) - syv);
        f[q] = (R(1) - sg) * sv + sg * a.a(kOPP[q], x);
    }
    f[18] = a.a(18, x);
}

// grey_pull that also returns the links: sig[q] the link fraction, src[q] the source value,
// sy[q] the source's solid fraction (0 on a ghost), kind[q] 0 interior, 1 inlet, 2 outlet.  The
// same arithmetic as grey_link + grey_pull, written out: the FP32 adjoint collide built on it
// takes 166 registers per lane, built on grey_link 186 (three against two wavefronts per SIMD).
template <typename R>
__device__ __forceinline__ void grey_pull_links(const GArgs<R> &a, const Nbr &nb, int i, int j, int k, int64_t x, R *f,
                                          R *sig, R *src, R *sy, int *kind)
{
    const R sx = a.s[x];
#pragma unroll
    for (int q = 0; q < 18; ++q) {
        const int64_t y = nb.at(i, j, k, -kCX[q], -kCY[q], -kCZ[q]);
        R sv, syv;
        int kd;
        if (kCX[q] == 1 && y == -1) { // only c_x = +1 pulls from the inlet ghost
            sv = a.inl[kCX[q] == 1 ? q : 0];
            syv = R(0);
            kd = 1;
        } else if (kCX[q] == -1 && y == -2) { // only c_x = -1 pulls from the outlet ghost
            const int64_t z = nb.at(a.l - 1, j, k, 0, -kCY[q], -kCZ[q]);
            R yv[19], o[5];
#pragma unroll
            for (int p = 0; p < 19; ++p) yv[p] = a.a(p, z);
            outlet_eq<R>(yv, o);
            sv = o[kCX[q] == -1 ? q - 9 : 0];
            syv = R(0);
            kd = 2;
        } else {
            sv = a.a(q, y);
            syv = a.s[y];
            kd = 0;
        }
        const R sg = R(1) - (R(1) - sx) * (R(1) - syv);
        sig[q] = sg;
        src[q] = sv;
        sy[q] = syv;
        kind[q] = kd;
        f[q] = (R(1) - sg) * sv + sg * a.a(kOPP[q], x);
    }
    f[18] = a.a(18, x);
}

template <typename R> __global__ __launch_bounds__(256) void k_grey_step(GArgs<R> a)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t N = (int64_t)a.l * a.m * a.n;
    if (x >= N) return;
    const int i = (int)(x % a.l);
    const int64_t r = x / a.l;
    const int j = (int)(r % a.m), k = (int)(r / a.m);
    Nbr nb{a.l, a.m, a.n, a.xopen};
    R f[19];
    grey_pull<R>(a, nb, i, j, k, x, f);
    collide<R>(f, f, a.omega, R(0));
#pragma unroll
    for (int q = 0; q < 19; ++q) a.B[q * a.stride + x] = f[q];
}

__device__ __forceinline__ double wsum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// block partial sums of the force (3 doubles per block)
template <typename R> __global__ __launch_bounds__(256) void k_grey_force(GArgs<R> a, double *partial)
{
    const int64_t N = (int64_t)a.l * a.m * a.n;
    Nbr nb{a.l, a.m, a.n, a.xopen};
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < N; x += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(x % a.l);
        const int64_t r = x / a.l;
        const int j = (int)(r % a.m), k = (int)(r / a.m);
        const R sx = a.s[x];
#pragma unroll
        for (int q = 0; q < 18; ++q) {
            R sv, syv;
            int kd;
            grey_link<R>(a, nb, i, j, k, q, sv, syv, kd);
            if (kd != 0) continue;
            const R sg = R(1) - (R(1) - sx) * (R(1) - syv);
            const double d = (double)sg * ((double)sv - (double)a.a(kOPP[q], x));
            v[0] += kCX[q] * d;
            v[1] += kCY[q] * d;
            v[2] += kCZ[q] * d;
        }
    }
    __shared__ double sh[3][4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int c = 0; c < 3; ++c) {
        double t = wsum(v[c]);
        if (lane == 0) sh[c][w] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) partial[blockIdx.x * 3 + c] = sh[c][0] + sh[c][1] + sh[c][2] + sh[c][3];
}

// acc[c] += weight * sum of partials (single block, fixed order)
__global__ void k_force_acc(const double *partial, int nb, double weight, double *acc, double *last)
{
    __shared__ double sh[3][256];
    double v[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        for (int c = 0; c < 3; ++c) v[c] += partial[b * 3 + c];
    for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) {
            last[c] = sh[c][0];
            acc[c] += weight * sh[c][0];
        }
}

// transposed BGK Jacobian at pre-collision f applied to lam
template <typename R> __device__ __forceinline__ void collide_adj(const R *f, const R *lam, R omega, R *out)
{
    R rho = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        rho += f[q];
        jx += kCX[q] * f[q];
        jy += kCY[q] * f[q];
        jz += kCZ[q] * f[q];
    }
    const R ux = jx / rho, uy = jy / rho, uz = jz / rho, usq = ux * ux + uy * uy + uz * uz;
    R A = 0, Bx = 0, By = 0, Bz = 0;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        const R w = wq<R>(q);
        const R cu = kCX[q] * ux + kCY[q] * uy + kCZ[q] * uz;
        A += lam[q] * w * (R(1) - R(4.5) * cu * cu + R(1.5) * usq);
        const R g = lam[q] * w * (R(3) + R(9) * cu);
        Bx += g * kCX[q] - lam[q] * w * R(3) * ux;
        By += g * kCY[q] - lam[q] * w * R(3) * uy;
        Bz += g * kCZ[q] - lam[q] * w * R(3) * uz;
    }
#pragma unroll
    for (int p = 0; p < 19; ++p)
        out[p] = (R(1) - omega) * lam[p] + omega * (A + Bx * kCX[p] + By * kCY[p] + Bz * kCZ[p]);
}

template <typename R> struct AdjArgs {
    GArgs<R> g;
    const R *lam_in; // lam^{t+1} (the state layout of g)
    R *lin;          // J_C^T lam^{t+1} at every site
    R *lam_out;      // lam^t
    double *gs;
    double wobj;     // objective weight at time t (0 outside the window)
};

// lin(x) = J_C(f(x))^T lam^{t+1}(x) from lam[19] = lam^{t+1}(x) (overwritten), stored to a.lin;
// g gets the gradient terms of the links that x pulls through (x as the pulling site:
// d f_q(x) / d s(x) on every link, plus the objective's force term)
template <typename R>
__device__ __forceinline__ void adj_collide_site(const AdjArgs<R> &a, const Nbr &nb, int i, int j, int k, int64_t x,
                                                 R *lam, double &g)
{
    const int64_t st = a.g.stride;
    // FP32 keeps each link's source value, source fraction and kind from the pull for the
    // gradient terms; FP64 reads them again (cache hits) instead, which lowers its registers from
    // 306 to 248 per lane (one to two wavefronts per SIMD).  Measured on MI355X (128 x 64 x 64):
    // FP64 117.5 -> 112.6 us, while the FP32 kernel would go from 42 to 51-62 us.
    const double w = a.wobj;
    if constexpr (sizeof(R) == 4) {
        R f[19], sig[18], src[18], sy[18], out[19];
        int kind[18];
        grey_pull_links<R>(a.g, nb, i, j, k, x, f, sig, src, sy, kind);
        collide_adj<R>(f, lam, a.g.omega, out);
#pragma unroll
        for (int q = 0; q < 19; ++q) a.lin[q * st + x] = out[q];
#pragma unroll
        for (int q = 0; q < 18; ++q) {
            const double L = (double)out[q];
            const double abar = (double)a.g.a(kOPP[q], x);
            const double syq = (double)sy[q];
            g += L * (abar - (double)src[q]) * (1.0 - syq);
            if (kind[q] == 0 && w != 0.0) g += w * kCX[q] * ((double)src[q] - abar) * (1.0 - syq);
        }
    } else {
        R f[19];
        grey_pull<R>(a.g, nb, i, j, k, x, f);
        collide_adj<R>(f, lam, a.g.omega, lam); // in place: out[p] reads only lam[p] after the sums
#pragma unroll
        for (int q = 0; q < 19; ++q) a.lin[q * st + x] = lam[q];
#pragma unroll
        for (int q = 0; q < 18; ++q) {
            R sv, syv;
            int kd;
            grey_link<R>(a.g, nb, i, j, k, q, sv, syv, kd);
            const double L = (double)lam[q];
            const double abar = (double)a.g.a(kOPP[q], x);
            const double syq = (double)syv;
            g += L * (abar - (double)sv) * (1.0 - syq);
            if (kd == 0 && w != 0.0) g += w * kCX[q] * ((double)sv - abar) * (1.0 - syq);
        }
    }
}

// lam^t(z) into lam[19]: the transposed pull gathered from a.lin at the sites z + c_p that pull
// from z, the outlet-equilibrium transpose on the last x column and the objective's force term;
// g gets the gradient terms of the links of which z is the source
template <typename R>
__device__ __forceinline__ void adj_stream_site(const AdjArgs<R> &a, const Nbr &nb, int i, int j, int k, int64_t z,
                                                R *lam, double &g)
{
    const int l = a.g.l;
    const int64_t st = a.g.stride;
    const double w = a.wobj;
    const R sz = a.g.s[z];
#pragma unroll
    for (int p = 0; p < 18; ++p) {
        const int pb = kOPP[p];
        const int64_t zp = nb.at(i, j, k, kCX[p], kCY[p], kCZ[p]); // the site pulling p from z
        double lp;
        if (zp >= 0) {
            const R s2 = a.g.s[zp];
            const double sl = 1.0 - (1.0 - (double)sz) * (1.0 - (double)s2);
            const double Lp = (double)a.lin[p * st + zp];
            lp = (1.0 - sl) * Lp + sl * (double)a.lin[pb * st + z];
            if (w != 0.0) lp += 2.0 * w * sl * kCX[p];
            const double d = (double)a.g.a(pb, zp) - (double)a.g.a(p, z);
            g += Lp * d * (1.0 - (double)s2);
            if (w != 0.0) g -= w * kCX[p] * d * (1.0 - (double)s2);
        } else {
            lp = (double)sz * (double)a.lin[pb * st + z]; // ghost link: sigma = s(z)
        }
        lam[p] = (R)lp;
    }
    lam[18] = a.lin[18 * st + z];
    // outlet equilibrium transpose: z on the last x column feeds the ghost values read by the
    // sites x = (l-1, j + cy, k + cz) for q = 10..14 (0-based 9..13)
    if (a.g.xopen && i == l - 1) {
        double Jx = 0, Jy = 0, Jz = 0;
#pragma unroll
        for (int p = 0; p < 19; ++p) {
            const double yp = (double)a.g.a(p, z);
            Jx += kCX[p] * yp;
            Jy += kCY[p] * yp;
            Jz += kCZ[p] * yp;
        }
        double Lw[5], cJ[5];
#pragma unroll
        for (int q = 9; q < 14; ++q) {
            const int64_t x = nb.at(l - 1, j, k, 0, kCY[q], kCZ[q]);
            const double sg = (double)a.g.s[x]; // sigma of the ghost link
            const double L = (1.0 - sg) * (double)a.lin[q * st + x];
            Lw[q - 9] = L * (double)wq<double>(q);
            cJ[q - 9] = kCX[q] * Jx + kCY[q] * Jy + kCZ[q] * Jz;
        }
        // per population p the five terms in the order q = 10..14, one p at a time
#pragma unroll
        for (int p = 0; p < 19; ++p) {
            const double Jc = Jx * kCX[p] + Jy * kCY[p] + Jz * kCZ[p];
            double acc = 0.0;
#pragma unroll
            for (int q = 9; q < 14; ++q) {
                const double cc = kCX[q] * kCX[p] + kCY[q] * kCY[p] + kCZ[q] * kCZ[p];
                acc += Lw[q - 9] * (3.0 * cc + 9.0 * cJ[q - 9] * cc - 3.0 * Jc);
            }
            lam[p] = (R)((double)lam[p] + acc);
        }
    }
}

// one thread per site: (i, j, k) of site x, false past the lattice
__device__ __forceinline__ bool site_of(int l, int m, int n, int64_t &x, int &i, int &j, int &k)
{
    x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (int64_t)l * m * n) return false;
    i = (int)(x % l);
    const int64_t r = x / l;
    j = (int)(r % m);
    k = (int)(r / m);
    return true;
}

// the collide half of an adjoint step on its own: lin = J_C^T lam_in, and the gradient terms of
// the links that x pulls through.  The same arithmetic as adj_collide_site written out with lam
// read after the pull: built on adj_collide_site the FP32 kernel takes 170 registers per lane
// instead of 166 (two wavefronts per SIMD instead of three; 47 against 42 us on 128 x 64 x 64).
template <typename R> __global__ __launch_bounds__(256) void k_adj_collide(AdjArgs<R> a)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t N = (int64_t)a.g.l * a.g.m * a.g.n;
    if (x >= N) return;
    const int i = (int)(x % a.g.l);
    const int64_t r = x / a.g.l;
    const int j = (int)(r % a.g.m), k = (int)(r / a.g.m);
    const int64_t st = a.g.stride;
    Nbr nb{a.g.l, a.g.m, a.g.n, a.g.xopen};
    const double w = a.wobj;
    double g = 0.0;
    if constexpr (sizeof(R) == 4) {
        R f[19], sig[18], src[18], sy[18], lam[19], out[19];
        int kind[18];
        grey_pull_links<R>(a.g, nb, i, j, k, x, f, sig, src, sy, kind);
#pragma unroll
        for (int q = 0; q < 19; ++q) lam[q] = a.lam_in[q * st + x];
        collide_adj<R>(f, lam, a.g.omega, out);
#pragma unroll
        for (int q = 0; q < 19; ++q) a.lin[q * st + x] = out[q];
#pragma unroll
        for (int q = 0; q < 18; ++q) {
            const double L = (double)out[q];
            const double abar = (double)a.g.a(kOPP[q], x);
            const double syq = (double)sy[q];
            g += L * (abar - (double)src[q]) * (1.0 - syq);
            if (kind[q] == 0 && w != 0.0) g += w * kCX[q] * ((double)src[q] - abar) * (1.0 - syq);
        }
    } else {
        R f[19], lam[19];
        grey_pull<R>(a.g, nb, i, j, k, x, f);
#pragma unroll
        for (int q = 0; q < 19; ++q) lam[q] = a.lam_in[q * st + x];
        collide_adj<R>(f, lam, a.g.omega, lam); // in place: out[p] reads only lam[p] after the sums
#pragma unroll
        for (int q = 0; q < 19; ++q) a.lin[q * st + x] = lam[q];
#pragma unroll
        for (int q = 0; q < 18; ++q) {
            R sv, syv;
            int kd;
            grey_link<R>(a.g, nb, i, j, k, q, sv, syv, kd);
            const double L = (double)lam[q];
            const double abar = (double)a.g.a(kOPP[q], x);
            const double syq = (double)syv;
            g += L * (abar - (double)sv) * (1.0 - syq);
            if (kd == 0 && w != 0.0) g += w * kCX[q] * ((double)sv - abar) * (1.0 - syq);
        }
    }
    a.gs[x] += g;
}

// the stream half of an adjoint step on its own: lam_out = P^T lin (+ objective terms)
template <typename R> __global__ __launch_bounds__(256) void k_adj_stream(AdjArgs<R> a)
{
    int64_t z;
    int i, j, k;
    if (!site_of(a.g.l, a.g.m, a.g.n, z, i, j, k)) return;
    const int64_t st = a.g.stride;
    Nbr nb{a.g.l, a.g.m, a.g.n, a.g.xopen};
    R lam[19];
    double g = 0.0;
    adj_stream_site<R>(a, nb, i, j, k, z, lam, g);
#pragma unroll
    for (int p = 0; p < 19; ++p) a.lam_out[p * st + z] = lam[p];
    a.gs[z] += g;
}

// the stream half of step t (sa: state a^t, lin^t in sa.lin) and the collide half of step t - 1
// (ca: state a^{t-1}, lin^{t-1} into ca.lin) in one pass: lam^t stays in registers, so the
// adjoint field is read and written once per step instead of twice.  The gradient sum of each
// site takes the two halves' terms in the same order as the split kernels, so the result is
// bit-identical to running k_adj_stream(t) and k_adj_collide(t - 1).
template <typename R> __global__ __launch_bounds__(256) void k_adj_fused(AdjArgs<R> sa, AdjArgs<R> ca)
{
    int64_t z;
    int i, j, k;
    if (!site_of(sa.g.l, sa.g.m, sa.g.n, z, i, j, k)) return;
    Nbr nb{sa.g.l, sa.g.m, sa.g.n, sa.g.xopen};
    R lam[19];
    double g1 = 0.0, g2 = 0.0;
    adj_stream_site<R>(sa, nb, i, j, k, z, lam, g1);
    adj_collide_site<R>(ca, nb, i, j, k, z, lam, g2);
    double v = sa.gs[z];
    v += g1;
    v += g2;
    sa.gs[z] = v;
}

// objective seed at t1: lam = w dF/da, gs += w dF/ds
template <typename R> __global__ __launch_bounds__(256) void k_force_adj_seed(AdjArgs<R> a)
{
    const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int l = a.g.l, m = a.g.m, n = a.g.n;
    const int64_t N = (int64_t)l * m * n;
    if (z >= N) return;
    const int i = (int)(z % l);
    const int64_t r = z / l;
    const int j = (int)(r % m), k = (int)(r / m);
    const int64_t st = a.g.stride;
    Nbr nb{l, m, n, a.g.xopen};
    const double w = a.wobj;
    const double sz = (double)a.g.s[z];
    double g = 0.0;
#pragma unroll
    for (int q = 0; q < 18; ++q) { // z as x: pair (z, q) with interior source y = z - c_q
        const int64_t y = nb.at(i, j, k, -kCX[q], -kCY[q], -kCZ[q]);
        if (y < 0) continue;
        const double d = (double)a.g.a(q, y) - (double)a.g.a(kOPP[q], z);
        g += w * kCX[q] * d * (1.0 - (double)a.g.s[y]);
    }
#pragma unroll
    for (int p = 0; p < 18; ++p) {
        const int64_t zp = nb.at(i, j, k, kCX[p], kCY[p], kCZ[p]);
        double lam = 0.0;
        if (zp >= 0) {
            const double s2 = (double)a.g.s[zp];
            const double sl = 1.0 - (1.0 - sz) * (1.0 - s2);
            lam = 2.0 * w * sl * kCX[p];
            const double d = (double)a.g.a(p, z) - (double)a.g.a(kOPP[p], zp);
            g += w * kCX[p] * d * (1.0 - s2);
        }
        a.lam_out[p * st + z] = (R)lam;
    }
    a.lam_out[18 * st + z] = R(0);
    a.gs[z] += g;
}

template <typename R> GArgs<R> gargs(Handle *h, const void *from, void *to)
{
    GArgs<R> a;
    a.A = static_cast<const R *>(from);
    a.B = static_cast<R *>(to);
    a.stride = h->stride;
    a.s = static_cast<const R *>(h->grey);
    R inl[5];
    inlet_eq<R>((R)h->cfg.u_inflow, inl);
    for (int u = 0; u < 5; ++u) a.inl[u] = inl[u];
    a.omega = (R)h->cfg.omega;
    a.l = h->g.l;
    a.m = h->g.m;
    a.n = h->g.n;
    a.xopen = h->cfg.front != EUP_FACE_PERIODIC;
    return a;
}

template <typename R> unsigned grid_sites(Handle *h) { return (unsigned)((h->g.sites + 255) / 256); }

constexpr int kForceBlocks = 1024;

template <typename R> int grey_step_t(Handle *h)
{
    GArgs<R> a = gargs<R>(h, h->buf[h->cur], h->buf[h->cur ^ 1]);
    hipLaunchKernelGGL((k_grey_step<R>), dim3(grid_sites<R>(h)), dim3(256), 0, h->stream, a);
    if (hip_check(h, hipGetLastError(), "grey step")) return EUP_EHIP;
    h->cur ^= 1;
    return 0;
}

template <typename R> int force_into(Handle *h, const void *state, double weight, double *acc, double *last)
{
    GArgs<R> a = gargs<R>(h, state, nullptr);
    hipLaunchKernelGGL((k_grey_force<R>), dim3(kForceBlocks), dim3(256), 0, h->stream, a, h->diag);
    hipLaunchKernelGGL(k_force_acc, dim3(1), dim3(256), 0, h->stream, h->diag, kForceBlocks, weight, acc, last);
    return hip_check(h, hipGetLastError(), "grey force") ? EUP_EHIP : 0;
}

template <typename R>
int objective_t(Handle *h, int t0, int t1, int K, double norm, double *J, void *grad_dev, int gdtype,
                double *grad_host)
{
    const int64_t N = h->g.sites;
    const size_t S = sizeof(R) * (size_t)h->stride * 19; // one state
    const int nck = t1 / K + 1;
    const bool want_grad = grad_dev || grad_host;
    std::vector<void *> mem;
    auto alloc = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        mem.push_back(p);
        return p;
    };
    auto cleanup = [&]() {
        for (void *p : mem) hipFree(p);
    };
    if (h->diag_len < 3 * kForceBlocks + 16) {
        hipFree(h->diag);
        h->diag = nullptr;
        h->diag_len = 0;
        if (hip_check(h, hipMalloc((void **)&h->diag, sizeof(double) * (3 * kForceBlocks + 16)), "hipMalloc"))
            return EUP_EHIP;
        h->diag_len = 3 * kForceBlocks + 16;
    }
    double *acc = h->diag + 3 * kForceBlocks;     // [0..2] accumulator, [3..5] last force
    char *ck = static_cast<char *>(alloc(S * nck));
    void *work[2] = {alloc(S), alloc(S)};
    if (!ck || !work[0] || !work[1]) {
        cleanup();
        return fail(h, EUP_EHIP, "eup_grey_objective: out of device memory for checkpoints");
    }
    if (hip_check(h, hipMemsetAsync(acc, 0, sizeof(double) * 6, h->stream), "memset")) {
        cleanup();
        return EUP_EHIP;
    }
    // forward sweep from the current state
    void *init = h->buf[h->cur];
    if (hip_check(h, hipMemcpyAsync(work[0], init, S, hipMemcpyDeviceToDevice, h->stream), "copy")) {
        cleanup();
        return EUP_EHIP;
    }
    int c = 0;
    for (int t = 0; t <= t1; ++t) {
        void *cur = work[c];
        if (t % K == 0 && hip_check(h, hipMemcpyAsync(ck + S * (size_t)(t / K), cur, S, hipMemcpyDeviceToDevice,
                                                      h->stream), "checkpoint copy")) {
            cleanup();
            return EUP_EHIP;
        }
        if (t >= t0 && force_into<R>(h, cur, 1.0, acc, acc + 3)) {
            cleanup();
            return EUP_EHIP;
        }
        if (t == t1) break;
        GArgs<R> a = gargs<R>(h, cur, work[c ^ 1]);
        hipLaunchKernelGGL((k_grey_step<R>), dim3(grid_sites<R>(h)), dim3(256), 0, h->stream, a);
        c ^= 1;
    }
    double sum[3];
    if (hip_check(h, hipGetLastError(), "forward sweep") ||
        hip_check(h, hipMemcpyAsync(sum, acc, sizeof(double) * 3, hipMemcpyDeviceToHost, h->stream), "copy") ||
        hip_check(h, hipStreamSynchronize(h->stream), "sync")) {
        cleanup();
        return EUP_EHIP;
    }
    const double wobj = 1.0 / ((double)(t1 - t0 + 1) * norm);
    *J = sum[0] * wobj;
    if (want_grad) {
        // a segment holds at most min(K, t1) forward states (t - tb < te - tb <= min(K, t1))
        char *seg = static_cast<char *>(alloc(S * (size_t)(K < t1 ? K : t1)));
        void *lam = alloc(S);                    // lam^t between the halves of a step
        void *linb[2] = {alloc(S), alloc(S)};    // lin^t, lin^{t-1}
        double *gs = static_cast<double *>(alloc(sizeof(double) * N));
        if (!seg || !lam || !linb[0] || !linb[1] || !gs) {
            cleanup();
            return fail(h, EUP_EHIP, "eup_grey_objective: out of device memory for the adjoint");
        }
        if (hip_check(h, hipMemsetAsync(gs, 0, sizeof(double) * N, h->stream), "memset gradient")) {
            cleanup();
            return EUP_EHIP;
        }
        auto adj = [&](const void *state, const void *lam_in, void *lin, void *lam_out, double w) {
            AdjArgs<R> q;
            q.g = gargs<R>(h, state, nullptr);
            q.lam_in = static_cast<const R *>(lam_in);
            q.lin = static_cast<R *>(lin);
            q.lam_out = static_cast<R *>(lam_out);
            q.gs = gs;
            q.wobj = w;
            return q;
        };
        // seed at t1: recompute a^{t1} from its checkpoint
        {
            const int cc = t1 / K;
            if (hip_check(h, hipMemcpyAsync(work[0], ck + S * (size_t)cc, S, hipMemcpyDeviceToDevice, h->stream),
                          "checkpoint restore")) {
                cleanup();
                return EUP_EHIP;
            }
            int w = 0;
            for (int t = cc * K; t < t1; ++t) {
                GArgs<R> a = gargs<R>(h, work[w], work[w ^ 1]);
                hipLaunchKernelGGL((k_grey_step<R>), dim3(grid_sites<R>(h)), dim3(256), 0, h->stream, a);
                w ^= 1;
            }
            AdjArgs<R> q = adj(work[w], nullptr, nullptr, lam, wobj);
            hipLaunchKernelGGL((k_force_adj_seed<R>), dim3(grid_sites<R>(h)), dim3(256), 0, h->stream, q);
        }
        // Per segment: the collide half of its last step, then the stream half of step t fused
        // with the collide half of step t - 1 (both states are in the recomputed segment), then
        // the stream half of its first step.  Fused by default in FP64 (MI355X, 128 x 64 x 64:
        // 154 us against 102 + 71 us split); FP32 runs the halves on their own by default (fused
        // 97 us against 47 + 44 us: the fused kernel's 173 registers per lane allow two
        // wavefronts per SIMD, the split collide's three).  EUP_ADJ_FUSED=0 / 1 forces either
        // form; both give the same results bit for bit.
        bool fuse = sizeof(R) == 8;
        if (const char *e = getenv("EUP_ADJ_FUSED")) fuse = atoi(e) != 0;
        const dim3 grid(grid_sites<R>(h));
        for (int cc = (t1 - 1) / K; cc >= 0; --cc) {
            const int tb = cc * K, te = (cc + 1) * K < t1 ? (cc + 1) * K : t1;
            if (te <= tb) continue; // no step in this segment (t1 >= 1 keeps every segment non-empty)
            if (hip_check(h, hipMemcpyAsync(seg, ck + S * (size_t)cc, S, hipMemcpyDeviceToDevice, h->stream),
                          "checkpoint restore")) {
                cleanup();
                return EUP_EHIP;
            }
            for (int t = tb + 1; t < te; ++t) {
                GArgs<R> a = gargs<R>(h, seg + S * (size_t)(t - 1 - tb), seg + S * (size_t)(t - tb));
                hipLaunchKernelGGL((k_grey_step<R>), grid, dim3(256), 0, h->stream, a);
            }
            auto at = [&](int t) { return static_cast<const void *>(seg + S * (size_t)(t - tb)); };
            auto wt = [&](int t) { return t >= t0 ? wobj : 0.0; };
            int b = 0; // linb[b] holds lin^t
            hipLaunchKernelGGL((k_adj_collide<R>), grid, dim3(256), 0, h->stream,
                               adj(at(te - 1), lam, linb[b], nullptr, wt(te - 1)));
            for (int t = te - 1; t > tb; --t) {
                if (fuse) {
                    hipLaunchKernelGGL((k_adj_fused<R>), grid, dim3(256), 0, h->stream,
                                       adj(at(t), nullptr, linb[b], nullptr, wt(t)),
                                       adj(at(t - 1), nullptr, linb[b ^ 1], nullptr, wt(t - 1)));
                    b ^= 1;
                } else {
                    hipLaunchKernelGGL((k_adj_stream<R>), grid, dim3(256), 0, h->stream,
                                       adj(at(t), nullptr, linb[b], lam, wt(t)));
                    hipLaunchKernelGGL((k_adj_collide<R>), grid, dim3(256), 0, h->stream,
                                       adj(at(t - 1), lam, linb[b], nullptr, wt(t - 1)));
                }
            }
            hipLaunchKernelGGL((k_adj_stream<R>), grid, dim3(256), 0, h->stream,
                               adj(at(tb), nullptr, linb[b], lam, wt(tb)));
        }
        if (hip_check(h, hipGetLastError(), "adjoint sweep")) {
            cleanup();
            return EUP_EHIP;
        }
        int rc = 0;
        if (grad_host) {
            if (hip_check(h, hipMemcpyAsync(grad_host, gs, sizeof(double) * N, hipMemcpyDeviceToHost, h->stream), "copy grad"))
                rc = EUP_EHIP;
        } else if (gdtype == EUP_STORE_F64) {
            if (hip_check(h, hipMemcpyAsync(grad_dev, gs, sizeof(double) * N, hipMemcpyDeviceToDevice, h->stream), "copy grad"))
                rc = EUP_EHIP;
        } else {
            rc = 0;
            float *gd = static_cast<float *>(grad_dev);
            std::vector<double> tmp(N);
            if (hip_check(h, hipMemcpyAsync(tmp.data(), gs, sizeof(double) * N, hipMemcpyDeviceToHost, h->stream), "copy") ||
                hip_check(h, hipStreamSynchronize(h->stream), "sync"))
                rc = EUP_EHIP;
            std::vector<float> tf(N);
            for (int64_t x = 0; x < N; ++x) tf[x] = (float)tmp[x];
            if (!rc && hip_check(h, hipMemcpy(gd, tf.data(), sizeof(float) * N, hipMemcpyHostToDevice), "copy"))
                rc = EUP_EHIP;
        }
        if (!rc && hip_check(h, hipStreamSynchronize(h->stream), "sync")) rc = EUP_EHIP;
        if (rc) {
            cleanup();
            return rc;
        }
    }
    cleanup();
    return 0;
}

template <typename F> int rdispatch(Handle *h, F &&f)
{
    if (h->cfg.storage == EUP_STORE_F64) return f((double *)nullptr);
    return f((float *)nullptr);
}

template <typename R>
__global__ void k_grey_from(const void *src, int dtype, R *dst, int64_t N)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= N) return;
    double v = dtype == EUP_STORE_F64 ? static_cast<const double *>(src)[x] : (double)static_cast<const float *>(src)[x];
    dst[x] = (R)v;
}

} // namespace

int grey_step(Handle *h)
{
    return rdispatch(h, [&](auto *p) { return grey_step_t<std::remove_pointer_t<decltype(p)>>(h); });
}

int grey_set_field(Handle *h, const void *s, int dtype, bool on_device)
{
    const int64_t N = h->g.sites;
    const size_t rb = h->cfg.storage == EUP_STORE_F64 ? 8 : 4;
    if (!h->grey && hip_check(h, hipMalloc(&h->grey, rb * N), "hipMalloc grey")) return EUP_EHIP;
    const void *src = s;
    void *tmp = nullptr;
    if (!on_device) {
        if (hip_check(h, hipMalloc(&tmp, sizeof(double) * N), "hipMalloc") ||
            hip_check(h, hipMemcpy(tmp, s, sizeof(double) * N, hipMemcpyHostToDevice), "copy grey")) {
            hipFree(tmp);
            return EUP_EHIP;
        }
        src = tmp;
        dtype = EUP_STORE_F64;
    }
    int rc = rdispatch(h, [&](auto *p) {
        using R = std::remove_pointer_t<decltype(p)>;
        hipLaunchKernelGGL((k_grey_from<R>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, h->stream, src,
                           dtype, static_cast<R *>(h->grey), N);
        return hip_check(h, hipGetLastError(), "grey field") ? EUP_EHIP : 0;
    });
    if (!rc && hip_check(h, hipStreamSynchronize(h->stream), "sync")) rc = EUP_EHIP;
    hipFree(tmp);
    return rc;
}

int grey_force(Handle *h, double F[3])
{
    if (h->diag_len < 3 * kForceBlocks + 16) {
        hipFree(h->diag);
        h->diag = nullptr;
        h->diag_len = 0;
        if (hip_check(h, hipMalloc((void **)&h->diag, sizeof(double) * (3 * kForceBlocks + 16)), "hipMalloc"))
            return EUP_EHIP;
        h->diag_len = 3 * kForceBlocks + 16;
    }
    double *acc = h->diag + 3 * kForceBlocks;
    if (hip_check(h, hipMemsetAsync(acc, 0, sizeof(double) * 6, h->stream), "memset")) return EUP_EHIP;
    int rc = rdispatch(h, [&](auto *p) {
        return force_into<std::remove_pointer_t<decltype(p)>>(h, h->buf[h->cur], 1.0, acc, acc + 3);
    });
    if (rc) return rc;
    if (hip_check(h, hipMemcpyAsync(F, acc + 3, sizeof(double) * 3, hipMemcpyDeviceToHost, h->stream), "copy") ||
        hip_check(h, hipStreamSynchronize(h->stream), "sync"))
        return EUP_EHIP;
    return 0;
}

int grey_objective(Handle *h, int t0, int t1, int K, double norm, double *J, void *grad_dev, int gdtype,
                   double *grad_host)
{
    return rdispatch(h, [&](auto *p) {
        return objective_t<std::remove_pointer_t<decltype(p)>>(h, t0, t1, K, norm, J, grad_dev, gdtype,
                                                               grad_host);
    });
}

} // namespace eup
