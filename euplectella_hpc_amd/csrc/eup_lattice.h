// D3Q19 lattice arithmetic shared by the HIP kernels (host- and device-callable).
//
// Direction numbering and weights: CODE/SRC/bgk3d.F90:31-49, storage.F90:77-81.  The collision,
// inlet and outlet formulas are written in the reference's evaluation order (col.F90:127-229,
// bcond.bc.F90:138-164 and :300-314); compiled with -ffp-contract=off the FP64 instantiation is
// bit-identical to the reference's serial path.  collide_dev is the same collision rearranged
// for single precision (explicit fused multiply-adds) on deviation-form populations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define EUP_HD __host__ __device__ __forceinline__

#include "eup_udiv.h"

namespace eup {

// Direction table as compile-time constants, 0-based population index Q = 0..18 for the
// reference's directions 1..19; opp = index of -c_Q.
template <int Q> struct Dir;
#define EUP_DIR(Q, X, Y, Z, O)                                                                   \
    template <> struct Dir<Q> {                                                                  \
        static constexpr int cx = X, cy = Y, cz = Z, opp = O;                                    \
    };
EUP_DIR(0, 1, -1, 0, 11)
EUP_DIR(1, 1, 0, -1, 12)
EUP_DIR(2, 1, 1, 0, 9)
EUP_DIR(3, 1, 0, 1, 10)
EUP_DIR(4, 1, 0, 0, 13)
EUP_DIR(5, 0, 0, 1, 14)
EUP_DIR(6, 0, 1, 1, 15)
EUP_DIR(7, 0, 1, 0, 16)
EUP_DIR(8, 0, 1, -1, 17)
EUP_DIR(9, -1, -1, 0, 2)
EUP_DIR(10, -1, 0, -1, 3)
EUP_DIR(11, -1, 1, 0, 0)
EUP_DIR(12, -1, 0, 1, 1)
EUP_DIR(13, -1, 0, 0, 4)
EUP_DIR(14, 0, 0, -1, 5)
EUP_DIR(15, 0, -1, -1, 6)
EUP_DIR(16, 0, -1, 0, 7)
EUP_DIR(17, 0, -1, 1, 8)
EUP_DIR(18, 0, 0, 0, 18)
#undef EUP_DIR

template <typename R> struct W {
    static constexpr R p0 = R(1) / R(3);
    static constexpr R p1 = R(1) / R(18);
    static constexpr R p2 = R(1) / R(36);
};

template <typename R> EUP_HD constexpr R weight(int q)
{
    return q == 18 ? W<R>::p0 : ((q == 4 || q == 5 || q == 7 || q == 13 || q == 14 || q == 16)
                                     ? W<R>::p1
                                     : W<R>::p2);
}

// BGK collision of one site, x[19] pre-collision -> out[19] (may alias); col.F90:127-229
template <typename R>
EUP_HD void collide(const R *x, R *out, R omega, R fgrad)
{
    const R one = R(1), three = R(3), rf = R(3), qf = R(1.5);
    const R omega1 = one - omega;
    R rho = x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9] + x[10] + x[11]
            + x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18];
    R rhoinv = one / rho;
    R vx = (x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rhoinv;
    R vy = (x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rhoinv;
    R vz = (x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rhoinv;
    R vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
    R vsq = vx2 + vy2 + vz2;
    R vxpy = vx + vy;
    R qxpy = one + qf * (three * vxpy * vxpy - vsq);
    R vxmy = vx - vy;
    R qxmy = one + qf * (three * vxmy * vxmy - vsq);
    R vxpz = vx + vz;
    R qxpz = one + qf * (three * vxpz * vxpz - vsq);
    R vxmz = vx - vz;
    R qxmz = one + qf * (three * vxmz * vxmz - vsq);
    R vypz = vy + vz;
    R qypz = one + qf * (three * vypz * vypz - vsq);
    R vymz = vy - vz;
    R qymz = one + qf * (three * vymz * vymz - vsq);
    R qx = one + qf * (three * vx2 - vsq);
    R qy = one + qf * (three * vy2 - vsq);
    R qz = one + qf * (three * vz2 - vsq);
    vx = rf * vx;
    vy = rf * vy;
    vz = rf * vz;
    vxpy = rf * vxpy;
    vxmy = rf * vxmy;
    vxpz = rf * vxpz;
    vxmz = rf * vxmz;
    vypz = rf * vypz;
    vymz = rf * vymz;
    const R rp1 = rho * W<R>::p1;
    const R rp2 = rho * W<R>::p2;
    const R fx = fgrad * rho; // FORCING_X (col.F90:173-187)
    const R e0 = rp2 * (vxmy + qxmy);
    const R e1 = rp2 * (vxmz + qxmz);
    const R e2 = rp2 * (vxpy + qxpy);
    const R e3 = rp2 * (vxpz + qxpz);
    const R e4 = rp1 * (vx + qx);
    const R e5 = rp1 * (vz + qz);
    const R e6 = rp2 * (vypz + qypz);
    const R e7 = rp1 * (vy + qy);
    const R e8 = rp2 * (vymz + qymz);
    const R e9 = rp2 * (-vxpy + qxpy);
    const R e10 = rp2 * (-vxpz + qxpz);
    const R e11 = rp2 * (-vxmy + qxmy);
    const R e12 = rp2 * (-vxmz + qxmz);
    const R e13 = rp1 * (-vx + qx);
    const R e14 = rp1 * (-vz + qz);
    const R e15 = rp2 * (-vypz + qypz);
    const R e16 = rp1 * (-vy + qy);
    const R e17 = rp2 * (-vymz + qymz);
    const R e18 = rho * W<R>::p0 * (one - qf * vsq);
    const R z = R(0);
    // (omega*e + fx) -/+ fy -/+ fz with fy = fz = 0, exactly as the reference's expressions
    out[0] = omega1 * x[0] + (omega * e0 + fx - z);
    out[1] = omega1 * x[1] + (omega * e1 + fx - z);
    out[2] = omega1 * x[2] + (omega * e2 + fx + z);
    out[3] = omega1 * x[3] + (omega * e3 + fx + z);
    out[4] = omega1 * x[4] + (omega * e4 + fx);
    out[5] = omega1 * x[5] + (omega * e5 + z);
    out[6] = omega1 * x[6] + (omega * e6 + z + z);
    out[7] = omega1 * x[7] + (omega * e7 + z);
    out[8] = omega1 * x[8] + (omega * e8 + z - z);
    out[9] = omega1 * x[9] + (omega * e9 - fx - z);
    out[10] = omega1 * x[10] + (omega * e10 - fx - z);
    out[11] = omega1 * x[11] + (omega * e11 - fx + z);
    out[12] = omega1 * x[12] + (omega * e12 - fx + z);
    out[13] = omega1 * x[13] + (omega * e13 - fx);
    out[14] = omega1 * x[14] + (omega * e14 - z);
    out[15] = omega1 * x[15] + (omega * e15 - z - z);
    out[16] = omega1 * x[16] + (omega * e16 - z);
    out[17] = omega1 * x[17] + (omega * e17 - z + z);
    out[18] = omega1 * x[18] + (omega * e18);
}

// Constants of the single-precision collision on deviations (collide_dev).  The populations are
// held as d_q = f_q - r_q with a constant r_q per population (r_q = 0 for full-value storage).
struct BgkF {
    float r[19];   // r_q
    float wr[19];  // omega * r_q
    float bb[19];  // r_opp(q) - r_q: added when population q bounces back from the site's own opp(q)
    float rsum;    // sum_q r_q
    float jr[3];   // sum_q c_q r_q
    float omega1;  // 1 - omega
    float ow1, ow2;      // omega * p1, omega * p2
    float fgrad;
};

// The BGK collision of col.F90:127-229 in value, written for single precision with fewer
// operations, applied to deviations d[19] in place (the rest population from the mass balance):
//   rho = sum d + sum r, j = sum c d + sum c r, u = j / rho (hardware reciprocal, 1 ulp),
//   omega e_q = W (s + 3 c_q.u) with W = omega w_q rho and s = 1 - 1.5 u^2 + 4.5 (c_q.u)^2, which
//   is shared by q and its opposite (whose 3 c.u has the other sign),
//   d'_q = (1 - omega) d_q + (omega e_q - omega r_q) +/- fgrad rho (c_x = +/-1, FORCING_X).
// The moments are summed by planes of c_x (the x-momentum is the difference of two of them).
// The FP64 path keeps the reference's own evaluation order (collide); this form is for the
// FP32-compute storages, whose results are checked against FP64 with tolerances.
template <bool DEV>
__device__ __forceinline__ void collide_dev(float *d, const BgkF &c)
{
    const float sp = ((d[0] + d[1]) + (d[2] + d[3])) + d[4];     // c_x = +1
    const float sm = ((d[9] + d[10]) + (d[11] + d[12])) + d[13]; // c_x = -1
    const float s0 = (((d[5] + d[6]) + (d[7] + d[8])) + ((d[14] + d[15]) + (d[16] + d[17]))) + d[18];
    float rho = (sp + sm) + s0;
    float jx = sp - sm;
    float jy = (((d[2] + d[6]) + (d[7] + d[8])) + d[11]) - (((d[0] + d[9]) + (d[15] + d[16])) + d[17]);
    float jz = (((d[3] + d[5]) + (d[6] + d[12])) + d[17]) - (((d[1] + d[8]) + (d[10] + d[14])) + d[15]);
    if constexpr (DEV) {
        rho += c.rsum;
        jx += c.jr[0];
        jy += c.jr[1];
        jz += c.jr[2];
    }
    const float rinv = __builtin_amdgcn_rcpf(rho);
    const float ux = jx * rinv, uy = jy * rinv, uz = jz * rinv;
    const float base = __builtin_fmaf(-1.5f, __builtin_fmaf(ux, ux, __builtin_fmaf(uy, uy, uz * uz)), 1.0f);
    const float x3 = 3.0f * ux, y3 = 3.0f * uy, z3 = 3.0f * uz;
    const float xh = 1.5f * ux, yh = 1.5f * uy, zh = 1.5f * uz;
    const float W1 = c.ow1 * rho, W2 = c.ow2 * rho;
    const float fx = c.fgrad * rho;
    const float o1 = c.omega1;
    // pair (q, qb = opp(q)) with t3 = 3 c_q.u, th = 1.5 c_q.u; F = forcing of q (-F for qb)
    auto pair = [&](int q, int qb, float W, float t3, float th, float F, bool forced) {
        const float s = __builtin_fmaf(t3, th, base);
        float A, Ab;
        if constexpr (DEV) {
            A = __builtin_fmaf(W, s, -c.wr[q]);
            Ab = __builtin_fmaf(W, s, -c.wr[qb]);
        } else {
            A = W * s;
            Ab = A;
        }
        const float B = forced ? __builtin_fmaf(W, t3, F) : W * t3;
        d[q] = __builtin_fmaf(o1, d[q], A + B);
        d[qb] = __builtin_fmaf(o1, d[qb], Ab - B);
    };
    pair(0, 11, W2, x3 - y3, xh - yh, fx, true);  // (1,-1, 0)
    pair(1, 12, W2, x3 - z3, xh - zh, fx, true);  // (1, 0,-1)
    pair(2, 9, W2, x3 + y3, xh + yh, fx, true);   // (1, 1, 0)
    pair(3, 10, W2, x3 + z3, xh + zh, fx, true);  // (1, 0, 1)
    pair(4, 13, W1, x3, xh, fx, true);            // (1, 0, 0)
    pair(5, 14, W1, z3, zh, 0.0f, false);         // (0, 0, 1)
    pair(6, 15, W2, y3 + z3, yh + zh, 0.0f, false); // (0, 1, 1)
    pair(7, 16, W1, y3, yh, 0.0f, false);         // (0, 1, 0)
    pair(8, 17, W2, y3 - z3, yh - zh, 0.0f, false); // (0, 1,-1)
    // The rest population from the mass balance: sum_q d'_q = sum_q d_q is BGK's mass
    // conservation, so d'_18 = sum_q d_q - sum_{q<18} d'_q is the collision's value in exact
    // arithmetic.  Written as (1 - omega) d_18 + omega w_0 rho s_0 instead, the single-precision
    // constants (1 - omega, omega w_q, their deviation sums) do not cancel exactly and every step
    // would scale the mass by 1 + O(1e-8): over the 2000-step start-up that is a density drift of
    // 1e-4 (measured: an FP32 free stream moves 2.4e-6 from equilibrium in 200 steps, 5.5e-7
    // this way; FP32 drag and probe errors against FP64 at tau 0.505 drop from 2.1e-4 / 3.1e-5
    // to 1.2e-4 / 4.0e-6).
    const float dsum = (sp + sm) + s0;
    const float t18 = ((((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + d[7]))) +
                       (((d[8] + d[9]) + (d[10] + d[11])) + ((d[12] + d[13]) + (d[14] + d[15])))) +
                      (d[16] + d[17]);
    d[18] = dsum - t18;
}

// outlet ghost values o[0..4] = q10..q14 from the momentum (not divided by rho) of x[19];
// bcond.bc.F90:138-164
template <typename R>
EUP_HD void outlet_eq(const R *x, R *o)
{
    const R rf = R(3), qf = R(1.5), three = R(3), one = R(1);
    R xj = +(x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]);
    R yj = +(x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]);
    R zj = +(x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]);
    R cvsq = xj * xj + yj * yj + zj * zj;
    R c10 = rf * (-xj - yj) + qf * (three * (xj + yj) * (xj + yj) - cvsq);
    R c11 = rf * (-xj - zj) + qf * (three * (xj + zj) * (xj + zj) - cvsq);
    R c12 = rf * (-xj + yj) + qf * (three * (xj - yj) * (xj - yj) - cvsq);
    R c13 = rf * (-xj + zj) + qf * (three * (xj - zj) * (xj - zj) - cvsq);
    R c14 = rf * (-xj) + qf * (three * (xj) * (xj)-cvsq);
    o[0] = W<R>::p2 * (one + c10);
    o[1] = W<R>::p2 * (one + c11);
    o[2] = W<R>::p2 * (one + c12);
    o[3] = W<R>::p2 * (one + c13);
    o[4] = W<R>::p1 * (one + c14);
}

// inlet ghost values o[0..4] = q1..q5 at u = (u_in, 0, 0); bcond.bc.F90:300-314
template <typename R>
EUP_HD void inlet_eq(R u_in, R *o)
{
    const R rf = R(3), qf = R(1.5), three = R(3), one = R(1);
    R xj = u_in, yj = R(0), zj = R(0);
    R cvsq = xj * xj + yj * yj + zj * zj;
    R c1 = rf * (xj - yj) + qf * (three * (xj - yj) * (xj - yj) - cvsq);
    R c2 = rf * (xj - zj) + qf * (three * (xj - zj) * (xj - zj) - cvsq);
    R c3 = rf * (xj + yj) + qf * (three * (xj + yj) * (xj + yj) - cvsq);
    R c4 = rf * (xj + zj) + qf * (three * (xj + zj) * (xj + zj) - cvsq);
    R c5 = rf * (xj) + qf * (three * (xj) * (xj)-cvsq);
    o[0] = W<R>::p2 * (one + c1);
    o[1] = W<R>::p2 * (one + c2);
    o[2] = W<R>::p2 * (one + c3);
    o[3] = W<R>::p2 * (one + c4);
    o[4] = W<R>::p1 * (one + c5);
}

// ---------------------------------------------------------------------------------------------
// storage traits: how a population is held in HBM and converted to the collision type
// ---------------------------------------------------------------------------------------------
struct bf16_t {
    uint16_t bits;
};

template <typename S> struct Store;

// `ref` is the reference value of the population for deviation storage (w_q, or the free-stream
// equilibrium of that population); full-value storages ignore it.  `q` is the population (0..18):
// only the weight-scaled format (e3m12_t) uses it, as a compile-time constant at every call site.
// load64 / store64 convert between the stored value and a full value in double (state set / get,
// equilibrium start, diagnostics): the deviation storages form ref + d and f - ref in double, where
// both are exact for the values they produce, so that storing what load64 returned gives back the
// stored code (a restart through the state changes nothing) -- for every code of FP16 and e3m12,
// and for BF16 deviations above 2^-53 of the reference (below, ref + d rounds to ref).
template <> struct Store<double> {
    using compute = double;
    static constexpr bool deviation = false;
    static __device__ __forceinline__ double load64(double v, float, int) { return v; }
    static __device__ __forceinline__ double store64(double f, float, int) { return f; }
    static __device__ __forceinline__ double load(double v, float, int) { return v; }
    static __device__ __forceinline__ double store(double f, float, int) { return f; }
    static __device__ __forceinline__ double raw_load(double v, int) { return v; }
    static __device__ __forceinline__ double raw_store(double f, int) { return f; }
};

template <> struct Store<float> {
    using compute = float;
    static constexpr bool deviation = false;
    static __device__ __forceinline__ double load64(float v, float, int) { return (double)v; }
    static __device__ __forceinline__ float store64(double f, float, int) { return (float)f; }
    static __device__ __forceinline__ float load(float v, float, int) { return v; }
    static __device__ __forceinline__ float store(float f, float, int) { return f; }
    static __device__ __forceinline__ float raw_load(float v, int) { return v; }
    static __device__ __forceinline__ float raw_store(float f, int) { return f; }
};

// deviation f - ref in binary16; float -> half conversion rounds to nearest even
template <> struct Store<_Float16> {
    using compute = float;
    static constexpr bool deviation = true;
    static __device__ __forceinline__ float load(_Float16 v, float ref, int) { return ref + (float)v; }
    static __device__ __forceinline__ _Float16 store(float f, float ref, int) { return (_Float16)(f - ref); }
    // the stored deviation itself (kernels that compute on deviations)
    static __device__ __forceinline__ float raw_load(_Float16 v, int) { return (float)v; }
    static __device__ __forceinline__ _Float16 raw_store(float d, int) { return (_Float16)d; }
    static __device__ __forceinline__ double load64(_Float16 v, float ref, int) { return (double)ref + (double)(float)v; }
    static __device__ __forceinline__ _Float16 store64(double f, float ref, int) { return (_Float16)(float)(f - (double)ref); }
};

// deviation f - ref in bfloat16 with round-to-nearest-even
template <> struct Store<bf16_t> {
    using compute = float;
    static constexpr bool deviation = true;
    static __device__ __forceinline__ float to_float(bf16_t v) { return __uint_as_float((uint32_t)v.bits << 16); }
    // hardware conversion (v_cvt_pk_bf16_f32, round to nearest even in the default mode)
    static __device__ __forceinline__ bf16_t from_float(float f)
    {
        const __bf16 h = (__bf16)f;
        bf16_t b;
        b.bits = __builtin_bit_cast(uint16_t, h);
        return b;
    }
    static __device__ __forceinline__ float load(bf16_t v, float ref, int) { return ref + to_float(v); }
    static __device__ __forceinline__ bf16_t store(float f, float ref, int) { return from_float(f - ref); }
    static __device__ __forceinline__ float raw_load(bf16_t v, int) { return to_float(v); }
    static __device__ __forceinline__ bf16_t raw_store(float d, int) { return from_float(d); }
    static __device__ __forceinline__ double load64(bf16_t v, float ref, int) { return (double)ref + (double)to_float(v); }
    static __device__ __forceinline__ bf16_t store64(double f, float ref, int) { return from_float((float)(f - (double)ref)); }
};

// Deviation f_q - ref_q divided by w_q, as a 16-bit float with 1 sign, 3 exponent and 12 mantissa
// bits: binades [2^-5, 2^-4) .. [2, 4) (in units of w_q) and below them a fixed step of 2^-17 w_q.
// Deviations of BGK populations stay well inside |f_q - ref_q| < 2 w_q, where IEEE binary16 spends
// 5 exponent bits on 30 binades it never uses; this keeps 12 significant bits instead of 11 and
// an absolute step of w_q 2^-17 near zero.  The top code (all ones) is NaN: a deviation of 4 w_q or
// more, or a non-finite one, stores NaN rather than a saturated value, so a run that leaves the
// format's range shows up as non-finite.
//   encode: v = d 2^-121 / w_q puts the format's subnormal range onto FP32 denormals (kept by the
//           kernels: float denormal mode preserved), so the FP32 bits of |v| are the code scaled by
//           2^11: clamp, round to nearest with ties away from zero at bit 11, and shift -- built in
//           the upper half of a 32-bit word, where FP32's own sign bit already sits (4 VALU
//           operations; a tie needs the 11 dropped bits to be exactly 0x400, and rounding it away
//           from zero biases a magnitude by 2^-13 of its last place at most);
//   decode: the code shifted back into FP32 bits (sign-extending 16-bit load, shift, mask), times
//           w_q 2^121.
struct e3m12_t {
    uint16_t bits;
};

template <> struct Store<e3m12_t> {
    using compute = float;
    static constexpr bool deviation = true;
    // largest FP32 bit pattern of |v| whose rounded code is below the NaN code's rounding range
    static constexpr uint32_t kTop = 0x03FFFBFFu;
    // the FP32 bits of the NaN code's value, 0x7FFF << 11
    static constexpr uint32_t kNanBits = 0x03FFF800u;
    // the code in the upper 16 bits of the result (the lower half is not part of it)
    static __device__ __forceinline__ uint32_t code_hi(float d, int q)
    {
        // |d| through the multiply's input modifier; a NaN or infinity has bits above kTop
        uint32_t a = __float_as_uint(__builtin_fabsf(d) * (0x1p-121f / weight<float>(q)));
        a = a < kTop ? a : kTop;
        uint32_t u = (a << 5) + 0x8000u; // v_lshl_add_u32
        // hide the known zero sign bit of u, which would let the compiler move the rounding add
        // behind the sign merge (3 operations instead of v_and_or_b32)
        asm volatile("" : "+v"(u));
        return (__float_as_uint(d) & 0x80000000u) | u;
    }
    // FP32 bits of the stored value before the scale w_q 2^121 (sign and magnitude in place)
    static __device__ __forceinline__ uint32_t value_bits(e3m12_t x)
    {
        const uint32_t s = (uint32_t)(int32_t)(int16_t)x.bits;
        return (s << 11) & 0x83FFF800u;
    }
    static __device__ __forceinline__ e3m12_t raw_store(float d, int q)
    {
        e3m12_t r;
        r.bits = (uint16_t)(code_hi(d, q) >> 16);
        return r;
    }
    static __device__ __forceinline__ float raw_load(e3m12_t x, int q)
    {
        const uint32_t b = value_bits(x);
        const float v = __uint_as_float(b) * (weight<float>(q) * 0x1p121f);
        return (b & 0x7FFFFFFFu) == kNanBits ? __builtin_nanf("") : v;
    }
    // The step kernel's load: the NaN code is not tested per value; the largest magnitude of a
    // site's loads is kept in `top` and tested once (nan_if_top), so a NaN code among them makes
    // the site's update NaN -- also one it loads without using (the own populations of a mixed
    // tile's site, the row-end lanes' clamped pulls), which a NaN state does not mind.
    static __device__ __forceinline__ float step_load(e3m12_t x, int q, float &top)
    {
        const float h = __uint_as_float(value_bits(x));
        top = __builtin_fmaxf(top, __builtin_fabsf(h));
        return h * (weight<float>(q) * 0x1p121f);
    }
    // the stored value with the NaN code read as its magnitude: the outlet ghost plane, which the
    // step kernel and the host's refresh compute alike (a NaN code at x = l makes the site's own
    // update NaN through nan_if_top; the plane itself stays finite)
    static __device__ __forceinline__ float load_magnitude(e3m12_t x, int q)
    {
        return __uint_as_float(value_bits(x)) * (weight<float>(q) * 0x1p121f);
    }
    static __device__ __forceinline__ float nan_if_top(float top, float v)
    {
        return top >= __uint_as_float(kNanBits) ? __builtin_nanf("") : v;
    }
    static __device__ __forceinline__ float load(e3m12_t v, float ref, int q) { return ref + raw_load(v, q); }
    static __device__ __forceinline__ e3m12_t store(float f, float ref, int q) { return raw_store(f - ref, q); }
    static __device__ __forceinline__ double load64(e3m12_t v, float ref, int q) { return (double)ref + (double)raw_load(v, q); }
    static __device__ __forceinline__ e3m12_t store64(double f, float ref, int q) { return raw_store((float)(f - (double)ref), q); }
};

} // namespace eup

namespace eup {
// reference values of the 19 populations for deviation storage (kernel argument)
struct DevRef {
    float r[19];
};
} // namespace eup
