// device_common.hpp -- device helpers shared by the PIMD-B hot-path kernels
// (gfx950, wave64).  All arithmetic is FP64.
//
// Bit-level notes: the distance helpers disable FP contraction so that a
// min-imaged displacement, its squared norm and the norm are bit-identical to
// the reference's (src/common.cpp:41-43, include/common.h:269-275); the
// potentials follow the reference's operation order but let the compiler
// contract, which moves results by at most a few ulp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pimdb {

constexpr int WAVE = 64;

enum PotKind : int { POT_FREE = 0, POT_HARMONIC = 1, POT_DIPOLE = 2, POT_AZIZ = 3, POT_DOUBLE_WELL = 4, POT_COSINE = 5 };

// Aziz HFDHE2 constants, src/potentials/aziz.cpp:6-13 (atomic units).
struct Aziz {
    static constexpr double rm = 5.60738;
    static constexpr double A = 0.5448504e6;
    static constexpr double epsilon = 3.42016E-5;
    static constexpr double alpha = 13.353384;
    static constexpr double D = 1.241314;
    static constexpr double C6 = 1.3732412;
    static constexpr double C8 = 0.4253785;
    static constexpr double C10 = 0.1781;
};
constexpr double HARD_CORE_EPS = 1.0E-7;   // include/common.h:193

// dx - L floor(dx/L + 1/2), src/common.cpp:41-43.
__device__ __forceinline__ double min_image(double dx, double L) {
#pragma clang fp contract(off)
    return dx - L * floor(dx / L + 0.5);
}

// Displacement b - a along one axis, min-imaged when pbc.
__device__ __forceinline__ double disp(double a, double b, bool pbc, double L) {
#pragma clang fp contract(off)
    double d = b - a;
    return pbc ? min_image(d, L) : d;
}

// Same image as min_image(dx, L) without a division on the common path:
// floor(dx * (1/L) + 1/2) equals floor(dx / L + 1/2) unless the argument lies
// within a few ulp of an integer, which the guard sends to the exact path.
__device__ __forceinline__ double min_image_fast(double dx, double L, double invL) {
#pragma clang fp contract(off)
    const double q = dx * invL + 0.5;
    double fl = floor(q);
    const double fr = q - fl;
    if (fr < 1e-12 || fr > 1.0 - 1e-12) fl = floor(dx / L + 0.5);
    return dx - L * fl;
}
__device__ __forceinline__ double disp_fast(double a, double b, bool pbc, double L, double invL) {
#pragma clang fp contract(off)
    double d = b - a;
    return pbc ? min_image_fast(d, L, invL) : d;
}

// Squared norm, summed over axes in order without contraction.
template <int D>
__device__ __forceinline__ double norm2(const double (&d)[D]) {
#pragma clang fp contract(off)
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) s += d[a] * d[a];
    return s;
}

// Lane l receives what lane (l+1) mod 64 held (DPP wave_rol:1).  Call with
// the whole wave active.
__device__ __forceinline__ int lane_rotate(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x134, 0xF, 0xF, false); }
__device__ __forceinline__ double lane_rotate(double v) {
    const int lo = lane_rotate(__double2loint(v));
    const int hi = lane_rotate(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Squared norm with fused multiply-adds (the pair loop and the exchange's
// cycle energies on the step path).
template <int D>
__device__ __forceinline__ double norm2_fma(const double (&d)[D]) {
    double s = d[0] * d[0];
#pragma unroll
    for (int a = 1; a < D; ++a) s = fma(d[a], d[a], s);
    return s;
}

// A copy of a wave-uniform double in vector registers, for an operand that
// meets another scalar one in the same instruction (one scalar operand per
// VALU instruction): the compiler would otherwise copy it to a VGPR at every
// use, and under scalar register pressure reload it from a spill first.
__device__ __forceinline__ double vgpr_copy(double s) {
    double v;
    asm volatile("v_mov_b64 %0, %1" : "=v"(v) : "s"(s));
    return v;
}

// a * b + c with c in a scalar register.  The compiler turns p = fma(p, r, c)
// with a loop-invariant c into a copy of c plus v_fmac; a three-operand
// v_fma_f64 with the constant in an SGPR needs neither.
__device__ __forceinline__ double fma_sconst(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}

// e^x for x <= 709 (the pair potentials evaluate it at x <= 0), as 2^k e^r
// with k = rint(x / ln 2) and r = x - k ln 2 reduced with the two-part ln 2 of
// Cody and Waite (1980), |r| <= ln(2)/2.  e^r is the degree-11 polynomial of
// best relative approximation on that interval (Remez exchange; coefficients
// from gen_exp_minimax.py beside this file): its approximation error is
// 2.3e-17 relative with the coefficients rounded to double, so the result is
// within about 1 ulp of the correctly rounded value.  Arguments below -746
// give 0 (e^-746 is below half the smallest subnormal).
__device__ __forceinline__ double exp_minimax(double x) {
    const double LOG2E = 1.4426950408889634;
    const double LN2_HI = 6.93147180369123816490e-01;
    const double LN2_LO = 1.90821492927058770002e-10;
    x = fmax(x, -746.0);
    const double k = rint(x * LOG2E);
    double r = fma(-k, LN2_HI, x);
    r = fma(-k, LN2_LO, r);
    double p = fma_sconst(2.4994304015873615e-08, r, 2.7632293297646363e-07);
    p = fma_sconst(p, r, 2.7557622533560847e-06);
    p = fma_sconst(p, r, 2.4801486521375564e-05);
    p = fma_sconst(p, r, 1.9841269432676259e-04);
    p = fma_sconst(p, r, 1.388888895122404e-03);
    p = fma_sconst(p, r, 8.333333333559272e-03);
    p = fma_sconst(p, r, 4.166666666649277e-02);
    p = fma_sconst(p, r, 1.666666666666617e-01);
    p = fma_sconst(p, r, 5.000000000000018e-01);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)k);
}

// e^x for x <= 0 to within 5e-14 relative, for the pair force's
// exponentials (whose results enter forces compared at 1e-10 relative):
// Cody-Waite reduction with ln 2 in one double (its 2.3e-17 relative error,
// times |k| <= 1076, bounds the reduction's share by 2.5e-14; 1.7e-15 for
// x >= -51, the Aziz repulsion's range inside L/2) and a degree-9 minimax
// polynomial on |r| <= ln(2)/2 (pimd-b_amd/csrc/gen_exp_minimax.py 9: 1.36e-14
// relative with the coefficients rounded to double).  Three instructions
// fewer than exp_minimax.  Arguments below -746 give 0.
__device__ __forceinline__ double exp_pair(double x) {
    const double LOG2E = 1.4426950408889634;
    const double LN2 = 6.931471805599453e-01;
    x = fmax(x, -746.0);
    const double k = rint(x * LOG2E);
    const double r = fma(-k, LN2, x);
    double p = fma_sconst(2.7482187570680464e-06, r, 2.488401002638418e-05);
    p = fma_sconst(p, r, 1.9841596395525616e-04);
    p = fma_sconst(p, r, 1.3888802828792633e-03);
    p = fma_sconst(p, r, 8.333333002823904e-03);
    p = fma_sconst(p, r, 4.166666703246359e-02);
    p = fma_sconst(p, r, 1.6666666667826951e-01);
    p = fma_sconst(p, r, 4.999999999945797e-01);
    p = fma_sconst(p, r, 9.999999999998928e-01);
    p = fma_sconst(p, r, 1.0000000000000127);
    return ldexp(p, (int)k);
}

// 2^x for finite x, to within 1.4e-14 relative (plus a few ulp from the
// folded constants): k = rint(x), f = x - k exactly (|f| <= 1/2), and 2^f
// from exp_pair's degree-9 polynomial taken at r = f ln 2 (coefficient j times
// (ln 2)^j).  For the batched exchange, whose exponents are kept in base 2.
// k below -1074 gives 0; x must not be infinite or NaN (-inf gives NaN).
__device__ __forceinline__ double exp2_pair(double x) {
    constexpr double l1 = 6.931471805599453e-01, l2 = l1 * l1, l3 = l2 * l1, l4 = l3 * l1, l5 = l4 * l1,
                     l6 = l5 * l1, l7 = l6 * l1, l8 = l7 * l1, l9 = l8 * l1;
    const double k = rint(x);
    const double f = x - k;
    double p = fma_sconst(2.7482187570680464e-06 * l9, f, 2.488401002638418e-05 * l8);
    p = fma_sconst(p, f, 1.9841596395525616e-04 * l7);
    p = fma_sconst(p, f, 1.3888802828792633e-03 * l6);
    p = fma_sconst(p, f, 8.333333002823904e-03 * l5);
    p = fma_sconst(p, f, 4.166666703246359e-02 * l4);
    p = fma_sconst(p, f, 1.6666666667826951e-01 * l3);
    p = fma_sconst(p, f, 4.999999999945797e-01 * l2);
    p = fma_sconst(p, f, 9.999999999998928e-01 * l1);
    p = fma_sconst(p, f, 1.0000000000000127);
    return ldexp(p, (int)k);
}

// Scalar factor g(r) of the interaction gradient, gradV(d) = g * d
// (harmonic.cpp:21-23, dipole.cpp:24-45, aziz.cpp:56-106).  `norm` is |d|.
template <int KIND>
__device__ __forceinline__ double pair_grad_factor(double norm, double kpair, double strength) {
    if constexpr (KIND == POT_HARMONIC) {
        return kpair;
    } else if constexpr (KIND == POT_DIPOLE) {
        return -3.0 * strength / (norm * norm * norm * norm * norm);
    } else if constexpr (KIND == POT_AZIZ) {
        const double r = norm / Aziz::rm;
        const double T1 = -Aziz::A * Aziz::alpha * exp(-Aziz::alpha * r);
        double g;
        if (r > HARD_CORE_EPS && r < 0.01) {
            g = T1 * (Aziz::epsilon / Aziz::rm);
        } else {
            const double ix = 1.0 / r;
            const double ix2 = ix * ix;
            const double ix6 = ix2 * ix2 * ix2;
            const double ix7 = ix6 * ix;
            const double ix8 = ix6 * ix2;
            const double ix9 = ix8 * ix;
            const double ix10 = ix8 * ix2;
            const double ix11 = ix10 * ix;
            double F = 1.0, dF = 0.0;
            if (r < Aziz::D) {
                // aziz.h:25-43: F uses D/x, dF uses D*(1/x); one exponential
                // serves both (the two arguments differ by at most an ulp).
                const double t = Aziz::D * ix - 1.0;
                const double ex = exp(-t * t);
                F = ex;
                dF = 2.0 * Aziz::D * ix * ix * t * ex;
            }
            const double T2 = (6.0 * Aziz::C6 * ix7 + 8.0 * Aziz::C8 * ix9 + 10.0 * Aziz::C10 * ix11) * F;
            const double T3 = -(Aziz::C6 * ix6 + Aziz::C8 * ix8 + Aziz::C10 * ix10) * dF;
            g = (Aziz::epsilon / Aziz::rm) * (T1 + T2 + T3);
        }
        return g / norm;
    } else {
        return 0.0;
    }
}

// g(r) as above from |d| and 1/|d| with products in place of the
// reference's divisions (norm/rm, 1/r, g/norm), the dispersion terms
// factored as polynomials in ix^2, and exp_pair in place of the library exp:
// within a few 1e-14 relative.  For Aziz the factor epsilon/rm is folded into the
// constants, the repulsive exponent is taken from |d| directly
// (-alpha/rm |d|), and the range tests on r = |d|/rm are made on |d|
// against rm-scaled bounds (at x = D the damping factor joins F = 1 with a
// continuous first derivative).
template <int KIND>
__device__ __forceinline__ double pair_grad_factor_fast(double norm, double inv_norm, double kpair, double strength) {
    if constexpr (KIND == POT_HARMONIC) {
        return kpair;
    } else if constexpr (KIND == POT_DIPOLE) {
        const double i2 = inv_norm * inv_norm;
        return -3.0 * strength * (i2 * i2 * inv_norm);
    } else if constexpr (KIND == POT_AZIZ) {
        constexpr double ER = Aziz::epsilon / Aziz::rm;
        // (epsilon/rm) T1, T1 = -A alpha exp(-alpha r)
        const double T1 = (-Aziz::A * Aziz::alpha * ER) * exp_pair((-Aziz::alpha / Aziz::rm) * norm);
        // T2 + T3 = ix^6 [ix (6C6 + 8C8 ix^2 + 10C10 ix^4) F
        //                 - (C6 + C8 ix^2 + C10 ix^4) dF], dF = 2 D ix^2 t F
        // (times epsilon/rm, folded into the coefficients)
        const double ix = Aziz::rm * inv_norm;
        const double ix2 = ix * ix;
        const double ix6 = ix2 * ix2 * ix2;
        const double pa =
            fma_sconst(fma_sconst(ER * 10.0 * Aziz::C10, ix2, ER * 8.0 * Aziz::C8), ix2, ER * 6.0 * Aziz::C6);
        double attr = ix * pa;
        if (norm < Aziz::D * Aziz::rm) {
            // the reference's hard-core range (r in (eps, 0.01), inside the
            // damping range) keeps the repulsive term alone: attr = 0 gives
            // g = T1 exactly (ix6 is finite there)
            if (norm > HARD_CORE_EPS * Aziz::rm && norm < 0.01 * Aziz::rm) {
                attr = 0.0;
            } else {
                const double pb = fma_sconst(fma_sconst(ER * Aziz::C10, ix2, ER * Aziz::C8), ix2, ER * Aziz::C6);
                const double t = Aziz::D * ix - 1.0;
                const double ex = exp_pair(-t * t);
                attr = (attr - (2.0 * Aziz::D) * ix2 * t * pb) * ex;
            }
        }
        const double g = fma(ix6, attr, T1);
        return g * inv_norm;
    } else {
        return 0.0;
    }
}

// External potential of one particle (Simulation::initializePotential,
// simulation.cpp:612-645):
//   harmonic     a = m w^2                          (harmonic.cpp:3-27)
//   double well  a = 4 m s, b = loc^2, c = m s       (double_well.cpp:6-40)
//   cosine       a = -A k, b = k = 2 pi / L, c = phase, d = A   (cosine.cpp:4-36)
// The double well's energy sums (x_a^2 - loc^2)^2 over axes while its
// gradient is the isotropic 4 m s (|x|^2 - loc^2) x, as in the reference.
struct ExtPot {
    int kind;
    double a, b, c, d;
};

// Force -gradV_ext at one particle, computed as gradV * (-1) like
// Simulation::updatePhysicalForces (simulation.cpp:430).
template <int D>
__device__ __forceinline__ void ext_force(const ExtPot& e, const double (&x)[D], double (&F)[D]) {
    if (e.kind == POT_HARMONIC) {
#pragma unroll
        for (int a = 0; a < D; ++a) F[a] = (x[a] * e.a) * (-1.0);
    } else if (e.kind == POT_DOUBLE_WELL) {
        double r2 = 0.0;
#pragma unroll
        for (int a = 0; a < D; ++a) r2 += x[a] * x[a];
        const double pref = e.a * (r2 - e.b);
#pragma unroll
        for (int a = 0; a < D; ++a) F[a] = (pref * x[a]) * (-1.0);
    } else if (e.kind == POT_COSINE) {
#pragma unroll
        for (int a = 0; a < D; ++a) F[a] = (e.a * sin(e.b * x[a] + e.c)) * (-1.0);
    } else {
#pragma unroll
        for (int a = 0; a < D; ++a) F[a] = -0.0;
    }
}

// The particle's share of V_ext of its bead.
template <int D>
__device__ __forceinline__ double ext_energy(const ExtPot& e, const double (&x)[D]) {
    double s = 0.0;
    if (e.kind == POT_HARMONIC) {
#pragma unroll
        for (int a = 0; a < D; ++a) s += x[a] * x[a];
        return 0.5 * e.a * s;
    } else if (e.kind == POT_DOUBLE_WELL) {
#pragma unroll
        for (int a = 0; a < D; ++a) s += (x[a] * x[a] - e.b) * (x[a] * x[a] - e.b);
        return s * e.c;
    } else if (e.kind == POT_COSINE) {
#pragma unroll
        for (int a = 0; a < D; ++a) s += cos(e.b * x[a] + e.c);
        return s * e.d;
    }
    return 0.0;
}

// Pair potential energy V(d) (harmonic.cpp:9-19, dipole.cpp:5-22, aziz.cpp:18-54).
template <int KIND>
__device__ __forceinline__ double pair_energy(double norm, double r2, double kpair, double strength) {
    if constexpr (KIND == POT_HARMONIC) {
        return 0.5 * kpair * r2;
    } else if constexpr (KIND == POT_DIPOLE) {
        return strength / (norm * norm * norm);
    } else if constexpr (KIND == POT_AZIZ) {
        const double r = norm / Aziz::rm;
        const double Urep = Aziz::A * exp(-Aziz::alpha * r);
        if (r > HARD_CORE_EPS && r < 0.01) return Aziz::epsilon * Urep;
        const double ix2 = 1.0 / (r * r);
        const double ix6 = ix2 * ix2 * ix2;
        const double ix8 = ix6 * ix2;
        const double ix10 = ix8 * ix2;
        const double F = (r < Aziz::D) ? exp(-(Aziz::D / r - 1.0) * (Aziz::D / r - 1.0)) : 1.0;
        const double Uatt = -(Aziz::C6 * ix6 + Aziz::C8 * ix8 + Aziz::C10 * ix10) * F;
        return Aziz::epsilon * (Urep + Uatt);
    } else {
        return 0.0;
    }
}

// ---- wave-level reductions (DPP/permute through the compiler's shuffles) ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, WAVE));
    return v;
}
// Broadcast a double held by lane `src` to the whole wave.
__device__ __forceinline__ double wave_bcast(double v, int src) {
    return __shfl(v, src, WAVE);
}
// Same for a wave-uniform lane index: two v_readlane into scalar registers.
__device__ __forceinline__ double wave_bcast_uniform(double v, int src) {
    const uint64_t bits = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)bits, src);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(bits >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ----
struct U4 { uint32_t x, y, z, w; };
__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c.x;
        const uint64_t p1 = (uint64_t)M1 * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// ln u for u in (0, 1] (Box-Muller's radius).  u = m 2^e with m in
// [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m - 1)/(m + 1), |s| <= 0.1716,
// by its Taylor series sum_k s^(2k+1)/(2k+1) through s^21 (the rest is below
// 1e-18 relative); e ln 2 with ln 2 in the two parts of exp_minimax.
// FAST: s from the hardware reciprocal estimate of m + 1 (about 2^-23
// relative), one Newton step and one correction of the quotient, within an
// ulp or so of the correctly rounded division.
template <bool FAST = false>
__device__ __forceinline__ double log_unit(double u) {
    const double LN2_HI = 6.93147180369123816490e-01;
    const double LN2_LO = 1.90821492927058770002e-10;
    int e = __builtin_amdgcn_frexp_exp(u);
    double m = __builtin_amdgcn_frexp_mant(u);
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double f = m - 1.0;
    double s;
    if constexpr (FAST) {
        const double d = 2.0 + f;
        double rc = __builtin_amdgcn_rcp(d);
        rc = fma(rc, fma(-d, rc, 1.0), rc);
        s = f * rc;
        s = fma(fma(-d, s, f), rc, s);
    } else {
        s = f / (2.0 + f);
    }
    const double z = s * s;
    double p = 1.0 / 21.0;
    p = fma_sconst(p, z, 1.0 / 19.0);
    p = fma_sconst(p, z, 1.0 / 17.0);
    p = fma_sconst(p, z, 1.0 / 15.0);
    p = fma_sconst(p, z, 1.0 / 13.0);
    p = fma_sconst(p, z, 1.0 / 11.0);
    p = fma_sconst(p, z, 1.0 / 9.0);
    p = fma_sconst(p, z, 1.0 / 7.0);
    p = fma_sconst(p, z, 1.0 / 5.0);
    p = fma_sconst(p, z, 1.0 / 3.0);
    p = fma_sconst(p, z, 1.0);
    const double ed = (double)e;
    return fma(ed, LN2_HI, fma(ed, LN2_LO, (2.0 * s) * p));
}

// sin and cos of 2 pi u for u in [0, 1): 4u = q + r with q = rint(4u) and
// |r| <= 1/2 (both exact), x = r pi/2, and the Taylor series of sin x through
// x^17 and of cos x through x^18 (|x| <= pi/4: the rest is below 1e-17),
// rotated by the quadrant q.
__device__ __forceinline__ void sincos_2pi(double u, double& sn, double& cs) {
    const double t = 4.0 * u;
    const double q = rint(t);
    const double x = (t - q) * 1.57079632679489661923;
    const double z = x * x;
    double ps = -1.0 / 355687428096000.0;          // -1/17!
    ps = fma_sconst(ps, z, 1.0 / 1307674368000.0);        //  1/15!
    ps = fma_sconst(ps, z, -1.0 / 6227020800.0);          // -1/13!
    ps = fma_sconst(ps, z, 1.0 / 39916800.0);             //  1/11!
    ps = fma_sconst(ps, z, -1.0 / 362880.0);              // -1/9!
    ps = fma_sconst(ps, z, 1.0 / 5040.0);                 //  1/7!
    ps = fma_sconst(ps, z, -1.0 / 120.0);                 // -1/5!
    ps = fma_sconst(ps, z, 1.0 / 6.0);                    //  1/3!, sign below
    const double sx = fma(-x * z, ps, x);
    double pc = 1.0 / 6402373705728000.0;          //  1/18!
    pc = fma_sconst(pc, z, -1.0 / 20922789888000.0);      // -1/16!
    pc = fma_sconst(pc, z, 1.0 / 87178291200.0);          //  1/14!
    pc = fma_sconst(pc, z, -1.0 / 479001600.0);           // -1/12!
    pc = fma_sconst(pc, z, 1.0 / 3628800.0);              //  1/10!
    pc = fma_sconst(pc, z, -1.0 / 40320.0);               // -1/8!
    pc = fma_sconst(pc, z, 1.0 / 720.0);                  //  1/6!
    pc = fma_sconst(pc, z, -1.0 / 24.0);                  // -1/4!
    pc = fma_sconst(pc, z, 0.5);                          //  1/2!, sign below
    const double cx = fma(-z, pc, 1.0);
    const int qi = (int)q & 3;
    const double a = (qi & 1) ? cx : sx, b = (qi & 1) ? sx : cx;
    sn = (qi & 2) ? -a : a;
    cs = ((qi + 1) & 2) ? -b : b;
}

// Two standard normal deviates from one Philox block (Box-Muller on 53-bit
// uniforms in (0,1]).  Counter = (element, step, ostep|stream, replica).
// FAST (the batched form): the same draws with log_unit<true> and the square
// root as l / sqrt(l) from the hardware estimate and one third-order
// correction, each within an ulp or so of the correctly rounded operation,
// so the deviates agree with the other form's to a few units in the last
// place.
template <bool FAST = false>
__device__ __forceinline__ void gauss2(uint32_t elem, uint32_t step, uint32_t ostep,
                                       uint32_t replica, uint32_t seed, double& g0, double& g1) {
    const U4 r = philox4x32_10(U4{elem, step, ostep, replica}, seed, 0x5EEDB05Eu);
    const double TWO53 = 9007199254740992.0;
    const uint64_t a = ((uint64_t)r.x << 21) ^ (uint64_t)(r.y >> 11);
    const uint64_t b = ((uint64_t)r.z << 21) ^ (uint64_t)(r.w >> 11);
    const double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) / TWO53;   // (0, 1]
    const double u2 = (double)(b & ((1ull << 53) - 1)) / TWO53;           // [0, 1)
    double rad;
    if constexpr (FAST) {
        const double l = -2.0 * log_unit<true>(u1);   // >= 0; 0 at u1 = 1
        const double y0 = __builtin_amdgcn_rsq(l);
        const double e = fma(-l, y0 * y0, 1.0);
        const double y = fma(y0 * e, fma(e, 0.375, 0.5), y0);
        rad = l > 0.0 ? l * y : 0.0;
    } else {
        rad = sqrt(-2.0 * log_unit(u1));
    }
    double s, c;
    sincos_2pi(u2, s, c);
    g0 = rad * c;
    g1 = rad * s;
}

}  // namespace pimdb
