// sdm_device.h -- device-side helpers of the SDM HIP path (gfx950, wave64).
//
// Physics conventions (OpenMM NonbondedForce, CutoffPeriodic, reaction field; the
// force the reference evaluates at LangevinIntegratorSDM.cpp:160,168):
//   E_LJ = 4 eps_ij [(s_ij/r)^12 - (s_ij/r)^6], s_ij = (s_i+s_j)/2, eps_ij = sqrt(e_i e_j)
//   E_C  = k_e q_i q_j (1/r + k_rf r^2 - c_rf), k_rf = (e_rf-1)/((2e_rf+1) rc^3),
//          c_rf = 3 e_rf/((2e_rf+1) rc), k_e = 138.935456 kJ nm/(mol e^2)
//   r < rc only, excluded pairs skipped, no switching, no dispersion correction.
// Forces are accumulated as int64 fixed point with 2^32 units per kJ/mol/nm
// (the layout of platforms/cuda/src/kernels/langevin.cu:7-21).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SDM_ONE_4PI_EPS0 138.935456
#define SDM_FIX_SCALE 4294967296.0       // 2^32
#define SDM_FIX_LIMIT 2147483647.0       // |F| must stay below 2^31 kJ/mol/nm
#define SDM_BOLTZ (1.380658e-23 * 6.0221367e23 / 1000.0)
#define SDM_MAX_EXCL 32                  // exclusion partners per atom
#define SDM_MAX_CLASSES 8                // distinct nonzero displacement vectors

namespace sdm {

// Per-replica parameter block on the device (sdm_params + the integrator clock).
struct DevState {
    int32_t bias_method, softcore_method;
    double lambda, gamma, wbcoeff, w0coeff, lambda1, lambda2, alpha, u0;
    double umax, acore, ubcore;
    int32_t nonequilibrium;
    double noneq_tmax, work;
    double m_l1, m_l2, m_u0, m_w0, b_l1, b_l2, b_u0, b_w0;
    double temperature, friction, step_size;
    uint64_t seed;
    double m_alpha;
    // clock (not part of sdm_params)
    double time;
    int64_t step;
    // u' records taken since the history was configured (sdm_set_u_history)
    int64_t hist_n;
};

// Per-replica step record written by the scalar stage.
struct DevRecord {
    double U0, U1, u, Ub, bind_e, fp, W, Wp, sp, pot_energy;
    double lambda, lambda1, lambda2, u0, w0, work, time;
    int64_t step;
    int32_t flags, pad;
    // for the Langevin kernel of the same step
    double vscale, fscale, noisescale, dt;
    int64_t noise_step;
};

__device__ __forceinline__ long long to_fixed(double f, int* bad) {
    if (!(fabs(f) < SDM_FIX_LIMIT)) {
        *bad = 1;
        f = (f != f) ? 0.0 : (f > 0 ? SDM_FIX_LIMIT : -SDM_FIX_LIMIT);
    }
    return __double2ll_rn(f * SDM_FIX_SCALE);
}

// The same fixed point from an FP32 value: floor(f) is the high word and the fraction
// times 2^32 (exact in FP32, truncated below 2^-32) the low word.  Out of range or NaN
// sets *bad and clamps like to_fixed.
__device__ __forceinline__ unsigned long long fixed_from_float(float f, int* bad) {
    if (!(fabsf(f) < 2147483520.0f)) {  // largest FP32 below 2^31
        *bad = 1;
        f = (f != f) ? 0.0f : copysignf(2147483520.0f, f);
    }
    float fl = floorf(f);
    unsigned hi = (unsigned)(int)fl;
    unsigned lo = (unsigned)((f - fl) * 4294967296.0f);
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ double from_fixed(long long v) { return (double)v * (1.0 / SDM_FIX_SCALE); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_min_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double min_image(double d, double L, double iL) {
    return d - L * floor(d * iL + 0.5);
}

// One pair in FP64 (displaced-state path).  d = x_i - x_j after minimum image.
// qq includes k_e; eps4 = 4 sqrt(e_i e_j).  Returns 1 inside the cutoff.
__device__ __forceinline__ int pair64(double dx, double dy, double dz, double qq, double sig,
                                      double eps4, double rc2, double krf, double crf,
                                      double* e, double* fr) {
    double r2 = dx * dx + dy * dy + dz * dz;
    if (!(r2 < rc2)) {
        *e = 0.0;
        *fr = 0.0;
        return 0;
    }
    // rsqrt (the device library's FP64 reciprocal square root, refined to full precision)
    // in place of an IEEE division and a square root: a third fewer instructions per call
    double ir = rsqrt(r2);
    double ir2 = ir * ir;
    double s2 = sig * sig * ir2;
    double s6 = s2 * s2 * s2;
    *e = eps4 * s6 * (s6 - 1.0) + qq * (ir + krf * r2 - crf);
    *fr = eps4 * s6 * (12.0 * s6 - 6.0) * ir2 + qq * (ir * ir2 - 2.0 * krf);
    return 1;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t lo0 = 0xD2511F53u * c[0];
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo1 = 0xCD9E8D57u * c[2];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

// Three standard normals for (seed, replica, step, atom): Box-Muller on one Philox
// block, FP64 throughout.
__device__ __forceinline__ void gauss3(uint64_t seed, uint32_t replica, uint64_t step,
                                       uint32_t atom, double g[3]) {
    uint32_t c[4] = {atom, replica, (uint32_t)step, (uint32_t)(step >> 32)};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double twom32 = 2.3283064365386963e-10;
    const double twopi = 6.283185307179586;
    double u1 = ((double)c[0] + 1.0) * twom32;
    double u2 = (double)c[1] * twom32;
    double u3 = ((double)c[2] + 1.0) * twom32;
    double u4 = (double)c[3] * twom32;
    double r1 = sqrt(-2.0 * log(u1));
    double r2 = sqrt(-2.0 * log(u3));
    double s, co;
    sincos(twopi * u2, &s, &co);
    g[0] = r1 * co;
    g[1] = r1 * s;
    g[2] = r2 * cos(twopi * u4);
}

// ln(1 + e^z) without overflow.
__device__ __forceinline__ double softplus(double z) {
    return fmax(z, 0.0) + log1p(exp(-fabs(z)));
}
// 1 / (1 + e^z) without overflow.
__device__ __forceinline__ double inv_one_plus_exp(double z) {
    if (z > 0) {
        double t = exp(-z);
        return t / (1.0 + t);
    }
    return 1.0 / (1.0 + exp(z));
}

}  // namespace sdm
