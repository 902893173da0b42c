// dev32.h -- FP32-deviation arithmetic for the bulk lattice tiles.
//
// Populations are stored as g_i = f_i - w_i (rho0 = 1) in FP32.  Every quantity the collision
// needs is formed from the deviations directly, so the rounding error scales with |g| (about
// 1e-3 of w at the workloads here) instead of with |f|:
//   rho = 1 + sum g,   rho*u = sum g e   (sum_i w_i e_i = 0),
//   feq_i - w_i = w_i (drho + rho X_i),  X_i = 3 e.u + 4.5 (e.u)^2 - 1.5 u^2  (compressible),
//               = w_i (drho + X_i)                                              (incompressible).
// The central-moment MRT is linear in f for a fixed u, so its deviation form subtracts the
// central moments of the weight set w about u (closed-form polynomials in u) before relaxing.
#pragma once
#include "lbm_math.h"

namespace iblbm {

__host__ __device__ constexpr float Wf(int i) { return (float)W(i); }
__host__ __device__ constexpr float EXf(int i) { return (float)EXi[i]; }
__host__ __device__ constexpr float EYf(int i) { return (float)EYi[i]; }

struct CollParams32 {
    float inv_tau, s_plus, s_minus;
    float S[9];
};

// FP32 arithmetic may contract a * b + c into one fused multiply-add (the FP32 path is held
// to a tolerance against FP64, not to bit equality); terms with a zero lattice-velocity
// coordinate are left out.
__device__ inline void moments32(const float *g, bool inc, float &drho, float &ux, float &uy) {
#pragma clang fp contract(fast)
    float s = g[0];
#pragma unroll
    for (int i = 1; i < 9; ++i) s += g[i];
    const float jx = g[1] - g[3] + g[5] - g[6] - g[7] + g[8];
    const float jy = g[2] - g[4] + g[5] + g[6] - g[7] - g[8];
    drho = s;
    if (inc) {
        ux = jx;
        uy = jy;
    } else {
        const float r = 1.f / (1.f + s);
        ux = jx * r;
        uy = jy * r;
    }
}

// Deviation equilibrium; directions 3, 4, 7, 8 reuse the terms of their opposites 1, 2, 5, 6
// (eU changes sign, eU^2 does not).
__device__ inline void geq32(float drho, float ux, float uy, bool inc, float *ge) {
#pragma clang fp contract(fast)
    const float a = 1.5f * (ux * ux + uy * uy);
    const float rho = 1.f + drho;
    ge[0] = Wf(0) * (drho + (inc ? -a : rho * -a));
    const float E[4] = {ux, uy, ux + uy, uy - ux};
    const int PI[4] = {1, 2, 5, 6}, MI[4] = {3, 4, 7, 8};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float pe = 3.f * E[k], t = 4.5f * E[k] * E[k] - a;
        const float Xp = pe + t, Xm = t - pe;
        ge[PI[k]] = Wf(PI[k]) * (drho + (inc ? Xp : rho * Xp));
        ge[MI[k]] = Wf(MI[k]) * (drho + (inc ? Xm : rho * Xm));
    }
}

// Unforced collision in deviation form; model is the collision id (0 BGK, 1 TRT, 2 CM-MRT).
template <int MODEL>
__device__ inline void collide32(const CollParams32 &p, const float *g, float drho, float ux,
                                 float uy, bool inc, float *out) {
#pragma clang fp contract(fast)
    if constexpr (MODEL == 0) {
        float ge[9];
        geq32(drho, ux, uy, inc, ge);
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = g[i] - p.inv_tau * (g[i] - ge[i]);
    } else if constexpr (MODEL == 1) {
        float ge[9];
        geq32(drho, ux, uy, inc, ge);
        out[0] = g[0] - p.s_plus * (g[0] - ge[0]);
        const int PI[4] = {1, 2, 5, 6}, MI[4] = {3, 4, 7, 8};
#pragma unroll
        for (int k = 0; k < 4; ++k) {  // the pair (i, OPP[i]) shares gp - ep, negates gm - em
            const int i = PI[k], o = MI[k];
            const float gp = 0.5f * (g[i] + g[o]), gm = 0.5f * (g[i] - g[o]);
            const float ep = 0.5f * (ge[i] + ge[o]), em = 0.5f * (ge[i] - ge[o]);
            const float A = p.s_plus * (gp - ep), B = p.s_minus * (gm - em);
            out[i] = g[i] - (A + B);
            out[o] = g[o] - (A - B);
        }
    } else {
        // central moments of g about u
        float k[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const float cx = EXf(i) - ux, cy = EYf(i) - uy;
            const float cx2 = cx * cx, cy2 = cy * cy, v = g[i];
            k[0] += v;
            k[1] += v * cx;
            k[2] += v * cy;
            k[3] += v * cx2;
            k[4] += v * cy2;
            k[5] += v * cx * cy;
            k[6] += v * cx2 * cy;
            k[7] += v * cx * cy2;
            k[8] += v * cx2 * cy2;
        }
        // (central moments of w about u) - (equilibrium central moments), analytically
        const float ux2 = ux * ux, uy2 = uy * uy, rho = 1.f + drho;
        float dw[9];
        dw[0] = -drho;                                      // 1 - rho
        dw[1] = -ux;                                        // k10(w) - 0
        dw[2] = -uy;
        dw[3] = ux2 - drho * (1.f / 3.f);                   // (1/3 + ux^2) - rho/3
        dw[4] = uy2 - drho * (1.f / 3.f);
        dw[5] = ux * uy;
        dw[6] = -uy * (1.f / 3.f) - ux2 * uy;
        dw[7] = -ux * (1.f / 3.f) - ux * uy2;
        dw[8] = (ux2 + uy2) * (1.f / 3.f) + ux2 * uy2 - drho * (1.f / 9.f);
        (void)rho;
        float c[9];
#pragma unroll
        for (int m = 0; m < 9; ++m) c[m] = k[m] - p.S[m] * (k[m] + dw[m]);
        // shift back to raw moments and rebuild populations (cm_mrt.py:137-163 algebra)
        const float k0 = c[0];
        const float k1 = c[1] + ux * c[0];
        const float k2 = c[2] + uy * c[0];
        const float k3 = c[3] + 2.f * ux * c[1] + ux2 * c[0];
        const float k4 = c[4] + 2.f * uy * c[2] + uy2 * c[0];
        const float k5 = c[5] + uy * c[1] + ux * c[2] + ux * uy * c[0];
        const float k6 = c[6] + uy * c[3] + 2.f * ux * c[5] + 2.f * ux * uy * c[1] + ux2 * c[2] +
                         ux2 * uy * c[0];
        const float k7 = c[7] + 2.f * uy * c[5] + ux * c[4] + 2.f * ux * uy * c[2] + uy2 * c[1] +
                         ux * uy2 * c[0];
        const float k8 = c[8] + 2.f * uy * c[6] + uy2 * c[3] + 2.f * ux * c[7] +
                         4.f * ux * uy * c[5] + 2.f * ux * uy2 * c[1] + ux2 * c[4] +
                         2.f * ux2 * uy * c[2] + ux2 * uy2 * c[0];
        out[0] = k0 - k3 - k4 + k8;
        out[1] = 0.5f * (k1 + k3 - k7 - k8);
        out[2] = 0.5f * (k2 + k4 - k6 - k8);
        out[3] = 0.5f * (-k1 + k3 + k7 - k8);
        out[4] = 0.5f * (-k2 + k4 + k6 - k8);
        out[5] = 0.25f * (k5 + k6 + k7 + k8);
        out[6] = 0.25f * (-k5 + k6 - k7 + k8);
        out[7] = 0.25f * (k5 - k6 - k7 + k8);
        out[8] = 0.25f * (-k5 - k6 + k7 + k8);
    }
}

// Neumann copy in deviation form: g_out = geq(rho, u) + (g_in - geq(rho, u)), rho, u of g_in
__device__ inline void neumann_copy32(const float *gin, float *gout) {
    float drho, ux, uy, ge[9];
    moments32(gin, false, drho, ux, uy);
    geq32(drho, ux, uy, false, ge);
#pragma unroll
    for (int i = 0; i < 9; ++i) gout[i] = ge[i] + (gin[i] - ge[i]);
}

__device__ inline void settling_face32(int face, float *f) {
    // identical to the FP64 closure: w is symmetric under every swap it performs
    settling_face_t(face, f);
}

}  // namespace iblbm
