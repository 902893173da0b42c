// lbm_math.h -- D2Q9 per-node arithmetic shared by the fused step and the operator kernels.
//
// Every expression keeps the operand order of the reference's CPU kernels and the library is
// compiled with -ffp-contract=off, so each FP64 result is bit-identical to the reference's
// (no fused multiply-add, IEEE-rounded division and square root).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace iblbm {

// iblbm/lbm/lattice.py:43-56
constexpr int EXi[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int EYi[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
__host__ __device__ constexpr double EX(int i) { return (double)EXi[i]; }
__host__ __device__ constexpr double EY(int i) { return (double)EYi[i]; }
// weights as the reference builds them: Python float divisions 4/9, 1/9, 1/36
__host__ __device__ constexpr double W(int i) {
    return i == 0 ? 4.0 / 9.0 : (i < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
}

// iblbm/lbm/equilibrium.py:33-41
__host__ __device__ inline void feq9(double ro, double u0, double u1, bool inc, double *o) {
    const double u2 = u0 * u0 + u1 * u1;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const double eU = EX(i) * u0 + EY(i) * u1;
        if (inc)
            o[i] = W(i) * (ro + 3.0 * eU + 4.5 * eU * eU - 1.5 * u2);
        else
            o[i] = ro * W(i) * (1.0 + 3.0 * eU + 4.5 * eU * eU - 1.5 * u2);
    }
}

// iblbm/lbm/macroscopic.py:35-48 (sequential i = 0..8)
__host__ __device__ inline void macro9(const double *f, bool inc, double &ro, double &u0,
                                       double &u1) {
    double s = 0.0, ux = 0.0, uy = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        s += f[i];
        ux += f[i] * EX(i);
        uy += f[i] * EY(i);
    }
    ro = s;
    if (inc) {
        u0 = ux;
        u1 = uy;
    } else {
        u0 = ux / s;
        u1 = uy / s;
    }
}

// NumPy's reduction of a contiguous axis of length 9:
// (((a0+a1)+(a2+a3))+((a4+a5)+(a6+a7)))+a8
__host__ __device__ inline double np_sum9(const double *a) {
    return (((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]))) + a[8];
}

// _macro_from_streamed + the Neumann copy of iblbm/boundary/dirichlet_neumann.py:61-67,
// :201-206: f_out = f_eq(rho, u) + (f_in - f_eq(rho, u)) with (rho, u) from f_in.
__host__ __device__ inline void neumann_copy(const double *fin, double *fout) {
    double px[9], py[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        px[i] = fin[i] * EX(i);
        py[i] = fin[i] * EY(i);
    }
    const double s = np_sum9(fin);
    const double u0 = np_sum9(px) / s;
    const double u1 = np_sum9(py) / s;
    double q[9];
    feq9(s, u0, u1, false, q);
#pragma unroll
    for (int i = 0; i < 9; ++i) fout[i] = q[i] + (fin[i] - q[i]);
}

// iblbm/boundary/settling_channel.py:65-127: Zou-He no-slip closure of one face node.
// face: 0 left, 1 right, 2 top, 3 bottom.  Written as selects with fixed indices (no face
// writes a value another face's formula reads), so f stays in registers.
template <typename T>
__host__ __device__ inline void settling_face_t(int face, T *f) {
    const T h = (T)0.5;
    const T f1 = f[1], f2 = f[2], f3 = f[3], f4 = f[4];
    const T f5 = f[5], f6 = f[6], f7 = f[7], f8 = f[8];
    const T dv = h * (f2 - f4), dh = h * (f1 - f3);
    f[1] = face == 0 ? f3 : f1;
    f[3] = face == 1 ? f1 : f3;
    f[2] = face == 3 ? f4 : f2;
    f[4] = face == 2 ? f2 : f4;
    f[5] = face == 0 ? f7 - dv : (face == 3 ? f7 - dh : f5);
    f[6] = face == 1 ? f8 - dv : (face == 3 ? f8 + dh : f6);
    f[7] = face == 1 ? f5 + dv : (face == 2 ? f5 + dh : f7);
    f[8] = face == 0 ? f6 + dv : (face == 2 ? f6 - dh : f8);
}
__host__ __device__ inline void settling_face(int face, double *f) { settling_face_t(face, f); }

// Guo term without prefactor, w_i omitted: iblbm/lbm/collision/bgk.py:32-37
__host__ __device__ inline double guo_term(int i, double u0, double u1, double fx, double fy) {
    const double ax = EX(i) - u0;
    const double ay = EY(i) - u1;
    const double eU = EX(i) * u0 + EY(i) * u1;
    return ((3.0 * ax + 9.0 * eU * EX(i)) * fx + (3.0 * ay + 9.0 * eU * EY(i)) * fy);
}

struct CollParams {
    int model;  // 0 BGK, 1 TRT, 2 CM_MRT
    double inv_tau, guo_pref;  // BGK: 1/tau, 1 - 0.5/tau
    double s_plus, s_minus, gpp, gpm;  // TRT
    double S[9];  // CM_MRT
    double dt;
};

// BGK (iblbm/lbm/collision/bgk.py:25-40).  FORCED = false drops the "+ F*dt" term, which
// is exactly +0.0 when fib = 0 (only the sign of an exact zero can differ).
template <bool FORCED>
__host__ __device__ inline void collide_bgk(const CollParams &p, const double *fs,
                                            const double *fe, double u0, double u1, double fx,
                                            double fy, double *out) {
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double v = fs[i] - p.inv_tau * (fs[i] - fe[i]);
        if (FORCED) {
            const double F = p.guo_pref * W(i) * guo_term(i, u0, u1, fx, fy);
            v = v + F * p.dt;
        }
        out[i] = v;
    }
}

// TRT (iblbm/lbm/collision/trt.py:38-67)
template <bool FORCED>
__host__ __device__ inline void collide_trt(const CollParams &p, const double *fs,
                                            const double *fe, double u0, double u1, double fx,
                                            double fy, double *out) {
    double bare[9];
    if (FORCED) {
#pragma unroll
        for (int i = 0; i < 9; ++i) bare[i] = W(i) * guo_term(i, u0, u1, fx, fy);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const int o = OPP[i];
        const double fp = 0.5 * (fs[i] + fs[o]);
        const double fm = 0.5 * (fs[i] - fs[o]);
        const double ep = 0.5 * (fe[i] + fe[o]);
        const double em = 0.5 * (fe[i] - fe[o]);
        const double coll = p.s_plus * (fp - ep) + p.s_minus * (fm - em);
        double v = fs[i] - coll;
        if (FORCED) {
            const double bp = 0.5 * (bare[i] + bare[o]);
            const double bm = 0.5 * (bare[i] - bare[o]);
            const double F = p.gpp * bp + p.gpm * bm;
            v = v + F * p.dt;
        }
        out[i] = v;
    }
}

// Central-moment MRT (iblbm/lbm/collision/cm_mrt.py:52-165); feq is not used.
template <bool FORCED>
__host__ __device__ inline void collide_cm(const CollParams &p, const double *fs, double ux,
                                           double uy, double fx, double fy, double *out) {
    double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0, c4 = 0.0, c5 = 0.0, c6 = 0.0, c7 = 0.0,
           c8 = 0.0;
    double F0 = 0.0, F1 = 0.0, F2 = 0.0, F3 = 0.0, F4 = 0.0, F5 = 0.0, F6 = 0.0, F7 = 0.0,
           F8 = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const double val = fs[i];
        const double cix = EX(i) - ux, ciy = EY(i) - uy;
        const double cix2 = cix * cix, ciy2 = ciy * ciy;
        c0 += val;
        c1 += val * cix;
        c2 += val * ciy;
        c3 += val * cix2;
        c4 += val * ciy2;
        c5 += val * cix * ciy;
        c6 += val * cix2 * ciy;
        c7 += val * cix * ciy2;
        c8 += val * cix2 * ciy2;
        if (FORCED) {
            const double Fb = W(i) * guo_term(i, ux, uy, fx, fy);
            F0 += Fb;
            F1 += Fb * cix;
            F2 += Fb * ciy;
            F3 += Fb * cix2;
            F4 += Fb * ciy2;
            F5 += Fb * cix * ciy;
            F6 += Fb * cix2 * ciy;
            F7 += Fb * cix * ciy2;
            F8 += Fb * cix2 * ciy2;
        }
    }
    const double rho = c0, cs2 = 1.0 / 3.0, cs4 = 1.0 / 9.0;
    const double *S = p.S;
    const double dt = p.dt;
    double p0 = c0 - S[0] * (c0 - rho);
    double p1 = c1 - S[1] * (c1 - 0.0);
    double p2 = c2 - S[2] * (c2 - 0.0);
    double p3 = c3 - S[3] * (c3 - rho * cs2);
    double p4 = c4 - S[4] * (c4 - rho * cs2);
    double p5 = c5 - S[5] * (c5 - 0.0);
    double p6 = c6 - S[6] * (c6 - 0.0);
    double p7 = c7 - S[7] * (c7 - 0.0);
    double p8 = c8 - S[8] * (c8 - rho * cs4);
    if (FORCED) {
        p0 = p0 + (1.0 - 0.5 * S[0]) * F0 * dt;
        p1 = p1 + (1.0 - 0.5 * S[1]) * F1 * dt;
        p2 = p2 + (1.0 - 0.5 * S[2]) * F2 * dt;
        p3 = p3 + (1.0 - 0.5 * S[3]) * F3 * dt;
        p4 = p4 + (1.0 - 0.5 * S[4]) * F4 * dt;
        p5 = p5 + (1.0 - 0.5 * S[5]) * F5 * dt;
        p6 = p6 + (1.0 - 0.5 * S[6]) * F6 * dt;
        p7 = p7 + (1.0 - 0.5 * S[7]) * F7 * dt;
        p8 = p8 + (1.0 - 0.5 * S[8]) * F8 * dt;
    }
    const double ux2 = ux * ux, uy2 = uy * uy;
    const double k0 = p0;
    const double k1 = p1 + ux * p0;
    const double k2 = p2 + uy * p0;
    const double k3 = p3 + 2.0 * ux * p1 + ux2 * p0;
    const double k4 = p4 + 2.0 * uy * p2 + uy2 * p0;
    const double k5 = p5 + uy * p1 + ux * p2 + ux * uy * p0;
    const double k6 = (p6 + uy * p3 + 2.0 * ux * p5 + 2.0 * ux * uy * p1 + ux2 * p2 +
                       ux2 * uy * p0);
    const double k7 = (p7 + 2.0 * uy * p5 + ux * p4 + 2.0 * ux * uy * p2 + uy2 * p1 +
                       ux * uy2 * p0);
    const double k8 = (p8 + 2.0 * uy * p6 + uy2 * p3 + 2.0 * ux * p7 + 4.0 * ux * uy * p5 +
                       2.0 * ux * uy2 * p1 + ux2 * p4 + 2.0 * ux2 * uy * p2 + ux2 * uy2 * p0);
    out[0] = k0 - k3 - k4 + k8;
    out[1] = 0.5 * (k1 + k3 - k7 - k8);
    out[2] = 0.5 * (k2 + k4 - k6 - k8);
    out[3] = 0.5 * (-k1 + k3 + k7 - k8);
    out[4] = 0.5 * (-k2 + k4 + k6 - k8);
    out[5] = 0.25 * (k5 + k6 + k7 + k8);
    out[6] = 0.25 * (-k5 + k6 - k7 + k8);
    out[7] = 0.25 * (k5 - k6 - k7 + k8);
    out[8] = 0.25 * (-k5 - k6 + k7 + k8);
}

// Unforced bulk node update: moments, equilibrium and BGK (MODEL 0) or TRT (MODEL 1)
// collision, equal bit for bit to macro9 + feq9 + collide_bgk / collide_trt<false> for finite,
// normal populations, with about half of their arithmetic.  Two exact identities do it:
//  - a term with a zero lattice-velocity coordinate (0 * x) is an exact zero, and adding an
//    exact zero changes nothing but the sign of an exactly-zero sum, which every later step
//    adds to a nonzero term;
//  - the opposite directions i and OPP[i] share their sums and negate their differences
//    (a + b = b + a and a - b = -(b - a) exactly in round-to-nearest; so do 3 * (-e) = -(3 e)
//    and (4.5 (-e)) (-e) = (4.5 e) e), so each of the 4 pairs is evaluated once.  The rest
//    direction needs no pair: 0.5 (f0 + f0) = f0 exactly.
// Returns rho.
template <int MODEL>
__host__ __device__ inline double bulk_collide64(const CollParams &p, const double *f, bool inc,
                                                 double *out) {
    // macroscopic.py:35-48, sequential over i = 0..8, zero terms left out
    double s = f[0];
#pragma unroll
    for (int i = 1; i < 9; ++i) s += f[i];
    double jx = f[1] - f[3];
    jx = jx + f[5];
    jx = jx - f[6];
    jx = jx - f[7];
    jx = jx + f[8];
    double jy = f[2] - f[4];
    jy = jy + f[5];
    jy = jy + f[6];
    jy = jy - f[7];
    jy = jy - f[8];
    const double u0 = inc ? jx : jx / s, u1 = inc ? jy : jy / s;
    // equilibrium.py:33-41: (c + 3 eU + 4.5 eU eU - 1.5 u2) scaled by rho w (c = 1) or by w
    // (incompressible, c = rho)
    const double a = 1.5 * (u0 * u0 + u1 * u1);
    const double c = inc ? s : 1.0;
    const double rw0 = inc ? W(0) : s * W(0), rw1 = inc ? W(1) : s * W(1),
                 rw5 = inc ? W(5) : s * W(5);
    double e[9];
    e[0] = rw0 * (c - a);
    // eU of directions 1, 2, 5, 6; their opposites 3, 4, 7, 8 have -eU
    const double E[4] = {u0, u1, u0 + u1, u1 - u0};
    const int PI[4] = {1, 2, 5, 6}, MI[4] = {3, 4, 7, 8};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double pe = 3.0 * E[k];
        const double t = 4.5 * E[k] * E[k];
        const double rw = k < 2 ? rw1 : rw5;
        e[PI[k]] = rw * (c + pe + t - a);
        e[MI[k]] = rw * (c - pe + t - a);
    }
    if constexpr (MODEL == 0) {  // bgk.py:25-40
#pragma unroll
        for (int i = 0; i < 9; ++i) out[i] = f[i] - p.inv_tau * (f[i] - e[i]);
    } else {  // trt.py:38-67
        out[0] = f[0] - p.s_plus * (f[0] - e[0]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = PI[k], o = MI[k];
            const double fp = 0.5 * (f[i] + f[o]), fm = 0.5 * (f[i] - f[o]);
            const double ep = 0.5 * (e[i] + e[o]), em = 0.5 * (e[i] - e[o]);
            const double A = p.s_plus * (fp - ep), B = p.s_minus * (fm - em);
            out[i] = f[i] - (A + B);
            out[o] = f[o] - (A - B);
        }
    }
    return s;
}

template <bool FORCED>
__host__ __device__ inline void collide(const CollParams &p, const double *fs, const double *fe,
                                        double u0, double u1, double fx, double fy,
                                        double *out) {
    if (p.model == 0)
        collide_bgk<FORCED>(p, fs, fe, u0, u1, fx, fy, out);
    else if (p.model == 1)
        collide_trt<FORCED>(p, fs, fe, u0, u1, fx, fy, out);
    else
        collide_cm<FORCED>(p, fs, u0, u1, fx, fy, out);
}

// iblbm/ibm/common.py:52-111 (type 0 = peskin4pt, 1 = hat)
__host__ __device__ inline double delta_phi(int type, double r) {
    const double ar = r < 0.0 ? -r : r;
    if (type == 1) {
        const double v = 1.0 - ar;
        return v > 0.0 ? v : 0.0;
    }
    if (ar <= 1.0) {
        const double s = 1.0 + 4.0 * ar - 4.0 * (ar * ar);
        return (1.0 / 8.0) * (3.0 - 2.0 * ar + sqrt(s > 0.0 ? s : 0.0));
    }
    if (ar <= 2.0) {
        const double s = -7.0 + 12.0 * ar - 4.0 * (ar * ar);
        return (1.0 / 8.0) * (5.0 - 2.0 * ar - sqrt(s > 0.0 ? s : 0.0));
    }
    return 0.0;
}

}  // namespace iblbm
