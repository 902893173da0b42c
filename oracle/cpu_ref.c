/*
 * cpu_ref.c -- FP64 CPU restatement of the reference IB-LBM per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP path is checked against and the
 * CPU baseline that the benchmark times beside it.  Nothing in the shipped library links it.
 *
 * It restates the reference's CPU backend (IBLBM_GPU=0: numba kernels in iblbm/lbm/*.py and
 * the vectorised NumPy IBM in iblbm/ibm/*.py) in plain C over the reference's own layout:
 * populations are AoS (N, 9) with node n = j*nx + i, U is (N, 2), Eulerian 2-D fields are
 * (ny, nx).  Every expression keeps the reference's operand order and is compiled with
 * -ffp-contract=off, so each value is the IEEE-double result the reference computes.
 * Where NumPy reduces (np.sum over an axis of 9, np.add.at, np.max) the NumPy order is
 * restated: a 9-term axis sum is NumPy's pairwise form ((a0+a1)+(a2+a3))+((a4+a5)+(a6+a7))+a8,
 * np.add.at accumulates in index order.
 *
 * Parallelism: OpenMP over nodes for the per-node kernels (each node is independent, so the
 * thread count does not change any value); the IBM stages are serial, as np.add.at is.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* iblbm/lbm/lattice.py:43-56 */
static const double EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const double EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static double W[9];

static void init_w(void) {
    W[0] = 4.0 / 9.0;
    for (int i = 1; i < 5; ++i) W[i] = 1.0 / 9.0;
    for (int i = 5; i < 9; ++i) W[i] = 1.0 / 36.0;
}

int ref_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* iblbm/lbm/equilibrium.py:27-42 (_compute_feq_nb) */
static inline void feq_node(double ro, double u0, double u1, int inc, double *out) {
    double u2 = u0 * u0 + u1 * u1;
    for (int i = 0; i < 9; ++i) {
        double eU = EX[i] * u0 + EY[i] * u1;
        if (inc)
            out[i] = W[i] * (ro + 3.0 * eU + 4.5 * eU * eU - 1.5 * u2);
        else
            out[i] = ro * W[i] * (1.0 + 3.0 * eU + 4.5 * eU * eU - 1.5 * u2);
    }
}

void ref_feq(const double *ro, const double *U, long N, int inc, double *feq) {
    init_w();
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; ++n) feq_node(ro[n], U[2 * n], U[2 * n + 1], inc, feq + 9 * n);
}

/* iblbm/lbm/macroscopic.py:27-49 (_macroscopic_nb) */
void ref_macro(const double *f, long N, int inc, double *ro, double *U) {
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; ++n) {
        double s = 0.0, ux = 0.0, uy = 0.0;
        const double *fn = f + 9 * n;
        for (int i = 0; i < 9; ++i) {
            s += fn[i];
            ux += fn[i] * EX[i];
            uy += fn[i] * EY[i];
        }
        ro[n] = s;
        if (inc) {
            U[2 * n] = ux;
            U[2 * n + 1] = uy;
        } else {
            U[2 * n] = ux / s;
            U[2 * n + 1] = uy / s;
        }
    }
}

/* iblbm/lbm/streaming.py:38-77: push streaming onto a copy of the previous fstar. */
void ref_stream(const double *fstar_old, const double *f, int nx, int ny, double *out) {
    long N = (long)nx * ny;
    memcpy(out, fstar_old, sizeof(double) * 9 * N);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            long n = (long)j * nx + i;
            for (int q = 0; q < 9; ++q) {
                int si = i - (int)EX[q], sj = j - (int)EY[q];
                if (si < 0 || si >= nx || sj < 0 || sj >= ny) continue;
                out[9 * n + q] = f[9 * ((long)sj * nx + si) + q];
            }
        }
    }
}

/* Guo forcing term w_i [3(e-u) + 9(e.u)e].fib without the prefactor
 * (iblbm/lbm/collision/bgk.py:31-37, trt.py:45-53, cm_mrt.py:96-101). */
static inline double guo_term(int i, double u0, double u1, double fx, double fy) {
    double ax = EX[i] - u0;
    double ay = EY[i] - u1;
    double eU = EX[i] * u0 + EY[i] * u1;
    return ((3.0 * ax + 9.0 * eU * EX[i]) * fx + (3.0 * ay + 9.0 * eU * EY[i]) * fy);
}

/* iblbm/lbm/collision/bgk.py:25-40 */
void ref_collide_bgk(const double *fstar, const double *feq, const double *U, const double *fib,
                     double tau, double dt, long N, double *out) {
    init_w();
    double inv_tau = 1.0 / tau;
    double guo_pref = 1.0 - 0.5 / tau;
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; ++n) {
        for (int i = 0; i < 9; ++i) {
            double term = guo_term(i, U[2 * n], U[2 * n + 1], fib[2 * n], fib[2 * n + 1]);
            double F = guo_pref * W[i] * term;
            out[9 * n + i] = fstar[9 * n + i] - inv_tau * (fstar[9 * n + i] - feq[9 * n + i]) + F * dt;
        }
    }
}

/* iblbm/lbm/collision/trt.py:38-67 (s_plus = 1/tau, s_minus = 1/tau_minus from trt.py:81-82) */
void ref_collide_trt(const double *fstar, const double *feq, const double *U, const double *fib,
                     double s_plus, double s_minus, double dt, long N, double *out) {
    init_w();
    double gpp = 1.0 - 0.5 * s_plus;
    double gpm = 1.0 - 0.5 * s_minus;
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; ++n) {
        double bare[9];
        const double *fs = fstar + 9 * n, *fe = feq + 9 * n;
        for (int i = 0; i < 9; ++i)
            bare[i] = W[i] * guo_term(i, U[2 * n], U[2 * n + 1], fib[2 * n], fib[2 * n + 1]);
        for (int i = 0; i < 9; ++i) {
            int o = OPP[i];
            double fp = 0.5 * (fs[i] + fs[o]);
            double fm = 0.5 * (fs[i] - fs[o]);
            double ep = 0.5 * (fe[i] + fe[o]);
            double em = 0.5 * (fe[i] - fe[o]);
            double coll = s_plus * (fp - ep) + s_minus * (fm - em);
            double bp = 0.5 * (bare[i] + bare[o]);
            double bm = 0.5 * (bare[i] - bare[o]);
            double F = gpp * bp + gpm * bm;
            out[9 * n + i] = fs[i] - coll + F * dt;
        }
    }
}

/* iblbm/lbm/collision/cm_mrt.py:52-165 (feq is not read by the reference kernel) */
void ref_collide_cm(const double *fstar, const double *U, const double *fib, const double *S,
                    double dt, long N, double *out) {
    init_w();
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; ++n) {
        double ux = U[2 * n], uy = U[2 * n + 1];
        double c[9] = {0}, F[9] = {0};
        for (int i = 0; i < 9; ++i) {
            double val = fstar[9 * n + i];
            double cix = EX[i] - ux, ciy = EY[i] - uy;
            double cix2 = cix * cix, ciy2 = ciy * ciy;
            c[0] += val;
            c[1] += val * cix;
            c[2] += val * ciy;
            c[3] += val * cix2;
            c[4] += val * ciy2;
            c[5] += val * cix * ciy;
            c[6] += val * cix2 * ciy;
            c[7] += val * cix * ciy2;
            c[8] += val * cix2 * ciy2;
            double Fb = W[i] * guo_term(i, ux, uy, fib[2 * n], fib[2 * n + 1]);
            F[0] += Fb;
            F[1] += Fb * cix;
            F[2] += Fb * ciy;
            F[3] += Fb * cix2;
            F[4] += Fb * ciy2;
            F[5] += Fb * cix * ciy;
            F[6] += Fb * cix2 * ciy;
            F[7] += Fb * cix * ciy2;
            F[8] += Fb * cix2 * ciy2;
        }
        double rho = c[0], cs2 = 1.0 / 3.0, cs4 = 1.0 / 9.0;
        double p0 = c[0] - S[0] * (c[0] - rho) + (1.0 - 0.5 * S[0]) * F[0] * dt;
        double p1 = c[1] - S[1] * (c[1] - 0.0) + (1.0 - 0.5 * S[1]) * F[1] * dt;
        double p2 = c[2] - S[2] * (c[2] - 0.0) + (1.0 - 0.5 * S[2]) * F[2] * dt;
        double p3 = c[3] - S[3] * (c[3] - rho * cs2) + (1.0 - 0.5 * S[3]) * F[3] * dt;
        double p4 = c[4] - S[4] * (c[4] - rho * cs2) + (1.0 - 0.5 * S[4]) * F[4] * dt;
        double p5 = c[5] - S[5] * (c[5] - 0.0) + (1.0 - 0.5 * S[5]) * F[5] * dt;
        double p6 = c[6] - S[6] * (c[6] - 0.0) + (1.0 - 0.5 * S[6]) * F[6] * dt;
        double p7 = c[7] - S[7] * (c[7] - 0.0) + (1.0 - 0.5 * S[7]) * F[7] * dt;
        double p8 = c[8] - S[8] * (c[8] - rho * cs4) + (1.0 - 0.5 * S[8]) * F[8] * dt;
        double ux2 = ux * ux, uy2 = uy * uy;
        double k0 = p0;
        double k1 = p1 + ux * p0;
        double k2 = p2 + uy * p0;
        double k3 = p3 + 2.0 * ux * p1 + ux2 * p0;
        double k4 = p4 + 2.0 * uy * p2 + uy2 * p0;
        double k5 = p5 + uy * p1 + ux * p2 + ux * uy * p0;
        double k6 = (p6 + uy * p3 + 2.0 * ux * p5 + 2.0 * ux * uy * p1 + ux2 * p2 + ux2 * uy * p0);
        double k7 = (p7 + 2.0 * uy * p5 + ux * p4 + 2.0 * ux * uy * p2 + uy2 * p1 + ux * uy2 * p0);
        double k8 = (p8 + 2.0 * uy * p6 + uy2 * p3 + 2.0 * ux * p7 + 4.0 * ux * uy * p5
                     + 2.0 * ux * uy2 * p1 + ux2 * p4 + 2.0 * ux2 * uy * p2 + ux2 * uy2 * p0);
        double *o = out + 9 * n;
        o[0] = k0 - k3 - k4 + k8;
        o[1] = 0.5 * (k1 + k3 - k7 - k8);
        o[2] = 0.5 * (k2 + k4 - k6 - k8);
        o[3] = 0.5 * (-k1 + k3 + k7 - k8);
        o[4] = 0.5 * (-k2 + k4 + k6 - k8);
        o[5] = 0.25 * (k5 + k6 + k7 + k8);
        o[6] = 0.25 * (-k5 + k6 - k7 + k8);
        o[7] = 0.25 * (k5 - k6 - k7 + k8);
        o[8] = 0.25 * (-k5 - k6 + k7 + k8);
    }
}

/* NumPy's reduction of a contiguous axis of length 9 (pairwise form). */
static inline double np_sum9(const double *a) {
    return (((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]))) + a[8];
}

/* _macro_from_streamed of iblbm/boundary/dirichlet_neumann.py:61-67 for one node. */
static inline void macro_streamed(const double *fn, double *ro, double *u0, double *u1) {
    double px[9], py[9];
    for (int i = 0; i < 9; ++i) {
        px[i] = fn[i] * EX[i];
        py[i] = fn[i] * EY[i];
    }
    double s = np_sum9(fn);
    *ro = s;
    *u0 = np_sum9(px) / s;
    *u1 = np_sum9(py) / s;
}

/* iblbm/boundary/dirichlet_neumann.py:39-118.  The macroscopic writes of the reference are
 * overwritten by the full macroscopic pass that follows, so only fstar is produced. */
void ref_bc_dirichlet_neumann(double *fs, int nx, int ny, double in_u, double dens) {
    init_w();
    double fe[9];
    feq_node(dens, in_u, 0.0, 0, fe);
    for (int j = 1; j < ny - 1; ++j) memcpy(fs + 9 * ((long)j * nx), fe, sizeof fe);          /* left */
    for (int i = 1; i < nx - 1; ++i) memcpy(fs + 9 * ((long)(ny - 1) * nx + i), fe, sizeof fe); /* top */
    for (int i = 1; i < nx - 1; ++i) memcpy(fs + 9 * (long)i, fe, sizeof fe);                   /* bottom */
    for (int j = 1; j < ny - 1; ++j) {                                                           /* right */
        const double *fin = fs + 9 * ((long)j * nx + nx - 2);
        double r, u0, u1, q[9];
        macro_streamed(fin, &r, &u0, &u1);
        feq_node(r, u0, u1, 0, q);
        double *fo = fs + 9 * ((long)j * nx + nx - 1);
        for (int i = 0; i < 9; ++i) fo[i] = q[i] + (fin[i] - q[i]);
    }
    memcpy(fs, fe, sizeof fe);
    memcpy(fs + 9 * ((long)nx * (ny - 1)), fe, sizeof fe);
    long rc[2] = {nx - 1, (long)nx * ny - 1}, rd[2] = {2L * nx - 2, (long)nx * (ny - 1) - 2};
    for (int c = 0; c < 2; ++c) {
        const double *fin = fs + 9 * rd[c];
        double r, u0, u1, q[9];
        macro_streamed(fin, &r, &u0, &u1);
        feq_node(r, u0, u1, 0, q);
        for (int i = 0; i < 9; ++i) fs[9 * rc[c] + i] = q[i] + (fin[i] - q[i]);
    }
}

/* iblbm/boundary/settling_channel.py:42-142 (face order left, right, top, bottom, corners). */
void ref_bc_settling(double *fs, int nx, int ny, double dens, int inc) {
    init_w();
    for (int j = 1; j < ny - 1; ++j) {
        double *f = fs + 9 * ((long)j * nx);
        f[1] = f[3];
        f[5] = f[7] - 0.5 * (f[2] - f[4]);
        f[8] = f[6] + 0.5 * (f[2] - f[4]);
    }
    for (int j = 1; j < ny - 1; ++j) {
        double *f = fs + 9 * ((long)j * nx + nx - 1);
        f[3] = f[1];
        f[7] = f[5] + 0.5 * (f[2] - f[4]);
        f[6] = f[8] - 0.5 * (f[2] - f[4]);
    }
    for (int i = 1; i < nx - 1; ++i) {
        double *f = fs + 9 * ((long)(ny - 1) * nx + i);
        f[4] = f[2];
        f[7] = f[5] + 0.5 * (f[1] - f[3]);
        f[8] = f[6] - 0.5 * (f[1] - f[3]);
    }
    for (int i = 1; i < nx - 1; ++i) {
        double *f = fs + 9 * (long)i;
        f[2] = f[4];
        f[5] = f[7] - 0.5 * (f[1] - f[3]);
        f[6] = f[8] + 0.5 * (f[1] - f[3]);
    }
    double fe[9];
    feq_node(dens, 0.0, 0.0, inc, fe);
    long c[4] = {0, nx - 1, (long)nx * (ny - 1), (long)nx * ny - 1};
    for (int k = 0; k < 4; ++k) memcpy(fs + 9 * c[k], fe, sizeof fe);
}

/* iblbm/ibm/common.py:52-111 (delta_type 0 = peskin4pt, 1 = hat) */
double ref_delta(int type, double r) {
    double ar = fabs(r);
    if (type == 1) {
        double v = 1.0 - ar;
        return v > 0.0 ? v : 0.0; /* np.maximum(1 - |r|, 0) */
    }
    if (ar <= 1.0) {
        double s = 1.0 + 4.0 * ar - 4.0 * (ar * ar);
        return (1.0 / 8.0) * (3.0 - 2.0 * ar + sqrt(s > 0.0 ? s : 0.0));
    }
    if (ar <= 2.0) {
        double s = -7.0 + 12.0 * ar - 4.0 * (ar * ar);
        return (1.0 / 8.0) * (5.0 - 2.0 * ar - sqrt(s > 0.0 ? s : 0.0));
    }
    return 0.0;
}

static inline long clampl(long v, long lo, long hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* iblbm/ibm/df.py:94-154 (CPU branch).  Eux/Euy/Ro are (ny, nx); fib is (N, 2), zero-filled
 * here.  Lux, Luy, Lfx, Lfy, R are (Lb,). */
void ref_df(const double *Lx, const double *Ly, const double *ud, int Lb, const double *Eux,
            const double *Euy, const double *Ro, double dx, double dy, double dt, double Larea,
            int ny, int nx, int dtype, int inc, double *fib, double *Lux, double *Luy,
            double *Lfx, double *Lfy, double *R) {
    int a = dtype == 1 ? 1 : 2;
    long N = (long)nx * ny;
    for (int k = 0; k < Lb; ++k) Lux[k] = Luy[k] = R[k] = 0.0;
    for (int di = -a; di <= a; ++di)
        for (int dj = -a; dj <= a; ++dj)
            for (int k = 0; k < Lb; ++k) {
                long ix0 = (long)floor(Lx[k] / dx + 0.5), iy0 = (long)floor(Ly[k] / dy + 0.5);
                long ei = ix0 + di, ej = iy0 + dj;
                long eic = clampl(ei, 0, nx - 1), ejc = clampl(ej, 0, ny - 1);
                double wx = ref_delta(dtype, (Lx[k] - (double)ei * dx) / dx);
                double wy = ref_delta(dtype, (Ly[k] - (double)ej * dy) / dy);
                double w = wx * wy;
                long m = ejc * nx + eic;
                R[k] += Ro[m] * w;
                Lux[k] += Eux[m] * w;
                Luy[k] += Euy[m] * w;
            }
    for (int k = 0; k < Lb; ++k) {
        if (inc) {
            Lfx[k] = 2.0 * (ud[2 * k] - Lux[k]) / dt;
            Lfy[k] = 2.0 * (ud[2 * k + 1] - Luy[k]) / dt;
        } else {
            Lfx[k] = 2.0 * R[k] * (ud[2 * k] - Lux[k]) / dt;
            Lfy[k] = 2.0 * R[k] * (ud[2 * k + 1] - Luy[k]) / dt;
        }
    }
    memset(fib, 0, sizeof(double) * 2 * N);
    for (int di = -a; di <= a; ++di)
        for (int dj = -a; dj <= a; ++dj)
            for (int k = 0; k < Lb; ++k) {
                long ix0 = (long)floor(Lx[k] / dx + 0.5), iy0 = (long)floor(Ly[k] / dy + 0.5);
                long ei = ix0 + di, ej = iy0 + dj;
                long eic = clampl(ei, 0, nx - 1), ejc = clampl(ej, 0, ny - 1);
                double wx = ref_delta(dtype, ((double)ei * dx - Lx[k]) / dx);
                double wy = ref_delta(dtype, ((double)ej * dy - Ly[k]) / dy);
                double w = wx * wy * Larea;
                long m = ejc * nx + eic;
                fib[2 * m] += Lfx[k] * w;
                fib[2 * m + 1] += Lfy[k] * w;
            }
}

/* iblbm/ibm/mdf.py:113-229, CPU branch without rotation coupling.  U is (N, 2), ro (N,).
 * stats = {iter_count, last_residual, diverged, converged}. */
void ref_mdf(const double *Lx, const double *Ly, const double *ud, int Lb, const double *U,
             const double *ro, double dx, double dy, double dt, double Larea, int ny, int nx,
             int n_iter, int min_iter, int dtype, double omega, double tol, double *fib_total,
             double *stats) {
    long N = (long)nx * ny;
    double *Ex = malloc(sizeof(double) * N), *Ey = malloc(sizeof(double) * N);
    double *vs = malloc(sizeof(double) * N), *fib = malloc(sizeof(double) * 2 * N);
    double *Lux = malloc(sizeof(double) * Lb), *Luy = malloc(sizeof(double) * Lb);
    double *Lfx = malloc(sizeof(double) * Lb), *Lfy = malloc(sizeof(double) * Lb);
    double *R = malloc(sizeof(double) * Lb);
    for (long n = 0; n < N; ++n) {
        Ex[n] = U[2 * n];
        Ey[n] = U[2 * n + 1];
        vs[n] = dt / (2.0 * ro[n]);
    }
    if (min_iter > n_iter) min_iter = n_iter;
    if (min_iter < 1) min_iter = 1;
    double prev = INFINITY, last = INFINITY;
    int iter_count = 0, diverged = 0, converged = 0;
    memset(fib_total, 0, sizeof(double) * 2 * N);
    for (int it = 0; it < n_iter; ++it) {
        ref_df(Lx, Ly, ud, Lb, Ex, Ey, ro, dx, dy, dt, Larea, ny, nx, dtype, 0, fib, Lux, Luy, Lfx,
               Lfy, R);
        double res = -INFINITY;
        for (int k = 0; k < Lb; ++k) {
            double vx = ud[2 * k] - Lux[k], vy = ud[2 * k + 1] - Luy[k];
            double s = sqrt(vx * vx + vy * vy);
            if (s > res || isnan(s)) res = s;
        }
        last = res;
        if (it > 0 && res > prev * 1.5) {
            diverged = 1;
            iter_count = it + 1;
            break;
        }
        for (long n = 0; n < N; ++n) {
            double fx = fib[2 * n] * omega, fy = fib[2 * n + 1] * omega;
            fib_total[2 * n] += fx;
            fib_total[2 * n + 1] += fy;
            Ex[n] += fx * vs[n];
            Ey[n] += fy * vs[n];
        }
        prev = res;
        iter_count = it + 1;
        if (it + 1 >= min_iter && res < tol) {
            converged = 1;
            break;
        }
    }
    stats[0] = iter_count;
    stats[1] = last;
    stats[2] = diverged;
    stats[3] = converged;
    free(Ex); free(Ey); free(vs); free(fib); free(Lux); free(Luy); free(Lfx); free(Lfy); free(R);
}

/* iblbm/ibm/dfc.py:126-186 (compute_lambda, rho_f passed in). */
void ref_dfc_lambda(const double *Lx, const double *Ly, int Lb, double rho_f, double Larea,
                    double dx, double dy, int ny, int nx, int dtype, double *lam) {
    int a = dtype == 1 ? 1 : 2;
    long N = (long)nx * ny;
    double *Wt = calloc(N, sizeof(double));
    double *Ws = calloc(Lb, sizeof(double));
    for (int pass = 0; pass < 2; ++pass)
        for (int di = -a; di <= a; ++di)
            for (int dj = -a; dj <= a; ++dj)
                for (int k = 0; k < Lb; ++k) {
                    long ix0 = (long)floor(Lx[k] / dx + 0.5), iy0 = (long)floor(Ly[k] / dy + 0.5);
                    long ei = ix0 + di, ej = iy0 + dj;
                    long m = clampl(ej, 0, ny - 1) * nx + clampl(ei, 0, nx - 1);
                    double w = ref_delta(dtype, (Lx[k] - (double)ei * dx) / dx) *
                               ref_delta(dtype, (Ly[k] - (double)ej * dy) / dy);
                    if (pass == 0)
                        Wt[m] += w;
                    else
                        Ws[k] += Wt[m] * w;
                }
    for (int k = 0; k < Lb; ++k) lam[k] = 1.0 / (2.0 * rho_f * Larea * Ws[k]);
    free(Wt);
    free(Ws);
}

/* iblbm/ibm/dfc.py:55-95, :98-123, :189-253, :319-341 (CPU branch, correction_scale = cs).
 * fstar (N, 9) is corrected in place; force is (Lb, 2). */
void ref_dfc_apply(const double *Lx, const double *Ly, const double *ud, int Lb, double *fstar,
                   const double *lam, double dx, double dy, double Larea, int ny, int nx,
                   int dtype, double cs, double *force) {
    init_w();
    int a = dtype == 1 ? 1 : 2;
    long N = (long)nx * ny;
    double *fi = calloc(9 * (size_t)Lb, sizeof(double));
    double *dev = malloc(sizeof(double) * 9 * Lb);
    double *dfl = malloc(sizeof(double) * 9 * Lb);
    double *Ec = calloc(9 * (size_t)N, sizeof(double));
    for (int di = -a; di <= a; ++di)
        for (int dj = -a; dj <= a; ++dj)
            for (int k = 0; k < Lb; ++k) {
                long ix0 = (long)floor(Lx[k] / dx + 0.5), iy0 = (long)floor(Ly[k] / dy + 0.5);
                long ei = ix0 + di, ej = iy0 + dj;
                long m = clampl(ej, 0, ny - 1) * nx + clampl(ei, 0, nx - 1);
                double w = ref_delta(dtype, (Lx[k] - (double)ei * dx) / dx) *
                           ref_delta(dtype, (Ly[k] - (double)ej * dy) / dy);
                for (int q = 0; q < 9; ++q) fi[9 * k + q] += fstar[9 * m + q] * w;
            }
    const double cs2 = 1.0 / 3.0;
    for (int k = 0; k < Lb; ++k)
        for (int q = 0; q < 9; ++q) {
            double edu = ud[2 * k] * EX[q] + ud[2 * k + 1] * EY[q];
            double fbb = fi[9 * k + OPP[q]] + 2.0 * W[q] * 1.0 * edu / cs2;
            dev[9 * k + q] = fbb - fi[9 * k + q];
            dfl[9 * k + q] = cs * lam[k] * dev[9 * k + q];
        }
    for (int di = -a; di <= a; ++di)
        for (int dj = -a; dj <= a; ++dj)
            for (int k = 0; k < Lb; ++k) {
                long ix0 = (long)floor(Lx[k] / dx + 0.5), iy0 = (long)floor(Ly[k] / dy + 0.5);
                long ei = ix0 + di, ej = iy0 + dj;
                long m = clampl(ej, 0, ny - 1) * nx + clampl(ei, 0, nx - 1);
                double w = ref_delta(dtype, (Lx[k] - (double)ei * dx) / dx) *
                           ref_delta(dtype, (Ly[k] - (double)ej * dy) / dy) * Larea;
                for (int q = 0; q < 9; ++q) Ec[9 * m + q] += dfl[9 * k + q] * w;
            }
    for (long n = 0; n < 9 * N; ++n) fstar[n] = fstar[n] + Ec[n];
    for (int k = 0; k < Lb; ++k) {
        double sx = 0.0, sy = 0.0;
        for (int q = 0; q < 9; ++q) {
            double sd = -(lam[k] * dev[9 * k + q] * Larea);
            sx += sd * EX[q];
            sy += sd * EY[q];
        }
        force[2 * k] = cs * sx;
        force[2 * k + 1] = cs * sy;
    }
    free(fi); free(dev); free(dfl); free(Ec);
}

/* iblbm/ibm/a_norm.py:75-172 in lattice units (dx = 1) as called from :204-266.
 * Returns ||A||_inf; Lx, Ly are already divided by the domain dx. */
double ref_a_norm_inf(const double *Lx, const double *Ly, int Lb, int dtype, double Larea) {
    int a = dtype == 1 ? 1 : 2;
    double cutoff = 2.0 * a * 1.0, factor = Larea / (1.0 * 1.0), best = -INFINITY;
    for (int l = 0; l < Lb; ++l) {
        double acc = 0.0;
        for (int m = 0; m < Lb; ++m) {
            double xl = Lx[l], yl = Ly[l], xm = Lx[m], ym = Ly[m];
            if (fabs(xl - xm) >= cutoff || fabs(yl - ym) >= cutoff) continue;
            double sums[2];
            int skip = 0;
            for (int ax = 0; ax < 2 && !skip; ++ax) {
                double pl = ax ? yl : xl, pm = ax ? ym : xm;
                double lo = (pl > pm ? pl : pm) - a * 1.0, hi = (pl < pm ? pl : pm) + a * 1.0;
                long ilo = (long)ceil(lo / 1.0), ihi = (long)floor(hi / 1.0) + 1;
                if (ilo >= ihi) { skip = 1; break; }
                double s = 0.0;
                for (long i = ilo; i < ihi; ++i) {
                    double xi = (double)i * 1.0;
                    s += ref_delta(dtype, (xi - pl) / 1.0) * ref_delta(dtype, (xi - pm) / 1.0);
                }
                if (s == 0.0) { skip = 1; break; }
                sums[ax] = s;
            }
            if (skip) continue;
            acc += factor * sums[0] * sums[1];
        }
        if (acc > best) best = acc;
    }
    return best;
}

/* ---------------------------------------------------------------------------------------
 * Zou-He face closures with prescribed or extrapolated velocities.  U is the state's
 * velocity of the previous step (N, 2), read only (the closures' writes of rho and U are
 * overwritten by the macroscopic pass that follows, so only fstar is produced here; the
 * velocities a closure sets are kept in locals).  Face index sets as
 * iblbm/boundary/indices.py:33-50. */

/* Left face at node f with velocity (u0, u1): rho from the known populations over 1 - u0. */
static inline void zh_left_rho(double *f, double u0, double u1) {
    const double ro = (1.0 / (1.0 - u0)) * (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7]));
    f[1] = f[3] + (2.0 / 3.0) * ro * u0;
    f[5] = f[7] - 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * ro * u0 + 0.5 * ro * u1;
    f[8] = f[6] + 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * ro * u0 - 0.5 * ro * u1;
}
/* Right face, rho given; u0 from the known populations (pressure form). */
static inline void zh_right_p(double *f, double ro, double u1) {
    const double u0 = -1.0 + (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])) / ro;
    f[3] = f[1] - (2.0 / 3.0) * ro * u0;
    f[7] = f[5] + 0.5 * (f[2] - f[4]) - (1.0 / 6.0) * ro * u0 - 0.5 * ro * u1;
    f[6] = f[8] - 0.5 * (f[2] - f[4]) - (1.0 / 6.0) * ro * u0 + 0.5 * ro * u1;
}

static inline double extrap(double a1, double a2) { return (4.0 * a1 - a2) / 3.0; }

/* iblbm/boundary/velocity_inlet.py:45-157 (left and right faces with their corners, then
 * top, bottom, then the corner refills). */
void ref_bc_velocity_inlet(double *fs, const double *U, int nx, int ny, double in_u,
                           double dens) {
    init_w();
    for (int j = 0; j < ny; ++j) zh_left_rho(fs + 9 * ((long)j * nx), in_u, 0.0);
    for (int j = 0; j < ny; ++j) {
        const long n = (long)j * nx + nx - 1;
        zh_right_p(fs + 9 * n, dens, extrap(U[2 * (n - 1) + 1], U[2 * (n - 2) + 1]));
    }
    for (int i = 1; i < nx - 1; ++i) {  /* top: no-slip, tangential u extrapolated */
        const long n = (long)(ny - 1) * nx + i;
        double *f = fs + 9 * n;
        const double u0 = extrap(U[2 * (n - nx)], U[2 * (n - 2L * nx)]), u1 = 0.0;
        const double ro = (1.0 / (1.0 + u1)) * (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6]));
        f[4] = f[2] - (2.0 / 3.0) * ro * u1;
        f[7] = f[5] + 0.5 * (f[1] - f[3]) - 0.5 * ro * u0;
        f[8] = f[6] - 0.5 * (f[1] - f[3]) + 0.5 * ro * u0;
    }
    for (int i = 1; i < nx - 1; ++i) {  /* bottom */
        const long n = i;
        double *f = fs + 9 * n;
        const double u0 = extrap(U[2 * (n + nx)], U[2 * (n + 2L * nx)]), u1 = 0.0;
        const double ro = (1.0 / (1.0 - u1)) * (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8]));
        f[2] = f[4] + (2.0 / 3.0) * ro * u1;
        f[5] = f[7] - 0.5 * (f[1] - f[3]) + 0.5 * ro * u0;
        f[6] = f[8] + 0.5 * (f[1] - f[3]) - 0.5 * ro * u0;
    }
    double fe[9];
    feq_node(dens, in_u, 0.0, 0, fe); /* inlet corners */
    memcpy(fs, fe, sizeof fe);
    memcpy(fs + 9 * ((long)(ny - 1) * nx), fe, sizeof fe);
    /* outlet corners: u0 extrapolated from the bottom / top face values just set */
    for (int c = 0; c < 2; ++c) {
        const int j = c == 0 ? 0 : ny - 1;
        const int jj = c == 0 ? 1 : ny - 2, jk = c == 0 ? 2 : ny - 3;
        const double b1 = extrap(U[2 * ((long)jj * nx + nx - 2)], U[2 * ((long)jk * nx + nx - 2)]);
        const double b2 = extrap(U[2 * ((long)jj * nx + nx - 3)], U[2 * ((long)jk * nx + nx - 3)]);
        feq_node(dens, extrap(b1, b2), 0.0, 0, fe);
        memcpy(fs + 9 * ((long)j * nx + nx - 1), fe, sizeof fe);
    }
}

/* iblbm/boundary/open_boundary.py:44-162: rho = dens on every face, normal velocity from the
 * known populations, tangential velocity extrapolated; corners f_eq(dens, U of the diagonal
 * inner node). */
void ref_bc_open(double *fs, const double *U, int nx, int ny, double dens) {
    init_w();
    const double ro = dens;
    for (int j = 1; j < ny - 1; ++j) {  /* left */
        const long n = (long)j * nx;
        double *f = fs + 9 * n;
        const double u0 = 1.0 - (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[7] + f[6])) / ro;
        const double u1 = extrap(U[2 * (n + 1) + 1], U[2 * (n + 2) + 1]);
        f[1] = f[3] + (2.0 / 3.0) * ro * u0;
        f[5] = f[7] - 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * ro * u0 + 0.5 * ro * u1;
        f[8] = f[6] + 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * ro * u0 - 0.5 * ro * u1;
    }
    for (int j = 1; j < ny - 1; ++j) {  /* right */
        const long n = (long)j * nx + nx - 1;
        zh_right_p(fs + 9 * n, ro, extrap(U[2 * (n - 1) + 1], U[2 * (n - 2) + 1]));
    }
    for (int i = 1; i < nx - 1; ++i) {  /* top */
        const long n = (long)(ny - 1) * nx + i;
        double *f = fs + 9 * n;
        const double u1 = -1.0 + (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6])) / ro;
        const double u0 = extrap(U[2 * (n - nx)], U[2 * (n - 2L * nx)]);
        f[4] = f[2] - (2.0 / 3.0) * ro * u1;
        f[7] = f[5] + 0.5 * (f[1] - f[3]) - (1.0 / 6.0) * ro * u1 - 0.5 * ro * u0;
        f[8] = f[6] - 0.5 * (f[1] - f[3]) - (1.0 / 6.0) * ro * u1 + 0.5 * ro * u0;
    }
    for (int i = 1; i < nx - 1; ++i) {  /* bottom */
        const long n = i;
        double *f = fs + 9 * n;
        const double u1 = 1.0 - (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8])) / ro;
        const double u0 = extrap(U[2 * (n + nx)], U[2 * (n + 2L * nx)]);
        f[2] = f[4] + (2.0 / 3.0) * ro * u1;
        f[5] = f[7] - 0.5 * (f[1] - f[3]) + (1.0 / 6.0) * ro * u1 + 0.5 * ro * u0;
        f[6] = f[8] + 0.5 * (f[1] - f[3]) + (1.0 / 6.0) * ro * u1 - 0.5 * ro * u0;
    }
    const long c[4] = {0, nx - 1, (long)nx * (ny - 1), (long)nx * ny - 1};
    const long d[4] = {nx + 1, 2L * nx - 2, (long)nx * (ny - 2) + 1, (long)nx * (ny - 1) - 2};
    for (int k = 0; k < 4; ++k) {
        double fe[9];
        feq_node(dens, U[2 * d[k]], U[2 * d[k] + 1], 0, fe);
        memcpy(fs + 9 * c[k], fe, sizeof fe);
    }
}

/* iblbm/boundary/tg_analytical.py:29-136: every face (left and right with their corners)
 * takes the analytic velocity; Ub holds it per boundary node in the order left (ny), right
 * (ny), top (nx - 2), bottom (nx - 2), corners (0,0), (nx-1,0), (0,ny-1), (nx-1,ny-1). */
void ref_bc_tg(double *fs, const double *Ub, int nx, int ny) {
    init_w();
    const double *UL = Ub, *UR = Ub + 2L * ny, *UT = Ub + 4L * ny, *UB = UT + 2L * (nx - 2);
    const double *UC = UB + 2L * (nx - 2);
    for (int j = 0; j < ny; ++j) zh_left_rho(fs + 9 * ((long)j * nx), UL[2 * j], UL[2 * j + 1]);
    for (int j = 0; j < ny; ++j) {
        double *f = fs + 9 * ((long)j * nx + nx - 1);
        const double u0 = UR[2 * j], u1 = UR[2 * j + 1];
        const double ro = (1.0 / (1.0 + u0)) * (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8]));
        f[3] = f[1] - (2.0 / 3.0) * ro * u0;
        f[7] = f[5] + 0.5 * (f[2] - f[4]) - (1.0 / 6.0) * ro * u0 - 0.5 * ro * u1;
        f[6] = f[8] - 0.5 * (f[2] - f[4]) - (1.0 / 6.0) * ro * u0 + 0.5 * ro * u1;
    }
    for (int i = 1; i < nx - 1; ++i) {
        double *f = fs + 9 * ((long)(ny - 1) * nx + i);
        const double u0 = UT[2 * (i - 1)], u1 = UT[2 * (i - 1) + 1];
        const double ro = (1.0 / (1.0 + u1)) * (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6]));
        f[4] = f[2] - (2.0 / 3.0) * ro * u1;
        f[7] = f[5] + 0.5 * (f[1] - f[3]) - (1.0 / 6.0) * ro * u1 - 0.5 * ro * u0;
        f[8] = f[6] - 0.5 * (f[1] - f[3]) - (1.0 / 6.0) * ro * u1 + 0.5 * ro * u0;
    }
    for (int i = 1; i < nx - 1; ++i) {
        double *f = fs + 9 * (long)i;
        const double u0 = UB[2 * (i - 1)], u1 = UB[2 * (i - 1) + 1];
        const double ro = (1.0 / (1.0 - u1)) * (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8]));
        f[2] = f[4] + (2.0 / 3.0) * ro * u1;
        f[5] = f[7] - 0.5 * (f[1] - f[3]) + (1.0 / 6.0) * ro * u1 + 0.5 * ro * u0;
        f[6] = f[8] + 0.5 * (f[1] - f[3]) + (1.0 / 6.0) * ro * u1 - 0.5 * ro * u0;
    }
    const long c[4] = {0, nx - 1, (long)nx * (ny - 1), (long)nx * ny - 1};
    for (int k = 0; k < 4; ++k) {
        double fe[9];
        feq_node(1.0, UC[2 * k], UC[2 * k + 1], 0, fe);
        memcpy(fs + 9 * c[k], fe, sizeof fe);
    }
}
