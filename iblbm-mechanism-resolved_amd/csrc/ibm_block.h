// ibm_block.h -- the immersed-boundary workgroup of the fused step (one per instance).
//
// It owns the window box: it streams and closes the box nodes itself, computes rho and U,
// runs DF (iblbm/ibm/df.py:94-154), MDF (iblbm/ibm/mdf.py:175-229) or DFC
// (iblbm/ibm/dfc.py:256-342) on them, applies the velocity correction
// (iblbm/ibm/dispatch.py:18-33), and finishes the collision of the box nodes.
//
// Spreading is a gather: every box slot owns the list of (stencil offset s, marker k)
// contributions that land on it, sorted by c = s*Lb + k, which is exactly the order in
// which np.add.at applies them (offset loop outer, markers inner), so the FP64 sums carry
// the reference's bits and the result does not depend on thread timing (no float atomics).
#pragma once
#include "step_kernel.h"

namespace iblbm {

__device__ inline int64_t floor_idx(double v) { return (int64_t)floor(v); }

// Contribution tables: weights, target slots and per-slot ordered lists.
__device__ void ibm_build_tables(Inst &I, int *redi) {
    const int t = threadIdx.x;
    const int a = I.delta == 1 ? 1 : 2, side = 2 * a + 1, K = side * side, Lb = I.Lb;
    const int nc = K * Lb, nbox = I.bw * I.bh;
    IbmScratch &S = I.s;
    for (int c = t; c < nc; c += NT) {
        const int s = c / Lb, k = c - s * Lb;
        const int di = s / side - a, dj = s % side - a;
        const double Lx = I.Lx[k], Ly = I.Ly[k];
        const int64_t ix0 = floor_idx(Lx / I.dx + 0.5), iy0 = floor_idx(Ly / I.dy + 0.5);
        const int64_t ei = ix0 + di, ej = iy0 + dj;
        const double wx = delta_phi(I.delta, (Lx - (double)ei * I.dx) / I.dx);
        const double wy = delta_phi(I.delta, (Ly - (double)ej * I.dy) / I.dy);
        S.cw[c] = wx * wy;
        const int64_t eic = ei < 0 ? 0 : (ei > I.nx - 1 ? I.nx - 1 : ei);
        const int64_t ejc = ej < 0 ? 0 : (ej > I.ny - 1 ? I.ny - 1 : ej);
        // The host sizes the box to hold every clamped stencil node.  A node outside it
        // would be a host bug: it is routed to slot 0 with a NaN weight so the step output
        // turns NaN (loudly) instead of touching memory outside the scratch arrays.
        if (eic < I.bx0 || eic >= I.bx0 + I.bw || ejc < I.by0 || ejc >= I.by0 + I.bh) {
            S.cslot[c] = 0;
            S.cw[c] = __builtin_nan("");
        } else {
            S.cslot[c] = (int)((ejc - I.by0) * I.bw + (eic - I.bx0));
        }
    }
    for (int k = t; k < nbox; k += NT) {
        S.cnt[k] = 0;
        S.cur[k] = 0;
    }
    __syncthreads();
    for (int c = t; c < nc; c += NT) {
        atomicAdd(&S.cnt[S.cslot[c]], 1);
    }
    __syncthreads();
    block_scan(S.cnt, S.off, nbox, redi);
    for (int c = t; c < nc; c += NT) {
        const int sl = S.cslot[c];
        const int pos = atomicAdd(&S.cur[sl], 1);
        S.list[S.off[sl] + pos] = c;
    }
    __threadfence_block();
    __syncthreads();
    for (int sl = t; sl < nbox; sl += NT) {  // insertion sort each slot's list by c
        int *L = S.list + S.off[sl];
        const int n = S.cnt[sl];
        for (int x = 1; x < n; ++x) {
            const int v = L[x];
            int y = x - 1;
            while (y >= 0 && L[y] > v) {
                L[y + 1] = L[y];
                --y;
            }
            L[y + 1] = v;
        }
    }
    __threadfence_block();
    __syncthreads();
}

// DFC lambda (iblbm/ibm/dfc.py:157-185), rho_f = 1.
__device__ void ibm_dfc_lambda(Inst &I) {
    const int t = threadIdx.x;
    const int a = I.delta == 1 ? 1 : 2, K = (2 * a + 1) * (2 * a + 1), Lb = I.Lb;
    const int nbox = I.bw * I.bh;
    IbmScratch &S = I.s;
    for (int sl = t; sl < nbox; sl += NT) {
        double w = 0.0;
        const int *L = S.list + S.off[sl];
        for (int x = 0, n = S.cnt[sl]; x < n; ++x) w += S.cw[L[x]];
        S.wt[sl] = w;
    }
    __threadfence_block();
    __syncthreads();
    for (int k = t; k < Lb; k += NT) {
        double ws = 0.0;
        for (int s = 0; s < K; ++s) {
            const int c = s * Lb + k;
            const int sl = S.cslot[c];
            ws += S.wt[sl] * S.cw[c];
        }
        S.lam[k] = 1.0 / (2.0 * 1.0 * I.Larea * ws);
    }
    __threadfence_block();
    __syncthreads();
}

// Interpolate rho and U (from ux/uy planes) to the markers and form the DF force.
// compressible: F = 2 R (U_d - U~)/dt; incompressible (df.py:256-259): F = 2 (U_d - U~)/dt.
// Returns the block-wide L-inf slip residual max_k |U_d - U~_k|.
__device__ double ibm_interp_force(Inst &I, const double *ux, const double *uy, bool inc,
                                   double *red) {
    const int t = threadIdx.x;
    const int a = I.delta == 1 ? 1 : 2, K = (2 * a + 1) * (2 * a + 1), Lb = I.Lb;
    IbmScratch &S = I.s;
    double rmax = -INFINITY;
    for (int k = t; k < Lb; k += NT) {
        double R = 0.0, lx = 0.0, ly = 0.0;
        for (int s = 0; s < K; ++s) {
            const int c = s * Lb + k;
            const int sl = S.cslot[c];
            const double w = S.cw[c];
            R += S.ro[sl] * w;
            lx += ux[sl] * w;
            ly += uy[sl] * w;
        }
        const double udx = I.ud[2 * k], udy = I.ud[2 * k + 1];
        S.R[k] = R;
        S.Lux[k] = lx;
        S.Luy[k] = ly;
        if (inc) {
            S.Lfx[k] = 2.0 * (udx - lx) / I.cp.dt;
            S.Lfy[k] = 2.0 * (udy - ly) / I.cp.dt;
        } else {
            S.Lfx[k] = 2.0 * R * (udx - lx) / I.cp.dt;
            S.Lfy[k] = 2.0 * R * (udy - ly) / I.cp.dt;
        }
        const double vx = udx - lx, vy = udy - ly;
        rmax = nanmax2(sqrt(vx * vx + vy * vy), rmax);
    }
    __threadfence_block();
    return block_max(rmax, red);  // includes the barrier that publishes Lf
}

// Gather the spread force of one slot in np.add.at order: sum_c Lf[k] * (w_c * Larea).
__device__ inline void ibm_gather_force(const Inst &I, int sl, double &ex, double &ey) {
    const IbmScratch &S = I.s;
    const int *L = S.list + S.off[sl];
    const int Lb = I.Lb;
    double sx = 0.0, sy = 0.0;
    for (int x = 0, n = S.cnt[sl]; x < n; ++x) {
        const int c = L[x];
        const int k = c % Lb;
        const double w = S.cw[c] * I.Larea;
        sx += S.Lfx[k] * w;
        sy += S.Lfy[k] * w;
    }
    ex = sx;
    ey = sy;
}

__device__ void ibm_role(Inst &I, int parity, int diag, double *red, int *redi) {
    const int t = threadIdx.x;
    const double *__restrict__ A = I.f[parity];
    double *__restrict__ B = I.f[parity ^ 1];
    IbmScratch &S = I.s;
    const int nbox = I.bw * I.bh;
    const bool inc = I.inc != 0;

    if (!I.tables_valid) {
        ibm_build_tables(I, redi);
        if (t == 0) I.tables_valid = I.fixed_body;
    }
    if (I.ibm == 3 && !I.lambda_valid) {
        ibm_dfc_lambda(I);
        if (t == 0) I.lambda_valid = I.fixed_body;
    }

    // P1: stream + close the box nodes, macroscopic moments
    for (int sl = t; sl < nbox; sl += NT) {
        const int i = I.bx0 + sl % I.bw, j = I.by0 + sl / I.bw;
        double fs[9];
        fstar_at(A, I, i, j, fs);
#pragma unroll
        for (int q = 0; q < 9; ++q) S.fs[q * I.box_cap + sl] = fs[q];
        double r, u0, u1;
        macro9(fs, inc, r, u0, u1);
        S.ro[sl] = r;
        S.ux0[sl] = u0;
        S.uy0[sl] = u1;
        S.fibx[sl] = 0.0;
        S.fiby[sl] = 0.0;
    }
    __threadfence_block();
    __syncthreads();

    if (I.ibm == 1) {  // DF: one interpolation + spread, iblbm/ibm/strategies/df.py:27-49
        ibm_interp_force(I, S.ux0, S.uy0, inc, red);
        for (int sl = t; sl < nbox; sl += NT) {
            if (S.cnt[sl] == 0) continue;
            double ex, ey;
            ibm_gather_force(I, sl, ex, ey);
            S.fibx[sl] = ex;
            S.fiby[sl] = ey;
        }
    } else if (I.ibm == 2) {  // MDF, iblbm/ibm/mdf.py:113-229 (compressible sub-steps)
        for (int sl = t; sl < nbox; sl += NT) {
            S.uwx[sl] = S.ux0[sl];
            S.uwy[sl] = S.uy0[sl];
            S.vs[sl] = I.cp.dt / (2.0 * S.ro[sl]);
        }
        __threadfence_block();
        __syncthreads();
        const int n_iter = I.n_iter;
        const int min_iter = max(1, min(I.min_iter, n_iter));
        double prev = INFINITY, last = INFINITY;
        int iter_count = 0, diverged = 0, converged = 0;
        for (int it = 0; it < n_iter; ++it) {
            const double res = ibm_interp_force(I, S.uwx, S.uwy, false, red);
            last = res;
            if (it > 0 && res > prev * 1.5) {
                diverged = 1;
                iter_count = it + 1;
                break;
            }
            for (int sl = t; sl < nbox; sl += NT) {
                if (S.cnt[sl] == 0) continue;
                double ex, ey;
                ibm_gather_force(I, sl, ex, ey);
                const double fx = ex * I.omega, fy = ey * I.omega;
                S.fibx[sl] = S.fibx[sl] + fx;
                S.fiby[sl] = S.fiby[sl] + fy;
                S.uwx[sl] = S.uwx[sl] + fx * S.vs[sl];
                S.uwy[sl] = S.uwy[sl] + fy * S.vs[sl];
            }
            __threadfence_block();
            __syncthreads();
            prev = res;
            iter_count = it + 1;
            if (it + 1 >= min_iter && res < I.tol) {
                converged = 1;
                break;
            }
        }
        if (t == 0) {
            I.stats[0] = iter_count;
            I.stats[1] = last;
            I.stats[2] = diverged;
            I.stats[3] = converged;
        }
    } else if (I.ibm == 3) {  // DFC, iblbm/ibm/dfc.py:319-341 (correction_scale = 1)
        const int a = I.delta == 1 ? 1 : 2, K = (2 * a + 1) * (2 * a + 1), Lb = I.Lb;
        const double cs2 = 1.0 / 3.0;
        for (int k = t; k < Lb; k += NT) {
            double fi[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) fi[q] = 0.0;
            for (int s = 0; s < K; ++s) {
                const int c = s * Lb + k;
                const int sl = S.cslot[c];
                const double w = S.cw[c];
#pragma unroll
                for (int q = 0; q < 9; ++q) fi[q] += S.fs[q * I.box_cap + sl] * w;
            }
            const double udx = I.ud[2 * k], udy = I.ud[2 * k + 1], lam = S.lam[k];
            double qx = 0.0, qy = 0.0;
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const double edu = udx * EX(q) + udy * EY(q);
                const double fbb = fi[OPP[q]] + 2.0 * W(q) * 1.0 * edu / cs2;
                const double dev = fbb - fi[q];
                S.dfl[9 * k + q] = 1.0 * lam * dev;
                const double sd = -(lam * dev * I.Larea);
                qx += sd * EX(q);
                qy += sd * EY(q);
            }
            S.q[2 * k] = 1.0 * qx;
            S.q[2 * k + 1] = 1.0 * qy;
        }
        __threadfence_block();
        __syncthreads();
        for (int sl = t; sl < nbox; sl += NT) {
            double fs[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) fs[q] = S.fs[q * I.box_cap + sl];
            if (S.cnt[sl] != 0) {
                double e[9];
#pragma unroll
                for (int q = 0; q < 9; ++q) e[q] = 0.0;
                const int *L = S.list + S.off[sl];
                for (int x = 0, n = S.cnt[sl]; x < n; ++x) {
                    const int c = L[x];
                    const int k = c % Lb;
                    const double w = S.cw[c] * I.Larea;
#pragma unroll
                    for (int q = 0; q < 9; ++q) e[q] += S.dfl[9 * k + q] * w;
                }
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    fs[q] = fs[q] + e[q];
                    S.fs[q * I.box_cap + sl] = fs[q];
                }
            }
            double r, u0, u1;
            macro9(fs, inc, r, u0, u1);
            S.ro[sl] = r;
            S.ux0[sl] = u0;
            S.uy0[sl] = u1;
        }
    }
    __threadfence_block();
    __syncthreads();

    // P3: velocity correction (DF/MDF), equilibrium, collision, store
    const bool correct = I.ibm == 1 || I.ibm == 2;
    double rho_acc = 0.0, fxs = 0.0, fys = 0.0, trq = 0.0;
    for (int sl = t; sl < nbox; sl += NT) {
        const int i = I.bx0 + sl % I.bw, j = I.by0 + sl / I.bw;
        double fs[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) fs[q] = S.fs[q * I.box_cap + sl];
        const double r = S.ro[sl], fx = S.fibx[sl], fy = S.fiby[sl];
        double u0 = S.ux0[sl], u1 = S.uy0[sl];
        if (correct) {
            if (inc) {
                u0 = u0 + 0.5 * fx * I.cp.dt;
                u1 = u1 + 0.5 * fy * I.cp.dt;
            } else {
                u0 = u0 + fx * I.cp.dt / (2.0 * r);
                u1 = u1 + fy * I.cp.dt / (2.0 * r);
            }
            S.ux0[sl] = u0;
            S.uy0[sl] = u1;
        }
        double fe[9], out[9];
        feq9(r, u0, u1, inc, fe);
        collide<true>(I.cp, fs, fe, u0, u1, fx, fy, out);
        store9(B, I, i, j, out);
        if (diag) {
            rho_acc += r;
            fxs += fx;
            fys += fy;
            trq += ((double)i - I.cxc) * fy - ((double)j - I.cyc) * fx;
        }
    }
    if (diag) {
        const double rs = block_sum_fixed(rho_acc, red);
        double f0 = block_sum_fixed(fxs, red), f1 = block_sum_fixed(fys, red);
        const double tq = block_sum_fixed(trq, red);
        double tqq = 0.0;
        if (I.ibm == 3) {
            double qx = 0.0, qy = 0.0, tk = 0.0;
            for (int k = t; k < I.Lb; k += NT) {
                qx += S.q[2 * k];
                qy += S.q[2 * k + 1];
                tk += (I.Lx[k] - I.cxd) / I.dx * S.q[2 * k + 1] -
                      (I.Ly[k] - I.cyd) / I.dy * S.q[2 * k];
            }
            f0 = block_sum_fixed(qx, red);
            f1 = block_sum_fixed(qy, red);
            tqq = block_sum_fixed(tk, red);
        }
        if (t == 0) {
            I.diag[0] = rs;
            I.diag[1] = f0;
            I.diag[2] = f1;
            I.diag[3] = tq;
            I.diag[4] = tqq;
        }
    }
}

}  // namespace iblbm
