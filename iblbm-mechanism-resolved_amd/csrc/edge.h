// edge.h -- the boundary closures that read more than the streamed populations: the
// previous step's velocity (velocity_inlet, open_boundary) or a time-dependent analytic
// velocity (tg_analytical).
//
// The bulk kernels leave the outermost ring of such an instance alone; k_edge runs it, one
// thread per node of the three outermost rings:
//  * every ring node pulls its populations from the previous step's buffer (a source outside
//    the grid reads 0: each closure below overwrites every such population), forms rho and u
//    and keeps u for the next step's closures (the "previous velocity" of the faces'
//    extrapolations is the state's velocity after the previous step's macroscopic pass, which
//    for these nodes, outside every marker window, is exactly this u);
//  * a node of the outermost ring first applies its face's closure (Zou-He, Zou & He 1997 sec.
//    3, as iblbm/boundary/{velocity_inlet,open_boundary,tg_analytical}.py write it), then the
//    equilibrium and the unforced collision of its instance, and stores the new populations.
// Marker windows of these instances stay three nodes away from every wall (checked when the
// markers are set), so no window node is a ring node.
#pragma once
#include "../../include/iblbm_hip.h"
#include "step_kernel.h"

namespace iblbm {

// Slot of node (i, j) among the nodes within three of a wall: the three bottom rows, the
// three top rows, then the three left and the three right columns of the rows between.
__host__ __device__ inline int ring_slot(int i, int j, int nx, int ny) {
    if (j < 3) return j * nx + i;
    if (j >= ny - 3) return 3 * nx + (j - (ny - 3)) * nx + i;
    const int base = 6 * nx, h = ny - 6;
    if (i < 3) return base + (j - 3) * 3 + i;
    if (i >= nx - 3) return base + 3 * h + (j - 3) * 3 + (i - (nx - 3));
    return -1;
}
__host__ __device__ inline int ring_count(int nx, int ny) { return 6 * nx + 6 * (ny - 6); }
__host__ __device__ inline void ring_node(int s, int nx, int ny, int &i, int &j) {
    if (s < 3 * nx) {
        j = s / nx;
        i = s - j * nx;
        return;
    }
    s -= 3 * nx;
    if (s < 3 * nx) {
        j = ny - 3 + s / nx;
        i = s - (s / nx) * nx;
        return;
    }
    s -= 3 * nx;
    const int h = ny - 6;
    if (s < 3 * h) {
        j = 3 + s / 3;
        i = s % 3;
        return;
    }
    s -= 3 * h;
    j = 3 + s / 3;
    i = nx - 3 + s % 3;
}

__device__ inline double extrap2(double a1, double a2) { return (4.0 * a1 - a2) / 3.0; }

// Left face with velocity (u0, u1): rho from the known populations over 1 - u0
// (velocity_inlet.py:69-86, tg_analytical.py:64-80).
__device__ inline void zh_left(double *f, double u0, double u1) {
    const double ro = (1.0 / (1.0 - u0)) * (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7]));
    f[1] = f[3] + (2.0 / 3.0) * ro * u0;
    f[5] = f[7] - 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * ro * u0 + 0.5 * ro * u1;
    f[8] = f[6] + 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * ro * u0 - 0.5 * ro * u1;
}
// Right face at pressure ro: u0 from the known populations (velocity_inlet.py:88-106,
// open_boundary.py:82-99).
__device__ inline void zh_right_p(double *f, double ro, double u1) {
    const double u0 = -1.0 + (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])) / ro;
    f[3] = f[1] - (2.0 / 3.0) * ro * u0;
    f[7] = f[5] + 0.5 * (f[2] - f[4]) - (1.0 / 6.0) * ro * u0 - 0.5 * ro * u1;
    f[6] = f[8] - 0.5 * (f[2] - f[4]) - (1.0 / 6.0) * ro * u0 + 0.5 * ro * u1;
}

// The previous step's velocity component c at node (i, j) of the rings.
__device__ inline double uprev(const Inst &I, const double *side, int i, int j, int c) {
    return side[2 * ring_slot(i, j, I.nx, I.ny) + c];
}

// The closure of boundary node (i, j) on its pulled populations f.
__device__ inline void edge_closure(const Inst &I, const double *side, const double *tgv,
                                    int i, int j, double *f) {
    const int nx = I.nx, ny = I.ny;
    const bool left = i == 0, right = i == nx - 1, bottom = j == 0, top = j == ny - 1;
    const bool corner = (left || right) && (bottom || top);
    const double dens = I.dens;
    if (I.bc == IBLBM_BC_VELOCITY_INLET) {  // velocity_inlet.py:45-157
        if (corner) {
            double u0 = I.in_u;
            if (right) {  // outlet corners: u0 from the bottom / top face values of this step
                const int jj = bottom ? 1 : ny - 2, jk = bottom ? 2 : ny - 3;
                const double b1 = extrap2(uprev(I, side, nx - 2, jj, 0), uprev(I, side, nx - 2, jk, 0));
                const double b2 = extrap2(uprev(I, side, nx - 3, jj, 0), uprev(I, side, nx - 3, jk, 0));
                u0 = extrap2(b1, b2);
            }
            feq9(dens, u0, 0.0, false, f);
        } else if (left) {
            zh_left(f, I.in_u, 0.0);
        } else if (right) {
            zh_right_p(f, dens, extrap2(uprev(I, side, nx - 2, j, 1), uprev(I, side, nx - 3, j, 1)));
        } else if (top) {
            const double u0 = extrap2(uprev(I, side, i, ny - 2, 0), uprev(I, side, i, ny - 3, 0));
            const double u1 = 0.0;
            const double ro =
                (1.0 / (1.0 + u1)) * (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6]));
            f[4] = f[2] - (2.0 / 3.0) * ro * u1;
            f[7] = f[5] + 0.5 * (f[1] - f[3]) - 0.5 * ro * u0;
            f[8] = f[6] - 0.5 * (f[1] - f[3]) + 0.5 * ro * u0;
        } else {
            const double u0 = extrap2(uprev(I, side, i, 1, 0), uprev(I, side, i, 2, 0));
            const double u1 = 0.0;
            const double ro =
                (1.0 / (1.0 - u1)) * (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8]));
            f[2] = f[4] + (2.0 / 3.0) * ro * u1;
            f[5] = f[7] - 0.5 * (f[1] - f[3]) + 0.5 * ro * u0;
            f[6] = f[8] + 0.5 * (f[1] - f[3]) - 0.5 * ro * u0;
        }
        return;
    }
    if (I.bc == IBLBM_BC_OPEN) {  // open_boundary.py:44-162
        const double ro = dens;
        if (corner) {  // f_eq(dens, u of the diagonal inner node)
            const int di = left ? 1 : nx - 2, dj = bottom ? 1 : ny - 2;
            feq9(dens, uprev(I, side, di, dj, 0), uprev(I, side, di, dj, 1), false, f);
        } else if (left) {
            const double u0 = 1.0 - (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[7] + f[6])) / ro;
            const double u1 = extrap2(uprev(I, side, 1, j, 1), uprev(I, side, 2, j, 1));
            f[1] = f[3] + (2.0 / 3.0) * ro * u0;
            f[5] = f[7] - 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * ro * u0 + 0.5 * ro * u1;
            f[8] = f[6] + 0.5 * (f[2] - f[4]) + (1.0 / 6.0) * ro * u0 - 0.5 * ro * u1;
        } else if (right) {
            zh_right_p(f, ro, extrap2(uprev(I, side, nx - 2, j, 1), uprev(I, side, nx - 3, j, 1)));
        } else if (top) {
            const double u1 = -1.0 + (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6])) / ro;
            const double u0 = extrap2(uprev(I, side, i, ny - 2, 0), uprev(I, side, i, ny - 3, 0));
            f[4] = f[2] - (2.0 / 3.0) * ro * u1;
            f[7] = f[5] + 0.5 * (f[1] - f[3]) - (1.0 / 6.0) * ro * u1 - 0.5 * ro * u0;
            f[8] = f[6] - 0.5 * (f[1] - f[3]) - (1.0 / 6.0) * ro * u1 + 0.5 * ro * u0;
        } else {
            const double u1 = 1.0 - (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8])) / ro;
            const double u0 = extrap2(uprev(I, side, i, 1, 0), uprev(I, side, i, 2, 0));
            f[2] = f[4] + (2.0 / 3.0) * ro * u1;
            f[5] = f[7] - 0.5 * (f[1] - f[3]) + (1.0 / 6.0) * ro * u1 + 0.5 * ro * u0;
            f[6] = f[8] + 0.5 * (f[1] - f[3]) + (1.0 / 6.0) * ro * u1 - 0.5 * ro * u0;
        }
        return;
    }
    // tg_analytical.py:29-136: the analytic velocity tgv = (u0, u1) of this node
    const double u0 = tgv[0], u1 = tgv[1];
    if (corner) {
        feq9(1.0, u0, u1, false, f);
    } else if (left) {
        zh_left(f, u0, u1);
    } else if (right) {
        const double ro = (1.0 / (1.0 + u0)) * (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8]));
        f[3] = f[1] - (2.0 / 3.0) * ro * u0;
        f[7] = f[5] + 0.5 * (f[2] - f[4]) - (1.0 / 6.0) * ro * u0 - 0.5 * ro * u1;
        f[6] = f[8] - 0.5 * (f[2] - f[4]) - (1.0 / 6.0) * ro * u0 + 0.5 * ro * u1;
    } else if (top) {
        const double ro = (1.0 / (1.0 + u1)) * (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6]));
        f[4] = f[2] - (2.0 / 3.0) * ro * u1;
        f[7] = f[5] + 0.5 * (f[1] - f[3]) - (1.0 / 6.0) * ro * u1 - 0.5 * ro * u0;
        f[8] = f[6] - 0.5 * (f[1] - f[3]) - (1.0 / 6.0) * ro * u1 + 0.5 * ro * u0;
    } else {
        const double ro = (1.0 / (1.0 - u1)) * (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8]));
        f[2] = f[4] + (2.0 / 3.0) * ro * u1;
        f[5] = f[7] - 0.5 * (f[1] - f[3]) + (1.0 / 6.0) * ro * u1 + 0.5 * ro * u0;
        f[6] = f[8] + 0.5 * (f[1] - f[3]) + (1.0 / 6.0) * ro * u1 - 0.5 * ro * u0;
    }
}

// Index of boundary node (i, j) in the analytic table order: left (ny), right (ny), top
// (nx - 2), bottom (nx - 2), corners (0,0), (nx-1,0), (0,ny-1), (nx-1,ny-1).
__device__ inline int tg_index(int i, int j, int nx, int ny) {
    const bool left = i == 0, right = i == nx - 1, bottom = j == 0, top = j == ny - 1;
    if ((left || right) && (bottom || top))
        return 2 * ny + 2 * (nx - 2) + (right ? 1 : 0) + (top ? 2 : 0);
    if (left) return j;
    if (right) return ny + j;
    if (top) return 2 * ny + (i - 1);
    return 2 * ny + (nx - 2) + (i - 1);
}

}  // namespace iblbm
