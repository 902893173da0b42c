// step_kernel.h -- the fused per-step kernel: bulk lattice tiles plus one immersed-boundary
// workgroup per instance, in one launch.
//
// Layout in HBM (per instance): populations are structure-of-arrays f[q][j][i] with a row
// pitch padded to a multiple of 32 elements, two buffers (source and destination swap every
// step).  The marker window is the box [bx0, bx0+bw) x [by0, by0+bh) that holds every
// clamped stencil node of every marker; box nodes are owned by the instance's IBM workgroup
// and skipped by the bulk tiles, so the two roles run concurrently and each node is read
// once and written once per step.
#pragma once
#include "lbm_math.h"

namespace iblbm {

constexpr int NT = 1024;  // threads per workgroup (16 wavefronts of 64)
constexpr int TILE_X = 64;
constexpr int TILE_Y = NT / TILE_X;

struct IbmScratch {
    // per box slot
    double *fs;     // 9 planes of box_cap: post-stream/boundary (and DFC-corrected) fstar
    double *ro, *ux0, *uy0;   // macroscopic rho, U from fstar
    double *uwx, *uwy, *vs;   // MDF working velocity and dt/(2 rho)
    double *fibx, *fiby;      // fib (DF) or fib_total (MDF); corrected U is written to ux0/uy0
    double *wt;               // DFC W_total
    int *cnt, *off, *cur;     // per-slot contribution count, offset (box_cap + 1), cursor
    // per contribution c = s*Lb + k (s = stencil offset index, k = marker)
    double *cw;    // wx*wy
    int *cslot;    // clamped target slot
    int *list;     // per-slot contribution ids sorted ascending
    // per marker
    double *Lux, *Luy, *R, *Lfx, *Lfy, *lam, *q;  // q: DFC force (Lb, 2)
    double *fi, *dfl;  // DFC interpolated f and delta f (Lb, 9)
};

struct Inst {
    // lattice
    int nx, ny, pitch, pad0;
    int64_t qs;  // elements per direction plane
    double *f[2];
    int bc, inc, tiles_x, tiles_y;
    CollParams cp;
    double in_u, dens;
    double feq_bc[9];  // dirichlet: f_eq(dens, (in_u, 0)); settling corners: f_eq(dens, 0)
    // immersed boundary
    int ibm, delta, Lb, lb_cap;
    int bx0, by0, bw, bh;
    int box_cap, n_iter, min_iter, fixed_body;
    int tables_valid, lambda_valid, pad1, pad2;
    double dx, dy, Larea, omega, tol;
    const double *Lx, *Ly, *ud;
    IbmScratch s;
    // outputs of the IBM workgroup
    double *stats;  // [iter_count, last_residual, diverged, converged]
    double *diag;   // [rho_box, force_x, force_y, torque_fib, torque_q]
    double cxc, cyc;  // torque centre in lattice units (cx/dx, cy/dy)
    double cxd, cyd;  // torque centre in domain units
};

struct LaunchTable {
    Inst *insts;
    const int *ibm_inst;     // instance id of each IBM workgroup
    const int *tile_prefix;  // tile_prefix[i]: first bulk tile of instance i (n_inst + 1 entries)
    double *rho_partial;     // one FP64 partial rho sum per workgroup (diagnostic steps)
    int n_inst, n_ibm;
    int parity;              // source buffer index of this step
    int diag;                // take diagnostic reductions this step
};

// ---------------------------------------------------------------- node-level helpers

__device__ inline const double *src_buf(const Inst &I, int parity) { return I.f[parity]; }

// Pull the 9 post-stream values of an interior node (all sources inside the grid).
__device__ inline void pull_interior(const double *__restrict__ A, const Inst &I, int i, int j,
                                     double *fs) {
#pragma unroll
    for (int q = 0; q < 9; ++q)
        fs[q] = A[q * I.qs + (int64_t)(j - EYi[q]) * I.pitch + (i - EXi[q])];
}

// Post-stream + boundary-closure populations of node (i, j): the value of state.fstar after
// iblbm/runtime/step.py:49-56 (streaming.py push + the registered closure).
__device__ inline void fstar_at(const double *__restrict__ A, const Inst &I, int i, int j,
                                double *fs) {
    const int nx = I.nx, ny = I.ny;
    const bool edge = (i == 0) | (j == 0) | (i == nx - 1) | (j == ny - 1);
    if (!edge) {
        pull_interior(A, I, i, j, fs);
        return;
    }
    if (I.bc == 0) {  // dirichlet_neumann.py:39-118
        if (i == nx - 1) {
            const int jj = j == 0 ? 1 : (j == ny - 1 ? ny - 2 : j);
            double fin[9];
            pull_interior(A, I, nx - 2, jj, fin);
            neumann_copy(fin, fs);
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) fs[q] = I.feq_bc[q];
        }
        return;
    }
    // settling_channel.py:42-142: corners f_eq(dens, 0); faces pull the known directions
    const bool corner = (i == 0 || i == nx - 1) && (j == 0 || j == ny - 1);
    if (corner) {
#pragma unroll
        for (int q = 0; q < 9; ++q) fs[q] = I.feq_bc[q];
        return;
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int si = i - EXi[q], sj = j - EYi[q];
        fs[q] = (si >= 0 && si < nx && sj >= 0 && sj < ny)
                    ? A[q * I.qs + (int64_t)sj * I.pitch + si]
                    : 0.0;
    }
    const int face = i == 0 ? 0 : (i == nx - 1 ? 1 : (j == ny - 1 ? 2 : 3));
    settling_face(face, fs);
}

__device__ inline void store9(double *__restrict__ B, const Inst &I, int i, int j,
                              const double *f) {
#pragma unroll
    for (int q = 0; q < 9; ++q) B[q * I.qs + (int64_t)j * I.pitch + i] = f[q];
}

// ---------------------------------------------------------------- block-wide helpers

__device__ inline double block_sum_fixed(double v, double *red) {
    // deterministic tree over the NT lanes (fixed order, independent of timing)
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) red[t] = red[t] + red[t + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__device__ inline double nanmax2(double a, double b) {
    return (a != a || a > b) ? a : b;  // np.max semantics: NaN propagates
}

__device__ inline double block_max(double v, double *red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) red[t] = nanmax2(red[t], red[t + s]);
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// exclusive scan of cnt[0..n) into off[0..n], off[n] = total (block-wide, chunked)
__device__ inline void block_scan(const int *cnt, int *off, int n, int *red) {
    const int t = threadIdx.x;
    const int chunk = (n + NT - 1) / NT;
    const int b = t * chunk, e = min(n, b + chunk);
    int s = 0;
    for (int k = b; k < e; ++k) s += cnt[k];
    red[t] = s;
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int k = 0; k < NT; ++k) {
            const int v = red[k];
            red[k] = acc;
            acc += v;
        }
        off[n] = acc;
    }
    __syncthreads();
    int acc = red[t];
    for (int k = b; k < e; ++k) {
        off[k] = acc;
        acc += cnt[k];
    }
    __syncthreads();
}

}  // namespace iblbm
