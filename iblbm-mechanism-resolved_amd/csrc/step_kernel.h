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
#include "dev32.h"
#include "lbm_math.h"

namespace iblbm {

constexpr int NT_BULK = 256;  // bulk tile workgroup: 4 wavefronts, a 64 x 4 node tile
constexpr int TILE_X = 64;
constexpr int TILE_Y = NT_BULK / TILE_X;
constexpr int NTI = 512;
constexpr int NTF = 1024;              // threads of a fused-step workgroup (16 wavefronts)
constexpr int FAST_T = NTF - 64;       // touched-node capacity of the fused window (1 per thread)
constexpr int FAST_G = 8;              // gather entries per touched node on that path
constexpr int FAST_LB = 128;           // markers on that path (up to 2 marker wavefronts)
constexpr int FAST_K = 25;             // stencil points per marker (Peskin 4-point: 5 x 5)

constexpr int IBM_MAXP = 8;  // particles of one DF / MDF window

struct IbmScratch {
    // per box slot (box_cap)
    double *fs;               // 9 planes: post-stream/boundary (and DFC-corrected) fstar
    double *ro, *ux0, *uy0;   // macroscopic rho, U (velocity-corrected U after the step)
    double *fibx, *fiby;      // fib of the step (0 on untouched nodes)
    int *cnt, *cur;           // contributions per slot, fill cursor
    int *tindex;              // slot -> touched index or -1 (box_cap + 1 entries)
    // per touched node (<= box_cap)
    int *tslot;               // touched index -> box slot
    int *toff;                // gather-list offsets (box_cap + 1)
    double *uwx, *uwy, *vs, *tfibx, *tfiby, *rot;  // global fallback of the LDS arrays
    double *wt;               // DFC W_total
    // per contribution, marker-major c_m = k*K + s (K = stencil points, lb_cap*25)
    double *cw;               // wx*wy
    int *cslot, *cti;         // box slot, touched index
    // gather lists (same length), np.add.at order per touched node
    int *gc, *gk;             // order key s*Lb + k, marker k
    double *gw;               // wx*wy*Larea
    // per marker (lb_cap)
    double *Lux, *Luy, *R, *Lfx, *Lfy, *lam, *q;  // q: DFC force (Lb, 2)
    // several particles' DF / MDF windows: per particle window fib [p][x|y][box_cap]
    // (coupled MDF) and the marker forces of the last pass (Lb, 2)
    double *pfb, *mf;
    double *dfl;              // DFC delta f (Lb, 9)
    // fused window path: gather lists padded to FAST_G entries, entry-major [e * FAST_T + ti]
    // (padding: marker Lb, weight 0), and the marker stencils stencil-major [s * Lb + k]
    int *pk, *mti;
    double *pw, *mw;
    // learned MDF force predictor (allocated when enabled): FP32 parameters, normalisation,
    // BF16 weight blocks, training samples (features, target), previous converged marker
    // forces, counters [samples written, previous force valid] and statistics [applied
    // corrections, steps, first residual, fallbacks]
    float *mlp_p, *mlp_norm, *data;
    unsigned short *mlp_wb;
    double *fprev, *pstat;
    long long *pctr;
    // the generic window's inference tiles: BF16 feature rows (16 per marker) and outputs
    // (MLP_OUT per marker), lb_cap rounded up to whole 16-marker tiles
    unsigned short *mlpx;
    float *mlpy;
};

struct Inst {
    // lattice
    int nx, ny, pitch, pad0;
    int64_t qs;  // elements per direction plane
    double *f[2];  // float* when prec == 1 (FP32 deviations g = f - w)
    int bc, inc, tiles_x, tiles_y;
    int prec, flat;  // flat: bulk rows are 64-node chunks of the row-major node order
    CollParams cp;
    CollParams32 cp32;
    float gbc[9];      // feq_bc - w in FP32 (deviation-form boundary constants)
    double in_u, dens;
    double feq_bc[9];  // dirichlet: f_eq(dens, (in_u, 0)); settling corners: f_eq(dens, 0)
    // immersed boundary
    int ibm, delta, Lb, lb_cap;
    int bx0, by0, bw, bh;
    int box_cap, n_iter, min_iter, fixed_body;
    int nT, fast;  // touched nodes; 1 when the register-resident window path applies
    int pmode, data_cap;  // predictor mode (0 off, 1 collect, 2 warm start, 3 learned), capacity
    double dx, dy, Larea, omega, tol, dfc_scale;
    const double *Lx, *Ly, *ud;
    // particles of a DF / MDF window (markers concatenated in particle order); with more
    // than one, contributions are spread per particle and the MDF iteration is coupled
    int np_part;
    int df_multi;  // a DF window of a particle list (any length): per-particle closures, marker
                   // forces kept (multi_particle.py:347-421)
    int poff[IBM_MAXP + 1];     // first marker of each particle
    double plarea[IBM_MAXP];    // arc length per marker of each particle
    IbmScratch s;
    // outputs of the IBM workgroup
    double *stats;  // [iter_count, last_residual, diverged, converged, guard_trips, omega]
    double *diag;   // [rho_box, force_x, force_y, torque_fib, torque_q]
    double cxc, cyc;  // torque centre in lattice units (cx/dx, cy/dy)
    double cxd, cyd;  // torque centre in domain units
    // closures run by k_edge (bc >= 2, edge.h): the three outermost rings' velocities of the
    // previous and of this step (2 x nring x 2), the outer ring's post-closure populations
    // of the last step (nring x 9, for downloads), the analytic-velocity coefficients
    // (A, B per boundary node: u = (A d(t), B d(t))) and decay table d(t), t < tg_nt
    double *side, *edge_fs;
    const double *tgc, *tgd;
    int nring, tg_nt;
};

// Arguments of the persistent step launch: the grid barrier's arrival counter and error flag,
// the number of steps, the diagnostic interval, the step counter and parity it starts from.
struct PersistArgs {
    unsigned *bar;
    int *err;
    int nsteps, diag_every;
    long long step0;
    int parity0, pad;
};

struct LaunchTable {
    Inst *insts;
    const int *ibm_inst;     // instance ids of the window workgroups: fast-path ones first
    const int *tile_prefix;  // tile_prefix[i]: first bulk tile of instance i (n_inst + 1 entries)
    double *rho_partial;     // one FP64 partial rho sum per bulk workgroup (diagnostic steps)
    int n_inst, n_ibm, n_fast, n_tiles;
    int parity;              // source buffer index of this step
    int diag;                // take diagnostic reductions this step
    int nrows;               // rows a bulk wavefront keeps in flight: 1, 2, 4 (FP32), or
                             // 3: one row with the next row's loads issued ahead
    int blocked;             // fused launch: each bulk wavefront takes one block of rows
    int bulk_delay;          // fused launch: bulk start delay in 10 ns ticks
    // boundary kernel (k_edge): per workgroup (instance, first ring slot), and its rho
    // partial sums; the step number t of this step is stepc[parity] (the kernel writes t + 1
    // to the other slot for the next step)
    const int2 *edge_blocks;
    double *edge_rho;
    long long *stepc;
    int n_edge_blocks;
};

// ---------------------------------------------------------------- storage traits
// FP64 storage holds f itself; FP32 storage holds g = f - w.  ld() returns f in FP64 and st()
// stores an FP64 f, so the immersed-boundary workgroup computes in FP64 for both.
template <typename S>
struct Store;
template <>
struct Store<double> {
    __device__ static double ld(const double *__restrict__ A, int64_t idx, int) { return A[idx]; }
    __device__ static double st(double f, int) { return f; }
};
template <>
struct Store<float> {
    __device__ static double ld(const float *__restrict__ A, int64_t idx, int q) {
        return W(q) + (double)A[idx];
    }
    __device__ static float st(double f, int q) { return (float)(f - W(q)); }
};

template <typename S>
__device__ inline S *buf(const Inst &I, int k) {
    return reinterpret_cast<S *>(I.f[k]);
}

// ---------------------------------------------------------------- node-level helpers

// Pull the 9 post-stream values of an interior node (all sources inside the grid).
template <typename S>
__device__ inline void pull_interior(const S *__restrict__ A, const Inst &I, int i, int j,
                                     double *fs) {
#pragma unroll
    for (int q = 0; q < 9; ++q)
        fs[q] = Store<S>::ld(A, q * I.qs + (int64_t)(j - EYi[q]) * I.pitch + (i - EXi[q]), q);
}

// Post-stream + boundary-closure populations of node (i, j): the value of state.fstar after
// iblbm/runtime/step.py:49-56 (streaming.py push + the registered closure).
template <typename S>
__device__ inline void fstar_at(const S *__restrict__ A, const Inst &I, int i, int j,
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
                    ? Store<S>::ld(A, q * I.qs + (int64_t)sj * I.pitch + si, q)
                    : 0.0;
    }
    const int face = i == 0 ? 0 : (i == nx - 1 ? 1 : (j == ny - 1 ? 2 : 3));
    settling_face(face, fs);
}

template <typename S>
__device__ inline void store9(S *__restrict__ B, const Inst &I, int i, int j, const double *f) {
#pragma unroll
    for (int q = 0; q < 9; ++q) B[q * I.qs + (int64_t)j * I.pitch + i] = Store<S>::st(f[q], q);
}

// FP32-deviation post-stream + closure (same closures as fstar_at, on g = f - w).
__device__ inline void fstar_dev32(const float *__restrict__ A, const Inst &I, int i, int j,
                                   float *g) {
    const int nx = I.nx, ny = I.ny;
    const bool edge = (i == 0) | (j == 0) | (i == nx - 1) | (j == ny - 1);
    if (!edge) {
#pragma unroll
        for (int q = 0; q < 9; ++q)
            g[q] = A[q * I.qs + (int64_t)(j - EYi[q]) * I.pitch + (i - EXi[q])];
        return;
    }
    if (I.bc == 0) {
        if (i == nx - 1) {
            const int jj = j == 0 ? 1 : (j == ny - 1 ? ny - 2 : j);
            float gin[9];
#pragma unroll
            for (int q = 0; q < 9; ++q)
                gin[q] = A[q * I.qs + (int64_t)(jj - EYi[q]) * I.pitch + (nx - 2 - EXi[q])];
            neumann_copy32(gin, g);
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) g[q] = I.gbc[q];
        }
        return;
    }
    const bool corner = (i == 0 || i == nx - 1) && (j == 0 || j == ny - 1);
    if (corner) {
#pragma unroll
        for (int q = 0; q < 9; ++q) g[q] = I.gbc[q];
        return;
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int si = i - EXi[q], sj = j - EYi[q];
        g[q] = (si >= 0 && si < nx && sj >= 0 && sj < ny) ? A[q * I.qs + (int64_t)sj * I.pitch + si]
                                                          : 0.f;
    }
    settling_face32(i == 0 ? 0 : (i == nx - 1 ? 1 : (j == ny - 1 ? 2 : 3)), g);
}

// ---------------------------------------------------------------- block-wide helpers

__device__ inline double nanmax2(double a, double b) {
    return (a != a || a > b) ? a : b;  // np.max semantics: NaN propagates
}


}  // namespace iblbm
