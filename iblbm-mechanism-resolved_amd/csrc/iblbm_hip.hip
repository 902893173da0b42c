// iblbm_hip.hip -- gfx950 kernels and the C ABI of libiblbm_hip.so (see include/iblbm_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <climits>
#include <map>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/iblbm_hip.h"
#include "edge.h"
#include "fused_step.h"
#include "scratch_alloc.h"

using namespace iblbm;

// ------------------------------------------------------------------ kernels

// The instance that a wavefront's current rows belong to, with every field the bulk path needs
// to address a row, held in scalar registers for as long as consecutive rows stay in that
// instance.  Rows of one wavefront are mostly consecutive (one block, or a stride through one
// lattice), so the instance lookup and the descriptor reads happen about once per wavefront
// instead of once per row, and nothing on a row's address chain waits for a scalar load.
struct RowRun {
    const Inst *I;
    int tb_lo, tb_hi;  // the instance's bulk tiles [tb_lo, tb_hi)
    int tiles_x, nx, ny, ibm, bx0, by0, bw, bh, pitch, bc;
    int64_t qs;
    const int *tindex;
    const void *A;  // source planes of this step
    void *B;        // destination planes
};

// Reads are scalar loads through the constant address space: the tables are read-only
// during a step.
__device__ __attribute__((always_inline)) void row_run_load(const LaunchTable &T, int tb,
                                                           RowRun &R) {
    typedef __attribute__((address_space(4))) const int CInt;
    typedef __attribute__((address_space(4))) Inst CInst;
    const CInt *prefix = (const CInt *)T.tile_prefix;
    int lo = 0, hi = T.n_inst - 1;  // the last instance whose first tile is <= tb
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= tb)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int inst = __builtin_amdgcn_readfirstlane(lo);
    const CInst *C = (const CInst *)T.insts + inst;
    R.I = (const Inst *)C;
    R.tb_lo = prefix[inst];
    R.tb_hi = prefix[inst + 1];
    R.tiles_x = C->tiles_x;
    R.nx = C->nx;
    R.ny = C->ny;
    R.ibm = C->ibm;
    R.bx0 = C->bx0;
    R.by0 = C->by0;
    R.bw = C->bw;
    R.bh = C->bh;
    R.pitch = C->pitch;
    R.bc = C->bc;
    R.qs = C->qs;
    R.tindex = C->s.tindex;
    R.A = C->f[T.parity];
    R.B = C->f[T.parity ^ 1];
}

// Bulk lattice node in FP64 (reference arithmetic, bit-exact), in two halves: the pull
// (loads) and the moments + collision + store, so that a thread can have several nodes' loads
// in flight.  The second half stores only when `own` holds: on an interior row every lane runs
// it and only the store is masked, so no branch can leave a load outstanding at the end of a
// row.  (gfx9 counts loads and stores in one vmcnt: a load left outstanding on some path makes
// the compiler wait for vmcnt(0) before the next row's loads, i.e. for the previous row's
// stores to be acknowledged.)
__device__ __attribute__((always_inline)) void bulk_pull(const Inst &I, int parity, int i, int j,
                                                        double *fs) {
    fstar_at(buf<double>(I, parity), I, i, j, fs);
}
// nt: non-temporal stores (the populations written this step are read next step only, so
// for a working set larger than the last-level cache they bypass it)
template <int MODEL, bool NT = false>
__device__ __attribute__((always_inline)) double bulk_finish(const Inst &I, const RowRun &R, int i,
                                                            int j, double *fs, bool own = true) {
    double *__restrict__ B = (double *)R.B;
    double out[9];
    double r;
    const bool inc = I.inc != 0;
    if constexpr (MODEL == 0 || MODEL == 1) {
        r = bulk_collide64<MODEL>(I.cp, fs, inc, out);
    } else {
        const int model = MODEL == 2 ? 2 : I.cp.model;
        if (model == 0) {
            r = bulk_collide64<0>(I.cp, fs, inc, out);
        } else if (model == 1) {
            r = bulk_collide64<1>(I.cp, fs, inc, out);
        } else {
            double u0, u1;
            macro9(fs, inc, r, u0, u1);
            collide_cm<false>(I.cp, fs, u0, u1, 0.0, 0.0, out);
        }
    }
    if (own) {
        if constexpr (NT) {
#pragma unroll
            for (int q = 0; q < 9; ++q)
                __builtin_nontemporal_store(out[q], &B[q * R.qs + (int64_t)j * R.pitch + i]);
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) B[q * R.qs + (int64_t)j * R.pitch + i] = out[q];
        }
    }
    return r;
}

// The same in FP32 deviation form.
__device__ __attribute__((always_inline)) void bulk_pull(const Inst &I, int parity, int i, int j,
                                                        float *g) {
    fstar_dev32(buf<float>(I, parity), I, i, j, g);
}
template <int MODEL, bool NT = false>
__device__ __attribute__((always_inline)) double bulk_finish(const Inst &I, const RowRun &R, int i,
                                                            int j, float *g, bool own = true) {
    float *__restrict__ B = (float *)R.B;
    float out[9];
    float drho, ux, uy;
    const bool inc = I.inc != 0;
    moments32(g, inc, drho, ux, uy);
    if constexpr (MODEL >= 0) {
        collide32<MODEL>(I.cp32, g, drho, ux, uy, inc, out);
    } else {
        if (I.cp.model == 0)
            collide32<0>(I.cp32, g, drho, ux, uy, inc, out);
        else if (I.cp.model == 1)
            collide32<1>(I.cp32, g, drho, ux, uy, inc, out);
        else
            collide32<2>(I.cp32, g, drho, ux, uy, inc, out);
    }
    if (own) {
        if constexpr (NT) {
#pragma unroll
            for (int q = 0; q < 9; ++q)
                __builtin_nontemporal_store(out[q], &B[q * R.qs + (int64_t)j * R.pitch + i]);
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) B[q * R.qs + (int64_t)j * R.pitch + i] = out[q];
        }
    }
    return 1.0 + (double)drho;
}

// Pull of a node whose 8 neighbours all lie inside the grid (no closure, no branches).
__device__ __attribute__((always_inline)) void bulk_pull_interior(const RowRun &R, int i, int j,
                                                                 double *fs) {
    const double *__restrict__ A = (const double *)R.A;
#pragma unroll
    for (int q = 0; q < 9; ++q) fs[q] = A[q * R.qs + (int64_t)(j - EYi[q]) * R.pitch + (i - EXi[q])];
}
__device__ __attribute__((always_inline)) void bulk_pull_interior(const RowRun &R, int i, int j,
                                                                 float *g) {
    const float *__restrict__ A = (const float *)R.A;
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = A[q * R.qs + (int64_t)(j - EYi[q]) * R.pitch + (i - EXi[q])];
}

// Pull of a row that touches the grid's boundary or runs past it.  Every lane loads its 9
// values from addresses clamped into the grid (a node of the Dirichlet-Neumann outlet column
// loads its inner neighbour's, whose sources all lie inside), so a row's loads issue together
// with no branch around them; lanes on the boundary then take the closure's values in
// registers, exactly as fstar_at / fstar_dev32 form them.  Lanes past the grid load valid
// addresses and are masked by the caller.
// The node whose pull a lane at (i, j) takes: itself clamped into the grid, or for the
// Dirichlet-Neumann outlet column its inner neighbour (dirichlet_neumann.py:39-118: pull of
// (nx-2, jj)).
__device__ __attribute__((always_inline)) void edge_base(const RowRun &R, int i, int j, int &ci,
                                                        int &cj) {
    const int nx = R.nx, ny = R.ny;
    ci = i < nx - 1 ? i : nx - 1;
    cj = j < ny - 1 ? j : ny - 1;
    if (R.bc == 0 && ci == nx - 1) {
        ci = nx - 2;
        cj = cj == 0 ? 1 : (cj == ny - 1 ? ny - 2 : cj);
    }
}
// Offset of population q's source for that node, clamped into the grid.
__device__ __attribute__((always_inline)) int64_t edge_off(const RowRun &R, int ci, int cj,
                                                          int q) {
    const int nx = R.nx, ny = R.ny;
    int si = ci - EXi[q], sj = cj - EYi[q];
    si = si < 0 ? 0 : (si > nx - 1 ? nx - 1 : si);
    sj = sj < 0 ? 0 : (sj > ny - 1 ? ny - 1 : sj);
    return q * R.qs + (int64_t)sj * R.pitch + si;
}
template <typename S, typename V>
__device__ __attribute__((always_inline)) void edge_loads(const RowRun &R, int i, int j, V *v) {
    const S *__restrict__ A = (const S *)R.A;
    int ci, cj;
    edge_base(R, i, j, ci, cj);
#pragma unroll
    for (int q = 0; q < 9; ++q) v[q] = A[edge_off(R, ci, cj, q)];
}
__device__ __attribute__((always_inline)) void edge_fix(const RowRun &R, int i, int j,
                                                       double *fs) {
    const int nx = R.nx, ny = R.ny;
    if (i >= nx || j >= ny) return;
    if (!((i == 0) | (j == 0) | (i == nx - 1) | (j == ny - 1))) return;
    const Inst &I = *R.I;
    if (R.bc == 0) {
        if (i == nx - 1) {
            double fin[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) fin[q] = fs[q];
            neumann_copy(fin, fs);
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) fs[q] = I.feq_bc[q];
        }
        return;
    }
    if ((i == 0 || i == nx - 1) && (j == 0 || j == ny - 1)) {
#pragma unroll
        for (int q = 0; q < 9; ++q) fs[q] = I.feq_bc[q];
        return;
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int si = i - EXi[q], sj = j - EYi[q];
        if (!(si >= 0 && si < nx && sj >= 0 && sj < ny)) fs[q] = 0.0;
    }
    settling_face(i == 0 ? 0 : (i == nx - 1 ? 1 : (j == ny - 1 ? 2 : 3)), fs);
}
__device__ __attribute__((always_inline)) void edge_fix(const RowRun &R, int i, int j,
                                                       float *g) {
    const int nx = R.nx, ny = R.ny;
    if (i >= nx || j >= ny) return;
    if (!((i == 0) | (j == 0) | (i == nx - 1) | (j == ny - 1))) return;
    const Inst &I = *R.I;
    if (R.bc == 0) {
        if (i == nx - 1) {
            float gin[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) gin[q] = g[q];
            neumann_copy32(gin, g);
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) g[q] = I.gbc[q];
        }
        return;
    }
    if ((i == 0 || i == nx - 1) && (j == 0 || j == ny - 1)) {
#pragma unroll
        for (int q = 0; q < 9; ++q) g[q] = I.gbc[q];
        return;
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int si = i - EXi[q], sj = j - EYi[q];
        if (!(si >= 0 && si < nx && sj >= 0 && sj < ny)) g[q] = 0.f;
    }
    settling_face32(i == 0 ? 0 : (i == nx - 1 ? 1 : (j == ny - 1 ? 2 : 3)), g);
}
template <typename V>
__device__ __attribute__((always_inline)) void bulk_pull_edge(const RowRun &R, int i, int j,
                                                             V *v) {
    edge_loads<V, V>(R, i, j, v);
}

// The bulk work is a list of rows: row r is row r % TILE_Y of tile r / TILE_Y (TILE_X nodes,
// one per lane of a wavefront).  A lane's node: coordinates, and whether the node is the
// bulk's (inside the grid, not touched by a marker stencil).
struct TileNode {
    int i, j;  // j * pitch + i is the node's offset in a plane; (i, j) is the lattice node,
               // except for a lane of a flat chunk past a row's end (see lattice_ij)
    bool mine;
    bool interior;  // (uniform) every node of the row and all its neighbours are in the grid
};
// row is uniform over the wavefront; R is reloaded only when the row leaves R's instance.
// FL: the context's row layout (0: 64-column tiles of lattice rows; 1: flat chunks of 64
// row-major nodes, every instance at least 64 nodes wide), fixed per kernel so that each
// kernel carries one mapping.
template <int FL>
__device__ __attribute__((always_inline)) TileNode row_node(const LaunchTable &T, RowRun &R,
                                                           int row, int lane) {
    row = __builtin_amdgcn_readfirstlane(row);
    const int tb = row / TILE_Y;
    if (tb < R.tb_lo || tb >= R.tb_hi) row_run_load(T, tb, R);
    const int lt = tb - R.tb_lo;
    TileNode n;
    bool in_grid;
    int ti, tj;  // the lane's lattice node
    if constexpr (FL == 1) {  // chunk c holds nodes 64c .. 64c + 63 in row-major order
        const int k0 = (lt * TILE_Y + row % TILE_Y) * TILE_X;
        // floor(k0 / nx): a single-precision estimate (off by less than one for any grid
        // below 2^20 rows), then one correction either way
        int j0 = (int)((float)k0 * __builtin_amdgcn_rcpf((float)R.nx));
        j0 -= j0 * R.nx > k0 ? 1 : 0;
        j0 += (j0 + 1) * R.nx <= k0 ? 1 : 0;
        j0 = __builtin_amdgcn_readfirstlane(j0);
        const int i0 = k0 - j0 * R.nx;
        const int il = i0 + lane;
        const bool wrap = il >= R.nx;  // at most one wrap: nx >= 64
        // n.j stays the chunk's first row (uniform); a wrapped lane's column is moved on by
        // pitch - nx, so that j * pitch + i is still its node's address
        n.i = wrap ? il - R.nx + R.pitch : il;
        n.j = j0;
        ti = wrap ? il - R.nx : il;
        tj = wrap ? j0 + 1 : j0;
        in_grid = tj < R.ny;
        n.interior = i0 >= 1 && i0 + TILE_X <= R.nx - 1 && j0 >= 1 && j0 <= R.ny - 2;
    } else {
        const int tx = lt % R.tiles_x, ty = lt / R.tiles_x;
        n.i = tx * TILE_X + lane;
        n.j = ty * TILE_Y + row % TILE_Y;
        ti = n.i;
        tj = n.j;
        in_grid = n.i < R.nx && n.j < R.ny;
        n.interior = tx * TILE_X >= 1 && tx * TILE_X + TILE_X <= R.nx - 1 && n.j >= 1 &&
                     n.j <= R.ny - 2;
    }
    // nodes a marker stencil touches belong to the window workgroup; the outermost ring of
    // an instance whose closure k_edge runs belongs to k_edge
    const bool in_box = R.ibm != 0 & ti >= R.bx0 & ti < R.bx0 + R.bw & tj >= R.by0 &
                        tj < R.by0 + R.bh;
    bool touched = false;
    if (in_box) touched = R.tindex[(tj - R.by0) * R.bw + (ti - R.bx0)] >= 0;
    const bool ring0 = (R.bc >= 2) & (ti == 0 | tj == 0 | ti == R.nx - 1 | tj == R.ny - 1);
    n.mine = in_grid && !touched && !ring0;
    return n;
}

// The lattice node of a lane from its row coordinates (n.i may carry a flat chunk's wrap).
template <int FL>
__device__ __attribute__((always_inline)) void lattice_ij(const RowRun &R, const TileNode &n,
                                                         int &i, int &j) {
    if constexpr (FL == 1) {
        const bool wrap = n.i >= R.nx;
        i = wrap ? n.i - R.pitch : n.i;
        j = wrap ? n.j + 1 : n.j;
    } else {
        i = n.i;
        j = n.j;
    }
}

// Bulk rows of one time step (every node that no marker stencil touches): a wavefront takes
// rows first, first + nw, ... below stop (all rows when stop < 0), NR rows at a time with all
// NR rows' loads issued before any of them is finished (NR = 1, 2 or 4: four FP32 rows hold
// the bytes of two FP64 rows in flight).  Each row's FP64 partial rho sum is a fixed shuffle
// tree over the wavefront.
// EB: edge rows take the batched loads of edge_loads / edge_fix (else the per-lane general
// pull); chosen per kernel where it does not raise the register count.
template <typename S, int MODEL, int NR, bool EB, int FL, bool NT = false>
__device__ __attribute__((always_inline)) void bulk_rows(const LaunchTable &T, int first, int nw,
                                                        int stop = -1) {
    static_assert(NR == 1 || NR == 2 || NR == 4, "rows in flight: 1, 2 or 4");
    const int lane = (int)threadIdx.x & 63;
    const int n_rows = stop >= 0 ? stop : T.n_tiles * TILE_Y;
    RowRun R[NR];  // empty tile ranges: the first row of each slot loads it
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        R[k].tb_lo = 1;
        R[k].tb_hi = 0;
    }
    for (int r0 = first; r0 < n_rows; r0 += NR * nw) {
        int rr[NR];
        bool vv[NR];
        TileNode nn[NR];
        bool all_interior = true;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            vv[k] = r0 + k * nw < n_rows;
            rr[k] = vv[k] ? r0 + k * nw : r0;  // a missing row repeats the first, unused
            nn[k] = row_node<FL>(T, R[k], rr[k], lane);
            nn[k].mine = nn[k].mine && vv[k];
            all_interior = all_interior && nn[k].interior;
        }
        S a[NR][9];
        double rho[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) rho[k] = 0.0;
        if (all_interior) {  // uniform: every row's loads in one batch
#pragma unroll
            for (int k = 0; k < NR; ++k) bulk_pull_interior(R[k], nn[k].i, nn[k].j, a[k]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const double v = bulk_finish<MODEL, NT>(*R[k].I, R[k], nn[k].i, nn[k].j, a[k],
                                                        nn[k].mine);
                rho[k] = nn[k].mine ? v : 0.0;
            }
        } else if constexpr (EB) {  // edge rows: the same batched loads, then the closure
            int ii[NR], jj[NR];
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                lattice_ij<FL>(R[k], nn[k], ii[k], jj[k]);
                bulk_pull_edge(R[k], ii[k], jj[k], a[k]);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                edge_fix(R[k], ii[k], jj[k], a[k]);
                const double v = bulk_finish<MODEL, NT>(*R[k].I, R[k], nn[k].i, nn[k].j, a[k],
                                                        nn[k].mine);
                rho[k] = nn[k].mine ? v : 0.0;
            }
        } else {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                if (nn[k].mine) {
                    int i0, j0;
                    lattice_ij<FL>(R[k], nn[k], i0, j0);
                    bulk_pull(*R[k].I, T.parity, i0, j0, a[k]);
                    rho[k] = bulk_finish<MODEL, NT>(*R[k].I, R[k], nn[k].i, nn[k].j, a[k], true);
                }
            }
        }
        if (T.diag) {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const double w = wave_sum(rho[k]);
                if (lane == 0 && vv[k]) T.rho_partial[rr[k]] = w;
            }
        }
    }
}

// One row at a time, software-pipelined: a wavefront issues its next row's loads before the
// current row's collision and stores, so a row's loads are in flight while it computes (the
// register footprint of two rows in flight, without their lock-step).  Only interior rows are
// loaded ahead; an edge row takes its batched (EB) or per-lane pull at its turn.
template <typename S, int MODEL, bool EB, int FL, bool NT = false>
__device__ __attribute__((always_inline)) void bulk_rows_pipe(const LaunchTable &T, int first,
                                                             int nw, int stop = -1) {
    const int lane = (int)threadIdx.x & 63;
    const int n_rows = stop >= 0 ? stop : T.n_tiles * TILE_Y;
    if (first >= n_rows) return;
    RowRun R[2];
    R[0].tb_lo = R[1].tb_lo = 1;
    R[0].tb_hi = R[1].tb_hi = 0;
    S a[2][9];
    TileNode nn[2];
    int rr[2];
    bool pre[2];
    rr[0] = first;
    nn[0] = row_node<FL>(T, R[0], first, lane);
    pre[0] = nn[0].interior;
    if (pre[0]) bulk_pull_interior(R[0], nn[0].i, nn[0].j, a[0]);
    // slot C holds the current row, slot 1 - C receives the next; both indices are constants
    // in each instantiation, so the arrays stay in registers
    auto body = [&](auto cc) -> bool {
        constexpr int C = decltype(cc)::value, D = 1 - C;
        const int rn = rr[C] + nw;
        const bool more = rn < n_rows;
        if (more) {
            rr[D] = rn;
            nn[D] = row_node<FL>(T, R[D], rn, lane);
            pre[D] = nn[D].interior;
            if (pre[D]) bulk_pull_interior(R[D], nn[D].i, nn[D].j, a[D]);
        }
        __builtin_amdgcn_sched_barrier(0);
        double rho = 0.0;
        if (pre[C]) {
            const double v = bulk_finish<MODEL, NT>(*R[C].I, R[C], nn[C].i, nn[C].j, a[C],
                                                    nn[C].mine);
            rho = nn[C].mine ? v : 0.0;
        } else if constexpr (EB) {
            int i0, j0;
            lattice_ij<FL>(R[C], nn[C], i0, j0);
            bulk_pull_edge(R[C], i0, j0, a[C]);
            edge_fix(R[C], i0, j0, a[C]);
            const double v = bulk_finish<MODEL, NT>(*R[C].I, R[C], nn[C].i, nn[C].j, a[C],
                                                    nn[C].mine);
            rho = nn[C].mine ? v : 0.0;
        } else if (nn[C].mine) {
            int i0, j0;
            lattice_ij<FL>(R[C], nn[C], i0, j0);
            bulk_pull(*R[C].I, T.parity, i0, j0, a[C]);
            rho = bulk_finish<MODEL, NT>(*R[C].I, R[C], nn[C].i, nn[C].j, a[C], true);
        }
        if (T.diag) {
            const double w = wave_sum(rho);
            if (lane == 0) T.rho_partial[rr[C]] = w;
        }
        return more;
    };
    while (body(std::integral_constant<int, 0>{}) && body(std::integral_constant<int, 1>{})) {
    }
}

// 4 bytes per lane from global memory straight into LDS (global_load_lds_dword): lane l of the
// calling wavefront writes lds + 4 l.  The copy counts in vmcnt; it is in LDS once the wave's
// vmcnt has reached it.
__device__ __attribute__((always_inline)) void lds_dma4(const void *src, void *lds) {
    typedef __attribute__((address_space(1))) void GVoid;
    typedef __attribute__((address_space(3))) void LVoid;
    __builtin_amdgcn_global_load_lds((GVoid *)src, (LVoid *)lds, 4, 0, 0);
}
// A row's 9 pull planes into LDS, plane q at buf + 64 q (value of lane l at buf[64 q + l]):
// an interior row's plain pulls, an edge row's clamped ones (edge_base / edge_off, the
// addresses edge_loads reads).  FP64 copies a dword per lane, so lane l of the copy k = 0, 1
// moves dword l & 1 of the row's value 32 k + l / 2: an interior tile row's plane is 512
// contiguous bytes, an edge row's values are taken one by one.  FP32 copies one value per lane
// from its own address (any row layout).
template <typename S, int FL>
__device__ __attribute__((always_inline)) void row_to_lds(const RowRun &R, const TileNode &n,
                                                         int lane, S *buf) {
    const S *__restrict__ A = (const S *)R.A;
    if (n.interior) {
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            if constexpr (sizeof(S) == 8) {
                const char *src = (const char *)(A + q * R.qs +
                                                 (int64_t)(n.j - EYi[q]) * R.pitch +
                                                 (n.i - lane - EXi[q]));
                char *dst = (char *)(buf + 64 * q);
                lds_dma4(src + 4 * lane, dst);
                lds_dma4(src + 256 + 4 * lane, dst + 256);
            } else {
                lds_dma4(&A[q * R.qs + (int64_t)(n.j - EYi[q]) * R.pitch + (n.i - EXi[q])],
                         buf + 64 * q);
            }
        }
    } else if constexpr (sizeof(S) == 8) {  // tiles: the row's values are columns i0 + e
        int ci[2], cj[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) edge_base(R, n.i - lane + 32 * k + (lane >> 1), n.j, ci[k], cj[k]);
#pragma unroll
        for (int q = 0; q < 9; ++q)
#pragma unroll
            for (int k = 0; k < 2; ++k)
                lds_dma4((const char *)(A + edge_off(R, ci[k], cj[k], q)) + 4 * (lane & 1),
                         (char *)(buf + 64 * q) + 256 * k);
    } else {
        int i0, j0, ci, cj;
        lattice_ij<FL>(R, n, i0, j0);
        edge_base(R, i0, j0, ci, cj);
#pragma unroll
        for (int q = 0; q < 9; ++q) lds_dma4(&A[edge_off(R, ci, cj, q)], buf + 64 * q);
    }
}

// Rows taken one at a time from the workgroup's counter (below r1), with the next row's
// populations copied into LDS (global_load_lds) while the current row computes: the copy
// holds no registers, so FP64 gets the look-ahead that its registers cannot hold.  Interior
// rows and, with batched edge rows (EB), edge rows are copied ahead (an edge row's closure
// then runs on the copied values); otherwise an edge row takes its per-lane pull at its turn.
// lds: this wavefront's two row buffers of 9 x 64 values.
template <typename S, int MODEL, bool EB, int FL>
__device__ __attribute__((always_inline)) void bulk_rows_lds(const LaunchTable &T, int *queue,
                                                            int r1, S *lds) {
    const int lane = (int)threadIdx.x & 63;
    auto take = [&]() {
        int r = 0;
        if (lane == 0) r = atomicAdd(queue, 1);
        return __builtin_amdgcn_readfirstlane(__shfl(r, 0, 64));
    };
    RowRun R[2];
    R[0].tb_lo = R[1].tb_lo = 1;
    R[0].tb_hi = R[1].tb_hi = 0;
    TileNode nn[2];
    int rr[2];
    bool pre[2];
    rr[0] = take();
    if (rr[0] >= r1) return;
    nn[0] = row_node<FL>(T, R[0], rr[0], lane);
    pre[0] = EB || nn[0].interior;
    if (pre[0]) row_to_lds<S, FL>(R[0], nn[0], lane, lds);
    auto body = [&](auto cc) -> bool {
        constexpr int C = decltype(cc)::value, D = 1 - C;
        S a[9];
        if (pre[C]) {
            // every copy this wavefront issued has landed (the only ones outstanding are this
            // row's, and the previous row's stores)
            __builtin_amdgcn_sched_barrier(0);
            // s_waitcnt vmcnt(0) (gfx9 encoding: vmcnt in bits 3:0 and 15:14 = 0, expcnt in
            // 6:4 and lgkmcnt in 11:8 at their maxima, i.e. not waited on)
            __builtin_amdgcn_s_waitcnt(0x0F70);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 9; ++q) a[q] = lds[C * 9 * 64 + 64 * q + lane];
        }
        const int rn = take();
        const bool more = rn < r1;
        if (more) {
            rr[D] = rn;
            nn[D] = row_node<FL>(T, R[D], rn, lane);
            pre[D] = EB || nn[D].interior;
            if (pre[D]) row_to_lds<S, FL>(R[D], nn[D], lane, lds + D * 9 * 64);
        }
        __builtin_amdgcn_sched_barrier(0);
        double rho = 0.0;
        if (pre[C]) {
            if (!nn[C].interior) {  // the closure on the copied (clamped) pulls
                int i0, j0;
                lattice_ij<FL>(R[C], nn[C], i0, j0);
                edge_fix(R[C], i0, j0, a);
            }
            const double v = bulk_finish<MODEL>(*R[C].I, R[C], nn[C].i, nn[C].j, a, nn[C].mine);
            rho = nn[C].mine ? v : 0.0;
        } else if (nn[C].mine) {
            int i0, j0;
            lattice_ij<FL>(R[C], nn[C], i0, j0);
            bulk_pull(*R[C].I, T.parity, i0, j0, a);
            rho = bulk_finish<MODEL>(*R[C].I, R[C], nn[C].i, nn[C].j, a, true);
        }
        if (T.diag) {
            const double w = wave_sum(rho);
            if (lane == 0) T.rho_partial[rr[C]] = w;
        }
        return more;
    };
    while (body(std::integral_constant<int, 0>{}) && body(std::integral_constant<int, 1>{})) {
    }
}

// Bulk kernel of the two-kernel step.  Any workgroup size up to 1024 (a tile of 64 x
// blockDim / 64 rows per workgroup); the grid either covers every row once or walks them
// with a grid stride (the schedule picks both).
// NT: non-temporal stores (a compile-time choice: as a run-time flag the compiler merges the
// two store forms and drops the hint)
template <typename S, int MODEL, int NR, int FL, bool NT>
__global__ __launch_bounds__(1024) void k_bulk(LaunchTable T) {
    const int wpb = (int)blockDim.x >> 6;
    constexpr bool EB = sizeof(S) == 4;
    const int first = (int)blockIdx.x * wpb + (int)(threadIdx.x >> 6);
    if constexpr (NR == 3 && sizeof(S) == 4)
        bulk_rows_pipe<S, MODEL, EB, FL, NT>(T, first, (int)gridDim.x * wpb);
    else
        bulk_rows<S, MODEL, NR, EB, FL, NT>(T, first, (int)gridDim.x * wpb);
}

// Immersed-boundary workgroups of one time step (one per instance with markers).  PRED:
// with the learned predictor's modes (picked when an instance has one set).
template <typename S, int MODEL, bool PRED>
__global__ __launch_bounds__(NTI) void k_ibm(LaunchTable T) {
    __shared__ double red[NTI];
    Inst &I = T.insts[T.ibm_inst[blockIdx.x]];
    ibm_role<S, MODEL, PRED>(I, T.parity, T.diag, red);
}

// The three outermost rings of the instances whose closure reads the previous velocity or an
// analytic velocity (edge.h).  One thread per ring node; workgroup b covers ring slots
// [edge_blocks[b].y, + 256) of instance edge_blocks[b].x.
template <typename S>
__global__ __launch_bounds__(256) void k_edge(LaunchTable T) {
    __shared__ double red[4];
    const int2 eb = T.edge_blocks[blockIdx.x];
    const Inst &I = T.insts[eb.x];
    const int s = eb.y + (int)threadIdx.x;
    const long long t = T.stepc[T.parity];  // this step's number (1-based)
    if (blockIdx.x == 0 && threadIdx.x == 0) T.stepc[T.parity ^ 1] = t + 1;
    const double *side_prev = I.side + (size_t)2 * I.nring * T.parity;
    double *side_next = I.side + (size_t)2 * I.nring * (T.parity ^ 1);
    double rho_b = 0.0;
    if (s < I.nring) {
        int i, j;
        ring_node(s, I.nx, I.ny, i, j);
        const S *__restrict__ A = reinterpret_cast<const S *>(I.f[T.parity]);
        double fs[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const int si = i - EXi[q], sj = j - EYi[q];
            fs[q] = (si >= 0 && si < I.nx && sj >= 0 && sj < I.ny)
                        ? Store<S>::ld(A, q * I.qs + (int64_t)sj * I.pitch + si, q)
                        : 0.0;
        }
        const bool outer = i == 0 || j == 0 || i == I.nx - 1 || j == I.ny - 1;
        if (outer) {
            double tgv[2] = {0.0, 0.0};
            if (I.bc == IBLBM_BC_TG_ANALYTICAL) {
                const int k = tg_index(i, j, I.nx, I.ny);
                const double d = I.tgd[t < I.tg_nt ? t : I.tg_nt - 1];
                tgv[0] = I.tgc[2 * k] * d;
                tgv[1] = I.tgc[2 * k + 1] * d;
            }
            edge_closure(I, side_prev, tgv, i, j, fs);
#pragma unroll
            for (int q = 0; q < 9; ++q) I.edge_fs[9 * s + q] = fs[q];
        }
        double r, u0, u1;
        macro9(fs, I.inc != 0, r, u0, u1);
        side_next[2 * s] = u0;
        side_next[2 * s + 1] = u1;
        if (outer) {
            double fe[9], out[9];
            if (I.cp.model != 2) feq9(r, u0, u1, I.inc != 0, fe);
            collide<false>(I.cp, fs, fe, u0, u1, 0.0, 0.0, out);
            store9(reinterpret_cast<S *>(I.f[T.parity ^ 1]), I, i, j, out);
            rho_b = r;
        }
    }
    if (T.diag) {  // the outer ring's rho, a fixed shuffle tree, then the 4 wavefronts
        const double w = wave_sum(rho_b);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
        __syncthreads();
        if (threadIdx.x == 0) T.edge_rho[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
    }
}

// One whole time step in one launch, for contexts whose windows all take the fused path:
// workgroups [0, n_fast) run the windows (fused_step.h), the others run four bulk tiles each
// (one node per thread).  One launch per step leaves no cross-stream wait between steps.
// LA: the row queue's LDS look-ahead form (rows = 3) is compiled in; it is left out where its
// registers would add spills to the kernel (the central-moment and mixed collisions, and the
// persistent launch), which then takes two rows at a time.
template <typename S, int MODEL, bool PRED, int FL,
          bool LA = !PRED && (MODEL == 0 || MODEL == 1)>
__device__ __forceinline__ void step_body(const LaunchTable &T, double *red, FusedShared &sh) {
    if ((int)blockIdx.x < T.n_fast) {
        __builtin_amdgcn_s_setprio(3);  // issue ahead of bulk wavefronts sharing the CU
        // The instance table and descriptors are read-only during a step: reading them through
        // the constant address space with a uniform index makes them scalar (K$) loads.
        typedef __attribute__((address_space(4))) const int CInt;
        typedef __attribute__((address_space(4))) Inst CInst;
        const int id = __builtin_amdgcn_readfirstlane(((CInt *)T.ibm_inst)[blockIdx.x]);
        Inst &I = *(Inst *)((CInst *)T.insts + id);
        ibm_fused_dispatch<S, MODEL, PRED>(I, T.parity, T.diag, red, sh);
        return;
    }
    // bulk: every wavefront walks the rows with a grid stride (no workgroup barrier: a
    // wavefront moves on as soon as its rows are done).  With a start delay the windows'
    // first loads (their critical path) meet an idle memory system; the wait is bounded by
    // the real-time clock and by an iteration cap.
    if (T.bulk_delay > 0) {
        const long long t0 = (long long)wall_clock64();
        for (int k = 0; k < 100000 && (long long)wall_clock64() - t0 < T.bulk_delay; ++k)
            __builtin_amdgcn_s_sleep(1);
    }
    constexpr bool EB = sizeof(S) == 4 || MODEL >= 0;
    // the pipelined rows fit the step kernel's registers with FP32 storage and the BGK / TRT
    // collision only (FP64 and the central-moment collision would spill)
    constexpr bool PIPE = sizeof(S) == 4 && (MODEL == 0 || MODEL == 1);
    const int first = ((int)blockIdx.x - T.n_fast) * (NTF / 64) + (int)(threadIdx.x >> 6);
    const int nw = ((int)gridDim.x - T.n_fast) * (NTF / 64);
    if (T.blocked < 0) {
        // one block of adjacent rows per workgroup, its wavefronts taking the next rows from a
        // counter in LDS: the SIMDs issue their oldest wavefronts first, so equal shares per
        // wavefront leave the last step of a workgroup to one or two wavefronts per SIMD
        // (fewer loads in flight); a shared counter keeps every wavefront busy to the end
        const int n_rows = T.n_tiles * TILE_Y;
        const int b = (int)blockIdx.x - T.n_fast, nb = (int)gridDim.x - T.n_fast;
        const int r0 = (int)((int64_t)b * n_rows / nb);
        const int r1 = (int)((int64_t)(b + 1) * n_rows / nb);
        int *queue = reinterpret_cast<int *>(red);  // bulk workgroups use no reductions
        if (threadIdx.x == 0) *queue = r0;
        __syncthreads();
        if constexpr (LA) {
            if (T.nrows == 3) {
                // one row at a time, the next row's populations copied into LDS ahead: the
                // window's shared memory is unused in a bulk workgroup, and holds two row
                // buffers per wavefront
                static_assert(sizeof(FusedShared) >= NTF / 64 * 2 * 9 * 64 * sizeof(double),
                              "row buffers");
                S *buf = reinterpret_cast<S *>(&sh) + (threadIdx.x >> 6) * 2 * 9 * 64;
                bulk_rows_lds<S, MODEL, EB, FL>(T, queue, r1, buf);
                return;
            }
        }
        const int grab = T.nrows >= 2 ? 2 : 1;
        for (;;) {
            int r = 0;
            if ((threadIdx.x & 63) == 0) r = atomicAdd(queue, grab);
            r = __builtin_amdgcn_readfirstlane(__shfl(r, 0, 64));
            if (r >= r1) break;
            if (grab == 2)
                bulk_rows<S, MODEL, 2, EB, FL>(T, r, 1, min(r + 2, r1));
            else
                bulk_rows<S, MODEL, 1, EB, FL>(T, r, 1, min(r + 1, r1));
        }
    } else if (T.blocked) {  // one block of adjacent rows per wavefront
        const int n_rows = T.n_tiles * TILE_Y;
        int r0, r1;
        if (T.blocked == 1) {  // every wavefront, block sizes differing by one
            r0 = (int)((int64_t)first * n_rows / nw);
            r1 = (int)((int64_t)(first + 1) * n_rows / nw);
        } else {  // blocks of T.blocked rows; the wavefronts past the last block stay idle
            r0 = (int)min((int64_t)n_rows, (int64_t)first * T.blocked);
            r1 = min(n_rows, r0 + T.blocked);
        }
        if (sizeof(S) == 4 && T.nrows == 4)
            bulk_rows<S, MODEL, (sizeof(S) == 4 ? 4 : 2), EB, FL>(T, r0, 1, r1);
        else if (PIPE && T.nrows == 3)
            bulk_rows_pipe<S, MODEL, EB, FL>(T, r0, 1, r1);
        else if (T.nrows >= 2)
            bulk_rows<S, MODEL, 2, EB, FL>(T, r0, 1, r1);
        else
            bulk_rows<S, MODEL, 1, EB, FL>(T, r0, 1, r1);
    } else if (sizeof(S) == 4 && T.nrows == 4) {
        bulk_rows<S, MODEL, (sizeof(S) == 4 ? 4 : 2), EB, FL>(T, first, nw);
    } else if (PIPE && T.nrows == 3) {
        bulk_rows_pipe<S, MODEL, EB, FL>(T, first, nw);
    } else if (T.nrows >= 2) {
        bulk_rows<S, MODEL, 2, EB, FL>(T, first, nw);
    } else {
        bulk_rows<S, MODEL, 1, EB, FL>(T, first, nw);
    }
}
template <typename S, int MODEL, bool PRED, int FL>
__global__ __launch_bounds__(NTF) void k_step(LaunchTable T) {
    __shared__ double red[NTF / 64 + 3];  // wavefront partials, then 3 residual slots
    __shared__ FusedShared sh;
    step_body<S, MODEL, PRED, FL>(T, red, sh);
}

// Grid-wide barrier of a cooperative launch (every workgroup resident): after a workgroup
// barrier (each wavefront's stores complete), thread 0 arrives with an agent-scope release
// (the step's writes written back for every XCD), polls the counter with relaxed agent-scope
// loads until all workgroups of this step have arrived, then takes one agent-scope acquire
// fence (an acquire on every poll would invalidate the caches under the workgroups still
// working).  The wait is bounded: past the spin limit it records an error and goes on, so the
// launch always ends.
__device__ inline void grid_barrier(unsigned *bar, unsigned target, int *err) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        int k = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++k > (1 << 22)) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Several time steps in one persistent launch: the fused step's body, then a grid-wide
// barrier, step after step (parity and diagnostic steps as the per-step launches have them).
template <typename S, int MODEL, bool PRED, int FL>
__global__ __launch_bounds__(NTF) void k_step_persist(LaunchTable T, PersistArgs A) {
    __shared__ double red[NTF / 64 + 3];
    __shared__ FusedShared sh;
    for (int s = 0; s < A.nsteps; ++s) {
        LaunchTable Ts = T;
        Ts.parity = (A.parity0 + s) & 1;
        const long long stepno = A.step0 + s + 1;
        Ts.diag = (A.diag_every > 0 && stepno % A.diag_every == 0) ? 1 : 0;
        step_body<S, MODEL, PRED, FL, false>(Ts, red, sh);
        if (s + 1 < A.nsteps) grid_barrier(A.bar, gridDim.x * (unsigned)(s + 1), A.err);
    }
}

// Seeded synthetic initial state f = f_eq(rho, u(1 + amp xi)) in both population buffers;
// xi in (-1, 1) per node and component from a splitmix64 hash of (seed, node, component).
__device__ inline double hash_sym(unsigned long long seed, long long key) {
    unsigned long long z = seed * 0xD1B54A32D192ED03ull + (unsigned long long)key;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-52 - 1.0;  // [-1, 1)
}
template <typename S>
__global__ __launch_bounds__(256) void k_init_flow(Inst I, double rho, double ux, double uy,
                                                   double amp, unsigned long long seed) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)I.nx * I.ny) return;
    const int i = (int)(n % I.nx), j = (int)(n / I.nx);
    const double u0 = ux * (1.0 + amp * hash_sym(seed, 2 * n));
    const double u1 = uy * (1.0 + amp * hash_sym(seed, 2 * n + 1));
    double fe[9];
    feq9(rho, u0, u1, I.inc != 0, fe);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int64_t at = q * I.qs + (int64_t)j * I.pitch + i;
        const auto v = Store<S>::st(fe[q], q);
        reinterpret_cast<S *>(I.f[0])[at] = v;
        reinterpret_cast<S *>(I.f[1])[at] = v;
    }
}

// Window tables of one instance for newly set markers (one workgroup).
__global__ __launch_bounds__(NTI) void k_tables(Inst *Ip) {
    __shared__ int redi[NTI];
    ibm_tables_role(*Ip, redi);
}

// Batched marker update: workgroup b copies instance ids[b]'s markers out of the staging
// buffer [Lx (total) | Ly (total) | ud (2 total)] at marker offset off[b].
__global__ __launch_bounds__(256) void k_scatter_markers(const Inst *insts, const int *ids,
                                                         const int *off, const double *stage,
                                                         int total) {
    const Inst &I = insts[ids[blockIdx.x]];
    const int o = off[blockIdx.x], lb = off[blockIdx.x + 1] - o;
    double *Lx = const_cast<double *>(I.Lx), *Ly = const_cast<double *>(I.Ly);
    double *ud = const_cast<double *>(I.ud);
    for (int k = threadIdx.x; k < lb; k += blockDim.x) {
        Lx[k] = stage[o + k];
        Ly[k] = stage[total + o + k];
        ud[2 * k] = stage[2 * total + 2 * (o + k)];
        ud[2 * k + 1] = stage[2 * total + 2 * (o + k) + 1];
    }
}

// Window tables of several instances, one workgroup each; the touched-node count and the
// path flag of each go to nt_fast[2b], nt_fast[2b + 1] for the host descriptors.
__global__ __launch_bounds__(NTI) void k_tables_batch(Inst *insts, const int *ids,
                                                      int *nt_fast) {
    __shared__ int redi[NTI];
    Inst &I = insts[ids[blockIdx.x]];
    ibm_tables_role(I, redi);
    __syncthreads();
    if (threadIdx.x == 0) {
        nt_fast[2 * blockIdx.x] = I.nT;
        nt_fast[2 * blockIdx.x + 1] = I.fast;
    }
}

// Batched force read: workgroup b copies the (Lb, 2) marker forces of instance ids[b] (the
// DFC force q, or the last pass's marker forces of a several-particle DF window) to out at
// marker offset off[b].
__global__ __launch_bounds__(256) void k_gather_forces(const Inst *insts, const int *ids,
                                                       const int *off, double *out) {
    const Inst &I = insts[ids[blockIdx.x]];
    const double *src = I.ibm == IBLBM_IBM_DFC ? I.s.q : I.s.mf;
    const int o = off[blockIdx.x], n2 = 2 * (off[blockIdx.x + 1] - o);
    for (int k = threadIdx.x; k < n2; k += blockDim.x) out[2 * o + k] = src[k];
}

// state.fstar / ro / U of the last step (source buffer A = that step's input) on the box
// [x0, x0 + w) x [y0, y0 + h) (box-major outputs; fstar may be null); touched nodes come from
// the IBM scratch (DFC-corrected fstar, velocity-corrected U).
template <typename S>
__global__ __launch_bounds__(256) void k_state(const Inst *Ip, int parity, int x0, int y0, int w,
                                               int h, double *fstar, double *ro, double *U) {
    const Inst &I = *Ip;
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)w * h) return;
    const int i = x0 + (int)(n % w), j = y0 + (int)(n / w);
    double fs[9], r, u0, u1;
    const bool in_box = I.ibm != 0 && i >= I.bx0 && i < I.bx0 + I.bw && j >= I.by0 &&
                        j < I.by0 + I.bh;
    const int sl = in_box ? (j - I.by0) * I.bw + (i - I.bx0) : 0;
    const bool outer = i == 0 || j == 0 || i == I.nx - 1 || j == I.ny - 1;
    if (in_box && I.s.tindex[sl] >= 0) {
        for (int q = 0; q < 9; ++q) fs[q] = I.s.fs[q * I.box_cap + sl];
        r = I.s.ro[sl];
        u0 = I.s.ux0[sl];
        u1 = I.s.uy0[sl];
    } else if (I.bc >= 2 && outer) {  // the boundary kernel's post-closure populations
        const int rs = ring_slot(i, j, I.nx, I.ny);
        for (int q = 0; q < 9; ++q) fs[q] = I.edge_fs[9 * rs + q];
        macro9(fs, I.inc != 0, r, u0, u1);
    } else {
        fstar_at(buf<S>(I, parity), I, i, j, fs);
        macro9(fs, I.inc != 0, r, u0, u1);
    }
    if (fstar)
        for (int q = 0; q < 9; ++q) fstar[9 * n + q] = fs[q];
    ro[n] = r;
    U[2 * n] = u0;
    U[2 * n + 1] = u1;
}

// AoS (N,9) FP64 <-> SoA f[q][j][pitch] (FP64 f, or FP32 g = f - w)
template <typename S>
__global__ void k_aos_to_soa(const double *aos, S *soa, int nx, int ny, int pitch, int64_t qs) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)nx * ny) return;
    const int i = (int)(n % nx), j = (int)(n / nx);
    for (int q = 0; q < 9; ++q) soa[q * qs + (int64_t)j * pitch + i] = Store<S>::st(aos[9 * n + q], q);
}
template <typename S>
__global__ void k_soa_to_aos(const S *soa, double *aos, int nx, int ny, int pitch, int64_t qs) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)nx * ny) return;
    const int i = (int)(n % nx), j = (int)(n / nx);
    for (int q = 0; q < 9; ++q) aos[9 * n + q] = Store<S>::ld(soa, q * qs + (int64_t)j * pitch + i, q);
}

// ------------------------------------------------------------------ host side

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(IBLBM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));  \
    } while (0)

struct DevAlloc {
    std::vector<void *> ptrs;
    hipError_t alloc(void **p, size_t bytes) {
        hipError_t e = hipMalloc(p, bytes < 8 ? 8 : bytes);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
    void release() {
        for (void *p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
};

struct HostInst {
    iblbm_case c;
    Inst d;  // shadow of the device descriptor
    int Lb = 0;
    double Larea = 0.0;
    int pfb_np = 0;  // particles the per-particle window fields are sized for
};

struct Ctx {
    int device = 0;
    int n_cu = 256;
    unsigned *d_bar = nullptr;  // persistent launch: grid-barrier counter, then its error flag
    double *d_box = nullptr;    // download_state_box staging (ro, U of a box), box_cap doubles
    size_t box_cap = 0;
    hipStream_t stream = nullptr, side = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_main = nullptr, ev_side = nullptr;
    bool concurrent = true;
    DevAlloc mem;
    std::vector<HostInst> inst;
    Inst *d_insts = nullptr;
    int *d_ibm_inst = nullptr, *d_tile_prefix = nullptr;
    double *d_rho_partial = nullptr;
    int cap_inst = 0, cap_blocks = 0;
    bool tables_dirty = true;
    int n_ibm = 0, n_fast = 0, n_tiles = 0;
    std::vector<int> h_tile_prefix, h_ibm_inst;
    int64_t step = 0;
    int parity = 0;
    float last_ms = 0.0f;
    int last_steps = 0;
    int64_t diag_step = -1;
    std::map<std::tuple<int, int, int, int>, hipGraphExec_t> graphs;
    iblbm_schedule sched{IBLBM_MODE_AUTO, 256, 1, 0, 2, 0, 0, 0, 0};
    // asynchronous diagnostics: two pinned host slots [row partials | per-instance box sums]
    // filled by copies queued behind the steps, each with its completion event
    double *h_diag[2] = {nullptr, nullptr};
    size_t h_diag_cap = 0;
    hipEvent_t ev_diag[2] = {nullptr, nullptr};
    int64_t diag_step_slot[2] = {-1, -1};
    hipEvent_t ev_mark[4] = {nullptr, nullptr, nullptr, nullptr};  // caller's timing marks
    // batched marker updates and force reads of many instances: pinned host staging and its
    // device twin (doubles: markers or forces; ints: instance ids, offsets, nT/fast pairs)
    double *h_bd = nullptr, *d_bd = nullptr;
    int *h_bi = nullptr, *d_bi = nullptr;
    size_t bd_cap = 0, bi_cap = 0;
    // boundary kernel of the closures that read the previous velocity (edge.h): its
    // workgroups (instance, first ring slot), their rho partials, the first workgroup of each
    // instance, and the device step counter it reads
    int2 *d_edge_blocks = nullptr;
    double *d_edge_rho = nullptr;
    long long *d_stepc = nullptr;
    int n_edge_blocks = 0, cap_edge_blocks = 0;
    std::vector<int> h_edge_prefix;
};

void clear_graphs(Ctx &C) {
    for (auto &g : C.graphs) (void)hipGraphExecDestroy(g.second);
    C.graphs.clear();
}

// The launch mode a step takes under the context's schedule: fused needs every window on
// the fast path and the concurrent setting; an explicit mode is honoured where it applies.
int effective_mode(const Ctx &C) {
    const bool fused_ok = C.n_fast == C.n_ibm;
    if (!C.concurrent) return IBLBM_MODE_SERIAL;
    switch (C.sched.mode) {
        case IBLBM_MODE_FUSED: return fused_ok ? IBLBM_MODE_FUSED : IBLBM_MODE_TWO_STREAMS;
        case IBLBM_MODE_TWO_STREAMS: return C.n_ibm > 0 ? IBLBM_MODE_TWO_STREAMS : IBLBM_MODE_SERIAL;
        case IBLBM_MODE_SERIAL: return IBLBM_MODE_SERIAL;
        default: break;
    }
    if (fused_ok) return IBLBM_MODE_FUSED;
    return C.n_ibm > 0 ? IBLBM_MODE_TWO_STREAMS : IBLBM_MODE_SERIAL;
}

// Bulk row geometry of an instance: 64-column tiles of TILE_Y lattice rows, or (flat)
// tiles of TILE_Y chunks of 64 consecutive row-major nodes.
void set_row_layout(Inst &d, bool flat) {
    d.flat = flat ? 1 : 0;
    if (d.flat) {
        const int64_t chunks = ((int64_t)d.nx * d.ny + TILE_X - 1) / TILE_X;
        d.tiles_x = 1;
        d.tiles_y = (int)((chunks + TILE_Y - 1) / TILE_Y);
    } else {
        d.tiles_x = (d.nx + TILE_X - 1) / TILE_X;
        d.tiles_y = (d.ny + TILE_Y - 1) / TILE_Y;
    }
}

// Applies the schedule's row layout to every instance (the row list changes: rebuild it).
// The flat layout applies to the whole context when its populations are FP32 deviations and
// every instance is at least 64 nodes wide (a chunk then wraps at most once); otherwise the
// context keeps tiles.  (In FP64 the fused step kernel sits at its register limit, and the
// per-lane row index of a wrapping chunk would make it spill.)
bool flat_eligible(const Ctx &C) {
    bool ok = !C.inst.empty();
    for (const HostInst &h : C.inst) ok = ok && h.d.nx >= TILE_X && h.d.prec == 1;
    return ok;
}
void apply_row_layout(Ctx &C) {
    const bool flat = C.sched.row_layout == 1 && flat_eligible(C);
    bool changed = false;
    for (HostInst &h : C.inst) {
        const int f = h.d.flat, tx = h.d.tiles_x, ty = h.d.tiles_y;
        set_row_layout(h.d, flat);
        changed |= f != h.d.flat || tx != h.d.tiles_x || ty != h.d.tiles_y;
    }
    if (changed) C.tables_dirty = true;
}

int ensure_tables(Ctx &C) {
    if (!C.tables_dirty) return 0;
    const int n = (int)C.inst.size();
    if (n == 0) return 0;  // built once the first instance is added
    C.h_tile_prefix.assign(n + 1, 0);
    C.h_ibm_inst.clear();
    for (int k = 0; k < n; ++k) {
        const Inst &d = C.inst[k].d;
        C.h_tile_prefix[k + 1] = C.h_tile_prefix[k] + d.tiles_x * d.tiles_y;
        if (d.ibm != 0 && d.fast) C.h_ibm_inst.push_back(k);
    }
    C.n_fast = (int)C.h_ibm_inst.size();
    for (int k = 0; k < n; ++k) {
        const Inst &d = C.inst[k].d;
        if (d.ibm != 0 && !d.fast) C.h_ibm_inst.push_back(k);
    }
    C.n_tiles = C.h_tile_prefix[n];
    C.n_ibm = (int)C.h_ibm_inst.size();
    const int blocks = C.n_tiles * TILE_Y + C.n_ibm;  // row partials
    if (n > C.cap_inst) {
        int cap = n < 16 ? 16 : 2 * n;
        void *p;
        HIPCHK(C.mem.alloc(&p, sizeof(Inst) * cap));
        C.d_insts = (Inst *)p;
        HIPCHK(C.mem.alloc(&p, sizeof(int) * (cap + 1)));
        C.d_tile_prefix = (int *)p;
        HIPCHK(C.mem.alloc(&p, sizeof(int) * (cap + 1)));
        C.d_ibm_inst = (int *)p;
        C.cap_inst = cap;
    }
    if (blocks > C.cap_blocks) {
        void *p;
        HIPCHK(C.mem.alloc(&p, sizeof(double) * blocks));
        C.d_rho_partial = (double *)p;
        C.cap_blocks = blocks;
    }
    // the boundary kernel's workgroups: 256 ring slots each, instance by instance
    std::vector<int2> eb;
    C.h_edge_prefix.assign(n + 1, 0);
    for (int k = 0; k < n; ++k) {
        const Inst &d = C.inst[k].d;
        if (d.bc >= 2)
            for (int s0 = 0; s0 < d.nring; s0 += 256) eb.push_back(make_int2(k, s0));
        C.h_edge_prefix[k + 1] = (int)eb.size();
    }
    C.n_edge_blocks = (int)eb.size();
    if (C.n_edge_blocks > C.cap_edge_blocks) {
        void *p;
        HIPCHK(C.mem.alloc(&p, sizeof(int2) * C.n_edge_blocks));
        C.d_edge_blocks = (int2 *)p;
        HIPCHK(C.mem.alloc(&p, sizeof(double) * C.n_edge_blocks));
        C.d_edge_rho = (double *)p;
        HIPCHK(hipMemsetAsync(C.d_edge_rho, 0, sizeof(double) * C.n_edge_blocks, C.stream));
        C.cap_edge_blocks = C.n_edge_blocks;
    }
    if (C.n_edge_blocks && !C.d_stepc) {
        void *p;
        HIPCHK(C.mem.alloc(&p, sizeof(long long) * 2));
        C.d_stepc = (long long *)p;
        HIPCHK(hipMemsetAsync(C.d_stepc, 0, sizeof(long long) * 2, C.stream));
    }
    if (C.n_edge_blocks)
        HIPCHK(hipMemcpyAsync(C.d_edge_blocks, eb.data(), sizeof(int2) * eb.size(),
                              hipMemcpyHostToDevice, C.stream));
    std::vector<Inst> hd(n);
    for (int k = 0; k < n; ++k) hd[k] = C.inst[k].d;
    HIPCHK(hipMemcpyAsync(C.d_insts, hd.data(), sizeof(Inst) * n, hipMemcpyHostToDevice,
                          C.stream));
    HIPCHK(hipMemcpyAsync(C.d_tile_prefix, C.h_tile_prefix.data(), sizeof(int) * (n + 1),
                          hipMemcpyHostToDevice, C.stream));
    if (C.n_ibm)
        HIPCHK(hipMemcpyAsync(C.d_ibm_inst, C.h_ibm_inst.data(), sizeof(int) * C.n_ibm,
                              hipMemcpyHostToDevice, C.stream));
    HIPCHK(hipStreamSynchronize(C.stream));
    clear_graphs(C);
    C.tables_dirty = false;
    return 0;
}

int upload_inst(Ctx &C, int id) {
    // replaces the device descriptor with the host shadow
    HIPCHK(hipMemcpyAsync(C.d_insts + id, &C.inst[id].d, sizeof(Inst), hipMemcpyHostToDevice,
                          C.stream));
    return 0;
}

LaunchTable make_table(Ctx &C, int parity, int diag) {
    LaunchTable T;
    T.insts = C.d_insts;
    T.ibm_inst = C.d_ibm_inst;
    T.tile_prefix = C.d_tile_prefix;
    T.rho_partial = C.d_rho_partial;
    T.n_inst = (int)C.inst.size();
    T.n_ibm = C.n_ibm;
    T.n_fast = C.n_fast;
    T.n_tiles = C.n_tiles;
    T.parity = parity;
    T.diag = diag;
    T.nrows = 1;
    T.blocked = 0;
    T.bulk_delay = 0;
    T.edge_blocks = C.d_edge_blocks;
    T.edge_rho = C.d_edge_rho;
    T.stepc = C.d_stepc;
    T.n_edge_blocks = C.n_edge_blocks;
    return T;
}

typedef void (*StepKernel)(LaunchTable);

template <typename S, int MODEL, int FL, bool NT>
StepKernel pick_bulk(int nr) {
    if constexpr (sizeof(S) == 4) {  // four rows in flight, or one pipelined: FP32 storage
        if (nr == 4) return k_bulk<S, MODEL, 4, FL, NT>;
        if (nr == 3) return k_bulk<S, MODEL, 3, FL, NT>;
    }
    return nr >= 2 ? k_bulk<S, MODEL, 2, FL, NT> : k_bulk<S, MODEL, 1, FL, NT>;
}
template <typename S, int MODEL, int FL>
void pick_layout(bool pred, int nr, bool nt, StepKernel &bulk, StepKernel &ibm,
                 StepKernel &step) {
    bulk = nt ? pick_bulk<S, MODEL, FL, true>(nr) : pick_bulk<S, MODEL, FL, false>(nr);
    ibm = pred ? k_ibm<S, MODEL, true> : k_ibm<S, MODEL, false>;
    step = pred ? k_step<S, MODEL, true, FL> : k_step<S, MODEL, false, FL>;
}
template <typename S, int MODEL>
void pick_pair(bool pred, int nr, bool flat, bool nt, StepKernel &bulk, StepKernel &ibm,
               StepKernel &step) {
    if constexpr (sizeof(S) == 4) {  // FP64 contexts keep tiles (see apply_row_layout)
        if (flat) return pick_layout<S, MODEL, 1>(pred, nr, nt, bulk, ibm, step);
    }
    pick_layout<S, MODEL, 0>(pred, nr, nt, bulk, ibm, step);
}

typedef void (*PersistKernel)(LaunchTable, PersistArgs);

// The persistent step kernel of a context whose instances share one collision model (the
// tile row layout; nullptr where no such kernel is built).
template <typename S>
PersistKernel pick_persist_model(int model, bool pred) {
    switch (model) {
        case 0: return pred ? k_step_persist<S, 0, true, 0> : k_step_persist<S, 0, false, 0>;
        case 1: return pred ? k_step_persist<S, 1, true, 0> : k_step_persist<S, 1, false, 0>;
        case 2: return pred ? k_step_persist<S, 2, true, 0> : k_step_persist<S, 2, false, 0>;
        default: return nullptr;
    }
}
PersistKernel pick_persist(const Ctx &C) {
    int model = C.inst[0].d.cp.model;
    bool pred = false;
    for (const HostInst &h : C.inst) {
        if (h.d.cp.model != model) model = -1;
        pred = pred || h.d.pmode != 0;
    }
    if (C.inst[0].d.flat != 0) return nullptr;
    return C.inst[0].d.prec == 1 ? pick_persist_model<float>(model, pred)
                                 : pick_persist_model<double>(model, pred);
}

void pick_kernels(const Ctx &C, StepKernel &bulk, StepKernel &ibm, StepKernel &step) {
    const int two = C.sched.bulk_rows;  // rows in flight of the bulk kernel
    int model = C.inst[0].d.cp.model;
    bool pred = false;  // the step kernel with the predictor modes only when one is set
    for (const HostInst &h : C.inst) {
        if (h.d.cp.model != model) model = -1;
        pred = pred || h.d.pmode != 0;
    }
    const bool f32 = C.inst[0].d.prec == 1;
    const bool fl = C.inst[0].d.flat != 0;  // one layout for the whole context
    const bool nt = C.sched.store_mode == 1;
    switch (model * 2 + (f32 ? 1 : 0)) {
        case 0: return pick_pair<double, 0>(pred, two, fl, nt, bulk, ibm, step);
        case 1: return pick_pair<float, 0>(pred, two, fl, nt, bulk, ibm, step);
        case 2: return pick_pair<double, 1>(pred, two, fl, nt, bulk, ibm, step);
        case 3: return pick_pair<float, 1>(pred, two, fl, nt, bulk, ibm, step);
        case 4: return pick_pair<double, 2>(pred, two, fl, nt, bulk, ibm, step);
        case 5: return pick_pair<float, 2>(pred, two, fl, nt, bulk, ibm, step);
        default: break;
    }
    if (f32)
        pick_pair<float, -1>(pred, two, fl, nt, bulk, ibm, step);
    else
        pick_pair<double, -1>(pred, two, fl, nt, bulk, ibm, step);
}

// Enqueues nsteps steps under the context's schedule.
//  fused: one k_step launch per step; the window workgroups come first, then one bulk
//    workgroup per remaining CU (a fused workgroup fills a CU's LDS) walking the rows.
//  two streams: the window kernel on a second stream, the bulk kernel on the context stream;
//    both depend on both kernels of the previous step through events, so the window
//    workgroups overlap the bulk tiles of the same step.
//  serial: window kernel then bulk kernel on the context stream.
// The device step counter the boundary kernel reads: slot `parity` holds the number (from 1)
// of the next step.  Launched outside any graph capture, before a batch of steps.
__global__ void k_set_step(long long *stepc, int parity, long long t) { stepc[parity] = t; }
int set_step_counter(Ctx &C) {
    if (!C.n_edge_blocks || !C.d_stepc) return 0;
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, C.stream, C.d_stepc, C.parity,
                       (long long)(C.step + 1));
    HIPCHK(hipGetLastError());
    return 0;
}

int launch_steps(Ctx &C, int nsteps, int diag_every, bool concurrent) {
    for (const HostInst &h : C.inst)
        if (h.d.bc == IBLBM_BC_TG_ANALYTICAL && (!h.d.tgc || !h.d.tgd || h.d.tg_nt < 1))
            return fail(IBLBM_EINVAL, "step: an analytic-boundary instance has no velocity "
                                      "table (iblbm_set_analytic_boundary)");
    StepKernel kb, ki, ks;
    pick_kernels(C, kb, ki, ks);
    const StepKernel ke = C.inst[0].d.prec == 1 ? (StepKernel)k_edge<float>
                                                : (StepKernel)k_edge<double>;
    auto edge = [&](const LaunchTable &T) {  // the closures that read the previous velocity
        if (C.n_edge_blocks)
            hipLaunchKernelGGL(ke, dim3(C.n_edge_blocks), dim3(256), 0, C.stream, T);
    };
    const int rows = C.n_tiles * TILE_Y;
    const int mode = concurrent ? effective_mode(C) : IBLBM_MODE_SERIAL;
    if (mode == IBLBM_MODE_FUSED) {
        const int bulk = std::max(1, std::min(C.n_cu - C.n_fast, (rows + NTF / 64 - 1) / (NTF / 64)));
        const int blocks = C.n_fast + bulk;
        for (int s = 0; s < nsteps; ++s) {
            const int64_t stepno = C.step + s + 1;
            const int diag = (diag_every > 0 && stepno % diag_every == 0) ? 1 : 0;
            LaunchTable T = make_table(C, (C.parity + s) & 1, diag);
            T.nrows = C.sched.step_rows;
            T.blocked = C.sched.step_order;
            T.bulk_delay = C.sched.bulk_delay;
            hipLaunchKernelGGL(ks, dim3(blocks), dim3(NTF), 0, C.stream, T);
            edge(T);
        }
        HIPCHK(hipGetLastError());
        return 0;
    }
    const int bt = C.sched.bulk_threads, wpb = bt / 64;
    const int full = (rows + wpb - 1) / wpb;  // one row per wavefront
    const int bblocks = C.sched.bulk_waves > 0
                            ? std::max(1, std::min(full, (C.sched.bulk_waves + wpb - 1) / wpb))
                            : full;
    const bool two = mode == IBLBM_MODE_TWO_STREAMS;
    if (two) {  // join the side stream to everything already queued on the main stream
        HIPCHK(hipEventRecord(C.ev_main, C.stream));
        HIPCHK(hipStreamWaitEvent(C.side, C.ev_main, 0));
    }
    for (int s = 0; s < nsteps; ++s) {
        const int64_t stepno = C.step + s + 1;
        const int diag = (diag_every > 0 && stepno % diag_every == 0) ? 1 : 0;
        LaunchTable T = make_table(C, (C.parity + s) & 1, diag);
        T.nrows = C.sched.bulk_rows;
        if (two) {
            if (s > 0) {
                HIPCHK(hipStreamWaitEvent(C.stream, C.ev_side, 0));
                HIPCHK(hipStreamWaitEvent(C.side, C.ev_main, 0));
            }
            hipLaunchKernelGGL(ki, dim3(C.n_ibm), dim3(NTI), 0, C.side, T);
            if (rows) hipLaunchKernelGGL(kb, dim3(bblocks), dim3(bt), 0, C.stream, T);
            edge(T);
            HIPCHK(hipEventRecord(C.ev_main, C.stream));
            HIPCHK(hipEventRecord(C.ev_side, C.side));
        } else {
            if (C.n_ibm) hipLaunchKernelGGL(ki, dim3(C.n_ibm), dim3(NTI), 0, C.stream, T);
            if (rows) hipLaunchKernelGGL(kb, dim3(bblocks), dim3(bt), 0, C.stream, T);
            edge(T);
        }
    }
    if (two) HIPCHK(hipStreamWaitEvent(C.stream, C.ev_side, 0));  // join back
    HIPCHK(hipGetLastError());
    return 0;
}

Ctx *as_ctx(void *p) { return static_cast<Ctx *>(p); }

bool valid_id(Ctx *C, int id) { return C && id >= 0 && id < (int)C->inst.size(); }

}  // namespace

extern "C" {

const char *iblbm_last_error(void) { return g_err.c_str(); }

__attribute__((visibility("hidden"))) int iblbm_set_op_error(int code, const char *m) {
    return fail(code, m);
}

int iblbm_device_name(int device, char *buf, int buflen) {
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, device));
    snprintf(buf, buflen, "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return 0;
}

int iblbm_ctx_create(int device, void **out) {
    if (!out) return fail(IBLBM_EINVAL, "ctx_create: null output pointer");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(IBLBM_EDEVICE, "ctx_create: no HIP device " + std::to_string(device));
    HIPCHK(hipSetDevice(device));
    Ctx *C = new Ctx();
    C->device = device;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            C->n_cu = prop.multiProcessorCount;
    }
    if (hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&C->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&C->ev0) != hipSuccess || hipEventCreate(&C->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&C->ev_diag[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&C->ev_diag[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&C->ev_mark[0]) != hipSuccess || hipEventCreate(&C->ev_mark[1]) != hipSuccess ||
        hipEventCreate(&C->ev_mark[2]) != hipSuccess || hipEventCreate(&C->ev_mark[3]) != hipSuccess ||
        hipEventCreateWithFlags(&C->ev_main, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&C->ev_side, hipEventDisableTiming) != hipSuccess) {
        delete C;
        return fail(IBLBM_EDEVICE, "ctx_create: stream/event creation failed");
    }
    *out = C;
    return 0;
}

int iblbm_ctx_destroy(void *p) {
    Ctx *C = as_ctx(p);
    if (!C) return 0;
    (void)hipSetDevice(C->device);
    (void)hipStreamSynchronize(C->stream);
    for (auto &g : C->graphs) (void)hipGraphExecDestroy(g.second);
    C->mem.release();
    (void)hipEventDestroy(C->ev0);
    (void)hipEventDestroy(C->ev1);
    for (int k = 0; k < 2; ++k) {
        (void)hipEventDestroy(C->ev_diag[k]);
        if (C->h_diag[k]) (void)hipHostFree(C->h_diag[k]);
    }
    for (int k = 0; k < 4; ++k) (void)hipEventDestroy(C->ev_mark[k]);
    if (C->h_bd) (void)hipHostFree(C->h_bd);
    if (C->h_bi) (void)hipHostFree(C->h_bi);
    (void)hipEventDestroy(C->ev_main);
    (void)hipEventDestroy(C->ev_side);
    (void)hipStreamDestroy(C->side);
    (void)hipStreamDestroy(C->stream);
    delete C;
    return 0;
}

int iblbm_instance_add(void *p, const iblbm_case *c, int *id) {
    Ctx *C = as_ctx(p);
    if (!C || !c || !id) return fail(IBLBM_EINVAL, "instance_add: null argument");
    if (c->nx < 3 || c->ny < 3) return fail(IBLBM_EINVAL, "instance_add: grid must be >= 3x3");
    if (c->collision < 0 || c->collision > 2)
        return fail(IBLBM_EKEY, "instance_add: unknown collision model");
    if (c->bc < 0 || c->bc > IBLBM_BC_TG_ANALYTICAL)
        return fail(IBLBM_EKEY, "instance_add: unknown bc_type");
    if (c->bc >= 2 && (c->nx < 7 || c->ny < 7))
        return fail(IBLBM_EINVAL, "instance_add: this boundary closure needs a grid >= 7x7");
    if (c->ibm < 0 || c->ibm > 3) return fail(IBLBM_EKEY, "instance_add: unknown ibm_method");
    if (c->delta < 0 || c->delta > 1) return fail(IBLBM_EKEY, "instance_add: unknown delta_type");
    if (c->precision != IBLBM_FP64 && c->precision != IBLBM_FP32_DEV)
        return fail(IBLBM_EINVAL, "instance_add: unknown precision");
    if (!C->inst.empty() && C->inst[0].d.prec != c->precision)
        return fail(IBLBM_EINVAL, "instance_add: all instances of a context share one precision");
    if (c->ibm != 0 && (c->lb_capacity < 1 || c->box_capacity < 1))
        return fail(IBLBM_EINVAL, "instance_add: IBM needs lb_capacity and box_capacity");
    if (c->ibm == IBLBM_IBM_MDF && (c->mdf_iterations < 1 || c->mdf_min_iterations < 1))
        return fail(IBLBM_EINVAL, "instance_add: mdf iterations must be >= 1");
    HIPCHK(hipSetDevice(C->device));
    HostInst h;
    h.c = *c;
    Inst &d = h.d;
    memset(&d, 0, sizeof d);
    d.nx = c->nx;
    d.ny = c->ny;
    d.pitch = (c->nx + 31) / 32 * 32;
    d.qs = (int64_t)d.pitch * d.ny;
    d.bc = c->bc;
    d.inc = c->incompressible;
    set_row_layout(d, false);  // the context's layout is applied once the instance is in
    CollParams &cp = d.cp;
    cp.model = c->collision;
    cp.dt = c->dt;
    cp.inv_tau = 1.0 / c->tau;             // bgk.py:28-29
    cp.guo_pref = 1.0 - 0.5 / c->tau;
    cp.s_plus = 1.0 / c->tau;              // trt.py:81-82
    cp.s_minus = 1.0 / c->tau_minus;
    cp.gpp = 1.0 - 0.5 * cp.s_plus;        // trt.py:42-43
    cp.gpm = 1.0 - 0.5 * cp.s_minus;
    for (int q = 0; q < 9; ++q) cp.S[q] = c->s_diag[q];
    d.prec = c->precision;
    d.cp32.inv_tau = (float)cp.inv_tau;
    d.cp32.s_plus = (float)cp.s_plus;
    d.cp32.s_minus = (float)cp.s_minus;
    for (int q = 0; q < 9; ++q) d.cp32.S[q] = (float)cp.S[q];
    d.in_u = c->in_u;
    d.dens = c->dens;
    if (c->bc == IBLBM_BC_DIRICHLET_NEUMANN)
        feq9(c->dens, c->in_u, 0.0, false, d.feq_bc);  // dirichlet_neumann.py:186-199
    else
        feq9(c->dens, 0.0, 0.0, c->incompressible != 0, d.feq_bc);  // settling_channel.py:137-142
    for (int q = 0; q < 9; ++q) d.gbc[q] = (float)(d.feq_bc[q] - W(q));
    d.ibm = c->ibm;
    d.delta = c->delta;
    d.dx = c->dx;
    d.dy = c->dy;
    d.omega = c->mdf_omega;
    d.tol = c->mdf_tolerance;
    d.n_iter = c->mdf_iterations;
    d.min_iter = c->mdf_min_iterations;
    d.fixed_body = c->fixed_body;
    d.np_part = 1;
    void *ptr;
    const size_t fbytes = (d.prec == 1 ? sizeof(float) : sizeof(double)) * 9 * (size_t)d.qs;
    for (int k = 0; k < 2; ++k) {
        if (C->mem.alloc(&ptr, fbytes) != hipSuccess)
            return fail(IBLBM_ENOMEM, "instance_add: population buffers do not fit");
        d.f[k] = (double *)ptr;
        HIPCHK(hipMemsetAsync(ptr, 0, fbytes, C->stream));
    }
    const int L = c->ibm ? c->lb_capacity : 1, B = c->ibm ? c->box_capacity : 1;
    d.dfc_scale = 1.0;
    const bool ok = alloc_ibm_scratch(d, L, B, [&](size_t bytes) -> void * {
        void *q = nullptr;
        return C->mem.alloc(&q, bytes) == hipSuccess ? q : nullptr;
    });
    if (!ok) return fail(IBLBM_ENOMEM, "instance_add: IBM scratch allocation failed");
    HIPCHK(hipMemsetAsync(d.s.fibx, 0, sizeof(double) * B, C->stream));
    HIPCHK(hipMemsetAsync(d.s.fiby, 0, sizeof(double) * B, C->stream));
    HIPCHK(hipMemsetAsync(d.stats, 0, sizeof(double) * 4, C->stream));
    HIPCHK(hipMemsetAsync(d.diag, 0, sizeof(double) * 8, C->stream));
    if (c->bc >= 2) {  // ring velocities (two steps) and the outer ring's populations
        d.nring = ring_count(d.nx, d.ny);
        if (C->mem.alloc(&ptr, sizeof(double) * 4 * (size_t)d.nring) != hipSuccess)
            return fail(IBLBM_ENOMEM, "instance_add: boundary buffers do not fit");
        d.side = (double *)ptr;
        HIPCHK(hipMemsetAsync(ptr, 0, sizeof(double) * 4 * (size_t)d.nring, C->stream));
        if (C->mem.alloc(&ptr, sizeof(double) * 9 * (size_t)d.nring) != hipSuccess)
            return fail(IBLBM_ENOMEM, "instance_add: boundary buffers do not fit");
        d.edge_fs = (double *)ptr;
        HIPCHK(hipMemsetAsync(ptr, 0, sizeof(double) * 9 * (size_t)d.nring, C->stream));
    }
    d.Lb = 0;
    if (c->ibm != 0) d.ibm = 0;  // inactive until markers arrive
    h.c.ibm = c->ibm;
    C->inst.push_back(h);
    C->tables_dirty = true;
    apply_row_layout(*C);
    *id = (int)C->inst.size() - 1;
    return 0;
}

int iblbm_marker_box(const double *Lx, const double *Ly, int Lb, double dx, double dy, int nx,
                     int ny, int delta, int box[4]) {
    if (!Lx || !Ly || !box || Lb < 1 || nx < 1 || ny < 1 || !(dx > 0.0) || !(dy > 0.0))
        return fail(IBLBM_EINVAL, "marker_box: bad argument");
    const int a = delta == IBLBM_DELTA_HAT ? 1 : 2;
    int64_t x0 = INT64_MAX, x1 = INT64_MIN, y0 = INT64_MAX, y1 = INT64_MIN;
    for (int k = 0; k < Lb; ++k) {
        if (!std::isfinite(Lx[k]) || !std::isfinite(Ly[k]))
            return fail(IBLBM_EINVAL, "marker_box: non-finite marker coordinate");
        const int64_t ix0 = (int64_t)std::floor(Lx[k] / dx + 0.5);  // df.py:229-230
        const int64_t iy0 = (int64_t)std::floor(Ly[k] / dy + 0.5);
        x0 = std::min(x0, ix0 - a);
        x1 = std::max(x1, ix0 + a);
        y0 = std::min(y0, iy0 - a);
        y1 = std::max(y1, iy0 + a);
    }
    auto cl = [](int64_t v, int64_t hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
    x0 = cl(x0, nx - 1);
    x1 = cl(x1, nx - 1);
    y0 = cl(y0, ny - 1);
    y1 = cl(y1, ny - 1);
    box[0] = (int)x0;
    box[1] = (int)y0;
    box[2] = (int)(x1 - x0 + 1);
    box[3] = (int)(y1 - y0 + 1);
    return 0;
}

int iblbm_set_markers(void *p, int id, const double *Lx, const double *Ly, const double *ud,
                      int Lb, double Larea) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id)) return fail(IBLBM_EINVAL, "set_markers: bad instance id");
    HostInst &h = C->inst[id];
    Inst &d = h.d;
    if (h.c.ibm == 0) return fail(IBLBM_EINVAL, "set_markers: instance has no IBM method");
    if (!Lx || !Ly || !ud) return fail(IBLBM_EINVAL, "set_markers: null marker array");
    if (Lb < 1 || Lb > d.lb_cap)
        return fail(IBLBM_EINVAL, "set_markers: Lb outside [1, lb_capacity]");
    int box[4];
    int brc = iblbm_marker_box(Lx, Ly, Lb, d.dx, d.dy, d.nx, d.ny, d.delta, box);
    if (brc) return brc;
    const int bx0 = box[0], by0 = box[1], bw = box[2], bh = box[3];
    if (bw < 1 || bh < 1 || (int64_t)bw * bh > d.box_cap)
        return fail(IBLBM_EINVAL, "set_markers: window box exceeds box_capacity");
    if (bx0 < 0 || by0 < 0 || bx0 + bw > d.nx || by0 + bh > d.ny)
        return fail(IBLBM_EINVAL, "set_markers: window box outside the grid");
    if (d.bc >= 2 && (bx0 < 3 || by0 < 3 || bx0 + bw > d.nx - 3 || by0 + bh > d.ny - 3))
        return fail(IBLBM_EINVAL, "set_markers: with this boundary closure the window must stay "
                                  "three nodes away from every wall");
    if (!(Larea > 0.0)) return fail(IBLBM_EINVAL, "set_markers: Larea must be positive");
    if (d.np_part > 1 && d.poff[d.np_part] != Lb)
        return fail(IBLBM_EINVAL, "set_markers: Lb differs from the particles' marker total");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipMemcpyAsync((void *)d.Lx, Lx, sizeof(double) * Lb, hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipMemcpyAsync((void *)d.Ly, Ly, sizeof(double) * Lb, hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipMemcpyAsync((void *)d.ud, ud, sizeof(double) * 2 * Lb, hipMemcpyHostToDevice,
                          C->stream));
    const bool was_active = d.ibm != 0;
    d.ibm = h.c.ibm;
    d.Lb = Lb;
    d.Larea = Larea;
    d.bx0 = bx0;
    d.by0 = by0;
    d.bw = bw;
    d.bh = bh;
    h.Lb = Lb;
    h.Larea = Larea;
    if (!was_active || C->tables_dirty) {
        C->tables_dirty = true;
        int rc = ensure_tables(*C);
        if (rc) return rc;
    } else {
        int rc = upload_inst(*C, id);
        if (rc) return rc;
    }
    // window tables, DFC lambda and a zeroed box force, ordered after every queued step
    hipLaunchKernelGGL(k_tables, dim3(1), dim3(NTI), 0, C->stream, C->d_insts + id);
    HIPCHK(hipGetLastError());
    // the touched-node count and path flag go back into the host descriptor, so later
    // uploads keep them; the synchronisation also retires the pageable marker copies
    static_assert(offsetof(Inst, fast) == offsetof(Inst, nT) + sizeof(int), "nT, fast adjacent");
    const int fast_before = d.fast;
    HIPCHK(hipMemcpyAsync(&d.nT, (const char *)(C->d_insts + id) + offsetof(Inst, nT),
                          2 * sizeof(int), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    if (d.fast != fast_before) {  // the window workgroup changes kernel: rebuild the lists
        C->tables_dirty = true;
        int rc = ensure_tables(*C);
        if (rc) return rc;
    }
    return 0;
}

namespace {

// Pinned host and device staging of the batched calls, grown geometrically (a superseded
// device buffer stays with the context's allocations until it is destroyed).  Every batched
// call ends with a stream synchronisation, so no copy from or into it is pending here.  The
// capacity is cleared before a buffer is released, so a failed allocation leaves none.
int grow_batch(Ctx &C, size_t nd, size_t ni) {
    if (nd > C.bd_cap) {
        const size_t cap = nd < 4096 ? 4096 : 2 * nd;
        C.bd_cap = 0;
        if (C.h_bd) HIPCHK(hipHostFree(C.h_bd));
        C.h_bd = nullptr;
        HIPCHK(hipHostMalloc((void **)&C.h_bd, sizeof(double) * cap, hipHostMallocDefault));
        void *q;
        if (C.mem.alloc(&q, sizeof(double) * cap) != hipSuccess)
            return fail(IBLBM_ENOMEM, "batch staging does not fit");
        C.d_bd = (double *)q;
        C.bd_cap = cap;
    }
    if (ni > C.bi_cap) {
        const size_t cap = ni < 1024 ? 1024 : 2 * ni;
        C.bi_cap = 0;
        if (C.h_bi) HIPCHK(hipHostFree(C.h_bi));
        C.h_bi = nullptr;
        HIPCHK(hipHostMalloc((void **)&C.h_bi, sizeof(int) * cap, hipHostMallocDefault));
        void *q;
        if (C.mem.alloc(&q, sizeof(int) * cap) != hipSuccess)
            return fail(IBLBM_ENOMEM, "batch staging does not fit");
        C.d_bi = (int *)q;
        C.bi_cap = cap;
    }
    return 0;
}

}  // namespace

int iblbm_set_markers_batch(void *p, int n, const int *ids, const int *lb, const double *Lx,
                            const double *Ly, const double *ud, const double *larea,
                            const double *center) {
    Ctx *C = as_ctx(p);
    if (!C || n < 1 || !ids || !lb || !Lx || !Ly || !ud || !larea)
        return fail(IBLBM_EINVAL, "set_markers_batch: need n >= 1 and every array");
    // every instance is checked before any state changes
    std::vector<int> box(4 * (size_t)n), off((size_t)n + 1, 0);
    std::vector<char> seen(C->inst.size(), 0);
    for (int b = 0; b < n; ++b) {
        const int id = ids[b];
        if (!valid_id(C, id)) return fail(IBLBM_EINVAL, "set_markers_batch: bad instance id");
        if (seen[id]) return fail(IBLBM_EINVAL, "set_markers_batch: instance listed twice");
        seen[id] = 1;
        const HostInst &h = C->inst[id];
        const Inst &d = h.d;
        if (h.c.ibm == 0)
            return fail(IBLBM_EINVAL, "set_markers_batch: instance has no IBM method");
        if (lb[b] < 1 || lb[b] > d.lb_cap)
            return fail(IBLBM_EINVAL, "set_markers_batch: Lb outside [1, lb_capacity]");
        if (d.np_part > 1 && d.poff[d.np_part] != lb[b])
            return fail(IBLBM_EINVAL,
                        "set_markers_batch: Lb differs from the particles' marker total");
        if (!(larea[b] > 0.0)) return fail(IBLBM_EINVAL, "set_markers_batch: Larea must be positive");
        int *bx = &box[4 * (size_t)b];
        int brc = iblbm_marker_box(Lx + off[b], Ly + off[b], lb[b], d.dx, d.dy, d.nx, d.ny,
                                   d.delta, bx);
        if (brc) return brc;
        if (bx[2] < 1 || bx[3] < 1 || (int64_t)bx[2] * bx[3] > d.box_cap)
            return fail(IBLBM_EINVAL, "set_markers_batch: window box exceeds box_capacity");
        if (bx[0] < 0 || bx[1] < 0 || bx[0] + bx[2] > d.nx || bx[1] + bx[3] > d.ny)
            return fail(IBLBM_EINVAL, "set_markers_batch: window box outside the grid");
        off[b + 1] = off[b] + lb[b];
    }
    const int total = off[n];
    HIPCHK(hipSetDevice(C->device));
    // int staging: ids[n], off[n + 1], then the (nT, fast) pairs k_tables_batch writes back
    const size_t n_int = 4 * (size_t)n + 1;
    int rc = grow_batch(*C, 4 * (size_t)total, n_int);
    if (rc) return rc;
    if (C->bi_cap < n_int || C->bd_cap < 4 * (size_t)total)
        return fail(IBLBM_EDEVICE, "set_markers_batch: staging smaller than the call's layout");
    std::memcpy(C->h_bd, Lx, sizeof(double) * total);
    std::memcpy(C->h_bd + total, Ly, sizeof(double) * total);
    std::memcpy(C->h_bd + 2 * total, ud, sizeof(double) * 2 * total);
    std::memcpy(C->h_bi, ids, sizeof(int) * n);
    std::memcpy(C->h_bi + n, off.data(), sizeof(int) * (n + 1));
    bool activated = false;
    for (int b = 0; b < n; ++b) {
        HostInst &h = C->inst[ids[b]];
        Inst &d = h.d;
        activated |= d.ibm == 0;
        d.ibm = h.c.ibm;
        d.Lb = lb[b];
        d.Larea = larea[b];
        d.bx0 = box[4 * b];
        d.by0 = box[4 * b + 1];
        d.bw = box[4 * b + 2];
        d.bh = box[4 * b + 3];
        h.Lb = lb[b];
        h.Larea = larea[b];
        if (center) {
            d.cxd = center[2 * b];
            d.cyd = center[2 * b + 1];
            d.cxc = d.cxd / d.dx;
            d.cyc = d.cyd / d.dy;
        }
    }
    if (activated) C->tables_dirty = true;
    std::vector<Inst> hd;  // the pageable source lives until the synchronisation below
    if (C->tables_dirty) {
        rc = ensure_tables(*C);  // uploads every descriptor
        if (rc) return rc;
    } else {
        hd.resize(C->inst.size());
        for (size_t k = 0; k < hd.size(); ++k) hd[k] = C->inst[k].d;
        HIPCHK(hipMemcpyAsync(C->d_insts, hd.data(), sizeof(Inst) * hd.size(),
                              hipMemcpyHostToDevice, C->stream));
    }
    HIPCHK(hipMemcpyAsync(C->d_bd, C->h_bd, sizeof(double) * 4 * total, hipMemcpyHostToDevice,
                          C->stream));
    HIPCHK(hipMemcpyAsync(C->d_bi, C->h_bi, sizeof(int) * (2 * n + 1), hipMemcpyHostToDevice,
                          C->stream));
    int *d_ntf = C->d_bi + 2 * n + 1;
    hipLaunchKernelGGL(k_scatter_markers, dim3(n), dim3(256), 0, C->stream, C->d_insts, C->d_bi,
                       C->d_bi + n, C->d_bd, total);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_tables_batch, dim3(n), dim3(NTI), 0, C->stream, C->d_insts, C->d_bi,
                       d_ntf);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(C->h_bi + 2 * n + 1, d_ntf, sizeof(int) * 2 * n, hipMemcpyDeviceToHost,
                          C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    bool path_changed = false;
    for (int b = 0; b < n; ++b) {
        Inst &d = C->inst[ids[b]].d;
        const int fast_before = d.fast;
        d.nT = C->h_bi[2 * n + 1 + 2 * b];
        d.fast = C->h_bi[2 * n + 2 + 2 * b];
        path_changed |= d.fast != fast_before;
    }
    if (path_changed) {  // a window workgroup changes kernel: rebuild the launch lists
        C->tables_dirty = true;
        rc = ensure_tables(*C);
        if (rc) return rc;
    }
    return 0;
}

int iblbm_download_marker_forces_batch(void *p, int n, const int *ids, double *F,
                                       int64_t capacity) {
    Ctx *C = as_ctx(p);
    if (!C || n < 1 || !ids || !F)
        return fail(IBLBM_EINVAL, "download_marker_forces_batch: need n >= 1 and every array");
    std::vector<int> off((size_t)n + 1, 0);
    for (int b = 0; b < n; ++b) {
        if (!valid_id(C, ids[b]))
            return fail(IBLBM_EINVAL, "download_marker_forces_batch: bad instance id");
        const Inst &d = C->inst[ids[b]].d;
        const bool dfc = d.ibm == IBLBM_IBM_DFC;
        const bool df_multi = d.df_multi != 0;
        if (!dfc && !df_multi)
            return fail(IBLBM_EINVAL, "download_marker_forces_batch: needs DFC instances or DF "
                                      "windows of several particles");
        off[b + 1] = off[b] + d.Lb;
    }
    const int total = off[n];
    if ((int64_t)total > capacity)
        return fail(IBLBM_EINVAL, "download_marker_forces_batch: F holds " +
                                      std::to_string(capacity) + " marker rows, the instances have " +
                                      std::to_string(total));
    HIPCHK(hipSetDevice(C->device));
    int rc = grow_batch(*C, 2 * (size_t)total, 2 * (size_t)n + 1);
    if (rc) return rc;
    std::memcpy(C->h_bi, ids, sizeof(int) * n);
    std::memcpy(C->h_bi + n, off.data(), sizeof(int) * (n + 1));
    HIPCHK(hipMemcpyAsync(C->d_bi, C->h_bi, sizeof(int) * (2 * n + 1), hipMemcpyHostToDevice,
                          C->stream));
    hipLaunchKernelGGL(k_gather_forces, dim3(n), dim3(256), 0, C->stream, C->d_insts, C->d_bi,
                       C->d_bi + n, C->d_bd);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(C->h_bd, C->d_bd, sizeof(double) * 2 * total, hipMemcpyDeviceToHost,
                          C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    std::memcpy(F, C->h_bd, sizeof(double) * 2 * total);
    return 0;
}

int iblbm_set_particles(void *p, int id, int np, const int *lb, const double *larea) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id)) return fail(IBLBM_EINVAL, "set_particles: bad instance id");
    if (np < 1 || np > IBM_MAXP || !lb || !larea)
        return fail(IBLBM_EINVAL, "set_particles: need 1 <= np <= 8 and both arrays");
    HostInst &h = C->inst[id];
    Inst &d = h.d;
    int off = 0;
    for (int q = 0; q < np; ++q) {
        if (lb[q] < 1 || !(larea[q] > 0.0))
            return fail(IBLBM_EINVAL, "set_particles: markers and Larea must be positive");
        off += lb[q];
    }
    if (off > d.lb_cap) return fail(IBLBM_EINVAL, "set_particles: markers exceed lb_capacity");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    if (np > 1 && (h.c.ibm == IBLBM_IBM_DF || h.c.ibm == IBLBM_IBM_MDF) && np > h.pfb_np) {
        void *q = nullptr;
        const size_t bytes = sizeof(double) * 2 * (size_t)np * d.box_cap;
        if (C->mem.alloc(&q, bytes) != hipSuccess)
            return fail(IBLBM_ENOMEM, "set_particles: per-particle window fields do not fit");
        HIPCHK(hipMemsetAsync(q, 0, bytes, C->stream));
        d.s.pfb = (double *)q;
        h.pfb_np = np;
    }
    d.np_part = np;
    d.df_multi = h.c.ibm == IBLBM_IBM_DF ? 1 : 0;
    d.poff[0] = 0;
    for (int q = 0; q < np; ++q) {
        d.poff[q + 1] = d.poff[q] + lb[q];
        d.plarea[q] = larea[q];
    }
    C->tables_dirty = true;
    return ensure_tables(*C);
}

int iblbm_download_particle_fib(void *p, int id, int part, double *fibx, double *fiby,
                                int box[4]) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !fibx || !fiby || !box)
        return fail(IBLBM_EINVAL, "download_particle_fib: bad argument");
    const HostInst &h = C->inst[id];
    const Inst &d = h.d;
    if (d.ibm != IBLBM_IBM_MDF || d.np_part < 2 || part < 0 || part >= d.np_part || !d.s.pfb)
        return fail(IBLBM_EINVAL, "download_particle_fib: needs a coupled MDF window and a "
                                  "particle index below np");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    const int nb = d.bw * d.bh;
    HIPCHK(hipMemcpy(fibx, d.s.pfb + (size_t)(2 * part) * d.box_cap, sizeof(double) * nb,
                     hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fiby, d.s.pfb + (size_t)(2 * part + 1) * d.box_cap, sizeof(double) * nb,
                     hipMemcpyDeviceToHost));
    box[0] = d.bx0;
    box[1] = d.by0;
    box[2] = d.bw;
    box[3] = d.bh;
    return 0;
}

int iblbm_download_marker_force(void *p, int id, double *F) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !F) return fail(IBLBM_EINVAL, "download_marker_force: bad argument");
    const Inst &d = C->inst[id].d;
    if (!d.df_multi)
        return fail(IBLBM_EINVAL, "download_marker_force: needs a DF window of a particle list "
                                  "(iblbm_set_particles)");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipMemcpy(F, d.s.mf, sizeof(double) * 2 * d.Lb, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_set_torque_center(void *p, int id, double cx, double cy) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id)) return fail(IBLBM_EINVAL, "set_torque_center: bad instance id");
    Inst &d = C->inst[id].d;
    d.cxd = cx;
    d.cyd = cy;
    d.cxc = cx / d.dx;
    d.cyc = cy / d.dy;
    if (C->tables_dirty) return ensure_tables(*C);
    int rc = upload_inst(*C, id);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(C->stream));
    return 0;
}

int iblbm_upload_f(void *p, int id, const double *f_aos) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !f_aos) return fail(IBLBM_EINVAL, "upload_f: bad argument");
    HIPCHK(hipSetDevice(C->device));
    int rc = ensure_tables(*C);
    if (rc) return rc;
    const Inst &d = C->inst[id].d;
    const int64_t N = (int64_t)d.nx * d.ny;
    void *tmp;
    HIPCHK(hipMalloc(&tmp, sizeof(double) * 9 * N));
    hipError_t e = hipMemcpy(tmp, f_aos, sizeof(double) * 9 * N, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (d.prec == 1)
            hipLaunchKernelGGL(k_aos_to_soa<float>, dim3((unsigned)((N + 255) / 256)), dim3(256),
                               0, C->stream, (const double *)tmp, (float *)d.f[C->parity], d.nx,
                               d.ny, d.pitch, d.qs);
        else
            hipLaunchKernelGGL(k_aos_to_soa<double>, dim3((unsigned)((N + 255) / 256)), dim3(256),
                               0, C->stream, (const double *)tmp, d.f[C->parity], d.nx, d.ny,
                               d.pitch, d.qs);
        e = hipStreamSynchronize(C->stream);
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(IBLBM_EDEVICE, std::string("upload_f: ") + hipGetErrorString(e));
    return 0;
}

int iblbm_download_f(void *p, int id, double *f_aos) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !f_aos) return fail(IBLBM_EINVAL, "download_f: bad argument");
    HIPCHK(hipSetDevice(C->device));
    const Inst &d = C->inst[id].d;
    const int64_t N = (int64_t)d.nx * d.ny;
    void *tmp;
    HIPCHK(hipMalloc(&tmp, sizeof(double) * 9 * N));
    if (d.prec == 1)
        hipLaunchKernelGGL(k_soa_to_aos<float>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                           C->stream, (const float *)d.f[C->parity], (double *)tmp, d.nx, d.ny,
                           d.pitch, d.qs);
    else
        hipLaunchKernelGGL(k_soa_to_aos<double>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                           C->stream, (const double *)d.f[C->parity], (double *)tmp, d.nx, d.ny,
                           d.pitch, d.qs);
    hipError_t e = hipStreamSynchronize(C->stream);
    if (e == hipSuccess) e = hipMemcpy(f_aos, tmp, sizeof(double) * 9 * N, hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(IBLBM_EDEVICE, std::string("download_f: ") + hipGetErrorString(e));
    return 0;
}

int iblbm_download_state(void *p, int id, double *fstar, double *ro, double *U) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !fstar || !ro || !U)
        return fail(IBLBM_EINVAL, "download_state: bad argument");
    if (C->step == 0) return fail(IBLBM_EINVAL, "download_state: no step has run yet");
    HIPCHK(hipSetDevice(C->device));
    const Inst &d = C->inst[id].d;
    const int64_t N = (int64_t)d.nx * d.ny;
    void *tmp;
    HIPCHK(hipMalloc(&tmp, sizeof(double) * 12 * N));
    double *tf = (double *)tmp, *tr = tf + 9 * N, *tu = tr + N;
    if (d.prec == 1)
        hipLaunchKernelGGL(k_state<float>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                           C->stream, (const Inst *)(C->d_insts + id), C->parity ^ 1, 0, 0, d.nx,
                           d.ny, tf, tr, tu);
    else
        hipLaunchKernelGGL(k_state<double>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                           C->stream, (const Inst *)(C->d_insts + id), C->parity ^ 1, 0, 0, d.nx,
                           d.ny, tf, tr, tu);
    hipError_t e = hipStreamSynchronize(C->stream);
    if (e == hipSuccess) e = hipMemcpy(fstar, tf, sizeof(double) * 9 * N, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ro, tr, sizeof(double) * N, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(U, tu, sizeof(double) * 2 * N, hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(IBLBM_EDEVICE, std::string("download_state: ") + hipGetErrorString(e));
    return 0;
}

int iblbm_download_state_box(void *p, int id, int x0, int y0, int w, int h, double *ro,
                             double *U) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !ro || !U || w < 1 || h < 1)
        return fail(IBLBM_EINVAL, "download_state_box: bad argument");
    const Inst &d = C->inst[id].d;
    if (x0 < 0 || y0 < 0 || x0 + w > d.nx || y0 + h > d.ny)
        return fail(IBLBM_EINVAL, "download_state_box: the box leaves the grid");
    if (C->step == 0) return fail(IBLBM_EINVAL, "download_state_box: no step has run yet");
    HIPCHK(hipSetDevice(C->device));
    const int64_t N = (int64_t)w * h;
    if ((size_t)(3 * N) > C->box_cap) {  // kept for the next calls (a larger box adds one)
        void *q = nullptr;
        if (C->mem.alloc(&q, sizeof(double) * 3 * N) != hipSuccess)
            return fail(IBLBM_ENOMEM, "download_state_box: staging");
        C->d_box = (double *)q;
        C->box_cap = (size_t)(3 * N);
    }
    double *tr = C->d_box, *tu = tr + N;
    if (d.prec == 1)
        hipLaunchKernelGGL(k_state<float>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                           C->stream, (const Inst *)(C->d_insts + id), C->parity ^ 1, x0, y0, w,
                           h, nullptr, tr, tu);
    else
        hipLaunchKernelGGL(k_state<double>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                           C->stream, (const Inst *)(C->d_insts + id), C->parity ^ 1, x0, y0, w,
                           h, nullptr, tr, tu);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipMemcpy(ro, tr, sizeof(double) * N, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(U, tu, sizeof(double) * 2 * N, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_download_fib(void *p, int id, double *fib) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !fib) return fail(IBLBM_EINVAL, "download_fib: bad argument");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    const HostInst &h = C->inst[id];
    const Inst &d = h.d;
    const int64_t N = (int64_t)d.nx * d.ny;
    memset(fib, 0, sizeof(double) * 2 * N);
    if (d.ibm == 0 || d.ibm == IBLBM_IBM_DFC) return 0;
    const int nb = d.bw * d.bh;
    std::vector<double> fx(nb), fy(nb);
    HIPCHK(hipMemcpy(fx.data(), d.s.fibx, sizeof(double) * nb, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fy.data(), d.s.fiby, sizeof(double) * nb, hipMemcpyDeviceToHost));
    for (int sl = 0; sl < nb; ++sl) {
        const int64_t n = (int64_t)(d.by0 + sl / d.bw) * d.nx + d.bx0 + sl % d.bw;
        fib[2 * n] = fx[sl];
        fib[2 * n + 1] = fy[sl];
    }
    return 0;
}

int iblbm_download_dfc_force(void *p, int id, double *q) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !q) return fail(IBLBM_EINVAL, "download_dfc_force: bad argument");
    const HostInst &h = C->inst[id];
    if (h.d.ibm != IBLBM_IBM_DFC) return fail(IBLBM_EINVAL, "download_dfc_force: not a DFC instance");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipMemcpy(q, h.d.s.q, sizeof(double) * 2 * h.Lb, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_download_window(void *p, int id, double *fibx, double *fiby, int box[4]) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !fibx || !fiby || !box)
        return fail(IBLBM_EINVAL, "download_window: bad argument");
    const Inst &d = C->inst[id].d;
    if (d.ibm == 0) return fail(IBLBM_EINVAL, "download_window: instance has no markers");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    const int nb = d.bw * d.bh;
    HIPCHK(hipMemcpy(fibx, d.s.fibx, sizeof(double) * nb, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fiby, d.s.fiby, sizeof(double) * nb, hipMemcpyDeviceToHost));
    box[0] = d.bx0;
    box[1] = d.by0;
    box[2] = d.bw;
    box[3] = d.bh;
    return 0;
}

int iblbm_step(void *p, int nsteps, int diag_every) {
    Ctx *C = as_ctx(p);
    if (!C || nsteps < 0) return fail(IBLBM_EINVAL, "step: bad argument");
    if (C->inst.empty()) return fail(IBLBM_EINVAL, "step: no instance");
    HIPCHK(hipSetDevice(C->device));
    int rc = ensure_tables(*C);
    if (rc) return rc;
    rc = set_step_counter(*C);
    if (rc) return rc;
    HIPCHK(hipEventRecord(C->ev0, C->stream));
    rc = launch_steps(*C, nsteps, diag_every, C->concurrent);
    if (rc) return rc;
    HIPCHK(hipEventRecord(C->ev1, C->stream));
    if (diag_every > 0) {
        const int64_t last = ((C->step + nsteps) / diag_every) * diag_every;
        if (last > C->step) C->diag_step = last;
    }
    C->step += nsteps;
    C->parity = (C->parity + nsteps) & 1;
    C->last_steps = nsteps;
    return 0;
}

int iblbm_step_persistent(void *p, int nsteps, int diag_every) {
    Ctx *C = as_ctx(p);
    if (!C || nsteps < 1) return fail(IBLBM_EINVAL, "step_persistent: bad argument");
    if (C->inst.empty()) return fail(IBLBM_EINVAL, "step_persistent: no instance");
    HIPCHK(hipSetDevice(C->device));
    int rc = ensure_tables(*C);
    if (rc) return rc;
    if (effective_mode(*C) != IBLBM_MODE_FUSED || C->n_edge_blocks)
        return fail(IBLBM_EINVAL, "step_persistent: needs the fused launch (every window on the "
                                  "fused path) and no boundary kernel");
    const PersistKernel kern = pick_persist(*C);
    if (!kern)
        return fail(IBLBM_EINVAL, "step_persistent: needs one collision model and the tile "
                                  "row layout");
    if (!C->d_bar) {
        void *q = nullptr;
        if (C->mem.alloc(&q, 2 * sizeof(unsigned)) != hipSuccess)
            return fail(IBLBM_ENOMEM, "step_persistent: barrier counter");
        C->d_bar = (unsigned *)q;
    }
    HIPCHK(hipMemsetAsync(C->d_bar, 0, 2 * sizeof(unsigned), C->stream));
    const int rows = C->n_tiles * TILE_Y;
    const int bulk = std::max(1, std::min(C->n_cu - C->n_fast, (rows + NTF / 64 - 1) / (NTF / 64)));
    const int blocks = C->n_fast + bulk;  // one workgroup per CU (the window's LDS fills one)
    LaunchTable T = make_table(*C, C->parity, 0);
    T.nrows = C->sched.step_rows;
    T.blocked = C->sched.step_order;
    T.bulk_delay = C->sched.bulk_delay;
    PersistArgs A;
    A.bar = C->d_bar;
    A.err = (int *)(C->d_bar + 1);
    A.nsteps = nsteps;
    A.diag_every = diag_every;
    A.step0 = C->step;
    A.parity0 = C->parity;
    A.pad = 0;
    void *args[] = {&T, &A};
    HIPCHK(hipEventRecord(C->ev0, C->stream));
    HIPCHK(hipLaunchCooperativeKernel((const void *)kern, dim3(blocks), dim3(NTF), args, 0,
                                      C->stream));
    HIPCHK(hipEventRecord(C->ev1, C->stream));
    if (diag_every > 0) {
        const int64_t last = ((C->step + nsteps) / diag_every) * diag_every;
        if (last > C->step) C->diag_step = last;
    }
    C->step += nsteps;
    C->parity = (C->parity + nsteps) & 1;
    C->last_steps = nsteps;
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, C->d_bar + 1, sizeof err, hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    if (err) return fail(IBLBM_EDEVICE, "step_persistent: a grid barrier timed out");
    return 0;
}

int iblbm_graph_run(void *p, int nsteps, int diag_every) {
    Ctx *C = as_ctx(p);
    if (!C || nsteps < 1) return fail(IBLBM_EINVAL, "graph_run: bad argument");
    if (C->inst.empty()) return fail(IBLBM_EINVAL, "graph_run: no instance");
    HIPCHK(hipSetDevice(C->device));
    int rc = ensure_tables(*C);
    if (rc) return rc;
    const int phase = diag_every > 0 ? (int)(C->step % diag_every) : 0;
    auto key = std::make_tuple(nsteps, diag_every * 2 + (C->concurrent ? 1 : 0), C->parity, phase);
    auto it = C->graphs.find(key);
    hipGraphExec_t exec;
    if (it == C->graphs.end()) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(C->stream, hipStreamCaptureModeThreadLocal));
        rc = launch_steps(*C, nsteps, diag_every, C->concurrent);
        hipError_t e = hipStreamEndCapture(C->stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) return fail(IBLBM_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(IBLBM_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
        C->graphs[key] = exec;
    } else {
        exec = it->second;
    }
    rc = set_step_counter(*C);
    if (rc) return rc;
    HIPCHK(hipEventRecord(C->ev0, C->stream));
    HIPCHK(hipGraphLaunch(exec, C->stream));
    HIPCHK(hipEventRecord(C->ev1, C->stream));
    if (diag_every > 0) {
        const int64_t last = ((C->step + nsteps) / diag_every) * diag_every;
        if (last > C->step) C->diag_step = last;
    }
    C->step += nsteps;
    C->parity = (C->parity + nsteps) & 1;
    C->last_steps = nsteps;
    return 0;
}

int iblbm_set_concurrent(void *p, int on) {
    Ctx *C = as_ctx(p);
    if (!C) return fail(IBLBM_EINVAL, "set_concurrent: null context");
    C->concurrent = on != 0;
    return 0;
}

int iblbm_sync(void *p) {
    Ctx *C = as_ctx(p);
    if (!C) return fail(IBLBM_EINVAL, "sync: null context");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    return 0;
}

int iblbm_step_mode(void *p, int *mode) {
    Ctx *C = as_ctx(p);
    if (!C || !mode) return fail(IBLBM_EINVAL, "step_mode: bad argument");
    int rc = ensure_tables(*C);
    if (rc) return rc;
    *mode = effective_mode(*C);
    return 0;
}

int iblbm_get_schedule(void *p, iblbm_schedule *out) {
    Ctx *C = as_ctx(p);
    if (!C || !out) return fail(IBLBM_EINVAL, "get_schedule: bad argument");
    *out = C->sched;
    return 0;
}

int iblbm_row_layout_in_use(void *p, int *layout) {
    Ctx *C = as_ctx(p);
    if (!C || !layout) return fail(IBLBM_EINVAL, "row_layout_in_use: bad argument");
    *layout = C->inst.empty() ? 0 : C->inst[0].d.flat;
    return 0;
}

int iblbm_set_schedule(void *p, const iblbm_schedule *in) {
    Ctx *C = as_ctx(p);
    if (!C || !in) return fail(IBLBM_EINVAL, "set_schedule: bad argument");
    if (in->mode < IBLBM_MODE_AUTO || in->mode > IBLBM_MODE_SERIAL)
        return fail(IBLBM_EINVAL, "set_schedule: unknown mode");
    if (in->bulk_threads < 64 || in->bulk_threads > 1024 || in->bulk_threads % 64)
        return fail(IBLBM_EINVAL, "set_schedule: bulk_threads must be a multiple of 64 in [64, 1024]");
    auto rows_ok = [](int r) { return r >= 1 && r <= 4; };
    if (!rows_ok(in->bulk_rows) || !rows_ok(in->step_rows))
        return fail(IBLBM_EINVAL, "set_schedule: rows in flight must be 1, 2, 3 (one row, "
                                  "pipelined) or 4");
    if (in->bulk_waves < 0) return fail(IBLBM_EINVAL, "set_schedule: bulk_waves < 0");
    if (in->step_order < -1 || in->step_order > 64)
        return fail(IBLBM_EINVAL, "set_schedule: step_order must be -1 (workgroup queue) or in "
                                  "[0, 64]");
    if (in->row_layout != 0 && in->row_layout != 1)
        return fail(IBLBM_EINVAL, "set_schedule: row_layout must be 0 (tiles) or 1 (flat)");
    if (in->store_mode != 0 && in->store_mode != 1)
        return fail(IBLBM_EINVAL, "set_schedule: store_mode must be 0 or 1 (non-temporal)");
    if (in->bulk_delay < 0 || in->bulk_delay > 1000)
        return fail(IBLBM_EINVAL, "set_schedule: bulk_delay must be in [0, 1000] (10 ns ticks)");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    C->sched = *in;
    apply_row_layout(*C);
    clear_graphs(*C);
    return ensure_tables(*C);
}

int iblbm_autotune(void *p, int steps, double *best_ms, int *n_tried) {
    Ctx *C = as_ctx(p);
    if (!C || steps < 1) return fail(IBLBM_EINVAL, "autotune: bad argument");
    if (C->inst.empty()) return fail(IBLBM_EINVAL, "autotune: no instance");
    for (const HostInst &h : C->inst)
        if (h.d.pmode != 0)
            return fail(IBLBM_EINVAL, "autotune: switch the learned predictor off first");
    HIPCHK(hipSetDevice(C->device));
    int rc = ensure_tables(*C);
    if (rc) return rc;
    // candidates: the fused launch where every window qualifies, and the two-kernel launch
    // over workgroup sizes, rows in flight and grid (every row once, or 16 / 32 wavefronts
    // per CU walking the rows)
    std::vector<iblbm_schedule> cand;
    // over both row layouts (the flat one where every instance is at least 64 nodes wide)
    std::vector<int> layouts = {0};
    if (flat_eligible(*C)) layouts.push_back(1);
    // and both store modes where the populations outgrow the 256 MB last-level cache
    // (below that, the next step's loads hit in it after normal stores)
    size_t pop_bytes = 0;
    for (const HostInst &h : C->inst)
        pop_bytes += 2 * 9 * (size_t)h.d.qs * (h.d.prec == 1 ? sizeof(float) : sizeof(double));
    std::vector<int> stores = {0};
    if (pop_bytes > ((size_t)256 << 20)) stores.push_back(1);
    const bool f32_ctx = C->inst[0].d.prec == 1;  // rows 3 and 4: FP32 storage only
    for (int st : stores)
        for (int lay : layouts) {
            if (C->n_fast == C->n_ibm && st == 0)  // the fused step stores normally
                // -1: the workgroup row queue, whose rows = 3 is one row with the next one
                // copied into LDS ahead
                for (int order : {-1, 0, 1, 6, 8})
                    for (int r : {1, 2, 3, 4})
                        if ((r < 4 || f32_ctx) && (order >= 0 || r <= 3) &&
                            (r != 3 || order < 0))
                        for (int dl : {0, 100, 200})  // bulk start delay: 0, 1, 2 us
                            if (dl == 0 || C->n_fast > 0)
                                cand.push_back({IBLBM_MODE_FUSED, 256, 1, 0, r, order, lay, st,
                                                dl});
            // the two-kernel launch with the windows beside the bulk (two streams) or before
            // it (serial: no contention for the CUs, no overlap either)
            std::vector<int> modes2 = {IBLBM_MODE_SERIAL};
            if (C->n_ibm > 0) modes2.push_back(IBLBM_MODE_TWO_STREAMS);
            for (int mode2 : modes2)
                for (int bt : {256, 512, 1024})
                    for (int r : {1, 2, 3, 4})
                        if (r < 3 || f32_ctx)
                        for (int wv : {0, 16 * C->n_cu, 32 * C->n_cu})
                            cand.push_back({mode2, bt, r, wv, 2, 0, lay, st, 0});
        }
    // the trial steps start from a copy of every instance's source populations
    const int n = (int)C->inst.size();
    std::vector<void *> backup(n, nullptr);
    auto release = [&]() {
        for (void *b : backup)
            if (b) (void)hipFree(b);
    };
    auto fbytes = [&](const Inst &d) {
        return (d.prec == 1 ? sizeof(float) : sizeof(double)) * 9 * (size_t)d.qs;
    };
    for (int k = 0; k < n; ++k) {
        const Inst &d = C->inst[k].d;
        if (hipMalloc(&backup[k], fbytes(d)) != hipSuccess) {
            backup[k] = nullptr;
            release();
            return fail(IBLBM_ENOMEM, "autotune: no room for a copy of the populations");
        }
        HIPCHK(hipMemcpyAsync(backup[k], d.f[C->parity], fbytes(d), hipMemcpyDeviceToDevice,
                              C->stream));
    }
    const int64_t step0 = C->step;
    const int parity0 = C->parity;
    auto restore = [&]() -> int {
        for (int k = 0; k < n; ++k) {
            const Inst &d = C->inst[k].d;
            HIPCHK(hipMemcpyAsync(d.f[parity0], backup[k], fbytes(d), hipMemcpyDeviceToDevice,
                                  C->stream));
        }
        C->step = step0;
        C->parity = parity0;
        return 0;
    };
    const iblbm_schedule keep = C->sched;
    const bool conc = C->concurrent;
    C->concurrent = true;
    const int warm = 2, timed = steps + (steps & 1);  // even: the trial ends on the same parity
    // one trial: the candidate's time per step over `len` steps from the saved state
    auto trial = [&](const iblbm_schedule &sc, int len, double &ms_step) -> int {
        C->sched = sc;
        apply_row_layout(*C);
        int r = ensure_tables(*C);
        if (!r) r = restore();
        if (!r) r = launch_steps(*C, warm, 0, true);
        C->parity = parity0;
        if (!r) HIPCHK(hipEventRecord(C->ev0, C->stream));
        if (!r) r = launch_steps(*C, len, 0, true);
        if (!r) HIPCHK(hipEventRecord(C->ev1, C->stream));
        if (!r) HIPCHK(hipEventSynchronize(C->ev1));
        if (r) return r;
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, C->ev0, C->ev1));
        ms_step = (double)ms / len;
        return 0;
    };
    auto bail = [&](int r) {
        release();
        C->sched = keep;
        C->concurrent = conc;
        apply_row_layout(*C);
        (void)ensure_tables(*C);
        return r;
    };
    // every candidate once, then the four fastest again over four times the steps, twice in
    // turn, so that a few percent between the leaders is not decided by one short trial
    std::vector<std::pair<double, int>> first(cand.size());
    for (size_t k = 0; k < cand.size(); ++k) {
        double t = 0.0;
        rc = trial(cand[k], timed, t);
        if (rc) return bail(rc);
        first[k] = {t, (int)k};
    }
    std::sort(first.begin(), first.end());
    const int lead = (int)std::min<size_t>(4, first.size());
    std::vector<double> total(lead, 0.0);
    for (int pass = 0; pass < 2; ++pass)
        for (int k = 0; k < lead; ++k) {
            double t = 0.0;
            rc = trial(cand[first[k].second], 4 * timed, t);
            if (rc) return bail(rc);
            total[k] += t;
        }
    int win = 0;
    for (int k = 1; k < lead; ++k)
        if (total[k] < total[win]) win = k;
    const iblbm_schedule pick = cand[first[win].second];
    const double best = total[win] / 2;
    rc = restore();
    HIPCHK(hipStreamSynchronize(C->stream));
    release();
    C->concurrent = conc;
    C->sched = pick;
    apply_row_layout(*C);
    if (!rc) rc = ensure_tables(*C);
    clear_graphs(*C);
    if (rc) return rc;
    if (best_ms) *best_ms = best;
    if (n_tried) *n_tried = (int)cand.size();
    return 0;
}

int iblbm_set_boundary_velocity(void *p, int id, const double *U) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !U) return fail(IBLBM_EINVAL, "set_boundary_velocity: bad argument");
    Inst &d = C->inst[id].d;
    if (d.bc < 2)
        return fail(IBLBM_EINVAL, "set_boundary_velocity: the instance's closure reads no "
                                  "previous velocity");
    std::vector<double> v(2 * (size_t)d.nring);
    for (int s = 0; s < d.nring; ++s) {
        int i, j;
        ring_node(s, d.nx, d.ny, i, j);
        const int64_t n = (int64_t)j * d.nx + i;
        v[2 * s] = U[2 * n];
        v[2 * s + 1] = U[2 * n + 1];
    }
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipMemcpyAsync(d.side + 2 * (size_t)d.nring * C->parity, v.data(),
                          sizeof(double) * v.size(), hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return 0;
}

int iblbm_set_analytic_boundary(void *p, int id, const double *coef, int ncoef,
                                const double *decay, int nt) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !coef || !decay || nt < 1)
        return fail(IBLBM_EINVAL, "set_analytic_boundary: bad argument");
    Inst &d = C->inst[id].d;
    if (d.bc != IBLBM_BC_TG_ANALYTICAL)
        return fail(IBLBM_EINVAL, "set_analytic_boundary: the instance's closure is not analytic");
    const int want = 2 * d.ny + 2 * (d.nx - 2) + 4;
    if (ncoef != want)
        return fail(IBLBM_EINVAL, "set_analytic_boundary: ncoef must be 2 ny + 2 (nx - 2) + 4 = " +
                                      std::to_string(want));
    HIPCHK(hipSetDevice(C->device));
    void *a, *b;
    if (C->mem.alloc(&a, sizeof(double) * 2 * (size_t)ncoef) != hipSuccess ||
        C->mem.alloc(&b, sizeof(double) * (size_t)nt) != hipSuccess)
        return fail(IBLBM_ENOMEM, "set_analytic_boundary: tables do not fit");
    HIPCHK(hipMemcpyAsync(a, coef, sizeof(double) * 2 * (size_t)ncoef, hipMemcpyHostToDevice,
                          C->stream));
    HIPCHK(hipMemcpyAsync(b, decay, sizeof(double) * (size_t)nt, hipMemcpyHostToDevice,
                          C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    d.tgc = (const double *)a;
    d.tgd = (const double *)b;
    d.tg_nt = nt;
    C->tables_dirty = true;
    return ensure_tables(*C);
}

int iblbm_set_desired_velocity(void *p, int id, const double *ud) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !ud) return fail(IBLBM_EINVAL, "set_desired_velocity: bad argument");
    const Inst &d = C->inst[id].d;
    if (d.ibm == 0 || d.Lb < 1)
        return fail(IBLBM_EINVAL, "set_desired_velocity: the instance has no markers yet");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipMemcpyAsync((void *)d.ud, ud, sizeof(double) * 2 * (size_t)d.Lb,
                          hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return 0;
}

int iblbm_init_flow(void *p, int id, double rho, double ux, double uy, double amp,
                    uint64_t seed) {
    Ctx *C = as_ctx(p);
    if (valid_id(C, id) && C->inst[id].d.bc >= 2)
        return fail(IBLBM_EINVAL, "init_flow: this boundary closure needs the host's initial "
                                  "velocity (iblbm_set_boundary_velocity)");
    if (!valid_id(C, id)) return fail(IBLBM_EINVAL, "init_flow: bad instance id");
    HIPCHK(hipSetDevice(C->device));
    const Inst &d = C->inst[id].d;
    const int64_t N = (int64_t)d.nx * d.ny;
    const int blocks = (int)((N + 255) / 256);
    if (d.prec == 1)
        hipLaunchKernelGGL(k_init_flow<float>, dim3(blocks), dim3(256), 0, C->stream, d, rho, ux,
                           uy, amp, (unsigned long long)seed);
    else
        hipLaunchKernelGGL(k_init_flow<double>, dim3(blocks), dim3(256), 0, C->stream, d, rho, ux,
                           uy, amp, (unsigned long long)seed);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    return 0;
}

int iblbm_last_timing(void *p, double *ms, int *steps) {
    Ctx *C = as_ctx(p);
    if (!C || !ms || !steps) return fail(IBLBM_EINVAL, "last_timing: bad argument");
    HIPCHK(hipEventSynchronize(C->ev1));
    float t = 0.0f;
    HIPCHK(hipEventElapsedTime(&t, C->ev0, C->ev1));
    *ms = t;
    *steps = C->last_steps;
    return 0;
}

int64_t iblbm_step_counter(void *p) {
    Ctx *C = as_ctx(p);
    return C ? C->step : -1;
}

int iblbm_diag_fetch(void *p, int slot) {
    Ctx *C = as_ctx(p);
    if (!C || slot < 0 || slot > 1) return fail(IBLBM_EINVAL, "diag_fetch: bad argument");
    HIPCHK(hipSetDevice(C->device));
    int rc = ensure_tables(*C);
    if (rc) return rc;
    const size_t rows = (size_t)C->n_tiles * TILE_Y, n = C->inst.size();
    const size_t need = rows + 8 * n + (size_t)C->n_edge_blocks;
    if (need > C->h_diag_cap) {  // (re)size both slots once no copy into them is pending
        HIPCHK(hipStreamSynchronize(C->stream));
        for (int k = 0; k < 2; ++k) {
            if (C->h_diag[k]) HIPCHK(hipHostFree(C->h_diag[k]));
            C->h_diag[k] = nullptr;
            HIPCHK(hipHostMalloc((void **)&C->h_diag[k], sizeof(double) * need,
                                 hipHostMallocDefault));
        }
        C->h_diag_cap = need;
    }
    HIPCHK(hipEventSynchronize(C->ev_diag[slot]));  // the slot's previous copy has landed
    double *h = C->h_diag[slot];
    if (rows)
        HIPCHK(hipMemcpyAsync(h, C->d_rho_partial, sizeof(double) * rows,
                              hipMemcpyDeviceToHost, C->stream));
    for (size_t k = 0; k < n; ++k) {
        const Inst &d = C->inst[k].d;
        if (d.ibm != 0)
            HIPCHK(hipMemcpyAsync(h + rows + 8 * k, d.diag, sizeof(double) * 5,
                                  hipMemcpyDeviceToHost, C->stream));
        else
            for (int q = 0; q < 5; ++q) h[rows + 8 * k + q] = 0.0;
    }
    if (C->n_edge_blocks) {  // the boundary kernel's outer-ring partials
        HIPCHK(hipMemcpyAsync(h + rows + 8 * n, C->d_edge_rho, sizeof(double) * C->n_edge_blocks,
                              hipMemcpyDeviceToHost, C->stream));
    }
    HIPCHK(hipEventRecord(C->ev_diag[slot], C->stream));
    C->diag_step_slot[slot] = C->diag_step;
    return 0;
}

int iblbm_get_diag_slot(void *p, int id, int slot, iblbm_diag *out) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !out || slot < 0 || slot > 1 || !C->h_diag[slot])
        return fail(IBLBM_EINVAL, "get_diag_slot: bad argument or slot never fetched");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipEventSynchronize(C->ev_diag[slot]));
    const size_t rows = (size_t)C->n_tiles * TILE_Y;
    const double *h = C->h_diag[slot];
    const int b0 = TILE_Y * C->h_tile_prefix[id], b1 = TILE_Y * C->h_tile_prefix[id + 1];
    double rs = 0.0;
    for (int b = b0; b < b1; ++b) rs += h[b];  // the order of iblbm_get_diag
    const double *box = h + rows + 8 * (size_t)id;
    rs += box[0];
    if (!C->h_edge_prefix.empty())
        for (int e = C->h_edge_prefix[id]; e < C->h_edge_prefix[id + 1]; ++e)
            rs += h[rows + 8 * C->inst.size() + e];
    out->rho_sum = rs;
    out->force[0] = box[1];
    out->force[1] = box[2];
    out->torque_fib = box[3];
    out->torque_q = box[4];
    out->step = C->diag_step_slot[slot];
    return 0;
}

int iblbm_event_record(void *p, int k) {
    Ctx *C = as_ctx(p);
    if (!C || k < 0 || k > 3) return fail(IBLBM_EINVAL, "event_record: bad argument");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipEventRecord(C->ev_mark[k], C->stream));
    return 0;
}

int iblbm_event_elapsed(void *p, int a, int b, double *ms) {
    Ctx *C = as_ctx(p);
    if (!C || a < 0 || a > 3 || b < 0 || b > 3 || !ms)
        return fail(IBLBM_EINVAL, "event_elapsed: bad argument");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipEventSynchronize(C->ev_mark[b]));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, C->ev_mark[a], C->ev_mark[b]));
    *ms = t;
    return 0;
}

int iblbm_get_mdf_stats(void *p, int id, iblbm_mdf_stats *out) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !out) return fail(IBLBM_EINVAL, "mdf_stats: bad argument");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    double s[6];
    HIPCHK(hipMemcpy(s, C->inst[id].d.stats, sizeof s, hipMemcpyDeviceToHost));
    out->iter_count = (int32_t)s[0];
    out->last_residual = s[1];
    out->diverged = (int32_t)s[2];
    out->converged = (int32_t)s[3];
    const bool coupled = C->inst[id].d.np_part > 1;
    out->guard_trips = coupled ? (int32_t)s[4] : 0;
    out->omega_used = coupled ? s[5] : C->inst[id].d.omega;
    return 0;
}

int iblbm_get_diag(void *p, int id, iblbm_diag *out) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !out) return fail(IBLBM_EINVAL, "get_diag: bad argument");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    const Inst &d = C->inst[id].d;
    const int b0 = TILE_Y * C->h_tile_prefix[id], b1 = TILE_Y * C->h_tile_prefix[id + 1];  // rows
    std::vector<double> part((size_t)(b1 - b0));
    if (b1 > b0)
        HIPCHK(hipMemcpy(part.data(), C->d_rho_partial + b0, sizeof(double) * part.size(),
                         hipMemcpyDeviceToHost));
    double box[5] = {0, 0, 0, 0, 0};
    if (d.ibm != 0) HIPCHK(hipMemcpy(box, d.diag, sizeof box, hipMemcpyDeviceToHost));
    double rs = 0.0;
    for (double v : part) rs += v;
    rs += box[0];
    const int e0 = C->h_edge_prefix.empty() ? 0 : C->h_edge_prefix[id];
    const int e1 = C->h_edge_prefix.empty() ? 0 : C->h_edge_prefix[id + 1];
    if (e1 > e0) {  // the boundary kernel's outer-ring partials
        std::vector<double> ep((size_t)(e1 - e0));
        HIPCHK(hipMemcpy(ep.data(), C->d_edge_rho + e0, sizeof(double) * ep.size(),
                         hipMemcpyDeviceToHost));
        for (double v : ep) rs += v;
    }
    out->rho_sum = rs;
    out->force[0] = box[1];
    out->force[1] = box[2];
    out->torque_fib = box[3];
    out->torque_q = box[4];
    out->step = C->diag_step;
    return 0;
}

// ------------------------------------------------------------------ learned MDF predictor

int iblbm_predictor_mode(void *p, int id, int mode, int64_t capacity) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id)) return fail(IBLBM_EINVAL, "predictor_mode: bad instance id");
    if (mode < 0 || mode > 3) return fail(IBLBM_EKEY, "predictor_mode: unknown mode");
    HostInst &h = C->inst[id];
    Inst &d = h.d;
    if (mode != 0 && h.c.ibm != IBLBM_IBM_MDF)
        return fail(IBLBM_EINVAL, "predictor_mode: the predictor drives MDF instances only");
    // the fused window or the generic one (one particle, markers in the threads' registers)
    if (mode != 0 && (d.np_part > 1 || d.lb_cap > NTI))
        return fail(IBLBM_EINVAL, "predictor_mode: the predictor drives one-particle windows of "
                                  "at most " + std::to_string(NTI) + " markers");
    HIPCHK(hipSetDevice(C->device));
    IbmScratch &S = d.s;
    if (mode != 0 && !S.mlp_p) {  // first use: allocate and initialise
        if (capacity < 1) return fail(IBLBM_EINVAL, "predictor_mode: capacity must be >= 1");
        void *q;
        auto A = [&](size_t bytes) -> void * {
            return C->mem.alloc(&q, bytes) == hipSuccess ? q : nullptr;
        };
        S.mlp_p = (float *)A(sizeof(float) * 4 * MLP_NP);  // params, m, v, grad
        S.mlp_norm = (float *)A(sizeof(float) * (MLP_NORM + 2));  // 16-byte copies
        S.mlp_wb = (unsigned short *)A(sizeof(unsigned short) * MLP_NBLK * 256);
        S.data = (float *)A(sizeof(float) * MLP_ROW * (size_t)capacity);
        S.fprev = (double *)A(sizeof(double) * 4 * d.lb_cap);  // forces, then slips
        S.pstat = (double *)A(sizeof(double) * 4);
        S.pctr = (long long *)A(sizeof(long long) * 2);
        const size_t tiles = (size_t)(d.lb_cap + 15) / 16;  // the generic window's tiles
        S.mlpx = (unsigned short *)A(sizeof(unsigned short) * 256 * tiles);
        S.mlpy = (float *)A(sizeof(float) * 16 * MLP_OUT * tiles);
        if (!S.mlp_p || !S.mlp_norm || !S.mlp_wb || !S.data || !S.fprev || !S.pstat || !S.pctr ||
            !S.mlpx || !S.mlpy)
            return fail(IBLBM_ENOMEM, "predictor_mode: allocation failed");
        d.data_cap = capacity > INT32_MAX ? INT32_MAX : (int)capacity;
        // Glorot-uniform weights from a fixed seed, zero biases and Adam moments
        std::vector<float> hp(4 * MLP_NP, 0.0f);
        uint64_t z = 0x2545F4914F6CDD1Dull;
        auto uni = [&]() {
            z ^= z << 13;
            z ^= z >> 7;
            z ^= z << 17;
            return (double)(z >> 11) * (1.0 / 9007199254740992.0);
        };
        auto fill = [&](int off, int fin, int fout) {
            const double a = std::sqrt(6.0 / (fin + fout));
            for (int k = 0; k < fin * fout; ++k) hp[off + k] = (float)((2.0 * uni() - 1.0) * a);
        };
        fill(MLP_W1, MLP_IN, MLP_H);
        fill(MLP_W2, MLP_H, MLP_H);
        fill(MLP_W3, MLP_H, MLP_OUT);
        std::vector<float> hn(MLP_NORM, 1.0f);
        for (int k = 0; k < MLP_IN; ++k) hn[k] = 0.0f;
        HIPCHK(hipMemcpy(S.mlp_p, hp.data(), sizeof(float) * hp.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(S.mlp_norm, hn.data(), sizeof(float) * MLP_NORM, hipMemcpyHostToDevice));
        HIPCHK(hipMemset(S.pstat, 0, sizeof(double) * 4));
        HIPCHK(hipMemset(S.pctr, 0, sizeof(long long) * 2));
        hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, C->stream, S.mlp_p, S.mlp_wb);
        HIPCHK(hipGetLastError());
    }
    // The previous force is stale after the predictor was off, and a training sample needs a
    // previous force from plain MDF: its target is the change between two plain-MDF forces.
    if ((mode != 0 && d.pmode == 0) || (mode == 1 && d.pmode >= 2)) {
        HIPCHK(hipMemsetAsync(S.pctr + 1, 0, sizeof(long long), C->stream));
    }
    if ((mode != 0) != (d.pmode != 0)) {  // captured steps may hold the other step kernel
        HIPCHK(hipStreamSynchronize(C->stream));
        for (auto &g : C->graphs) (void)hipGraphExecDestroy(g.second);
        C->graphs.clear();
    }
    d.pmode = mode;
    int rc = C->tables_dirty ? ensure_tables(*C) : upload_inst(*C, id);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(C->stream));
    return 0;
}

int iblbm_predictor_samples_written(void *p, int id, int64_t *n) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !n) return fail(IBLBM_EINVAL, "predictor_samples_written: bad argument");
    const Inst &d = C->inst[id].d;
    *n = 0;
    if (!d.s.pctr) return 0;
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    long long c = 0;
    HIPCHK(hipMemcpy(&c, d.s.pctr, sizeof c, hipMemcpyDeviceToHost));
    *n = c;
    return 0;
}

int iblbm_predictor_samples(void *p, int id, int64_t *n) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !n) return fail(IBLBM_EINVAL, "predictor_samples: bad argument");
    const Inst &d = C->inst[id].d;
    *n = 0;
    if (!d.s.pctr) return 0;
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    long long c = 0;
    HIPCHK(hipMemcpy(&c, d.s.pctr, sizeof c, hipMemcpyDeviceToHost));
    *n = c < d.data_cap ? c : d.data_cap;
    return 0;
}

int iblbm_predictor_train(void *p, int id, int steps, double lr, uint32_t seed, float *losses) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || steps < 1 || !(lr > 0.0))
        return fail(IBLBM_EINVAL, "predictor_train: bad argument");
    Inst &d = C->inst[id].d;
    IbmScratch &S = d.s;
    if (!S.mlp_p) return fail(IBLBM_EINVAL, "predictor_train: predictor not enabled");
    int64_t n = 0;
    int rc = iblbm_predictor_samples(p, id, &n);
    if (rc) return rc;
    if (n < MLP_B) return fail(IBLBM_EINVAL, "predictor_train: fewer samples than one minibatch");
    HIPCHK(hipSetDevice(C->device));
    float *dl = nullptr;
    HIPCHK(hipMalloc(&dl, sizeof(float) * steps));
    hipLaunchKernelGGL(k_mlp_norm, dim3(1), dim3(MLP_TT), 0, C->stream, S.data, (long long)n,
                       S.mlp_norm);
    // fresh Adam moments, then chunks of at most 256 steps per launch
    hipError_t e = hipMemsetAsync(S.mlp_p + MLP_NP, 0, sizeof(float) * 2 * MLP_NP, C->stream);
    for (int s0 = 0; e == hipSuccess && s0 < steps; s0 += 256) {
        const int k = std::min(256, steps - s0);
        hipLaunchKernelGGL(k_mlp_train, dim3(1), dim3(MLP_TT), 0, C->stream, S.mlp_p,
                           S.mlp_p + MLP_NP, S.mlp_p + 2 * MLP_NP, S.mlp_p + 3 * MLP_NP,
                           (const float *)S.mlp_norm, (const float *)S.data, (long long)n, k, s0,
                           (float)lr, seed, dl + s0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, C->stream, S.mlp_p, S.mlp_wb);
        e = hipStreamSynchronize(C->stream);
    }
    if (e == hipSuccess && losses)
        e = hipMemcpy(losses, dl, sizeof(float) * steps, hipMemcpyDeviceToHost);
    (void)hipFree(dl);
    if (e != hipSuccess) return fail(IBLBM_EDEVICE, std::string("predictor_train: ") + hipGetErrorString(e));
    return 0;
}

int iblbm_predictor_params(void *p, int id, float *params, float *norm) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !params || !norm) return fail(IBLBM_EINVAL, "predictor_params: bad argument");
    const IbmScratch &S = C->inst[id].d.s;
    if (!S.mlp_p) return fail(IBLBM_EINVAL, "predictor_params: predictor not enabled");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipMemcpy(params, S.mlp_p, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(norm, S.mlp_norm, sizeof(float) * MLP_NORM, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_predictor_set_params(void *p, int id, const float *params, const float *norm) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !params || !norm)
        return fail(IBLBM_EINVAL, "predictor_set_params: bad argument");
    IbmScratch &S = C->inst[id].d.s;
    if (!S.mlp_p) return fail(IBLBM_EINVAL, "predictor_set_params: predictor not enabled");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipMemcpy(S.mlp_p, params, sizeof(float) * MLP_NP, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(S.mlp_norm, norm, sizeof(float) * MLP_NORM, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, C->stream, S.mlp_p, S.mlp_wb);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    return 0;
}

int iblbm_predictor_data(void *p, int id, int64_t first, int64_t n, float *rows) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !rows || first < 0 || n < 0)
        return fail(IBLBM_EINVAL, "predictor_data: bad argument");
    const Inst &d = C->inst[id].d;
    if (!d.s.data) return fail(IBLBM_EINVAL, "predictor_data: predictor not enabled");
    if (n > d.data_cap) return fail(IBLBM_EINVAL, "predictor_data: more rows than the capacity");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    for (int64_t r = 0; r < n;) {  // the samples form a ring of data_cap rows
        const int64_t at = (first + r) % d.data_cap, k = std::min<int64_t>(n - r, d.data_cap - at);
        HIPCHK(hipMemcpy(rows + r * MLP_ROW, d.s.data + at * MLP_ROW, sizeof(float) * MLP_ROW * k,
                         hipMemcpyDeviceToHost));
        r += k;
    }
    return 0;
}

int iblbm_predictor_previous_force(void *p, int id, double *F, int set) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !F) return fail(IBLBM_EINVAL, "predictor_previous_force: bad argument");
    const HostInst &h = C->inst[id];
    const IbmScratch &S = h.d.s;
    if (!S.fprev) return fail(IBLBM_EINVAL, "predictor_previous_force: predictor not enabled");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    const size_t bytes = sizeof(double) * 2 * h.Lb;
    double *slip = S.fprev + 2 * h.d.lb_cap;
    if (set) {
        HIPCHK(hipMemcpy(S.fprev, F, bytes, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(slip, F + 2 * h.Lb, bytes, hipMemcpyHostToDevice));
        const long long one = 1;
        HIPCHK(hipMemcpy(S.pctr + 1, &one, sizeof one, hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMemcpy(F, S.fprev, bytes, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(F + 2 * h.Lb, slip, bytes, hipMemcpyDeviceToHost));
    }
    return 0;
}

int iblbm_predictor_stats(void *p, int id, double out[4], int reset) {
    Ctx *C = as_ctx(p);
    if (!valid_id(C, id) || !out) return fail(IBLBM_EINVAL, "predictor_stats: bad argument");
    const IbmScratch &S = C->inst[id].d.s;
    for (int k = 0; k < 4; ++k) out[k] = 0.0;
    if (!S.pstat) return 0;
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipMemcpy(out, S.pstat, sizeof(double) * 4, hipMemcpyDeviceToHost));
    if (reset) {
        HIPCHK(hipMemset(S.pstat, 0, sizeof(double) * 4));
        HIPCHK(hipDeviceSynchronize());
    }
    return 0;
}

}  // extern "C"
