// ibm_block.h -- the immersed-boundary workgroup of the fused step (one per instance).
//
// It owns the nodes some clamped marker stencil lands on ("touched" nodes; the bulk tiles
// skip exactly those): it streams and closes them, computes rho and U, runs DF
// (iblbm/ibm/df.py:94-154), MDF (iblbm/ibm/mdf.py:175-229) or DFC (iblbm/ibm/dfc.py:256-342)
// on them, applies the velocity correction (iblbm/ibm/dispatch.py:18-33) and finishes their
// collision.  Every other node of the window box carries fib = 0 exactly and is collided by
// the bulk tiles with the same bits.
//
// The tables (contributions, touched compaction, gather lists, DFC lambda) are built once per
// marker set by k_tables.  During a step the touched-node working arrays (MDF velocity,
// dt/(2 rho), accumulated fib), the marker forces and the gather lists live in LDS, and each
// marker's stencil (touched index, weight) lives in registers, so an MDF sub-step is two
// short latency chains (marker interpolation, node gather) and one wavefront-shuffle reduction.
//
// Spreading is a gather: every touched node owns the list of (stencil offset s, marker k)
// contributions that land on it, sorted by c = s*Lb + k, which is exactly the order in which
// np.add.at applies them (offset loop outer, markers inner), so the FP64 sums carry the
// reference's bits and do not depend on thread timing (no floating-point atomics).
#pragma once
#include <type_traits>

#include "mlp.h"
#include "step_kernel.h"

namespace iblbm {

constexpr int MAX_T = 1536;   // touched nodes held in LDS (beyond: global scratch)
constexpr int MAX_LB = 512;   // markers held in LDS (beyond: global scratch)
constexpr int MAX_G = 4096;   // gather-list entries held in LDS (beyond: global scratch)

struct IbmShared {
    double uwx[MAX_T], uwy[MAX_T], vs[MAX_T], fibx[MAX_T], fiby[MAX_T], rot[MAX_T];
    double Lfx[MAX_LB], Lfy[MAX_LB];
    alignas(16) double gw[MAX_G];
    alignas(16) int gk[MAX_G];
    alignas(16) int toff[MAX_T + 4];  // 16-byte copy granules
};

__device__ inline int64_t floor_idx(double v) { return (int64_t)floor(v); }

// ------------------------------------------------------------------ block-wide primitives

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = nanmax2(v, __shfl_xor(v, o, 64));
    return v;
}
// Unsigned 64-bit max over the wavefront by DPP row shifts and row broadcasts (a VALU
// sequence, no LDS round trips); the full max ends in lane 63.  Zero is the identity.
__device__ inline unsigned long long dpp_max_u64_lane63(unsigned long long v) {
    auto step = [](unsigned long long x, auto ctrl, auto rmask) -> unsigned long long {
        constexpr int C = decltype(ctrl)::value, RM = decltype(rmask)::value;
        const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)x, C, RM, 0xf, true);
        const unsigned hi =
            (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), C, RM, 0xf, true);
        const unsigned long long y = ((unsigned long long)hi << 32) | lo;
        return y > x ? y : x;
    };
    using std::integral_constant;
    v = step(v, integral_constant<int, 0x111>(), integral_constant<int, 0xf>());  // row_shr:1
    v = step(v, integral_constant<int, 0x112>(), integral_constant<int, 0xf>());  // row_shr:2
    v = step(v, integral_constant<int, 0x114>(), integral_constant<int, 0xf>());  // row_shr:4
    v = step(v, integral_constant<int, 0x118>(), integral_constant<int, 0xf>());  // row_shr:8
    v = step(v, integral_constant<int, 0x142>(), integral_constant<int, 0xa>());  // row_bcast:15
    v = step(v, integral_constant<int, 0x143>(), integral_constant<int, 0xc>());  // row_bcast:31
    return v;
}

// Largest value over the wavefront (every lane gets it).
__device__ inline float wave_maxf(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// Deterministic block sum/max (fixed shuffle tree per wavefront, then over wavefronts).
__device__ inline double block_reduce(double v, double *red, bool is_max) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    v = is_max ? wave_max(v) : wave_sum(v);
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (wv == 0) {
        double x = lane < NTI / 64 ? red[lane] : (is_max ? -INFINITY : 0.0);
        x = is_max ? wave_max(x) : wave_sum(x);
        if (lane == 0) red[NTI / 64] = x;
    }
    __syncthreads();
    const double r = red[NTI / 64];
    __syncthreads();
    return r;
}
// Exclusive scan of in[0..n) into out[0..n], out[n] = total (in and out may alias).
__device__ inline void block_scan_int(const int *in, int *out, int n, int *redi) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int chunk = (n + NTI - 1) / NTI;
    const int b = min(n, t * chunk), e = min(n, b + chunk);
    int s = 0;
    for (int k = b; k < e; ++k) s += in[k];
    int incl = s;  // inclusive scan of s across the wavefront
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) redi[wv] = incl;
    __syncthreads();
    if (wv == 0) {
        const int x = lane < NTI / 64 ? redi[lane] : 0;
        int xi = x;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(xi, o, 64);
            if (lane >= o) xi += y;
        }
        if (lane < NTI / 64) redi[NTI / 64 + lane] = xi - x;  // exclusive wave offsets
        if (lane == NTI / 64 - 1) redi[2 * (NTI / 64)] = xi;
    }
    __syncthreads();
    int acc = redi[NTI / 64 + wv] + incl - s;
    const int total = redi[2 * (NTI / 64)];
    __syncthreads();
    for (int k = b; k < e; ++k) {
        const int v = in[k];
        out[k] = acc;
        acc += v;
    }
    if (t == 0) out[n] = total;
    __threadfence_block();
    __syncthreads();
}

// ------------------------------------------------------------------ window tables

// Contribution tables (marker-major c_m = k*K + s), touched-node compaction and per-node
// gather lists in np.add.at order.
// Particle of marker k in a window of several particles (markers in particle order).
__device__ inline int particle_of(const Inst &I, int k) {
    int p = 0;
#pragma unroll
    for (int q = 1; q < IBM_MAXP; ++q) p += (q < I.np_part && k >= I.poff[q]) ? 1 : 0;
    return p;
}

template <int A>
__device__ void ibm_build_tables(Inst &I, int *redi) {
    constexpr int side = 2 * A + 1, K = side * side;
    const int t = threadIdx.x, Lb = I.Lb, nc = K * Lb, nbox = I.bw * I.bh;
    IbmScratch &S = I.s;
    for (int k = t; k < nbox; k += NTI) {
        S.cnt[k] = 0;
        S.cur[k] = 0;
    }
    __syncthreads();
    for (int cm = t; cm < nc; cm += NTI) {
        const int k = cm / K, s = cm - k * K;
        const int di = s / side - A, dj = s % side - A;
        const double Lx = I.Lx[k], Ly = I.Ly[k];
        const int64_t ix0 = floor_idx(Lx / I.dx + 0.5), iy0 = floor_idx(Ly / I.dy + 0.5);
        const int64_t ei = ix0 + di, ej = iy0 + dj;
        const double wx = delta_phi(I.delta, (Lx - (double)ei * I.dx) / I.dx);
        const double wy = delta_phi(I.delta, (Ly - (double)ej * I.dy) / I.dy);
        double w = wx * wy;
        const int64_t eic = ei < 0 ? 0 : (ei > I.nx - 1 ? I.nx - 1 : ei);
        const int64_t ejc = ej < 0 ? 0 : (ej > I.ny - 1 ? I.ny - 1 : ej);
        int slot;
        // The host sizes the box to hold every clamped stencil node.  A node outside it
        // would be a host bug: it is routed to slot 0 with a NaN weight so the step output
        // turns NaN (loudly) instead of touching memory outside the scratch arrays.
        if (eic < I.bx0 || eic >= I.bx0 + I.bw || ejc < I.by0 || ejc >= I.by0 + I.bh) {
            slot = 0;
            w = __builtin_nan("");
        } else {
            slot = (int)((ejc - I.by0) * I.bw + (eic - I.bx0));
        }
        S.cw[cm] = w;
        S.cslot[cm] = slot;
        atomicAdd(&S.cnt[slot], 1);
    }
    __threadfence_block();
    __syncthreads();
    // touched flags -> compact index (tindex has nbox + 1 entries)
    for (int k = t; k < nbox; k += NTI) S.tindex[k] = S.cnt[k] > 0 ? 1 : 0;
    __threadfence_block();
    __syncthreads();
    block_scan_int(S.tindex, S.tindex, nbox, redi);
    const int nT = S.tindex[nbox];
    for (int k = t; k < nbox; k += NTI)
        if (S.cnt[k] > 0) S.tslot[S.tindex[k]] = k;
    __threadfence_block();
    __syncthreads();
    for (int k = t; k < nbox; k += NTI)
        if (S.cnt[k] == 0) S.tindex[k] = -1;
    if (t == 0) I.nT = nT;
    // gather offsets over touched nodes (toff has nT + 1 entries)
    for (int ti = t; ti < nT; ti += NTI) S.toff[ti] = S.cnt[S.tslot[ti]];
    __threadfence_block();
    __syncthreads();
    block_scan_int(S.toff, S.toff, nT, redi);
    // np.add.at order key s * Lb + k; a DF / MDF window of several particles spreads each
    // particle on its own (multi_particle.py:139-160, :400-421), so its key leads with the
    // particle: (p * K + s) * Lb + k orders by particle, then as one particle's np.add.at
    const bool by_part = (I.ibm == 1 || I.ibm == 2) && I.np_part > 1;
    for (int cm = t; cm < nc; cm += NTI) {
        const int sl = S.cslot[cm];
        const int ti = S.tindex[sl];
        S.cti[cm] = ti;
        const int k = cm / K, s = cm - k * K;
        const int p = by_part ? particle_of(I, k) : 0;
        const int pos = atomicAdd(&S.cur[sl], 1);
        S.gc[S.toff[ti] + pos] = (p * K + s) * Lb + k;
    }
    __threadfence_block();
    __syncthreads();
    for (int ti = t; ti < nT; ti += NTI) {  // sort each list by the np order key
        int *L = S.gc + S.toff[ti];
        const int n = S.toff[ti + 1] - S.toff[ti];
        for (int x = 1; x < n; ++x) {
            const int v = L[x];
            int y = x - 1;
            while (y >= 0 && L[y] > v) {
                L[y + 1] = L[y];
                --y;
            }
            L[y + 1] = v;
        }
        for (int x = 0; x < n; ++x) {  // (marker, w*Larea) per entry
            const int c = L[x], s = (c / Lb) % K, k = c % Lb;
            S.gk[S.toff[ti] + x] = k;
            S.gw[S.toff[ti] + x] = S.cw[k * K + s] * (by_part ? I.plarea[particle_of(I, k)] : I.Larea);
        }
    }
    __threadfence_block();
    __syncthreads();
}

// DFC lambda (iblbm/ibm/dfc.py:157-185), rho_f = 1: W_total per touched node in np order,
// then W_sum per marker in stencil order.
template <int A>
__device__ void ibm_dfc_lambda(Inst &I, int nT, int Lb) {
    constexpr int K = (2 * A + 1) * (2 * A + 1);
    const int t = threadIdx.x;
    IbmScratch &S = I.s;
    for (int ti = t; ti < nT; ti += NTI) {
        double w = 0.0;
        for (int x = S.toff[ti], e = S.toff[ti + 1]; x < e; ++x) {
            const int k = S.gk[x];
            const int s = S.gc[x] / Lb;
            w += S.cw[k * K + s];
        }
        S.wt[ti] = w;
    }
    __threadfence_block();
    __syncthreads();
    for (int k = t; k < Lb; k += NTI) {
        double ws = 0.0;
#pragma unroll
        for (int s = 0; s < K; ++s) ws += S.wt[S.cti[k * K + s]] * S.cw[k * K + s];
        S.lam[k] = 1.0 / (2.0 * 1.0 * I.Larea * ws);
    }
    __threadfence_block();
    __syncthreads();
}

// ------------------------------------------------------------------ the IBM core

// The MDF loop decisions (mdf.py:175-229), made identically by every wavefront of the window
// from the block-wide residual.  In the warm-start modes the loop needs only one true
// correction.  It restarts plain MDF from the uncorrected velocity once when the predicted
// start is no better than no forcing at all (the first residual after the prediction is not
// below plain MDF's first residual, or is not finite: a wrong or corrupted network), or when
// the divergence guard trips.
struct MdfCtl {
    bool mdf, warm;
    int n_iter, min_iter, min_plain;
    double tol, prev = INFINITY, last = INFINITY;
    double plain0 = INFINITY;  // warm modes: plain MDF's first residual (slip before forcing)
    int iter_count = 0, diverged = 0, converged = 0, restarted = 0, applied = 0;
    // after barrier A: 0 apply this sub-step, 1 stop (divergence guard), 2 restart plain MDF
    __device__ int at_a(double res, int it) {
        if (!mdf) return 0;
        last = res;
        if (warm && !restarted && it == 0 && !(res < plain0)) {  // NaN compares false
            restarted = 1;
            prev = INFINITY;
            min_iter = min_plain;
            return 2;
        }
        if (it > 0 && res > prev * 1.5) {  // divergence guard, mdf.py:192-195
            if (warm && !restarted) {
                restarted = 1;
                prev = INFINITY;
                min_iter = min_plain;
                return 2;
            }
            diverged = 1;
            iter_count = it + 1;
            return 1;
        }
        return 0;
    }
    // after barrier B of an applied sub-step: true when the loop has converged
    __device__ bool at_b(double res, int it) {
        ++applied;
        if (!mdf) return true;
        prev = res;
        iter_count = it + 1;
        if (it + 1 >= min_iter && res < tol) {
            converged = 1;
            return true;
        }
        return false;
    }
};


// Touched-node working arrays (MDF velocity, dt/(2 rho), fib, rho), marker forces and gather
// lists.  LDS or global scratch is a compile-time choice, so that LDS accesses compile to
// ds_* instructions rather than flat ones.
struct IbmPtrs {
    double *uwx, *uwy, *vs, *fbx, *fby, *rot, *Lfx, *Lfy, *gw;
    int *gk, *toff;
};
template <bool LDS>
__device__ inline IbmPtrs ibm_ptrs(Inst &I, IbmShared &sh) {
    IbmScratch &S = I.s;
    if constexpr (LDS)
        return {sh.uwx, sh.uwy, sh.vs, sh.fibx, sh.fiby, sh.rot, sh.Lfx, sh.Lfy, sh.gw, sh.gk,
                sh.toff};
    else
        return {S.uwx, S.uwy, S.vs, S.tfibx, S.tfiby, S.rot, S.Lfx, S.Lfy, S.gw, S.gk, S.toff};
}
__device__ inline bool ibm_fits_lds(const Inst &I, int nT) {
    return nT <= MAX_T && I.Lb <= MAX_LB && I.s.toff[nT] <= MAX_G;
}

// Copies the gather lists of the window (gw, gk, toff; the global copies are padded to 16
// bytes) into LDS in 16-byte granules, several granules in flight per thread.
__device__ inline void ibm_stage_lists(const Inst &I, int nT, IbmShared &sh) {
    const IbmScratch &S = I.s;
    const int ng = S.toff[nT];
    const int ugw = (ng + 1) / 2, ugk = (ng + 3) / 4, uto = (nT + 4) / 4;
    const int U = ugw + ugk + uto;
#pragma unroll 8
    for (int u = threadIdx.x; u < U; u += NTI) {
        const uint4 *src;
        uint4 *dst;
        if (u < ugw) {
            src = reinterpret_cast<const uint4 *>(S.gw) + u;
            dst = reinterpret_cast<uint4 *>(sh.gw) + u;
        } else if (u < ugw + ugk) {
            src = reinterpret_cast<const uint4 *>(S.gk) + (u - ugw);
            dst = reinterpret_cast<uint4 *>(sh.gk) + (u - ugw);
        } else {
            src = reinterpret_cast<const uint4 *>(S.toff) + (u - ugw - ugk);
            dst = reinterpret_cast<uint4 *>(sh.toff) + (u - ugw - ugk);
        }
        *dst = *src;
    }
}

// One marker's stencil (touched index, weight) and desired velocity, held in registers by the
// thread that owns the marker (Lb <= NTI).
template <int A>
struct MarkerRegs {
    static constexpr int K = (2 * A + 1) * (2 * A + 1);
    int ti[K];
    double w[K];
    double udx, udy;
};
template <int A>
__device__ inline void load_marker_regs(const Inst &I, MarkerRegs<A> &M) {
    constexpr int K = MarkerRegs<A>::K;
    const int t = threadIdx.x;
    if (I.Lb <= NTI && t < I.Lb) {
#pragma unroll
        for (int s = 0; s < K; ++s) {
            M.ti[s] = I.s.cti[t * K + s];
            M.w[s] = I.s.cw[t * K + s];
        }
        M.udx = I.ud[2 * t];
        M.udy = I.ud[2 * t + 1];
    }
}

// Spread of one touched node as a gather in np.add.at order: sum_x Lf[k_x] * w_x.
__device__ inline void gather2(const IbmPtrs &P, int ti, double &ex, double &ey) {
    int x = P.toff[ti];
    const int e = P.toff[ti + 1];
    double sx = 0.0, sy = 0.0;
    for (; x + 4 <= e; x += 4) {
        const int k0 = P.gk[x], k1 = P.gk[x + 1], k2 = P.gk[x + 2], k3 = P.gk[x + 3];
        const double w0 = P.gw[x], w1 = P.gw[x + 1], w2 = P.gw[x + 2], w3 = P.gw[x + 3];
        const double a0 = P.Lfx[k0], a1 = P.Lfx[k1], a2 = P.Lfx[k2], a3 = P.Lfx[k3];
        const double b0 = P.Lfy[k0], b1 = P.Lfy[k1], b2 = P.Lfy[k2], b3 = P.Lfy[k3];
        sx += a0 * w0;
        sy += b0 * w0;
        sx += a1 * w1;
        sy += b1 * w1;
        sx += a2 * w2;
        sy += b2 * w2;
        sx += a3 * w3;
        sy += b3 * w3;
    }
    for (; x < e; ++x) {
        const int k = P.gk[x];
        const double w = P.gw[x];
        sx += P.Lfx[k] * w;
        sy += P.Lfy[k] * w;
    }
    ex = sx;
    ey = sy;
}

// Touched-node inputs of the core from box-slot arrays S.ro, S.ux0, S.uy0 (operator path).
__device__ inline void ibm_touched_init(const Inst &I, int nT, const IbmPtrs &P) {
    const IbmScratch &S = I.s;
    for (int ti = threadIdx.x; ti < nT; ti += NTI) {
        const int sl = S.tslot[ti];
        const double r = S.ro[sl];
        P.uwx[ti] = S.ux0[sl];
        P.uwy[ti] = S.uy0[sl];
        P.rot[ti] = r;
        P.vs[ti] = I.cp.dt / (2.0 * r);  // mdf.py:117
    }
    __threadfence_block();
    __syncthreads();
}

// DF / MDF / DFC on the window.  Inputs: P.uwx, P.uwy, P.rot, P.vs of the touched nodes
// (DF/MDF), or S.fs of the touched slots (DFC).  DF/MDF leave fib of the touched nodes in
// P.fbx/P.fby; DFC corrects S.fs and recomputes S.ro, S.ux0, S.uy0 of the touched slots and
// leaves the marker forces in S.q.
// DF / MDF of a window holding several particles (iblbm/ibm/multi_particle.py).
//  DF (multi_particle_df, :347-421): each particle's closure from the same velocity; the
//    window fib is the particles' spreads added in particle order; the marker forces of the
//    pass are kept (the host takes each particle's force and torque from them,
//    forcing.py:263-275).
//  MDF (multi_particle_mdf_coupled, :62-316): the particles share the working velocity;
//    omega starts at the host's min(0.8 omega_est, 0.8); an iteration whose residual exceeds
//    1.2 x the previous one is dropped and halves omega (floor omega / 8, reaching it ends the
//    loop as diverged); each particle's accumulated field is kept for its force and torque.
template <int A>
__device__ void ibm_core_multi(Inst &I, int nT, const IbmPtrs &P, double *red,
                               const MarkerRegs<A> &M) {
    constexpr int K = (2 * A + 1) * (2 * A + 1);
    const int t = threadIdx.x, Lb = I.Lb, bc = I.box_cap;
    IbmScratch &S = I.s;
    const bool inc = I.inc != 0;
    const double dt = I.cp.dt;
    const bool mdf = I.ibm == 2;
    const int npart = I.np_part;
    for (int ti = t; ti < nT; ti += NTI) {
        P.fbx[ti] = 0.0;
        P.fby[ti] = 0.0;
        if (mdf) {
            const int sl = S.tslot[ti];
            for (int p = 0; p < npart; ++p) {
                S.pfb[(2 * p) * bc + sl] = 0.0;
                S.pfb[(2 * p + 1) * bc + sl] = 0.0;
            }
        }
    }
    const bool regs = Lb <= NTI;
    double R = 0.0;
    if (regs && t < Lb) {
#pragma unroll
        for (int s = 0; s < K; ++s) R += P.rot[M.ti[s]] * M.w[s];
    }
    if (!regs) {
        for (int k = t; k < Lb; k += NTI) {
            double r = 0.0;
#pragma unroll
            for (int s = 0; s < K; ++s) r += P.rot[S.cti[k * K + s]] * S.cw[k * K + s];
            S.R[k] = r;
        }
    }
    __threadfence_block();
    __syncthreads();
    const int n_iter = mdf ? I.n_iter : 1;
    const int min_iter = max(1, min(I.min_iter, n_iter));
    const double tol = I.tol;
    double omega = I.omega;
    const double omega_min = omega * 0.125;
    double prev = INFINITY, last = INFINITY;
    int iter_count = 0, diverged = 0, converged = 0, trips = 0;
    for (int it = 0; it < n_iter; ++it) {
        double rmax = -INFINITY;
        for (int k = t; k < Lb; k += NTI) {  // interpolation + force (df.py:239-263)
            double lx = 0.0, ly = 0.0, Rk, udx, udy;
            if (regs) {
#pragma unroll
                for (int s = 0; s < K; ++s) {
                    lx += P.uwx[M.ti[s]] * M.w[s];
                    ly += P.uwy[M.ti[s]] * M.w[s];
                }
                Rk = R;
                udx = M.udx;
                udy = M.udy;
            } else {
#pragma unroll
                for (int s = 0; s < K; ++s) {
                    const int ti = S.cti[k * K + s];
                    const double w = S.cw[k * K + s];
                    lx += P.uwx[ti] * w;
                    ly += P.uwy[ti] * w;
                }
                Rk = S.R[k];
                udx = I.ud[2 * k];
                udy = I.ud[2 * k + 1];
            }
            if (inc) {
                P.Lfx[k] = 2.0 * (udx - lx) / dt;
                P.Lfy[k] = 2.0 * (udy - ly) / dt;
            } else {
                P.Lfx[k] = 2.0 * Rk * (udx - lx) / dt;
                P.Lfy[k] = 2.0 * Rk * (udy - ly) / dt;
            }
            if (!mdf) {
                S.mf[2 * k] = P.Lfx[k];
                S.mf[2 * k + 1] = P.Lfy[k];
            }
            const double vx = udx - lx, vy = udy - ly;
            rmax = nanmax2(sqrt(vx * vx + vy * vy), rmax);
        }
        __threadfence_block();
        const double res = block_reduce(rmax, red, true);  // its barrier publishes Lf
        if (mdf && it > 0 && res > prev * 1.2) {  // multi_particle.py:152-162
            ++trips;
            omega = fmax(omega * 0.5, omega_min);
            if (omega <= omega_min) {
                diverged = 1;
                break;
            }
            prev = res;
            continue;
        }
        // spread: each particle's contributions summed on their own, particles added in order
        for (int ti = t; ti < nT; ti += NTI) {
            const int sl = S.tslot[ti];
            double tx = 0.0, ty = 0.0;
            int x = P.toff[ti];
            const int end = P.toff[ti + 1];
            while (x < end) {
                const int p = particle_of(I, P.gk[x]);
                double ex = 0.0, ey = 0.0;
                for (; x < end && particle_of(I, P.gk[x]) == p; ++x) {
                    const int k = P.gk[x];
                    const double w = P.gw[x];
                    ex += P.Lfx[k] * w;
                    ey += P.Lfy[k] * w;
                }
                if (mdf) {
                    const double fx = ex * omega, fy = ey * omega;
                    tx = tx + fx;
                    ty = ty + fy;
                    S.pfb[(2 * p) * bc + sl] = S.pfb[(2 * p) * bc + sl] + fx;
                    S.pfb[(2 * p + 1) * bc + sl] = S.pfb[(2 * p + 1) * bc + sl] + fy;
                } else {
                    tx = tx + ex;
                    ty = ty + ey;
                }
            }
            if (mdf) {
                P.fbx[ti] = P.fbx[ti] + tx;
                P.fby[ti] = P.fby[ti] + ty;
                if (inc) {
                    P.uwx[ti] = P.uwx[ti] + 0.5 * tx * dt;
                    P.uwy[ti] = P.uwy[ti] + 0.5 * ty * dt;
                } else {
                    const double vs = P.vs[ti];
                    P.uwx[ti] = P.uwx[ti] + tx * vs;
                    P.uwy[ti] = P.uwy[ti] + ty * vs;
                }
            } else {
                P.fbx[ti] = tx;
                P.fby[ti] = ty;
            }
        }
        __threadfence_block();
        __syncthreads();
        if (mdf) {
            prev = res;
            ++iter_count;
            last = res;
            if (it + 1 >= min_iter && res < tol) {
                converged = 1;
                break;
            }
        }
    }
    if (mdf && t == 0) {
        I.stats[0] = iter_count;
        I.stats[1] = last;
        I.stats[2] = diverged;
        I.stats[3] = converged;
        I.stats[4] = trips;
        I.stats[5] = omega;
    }
    __threadfence_block();
    __syncthreads();
}

template <int A>
__device__ void ibm_core(Inst &I, int nT, const IbmPtrs &P, double *red, bool store_lu,
                         const MarkerRegs<A> &M) {
    constexpr int K = (2 * A + 1) * (2 * A + 1);
    const int t = threadIdx.x, Lb = I.Lb;
    IbmScratch &S = I.s;
    const bool inc = I.inc != 0;
    const double dt = I.cp.dt;
    if (I.df_multi || (I.ibm == 2 && I.np_part > 1)) {
        ibm_core_multi<A>(I, nT, P, red, M);
    } else if (I.ibm == 1 || I.ibm == 2) {
        for (int ti = t; ti < nT; ti += NTI) {
            P.fbx[ti] = 0.0;
            P.fby[ti] = 0.0;
        }
        const bool regs = Lb <= NTI;
        double R = 0.0;
        if (regs && t < Lb) {
#pragma unroll
            for (int s = 0; s < K; ++s) R += P.rot[M.ti[s]] * M.w[s];  // rho at the marker
        }
        if (!regs) {
            for (int k = t; k < Lb; k += NTI) {
                double r = 0.0;
#pragma unroll
                for (int s = 0; s < K; ++s) r += P.rot[S.cti[k * K + s]] * S.cw[k * K + s];
                S.R[k] = r;
            }
        }
        __threadfence_block();
        __syncthreads();
        const bool mdf = I.ibm == 2;
        const bool finc = mdf ? false : inc;  // MDF sub-steps are compressible (mdf.py:177-181)
        const int n_iter = mdf ? I.n_iter : 1;
        const int min_iter = max(1, min(I.min_iter, n_iter));
        const double omega = I.omega, tol = I.tol;
        double prev = INFINITY, last = INFINITY;
        int iter_count = 0, diverged = 0, converged = 0;
        for (int it = 0; it < n_iter; ++it) {
            // interpolation + force per marker (df.py:239-263)
            double rmax = -INFINITY;
            for (int k = t; k < Lb; k += NTI) {
                double lx = 0.0, ly = 0.0, Rk, udx, udy;
                if (regs) {
#pragma unroll
                    for (int s = 0; s < K; ++s) {
                        lx += P.uwx[M.ti[s]] * M.w[s];
                        ly += P.uwy[M.ti[s]] * M.w[s];
                    }
                    Rk = R;
                    udx = M.udx;
                    udy = M.udy;
                } else {
#pragma unroll
                    for (int s = 0; s < K; ++s) {
                        const int ti = S.cti[k * K + s];
                        const double w = S.cw[k * K + s];
                        lx += P.uwx[ti] * w;
                        ly += P.uwy[ti] * w;
                    }
                    Rk = S.R[k];
                    udx = I.ud[2 * k];
                    udy = I.ud[2 * k + 1];
                }
                if (finc) {
                    P.Lfx[k] = 2.0 * (udx - lx) / dt;
                    P.Lfy[k] = 2.0 * (udy - ly) / dt;
                } else {
                    P.Lfx[k] = 2.0 * Rk * (udx - lx) / dt;
                    P.Lfy[k] = 2.0 * Rk * (udy - ly) / dt;
                }
                if (store_lu) {
                    S.Lux[k] = lx;
                    S.Luy[k] = ly;
                }
                const double vx = udx - lx, vy = udy - ly;
                rmax = nanmax2(sqrt(vx * vx + vy * vy), rmax);
            }
            __threadfence_block();
            const double res = block_reduce(rmax, red, true);  // its barrier publishes Lf
            if (mdf) {
                last = res;
                if (it > 0 && res > prev * 1.5) {  // divergence guard, mdf.py:192-195
                    diverged = 1;
                    iter_count = it + 1;
                    break;
                }
            }
            // spread as a gather in np.add.at order (df.py:269-283)
            for (int ti = t; ti < nT; ti += NTI) {
                double ex, ey;
                gather2(P, ti, ex, ey);
                if (mdf) {  // mdf.py:197-200
                    const double fx = ex * omega, fy = ey * omega;
                    const double vs = P.vs[ti];
                    P.fbx[ti] = P.fbx[ti] + fx;
                    P.fby[ti] = P.fby[ti] + fy;
                    P.uwx[ti] = P.uwx[ti] + fx * vs;
                    P.uwy[ti] = P.uwy[ti] + fy * vs;
                } else {
                    P.fbx[ti] = ex;
                    P.fby[ti] = ey;
                }
            }
            __threadfence_block();
            __syncthreads();
            if (mdf) {
                prev = res;
                iter_count = it + 1;
                if (it + 1 >= min_iter && res < tol) {
                    converged = 1;
                    break;
                }
            }
        }
        if (mdf && t == 0) {
            I.stats[0] = iter_count;
            I.stats[1] = last;
            I.stats[2] = diverged;
            I.stats[3] = converged;
        }
    } else if (I.ibm == 3) {  // DFC, iblbm/ibm/dfc.py:319-341
        const double cs2 = 1.0 / 3.0, csc = I.dfc_scale;
        for (int k = t; k < Lb; k += NTI) {
            double fi[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) fi[q] = 0.0;
            for (int s = 0; s < K; ++s) {
                const int sl = S.tslot[S.cti[k * K + s]];
                const double w = S.cw[k * K + s];
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
                S.dfl[9 * k + q] = csc * lam * dev;
                const double sd = -(lam * dev * I.Larea);
                qx += sd * EX(q);
                qy += sd * EY(q);
            }
            S.q[2 * k] = csc * qx;
            S.q[2 * k + 1] = csc * qy;
        }
        __threadfence_block();
        __syncthreads();
        for (int ti = t; ti < nT; ti += NTI) {
            const int sl = S.tslot[ti];
            double e[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) e[q] = 0.0;
            for (int x = P.toff[ti], end = P.toff[ti + 1]; x < end; ++x) {
                const int k = P.gk[x];
                const double wl = P.gw[x];
#pragma unroll
                for (int q = 0; q < 9; ++q) e[q] += S.dfl[9 * k + q] * wl;
            }
            double fs[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                fs[q] = S.fs[q * I.box_cap + sl] + e[q];
                S.fs[q * I.box_cap + sl] = fs[q];
            }
            double r, u0, u1;  // strategies/dfc.py:50-54 macroscopic recompute
            macro9(fs, inc, r, u0, u1);
            S.ro[sl] = r;
            S.ux0[sl] = u0;
            S.uy0[sl] = u1;
        }
    }
    __threadfence_block();
    __syncthreads();
}

// MDF of a one-particle window with the learned force predictor (predictor modes 1-3), on
// the generic window: the same arithmetic, order and decisions as the fused window's
// predictor (fused_step.h), with the inference tiles' feature rows and outputs in global
// scratch (S.mlpx, S.mlpy) and the BF16 weight blocks and biases read from global memory.
// Each marker's stencil is in its thread's registers (Lb <= NTI).
//  pre-pass: the forcing slip q = R (ud - U) before forcing, its change since the previous
//    step, the FP32 features and sig (largest change over the marker's wavefront), the
//    training row (collect) or the network's predicted change of force (learned);
//  warm modes: the prediction F_prev (+ sig x net) is spread with weight 1 first, then at
//    least one true MDF correction follows and plain MDF restarts once from the uncorrected
//    velocity when the predicted start is no better than no forcing;
//  afterwards: the force before the last applied sub-step is kept as the next step's F_prev
//    (and its scaled change is the training target).
template <int A>
__device__ void ibm_core_learned(Inst &I, int nT, const IbmPtrs &P, double *red,
                                 const MarkerRegs<A> &M) {
    constexpr int K = (2 * A + 1) * (2 * A + 1);
    const int t = threadIdx.x, Lb = I.Lb;
    IbmScratch &S = I.s;
    const double dt = I.cp.dt, omega = I.omega;
    const int pmode = I.pmode;
    const bool warm = pmode >= 2, mk = t < Lb;
    const int n_iter = I.n_iter;
    const int min_iter = max(1, min(I.min_iter, n_iter));
    const long long pctr0 = S.pctr[0], pctr1 = S.pctr[1];
    const bool fprev_ok = pctr1 != 0;
    const bool record = pmode == 1 && fprev_ok;
    for (int ti = t; ti < nT; ti += NTI) {
        P.fbx[ti] = 0.0;
        P.fby[ti] = 0.0;
    }
    double R = 0.0;  // rho at the marker
    if (mk) {
#pragma unroll
        for (int s = 0; s < K; ++s) R += P.rot[M.ti[s]] * M.w[s];
    }
    // interpolation + force (df.py:239-263) of this marker from the working velocity
    auto interp_force = [&](double &lx, double &ly, double &Fx, double &Fy) {
        lx = 0.0;
        ly = 0.0;
#pragma unroll
        for (int s = 0; s < K; ++s) {
            lx += P.uwx[M.ti[s]] * M.w[s];
            ly += P.uwy[M.ti[s]] * M.w[s];
        }
        Fx = 2.0 * R * (M.udx - lx) / dt;
        Fy = 2.0 * R * (M.udy - ly) / dt;
    };
    // pre-pass: this marker's previous q and force, its q now and the features
    double pfx = 0.0, pfy = 0.0, ex = 0.0, ey = 0.0, ax = 0.0, ay = 0.0, rv0 = -INFINITY;
    if (mk) {
        double spx = 0.0, spy = 0.0;
        if (fprev_ok) {
            spx = S.fprev[2 * I.lb_cap + 2 * t];
            spy = S.fprev[2 * I.lb_cap + 2 * t + 1];
            if (warm) {
                pfx = S.fprev[2 * t];
                pfy = S.fprev[2 * t + 1];
            }
        }
        double lx, ly, Fx, Fy;
        interp_force(lx, ly, Fx, Fy);
        const double vx = M.udx - lx, vy = M.udy - ly;
        rv0 = sqrt(vx * vx + vy * vy);
        ax = R * (M.udx - lx);
        ay = R * (M.udy - ly);
        S.fprev[2 * I.lb_cap + 2 * t] = ax;  // the next step's previous q
        S.fprev[2 * I.lb_cap + 2 * t + 1] = ay;
        ex = ax - spx;
        ey = ay - spy;
    }
    __threadfence_block();
    // plain MDF's first residual, the bar a prediction must clear (warm modes)
    const double plain0 = warm ? block_reduce(rv0, red, true) : INFINITY;
    // the features feed a BF16 network: FP32 arithmetic from here
    const float fex = (float)ex, fey = (float)ey, fax = (float)ax, fay = (float)ay;
    const float a = sqrtf(fax * fax + fay * fay);
    const float sig =
        fmaxf(wave_maxf(sqrtf(fex * fex + fey * fey)) + 1e-9f * wave_maxf(a), 1e-30f);
    float feat[MLP_IN];
#pragma unroll
    for (int q = 0; q < MLP_IN; ++q) feat[q] = 0.f;
    if (mk) {
        feat[0] = fex / sig;
        feat[1] = fey / sig;
        feat[2] = a > 0.f ? fax / a : 0.f;
        feat[3] = a > 0.f ? fay / a : 0.f;
        const double mx = I.Lx[t], my = I.Ly[t];
        const float rx = (float)(mx - I.cxd), ry = (float)(my - I.cyd);
        const float rr = sqrtf(rx * rx + ry * ry);
        const double gx = mx / I.dx + 0.5, gy = my / I.dy + 0.5;
        feat[4] = rr > 0.f ? rx / rr : 0.f;
        feat[5] = rr > 0.f ? ry / rr : 0.f;
        feat[6] = (float)(gx - floor(gx) - 0.5);
        feat[7] = (float)(gy - floor(gy) - 0.5);
    }
    float *rowp = nullptr;  // this marker's training sample (collect mode)
    if (record && mk) {      // features now, sig in the target's place until the end
        rowp = S.data + ((pctr0 + t) % I.data_cap) * MLP_ROW;
#pragma unroll
        for (int q = 0; q < MLP_IN; ++q) rowp[q] = feat[q];
        rowp[MLP_IN] = sig;
    }
    if (pmode == 3) {  // inference: 16-marker tiles over the wavefronts
        const float *nrm = S.mlp_norm;
        const int ntile = (Lb + 15) >> 4;
        float xn[MLP_IN];
#pragma unroll
        for (int q = 0; q < MLP_IN; ++q) xn[q] = (feat[q] - nrm[q]) / nrm[MLP_IN + q];
        if (t < 16 * ntile) mlp_put_row(S.mlpx + 16 * t, xn);  // rows from Lb on: unused
        __threadfence_block();
        __syncthreads();  // feature rows published
        for (int tile = t >> 6; tile < ntile; tile += NTI / 64)
            mlp_tile16(reinterpret_cast<const mlp_short4 *>(S.mlp_wb), S.mlp_p + MLP_B1,
                       S.mlp_p + MLP_B2, S.mlp_p + MLP_B3, S.mlpx + 256 * tile,
                       S.mlpy + 16 * MLP_OUT * tile);
        __threadfence_block();
        __syncthreads();  // outputs published
        if (mk) {
            pfx += (double)sig * (double)(S.mlpy[MLP_OUT * t] * nrm[2 * MLP_IN]);
            pfy += (double)sig * (double)(S.mlpy[MLP_OUT * t + 1] * nrm[2 * MLP_IN + 1]);
        }
    }
    double facx = 0.0, facy = 0.0;  // this marker's force accumulated over the step
    if (warm) {  // the prediction, spread with weight 1 (not counted as a sub-step)
        if (mk) {
            P.Lfx[t] = pfx;
            P.Lfy[t] = pfy;
        }
        facx = pfx;
        facy = pfy;
        __threadfence_block();
        __syncthreads();
        for (int ti = t; ti < nT; ti += NTI) {
            double sx, sy;
            gather2(P, ti, sx, sy);
            const double vs = P.vs[ti];
            P.fbx[ti] = sx;
            P.fby[ti] = sy;
            P.uwx[ti] = P.uwx[ti] + sx * vs;
            P.uwy[ti] = P.uwy[ti] + sy * vs;
        }
        __threadfence_block();
        __syncthreads();
    }
    MdfCtl ctl;
    ctl.mdf = true;
    ctl.warm = warm;
    ctl.n_iter = n_iter;
    ctl.min_plain = min_iter;
    ctl.min_iter = warm ? 1 : min_iter;  // at least one true correction after a prediction
    ctl.tol = I.tol;
    ctl.plain0 = plain0;
    double lastx = 0.0, lasty = 0.0;  // the last applied sub-step's force increment
    bool first = true;
    for (int it = 0; it < n_iter;) {
        double Fx = 0.0, Fy = 0.0, rmax = -INFINITY;
        if (mk) {
            double lx, ly;
            interp_force(lx, ly, Fx, Fy);
            P.Lfx[t] = Fx;
            P.Lfy[t] = Fy;
            const double vx = M.udx - lx, vy = M.udy - ly;
            rmax = sqrt(vx * vx + vy * vy);
        }
        __threadfence_block();
        const double res = block_reduce(rmax, red, true);  // its barrier publishes Lf
        if (first && t == 0) S.pstat[2] += res;  // first residual after the prediction
        first = false;
        const int act = ctl.at_a(res, it);
        if (act == 1) break;
        if (act == 0) {
            lastx = omega * Fx;
            lasty = omega * Fy;
            facx += lastx;
            facy += lasty;
        } else {
            facx = 0.0;
            facy = 0.0;
            lastx = 0.0;
            lasty = 0.0;
        }
        for (int ti = t; ti < nT; ti += NTI) {
            if (act == 0) {  // mdf.py:197-200
                double ex2, ey2;
                gather2(P, ti, ex2, ey2);
                const double fx = ex2 * omega, fy = ey2 * omega;
                const double vs = P.vs[ti];
                P.fbx[ti] = P.fbx[ti] + fx;
                P.fby[ti] = P.fby[ti] + fy;
                P.uwx[ti] = P.uwx[ti] + fx * vs;
                P.uwy[ti] = P.uwy[ti] + fy * vs;
            } else {  // restart from the uncorrected velocity
                const int sl = S.tslot[ti];
                P.fbx[ti] = 0.0;
                P.fby[ti] = 0.0;
                P.uwx[ti] = S.ux0[sl];
                P.uwy[ti] = S.uy0[sl];
            }
        }
        __threadfence_block();
        __syncthreads();
        if (act == 2) {
            it = 0;
            continue;
        }
        if (ctl.at_b(res, it)) break;
        ++it;
    }
    if (t == 0) {
        I.stats[0] = ctl.iter_count;
        I.stats[1] = ctl.last;
        I.stats[2] = ctl.diverged;
        I.stats[3] = ctl.converged;
    }
    if (mk) {
        // the force before the last applied sub-step (the prediction itself after a learned
        // step that needed one correction)
        const double px = facx - lastx, py = facy - lasty;
        if (record) {  // training target: its scaled change since the previous step
            const double sg = (double)rowp[MLP_IN];
            rowp[MLP_IN] = (float)((px - S.fprev[2 * t]) / sg);
            rowp[MLP_IN + 1] = (float)((py - S.fprev[2 * t + 1]) / sg);
        }
        S.fprev[2 * t] = px;
        S.fprev[2 * t + 1] = py;
    }
    if (t == 0) {
        if (record) S.pctr[0] = pctr0 + Lb;
        S.pctr[1] = 1;
        S.pstat[0] += ctl.applied;
        S.pstat[1] += 1.0;
        S.pstat[3] += ctl.restarted;
    }
    __threadfence_block();
    __syncthreads();
}

// ------------------------------------------------------------------ the workgroup

// Window workgroup of one step.  It owns the touched nodes only (the bulk tiles skip exactly
// those): stream/close them, moments, the IBM stage, velocity correction, forced collision.
// The tables it reads were built by k_tables when the markers were set.
// PRED: compiled with the learned-predictor modes (one-particle MDF windows of at most NTI
// markers whose instance has a mode set; every other window runs ibm_core in either build).
template <typename ST, int MODEL, int A, bool LDS, bool PRED>
__device__ void ibm_role_impl(Inst &I, int parity, int diag, double *red, IbmShared &sh) {
    const int t = threadIdx.x;
    const ST *__restrict__ Asrc = buf<ST>(I, parity);
    ST *__restrict__ B = buf<ST>(I, parity ^ 1);
    IbmScratch &S = I.s;
    const bool inc = I.inc != 0;
    const double dt = I.cp.dt;
    const int Lb = I.Lb, nT = I.nT;
    const IbmPtrs P = ibm_ptrs<LDS>(I, sh);
    MarkerRegs<A> M;
    load_marker_regs<A>(I, M);  // in flight during P1
    if constexpr (LDS) ibm_stage_lists(I, nT, sh);

    // P1: stream + close the touched nodes, macroscopic moments, core inputs
    for (int ti = t; ti < nT; ti += NTI) {
        const int sl = S.tslot[ti];
        double fs[9];
        fstar_at(Asrc, I, I.bx0 + sl % I.bw, I.by0 + sl / I.bw, fs);
#pragma unroll
        for (int q = 0; q < 9; ++q) S.fs[q * I.box_cap + sl] = fs[q];
        double r, u0, u1;
        macro9(fs, inc, r, u0, u1);
        S.ro[sl] = r;
        S.ux0[sl] = u0;
        S.uy0[sl] = u1;
        P.uwx[ti] = u0;
        P.uwy[ti] = u1;
        P.rot[ti] = r;
        P.vs[ti] = dt / (2.0 * r);  // mdf.py:117
    }
    __threadfence_block();
    __syncthreads();
    if (PRED && I.pmode != 0 && I.ibm == 2 && I.np_part <= 1 && Lb <= NTI)
        ibm_core_learned<A>(I, nT, P, red, M);
    else
        ibm_core<A>(I, nT, P, red, false, M);

    // P3: velocity correction (DF/MDF), equilibrium, collision, store
    const bool correct = I.ibm == 1 || I.ibm == 2;
    double rho_acc = 0.0, fxs = 0.0, fys = 0.0, trq = 0.0;
    for (int t0 = t; t0 < nT; t0 += 2 * NTI) {
        const int t1 = t0 + NTI;
        const bool h1 = t1 < nT;
        const int sl0 = S.tslot[t0], sl1 = h1 ? S.tslot[t1] : sl0;
        double fa[9], fb[9], ra, rb = 1.0, ua0, ua1, ub0 = 0.0, ub1 = 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q) fa[q] = S.fs[q * I.box_cap + sl0];
        ra = S.ro[sl0];
        ua0 = S.ux0[sl0];
        ua1 = S.uy0[sl0];
        if (h1) {
#pragma unroll
            for (int q = 0; q < 9; ++q) fb[q] = S.fs[q * I.box_cap + sl1];
            rb = S.ro[sl1];
            ub0 = S.ux0[sl1];
            ub1 = S.uy0[sl1];
        }
        auto finish = [&](int ti, int sl, const double *fs, double r, double u0, double u1) {
            const int i = I.bx0 + sl % I.bw, j = I.by0 + sl / I.bw;
            const double fx = correct ? P.fbx[ti] : 0.0;
            const double fy = correct ? P.fby[ti] : 0.0;
            if (correct) {
                if (inc) {
                    u0 = u0 + 0.5 * fx * dt;
                    u1 = u1 + 0.5 * fy * dt;
                } else {
                    u0 = u0 + fx * dt / (2.0 * r);
                    u1 = u1 + fy * dt / (2.0 * r);
                }
                S.ux0[sl] = u0;
                S.uy0[sl] = u1;
            }
            S.fibx[sl] = fx;
            S.fiby[sl] = fy;
            double fe[9], out[9];
            if (MODEL != 2) feq9(r, u0, u1, inc, fe);
            if constexpr (MODEL == 0)
                collide_bgk<true>(I.cp, fs, fe, u0, u1, fx, fy, out);
            else if constexpr (MODEL == 1)
                collide_trt<true>(I.cp, fs, fe, u0, u1, fx, fy, out);
            else if constexpr (MODEL == 2)
                collide_cm<true>(I.cp, fs, u0, u1, fx, fy, out);
            else
                collide<true>(I.cp, fs, fe, u0, u1, fx, fy, out);
            store9(B, I, i, j, out);
            if (diag) {
                rho_acc += r;
                fxs += fx;
                fys += fy;
                trq += ((double)i - I.cxc) * fy - ((double)j - I.cyc) * fx;
            }
        };
        finish(t0, sl0, fa, ra, ua0, ua1);
        if (h1) finish(t1, sl1, fb, rb, ub0, ub1);
    }
    if (diag) {
        const double rs = block_reduce(rho_acc, red, false);
        double f0 = block_reduce(fxs, red, false), f1 = block_reduce(fys, red, false);
        const double tq = block_reduce(trq, red, false);
        double tqq = 0.0;
        if (I.ibm == 3) {
            double qx = 0.0, qy = 0.0, tk = 0.0;
            for (int k = t; k < Lb; k += NTI) {
                qx += S.q[2 * k];
                qy += S.q[2 * k + 1];
                tk += (I.Lx[k] - I.cxd) / I.dx * S.q[2 * k + 1] -
                      (I.Ly[k] - I.cyd) / I.dy * S.q[2 * k];
            }
            f0 = block_reduce(qx, red, false);
            f1 = block_reduce(qy, red, false);
            tqq = block_reduce(tk, red, false);
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

template <typename ST, int MODEL, bool PRED>
__device__ void ibm_role(Inst &I, int parity, int diag, double *red) {
    __shared__ IbmShared sh;
    const bool fits = ibm_fits_lds(I, I.nT);
    if (I.delta == 1) {
        if (fits)
            ibm_role_impl<ST, MODEL, 1, true, PRED>(I, parity, diag, red, sh);
        else
            ibm_role_impl<ST, MODEL, 1, false, PRED>(I, parity, diag, red, sh);
    } else {
        if (fits)
            ibm_role_impl<ST, MODEL, 2, true, PRED>(I, parity, diag, red, sh);
        else
            ibm_role_impl<ST, MODEL, 2, false, PRED>(I, parity, diag, red, sh);
    }
}

// Tables of the fused window path (see fused_step.h) and its flag: DF/MDF, Lb <= FAST_LB,
// one node thread per touched node, at most FAST_G contributions per node.
template <int A>
__device__ void ibm_fast_lists(Inst &I, int *redi) {
    constexpr int K = (2 * A + 1) * (2 * A + 1);
    IbmScratch &S = I.s;
    const int t = threadIdx.x, Lb = I.Lb;
    const int nT = S.tindex[I.bw * I.bh];
    __syncthreads();
    if (t == 0) redi[0] = 0;
    __syncthreads();
    int mx = 0;
    for (int ti = t; ti < nT; ti += NTI) mx = max(mx, S.toff[ti + 1] - S.toff[ti]);
    atomicMax(&redi[0], mx);
    __syncthreads();
    const int mwv = (Lb + 63) >> 6;  // marker wavefronts
    const int fast = (I.ibm == 1 || I.ibm == 2) && I.np_part <= 1 && !I.df_multi &&
                     Lb <= FAST_LB &&
                     nT <= NTF - 64 * mwv && redi[0] <= FAST_G;
    if (fast) {
        for (int ti = t; ti < nT; ti += NTI) {
            const int o = S.toff[ti], n = S.toff[ti + 1] - o;
            for (int e = 0; e < FAST_G; ++e) {
                S.pk[e * FAST_T + ti] = e < n ? S.gk[o + e] : Lb;  // padding: zero force
                S.pw[e * FAST_T + ti] = e < n ? S.gw[o + e] : 0.0;
            }
        }
        for (int c = t; c < K * Lb; c += NTI) {  // marker-major -> stencil-major
            const int k = c / K, sp = c - k * K;
            S.mti[sp * Lb + k] = S.cti[c];
            S.mw[sp * Lb + k] = S.cw[c];
        }
    }
    if (t == 0) I.fast = fast;
    __threadfence_block();
    __syncthreads();
}

// Window tables (and the DFC lambda) for new markers; zeroes the box fib.
__device__ inline void ibm_tables_role(Inst &I, int *redi) {
    if (I.delta == 1) {
        ibm_build_tables<1>(I, redi);
        if (I.ibm == 3) ibm_dfc_lambda<1>(I, I.s.tindex[I.bw * I.bh], I.Lb);
    } else {
        ibm_build_tables<2>(I, redi);
        if (I.ibm == 3) ibm_dfc_lambda<2>(I, I.s.tindex[I.bw * I.bh], I.Lb);
    }
    if (I.delta == 1)
        ibm_fast_lists<1>(I, redi);
    else
        ibm_fast_lists<2>(I, redi);
    for (int sl = threadIdx.x; sl < I.bw * I.bh; sl += NTI) {
        I.s.fibx[sl] = 0.0;
        I.s.fiby[sl] = 0.0;
    }
    if (I.s.pfb && I.np_part > 1 && I.ibm == 2)  // per-particle fields of the coupled MDF
        for (int p = 0; p < 2 * I.np_part; ++p)
            for (int sl = threadIdx.x; sl < I.bw * I.bh; sl += NTI)
                I.s.pfb[(size_t)p * I.box_cap + sl] = 0.0;
}

}  // namespace iblbm
