// fused_step.h -- the window role of the one-launch time step (k_step).
//
// A context whose windows all pass ibm_fast_lists runs each time step as ONE launch of
// NTF-thread workgroups: the first workgroups run one instance's window each, the others run
// bulk tiles.  One launch per step leaves no cross-stream wait between steps, and the
// register budget of 128 VGPRs (4 wavefronts per SIMD) keeps the bulk tiles streaming.
//
// Window workgroup: the first ceil(Lb/64) wavefronts are marker wavefronts (thread t < Lb owns
// marker t: interpolation, force, residual); every other thread owns at most one touched node
// (its gather list, MDF working velocity and accumulated force in registers).  Marker stencils,
// the working velocity, the marker forces and the post-stream populations live in LDS.  Both
// roles run the same barrier sequence and take the same (uniform) loop exits.  Arithmetic and
// summation orders are those of ibm_role_impl / ibm_core (DF: iblbm/ibm/df.py:94-154,
// MDF: iblbm/ibm/mdf.py:175-229, velocity correction: iblbm/ibm/dispatch.py:18-33).
#pragma once
#include "ibm_block.h"
#include "mlp.h"

namespace iblbm {

struct FusedShared {
    alignas(16) double uw[2 * FAST_T];         // MDF working velocity of the touched nodes
    alignas(16) double lf[2 * (FAST_LB + 1)];  // marker forces; entry Lb is the zero force
    double rot[FAST_T];                         // rho of the touched nodes
    double mw[FAST_K * FAST_LB];                // marker stencil weights, [s * Lb + k]
    int mti[FAST_K * FAST_LB];                  // marker stencil touched indices, [s * Lb + k]
    double fs[9 * FAST_T];                      // post-stream populations of the touched nodes
    alignas(16) unsigned short mlpx[FAST_LB * 16];  // predictor feature rows (bf16)
    float mlpy[FAST_LB * MLP_OUT];                   // predictor outputs
    MlpLds mlpw;                                     // predictor weights (learned mode)
    // predictor inputs, copied from global memory straight into LDS during P1: per marker
    // the previous step's force before its last sub-step and its forcing slip (fprev), the
    // position, and the normalisation
    alignas(16) double pf[2 * FAST_LB], pq[2 * FAST_LB], plx[FAST_LB], ply[FAST_LB];
    alignas(16) float pnrm[MLP_NORM + 2];
};

// Copies the descriptor fields the window reads into a local descriptor, in one batch at the
// start (reads of the global one after a store would be reissued one at a time).
__device__ inline void light_copy(const Inst &G, Inst &L) {
    L.nx = G.nx;
    L.ny = G.ny;
    L.pitch = G.pitch;
    L.qs = G.qs;
    L.f[0] = G.f[0];
    L.f[1] = G.f[1];
    L.bc = G.bc;
    L.inc = G.inc;
    L.cp = G.cp;
#pragma unroll
    for (int q = 0; q < 9; ++q) L.feq_bc[q] = G.feq_bc[q];
    L.ibm = G.ibm;
    L.Lb = G.Lb;
    L.bx0 = G.bx0;
    L.by0 = G.by0;
    L.bw = G.bw;
    L.bh = G.bh;
    L.box_cap = G.box_cap;
    L.n_iter = G.n_iter;
    L.min_iter = G.min_iter;
    L.nT = G.nT;
    L.omega = G.omega;
    L.tol = G.tol;
    L.ud = G.ud;
    L.s = G.s;
    L.stats = G.stats;
    L.diag = G.diag;
    L.cxc = G.cxc;
    L.cyc = G.cyc;
    L.cxd = G.cxd;
    L.cyd = G.cyd;
    L.dx = G.dx;
    L.dy = G.dy;
    L.Lx = G.Lx;
    L.Ly = G.Ly;
    L.pmode = G.pmode;
    L.lb_cap = G.lb_cap;
    L.data_cap = G.data_cap;
}

// Sum over the NTF threads: fixed shuffle tree per wavefront, then the wavefronts in order
// (the result is valid in thread 0).
__device__ inline double fused_block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        r = red[0];
        for (int k = 1; k < NTF / 64; ++k) r = r + red[k];
    }
    __syncthreads();
    return r;
}

// 16 bytes per lane from global memory straight into LDS (global_load_lds_dwordx4): lane l of
// the calling wavefront writes lds_wave + 16 l, where lds_wave is the same for the wavefront.
// The data arrive with the wavefront's other vector loads; the compiler waits for them before
// a barrier that follows.
__device__ inline void lds_dma16(const void *src, void *lds_wave) {
    typedef __attribute__((address_space(1))) void GVoid;
    typedef __attribute__((address_space(3))) void LVoid;
    __builtin_amdgcn_global_load_lds((GVoid *)src, (LVoid *)lds_wave, 16, 0, 0);
}

// One 16-marker tile of the learned predictor's inference on the calling wavefront.
__device__ inline void mlp_infer_tile(FusedShared &sh, int tile) {
    mlp_tile16(sh.mlpw.blk, sh.mlpw.bias, sh.mlpw.bias + MLP_H, sh.mlpw.bias + 2 * MLP_H,
               sh.mlpx + 256 * tile, sh.mlpy + 16 * MLP_OUT * tile);
}

// PRED: compiled with the learned-predictor modes (instances whose mode is off run plain MDF
// in either build; without PRED the predictor's registers stay out of the step kernel).
template <typename ST, int MODEL, int A, bool PRED>
__device__ __attribute__((always_inline)) void ibm_fused_role(Inst &Ig, int parity, int diag,
                                                             double *red, FusedShared &sh) {
    constexpr int K = (2 * A + 1) * (2 * A + 1);
    constexpr int GRP = 8;  // stencil points per batch of LDS loads
    Inst I;
    light_copy(Ig, I);
    const int t = threadIdx.x;
    IbmScratch &S = I.s;
    const bool inc = I.inc != 0;
    const double dt = I.cp.dt;
    const int Lb = I.Lb, nT = I.nT;
    const int mwv = (Lb + 63) >> 6;
    const bool marker_wave = (t >> 6) < mwv;
    const int n = t - 64 * mwv;  // touched node of a node thread
    const bool hn = !marker_wave && n < nT;
    double2 *uw = reinterpret_cast<double2 *>(sh.uw);
    double2 *lf = reinterpret_cast<double2 *>(sh.lf);
    unsigned long long *rbits = reinterpret_cast<unsigned long long *>(red) + NTF / 64;
    const bool mdf = I.ibm == 2;
    const bool finc = mdf ? false : inc;  // MDF sub-steps are compressible (mdf.py:177-181)
    const int n_iter = mdf ? I.n_iter : 1;
    const int min_iter = max(1, min(I.min_iter, n_iter));
    const double tol = I.tol;

    // Every global load of the prologue is issued before the first wait: a node's box slot
    // (first: P1's pulls wait for it alone), the marker stencils (for LDS), a marker's desired
    // velocity, a node's gather list.
    int sl = 0;
    if (hn) sl = S.tslot[n];
    constexpr int SU = (FAST_K * FAST_LB + NTF - 1) / NTF;  // stencil entries per thread
    int st_i[SU];
    double st_w[SU];
#pragma unroll
    for (int u = 0; u < SU; ++u) {
        const int e = t + u * NTF;
        if (e < K * Lb) {
            st_i[u] = S.mti[e];
            st_w[u] = S.mw[e];
        }
    }
    double udx = 0.0, udy = 0.0;
    // predictor inputs straight to LDS (no registers held across P1): marker wavefronts take
    // their markers' previous force and q; wavefront 0 the positions and normalisation; in
    // learned mode wavefronts 2.. the weight blocks and 3 more the biases
    long long pctr0 = 0, pctr1 = 0;  // the predictor's counters, read before P1's loads
    if (PRED && I.ibm == 2 && I.pmode != 0) {
        pctr0 = S.pctr[0];
        pctr1 = S.pctr[1];
        const int wv = t >> 6, ln = t & 63;
        if (t < Lb) {
            lds_dma16(S.fprev + 2 * t, sh.pf + 128 * wv);
            lds_dma16(S.fprev + 2 * I.lb_cap + 2 * t, sh.pq + 128 * wv);
        }
        if (wv == 0 && 2 * ln < Lb) {  // two positions per lane
            lds_dma16(I.Lx + 2 * ln, sh.plx);
            lds_dma16(I.Ly + 2 * ln, sh.ply);
        }
        if (wv == 0 && ln < (MLP_NORM + 3) / 4) lds_dma16(S.mlp_norm + 4 * ln, sh.pnrm);
        if (I.pmode == 3) {
            constexpr int NB = MLP_NBLK * 64 * 8 / 1024;  // 1 KB per wavefront instruction
            if (wv >= 2 && wv < 2 + NB)
                lds_dma16(reinterpret_cast<const char *>(S.mlp_wb) + 1024 * (wv - 2) + 16 * ln,
                          reinterpret_cast<char *>(sh.mlpw.blk) + 1024 * (wv - 2));
            // biases: b1 from lanes 0-15 of one wavefront; b2 from lanes 0-15 and b3 from lane
            // 16 of the next (lane 16 lands 64 floats after b2's start, where b3 belongs)
            static_assert(2 + NB + 1 < NTF / 64, "bias wavefronts");
            if (wv == 2 + NB && ln < 16) lds_dma16(S.mlp_p + MLP_B1 + 4 * ln, sh.mlpw.bias);
            if (wv == 3 + NB && ln <= 16)
                lds_dma16(S.mlp_p + (ln < 16 ? MLP_B2 + 4 * ln : MLP_B3), sh.mlpw.bias + MLP_H);
        }
    }
    if (t < Lb) {
        udx = I.ud[2 * t];
        udy = I.ud[2 * t + 1];
    }
    int kk[FAST_G];
    double ww[FAST_G];
    if (hn) {
#pragma unroll
        for (int e = 0; e < FAST_G; ++e) {
            kk[e] = S.pk[e * FAST_T + n];
            ww[e] = S.pw[e * FAST_T + n];
        }
    }

    // P1: each node thread pulls its own touched node (its box slot came with the prologue's
    // loads), and the stencil tables go to LDS
    const ST *__restrict__ Asrc = reinterpret_cast<const ST *>(parity ? I.f[1] : I.f[0]);
    const bool interior = I.bx0 >= 1 && I.by0 >= 1 && I.bx0 + I.bw <= I.nx - 1 &&
                          I.by0 + I.bh <= I.ny - 1;
    if (hn) {
        const int i = I.bx0 + sl % I.bw, j = I.by0 + sl / I.bw;
        double fsn[9];
        if (interior)
            pull_interior(Asrc, I, i, j, fsn);
        else
            fstar_at(Asrc, I, i, j, fsn);
        double r, u0, u1;
#pragma unroll
        for (int q = 0; q < 9; ++q) sh.fs[q * FAST_T + n] = fsn[q];
        macro9(fsn, inc, r, u0, u1);
        uw[n] = make_double2(u0, u1);
        sh.rot[n] = r;
    }
#pragma unroll
    for (int u = 0; u < SU; ++u) {
        const int e = t + u * NTF;
        if (e < K * Lb) {
            sh.mti[e] = st_i[u];
            sh.mw[e] = st_w[u];
        }
    }
    if (t < 3) rbits[t] = 0ull;
    if (t == 0) lf[Lb] = make_double2(0.0, 0.0);  // target of the padded gather entries
    // the predictor's copies into LDS come from other wavefronts too: every wavefront's own
    // have landed before it reaches the barrier
    if (PRED && I.ibm == 2 && I.pmode != 0) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();  // P1 done

    double rho_acc = 0.0, fxs = 0.0, fys = 0.0, trq = 0.0;
    const int pmode = (PRED && mdf) ? I.pmode : 0;  // predictor: 1 collect, 2 warm, 3 learned
    const bool warm = pmode >= 2;
    MdfCtl ctl;
    ctl.mdf = mdf;
    ctl.warm = warm;
    ctl.n_iter = n_iter;
    ctl.min_plain = min_iter;
    ctl.min_iter = warm ? 1 : min_iter;  // at least one true correction after a prediction
    ctl.tol = tol;
    int seq = 0;  // residual slot sequence (slots alternate)
    if (marker_wave) {
        // Predictor (MDF with a mode set), computed ahead of the loop while few registers are
        // taken.  The forcing slip q = R (ud - U) (U interpolated before forcing, R the marker
        // density, 1 for the incompressible form) is what the first MDF sub-step turns into a
        // force, and a fixed number of sub-steps maps q to the force linearly: the change of
        // force is a geometry-set operator on the change of q.  The force learned is plain
        // MDF's force before its last sub-step (N - 1 sub-steps from zero map q to it
        // linearly), so that the prediction plus the one true correction that follows
        // reproduces plain MDF's N-sub-step result, truncation included.  fprev holds per
        // marker the previous step's force before its last sub-step (2 Lb) and q (2 Lb after
        // it); only the marker's
        // own thread touches its entries.  Features: change of q since the previous step /
        // sig, direction of q, marker direction from the body centre, sub-cell offset.  The
        // network predicts (change of force) / sig with sig = max over the wavefront's markers
        // of |change of q| (+ 1e-9 max |q|), exact in scale while the change decays by orders of
        // magnitude along a run.
        // With a warm start the prediction F = F_prev (+ sig x the network's output) is
        // published here; the loop's first pass spreads it with weight 1 and does not count it
        // as a sub-step.
        const bool fprev_ok = pmode != 0 && pctr1 != 0;
        const bool record = pmode == 1 && fprev_ok;
        float *rowp = nullptr;  // this marker's training sample (collect mode)
        double facx = 0.0, facy = 0.0;
        int ti[K];
        if (t < Lb) {
#pragma unroll
            for (int s = 0; s < K; ++s) ti[s] = sh.mti[s * Lb + t];
        }
        double R = 0.0;
        if (t < Lb) {  // rho at the marker
            double rs[K];
#pragma unroll
            for (int s = 0; s < K; ++s) rs[s] = sh.rot[ti[s]];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int s = 0; s < K; ++s) R += rs[s] * sh.mw[s * Lb + t];
        }
        const bool unit_dt = dt == 1.0;  // a division by 1 is exact and skipped
        // interpolation + force (df.py:239-263) of this marker from uw, loads batched by GRP
        auto interp_force = [&](double &lx, double &ly, double &Fx, double &Fy) {
            lx = 0.0;
            ly = 0.0;
#pragma unroll
            for (int g0 = 0; g0 < K; g0 += GRP) {
                double2 u[GRP];
                double w[GRP];
#pragma unroll
                for (int s = g0; s < g0 + GRP && s < K; ++s) {
                    u[s - g0] = uw[ti[s]];
                    w[s - g0] = sh.mw[s * Lb + t];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int s = g0; s < g0 + GRP && s < K; ++s) {
                    lx += u[s - g0].x * w[s - g0];
                    ly += u[s - g0].y * w[s - g0];
                }
            }
            if (finc) {
                Fx = 2.0 * (udx - lx);
                Fy = 2.0 * (udy - ly);
            } else {
                Fx = 2.0 * R * (udx - lx);
                Fy = 2.0 * R * (udy - ly);
            }
            if (!unit_dt) {
                Fx = Fx / dt;
                Fy = Fy / dt;
            }
        };
        if (PRED && pmode != 0) {
            // this marker's inputs, already in LDS: previous q and force, position,
            // normalisation
            double spx = 0.0, spy = 0.0, pfx = 0.0, pfy = 0.0, mx = 0.0, my = 0.0;
            if (t < Lb) {
                if (fprev_ok) {
                    spx = sh.pq[2 * t];
                    spy = sh.pq[2 * t + 1];
                    if (warm) {
                        pfx = sh.pf[2 * t];
                        pfy = sh.pf[2 * t + 1];
                    }
                }
                mx = sh.plx[t];
                my = sh.ply[t];
            }
            const float *nrm = sh.pnrm;
            double ax = 0.0, ay = 0.0, ex = 0.0, ey = 0.0;
            double rv0 = 0.0;  // plain MDF's first residual at this marker
            if (t < Lb) {  // the loop's own interpolation (register stencils, batched loads)
                double lx, ly, Fx, Fy;
                interp_force(lx, ly, Fx, Fy);
                const double rm = finc ? 1.0 : R;
                const double vx = udx - lx, vy = udy - ly;
                rv0 = sqrt(vx * vx + vy * vy);
                ax = rm * (udx - lx);
                ay = rm * (udy - ly);
                S.fprev[2 * I.lb_cap + 2 * t] = ax;  // the next step's previous q
                S.fprev[2 * I.lb_cap + 2 * t + 1] = ay;
                ex = ax - spx;
                ey = ay - spy;
            }
            if (warm) {  // block max into the third residual slot, read after barrier A
                const unsigned long long wb0 = dpp_max_u64_lane63(
                    t < Lb ? (unsigned long long)__double_as_longlong(rv0) : 0ull);
                if ((t & 63) == 63) atomicMax(&rbits[2], wb0);
            }
            // the features feed a BF16 network: FP32 arithmetic from here
            const float fex = (float)ex, fey = (float)ey, fax = (float)ax, fay = (float)ay;
            const float a = sqrtf(fax * fax + fay * fay);
            const float sig =
                fmaxf(wave_maxf(sqrtf(fex * fex + fey * fey)) + 1e-9f * wave_maxf(a), 1e-30f);
            float feat[MLP_IN];
#pragma unroll
            for (int q = 0; q < MLP_IN; ++q) feat[q] = 0.f;
            if (t < Lb) {
                feat[0] = fex / sig;
                feat[1] = fey / sig;
                feat[2] = a > 0.f ? fax / a : 0.f;
                feat[3] = a > 0.f ? fay / a : 0.f;
                const float rx = (float)(mx - I.cxd), ry = (float)(my - I.cyd);
                const float rr = sqrtf(rx * rx + ry * ry);
                const double gx = mx / I.dx + 0.5, gy = my / I.dy + 0.5;
                feat[4] = rr > 0.f ? rx / rr : 0.f;
                feat[5] = rr > 0.f ? ry / rr : 0.f;
                feat[6] = (float)(gx - floor(gx) - 0.5);
                feat[7] = (float)(gy - floor(gy) - 0.5);
                if (record) {  // features now, sig in the target's place until the end
                    rowp = S.data + ((pctr0 + t) % I.data_cap) * MLP_ROW;
#pragma unroll
                    for (int q = 0; q < MLP_IN; ++q) rowp[q] = feat[q];
                    rowp[MLP_IN] = sig;
                }
            }
            if (pmode == 3) {  // inference: one 16-marker tile per wavefront (barriers X, Y)
                float xn[MLP_IN];
#pragma unroll
                for (int q = 0; q < MLP_IN; ++q) xn[q] = (feat[q] - nrm[q]) / nrm[MLP_IN + q];
                mlp_put_row(sh.mlpx + 16 * t, xn);  // rows from Lb on are computed, unused
                __syncthreads();  // X: feature rows published
                if ((t >> 6) < ((Lb + 15) >> 4)) mlp_infer_tile(sh, t >> 6);
                __syncthreads();  // Y: outputs published
                if (t < Lb) {
                    pfx += (double)sig * (double)(sh.mlpy[MLP_OUT * t] * nrm[2 * MLP_IN]);
                    pfy += (double)sig * (double)(sh.mlpy[MLP_OUT * t + 1] * nrm[2 * MLP_IN + 1]);
                }
            }
            if (warm) {  // published here; read by the node threads after barrier A
                if (t < Lb) lf[t] = make_double2(pfx, pfy);
                facx = pfx;
                facy = pfy;
            }
        }
        // the last applied sub-step's force increment: plain MDF's force before it is what the
        // network learns (collect mode), so that the one true correction a learned step always
        // applies lands on plain MDF's own answer, truncation included
        double lastx = 0.0, lasty = 0.0;
        for (int it = 0; it < n_iter; ++seq) {
            const bool predict = PRED && warm && seq == 0;
            double rv = 0.0, Fx = 0.0, Fy = 0.0;
            if (t < Lb && !predict) {
                double lx, ly;
                interp_force(lx, ly, Fx, Fy);
                lf[t] = make_double2(Fx, Fy);
                const double vx = udx - lx, vy = udy - ly;
                rv = sqrt(vx * vx + vy * vy);
            }
            // L-inf residual: an unsigned max over the bit patterns (non-negative doubles order
            // like their bits and every NaN pattern lies above +inf, so NaN propagates as in
            // np.max); a marker-less lane contributes 0, the identity; two slots alternate
            const unsigned long long wb = dpp_max_u64_lane63(
                t < Lb ? (unsigned long long)__double_as_longlong(rv) : 0ull);
            if ((t & 63) == 63) atomicMax(&rbits[seq & 1], wb);
            __syncthreads();  // A: forces and residual published
            const double res = __longlong_as_double((long long)rbits[seq & 1]);
            if (t == 0) {
                rbits[(seq + 1) & 1] = 0ull;  // its last reads preceded barrier B
                // first residual after the prediction (warm) or of plain MDF
                if (PRED && pmode != 0 && seq == (warm ? 1 : 0)) S.pstat[2] += res;
            }
            if (predict) {
                __syncthreads();  // B: predicted velocities applied
                continue;
            }
            if (warm) ctl.plain0 = __longlong_as_double((long long)rbits[2]);
            const int act = ctl.at_a(res, it);
            if (act == 1) break;
            if (act == 0) {
                lastx = I.omega * Fx;
                lasty = I.omega * Fy;
                facx += lastx;
                facy += lasty;
            } else {
                facx = 0.0;
                facy = 0.0;
                lastx = 0.0;
                lasty = 0.0;
            }
            __syncthreads();  // B: velocities updated
            if (act == 2) {
                it = 0;
                continue;
            }
            if (ctl.at_b(res, it)) break;
            ++it;
        }
        if (mdf && t == 0) {
            I.stats[0] = ctl.iter_count;
            I.stats[1] = ctl.last;
            I.stats[2] = ctl.diverged;
            I.stats[3] = ctl.converged;
        }
        if (pmode != 0) {
            if (t < Lb) {
                // the force before the last applied sub-step (the prediction itself after a
                // learned step that needed one correction)
                const double px = facx - lastx, py = facy - lasty;
                if (record) {  // training target: its scaled change since the previous step
                    const double sig = (double)rowp[MLP_IN];
                    rowp[MLP_IN] = (float)((px - S.fprev[2 * t]) / sig);
                    rowp[MLP_IN + 1] = (float)((py - S.fprev[2 * t + 1]) / sig);
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
        }
    } else {
        ST *__restrict__ B = reinterpret_cast<ST *>(parity ? I.f[0] : I.f[1]);
        const double omega = I.omega;
        double r = 1.0, u0 = 0.0, u1 = 0.0;
        if (hn) {
            const double2 v = uw[n];
            r = sh.rot[n];
            u0 = v.x;
            u1 = v.y;
        }
        double x = u0, y = u1, fx = 0.0, fy = 0.0;
        const double vs = dt / (2.0 * r);  // mdf.py:117
        // spread as a gather in np.add.at order (df.py:269-283)
        auto gather = [&](double &sx, double &sy) {
            double2 g[FAST_G];
#pragma unroll
            for (int e = 0; e < FAST_G; ++e) g[e] = lf[kk[e]];
            __builtin_amdgcn_sched_barrier(0);
            // padded entries point at the zero force with weight 0: they add +0.0, which
            // leaves the sum's bits unchanged (it starts at +0.0 and so is never -0.0)
            sx = 0.0;
            sy = 0.0;
#pragma unroll
            for (int e = 0; e < FAST_G; ++e) {
                sx += g[e].x * ww[e];
                sy += g[e].y * ww[e];
            }
        };
        if (PRED && pmode == 3) {  // the node wavefronts take the other inference tiles
            __syncthreads();  // X
            if ((t >> 6) < ((Lb + 15) >> 4)) mlp_infer_tile(sh, t >> 6);
            __syncthreads();  // Y
        }
        for (int it = 0; it < n_iter; ++seq) {
            __syncthreads();  // A
            if (PRED && warm && seq == 0) {  // the predicted force and the velocity it implies
                if (hn) {
                    double sx, sy;
                    gather(sx, sy);
                    fx = sx;
                    fy = sy;
                    x = u0 + sx * vs;
                    y = u1 + sy * vs;
                    uw[n] = make_double2(x, y);
                }
                __syncthreads();  // B
                continue;
            }
            const double res = __longlong_as_double((long long)rbits[seq & 1]);
            if (warm) ctl.plain0 = __longlong_as_double((long long)rbits[2]);
            const int act = ctl.at_a(res, it);
            if (act == 1) break;
            if (hn) {
                if (act == 0) {
                    double sx, sy;
                    gather(sx, sy);
                    if (mdf) {  // mdf.py:197-200
                        const double gx = sx * omega, gy = sy * omega;
                        fx = fx + gx;
                        fy = fy + gy;
                        x = x + gx * vs;
                        y = y + gy * vs;
                    } else {
                        fx = sx;
                        fy = sy;
                    }
                } else {  // restart from the uncorrected velocity
                    fx = 0.0;
                    fy = 0.0;
                    x = u0;
                    y = u1;
                }
                uw[n] = make_double2(x, y);
            }
            __syncthreads();  // B
            if (act == 2) {
                it = 0;
                continue;
            }
            if (ctl.at_b(res, it)) break;
            ++it;
        }
        // P3: velocity correction, equilibrium, forced collision, store
        if (hn) {
            const int i = I.bx0 + sl % I.bw, j = I.by0 + sl / I.bw;
            double fs[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                fs[q] = sh.fs[q * FAST_T + n];
                S.fs[q * I.box_cap + sl] = fs[q];  // the step's state, for downloads
            }
            double v0 = u0, v1 = u1;
            if (inc) {
                v0 = v0 + 0.5 * fx * dt;
                v1 = v1 + 0.5 * fy * dt;
            } else {
                v0 = v0 + fx * dt / (2.0 * r);
                v1 = v1 + fy * dt / (2.0 * r);
            }
            S.ro[sl] = r;
            S.ux0[sl] = v0;
            S.uy0[sl] = v1;
            S.fibx[sl] = fx;
            S.fiby[sl] = fy;
            double fe[9], out[9];
            if (MODEL != 2) feq9(r, v0, v1, inc, fe);
            if constexpr (MODEL == 0)
                collide_bgk<true>(I.cp, fs, fe, v0, v1, fx, fy, out);
            else if constexpr (MODEL == 1)
                collide_trt<true>(I.cp, fs, fe, v0, v1, fx, fy, out);
            else if constexpr (MODEL == 2)
                collide_cm<true>(I.cp, fs, v0, v1, fx, fy, out);
            else
                collide<true>(I.cp, fs, fe, v0, v1, fx, fy, out);
            store9(B, I, i, j, out);
            if (diag) {
                rho_acc = r;
                fxs = fx;
                fys = fy;
                trq = ((double)i - I.cxc) * fy - ((double)j - I.cyc) * fx;
            }
        }
    }
    if (diag) {
        const double rs = fused_block_sum(rho_acc, red);
        const double f0 = fused_block_sum(fxs, red);
        const double f1 = fused_block_sum(fys, red);
        const double tq = fused_block_sum(trq, red);
        if (t == 0) {
            I.diag[0] = rs;
            I.diag[1] = f0;
            I.diag[2] = f1;
            I.diag[3] = tq;
            I.diag[4] = 0.0;
        }
    }
}

template <typename ST, int MODEL, bool PRED>
__device__ __attribute__((always_inline)) void ibm_fused_dispatch(Inst &I, int parity, int diag,
                                                                 double *red, FusedShared &sh) {
    if (I.delta == 1)
        ibm_fused_role<ST, MODEL, 1, PRED>(I, parity, diag, red, sh);
    else
        ibm_fused_role<ST, MODEL, 2, PRED>(I, parity, diag, red, sh);
}

}  // namespace iblbm
