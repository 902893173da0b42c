// scratch_alloc.h -- host-side allocation of an instance's immersed-boundary scratch.
#pragma once
#include <hip/hip_runtime.h>

#include "step_kernel.h"

namespace iblbm {

// Allocates every IbmScratch array (plus markers, stats and diag) for lb_cap markers and a
// window of box_cap nodes with allocator `get(bytes) -> void*` (nullptr on failure).
template <typename Get>
inline bool alloc_ibm_scratch(Inst &d, int lb_cap, int box_cap, Get get) {
    const size_t L = (size_t)(lb_cap > 0 ? lb_cap : 1), B = (size_t)(box_cap > 0 ? box_cap : 1);
    const size_t KL = 25 * L;
    bool ok = true;
    auto D = [&](double **p, size_t n) {
        *p = (double *)get(sizeof(double) * n);
        ok = ok && *p;
    };
    auto I = [&](int **p, size_t n) {
        *p = (int *)get(sizeof(int) * n);
        ok = ok && *p;
    };
    IbmScratch &s = d.s;
    double *lx, *ly, *ud;
    D(&lx, L + 2);  // the predictor copies positions two at a time (16 bytes)
    D(&ly, L + 2);
    D(&ud, 2 * L);
    d.Lx = lx;
    d.Ly = ly;
    d.ud = ud;
    D(&s.fs, 9 * B);
    D(&s.ro, B);
    D(&s.ux0, B);
    D(&s.uy0, B);
    D(&s.fibx, B);
    D(&s.fiby, B);
    I(&s.cnt, B);
    I(&s.cur, B);
    I(&s.tindex, B + 1);
    I(&s.tslot, B);
    I(&s.toff, B + 4);  // padded to 16-byte copy granules
    D(&s.uwx, B);
    D(&s.uwy, B);
    D(&s.vs, B);
    D(&s.tfibx, B);
    D(&s.tfiby, B);
    D(&s.rot, B);
    D(&s.wt, B);
    D(&s.cw, KL);
    I(&s.cslot, KL);
    I(&s.cti, KL);
    I(&s.gc, KL);
    I(&s.gk, KL + 4);
    D(&s.gw, KL + 2);
    D(&s.Lux, L);
    D(&s.Luy, L);
    D(&s.R, L);
    D(&s.Lfx, L);
    D(&s.Lfy, L);
    D(&s.lam, L);
    D(&s.q, 2 * L);
    D(&s.dfl, 9 * L);
    D(&s.mf, 2 * L);
    s.pfb = nullptr;  // several particles only (iblbm_set_particles)
    I(&s.pk, (size_t)FAST_G * FAST_T);
    D(&s.pw, (size_t)FAST_G * FAST_T);
    I(&s.mti, (size_t)FAST_K * FAST_LB);
    D(&s.mw, (size_t)FAST_K * FAST_LB);
    D(&d.stats, 6);
    D(&d.diag, 8);
    d.lb_cap = (int)L;
    d.box_cap = (int)B;
    return ok;
}

}  // namespace iblbm
