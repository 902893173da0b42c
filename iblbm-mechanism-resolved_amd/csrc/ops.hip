// ops.hip -- operator-level entry points with the argument meaning of the reference's
// gpu_kernels.py wrappers (host arrays in and out), and the host-side omega estimate.
//
// They share the per-node arithmetic (lbm_math.h) and the window machinery (ibm_block.h) with
// the fused step, so a parity test of an operator also pins that piece of the fused kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/iblbm_hip.h"
#include "ibm_block.h"
#include "scratch_alloc.h"
#include "mlp.h"

using namespace iblbm;

namespace {

thread_local std::string g_op_err;

// shares the error slot with the context API through iblbm_last_error's fallback
int op_fail(int code, const std::string &m);

struct Dev {
    std::vector<void *> p;
    ~Dev() {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename T>
    T *get(size_t n) {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(T) * (n ? n : 1)) != hipSuccess) return nullptr;
        p.push_back(q);
        return (T *)q;
    }
    template <typename T>
    T *up(const T *h, size_t n) {
        T *d = get<T>(n);
        if (d && h && n && hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess)
            return nullptr;
        return d;
    }
};

#define OPCHK(expr)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return op_fail(IBLBM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define NEED(ptr)                                                                  \
    do {                                                                           \
        if (!(ptr)) return op_fail(IBLBM_ENOMEM, "device allocation failed");     \
    } while (0)

unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

extern "C" int iblbm_set_op_error(int code, const char *m);

namespace {
int op_fail(int code, const std::string &m) { return iblbm_set_op_error(code, m.c_str()); }
}  // namespace

// ------------------------------------------------------------------ lattice operators

__global__ void k_op_feq(const double *ro, const double *U, int64_t N, int inc, double *out) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    feq9(ro[n], U[2 * n], U[2 * n + 1], inc != 0, out + 9 * n);
}

__global__ void k_op_macro(const double *f, int64_t N, int inc, double *ro, double *U) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double r, u0, u1;
    macro9(f + 9 * n, inc != 0, r, u0, u1);
    ro[n] = r;
    U[2 * n] = u0;
    U[2 * n + 1] = u1;
}

// push streaming onto a copy of the previous fstar (streaming.py:38-77), written as a pull
__global__ void k_op_stream(const double *old, const double *f, int nx, int ny, double *out) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (int64_t)nx * ny) return;
    const int i = (int)(n % nx), j = (int)(n / nx);
    for (int q = 0; q < 9; ++q) {
        const int si = i - EXi[q], sj = j - EYi[q];
        out[9 * n + q] = (si >= 0 && si < nx && sj >= 0 && sj < ny)
                             ? f[9 * ((int64_t)sj * nx + si) + q]
                             : old[9 * n + q];
    }
}

__global__ void k_op_collide(CollParams cp, const double *fs, const double *fe, const double *U,
                             const double *fib, int64_t N, double *out) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    collide<true>(cp, fs + 9 * n, fe + 9 * n, U[2 * n], U[2 * n + 1], fib[2 * n], fib[2 * n + 1],
                  out + 9 * n);
}

// boundary closures in place on a post-stream field (one thread per boundary node)
__global__ void k_op_boundary(int bc, double *fs, int nx, int ny, double in_u, double dens,
                              int inc) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = 2LL * nx + 2LL * (ny - 2);
    if (k >= nb) return;
    int i, j;
    if (k < nx) {
        i = (int)k;
        j = 0;
    } else if (k < 2LL * nx) {
        i = (int)(k - nx);
        j = ny - 1;
    } else {
        const int64_t r = k - 2LL * nx;
        i = (r < ny - 2) ? 0 : nx - 1;
        j = 1 + (int)(r % (ny - 2));
    }
    double *f = fs + 9 * ((int64_t)j * nx + i);
    if (bc == IBLBM_BC_DIRICHLET_NEUMANN) {
        if (i == nx - 1) {
            const int jj = j == 0 ? 1 : (j == ny - 1 ? ny - 2 : j);
            neumann_copy(fs + 9 * ((int64_t)jj * nx + nx - 2), f);
        } else {
            feq9(dens, in_u, 0.0, false, f);
        }
        return;
    }
    const bool corner = (i == 0 || i == nx - 1) && (j == 0 || j == ny - 1);
    if (corner) {
        feq9(dens, 0.0, 0.0, inc != 0, f);
        return;
    }
    settling_face(i == 0 ? 0 : (i == nx - 1 ? 1 : (j == ny - 1 ? 2 : 3)), f);
}

// ------------------------------------------------------------------ IBM operators

// mode 0: DF fields, 1: MDF, 2: DFC lambda only, 3: DFC correction (lambda recomputed),
// 4: DFC correction with the caller's lambda, 5: DFC correction field alone into ef (N, 9)
// with the caller's lambda, 6: the same with lambda recomputed, 7: spread of the marker forces
// the host placed in S.Lfx, S.Lfy.  One workgroup; the window machinery is the fused step's
// (ibm_build_tables / ibm_core).
template <int A>
__device__ void op_ibm_impl(Inst &I, int mode, const double *Eux, const double *Euy,
                            const double *Ro, double *fstar, double *red, int *redi,
                            IbmShared &sh, int us = 1,  // us: element stride of Eux, Euy
                            double *ef = nullptr) {
    IbmScratch &S = I.s;
    const int t = threadIdx.x;
    const int nbox = I.bw * I.bh;
    ibm_build_tables<A>(I, redi);
    const int nT = S.tindex[nbox];
    const IbmPtrs P = ibm_ptrs<false>(I, sh);  // global working arrays on this path
    MarkerRegs<A> M;
    load_marker_regs<A>(I, M);
    if (mode == 7) {  // spread as a gather in np.add.at order (df.py:269-283)
        for (int ti = t; ti < nT; ti += NTI) {
            double ex, ey;
            gather2(P, ti, ex, ey);
            P.fbx[ti] = ex;
            P.fby[ti] = ey;
        }
        __threadfence_block();
        __syncthreads();
        for (int sl = t; sl < nbox; sl += NTI) {
            const int ti = S.tindex[sl];
            S.fibx[sl] = ti >= 0 ? P.fbx[ti] : 0.0;
            S.fiby[sl] = ti >= 0 ? P.fby[ti] : 0.0;
        }
        return;
    }
    if (mode >= 2) {
        if (mode != 4 && mode != 5) ibm_dfc_lambda<A>(I, nT, I.Lb);
        if (mode == 2) return;
        for (int sl = t; sl < nbox; sl += NTI) {
            const int64_t n = (int64_t)(I.by0 + sl / I.bw) * I.nx + I.bx0 + sl % I.bw;
            for (int q = 0; q < 9; ++q) S.fs[q * I.box_cap + sl] = fstar[9 * n + q];
            S.ro[sl] = 1.0;
            S.ux0[sl] = 0.0;
            S.uy0[sl] = 0.0;
        }
        __threadfence_block();
        __syncthreads();
        ibm_core<A>(I, nT, P, red, false, M);
        if (mode >= 5) {  // the Eulerian correction alone: the gather of dfc.py:208-229 from 0
            for (int ti = t; ti < nT; ti += NTI) {
                const int sl = S.tslot[ti];
                const int64_t n = (int64_t)(I.by0 + sl / I.bw) * I.nx + I.bx0 + sl % I.bw;
                double e[9];
                for (int q = 0; q < 9; ++q) e[q] = 0.0;
                for (int x = P.toff[ti], end = P.toff[ti + 1]; x < end; ++x) {
                    const int k = P.gk[x];
                    const double wl = P.gw[x];
                    for (int q = 0; q < 9; ++q) e[q] += S.dfl[9 * k + q] * wl;
                }
                for (int q = 0; q < 9; ++q) ef[9 * n + q] = e[q];
            }
            return;
        }
        for (int ti = t; ti < nT; ti += NTI) {
            const int sl = S.tslot[ti];
            const int64_t n = (int64_t)(I.by0 + sl / I.bw) * I.nx + I.bx0 + sl % I.bw;
            for (int q = 0; q < 9; ++q) fstar[9 * n + q] = S.fs[q * I.box_cap + sl];
        }
        return;
    }
    for (int sl = t; sl < nbox; sl += NTI) {
        const int64_t n = (int64_t)(I.by0 + sl / I.bw) * I.nx + I.bx0 + sl % I.bw;
        S.ro[sl] = Ro[n];
        S.ux0[sl] = Eux[us * n];
        S.uy0[sl] = Euy[us * n];
    }
    __threadfence_block();
    __syncthreads();
    ibm_touched_init(I, nT, P);
    ibm_core<A>(I, nT, P, red, mode == 0, M);
    for (int sl = t; sl < nbox; sl += NTI) {
        const int ti = S.tindex[sl];
        S.fibx[sl] = ti >= 0 ? P.fbx[ti] : 0.0;
        S.fiby[sl] = ti >= 0 ? P.fby[ti] : 0.0;
    }
}

__global__ __launch_bounds__(NTI) void k_op_ibm(Inst *Ip, int mode, const double *Eux,
                                               const double *Euy, const double *Ro,
                                               double *fstar, double *ef) {
    __shared__ double red[NTI];
    __shared__ int redi[NTI];
    __shared__ IbmShared sh;
    Inst &I = *Ip;
    if (I.delta == 1)
        op_ibm_impl<1>(I, mode, Eux, Euy, Ro, fstar, red, redi, sh, 1, ef);
    else
        op_ibm_impl<2>(I, mode, Eux, Euy, Ro, fstar, red, redi, sh, 1, ef);
}

// velocity correction in place (iblbm/ibm/dispatch.py:30-33): u + fib dt / (2 rho), or
// u + 0.5 fib dt for the incompressible form
__global__ void k_op_velcorr(double *U, const double *fib, const double *ro, int64_t N,
                             double dt, int inc) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int c = 0; c < 2; ++c) {
        const double f = fib[2 * n + c];
        U[2 * n + c] = inc ? U[2 * n + c] + 0.5 * f * dt : U[2 * n + c] + f * dt / (2.0 * ro[n]);
    }
}

namespace {

// Scratch for one operator call on a (ny, nx) grid with Lb markers.
struct OpIbm {
    Dev mem;
    Inst h;
    Inst *d = nullptr;
    int setup(const double *Lx, const double *Ly, const double *ud, int Lb, double dx, double dy,
              double dt, double Larea, int ny, int nx, int delta, int inc) {
        if (Lb < 1 || nx < 1 || ny < 1) return op_fail(IBLBM_EINVAL, "ibm op: empty input");
        if (delta != 0 && delta != 1) return op_fail(IBLBM_EKEY, "ibm op: unknown delta type");
        memset(&h, 0, sizeof h);
        int box[4];
        int rc = iblbm_marker_box(Lx, Ly, Lb, dx, dy, nx, ny, delta, box);
        if (rc) return rc;
        h.nx = nx;
        h.ny = ny;
        h.inc = inc;
        h.cp.dt = dt;
        h.delta = delta;
        h.Lb = Lb;
        h.lb_cap = Lb;
        h.bx0 = box[0];
        h.by0 = box[1];
        h.bw = box[2];
        h.bh = box[3];
        const int B = box[2] * box[3];
        h.dx = dx;
        h.dy = dy;
        h.Larea = Larea;
        h.dfc_scale = 1.0;
        if (!alloc_ibm_scratch(h, Lb, B, [&](size_t bytes) -> void * { return mem.get<char>(bytes); }))
            return op_fail(IBLBM_ENOMEM, "ibm op: scratch allocation failed");
        if (hipMemcpy((void *)h.Lx, Lx, sizeof(double) * Lb, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy((void *)h.Ly, Ly, sizeof(double) * Lb, hipMemcpyHostToDevice) != hipSuccess)
            return op_fail(IBLBM_EDEVICE, "ibm op: marker upload failed");
        if (ud && hipMemcpy((void *)h.ud, ud, sizeof(double) * 2 * Lb, hipMemcpyHostToDevice) !=
                      hipSuccess)
            return op_fail(IBLBM_EDEVICE, "ibm op: desired velocity upload failed");
        h.Lb = Lb;
        return 0;
    }
    int push() {
        d = mem.get<Inst>(1);
        if (!d) return op_fail(IBLBM_ENOMEM, "ibm op: allocation failed");
        OPCHK(hipMemcpy(d, &h, sizeof h, hipMemcpyHostToDevice));
        return 0;
    }
    // scatter box fib into a zeroed (N, 2) host array
    int fib_out(double *fib) {
        const int64_t N = (int64_t)h.nx * h.ny;
        const int B = h.bw * h.bh;
        std::vector<double> fx(B), fy(B);
        OPCHK(hipMemcpy(fx.data(), h.s.fibx, sizeof(double) * B, hipMemcpyDeviceToHost));
        OPCHK(hipMemcpy(fy.data(), h.s.fiby, sizeof(double) * B, hipMemcpyDeviceToHost));
        memset(fib, 0, sizeof(double) * 2 * N);
        for (int sl = 0; sl < B; ++sl) {
            const int64_t n = (int64_t)(h.by0 + sl / h.bw) * h.nx + h.bx0 + sl % h.bw;
            fib[2 * n] = fx[sl];
            fib[2 * n + 1] = fy[sl];
        }
        return 0;
    }
};

}  // namespace

// ------------------------------------------------------------------ no-allocation MDF
//
// iblbm_op_multi_direct_forcing_noalloc runs the MDF window of iblbm_op_multi_direct_forcing on
// caller-owned device arrays, with every working array of the window (the Inst descriptor,
// the marker copies, the contribution and gather tables, the touched-node fields and the
// statistics) carved from the file-scope array `scratch` below instead of hipMalloc.
constexpr int NA_LB_MAX = 512;       // markers per call
constexpr int NA_BOX_MAX = 16384;    // window box nodes (bw * bh) per call, e.g. 128 x 128
constexpr size_t NA_BYTES = 6u << 20;  // 6 MiB per GPU, held until the process ends
constexpr int NA_OUT = 16;           // [stats 0..5, unused, unused, status, box 0..3, ...]

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

__global__ void k_na_zero(double *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// k_na_put stores the call's descriptor (array pointers into the scratch) at the scratch's
// first slot.  k_na_mdf, one workgroup: copies the markers into the scratch, forms the window box (every clamped
// stencil node, as iblbm_marker_box does on the host), then runs the operator's MDF window
// and scatters the window's fib into the caller's zeroed (N, 2) array.  out[8] receives 0, or
// 1 for a non-finite or far out-of-range marker coordinate, or 2 for a box above NA_BOX_MAX;
// on 1 or 2 nothing else is written.
__global__ void k_na_put(Inst v, Inst *Ip) {
    if (threadIdx.x == 0) *Ip = v;
}

template <int A>
__global__ __launch_bounds__(NTI) void k_na_mdf(Inst *Ip, const double *Lx, const double *Ly,
                                                const double *ud, const double *U,
                                                const double *ro, double *fib, double *out) {
    __shared__ double red[NTI];
    __shared__ int redi[NTI];
    __shared__ IbmShared sh;
    __shared__ int bx[4], bad;
    const int t = threadIdx.x;
    if (t == 0) {
        bx[0] = INT_MAX;
        bx[1] = INT_MIN;
        bx[2] = INT_MAX;
        bx[3] = INT_MIN;
        bad = 0;
    }
    __syncthreads();
    Inst &I = *Ip;
    double *lx = (double *)I.Lx, *ly = (double *)I.Ly, *u2 = (double *)I.ud;
    const int nx = I.nx, ny = I.ny;
    const double dx = I.dx, dy = I.dy;
    for (int k = t; k < I.Lb; k += NTI) {
        const double x = Lx[k], y = Ly[k];
        lx[k] = x;
        ly[k] = y;
        u2[2 * k] = ud[2 * k];
        u2[2 * k + 1] = ud[2 * k + 1];
        const double fx = floor(x / dx + 0.5), fy = floor(y / dy + 0.5);
        if (!(fabs(fx) < 1.0e9) || !(fabs(fy) < 1.0e9)) {  // NaN fails both
            atomicOr(&bad, 1);
            continue;
        }
        atomicMin(&bx[0], (int)fx - A);
        atomicMax(&bx[1], (int)fx + A);
        atomicMin(&bx[2], (int)fy - A);
        atomicMax(&bx[3], (int)fy + A);
    }
    __syncthreads();
    const int x0 = min(max(bx[0], 0), nx - 1), x1 = min(max(bx[1], 0), nx - 1);
    const int y0 = min(max(bx[2], 0), ny - 1), y1 = min(max(bx[3], 0), ny - 1);
    const int bw = x1 - x0 + 1, bh = y1 - y0 + 1;
    const int status = bad ? 1 : ((int64_t)bw * bh > NA_BOX_MAX ? 2 : 0);
    if (t == 0) {
        out[8] = status;
        out[9] = x0;
        out[10] = y0;
        out[11] = bw;
        out[12] = bh;
    }
    if (status) return;  // uniform over the workgroup
    if (t == 0) {
        I.bx0 = x0;
        I.by0 = y0;
        I.bw = bw;
        I.bh = bh;
    }
    __threadfence_block();
    __syncthreads();
    op_ibm_impl<A>(I, 1, U, U + 1, ro, nullptr, red, redi, sh, 2);
    __threadfence_block();
    __syncthreads();
    for (int sl = t; sl < bw * bh; sl += NTI) {
        const int64_t n = (int64_t)(y0 + sl / bw) * nx + x0 + sl % bw;
        fib[2 * n] = I.s.fibx[sl];
        fib[2 * n + 1] = I.s.fiby[sl];
    }
}

namespace {

// Per device: the device address of `scratch` (hipGetSymbolAddress, taken once with that
// device current) and the descriptor whose array pointers are carved from it.
struct NaDevice {
    bool ready = false;
    unsigned char *base = nullptr;
    Inst tmpl;
};
std::mutex g_na_mu;
constexpr int NA_MAX_DEVICES = 64;
NaDevice g_na[NA_MAX_DEVICES];

int na_setup(int dev, NaDevice **out) {
    if (dev < 0 || dev >= NA_MAX_DEVICES)
        return op_fail(IBLBM_EDEVICE, "noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_na_mu);
    NaDevice &d = g_na[dev];
    if (!d.ready) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return op_fail(IBLBM_EDEVICE, std::string("noalloc: hipGetSymbolAddress(scratch): ") +
                                              hipGetErrorString(e));
        d.base = (unsigned char *)ptr;
        // bump allocation from the device address: the Inst first, then the window arrays
        size_t used = 0;
        auto get = [&](size_t bytes) -> void * {
            const size_t at = (used + 255) & ~(size_t)255;
            if (at + bytes > NA_BYTES) return nullptr;
            used = at + bytes;
            return d.base + at;
        };
        memset(&d.tmpl, 0, sizeof d.tmpl);
        if (!get(sizeof(Inst)) || !get(sizeof(double) * NA_OUT) ||
            !alloc_ibm_scratch(d.tmpl, NA_LB_MAX, NA_BOX_MAX, get))
            return op_fail(IBLBM_EDEVICE, "noalloc: the window arrays do not fit the scratch");
        d.ready = true;
    }
    *out = &d;
    return 0;
}

}  // namespace

extern "C" {

int iblbm_op_feq(const double *ro, const double *U, int64_t N, int inc, double *feq) {
    if (!ro || !U || !feq || N < 0) return op_fail(IBLBM_EINVAL, "op_feq: bad argument");
    if (N == 0) return 0;
    Dev m;
    const double *dr = m.up(ro, N), *du = m.up(U, 2 * N);
    double *dout = m.get<double>(9 * N);
    NEED(dr && du && dout);
    hipLaunchKernelGGL(k_op_feq, dim3(nblk(N, 256)), dim3(256), 0, 0, dr, du, N, inc, dout);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(feq, dout, sizeof(double) * 9 * N, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_op_macroscopic(const double *f, int64_t N, int inc, double *ro, double *U) {
    if (!f || !ro || !U || N < 0) return op_fail(IBLBM_EINVAL, "op_macroscopic: bad argument");
    if (N == 0) return 0;
    Dev m;
    const double *df = m.up(f, 9 * N);
    double *dr = m.get<double>(N), *du = m.get<double>(2 * N);
    NEED(df && dr && du);
    hipLaunchKernelGGL(k_op_macro, dim3(nblk(N, 256)), dim3(256), 0, 0, df, N, inc, dr, du);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(ro, dr, sizeof(double) * N, hipMemcpyDeviceToHost));
    OPCHK(hipMemcpy(U, du, sizeof(double) * 2 * N, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_op_streaming(const double *old, const double *f, int nx, int ny, double *out) {
    if (!old || !f || !out || nx < 1 || ny < 1)
        return op_fail(IBLBM_EINVAL, "op_streaming: bad argument");
    const int64_t N = (int64_t)nx * ny;
    Dev m;
    const double *dold = m.up(old, 9 * N), *df = m.up(f, 9 * N);
    double *dout = m.get<double>(9 * N);
    NEED(dold && df && dout);
    hipLaunchKernelGGL(k_op_stream, dim3(nblk(N, 256)), dim3(256), 0, 0, dold, df, nx, ny, dout);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(out, dout, sizeof(double) * 9 * N, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_op_collision(int model, const double *fstar, const double *feq, const double *U,
                       const double *fib, int64_t N, double tau, double tau_minus,
                       const double *s_diag, double dt, double *out) {
    if (model < 0 || model > 2) return op_fail(IBLBM_EKEY, "op_collision: unknown model");
    if (!fstar || !U || !fib || !out || N < 0 || (model != 2 && !feq) || (model == 2 && !s_diag))
        return op_fail(IBLBM_EINVAL, "op_collision: bad argument");
    if (N == 0) return 0;
    CollParams cp;
    memset(&cp, 0, sizeof cp);
    cp.model = model;
    cp.dt = dt;
    cp.inv_tau = 1.0 / tau;
    cp.guo_pref = 1.0 - 0.5 / tau;
    cp.s_plus = 1.0 / tau;
    cp.s_minus = 1.0 / tau_minus;
    cp.gpp = 1.0 - 0.5 * cp.s_plus;
    cp.gpm = 1.0 - 0.5 * cp.s_minus;
    if (model == 2)
        for (int q = 0; q < 9; ++q) cp.S[q] = s_diag[q];
    Dev m;
    const double *dfs = m.up(fstar, 9 * N), *dfe = model != 2 ? m.up(feq, 9 * N) : dfs;
    const double *du = m.up(U, 2 * N), *dfib = m.up(fib, 2 * N);
    double *dout = m.get<double>(9 * N);
    NEED(dfs && dfe && du && dfib && dout);
    hipLaunchKernelGGL(k_op_collide, dim3(nblk(N, 256)), dim3(256), 0, 0, cp, dfs, dfe, du, dfib,
                       N, dout);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(out, dout, sizeof(double) * 9 * N, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_op_boundary(int bc, double *fstar, int nx, int ny, double in_u, double dens, int inc) {
    if (bc != 0 && bc != 1) return op_fail(IBLBM_EKEY, "op_boundary: unknown bc_type");
    if (!fstar || nx < 3 || ny < 3) return op_fail(IBLBM_EINVAL, "op_boundary: bad argument");
    const int64_t N = (int64_t)nx * ny, nb = 2LL * nx + 2LL * (ny - 2);
    Dev m;
    double *d = m.up(fstar, 9 * N);
    NEED(d);
    hipLaunchKernelGGL(k_op_boundary, dim3(nblk(nb, 256)), dim3(256), 0, 0, bc, d, nx, ny, in_u,
                       dens, inc);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(fstar, d, sizeof(double) * 9 * N, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_op_direct_forcing(const double *Lx, const double *Ly, const double *ud, int Lb,
                            const double *Eux, const double *Euy, const double *Ro, double dx,
                            double dy, double dt, double Larea, int ny, int nx, int delta,
                            int inc, double *fib, double *Lux, double *Luy) {
    if (!Lx || !Ly || !ud || !Eux || !Euy || !Ro || !fib)
        return op_fail(IBLBM_EINVAL, "op_direct_forcing: null argument");
    OpIbm op;
    int rc = op.setup(Lx, Ly, ud, Lb, dx, dy, dt, Larea, ny, nx, delta, inc);
    if (rc) return rc;
    op.h.ibm = 1;
    rc = op.push();
    if (rc) return rc;
    const int64_t N = (int64_t)nx * ny;
    const double *dux = op.mem.up(Eux, N), *duy = op.mem.up(Euy, N), *dro = op.mem.up(Ro, N);
    NEED(dux && duy && dro);
    hipLaunchKernelGGL(k_op_ibm, dim3(1), dim3(NTI), 0, 0, op.d, 0, dux, duy, dro, (double *)nullptr,
                       (double *)nullptr);
    OPCHK(hipGetLastError());
    OPCHK(hipDeviceSynchronize());
    rc = op.fib_out(fib);
    if (rc) return rc;
    if (Lux) OPCHK(hipMemcpy(Lux, op.h.s.Lux, sizeof(double) * Lb, hipMemcpyDeviceToHost));
    if (Luy) OPCHK(hipMemcpy(Luy, op.h.s.Luy, sizeof(double) * Lb, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_op_multi_direct_forcing(const double *Lx, const double *Ly, const double *ud, int Lb,
                                  const double *U, const double *ro, double dx, double dy,
                                  double dt, double Larea, int ny, int nx, int n_iter,
                                  int min_iter, int delta, double omega, double tol,
                                  double *fib_total, iblbm_mdf_stats *stats) {
    if (!Lx || !Ly || !ud || !U || !ro || !fib_total)
        return op_fail(IBLBM_EINVAL, "op_multi_direct_forcing: null argument");
    if (n_iter < 1) return op_fail(IBLBM_EINVAL, "op_multi_direct_forcing: n_iter must be >= 1");
    OpIbm op;
    int rc = op.setup(Lx, Ly, ud, Lb, dx, dy, dt, Larea, ny, nx, delta, 0);
    if (rc) return rc;
    op.h.ibm = 2;
    op.h.n_iter = n_iter;
    op.h.min_iter = min_iter;
    op.h.omega = omega;
    op.h.tol = tol;
    rc = op.push();
    if (rc) return rc;
    const int64_t N = (int64_t)nx * ny;
    std::vector<double> ux(N), uy(N);
    for (int64_t n = 0; n < N; ++n) {
        ux[n] = U[2 * n];
        uy[n] = U[2 * n + 1];
    }
    const double *dux = op.mem.up(ux.data(), N), *duy = op.mem.up(uy.data(), N);
    const double *dro = op.mem.up(ro, N);
    NEED(dux && duy && dro);
    hipLaunchKernelGGL(k_op_ibm, dim3(1), dim3(NTI), 0, 0, op.d, 1, dux, duy, dro, (double *)nullptr,
                       (double *)nullptr);
    OPCHK(hipGetLastError());
    OPCHK(hipDeviceSynchronize());
    rc = op.fib_out(fib_total);
    if (rc) return rc;
    if (stats) {
        double s[4];
        OPCHK(hipMemcpy(s, op.h.stats, sizeof s, hipMemcpyDeviceToHost));
        stats->iter_count = (int32_t)s[0];
        stats->last_residual = s[1];
        stats->diverged = (int32_t)s[2];
        stats->converged = (int32_t)s[3];
        stats->guard_trips = 0;
        stats->omega_used = omega;
    }
    return 0;
}

int iblbm_op_multi_direct_forcing_noalloc(const double *Lx, const double *Ly, const double *ud,
                                          int Lb, const double *U, const double *ro, double dx,
                                          double dy, double dt, double Larea, int ny, int nx,
                                          int n_iter, int min_iter, int delta, double omega,
                                          double tol, double *fib_total, iblbm_mdf_stats *stats,
                                          void *stream) {
    if (!Lx || !Ly || !ud || !U || !ro || !fib_total)
        return op_fail(IBLBM_EINVAL, "op_multi_direct_forcing_noalloc: null argument");
    if (Lb < 1 || Lb > NA_LB_MAX)
        return op_fail(IBLBM_EINVAL, "op_multi_direct_forcing_noalloc: Lb = " + std::to_string(Lb) +
                                         " outside [1, " + std::to_string(NA_LB_MAX) + "]");
    if (nx < 1 || ny < 1 || !(dx > 0.0) || !(dy > 0.0))
        return op_fail(IBLBM_EINVAL, "op_multi_direct_forcing_noalloc: bad grid");
    if (n_iter < 1)
        return op_fail(IBLBM_EINVAL, "op_multi_direct_forcing_noalloc: n_iter must be >= 1");
    if (delta != 0 && delta != 1)
        return op_fail(IBLBM_EKEY, "op_multi_direct_forcing_noalloc: unknown delta type");
    int dev = 0;
    OPCHK(hipGetDevice(&dev));
    NaDevice *D = nullptr;
    int rc = na_setup(dev, &D);
    if (rc) return rc;
    Inst v = D->tmpl;  // array pointers into this device's scratch
    v.nx = nx;
    v.ny = ny;
    v.inc = 0;  // the MDF sub-step is compressible (mdf.py:177-181)
    v.cp.dt = dt;
    v.delta = delta;
    v.ibm = 2;
    v.Lb = Lb;
    v.dx = dx;
    v.dy = dy;
    v.Larea = Larea;
    v.dfc_scale = 1.0;
    v.n_iter = n_iter;
    v.min_iter = min_iter;
    v.omega = omega;
    v.tol = tol;
    Inst *Ip = (Inst *)D->base;  // the first scratch slot
    double *out = (double *)(D->base + ((sizeof(Inst) + 255) & ~(size_t)255));
    v.stats = out;
    hipStream_t st = (hipStream_t)stream;
    const int64_t n2 = 2LL * nx * ny;
    hipLaunchKernelGGL(k_na_zero, dim3((unsigned)std::min<int64_t>((n2 + 255) / 256, 8192)),
                       dim3(256), 0, st, fib_total, n2);
    OPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_na_put, dim3(1), dim3(64), 0, st, v, Ip);
    OPCHK(hipGetLastError());
    if (delta == 1)
        hipLaunchKernelGGL(k_na_mdf<1>, dim3(1), dim3(NTI), 0, st, Ip, Lx, Ly, ud, U, ro,
                           fib_total, out);
    else
        hipLaunchKernelGGL(k_na_mdf<2>, dim3(1), dim3(NTI), 0, st, Ip, Lx, Ly, ud, U, ro,
                           fib_total, out);
    OPCHK(hipGetLastError());
    double h[NA_OUT];
    OPCHK(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, st));
    OPCHK(hipStreamSynchronize(st));
    if (h[8] == 1.0)
        return op_fail(IBLBM_EINVAL,
                       "op_multi_direct_forcing_noalloc: non-finite or out-of-range marker");
    if (h[8] == 2.0)
        return op_fail(IBLBM_EINVAL, "op_multi_direct_forcing_noalloc: window box " +
                                         std::to_string((int)h[11]) + " x " +
                                         std::to_string((int)h[12]) + " exceeds " +
                                         std::to_string(NA_BOX_MAX) + " nodes");
    if (stats) {
        stats->iter_count = (int32_t)h[0];
        stats->last_residual = h[1];
        stats->diverged = (int32_t)h[2];
        stats->converged = (int32_t)h[3];
        stats->guard_trips = 0;
        stats->omega_used = omega;
    }
    return 0;
}

int iblbm_op_dfc_lambda(const double *Lx, const double *Ly, int Lb, double rho_f, double Larea,
                        double dx, double dy, int ny, int nx, int delta, double *lambda_k) {
    if (!Lx || !Ly || !lambda_k) return op_fail(IBLBM_EINVAL, "op_dfc_lambda: null argument");
    if (rho_f != 1.0) return op_fail(IBLBM_EINVAL, "op_dfc_lambda: only rho_f = 1 (dfc.py:297)");
    OpIbm op;
    int rc = op.setup(Lx, Ly, nullptr, Lb, dx, dy, 1.0, Larea, ny, nx, delta, 0);
    if (rc) return rc;
    op.h.ibm = 3;
    rc = op.push();
    if (rc) return rc;
    hipLaunchKernelGGL(k_op_ibm, dim3(1), dim3(NTI), 0, 0, op.d, 2, (const double *)nullptr,
                       (const double *)nullptr, (const double *)nullptr, (double *)nullptr,
                       (double *)nullptr);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(lambda_k, op.h.s.lam, sizeof(double) * Lb, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_op_dfc_correction(const double *Lx, const double *Ly, const double *ud, int Lb,
                            double *fstar, const double *lambda_k, double dx, double dy,
                            double Larea, int ny, int nx, int delta, double scale,
                            double *force) {
    if (!Lx || !Ly || !ud || !fstar || !force)
        return op_fail(IBLBM_EINVAL, "op_dfc_correction: null argument");
    OpIbm op;
    int rc = op.setup(Lx, Ly, ud, Lb, dx, dy, 1.0, Larea, ny, nx, delta, 0);
    if (rc) return rc;
    op.h.ibm = 3;
    op.h.dfc_scale = scale;
    rc = op.push();
    if (rc) return rc;
    const int64_t N = (int64_t)nx * ny;
    double *dfs = op.mem.up(fstar, 9 * N);
    NEED(dfs);
    // lambda: the caller's (the reference's lambda_cache) or recomputed (compute_lambda)
    if (lambda_k)
        OPCHK(hipMemcpy(op.h.s.lam, lambda_k, sizeof(double) * Lb, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_op_ibm, dim3(1), dim3(NTI), 0, 0, op.d, lambda_k ? 4 : 3,
                       (const double *)nullptr, (const double *)nullptr, (const double *)nullptr,
                       dfs, (double *)nullptr);
    OPCHK(hipGetLastError());
    OPCHK(hipDeviceSynchronize());
    OPCHK(hipMemcpy(fstar, dfs, sizeof(double) * 9 * N, hipMemcpyDeviceToHost));
    OPCHK(hipMemcpy(force, op.h.s.q, sizeof(double) * 2 * Lb, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_op_dfc_correction_field(const double *Lx, const double *Ly, const double *ud, int Lb,
                                  const double *fstar, const double *lambda_k, double dx,
                                  double dy, double Larea, int ny, int nx, int delta, double *ef,
                                  double *force) {
    if (!Lx || !Ly || !ud || !fstar || !ef || !force)
        return op_fail(IBLBM_EINVAL, "op_dfc_correction_field: null argument");
    OpIbm op;
    int rc = op.setup(Lx, Ly, ud, Lb, dx, dy, 1.0, Larea, ny, nx, delta, 0);
    if (rc) return rc;
    op.h.ibm = 3;
    op.h.dfc_scale = 1.0;
    rc = op.push();
    if (rc) return rc;
    const int64_t N = (int64_t)nx * ny;
    double *dfs = op.mem.up(fstar, 9 * N);
    double *def = op.mem.get<double>(9 * N);
    NEED(dfs && def);
    OPCHK(hipMemset(def, 0, sizeof(double) * 9 * N));
    if (lambda_k)
        OPCHK(hipMemcpy(op.h.s.lam, lambda_k, sizeof(double) * Lb, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_op_ibm, dim3(1), dim3(NTI), 0, 0, op.d, lambda_k ? 5 : 6,
                       (const double *)nullptr, (const double *)nullptr, (const double *)nullptr,
                       dfs, def);
    OPCHK(hipGetLastError());
    OPCHK(hipDeviceSynchronize());
    OPCHK(hipMemcpy(ef, def, sizeof(double) * 9 * N, hipMemcpyDeviceToHost));
    OPCHK(hipMemcpy(force, op.h.s.q, sizeof(double) * 2 * Lb, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_op_velocity_correction(double *U, const double *fib, const double *ro, int64_t N,
                                 double dt, int inc) {
    if (!U || !fib || !ro || N < 0)
        return op_fail(IBLBM_EINVAL, "op_velocity_correction: bad argument");
    if (N == 0) return 0;
    Dev m;
    double *du = m.up(U, 2 * N);
    const double *dfib = m.up(fib, 2 * N), *dro = m.up(ro, N);
    NEED(du && dfib && dro);
    hipLaunchKernelGGL(k_op_velcorr, dim3(nblk(N, 256)), dim3(256), 0, 0, du, dfib, dro, N, dt,
                       inc);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(U, du, sizeof(double) * 2 * N, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_op_interp_force(const double *Lx, const double *Ly, const double *ud, int Lb,
                          const double *Eux, const double *Euy, const double *Ro, double dx,
                          double dy, double dt, int ny, int nx, int delta, int inc, double *Lfx,
                          double *Lfy, double *Lux, double *Luy) {
    if (!Lx || !Ly || !ud || !Eux || !Euy || !Ro || !Lfx || !Lfy || !Lux || !Luy)
        return op_fail(IBLBM_EINVAL, "op_interp_force: null argument");
    OpIbm op;
    int rc = op.setup(Lx, Ly, ud, Lb, dx, dy, dt, 1.0, ny, nx, delta, inc);
    if (rc) return rc;
    op.h.ibm = 1;
    rc = op.push();
    if (rc) return rc;
    const int64_t N = (int64_t)nx * ny;
    const double *dux = op.mem.up(Eux, N), *duy = op.mem.up(Euy, N), *dro = op.mem.up(Ro, N);
    NEED(dux && duy && dro);
    hipLaunchKernelGGL(k_op_ibm, dim3(1), dim3(NTI), 0, 0, op.d, 0, dux, duy, dro,
                       (double *)nullptr, (double *)nullptr);
    OPCHK(hipGetLastError());
    OPCHK(hipDeviceSynchronize());
    OPCHK(hipMemcpy(Lfx, op.h.s.Lfx, sizeof(double) * Lb, hipMemcpyDeviceToHost));
    OPCHK(hipMemcpy(Lfy, op.h.s.Lfy, sizeof(double) * Lb, hipMemcpyDeviceToHost));
    OPCHK(hipMemcpy(Lux, op.h.s.Lux, sizeof(double) * Lb, hipMemcpyDeviceToHost));
    OPCHK(hipMemcpy(Luy, op.h.s.Luy, sizeof(double) * Lb, hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_op_spread(const double *Lfx, const double *Lfy, const double *Lx, const double *Ly,
                    int Lb, double dx, double dy, double Larea, int ny, int nx, int delta,
                    double *Efx, double *Efy) {
    if (!Lfx || !Lfy || !Lx || !Ly || !Efx || !Efy)
        return op_fail(IBLBM_EINVAL, "op_spread: null argument");
    OpIbm op;
    int rc = op.setup(Lx, Ly, nullptr, Lb, dx, dy, 1.0, Larea, ny, nx, delta, 0);
    if (rc) return rc;
    op.h.ibm = 1;
    rc = op.push();
    if (rc) return rc;
    OPCHK(hipMemcpy(op.h.s.Lfx, Lfx, sizeof(double) * Lb, hipMemcpyHostToDevice));
    OPCHK(hipMemcpy(op.h.s.Lfy, Lfy, sizeof(double) * Lb, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_op_ibm, dim3(1), dim3(NTI), 0, 0, op.d, 7, (const double *)nullptr,
                       (const double *)nullptr, (const double *)nullptr, (double *)nullptr,
                       (double *)nullptr);
    OPCHK(hipGetLastError());
    OPCHK(hipDeviceSynchronize());
    const int B = op.h.bw * op.h.bh;
    std::vector<double> fx(B), fy(B);
    OPCHK(hipMemcpy(fx.data(), op.h.s.fibx, sizeof(double) * B, hipMemcpyDeviceToHost));
    OPCHK(hipMemcpy(fy.data(), op.h.s.fiby, sizeof(double) * B, hipMemcpyDeviceToHost));
    for (int sl = 0; sl < B; ++sl) {  // accumulate into the caller's fields
        const int64_t n = (int64_t)(op.h.by0 + sl / op.h.bw) * nx + op.h.bx0 + sl % op.h.bw;
        Efx[n] = Efx[n] + fx[sl];
        Efy[n] = Efy[n] + fy[sl];
    }
    return 0;
}

// iblbm/ibm/a_norm.py:75-172 in lattice units, as called by :204-266
double iblbm_omega_estimate(const double *Lx, const double *Ly, int Lb, int delta, double Larea,
                            double dx) {
    if (!Lx || !Ly || Lb < 1 || !(dx > 0.0) || (delta != 0 && delta != 1)) return NAN;
    std::vector<double> x(Lb), y(Lb);
    for (int k = 0; k < Lb; ++k) {
        x[k] = Lx[k] / dx;
        y[k] = Ly[k] / dx;
    }
    const int a = delta == 1 ? 1 : 2;
    const double cutoff = 2.0 * a * 1.0, factor = Larea / (1.0 * 1.0);
    double best = -INFINITY;
    for (int l = 0; l < Lb; ++l) {
        double acc = 0.0;
        for (int m = 0; m < Lb; ++m) {
            if (std::fabs(x[l] - x[m]) >= cutoff || std::fabs(y[l] - y[m]) >= cutoff) continue;
            double sums[2];
            bool skip = false;
            for (int ax = 0; ax < 2 && !skip; ++ax) {
                const double pl = ax ? y[l] : x[l], pm = ax ? y[m] : x[m];
                const double lo = std::max(pl, pm) - a * 1.0, hi = std::min(pl, pm) + a * 1.0;
                const int64_t ilo = (int64_t)std::ceil(lo / 1.0);
                const int64_t ihi = (int64_t)std::floor(hi / 1.0) + 1;
                if (ilo >= ihi) {
                    skip = true;
                    break;
                }
                double s = 0.0;
                for (int64_t i = ilo; i < ihi; ++i) {
                    const double xi = (double)i * 1.0;
                    s += delta_phi(delta, (xi - pl) / 1.0) * delta_phi(delta, (xi - pm) / 1.0);
                }
                if (s == 0.0) {
                    skip = true;
                    break;
                }
                sums[ax] = s;
            }
            if (skip) continue;
            acc += factor * sums[0] * sums[1];
        }
        best = std::max(best, acc);
    }
    return 1.0 / best;
}

int iblbm_mlp_loss_grad(const float *params, const float *x, const float *y, int nb,
                        float *grad, float *loss) {
    if (!params || !x || !y || !grad || !loss || nb < 1 || nb > MLP_B)
        return op_fail(IBLBM_EINVAL, "mlp_loss_grad: bad argument (1 <= nb <= 128)");
    Dev m;
    const float *dp = m.up(params, MLP_NP), *dx = m.up(x, (size_t)nb * MLP_IN),
                *dy = m.up(y, (size_t)nb * MLP_OUT);
    float *dg = m.get<float>(MLP_NP), *dl = m.get<float>(1);
    NEED(dp && dx && dy && dg && dl);
    hipLaunchKernelGGL(k_mlp_grad, dim3(1), dim3(MLP_TT), 0, 0, dp, dx, dy, nb, dg, dl);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(grad, dg, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost));
    OPCHK(hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int iblbm_mlp_infer_bf16(const float *params, const float *x, int n, float *out) {
    if (!params || !x || !out || n < 1) return op_fail(IBLBM_EINVAL, "mlp_infer_bf16: bad argument");
    Dev m;
    const float *dp = m.up(params, MLP_NP), *dx = m.up(x, (size_t)n * MLP_IN);
    unsigned short *wb = m.get<unsigned short>(MLP_NBLK * 256);
    float *dout = m.get<float>((size_t)n * MLP_OUT);
    NEED(dp && dx && wb && dout);
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, 0, dp, wb);
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0, wb, dp, dx,
                       n, dout);
    OPCHK(hipGetLastError());
    OPCHK(hipMemcpy(out, dout, sizeof(float) * n * MLP_OUT, hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
