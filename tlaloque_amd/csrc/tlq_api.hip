// tlq_api.hip -- host side of the C-ABI declared in include/tlq.h.
//
// One context owns the device copies of the reference's global arrays
// (source/globals.f90:5-11) in structure-of-arrays layout, the problem
// description, and one HIP stream on which every call is ordered.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tlq.h"
#include "tlq_fused.hpp"
#include "tlq_gravity.hpp"
#include "tlq_launch.hpp"
#include "tlq_surrogate.hpp"

using namespace tlq;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(TLQ_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_));    \
    } while (0)

struct PendingTiming {
    int kernel;
    hipEvent_t a, b;
};

// "multigrid" times a whole self-gravity solve (many launches); "mg_relax0"
// the one-pass relaxation launches on the finest level.
const char *const kKernelNames[] = {"stage1", "stage2", "init", "cooling", "multigrid", "mg_relax0"};
constexpr int kNumKernels = 6;

}  // namespace

struct tlq_ctx {
    tlq_params prm{};
    Problem host{};            // host copy of the problem description
    Problem staged{};          // a setter's new problem while it is sent to the device
    Problem *dprob = nullptr;  // device copy
    Grid g{};
    hipStream_t stream = nullptr;
    int prec = TLQ_FP64;
    size_t esize = 8;  // bytes per element
    // Fused state: U ping-pong and U^{1/2}; per-routine arrays on demand.
    void *u[2] = {nullptr, nullptr};
    void *uh = nullptr;
    void *bufs[8] = {};  // TLQ_BUF_*; U aliases u[cur]
    int cur = 0;          // which u[] holds U^n
    StepState *st = nullptr;
    unsigned long long *scratch_max = nullptr;  // per-routine time_step maxima
    unsigned long long *scratch_cnt = nullptr;
    bool fused_ready = false;
    Tiling tiling1{}, tiling2{};  // tiles of the stage-1 and stage-2 kernels
    // timing
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ktiming = false;
    std::vector<PendingTiming> pending;
    double ktotal[kNumKernels] = {};
    long long kcount[kNumKernels] = {};
    double fallback = 0.0;
    // learned cooling surrogate
    float *sur_theta = nullptr, *sur_m = nullptr, *sur_v = nullptr, *sur_gout = nullptr;
    unsigned short *sur_thq = nullptr;
    float *sur_X = nullptr, *sur_Y = nullptr, *sur_partial = nullptr, *sur_loss = nullptr;
    long long sur_n = 0;
    int sur_t = 0;
    bool sur_ready = false;
    unsigned *cool_queue = nullptr;  // cells queued for the dense cooling pass
    unsigned *cool_fqueue = nullptr; // surrogate cells queued for the fallback
    unsigned *cool_seg = nullptr;    // per stage-2 tile: cells queued
    unsigned *cool_off = nullptr;    // their prefix sums (+ the total)
    unsigned *cool_fall = nullptr;   // cells of the surrogate's fallback queue
    float *gacc = nullptr;           // FP32 point-mass acceleration field (Problem::gacc)
    float *noalloc_scratch = nullptr;  // this device's copy of `scratch` (tlq_surrogate_grad_noalloc)
    // self-gravity: multigrid levels (level 0's E is PHI) and the step count
    // MultiGrid's mod(it, 10) test reads
    SgState sg;
    bool sg_on = false;
    int it_host = 0;
};

// Scratch of tlq_surrogate_grad_noalloc: the partial gradients (and loss) of
// up to TLQ_NOALLOC_MAX_SAMPLES / 64 blocks of the gradient kernel, in static
// device memory (4.6 MB on every GPU the library runs on, held until the
// process ends).  Host code never passes the array's name to a kernel: it is
// a host-side stand-in, not a device address.  Its address on each device is
// looked up once with hipGetSymbolAddress, with that device current, when the
// first context is created there, and kept in g_scratch_addr.
constexpr int kNoallocBlocks = TLQ_NOALLOC_MAX_SAMPLES / 64;
static __device__ __align__(256) float scratch[kNoallocBlocks * (mlp::NP + 1)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_addr[kMaxDevices] = {};

// The address of `scratch` on the current device (looked up on first use).
hipError_t scratch_address(float **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[dev] = static_cast<float *>(ptr);
    }
    *out = g_scratch_addr[dev];
    return hipSuccess;
}

bool valid_bc(int b) { return b == TLQ_BC_OUTFLOW || b == TLQ_BC_REFLECTIVE || b == TLQ_BC_PERIODIC; }

// The cooling queue (modes 2 and 3): one segment per stage-2 tile, of the
// tile's cell count.
long long cool_segment(const tlq_ctx *c);
long long cool_segments(const tlq_ctx *c);

long long cool_segment(const tlq_ctx *c) {
    if (c->prec == TLQ_FP32) {
        using S2 = TileShape<float>::S2;
        return (long long)S2::TX * S2::TY * S2::ZC;
    }
    using S2 = TileShape<double>::S2;
    return (long long)S2::TX * S2::TY * S2::ZC;
}
long long cool_segments(const tlq_ctx *c) { return c->tiling2.ntiles; }

long long ncell_box(const Grid &g) { return (long long)(g.nx + 4) * (g.ny + 4) * (g.nz + 4); }
long long ninterior(const Grid &g) { return (long long)g.nx * g.ny * g.nz; }

void *&buf_ref(tlq_ctx *c, int b) {
    if (b == TLQ_BUF_U) return c->u[c->cur];
    return c->bufs[b];
}

int ensure_buf(tlq_ctx *c, int b) {
    void *&p = buf_ref(c, b);
    if (p) return TLQ_OK;
    if (b == TLQ_BUF_UPP && c->uh) {  // U^{1/2} of the fused step is the reference's UPP
        p = c->uh;
        return TLQ_OK;
    }
    const size_t bytes = (size_t)5 * c->g.nc * c->esize;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        p = nullptr;
        return fail(TLQ_ERR_NOMEM, "hipMalloc failed for a state array");
    }
    if (hipMemsetAsync(p, 0, bytes, c->stream) != hipSuccess)
        return fail(TLQ_ERR_HIP, "hipMemsetAsync failed");
    return TLQ_OK;
}

int push_problem(tlq_ctx *c) {
    HIPCHK(hipMemcpyAsync(c->dprob, &c->host, sizeof(Problem), hipMemcpyHostToDevice, c->stream));
    return TLQ_OK;
}

// Makes `np` the context's problem: copies it to the device and, once that
// is enqueued, to the host copy.  On failure both copies stay as they were.
int commit_problem(tlq_ctx *c, const Problem &np) {
    c->staged = np;
    const hipError_t e = hipMemcpyAsync(c->dprob, &c->staged, sizeof(Problem),
                                        hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) {
        (void)push_problem(c);  // best effort: the device copy back to the host's
        return fail(TLQ_ERR_HIP, std::string("problem upload: ") + hipGetErrorString(e));
    }
    c->host = c->staged;
    return TLQ_OK;
}

bool all_finite(const double *a, int n) {
    for (int q = 0; q < n; ++q)
        if (!std::isfinite(a[q])) return false;
    return true;
}

template <class R> const R *cp(void *p) { return static_cast<const R *>(p); }
template <class R> R *mp(void *p) { return static_cast<R *>(p); }

// Time one launch when kernel timing is on.
struct KTimer {
    tlq_ctx *c;
    int k;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(tlq_ctx *ctx, int kernel) : c(ctx), k(kernel) {
        if (!c->ktiming) return;
        hipEventCreate(&a);
        hipEventCreate(&b);
        hipEventRecord(a, c->stream);
    }
    ~KTimer() {
        if (!c->ktiming) return;
        hipEventRecord(b, c->stream);
        c->pending.push_back({k, a, b});
    }
};

void resolve_timings(tlq_ctx *c) {
    if (c->pending.empty()) return;
    hipStreamSynchronize(c->stream);
    for (auto &p : c->pending) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, p.a, p.b);
        c->ktotal[p.kernel] += ms;
        c->kcount[p.kernel] += 1;
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    c->pending.clear();
}

// ---- precision-dispatched operations --------------------------------------

template <class R> int do_borders(tlq_ctx *c, void *buf, int order) {
    const Grid &g = c->g;
    const int n[3] = {g.nx, g.ny, g.nz};
    const long long plane[3] = {(long long)(g.ny + 4) * (g.nz + 4), (long long)(g.nx + 4) * (g.nz + 4),
                                (long long)(g.nx + 4) * (g.ny + 4)};
    auto copy = [&](int axis, int dst, int src, int neg) -> int {
        HIPCHK(Launch<R>::bc_plane(c->stream, c->dprob, mp<R>(buf), axis, dst, src, neg,
                                   plane[axis]));
        return TLQ_OK;
    };
    int rc;
    if (order == 1) {
        // boundariesI, boundaries.f90:179-259: x-, x+, y+, y-, z+, z-
        const int seq[6][2] = {{0, 0}, {0, 1}, {1, 1}, {1, 0}, {2, 1}, {2, 0}};
        for (auto &s : seq) {
            const int axis = s[0], side = s[1], bc = c->prm.bc[2 * axis + side];
            const int gl = side ? n[axis] + 1 : 0, in = side ? n[axis] : 1, wrap = side ? 1 : n[axis];
            if (bc == TLQ_BC_OUTFLOW) rc = copy(axis, gl, in, 0);
            else if (bc == TLQ_BC_REFLECTIVE) rc = copy(axis, gl, in, 1);
            else rc = copy(axis, gl, wrap, 0);
            if (rc) return rc;
        }
    } else {
        // boundariesII, boundaries.f90:432-533: x-, x+, y-, y+, z-, z+
        for (int axis = 0; axis < 3; ++axis)
            for (int side = 0; side < 2; ++side) {
                const int bc = c->prm.bc[2 * axis + side], N = n[axis];
                int d0, s0, d1, s1, neg = 0;
                if (side == 0) {
                    if (bc == TLQ_BC_PERIODIC) { d0 = -1; s0 = N - 1; d1 = 0; s1 = N; }
                    else if (bc == TLQ_BC_OUTFLOW) { d0 = 0; s0 = 1; d1 = -1; s1 = 1; }
                    else { d0 = 0; s0 = 1; d1 = -1; s1 = 2; neg = 1; }
                } else {
                    if (bc == TLQ_BC_PERIODIC) { d0 = N + 1; s0 = 1; d1 = N + 2; s1 = 2; }
                    else if (bc == TLQ_BC_OUTFLOW) { d0 = N + 1; s0 = N; d1 = N + 2; s1 = N; }
                    else { d0 = N + 1; s0 = N; d1 = N + 2; s1 = N - 1; neg = 1; }
                }
                if ((rc = copy(axis, d0, s0, neg))) return rc;
                if ((rc = copy(axis, d1, s1, neg))) return rc;
            }
    }
    if (c->host.nwinds > 0) {
        HIPCHK(Launch<R>::winds(c->stream, c->dprob, mp<R>(buf), -1, -1, -1, g.nx + 4, g.ny + 4,
                                g.nz + 4));
    }
    return TLQ_OK;
}

// MultiGrid(dens, PHI) on U^n's density at the current step count
// (mainHD.f90:50-52; HLLC.f90:16-21).
template <class R> int do_sg_solve(tlq_ctx *c) {
    if (!c->sg_on) return TLQ_OK;
    KTimer kt(c, 4);
    c->sg.time_relax0 = c->ktiming;
    HIPCHK(sg_solve<R>(c->stream, c->g, cp<R>(c->u[c->cur]), c->sg, c->it_host, c->dprob));
    if (c->ktiming) {
        c->ktotal[5] += c->sg.relax0_ms;
        c->kcount[5] += c->sg.relax0_n;
    }
    c->sg.relax0_ms = 0.0;
    c->sg.relax0_n = 0;
    return TLQ_OK;
}

template <class R> int do_init_maxima(tlq_ctx *c) {
    KTimer kt(c, 2);
    const int par = c->cur;  // parity of the step about to run
    HIPCHK(hipMemsetAsync(&c->st->maxima[par][0], 0, 4 * sizeof(unsigned long long), c->stream));
    HIPCHK(Launch<R>::maxima(c->stream, c->dprob, cp<R>(c->u[c->cur]), 0, &c->st->maxima[par][0],
                             ninterior(c->g)));
    c->fused_ready = true;
    return TLQ_OK;
}

template <class R> int do_advance(tlq_ctx *c, int nsteps) {
    if (!c->fused_ready) {
        int rc = do_init_maxima<R>(c);
        if (rc) return rc;
    }
    for (int s = 0; s < nsteps; ++s) {
        const int par = c->cur;
        // self-gravity: one solve per step from U^n, used by both stages
        // (the reference solves twice from the same U^n, mainHD.f90:80,88)
        if (int rc = do_sg_solve<R>(c)) return rc;
        {
            KTimer kt(c, 0);
            HIPCHK(Launch<R>::stage1(c->stream, c->dprob, c->tiling1, cp<R>(c->u[par]), mp<R>(c->uh),
                                     c->st, par, c->host.solver));
        }
        {
            KTimer kt(c, 1);
            HIPCHK(Launch<R>::stage2(c->stream, c->dprob, c->tiling2, cp<R>(c->u[par]),
                                     cp<R>(c->uh), mp<R>(c->u[par ^ 1]), c->st, par,
                                     c->host.limiter, c->host.solver, c->cool_queue,
                                     c->cool_seg));
        }
        if (c->host.cooling >= TLQ_COOL_SURROGATE) {
            KTimer kt(c, 3);
            HIPCHK(sur_cool_pass<R>(c->stream, c->dprob, (int)cool_segments(c), cool_segment(c),
                                    mp<R>(c->u[par ^ 1]), c->sur_thq, c->st, par, c->cool_queue,
                                    c->cool_seg, c->cool_off, c->cool_fqueue, c->cool_fall,
                                    c->host.cooling, c->host.ntable));
        }
        c->cur ^= 1;
        ++c->it_host;
    }
    return TLQ_OK;
}

template <class R> int do_upload(tlq_ctx *c, void *dst, const double *host) {
    const Grid &g = c->g;
    const long long plane = (long long)(g.nx + 4) * (g.ny + 4);
    const int nzt = g.nz + 4;
    int chunk = (int)std::max<long long>(1, (64ll << 20) / (plane * 5 * 8));
    if (chunk > nzt) chunk = nzt;
    double *stage = nullptr;
    HIPCHK(hipMalloc(&stage, (size_t)plane * chunk * 5 * 8));
    for (int k0 = 0; k0 < nzt; k0 += chunk) {
        const int nk = std::min(chunk, nzt - k0);
        HIPCHK(hipMemcpyAsync(stage, host + (size_t)5 * plane * k0, (size_t)5 * plane * nk * 8,
                              hipMemcpyHostToDevice, c->stream));
        HIPCHK(Launch<R>::aos_to_soa(c->stream, c->dprob, stage, mp<R>(dst), k0 - 1, nk,
                                     plane * nk));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(stage));
    return TLQ_OK;
}

template <class R> int do_download(tlq_ctx *c, const void *src, double *host) {
    const Grid &g = c->g;
    const long long plane = (long long)(g.nx + 4) * (g.ny + 4);
    const int nzt = g.nz + 4;
    int chunk = (int)std::max<long long>(1, (64ll << 20) / (plane * 5 * 8));
    if (chunk > nzt) chunk = nzt;
    double *stage = nullptr;
    HIPCHK(hipMalloc(&stage, (size_t)plane * chunk * 5 * 8));
    for (int k0 = 0; k0 < nzt; k0 += chunk) {
        const int nk = std::min(chunk, nzt - k0);
        HIPCHK(Launch<R>::soa_to_aos(c->stream, c->dprob, cp<R>(const_cast<void *>(src)), stage,
                                     k0 - 1, nk, plane * nk));
        HIPCHK(hipMemcpyAsync(host + (size_t)5 * plane * k0, stage, (size_t)5 * plane * nk * 8,
                              hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    HIPCHK(hipFree(stage));
    return TLQ_OK;
}

#define DISPATCH(ctx, fn, ...) \
    ((ctx)->prec == TLQ_FP32 ? fn<float>(ctx, ##__VA_ARGS__) : fn<double>(ctx, ##__VA_ARGS__))

template <class R> int do_calc_prims(tlq_ctx *c, int src) {
    int rc = ensure_buf(c, TLQ_BUF_UPRIM);
    if (rc) return rc;
    HIPCHK(Launch<R>::calc_prims(c->stream, c->dprob, cp<R>(buf_ref(c, src)),
                                 mp<R>(c->bufs[TLQ_BUF_UPRIM]), ncell_box(c->g)));
    return TLQ_OK;
}

template <class R> int do_time_step(tlq_ctx *c, double *dt) {
    int rc = ensure_buf(c, TLQ_BUF_UPRIM);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(c->scratch_max, 0, 4 * sizeof(unsigned long long), c->stream));
    HIPCHK(Launch<R>::maxima(c->stream, c->dprob, cp<R>(c->bufs[TLQ_BUF_UPRIM]), 1, c->scratch_max,
                             ninterior(c->g)));
    unsigned long long m[4];
    HIPCHK(hipMemcpyAsync(m, c->scratch_max, sizeof m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double a[3];
    for (int q = 0; q < 3; ++q) std::memcpy(&a[q], &m[q], 8);
    const Problem &P = c->host;
    double mm = P.dx / a[0];
    if (P.dy / a[1] < mm) mm = P.dy / a[1];
    if (P.dz / a[2] < mm) mm = P.dz / a[2];
    *dt = P.cfl * mm / std::sqrt(3.0);
    return TLQ_OK;
}

template <class R> int do_fluxes(tlq_ctx *c, int src, int choice) {
    for (int b : {TLQ_BUF_UPRIM, TLQ_BUF_F, TLQ_BUF_G, TLQ_BUF_H}) {
        int rc = ensure_buf(c, b);
        if (rc) return rc;
    }
    const Grid &g = c->g;
    HIPCHK(Launch<R>::fluxes(c->stream, c->dprob, cp<R>(buf_ref(c, src)),
                             cp<R>(c->bufs[TLQ_BUF_UPRIM]), mp<R>(c->bufs[TLQ_BUF_F]),
                             mp<R>(c->bufs[TLQ_BUF_G]), mp<R>(c->bufs[TLQ_BUF_H]), choice,
                             (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1)));
    return TLQ_OK;
}

template <class R> int do_godunov(tlq_ctx *c, int dst, double dt) {
    for (int b : {dst, TLQ_BUF_F, TLQ_BUF_G, TLQ_BUF_H}) {
        int rc = ensure_buf(c, b);
        if (rc) return rc;
    }
    const R *S = nullptr;
    if (c->host.nmasses > 0 || c->sg_on) {
        // with self-gravity every Godunov call solves the potential from
        // U^n first (HLLC.f90:16-23)
        int rc = do_sg_solve<R>(c);
        if (rc) return rc;
        rc = ensure_buf(c, TLQ_BUF_S);
        if (rc) return rc;
        if ((rc = ensure_buf(c, TLQ_BUF_UPRIM))) return rc;
        HIPCHK(Launch<R>::gravity(c->stream, c->dprob, cp<R>(c->bufs[TLQ_BUF_UPRIM]),
                                  mp<R>(c->bufs[TLQ_BUF_S]), ninterior(c->g)));
        S = cp<R>(c->bufs[TLQ_BUF_S]);
    }
    HIPCHK(Launch<R>::godunov(c->stream, c->dprob, mp<R>(buf_ref(c, dst)), cp<R>(c->u[c->cur]),
                              cp<R>(c->bufs[TLQ_BUF_F]), cp<R>(c->bufs[TLQ_BUF_G]),
                              cp<R>(c->bufs[TLQ_BUF_H]), S, dt, ninterior(c->g)));
    return TLQ_OK;
}

template <class R> int do_step(tlq_ctx *c) {
    for (int b : {TLQ_BUF_UP, TLQ_BUF_UPP}) {
        int rc = ensure_buf(c, b);
        if (rc) return rc;
    }
    void *U = c->u[c->cur];
    const long long n = 5 * c->g.nc;
    HIPCHK(Launch<R>::copy(c->stream, mp<R>(c->bufs[TLQ_BUF_UPP]), cp<R>(U), n));
    HIPCHK(Launch<R>::step(c->stream, c->dprob, mp<R>(U), cp<R>(c->bufs[TLQ_BUF_UP]),
                           cp<R>(c->bufs[TLQ_BUF_UPP]), ncell_box(c->g)));
    return TLQ_OK;
}

template <class R> int do_cooling(tlq_ctx *c, double dt) {
    int rc = ensure_buf(c, TLQ_BUF_UPRIM);
    if (rc) return rc;
    HIPCHK(Launch<R>::cooling(c->stream, c->dprob, mp<R>(c->bufs[TLQ_BUF_UPRIM]),
                              mp<R>(c->u[c->cur]), dt, ninterior(c->g)));
    return TLQ_OK;
}

template <class R> int do_check(tlq_ctx *c, long long *bad) {
    HIPCHK(hipMemsetAsync(c->scratch_cnt, 0, sizeof(unsigned long long), c->stream));
    HIPCHK(Launch<R>::check(c->stream, c->dprob, cp<R>(c->u[c->cur]), c->scratch_cnt,
                            ninterior(c->g)));
    unsigned long long v = 0;
    HIPCHK(hipMemcpyAsync(&v, c->scratch_cnt, sizeof v, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *bad = (long long)v;
    return TLQ_OK;
}

int check_ctx(tlq_ctx *c) { return c ? TLQ_OK : fail(TLQ_ERR_ARG, "null context"); }
int check_buf(int b) {
    return (b >= TLQ_BUF_U && b <= TLQ_BUF_S) ? TLQ_OK : fail(TLQ_ERR_ARG, "bad buffer id");
}

}  // namespace

extern "C" {

const char *tlq_last_error(void) { return g_err.c_str(); }

int tlq_create(const tlq_params *p, tlq_ctx **out) {
    if (!p || !out) return fail(TLQ_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->nx < 2 || p->ny < 2 || p->nz < 2) return fail(TLQ_ERR_ARG, "nx, ny, nz must be >= 2");
    for (int q = 0; q < 6; ++q) {
        if (p->bc[q] == TLQ_BC_INFLOW)
            return fail(TLQ_ERR_UNSUPPORTED,
                        "BC_INFLOW is an unimplemented stub in the reference (boundaries.f90:191)");
        if (!valid_bc(p->bc[q])) return fail(TLQ_ERR_ARG, "bad boundary id");
    }
    if (p->limiter < TLQ_LIM_NONE || p->limiter > TLQ_LIM_SUPERBEE)
        return fail(TLQ_ERR_ARG, "bad limiter id (the reference stops: HLLC.f90:630-632)");
    if (p->solver != TLQ_SOLVER_HLLC && p->solver != TLQ_SOLVER_HLL)
        return fail(TLQ_ERR_ARG, "bad solver id");
    if (p->precision != TLQ_FP64 && p->precision != TLQ_FP32)
        return fail(TLQ_ERR_ARG, "bad precision id");
    if (p->cooling < TLQ_COOL_OFF || p->cooling > TLQ_COOL_INTEGRATED)
        return fail(TLQ_ERR_ARG, "bad cooling id");
    if (!(p->gamma > 1.0) || !(p->cfl > 0.0) || !(p->xr > p->xl) || !(p->yr > p->yl) ||
        !(p->zr > p->zl) || !(p->mu0 > 0.0))
        return fail(TLQ_ERR_ARG, "bad physical parameters");
    int sg_levels = 0;
    if (p->self_gravity != TLQ_SG_OFF) {
        if (p->self_gravity != TLQ_SG_FIXED && p->self_gravity != TLQ_SG_MULTIPOLE)
            return fail(TLQ_ERR_ARG, "bad self_gravity id");
        if (p->mg_levels < 0 || p->mg_levels > SG_MAX_LEVELS) return fail(TLQ_ERR_ARG, "bad mg_levels");
        // poisson_solver and resid take the cell size from x alone (sources.f90:280, 322)
        const double dx = (p->xr - p->xl) / p->nx, dy = (p->yr - p->yl) / p->ny,
                     dz = (p->zr - p->zl) / p->nz;
        if (std::fabs(dy - dx) > 1e-12 * dx || std::fabs(dz - dx) > 1e-12 * dx)
            return fail(TLQ_ERR_UNSUPPORTED, "self-gravity needs cubic cells (dx = dy = dz)");
        sg_levels = p->mg_levels > 0 ? p->mg_levels : sg_levels_auto(p->nx, p->ny, p->nz);
        for (int l = 1; l <= sg_levels; ++l)
            if (p->nx % (1 << l) || p->ny % (1 << l) || p->nz % (1 << l))
                return fail(TLQ_ERR_ARG, "mg_levels: nx, ny, nz must be divisible by 2^mg_levels");
        if (sg_coarse_cells(p->nx, p->ny, p->nz, sg_levels) > SG_COARSE_MAX)
            return fail(TLQ_ERR_UNSUPPORTED, "coarsest multigrid level above 4096 cells");
        if (!(p->gconst >= 0.0)) return fail(TLQ_ERR_ARG, "bad gconst");
        if (!(p->mg_tol >= 0.0)) return fail(TLQ_ERR_ARG, "bad mg_tol");
    }

    tlq_ctx *c = new tlq_ctx();
    c->prm = *p;
    c->prec = p->precision;
    c->esize = p->precision == TLQ_FP32 ? 4 : 8;
    Grid &g = c->g;
    g.nx = p->nx;
    g.ny = p->ny;
    g.nz = p->nz;
    // i = 1 at offset 32 of every row; rows padded to a multiple of 32.
    const int ox = 31;  // offset of i = 0
    g.sx = ((ox + p->nx + 3 + 31) / 32) * 32;
    g.sxy = (long long)g.sx * (p->ny + 4);
    g.nc = g.sxy * (p->nz + 4);
    g.base = ox + (long long)g.sx * 1 + g.sxy * 1;  // cell (0,0,0): j = k = -1 at row/plane 0
    Problem &P = c->host;
    std::memset(&P, 0, sizeof P);
    P.g = g;
    P.xl = p->xl;
    P.yl = p->yl;
    P.zl = p->zl;
    P.dx = (p->xr - p->xl) / p->nx;
    P.dy = (p->yr - p->yl) / p->ny;
    P.dz = (p->zr - p->zl) / p->nz;
    P.gamma = p->gamma;
    P.mu0 = p->mu0;
    P.cfl = p->cfl;
    P.eta = p->eta;
    P.cooling_limit = p->cooling_limit;
    P.softening = p->softening;
    P.limiter = p->limiter;
    P.solver = p->solver;
    for (int q = 0; q < 6; ++q) P.bc[q] = p->bc[q];
    P.cooling = p->cooling;
    P.f_cT = (float)(p->mu0 * AMU / KB);
    P.f_inv_mu_amu = (float)(1.0 / (p->mu0 * AMU));
    P.f_inv_gm1 = (float)(1.0 / (p->gamma - 1.0));
    P.f_floor_f = (float)(1.0 - p->cooling_limit);
    P.net_min_dtc = (float)TLQ_SURROGATE_MIN_DTC;

    auto bail = [&](int code, const char *m) {
        tlq_destroy(c);
        return fail(code, m);
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(TLQ_ERR_HIP, "hipStreamCreate failed");
    if (hipMalloc(&c->dprob, sizeof(Problem)) != hipSuccess ||
        hipMalloc(&c->st, sizeof(StepState)) != hipSuccess ||
        hipMalloc(&c->scratch_max, 8 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&c->scratch_cnt, sizeof(unsigned long long)) != hipSuccess)
        return bail(TLQ_ERR_NOMEM, "hipMalloc failed");
    const size_t bytes = (size_t)5 * g.nc * c->esize;
    for (int q = 0; q < 2; ++q) {
        if (hipMalloc(&c->u[q], bytes) != hipSuccess) return bail(TLQ_ERR_NOMEM, "hipMalloc failed (U)");
        hipMemsetAsync(c->u[q], 0, bytes, c->stream);
    }
    if (hipMalloc(&c->uh, bytes) != hipSuccess) return bail(TLQ_ERR_NOMEM, "hipMalloc failed (U1/2)");
    hipMemsetAsync(c->uh, 0, bytes, c->stream);
    c->bufs[TLQ_BUF_UPP] = c->uh;
    if (c->prec == TLQ_FP32) {
        using S1 = TileShape<float>::S1;
        using S2 = TileShape<float>::S2;
        c->tiling1 = make_tiling(p->nx, p->ny, p->nz, S1::TX, S1::TY, S1::ZC);
        c->tiling2 = make_tiling(p->nx, p->ny, p->nz, S2::TX, S2::TY, S2::ZC);
    } else {
        using S1 = TileShape<double>::S1;
        using S2 = TileShape<double>::S2;
        c->tiling1 = make_tiling(p->nx, p->ny, p->nz, S1::TX, S1::TY, S1::ZC);
        c->tiling2 = make_tiling(p->nx, p->ny, p->nz, S2::TX, S2::TY, S2::ZC);
    }
    if (p->cooling >= TLQ_COOL_SURROGATE) {
        // queue of the dense cooling pass: one 32-bit cell offset per slot,
        // one segment of the tile's cell count per stage-2 tile
        if (g.nc >= (1ll << 32)) return bail(TLQ_ERR_UNSUPPORTED, "mesh too large for the cooling queue");
        const size_t nb = (size_t)cool_segments(c);
        const size_t qlen = nb * (size_t)cool_segment(c);
        if (hipMalloc(&c->cool_queue, qlen * sizeof(unsigned)) != hipSuccess ||
            hipMalloc(&c->cool_seg, 2 * nb * sizeof(unsigned)) != hipSuccess ||
            hipMalloc(&c->cool_off, 2 * (nb + 1) * sizeof(unsigned)) != hipSuccess ||
            hipMalloc(&c->cool_fall, sizeof(unsigned)) != hipSuccess)
            return bail(TLQ_ERR_NOMEM, "hipMalloc failed (cooling queue)");
        if (p->cooling == TLQ_COOL_SURROGATE &&
            hipMalloc(&c->cool_fqueue, qlen * sizeof(unsigned)) != hipSuccess)
            return bail(TLQ_ERR_NOMEM, "hipMalloc failed (fallback queue)");
        hipMemsetAsync(c->cool_seg, 0, 2 * nb * sizeof(unsigned), c->stream);
        hipMemsetAsync(c->cool_off, 0, 2 * (nb + 1) * sizeof(unsigned), c->stream);
        hipMemsetAsync(c->cool_fall, 0, sizeof(unsigned), c->stream);
    }
    if (p->self_gravity != TLQ_SG_OFF) {
        c->sg.boundary = p->self_gravity;
        c->sg.gconst = p->gconst;
        c->sg.tol = p->mg_tol > 0.0 ? p->mg_tol : 10.0;
        if (sg_alloc(c->sg, p->nx, p->ny, p->nz, sg_levels, p->xr - p->xl, c->stream) != hipSuccess)
            return bail(TLQ_ERR_NOMEM, "hipMalloc failed (multigrid levels)");
        c->sg_on = true;
        P.phi = c->sg.lv[0].e;
    }
    StepState s0{};
    hipMemcpyAsync(c->st, &s0, sizeof s0, hipMemcpyHostToDevice, c->stream);
    if (scratch_address(&c->noalloc_scratch) != hipSuccess)
        return bail(TLQ_ERR_HIP, "hipGetSymbolAddress failed for the allocation-free scratch");
    hipEventCreate(&c->ev0);
    hipEventCreate(&c->ev1);

    int rc = push_problem(c);
    if (rc) {
        tlq_destroy(c);
        return rc;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(TLQ_ERR_HIP, "init sync failed");
    *out = c;
    return TLQ_OK;
}

int tlq_destroy(tlq_ctx *c) {
    if (!c) return TLQ_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto &p : c->pending) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    for (int q = 0; q < 8; ++q)
        if (c->bufs[q] && c->bufs[q] != c->uh) hipFree(c->bufs[q]);
    for (int q = 0; q < 2; ++q)
        if (c->u[q]) hipFree(c->u[q]);
    if (c->uh) hipFree(c->uh);
    if (c->gacc) hipFree(c->gacc);
    if (c->dprob) hipFree(c->dprob);
    if (c->st) hipFree(c->st);
    if (c->scratch_max) hipFree(c->scratch_max);
    if (c->scratch_cnt) hipFree(c->scratch_cnt);
    if (c->cool_queue) hipFree(c->cool_queue);
    if (c->cool_fqueue) hipFree(c->cool_fqueue);
    if (c->cool_seg) hipFree(c->cool_seg);
    if (c->cool_off) hipFree(c->cool_off);
    if (c->cool_fall) hipFree(c->cool_fall);
    sg_free(c->sg);
    for (void *q : {(void *)c->sur_theta, (void *)c->sur_m, (void *)c->sur_v, (void *)c->sur_gout,
                    (void *)c->sur_thq, (void *)c->sur_X, (void *)c->sur_Y,
                    (void *)c->sur_partial, (void *)c->sur_loss})
        if (q) (void)hipFree(q);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return TLQ_OK;
}

int tlq_set_winds(tlq_ctx *c, const tlq_wind *w, int n) {
    if (int rc = check_ctx(c)) return rc;
    if (n < 0 || n > MAX_WINDS || (n > 0 && !w)) return fail(TLQ_ERR_ARG, "bad wind list");
    for (int q = 0; q < n; ++q)
        if (!all_finite(&w[q].xc, 11)) return fail(TLQ_ERR_ARG, "non-finite wind parameter");
    Problem np = c->host;
    for (int q = 0; q < n; ++q) {
        const tlq_wind &s = w[q];
        np.winds[q] = WindDev{s.xc, s.yc, s.zc, s.vx, s.vy, s.vz, s.radius, s.mdot, s.vinf,
                              s.temp, s.mu};
    }
    np.nwinds = n;
    return commit_problem(c, np);
}

// gravity_fast's point-mass acceleration of every interior cell, three
// fields in the grid's layout (the masses do not move during a run, as
// pointmass_potential's PHIP does not, pointsource_gravity.f90:12-39).
__global__ void k_gravity_field(const Problem *__restrict__ Pp, float *__restrict__ out) {
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const long long n = (long long)g.nx * g.ny * g.nz;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.nx) + 1;
        const long long r = t / g.nx;
        const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
        float a[3];
        gravity_fast(P, i, j, k, a);
        const long long cc = g.idx(i, j, k);
        out[cc] = a[0];
        out[g.nc + cc] = a[1];
        out[2 * g.nc + cc] = a[2];
    }
}

int tlq_set_point_masses(tlq_ctx *c, const double *xyzm, int n) {
    if (int rc = check_ctx(c)) return rc;
    if (n < 0 || n > MAX_MASSES || (n > 0 && !xyzm)) return fail(TLQ_ERR_ARG, "bad mass list");
    if (!all_finite(xyzm, 4 * n)) return fail(TLQ_ERR_ARG, "non-finite mass position or value");
    // FP32 mode: the acceleration of the fixed masses is computed once here
    // and read by the fused step in place of gravity_fast.  The field is
    // allocated and the new problem built before anything changes; a call
    // that fails leaves the context's masses and field as they were.
    const bool field = c->prec == TLQ_FP32 && n > 0;
    if (field && !c->gacc && hipMalloc(&c->gacc, 3 * (size_t)c->g.nc * sizeof(float)) != hipSuccess) {
        c->gacc = nullptr;
        return fail(TLQ_ERR_NOMEM, "hipMalloc failed (gravity field)");
    }
    Problem np = c->host;
    for (int q = 0; q < 4 * n; ++q) np.masses[q] = xyzm[q];
    np.nmasses = n;
    // FP32 gravity in units of L = dx (see gravity_fast)
    const double L = np.dx;
    for (int m = 0; m < n; ++m) {
        np.g_A[m] = (float)(2.0 * GR * xyzm[4 * m + 3] / (L * L));
        np.g_X0[m] = (float)((np.xl - np.dx / 2.0 - xyzm[4 * m + 0]) / L);
        np.g_Y0[m] = (float)((np.yl - np.dy / 2.0 - xyzm[4 * m + 1]) / L);
        np.g_Z0[m] = (float)((np.zl - np.dz / 2.0 - xyzm[4 * m + 2]) / L);
    }
    np.g_dY = (float)(np.dy / L);
    np.g_dZ = (float)(np.dz / L);
    np.g_e2 = (float)((np.softening / L) * (np.softening / L));
    np.gacc = field ? c->gacc : nullptr;
    const Problem old = c->host;
    if (int rc = commit_problem(c, np)) return rc;
    if (np.gacc) {
        const long long cells = ninterior(c->g);
        const unsigned nb = (unsigned)std::min<long long>((cells + 255) / 256, 65536);
        hipLaunchKernelGGL(k_gravity_field, dim3(nb), dim3(256), 0, c->stream, c->dprob, c->gacc);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            (void)commit_problem(c, old);
            return fail(TLQ_ERR_HIP, std::string("gravity field: ") + hipGetErrorString(e));
        }
    }
    return TLQ_OK;
}

int tlq_set_cooling_table(tlq_ctx *c, const double *T, const double *L, int n) {
    if (int rc = check_ctx(c)) return rc;
    if (n < 2 || n > MAX_TABLE || !T || !L) return fail(TLQ_ERR_ARG, "bad cooling table");
    // validate the whole table first: a rejected table leaves the current one intact
    for (int q = 0; q < n; ++q)
        if (!(T[q] > 0.0) || !std::isfinite(T[q]) || (q > 0 && !(T[q] > T[q - 1])))
            return fail(TLQ_ERR_ARG, "cooling table temperatures must be finite and increase");
    if (!all_finite(L, n)) return fail(TLQ_ERR_ARG, "non-finite log10 Lambda in the cooling table");
    Problem np = c->host;
    Problem &P = np;
    for (int q = 0; q < n; ++q) {
        P.tab_T[q] = T[q];
        P.tab_logT[q] = std::log10(T[q]);
        P.tab_L[q] = L[q];
    }
    P.ntable = n;
    P.tmin = T[0];
    for (int q = 0; q < n; ++q) {
        P.tab_Tf[q] = (float)T[q];
        P.tab_lgTf[q] = (float)P.tab_logT[q];
        P.tab_Lf[q] = (float)L[q];
        P.tab_slopef[q] = q ? (float)((L[q] - L[q - 1]) / (P.tab_logT[q] - P.tab_logT[q - 1])) : 0.f;
    }
    P.f_tmin = (float)T[0];
    P.f_tmax = (float)T[n - 1];
    {  // cumulative cooling integral Y at the nodes (closed form per power-law segment)
        double y = 0.0;
        P.tab_Ycum[0] = 0.0;
        P.tab_ycoef[0] = 0.f;
        for (int q = 1; q < n; ++q) {
            const double sl = (L[q] - L[q - 1]) / (P.tab_logT[q] - P.tab_logT[q - 1]);
            const double coef = T[q - 1] / std::pow(10.0, L[q - 1]);
            const double a = 1.0 - sl, ratio = T[q] / T[q - 1];
            y += std::fabs(a) > 1e-12 ? coef * (std::pow(ratio, a) - 1.0) / a : coef * std::log(ratio);
            P.tab_Ycum[q] = y;
            P.tab_ycoef[q] = (float)coef;
        }
    }
    // direct indexing when every node lies within 0.4 spacings of a uniform grid
    const double d = (P.tab_logT[n - 1] - P.tab_logT[0]) / (n - 1);
    double dev = 0.0;
    for (int q = 0; q < n; ++q)
        dev = std::max(dev, std::fabs(P.tab_logT[q] - (P.tab_logT[0] + q * d)));
    P.tab_uniform = dev < 0.4 * d ? 1 : 0;
    P.f_lgT0 = (float)P.tab_logT[0];
    P.f_inv_dlg = (float)(1.0 / d);
    return commit_problem(c, np);
}

int tlq_upload(tlq_ctx *c, int buf, const double *host) {
    if (int rc = check_ctx(c)) return rc;
    if (int rc = check_buf(buf)) return rc;
    if (!host) return fail(TLQ_ERR_ARG, "null host array");
    if (int rc = ensure_buf(c, buf)) return rc;
    if (buf == TLQ_BUF_U) c->fused_ready = false;
    return DISPATCH(c, do_upload, buf_ref(c, buf), host);
}

int tlq_download(tlq_ctx *c, int buf, double *host) {
    if (int rc = check_ctx(c)) return rc;
    if (int rc = check_buf(buf)) return rc;
    if (!host) return fail(TLQ_ERR_ARG, "null host array");
    if (!buf_ref(c, buf)) return fail(TLQ_ERR_STATE, "buffer not allocated yet");
    return DISPATCH(c, do_download, buf_ref(c, buf), host);
}

int tlq_calc_prims(tlq_ctx *c, int src) {
    if (int rc = check_ctx(c)) return rc;
    if (src != TLQ_BUF_U && src != TLQ_BUF_UP && src != TLQ_BUF_UPP)
        return fail(TLQ_ERR_ARG, "calc_prims source must be U, UP or UPP");
    if (!buf_ref(c, src)) return fail(TLQ_ERR_STATE, "source array not set");
    return DISPATCH(c, do_calc_prims, src);
}

int tlq_time_step(tlq_ctx *c, double *dt) {
    if (int rc = check_ctx(c)) return rc;
    if (!dt) return fail(TLQ_ERR_ARG, "null dt");
    return DISPATCH(c, do_time_step, dt);
}

int tlq_fluxes(tlq_ctx *c, int src, int choice) {
    if (int rc = check_ctx(c)) return rc;
    if (choice != 1 && choice != 2) return fail(TLQ_ERR_ARG, "choice must be 1 or 2");
    if (src != TLQ_BUF_U && src != TLQ_BUF_UPP) return fail(TLQ_ERR_ARG, "flux source must be U or UPP");
    return DISPATCH(c, do_fluxes, src, choice);
}

int tlq_godunov(tlq_ctx *c, int dst, double dt) {
    if (int rc = check_ctx(c)) return rc;
    if (dst != TLQ_BUF_UP && dst != TLQ_BUF_UPP) return fail(TLQ_ERR_ARG, "Godunov target must be UP or UPP");
    return DISPATCH(c, do_godunov, dst, dt);
}

int tlq_borders(tlq_ctx *c, int buf, int order) {
    if (int rc = check_ctx(c)) return rc;
    if (int rc = check_buf(buf)) return rc;
    if (order != 1 && order != 2) return fail(TLQ_ERR_ARG, "order must be 1 or 2");
    if (int rc = ensure_buf(c, buf)) return rc;
    if (buf == TLQ_BUF_U) c->fused_ready = false;
    return DISPATCH(c, do_borders, buf_ref(c, buf), order);
}

int tlq_step(tlq_ctx *c, double *t, int *it, double dt) {
    if (int rc = check_ctx(c)) return rc;
    int rc = DISPATCH(c, do_step);
    if (rc) return rc;
    c->fused_ready = false;
    ++c->it_host;
    if (t) *t = *t + dt;  // HydroCore.f90:208-209
    if (it) *it = *it + 1;
    return TLQ_OK;
}

int tlq_cooling(tlq_ctx *c, double dt) {
    if (int rc = check_ctx(c)) return rc;
    if (c->host.ntable < 2) return fail(TLQ_ERR_STATE, "cooling table not set");
    c->fused_ready = false;
    return DISPATCH(c, do_cooling, dt);
}

int tlq_init_state(tlq_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    int rc = DISPATCH(c, do_borders, c->u[c->cur], 1);  // mainHD.f90:38
    if (rc) return rc;
    // t = 0, it = 0 (mainHD.f90:22-27)
    StepState s0{};
    HIPCHK(hipMemcpyAsync(c->st, &s0, sizeof s0, hipMemcpyHostToDevice, c->stream));
    c->it_host = 0;
    rc = DISPATCH(c, do_sg_solve);                        // mainHD.f90:50-52
    if (rc) return rc;
    return DISPATCH(c, do_init_maxima);                   // calc_prims + first time_step
}

int tlq_advance_n(tlq_ctx *c, int n) {
    if (int rc = check_ctx(c)) return rc;
    if (n < 0) return fail(TLQ_ERR_ARG, "negative step count");
    if (c->host.cooling != TLQ_COOL_OFF && c->host.ntable < 2)
        return fail(TLQ_ERR_STATE, "cooling enabled but no table set");
    if (c->host.cooling == TLQ_COOL_SURROGATE && !c->sur_ready)
        return fail(TLQ_ERR_STATE, "surrogate cooling needs trained or loaded parameters");
    return DISPATCH(c, do_advance, n);
}

int tlq_resume(tlq_ctx *c, double t, int it, int fill_ghosts) {
    if (int rc = check_ctx(c)) return rc;
    if (!(t >= 0.0 && t < HUGE_VAL) || it < 0) return fail(TLQ_ERR_ARG, "bad time or step count");
    if (fill_ghosts != 0 && fill_ghosts != 1) return fail(TLQ_ERR_ARG, "fill_ghosts must be 0 or 1");
    // A dump with ghosts keeps them: U's ghosts are images of UP, not of U's
    // interior (HydroCore.f90:195-206), and a boundary fill would lose that.
    if (fill_ghosts) {
        int rc = DISPATCH(c, do_borders, c->u[c->cur], 1);
        if (rc) return rc;
    }
    StepState s0{};
    s0.time[c->cur] = t;
    s0.it = it;
    HIPCHK(hipMemcpyAsync(c->st, &s0, sizeof s0, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->it_host = it;
    if (int rc = DISPATCH(c, do_sg_solve)) return rc;     // from the PHI uploaded, if any
    return DISPATCH(c, do_init_maxima);
}

int tlq_upload_phi(tlq_ctx *c, const double *phi) {
    if (int rc = check_ctx(c)) return rc;
    if (!c->sg_on) return fail(TLQ_ERR_STATE, "self-gravity is off");
    if (!phi) return fail(TLQ_ERR_ARG, "null phi");
    const SgLevel &L = c->sg.lv[0];
    HIPCHK(hipMemcpyAsync(L.e, phi, sizeof(double) * (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2),
                          hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TLQ_OK;
}

int tlq_download_phi(tlq_ctx *c, double *phi) {
    if (int rc = check_ctx(c)) return rc;
    if (!c->sg_on) return fail(TLQ_ERR_STATE, "self-gravity is off");
    if (!phi) return fail(TLQ_ERR_ARG, "null phi");
    const SgLevel &L = c->sg.lv[0];
    HIPCHK(hipMemcpyAsync(phi, L.e, sizeof(double) * (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2),
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TLQ_OK;
}

int tlq_gravity_solve(tlq_ctx *c, int *cycles, double *dif) {
    if (int rc = check_ctx(c)) return rc;
    if (!c->sg_on) return fail(TLQ_ERR_STATE, "self-gravity is off");
    if (int rc = DISPATCH(c, do_sg_solve)) return rc;
    if (cycles) *cycles = c->sg.cycles;
    if (dif) *dif = c->sg.dif;
    return TLQ_OK;
}

int tlq_gravity_info(tlq_ctx *c, int *levels, int *cycles, double *dif, int *it) {
    if (int rc = check_ctx(c)) return rc;
    if (levels) *levels = c->sg_on ? c->sg.levels : -1;
    if (cycles) *cycles = c->sg.cycles;
    if (dif) *dif = c->sg.dif;
    if (it) *it = c->it_host;
    return TLQ_OK;
}

int tlq_get_time(tlq_ctx *c, double *dt, double *t, int *it) {
    if (int rc = check_ctx(c)) return rc;
    StepState s;
    HIPCHK(hipMemcpyAsync(&s, c->st, sizeof s, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (dt) *dt = s.dt;
    if (t) *t = s.time[c->cur];
    if (it) *it = (int)s.it;
    return TLQ_OK;
}

int tlq_advance(tlq_ctx *c, double *dt, double *t, int *it) {
    int rc = tlq_advance_n(c, 1);
    if (rc) return rc;
    if (dt || t || it) return tlq_get_time(c, dt, t, it);
    return TLQ_OK;
}

int tlq_sync(tlq_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return TLQ_OK;
}

int tlq_check_state(tlq_ctx *c, long long *bad) {
    if (int rc = check_ctx(c)) return rc;
    if (!bad) return fail(TLQ_ERR_ARG, "null output");
    return DISPATCH(c, do_check, bad);
}

int tlq_cooling_counts(tlq_ctx *c, long long *queued, long long *fallbacks) {
    if (int rc = check_ctx(c)) return rc;
    if (!queued || !fallbacks) return fail(TLQ_ERR_ARG, "null output");
    *queued = *fallbacks = 0;
    if (!c->cool_off) return TLQ_OK;
    // mode 2: the network's cells (off[nseg]) and its rejections; mode 3: the
    // integrator's cells (off[2 nseg + 1])
    const long long nseg = cool_segments(c);
    const bool net = c->host.cooling == TLQ_COOL_SURROGATE;
    unsigned q = 0, f = 0;
    HIPCHK(hipMemcpyAsync(&q, c->cool_off + (net ? nseg : 2 * nseg + 1), sizeof q,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&f, c->cool_fall, sizeof f, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *queued = q;
    *fallbacks = net ? f : 0;
    return TLQ_OK;
}

int tlq_fallback_fraction(tlq_ctx *c, double *frac) {
    if (!frac) return fail(TLQ_ERR_ARG, "null output");
    long long nq = 0, nf = 0;
    if (int rc = tlq_cooling_counts(c, &nq, &nf)) return rc;
    *frac = nq ? (double)nf / (double)nq : 0.0;
    return TLQ_OK;
}

int tlq_timer_start(tlq_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    return TLQ_OK;
}

int tlq_timer_stop(tlq_ctx *c, double *ms) {
    if (int rc = check_ctx(c)) return rc;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float f = 0.f;
    HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
    if (ms) *ms = f;
    return TLQ_OK;
}

int tlq_kernel_timing(tlq_ctx *c, int enable) {
    if (int rc = check_ctx(c)) return rc;
    resolve_timings(c);
    c->ktiming = enable != 0;
    for (int q = 0; q < kNumKernels; ++q) {
        c->ktotal[q] = 0.0;
        c->kcount[q] = 0;
    }
    return TLQ_OK;
}

int tlq_kernel_stats(tlq_ctx *c, int index, char *name, int name_len, double *total_ms,
                     long long *launches) {
    if (int rc = check_ctx(c)) return rc;
    if (index < 0 || index >= kNumKernels) return fail(TLQ_ERR_ARG, "bad kernel index");
    resolve_timings(c);
    if (name && name_len > 0) {
        std::strncpy(name, kKernelNames[index], name_len - 1);
        name[name_len - 1] = '\0';
    }
    if (total_ms) *total_ms = c->ktotal[index];
    if (launches) *launches = c->kcount[index];
    return TLQ_OK;
}

// ---- learned cooling surrogate -------------------------------------------

static int sur_alloc(tlq_ctx *c) {
    if (c->sur_theta) return TLQ_OK;
    const int np = mlp::NP;
    if (hipMalloc(&c->sur_theta, np * 4) != hipSuccess || hipMalloc(&c->sur_m, np * 4) != hipSuccess ||
        hipMalloc(&c->sur_v, np * 4) != hipSuccess || hipMalloc(&c->sur_gout, np * 4) != hipSuccess ||
        hipMalloc(&c->sur_thq, np * 2) != hipSuccess ||
        hipMalloc(&c->sur_partial, (size_t)1024 * (np + 1) * 4) != hipSuccess ||
        hipMalloc(&c->sur_loss, 4) != hipSuccess)
        return fail(TLQ_ERR_NOMEM, "surrogate allocation failed");
    HIPCHK(hipMemsetAsync(c->sur_m, 0, np * 4, c->stream));
    HIPCHK(hipMemsetAsync(c->sur_v, 0, np * 4, c->stream));
    return TLQ_OK;
}

int tlq_surrogate_set_params(tlq_ctx *c, const float *theta, int n) {
    if (int rc = check_ctx(c)) return rc;
    if (!theta || n != mlp::NP) return fail(TLQ_ERR_ARG, "expected TLQ_MLP_NPARAMS parameters");
    if (int rc = sur_alloc(c)) return rc;
    HIPCHK(hipMemcpyAsync(c->sur_theta, theta, n * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->sur_m, 0, n * 4, c->stream));
    HIPCHK(hipMemsetAsync(c->sur_v, 0, n * 4, c->stream));
    HIPCHK(sur_quantize(c->stream, c->sur_theta, c->sur_thq));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->sur_t = 0;
    c->sur_ready = true;
    return TLQ_OK;
}

int tlq_surrogate_get_params(tlq_ctx *c, float *theta, int n) {
    if (int rc = check_ctx(c)) return rc;
    if (!theta || n != mlp::NP) return fail(TLQ_ERR_ARG, "expected TLQ_MLP_NPARAMS parameters");
    if (!c->sur_theta) return fail(TLQ_ERR_STATE, "surrogate not initialised");
    HIPCHK(hipMemcpyAsync(theta, c->sur_theta, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TLQ_OK;
}

int tlq_surrogate_set_threshold(tlq_ctx *c, double dtc_min) {
    if (int rc = check_ctx(c)) return rc;
    if (!(dtc_min >= 0.0 && dtc_min < 1e30)) return fail(TLQ_ERR_ARG, "bad threshold");
    Problem np = c->host;
    np.net_min_dtc = (float)dtc_min;
    return commit_problem(c, np);
}

int tlq_surrogate_init(tlq_ctx *c, unsigned long long seed) {
    // Uniform initialisation with the fan-in scale (Glorot-style bounds), zero biases.
    std::vector<float> th(mlp::NP, 0.f);
    unsigned long long z = seed * 0x9E3779B97F4A7C15ull + 1;
    auto u01 = [&]() {
        z ^= z << 13;
        z ^= z >> 7;
        z ^= z << 17;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    const double a1 = std::sqrt(6.0 / (3 + 64)), a2 = std::sqrt(6.0 / (64 + 64)),
                 a3 = std::sqrt(6.0 / (64 + 1));
    for (int q = 0; q < 3 * 64; ++q) th[mlp::OW1 + q] = (float)(a1 * (2 * u01() - 1));
    for (int q = 0; q < 64 * 64; ++q) th[mlp::OW2 + q] = (float)(a2 * (2 * u01() - 1));
    for (int q = 0; q < 64; ++q) th[mlp::OW3 + q] = (float)(a3 * (2 * u01() - 1));
    return tlq_surrogate_set_params(c, th.data(), mlp::NP);
}

int tlq_surrogate_generate(tlq_ctx *c, long long n, unsigned long long seed) {
    if (int rc = check_ctx(c)) return rc;
    if (n < 64) return fail(TLQ_ERR_ARG, "need at least 64 samples");
    if (c->host.ntable < 2) return fail(TLQ_ERR_STATE, "cooling table not set");
    if (c->sur_X) {
        (void)hipFree(c->sur_X);
        (void)hipFree(c->sur_Y);
        c->sur_X = c->sur_Y = nullptr;
    }
    if (hipMalloc(&c->sur_X, (size_t)n * 12) != hipSuccess || hipMalloc(&c->sur_Y, (size_t)n * 4) != hipSuccess)
        return fail(TLQ_ERR_NOMEM, "sample allocation failed");
    HIPCHK(sur_generate(c->stream, c->dprob, n, seed, c->sur_X, c->sur_Y, 0.01, 100000));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->sur_n = n;
    return TLQ_OK;
}

int tlq_surrogate_samples(tlq_ctx *c, long long first, long long n, float *x, float *y) {
    if (int rc = check_ctx(c)) return rc;
    if (!c->sur_X || first < 0 || n < 0 || first + n > c->sur_n) return fail(TLQ_ERR_ARG, "bad range");
    if (x) HIPCHK(hipMemcpyAsync(x, c->sur_X + 3 * first, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream));
    if (y) HIPCHK(hipMemcpyAsync(y, c->sur_Y + first, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return TLQ_OK;
}

int tlq_surrogate_train(tlq_ctx *c, int steps, int batch, double lr_max, double lr_min,
                        double *final_loss) {
    if (int rc = check_ctx(c)) return rc;
    if (!c->sur_X) return fail(TLQ_ERR_STATE, "generate samples first");
    if (steps < 1 || batch < 64 || batch > c->sur_n || batch > 1024 * 64)
        return fail(TLQ_ERR_ARG, "bad steps/batch");
    if (!c->sur_theta) {
        int rc = tlq_surrogate_init(c, 1234);
        if (rc) return rc;
    }
    const int nb = std::min(1024, (batch + 63) / 64);
    const long long span = c->sur_n - batch;
    for (int s = 0; s < steps; ++s) {
        const double ph = steps > 1 ? (double)s / (steps - 1) : 1.0;
        const float lr = (float)(lr_min + 0.5 * (lr_max - lr_min) * (1.0 + std::cos(M_PI * ph)));
        const long long off = span > 0 ? ((long long)s * batch) % (span + 1) : 0;
        HIPCHK(sur_grad(c->stream, c->sur_thq, c->sur_X, c->sur_Y, off, batch, c->sur_partial, nb));
        HIPCHK(sur_adam(c->stream, c->sur_partial, nb, c->sur_theta, c->sur_thq, c->sur_m, c->sur_v,
                        ++c->sur_t, lr, c->sur_loss));
    }
    float l = 0.f;
    HIPCHK(hipMemcpyAsync(&l, c->sur_loss, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (final_loss) *final_loss = l;
    c->sur_ready = true;
    return TLQ_OK;
}

int tlq_surrogate_eval(tlq_ctx *c, const float *x, long long n, float *out) {
    if (int rc = check_ctx(c)) return rc;
    if (!c->sur_ready) return fail(TLQ_ERR_STATE, "surrogate has no parameters");
    if (!x || !out || n < 1) return fail(TLQ_ERR_ARG, "bad arguments");
    float *dx = nullptr, *dy = nullptr;
    HIPCHK(hipMalloc(&dx, (size_t)n * 12));
    HIPCHK(hipMalloc(&dy, (size_t)n * 4));
    HIPCHK(hipMemcpyAsync(dx, x, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
    HIPCHK(sur_eval(c->stream, c->sur_thq, dx, n, dy));
    HIPCHK(hipMemcpyAsync(out, dy, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(dx);
    (void)hipFree(dy);
    return TLQ_OK;
}

int tlq_surrogate_grad(tlq_ctx *c, const float *x, const float *y, int n, double *grad,
                       double *loss) {
    if (int rc = check_ctx(c)) return rc;
    if (!c->sur_ready) return fail(TLQ_ERR_STATE, "surrogate has no parameters");
    if (!x || !y || !grad || n < 1 || n > 1024 * 64) return fail(TLQ_ERR_ARG, "bad arguments");
    float *dx = nullptr, *dy = nullptr;
    HIPCHK(hipMalloc(&dx, (size_t)n * 12));
    HIPCHK(hipMalloc(&dy, (size_t)n * 4));
    HIPCHK(hipMemcpyAsync(dx, x, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(dy, y, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    const int nb = std::min(1024, (n + 63) / 64);
    HIPCHK(sur_grad(c->stream, c->sur_thq, dx, dy, 0, n, c->sur_partial, nb));
    HIPCHK(sur_adam(c->stream, c->sur_partial, nb, nullptr, nullptr, c->sur_gout, nullptr, 0, 0.f,
                    c->sur_loss));
    std::vector<float> g(mlp::NP);
    float l = 0.f;
    HIPCHK(hipMemcpyAsync(g.data(), c->sur_gout, mlp::NP * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&l, c->sur_loss, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < mlp::NP; ++q) grad[q] = g[q];
    if (loss) *loss = l;
    (void)hipFree(dx);
    (void)hipFree(dy);
    return TLQ_OK;
}

// tlq_surrogate_grad on device arrays, with its partial gradients in the
// static scratch of this device: no allocation, no synchronisation.
int tlq_surrogate_grad_noalloc(tlq_ctx *c, const float *x, const float *y, int n, float *grad,
                               float *loss) {
    if (int rc = check_ctx(c)) return rc;
    if (!c->sur_ready) return fail(TLQ_ERR_STATE, "surrogate has no parameters");
    if (!x || !y || !grad || !loss) return fail(TLQ_ERR_ARG, "null device pointer");
    if (n < 1 || n > TLQ_NOALLOC_MAX_SAMPLES)
        return fail(TLQ_ERR_ARG, "tlq_surrogate_grad_noalloc takes 1 to TLQ_NOALLOC_MAX_SAMPLES "
                                 "(16384) samples: its scratch holds 256 blocks of 64");
    if (!c->noalloc_scratch) return fail(TLQ_ERR_STATE, "no scratch address on this device");
    // the block count of tlq_surrogate_grad, so both give the same sums
    const int nb = (n + 63) / 64;
    HIPCHK(sur_grad(c->stream, c->sur_thq, x, y, 0, n, c->noalloc_scratch, nb));
    HIPCHK(sur_adam(c->stream, c->noalloc_scratch, nb, nullptr, nullptr, grad, nullptr, 0, 0.f,
                    loss));
    return TLQ_OK;
}

__global__ void k_stream_copy(const float4 *__restrict__ a, float4 *__restrict__ b, long long n) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x)
        b[t] = a[t];
}

// The same copy with UNR independent 16-byte loads in flight per thread
// (all issued before the first store) and, when NT, nontemporal loads and
// stores (the streamed data bypass the caches' normal retention).
extern "C++" {
typedef float f4v __attribute__((ext_vector_type(4)));

template <int UNR, bool NT>
__global__ void k_stream_copy_x(const f4v *__restrict__ a, f4v *__restrict__ b, long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (; t + (UNR - 1) * stride < n; t += UNR * stride) {
        f4v v[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if constexpr (NT) v[u] = __builtin_nontemporal_load(a + t + u * stride);
            else v[u] = a[t + u * stride];
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) {
            if constexpr (NT) __builtin_nontemporal_store(v[u], b + t + u * stride);
            else b[t + u * stride] = v[u];
        }
    }
    for (; t < n; t += stride) b[t] = a[t];
}

static hipError_t launch_copy(hipStream_t s, int unroll, bool nt, unsigned blocks, const float4 *a4,
                              float4 *b4, long long n) {
    const dim3 g(blocks), t(256);
    const f4v *a = reinterpret_cast<const f4v *>(a4);
    f4v *b = reinterpret_cast<f4v *>(b4);
    if (unroll == 1 && !nt) hipLaunchKernelGGL(k_stream_copy, g, t, 0, s, a4, b4, n);
    else if (unroll == 1) hipLaunchKernelGGL((k_stream_copy_x<1, true>), g, t, 0, s, a, b, n);
    else if (unroll == 2 && !nt) hipLaunchKernelGGL((k_stream_copy_x<2, false>), g, t, 0, s, a, b, n);
    else if (unroll == 2) hipLaunchKernelGGL((k_stream_copy_x<2, true>), g, t, 0, s, a, b, n);
    else if (unroll == 4 && !nt) hipLaunchKernelGGL((k_stream_copy_x<4, false>), g, t, 0, s, a, b, n);
    else if (unroll == 4) hipLaunchKernelGGL((k_stream_copy_x<4, true>), g, t, 0, s, a, b, n);
    else if (unroll == 8 && !nt) hipLaunchKernelGGL((k_stream_copy_x<8, false>), g, t, 0, s, a, b, n);
    else hipLaunchKernelGGL((k_stream_copy_x<8, true>), g, t, 0, s, a, b, n);
    return hipGetLastError();
}
}  // extern "C++"

int tlq_copy_probe_config(tlq_ctx *c, long long bytes, int iters, int unroll, int blocks,
                          int nontemporal, double *ms) {
    if (int rc = check_ctx(c)) return rc;
    if (bytes < 16 || iters < 1 || !ms) return fail(TLQ_ERR_ARG, "bad probe arguments");
    if (unroll != 1 && unroll != 2 && unroll != 4 && unroll != 8)
        return fail(TLQ_ERR_ARG, "probe unroll must be 1, 2, 4 or 8");
    if (blocks < 1 || blocks > (1 << 22)) return fail(TLQ_ERR_ARG, "probe blocks out of range");
    const long long n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return fail(TLQ_ERR_NOMEM, "probe alloc");
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return fail(TLQ_ERR_NOMEM, "probe alloc");
    }
    auto release = [&]() {
        (void)hipFree(a);
        (void)hipFree(b);
    };
    const bool nt = nontemporal != 0;
    if (hipMemsetAsync(a, 0, n * 16, c->stream) != hipSuccess ||
        launch_copy(c->stream, unroll, nt, (unsigned)blocks, a, b, n) != hipSuccess ||  // warm-up
        hipEventRecord(c->ev0, c->stream) != hipSuccess) {
        release();
        return fail(TLQ_ERR_HIP, "probe launch");
    }
    for (int q = 0; q < iters; ++q)
        if (launch_copy(c->stream, unroll, nt, (unsigned)blocks, a, b, n) != hipSuccess) {
            release();
            return fail(TLQ_ERR_HIP, "probe launch");
        }
    float f = 0.f;
    const bool ok = hipEventRecord(c->ev1, c->stream) == hipSuccess &&
                    hipEventSynchronize(c->ev1) == hipSuccess &&
                    hipEventElapsedTime(&f, c->ev0, c->ev1) == hipSuccess;
    release();
    if (!ok) return fail(TLQ_ERR_HIP, "probe timing");
    *ms = f / iters;
    return TLQ_OK;
}

int tlq_copy_probe(tlq_ctx *c, long long bytes, int iters, double *ms) {
    // the plain one-load-per-iteration copy over 8192 blocks of 256 threads
    return tlq_copy_probe_config(c, bytes, iters, 1, 256 * 8 * 4, 0, ms);
}

int tlq_layout(tlq_ctx *c, long long *cells, int *pitch_x, int *origin_x) {
    if (int rc = check_ctx(c)) return rc;
    if (cells) *cells = c->g.nc;
    if (pitch_x) *pitch_x = c->g.sx;
    if (origin_x) *origin_x = (int)(c->g.base - c->g.sx - c->g.sxy - 1);
    return TLQ_OK;
}

}  // extern "C"
