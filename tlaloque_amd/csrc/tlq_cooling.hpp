// tlq_cooling.hpp -- per-cell cooling arithmetic shared by the fused step
// (k_stage2, which classifies every cell in cooling modes 2 and 3) and the
// dense cooling passes of the surrogate (tlq_surrogate.hip).
//
// The converged integrator applies apply_cooling's per-cell formula
// (cooling.f90:67-112) in substeps of at most COOL_EPS t_cool, re-reading
// Lambda every substep; with one substep it is apply_cooling itself.
#pragma once

#include "tlq_device.hpp"

namespace tlq {

// ---------------------------------------------------------------------------
// Cooling integrated to convergence: apply_cooling's per-cell formula
// (cooling.f90:67-112) in substeps h <= eps t_cool with Lambda re-read at the
// current temperature, until dt is used or T < T_min (the reference's gate).
static __device__ double cool_sub_d(const Problem &P, double n, double T, double dt, double eps,
                             int max_sub) {
    double left = dt;
    for (int s = 0; left > 0.0 && T >= P.tmin && s < max_sub; ++s) {
        const double lam = find_coef(P, T);
        const double eth = n * KB * T / (P.gamma - 1.0);
        const double ce = lam * (n * n) / eth;
        double h = eps / ce;
        if (h > left) h = left;
        double f = exp(-ce * h);
        if (f < 1.0 - P.cooling_limit) f = 1.0 - P.cooling_limit;
        if (T * f < 10.0) f = 10.0 / T;
        T = T * f;
        left -= h;
    }
    return T;
}

// The FP32 cooling table as the dense cooling passes hold it in LDS: the rows
// of cool_fast_tab (TabRow: T, log10 T, log10 Lambda, slope of the segment
// ending at the row) and, for the closed-form cooling time, the node values
// of Y (FP64) with the segment coefficients.  The values are the Problem's
// FP32 table (and tab_Ycum, tab_ycoef), so a lookup gives what a lookup in
// global memory gives, at LDS latency.
struct YRow {
    double Ycum;  // Y at the row's temperature
    float ycoef;  // T_{i-1} / Lambda_{i-1} of the segment ending at the row
    float pad;
};

struct CoolTab {
    CoolF32 C;
    float cgk;     // (gamma - 1) / k_B
    float snap_T;  // SNAP T_min: where an integration that reaches the gate ends
    const TabRow *row;
    const YRow *y;  // null when only the rows are staged
};

// LDS bytes of a staged table of n rows (rows only, or rows and Y).
__host__ __device__ constexpr unsigned cool_tab_bytes(int n, bool with_y) {
    return (unsigned)n * (unsigned)(sizeof(TabRow) + (with_y ? sizeof(YRow) : 0));
}

// Fills row[] (and y[] when not null) from the Problem; the caller
// synchronises the block before use.
__device__ __forceinline__ void stage_cool_tab(const Problem &P, TabRow *row, YRow *y, int tid,
                                               int nthreads) {
    stage_table(P, row, tid, nthreads);
    if (y)
        for (int q = tid; q < P.ntable; q += nthreads) y[q] = YRow{P.tab_Ycum[q], P.tab_ycoef[q], 0.f};
}

constexpr double SNAP = 0.995;

__device__ __forceinline__ CoolTab cool_tab(const Problem &P, const TabRow *row, const YRow *y) {
    return CoolTab{cool_consts(P), (float)((P.gamma - 1.0) / KB), (float)(SNAP * P.tmin), row, y};
}

// Table segment of temperature T (first row with T_i > T, as find_coef),
// clamped to the last segment for T >= T_max.
__device__ __forceinline__ int table_row(const CoolTab &t, float temp, float lg) {
    const int n = t.C.ntable;
    if (!(temp < t.C.tmax)) return n - 1;
    int hi;
    if (t.C.uniform) {
        hi = (int)((lg - t.C.lgT0) * t.C.inv_dlg) + 1;
        hi = hi < 1 ? 1 : (hi > n - 1 ? n - 1 : hi);
        if (!(t.row[hi].T > temp) && hi < n - 1) ++hi;
        if (hi > 1 && t.row[hi - 1].T > temp) --hi;
    } else {
        int lo = 0;
        hi = n - 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (t.row[mid].T > temp) hi = mid;
            else lo = mid;
        }
    }
    return hi;
}

// Lambda(T) from the table in FP32 (find_coef's interpolation).
__device__ __forceinline__ float lambda_f(const CoolTab &t, float temp) {
    const float lg = __builtin_amdgcn_logf(temp) * 0.301029995663981f;
    if (!(temp < t.C.tmax)) return __builtin_amdgcn_exp2f(t.row[t.C.ntable - 1].L * 3.32192809488736f);
    const int hi = table_row(t, temp, lg);
    const float lgl = t.row[hi - 1].L + t.row[hi].slope * (lg - t.row[hi - 1].lgT);
    return __builtin_amdgcn_exp2f(lgl * 3.32192809488736f);
}

// Y(T) = int_{T_0}^{T} dT'/Lambda(T'): on each table segment Lambda is a
// power law (linear in log-log), so Y is closed form.  The time to cool from
// T0 to T1 at density n is t = k_B / ((gamma-1) n) (Y(T0) - Y(T1)).  Y is
// split into its value at the segment's lower node (double) and the part
// inside the segment (float), so that Y(T0) - Y(T1) keeps its accuracy when
// T1 is close to T0.
struct YSplit {
    double node;
    float part;
};
__device__ __forceinline__ YSplit cool_Y(const CoolTab &t, float temp) {
    if (!(temp > t.C.tmin)) return {0.0, 0.f};
    const float l2 = __builtin_amdgcn_logf(temp);
    const int hi = table_row(t, temp, l2 * 0.301029995663981f);
    const float a = 1.f - t.row[hi].slope;
    const float lr = l2 - __builtin_amdgcn_logf(t.row[hi - 1].T);  // log2(T/T_{hi-1})
    const float seg = fabsf(a) > 1e-4f ? (__builtin_amdgcn_exp2f(a * lr) - 1.f) / a
                                       : lr * 0.6931471805599453f;
    return {t.y[hi - 1].Ycum, t.y[hi].ycoef * seg};
}
__device__ __forceinline__ float cool_dY(const CoolTab &t, float T0, float T1) {
    const YSplit a = cool_Y(t, T0), b = cool_Y(t, T1);
    return (float)(a.node - b.node) + (a.part - b.part);
}

static __device__ float cool_sub_f(const CoolTab &t, float n, float T, float dt, float eps,
                                   int max_sub) {
    float left = dt;
    for (int s = 0; left > 0.f && T >= t.C.tmin && s < max_sub; ++s) {
        const float ce = lambda_f(t, T) * n * t.cgk * __builtin_amdgcn_rcpf(T);
        const float h = fminf(left, eps * __builtin_amdgcn_rcpf(ce));
        float f = __builtin_amdgcn_exp2f(-1.4426950408889634f * ce * h);
        if (f < t.C.floor_f) f = t.C.floor_f;
        if (T * f < 10.f) f = 10.f * __builtin_amdgcn_rcpf(T);
        T = T * f;
        left -= h;
    }
    return T;
}

// ---------------------------------------------------------------------------
// Exact integration of the cooling equation dT/dt = -(gamma-1) n Lambda(T)/k_B
// over dt for Lambda as find_coef interpolates it (a power law on each table
// segment, the last row's value above the table): the time to cool from T0 to
// T1 is k_B/((gamma-1) n) (Y(T0) - Y(T1)) with Y closed form (cool_Y), so T(dt)
// follows by inverting Y on the segment where Y(T0) - dt (gamma-1) n/k_B falls
// (Townsend 2009, ApJS 181, 391).  It is the eps -> 0 limit of the substepped
// integrator.  At the gate (the cooling time to T_min at most dt) the result is
// SNAP T_min, where the substepped integrator's last substep ends.

// FP64: the node values of Y from the host (tab_Ycum), the rest with the
// device's pow/log/exp.
static __device__ double cool_int_d(const Problem &P, double n, double T0, double dt) {
    if (!(T0 >= P.tmin)) return T0;
    const int last = P.ntable - 1;
    const double lam_top = pow(10.0, P.tab_L[last]);
    auto seg = [&](int i, double &a, double &coef) {
        a = 1.0 - (P.tab_L[i] - P.tab_L[i - 1]) / (P.tab_logT[i] - P.tab_logT[i - 1]);
        coef = P.tab_T[i - 1] / pow(10.0, P.tab_L[i - 1]);
    };
    double y0;
    if (T0 >= P.tab_T[last]) {
        y0 = P.tab_Ycum[last] + (T0 - P.tab_T[last]) / lam_top;
    } else {  // first row with T_i > T0 (find_coef's segment)
        int lo = 0, hi = last;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (P.tab_T[mid] > T0) hi = mid;
            else lo = mid;
        }
        double a, coef;
        seg(hi, a, coef);
        const double r = T0 / P.tab_T[hi - 1];
        y0 = P.tab_Ycum[hi - 1] + (fabs(a) > 1e-12 ? coef * (pow(r, a) - 1.0) / a : coef * log(r));
    }
    const double y1 = y0 - dt * (P.gamma - 1.0) * n / KB;
    if (!(y1 > 0.0)) return SNAP * P.tmin;
    if (y1 >= P.tab_Ycum[last]) return P.tab_T[last] + (y1 - P.tab_Ycum[last]) * lam_top;
    int lo = 0, hi = last;  // Ycum[lo] <= y1 < Ycum[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (P.tab_Ycum[mid] > y1) hi = mid;
        else lo = mid;
    }
    double a, coef;
    seg(hi, a, coef);
    const double s = (y1 - P.tab_Ycum[hi - 1]) / coef;
    return P.tab_T[hi - 1] * (fabs(a) > 1e-12 ? pow(1.0 + a * s, 1.0 / a) : exp(s));
}

// FP32: Y's node values in FP64, the segment parts in FP32 (cool_Y).
static __device__ float cool_int_f(const CoolTab &t, float n, float T0, float dt) {
    if (!(T0 >= t.C.tmin)) return T0;
    const int last = t.C.ntable - 1;
    const float lam_top = __builtin_amdgcn_exp2f(t.row[last].L * 3.32192809488736f);
    const double ytop = t.y[last].Ycum;
    double y0;
    if (!(T0 < t.C.tmax)) {
        y0 = ytop + (double)((T0 - t.C.tmax) / lam_top);
    } else {
        const YSplit q = cool_Y(t, T0);
        y0 = q.node + (double)q.part;
    }
    const double y1 = y0 - (double)dt * (double)(t.cgk * n);
    if (!(y1 > 0.0)) return t.snap_T;
    if (!(y1 < ytop)) return t.C.tmax + (float)(y1 - ytop) * lam_top;
    int lo = 0, hi = last;  // Ycum[lo] <= y1 < Ycum[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (t.y[mid].Ycum > y1) hi = mid;
        else lo = mid;
    }
    const float a = 1.f - t.row[hi].slope;
    const float s = (float)(y1 - t.y[hi - 1].Ycum) / t.y[hi].ycoef;
    const float r = fabsf(a) > 1e-4f
                        ? __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(fmaxf(1.f + a * s, 1e-30f)) / a)
                        : __builtin_amdgcn_exp2f(s * 1.4426950408889634f);
    return t.row[hi - 1].T * r;
}

template <class R>
__device__ __forceinline__ double cool_int(const Problem &P, const CoolTab &t, double n, double T,
                                           double dt) {
    if constexpr (sizeof(R) == 8) return cool_int_d(P, n, T, dt);
    else return (double)cool_int_f(t, (float)n, (float)T, (float)dt);
}

// ---------------------------------------------------------------------------
// Per-cell state and the cooling pass's bookkeeping.
constexpr float COOL_EPS = 0.01f;
constexpr int COOL_MAXSUB = 100000;

struct CellState {
    double rho, ek, p, nd, T;
    float rinv;  // 1/rho (FP32 state)
};

// Primitive state of one cell for cooling from its conserved values u (as
// stored): FP64 arithmetic for an FP64 state; for an FP32 state the guarded
// fast path of calc_prims (pressure recovery in FP64 with a refined
// reciprocal, the rest in FP32).
template <class R>
__device__ __forceinline__ CellState cell_state_vals(const Problem &P, const double u[5]) {
    CellState s;
    s.rho = u[0];
    if constexpr (sizeof(R) == 8) {
        s.ek = 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / s.rho;
        s.p = (P.gamma - 1.0) * (u[4] - s.ek);
        if (!(s.p > 1e-20)) s.p = 1e-20;
        s.nd = s.rho / (P.mu0 * AMU);
        s.T = s.p * P.mu0 * AMU / (s.rho * KB);
        s.rinv = 0.f;
    } else {
        s.rinv = __builtin_amdgcn_rcpf((float)u[0]);
        s.ek = 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * inv64((float)u[0]);
        s.p = (P.gamma - 1.0) * (u[4] - s.ek);
        if (!(s.p > 1e-20)) s.p = 1e-20;
        s.nd = (double)((float)u[0] * P.f_inv_mu_amu);
        s.T = (double)((float)s.p * P.f_cT * s.rinv);
    }
    return s;
}

// The same for the cell at element offset c of the SoA array U (u receives
// its values).
template <class R>
__device__ __forceinline__ CellState cell_state(const Problem &P, const R *__restrict__ U,
                                                long long nc, long long c, double u[5]) {
#pragma unroll
    for (int e = 0; e < 5; ++e) u[e] = (double)U[e * nc + c];
    return cell_state_vals<R>(P, u);
}

template <class R>
__device__ __forceinline__ double integrate(const Problem &P, const CoolTab &t, const CellState &s,
                                            double dt, int max_sub) {
    if constexpr (sizeof(R) == 8) return cool_sub_d(P, s.nd, s.T, dt, COOL_EPS, max_sub);
    else return cool_sub_f(t, (float)s.nd, (float)s.T, (float)dt, COOL_EPS, max_sub);
}

// A cell cooled to Tn at fixed density and momentum: its new total energy
// (returned) and its CFL maxima.
template <class R>
__device__ __forceinline__ double finish_vals(const Problem &P, const double u[5],
                                              const CellState &s, double Tn, double &mx,
                                              double &my, double &mz) {
    if constexpr (sizeof(R) == 8) {
        const double pn = s.p * (Tn / s.T);
        const double cs = sqrt(P.gamma * pn / s.rho);
        mx = fmax(mx, fabs(u[1] / s.rho) + cs);
        my = fmax(my, fabs(u[2] / s.rho) + cs);
        mz = fmax(mz, fabs(u[3] / s.rho) + cs);
        return s.ek + pn / (P.gamma - 1.0);
    } else {
        const float pn = (float)s.p * ((float)Tn * __builtin_amdgcn_rcpf((float)s.T));
        const float cs = __builtin_amdgcn_sqrtf((float)P.gamma * pn * s.rinv);
        mx = fmax(mx, (double)(fabsf((float)u[1] * s.rinv) + cs));
        my = fmax(my, (double)(fabsf((float)u[2] * s.rinv) + cs));
        mz = fmax(mz, (double)(fabsf((float)u[3] * s.rinv) + cs));
        return s.ek + (double)pn * (double)P.f_inv_gm1;
    }
}

// finish_vals for the cell at offset c of U, storing the new energy.
template <class R>
__device__ __forceinline__ void finish_cell(const Problem &P, R *__restrict__ U, long long nc,
                                            long long c, const double u[5], const CellState &s,
                                            double Tn, double &mx, double &my, double &mz) {
    U[4 * nc + c] = (R)finish_vals<R>(P, u, s, Tn, mx, my, mz);
}

// CFL maxima of a cell left as it is (below T_min)
template <class R>
__device__ __forceinline__ void cell_maxima(const Problem &P, const double u[5],
                                            const CellState &s, double &mx, double &my,
                                            double &mz) {
    if constexpr (sizeof(R) == 8) {
        const double cs = sqrt(P.gamma * s.p / s.rho);
        mx = fmax(mx, fabs(u[1] / s.rho) + cs);
        my = fmax(my, fabs(u[2] / s.rho) + cs);
        mz = fmax(mz, fabs(u[3] / s.rho) + cs);
    } else {
        const float cs = __builtin_amdgcn_sqrtf((float)P.gamma * (float)s.p * s.rinv);
        mx = fmax(mx, (double)(fabsf((float)u[1] * s.rinv) + cs));
        my = fmax(my, (double)(fabsf((float)u[2] * s.rinv) + cs));
        mz = fmax(mz, (double)(fabsf((float)u[3] * s.rinv) + cs));
    }
}

// log10 through the hardware log2 (FP32 inputs of the surrogate and its check)
__device__ __forceinline__ float lg10(float x) {
    return __builtin_amdgcn_logf(x) * 0.301029995663981f;
}


}  // namespace tlq
