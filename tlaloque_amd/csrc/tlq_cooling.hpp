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

// Table segment of temperature T (first row with T_i > T, as find_coef),
// clamped to the last segment for T >= T_max.
__device__ __forceinline__ int table_row(const Problem &P, float temp, float lg) {
    if (!(temp < P.f_tmax)) return P.ntable - 1;
    int hi;
    if (P.tab_uniform) {
        hi = (int)((lg - P.f_lgT0) * P.f_inv_dlg) + 1;
        hi = hi < 1 ? 1 : (hi > P.ntable - 1 ? P.ntable - 1 : hi);
        if (!(P.tab_Tf[hi] > temp) && hi < P.ntable - 1) ++hi;
        if (hi > 1 && P.tab_Tf[hi - 1] > temp) --hi;
    } else {
        int lo = 0;
        hi = P.ntable - 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (P.tab_Tf[mid] > temp) hi = mid;
            else lo = mid;
        }
    }
    return hi;
}

// Lambda(T) from the table in FP32 (find_coef's interpolation).
__device__ __forceinline__ float lambda_f(const Problem &P, float temp) {
    const float lg = __builtin_amdgcn_logf(temp) * 0.301029995663981f;
    if (!(temp < P.f_tmax)) return __builtin_amdgcn_exp2f(P.tab_Lf[P.ntable - 1] * 3.32192809488736f);
    const int hi = table_row(P, temp, lg);
    const float lgl = P.tab_Lf[hi - 1] + P.tab_slopef[hi] * (lg - P.tab_lgTf[hi - 1]);
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
__device__ __forceinline__ YSplit cool_Y(const Problem &P, float temp) {
    if (!(temp > P.f_tmin)) return {0.0, 0.f};
    const float l2 = __builtin_amdgcn_logf(temp);
    const int hi = table_row(P, temp, l2 * 0.301029995663981f);
    const float a = 1.f - P.tab_slopef[hi];
    const float lr = l2 - __builtin_amdgcn_logf(P.tab_Tf[hi - 1]);  // log2(T/T_{hi-1})
    const float seg = fabsf(a) > 1e-4f ? (__builtin_amdgcn_exp2f(a * lr) - 1.f) / a
                                       : lr * 0.6931471805599453f;
    return {P.tab_Ycum[hi - 1], P.tab_ycoef[hi] * seg};
}
__device__ __forceinline__ float cool_dY(const Problem &P, float T0, float T1) {
    const YSplit a = cool_Y(P, T0), b = cool_Y(P, T1);
    return (float)(a.node - b.node) + (a.part - b.part);
}

static __device__ float cool_sub_f(const Problem &P, float n, float T, float dt, float eps, int max_sub) {
    const float cgk = (float)((P.gamma - 1.0) / KB);
    const float floor_f = P.f_floor_f;
    float left = dt;
    for (int s = 0; left > 0.f && T >= P.f_tmin && s < max_sub; ++s) {
        const float ce = lambda_f(P, T) * n * cgk * __builtin_amdgcn_rcpf(T);
        const float h = fminf(left, eps * __builtin_amdgcn_rcpf(ce));
        float f = __builtin_amdgcn_exp2f(-1.4426950408889634f * ce * h);
        if (f < floor_f) f = floor_f;
        if (T * f < 10.f) f = 10.f * __builtin_amdgcn_rcpf(T);
        T = T * f;
        left -= h;
    }
    return T;
}

// ---------------------------------------------------------------------------
// Per-cell state and the cooling pass's bookkeeping.
constexpr float COOL_EPS = 0.01f;
constexpr int COOL_MAXSUB = 100000;
// A surrogate answer at the gate is set just below it, in the 1% band where
// the converged integrator's last substep ends (so the cell stays gated).
constexpr double SNAP = 0.995;

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

// dt / t_cool of a cell at T >= T_min
template <class R>
__device__ __forceinline__ double dt_over_tcool(const Problem &P, const CellState &s, double dt) {
    if constexpr (sizeof(R) == 8) {
        return find_coef(P, s.T) * s.nd * s.nd / (s.p / (P.gamma - 1.0)) * dt;
    } else {
        const float ce = lambda_f(P, (float)s.T) * (float)s.nd * (float)s.nd *
                         (float)(P.gamma - 1.0) * __builtin_amdgcn_rcpf((float)s.p);
        return (double)(ce * (float)dt);
    }
}

template <class R>
__device__ __forceinline__ double integrate(const Problem &P, const CellState &s, double dt,
                                            int max_sub) {
    if constexpr (sizeof(R) == 8) return cool_sub_d(P, s.nd, s.T, dt, COOL_EPS, max_sub);
    else return cool_sub_f(P, (float)s.nd, (float)s.T, (float)dt, COOL_EPS, max_sub);
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
