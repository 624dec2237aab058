// tlq_surrogate.hpp -- the learned radiative-cooling surrogate.
//
// The network maps (log10 n, log10 T, log10(dt/t_cool)) to g = r / r1 for one
// cooling step, r = log10(T_new/T) and r1 = -log10(e) d / (1 + d), d =
// dt/t_cool (r1 is r to first order in d and bounded for large d; learning
// the ratio makes the network's error relative to the change of T, which is
// what the per-cell check measures): an MLP 3 -> 64 -> 64 -> 1 with GELU (tanh form),
// bf16 weights and FP32 accumulation.  Its training targets come from the
// reference's cooling formula (apply_cooling, cooling.f90:67-112) integrated
// to convergence: the formula is applied in substeps of at most eps t_cool
// with Lambda re-read from the table every substep (one substep is exactly
// apply_cooling).  Layer 2 -- 4096 of the network's 4352 multiply-adds --
// runs on the MFMA units (v_mfma_f32_16x16x16_bf16), in training (forward
// and both backward products) and in inference.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "tlq_device.hpp"

namespace tlq {
namespace mlp {

constexpr int H = 64;
constexpr int OW1 = 0, OB1 = 3 * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H,
              OB3 = OW3 + H, NP = OB3 + 1;  // 4481 parameters
// input normalisation: x_n = (x - C) / S maps the training box onto [-1, 1]
// training box: n in [1e-3, 1e4] cm^-3, T in [1e4, 1e9] K (the table's range:
// nothing cools below T_min) and dt/t_cool in [eps, 100] (below eps one
// substep is the converged answer)
constexpr float C0 = 0.5f, C1 = 6.5f, C2 = 0.0f, S0 = 3.5f, S1 = 2.5f, S2 = 2.0f;
constexpr float X0LO = -3.f, X0HI = 4.f, X1LO = 4.f, X1HI = 9.f, X2LO = -2.f, X2HI = 2.f;

// r1(x2) = -log10(e) d / (1 + d), d = 10^x2: the scale of the network's output
__host__ __device__ inline double first_order(double x2) {
    const double d = std::pow(10.0, x2);
    return -0.4342944819032518 * d / (1.0 + d);
}

}  // namespace mlp

// Launchers (tlq_surrogate.hip).
hipError_t sur_generate(hipStream_t s, const Problem *P, long long n, unsigned long long seed,
                        float *X, float *Y, double eps, int max_sub);
hipError_t sur_grad(hipStream_t s, const unsigned short *thq, const float *X, const float *Y,
                    long long off, int B, float *partial, int nblocks);
hipError_t sur_adam(hipStream_t s, const float *partial, int nblocks, float *theta,
                    unsigned short *thq, float *m, float *v, int t, float lr, float *loss_out);
hipError_t sur_quantize(hipStream_t s, const float *theta, unsigned short *thq);
hipError_t sur_eval(hipStream_t s, const unsigned short *thq, const float *X, long long n,
                    float *out);

// Cooling after the fused corrector in modes 2 (surrogate with per-cell
// fallback) and 3 (the converged integrator): k_stage2 queues the cells with
// dt > 0.01 t_cool in one segment of C slots per stage-2 tile (nseg
// segments): from the front the cells for the network (mode 2, dt >=
// net_min_dtc t_cool), from the back those for the integrator; seg[b] and
// seg[nseg + b] hold the two counts of segment b.  This pass scans them into
// off (network: off[0..nseg], integrator: off[nseg+1..2 nseg+1]; the last of
// each is its total), runs the network on its cells (mode 2; rejected cells
// go to fqueue, *fcount of them) and the integrator on its cells and the
// fallbacks, and adds those cells' CFL maxima.
struct StepState;
template <class R>
hipError_t sur_cool_pass(hipStream_t s, const Problem *P, int nseg, long long C, R *U,
                         const unsigned short *thq, StepState *st, int parity, unsigned *queue,
                         const unsigned *seg, unsigned *off, unsigned *fqueue, unsigned *fcount,
                         int mode, int ntable);

}  // namespace tlq
