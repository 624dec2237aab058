// tlq_surrogate.hip -- kernels of the learned cooling surrogate (see
// tlq_surrogate.hpp): training data, MFMA training gradient, Adam, inference,
// and the cooling pass of the fused step.
#include "tlq_cooling.hpp"
#include "tlq_kernels.hpp"
#include "tlq_surrogate.hpp"

namespace tlq {

using namespace mlp;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) != 0x7f800000u) u = u + 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// GELU, tanh form: 0.5 x (1 + tanh(u/2)) = x sigmoid(u), u = 2 sqrt(2/pi) (x + 0.044715 x^3);
// exp(-u) = 2^(x (A + B x^2)) with the constants folded.
__device__ __forceinline__ float gelu(float x) {
    constexpr float A = -1.5957691216057308f * 1.4426950408889634f;
    constexpr float B = A * 0.044715f;
    const float e = __builtin_amdgcn_exp2f(x * __builtin_fmaf(B, x * x, A));
    return x * __builtin_amdgcn_rcpf(1.f + e);
}

// Four floats to bf16, rounded to nearest even (v_cvt_pk_bf16_f32).
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 b16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ short2 bf16_pair(float a, float b) {
    const b16x2 v = __builtin_convertvector(f32x2{a, b}, b16x2);
    return __builtin_bit_cast(short2, v);
}
__device__ __forceinline__ float gelu_d(float x) {
    const float x2 = x * x;
    const float u = 1.5957691216057308f * (x + 0.044715f * x2 * x);
    const float s = __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.4426950408889634f * u));
    const float du = 1.5957691216057308f * (1.f + 0.134145f * x2);
    return s + x * s * (1.f - s) * du;
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Training data: uniform samples of the training box, targets g = r / r1 from
// the FP64 converged integrator.  Counter-based generator (splitmix64 of
// seed, index).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double unif(unsigned long long seed, long long i, int c) {
    return (double)(mix64(seed ^ mix64((unsigned long long)i * 4 + c)) >> 11) * (1.0 / 9007199254740992.0);
}

__global__ void k_gen_samples(const Problem *__restrict__ Pp, long long n, unsigned long long seed,
                              float *__restrict__ X, float *__restrict__ Y, double eps,
                              int max_sub) {
    const Problem &P = *Pp;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const double x0 = X0LO + (X0HI - X0LO) * unif(seed, t, 0);
        const double x1 = X1LO + (X1HI - X1LO) * unif(seed, t, 1);
        const double x2 = X2LO + (X2HI - X2LO) * unif(seed, t, 2);
        const double nd = pow(10.0, x0), T = pow(10.0, x1);
        double r = 0.0;
        if (T >= P.tmin) {
            const double eth = nd * KB * T / (P.gamma - 1.0);
            const double ce = find_coef(P, T) * nd * nd / eth;
            const double dt = pow(10.0, x2) / ce;
            r = log10(cool_sub_d(P, nd, T, dt, eps, max_sub) / T) / first_order(x2);
        }
        X[3 * t + 0] = (float)x0;
        X[3 * t + 1] = (float)x1;
        X[3 * t + 2] = (float)x2;
        Y[t] = (float)r;
    }
}

// ---------------------------------------------------------------------------
// Network evaluation by one wavefront on 64 samples (one per lane).

struct MlpShared {
    unsigned short h1[64][72];   // layer-1 activations, bf16 [sample][unit]
    unsigned short h1t[64][72];  // transposed [unit][sample] (training only)
    unsigned short d[64][72];    // dz2 bf16 [sample][unit] (training only)
    unsigned short dt[64][72];   // transposed (training only)
    float f[64][65];             // FP32 staging [sample][unit]
    float x[64][4];              // normalised inputs
    float w1[3 * H], b1[H], b2[H], w3[H];
    float dout[64];
};
// Fragments of W2: lane holds W2[j][i..i+3] with j = 16 nt + lane%16,
// i = 16 kt + 4 (lane/16) -- the B operand of z2 = h1 W2^T (training) and the
// A operand of z2^T = W2 h1^T (inference, mlp_wave).
__device__ __forceinline__ void load_w2_fwd(const unsigned short *__restrict__ thq, int lane,
                                            bf16x4 w[4][4]) {
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int j = nt * 16 + (lane & 15), i = kt * 16 + 4 * (lane >> 4);
            const unsigned short *p = thq + OW2 + j * H + i;
            w[kt][nt] = bf16x4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
        }
}

// layer 1 for the lane's own sample -> h1 (bf16) and its transpose in LDS
__device__ __forceinline__ void layer1(MlpShared &sh, int lane, const float xn[3]) {
    for (int i = 0; i < H; ++i) {
        const float z = sh.b1[i] + sh.w1[3 * i] * xn[0] + sh.w1[3 * i + 1] * xn[1] +
                        sh.w1[3 * i + 2] * xn[2];
        const unsigned short q = f2bf(gelu(z));
        sh.h1[lane][i] = q;
        sh.h1t[i][lane] = q;
    }
}

// layer 2 (MFMA): f[sample][j] = sum_i h1[sample][i] W2[j][i]
__device__ __forceinline__ void layer2(MlpShared &sh, int lane, const bf16x4 w[4][4]) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                const bf16x4 a =
                    *reinterpret_cast<const bf16x4 *>(&sh.h1[mt * 16 + (lane & 15)][kt * 16 + 4 * (lane >> 4)]);
                acc = mfma16(a, w[kt][nt], acc);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) sh.f[mt * 16 + 4 * (lane >> 4) + v][nt * 16 + (lane & 15)] = acc[v];
        }
}

__device__ __forceinline__ void load_small(MlpShared &sh, const unsigned short *__restrict__ thq,
                                           int lane) {
    for (int q = lane; q < 3 * H; q += 64) sh.w1[q] = bf2f(thq[OW1 + q]);
    sh.b1[lane] = bf2f(thq[OB1 + lane]);
    sh.b2[lane] = bf2f(thq[OB2 + lane]);
    sh.w3[lane] = bf2f(thq[OW3 + lane]);
}

__device__ __forceinline__ void normalise(const float x[3], float xn[3]) {
    xn[0] = (x[0] - C0) * (1.f / S0);
    xn[1] = (x[1] - C1) * (1.f / S1);
    xn[2] = (x[2] - C2) * (1.f / S2);
}

// FP32 copies of the small parameter blocks (bf16-rounded), for inference.
struct MlpSmall {
    float w1[3 * H], b1[H], b2[H], w3[H];
};

__device__ __forceinline__ void load_small_f(MlpSmall &sm, const unsigned short *__restrict__ thq,
                                             int tid, int nthreads) {
    for (int q = tid; q < 3 * H; q += nthreads) sm.w1[q] = bf2f(thq[OW1 + q]);
    for (int q = tid; q < H; q += nthreads) {
        sm.b1[q] = bf2f(thq[OB1 + q]);
        sm.b2[q] = bf2f(thq[OB2 + q]);
        sm.w3[q] = bf2f(thq[OW3 + q]);
    }
}

// The network on the 64 samples of a wavefront, one per lane (xn: this
// lane's normalised inputs); returns this lane's output.  Layer 2 runs
// transposed, z2^T = W2 h1^T, so that the layer-1 activations each lane
// computes -- units 16 kt + 4 (lane/16) + v of sample 16 st + lane%16 -- are
// the MFMA B operand as they stand, and the layer-2 outputs -- units
// 16 jt + 4 (lane/16) + v of the same samples -- feed layer 3 in registers:
// no activation goes through LDS.  The four lanes that share a sample hold
// disjoint units of it; two cross-lane sums finish layer 3.  Same roundings
// as the training forward pass (bf16 h1 and weights, FP32 accumulation).
// Every lane of the wavefront must call it (the MFMA and shuffles are
// wavefront-wide).
__device__ __forceinline__ float mlp_wave(const MlpSmall &sm, const bf16x4 w2a[4][4], float b3,
                                          int lane, const float xn[3]) {
    const int q = lane >> 4;
    float r = 0.f;
    // one sample tile at a time keeps only its four activation fragments live
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        // The layer-1 and layer-3 parameters are read from LDS again for
        // every sample tile: hoisted out of this loop they hold about 100
        // registers and the dense pass runs two wavefronts per SIMD instead
        // of three (-4% for the pass).
        asm volatile("" ::: "memory");
        float xs[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) xs[c] = __shfl(xn[c], 16 * st + (lane & 15));
        bf16x4 hb[4];  // [kt]
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            float h[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 16 * kt + 4 * q + v;
                const float z = sm.b1[i] + sm.w1[3 * i] * xs[0] + sm.w1[3 * i + 1] * xs[1] +
                                sm.w1[3 * i + 2] * xs[2];
                h[v] = gelu(z);
            }
            const short2 lo = bf16_pair(h[0], h[1]), hi = bf16_pair(h[2], h[3]);
            hb[kt] = bf16x4{lo.x, lo.y, hi.x, hi.y};
        }
        float o = 0.f;
#pragma unroll
        for (int jt = 0; jt < 4; ++jt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) acc = mfma16(w2a[kt][jt], hb[kt], acc);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = 16 * jt + 4 * q + v;
                o += sm.w3[j] * gelu(acc[v] + sm.b2[j]);
            }
        }
        o += __shfl_xor(o, 16);
        o += __shfl_xor(o, 32);
        if (st == q) r = o;
    }
    return b3 + r;
}

// ---------------------------------------------------------------------------
// Training gradient of the mean squared error over B samples starting at
// `off`: every block (one wavefront) handles 64-sample chunks and writes its
// partial gradient (NP values) and partial loss to partial[block].
__global__ __launch_bounds__(64) void k_mlp_grad(const unsigned short *__restrict__ thq,
                                                 const float *__restrict__ X,
                                                 const float *__restrict__ Y, long long off, int B,
                                                 float *__restrict__ partial) {
    __shared__ MlpShared sh;
    const int lane = threadIdx.x;
    load_small(sh, thq, lane);
    const float b3 = bf2f(thq[OB3]);
    bf16x4 wf[4][4], wb[4][4];
    load_w2_fwd(thq, lane, wf);
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)  // B fragments of W2 for dh1 = dz2 W2
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int i = nt * 16 + (lane & 15), j = kt * 16 + 4 * (lane >> 4);
            wb[kt][nt] = bf16x4{(short)thq[OW2 + (j + 0) * H + i], (short)thq[OW2 + (j + 1) * H + i],
                                (short)thq[OW2 + (j + 2) * H + i], (short)thq[OW2 + (j + 3) * H + i]};
        }
    f32x4 gw2[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) gw2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    float gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, gb1 = 0.f, gw1[3] = {0.f, 0.f, 0.f}, loss = 0.f;
    const float invB = 1.f / (float)B;
    __syncthreads();
    const int nchunks = (B + 63) / 64;
    for (int ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
        const int b = ch * 64 + lane;
        const bool valid = b < B;
        float x[3] = {0.f, 0.f, 0.f}, y = 0.f;
        if (valid) {
            x[0] = X[3 * (off + b)];
            x[1] = X[3 * (off + b) + 1];
            x[2] = X[3 * (off + b) + 2];
            y = Y[off + b];
        }
        float xn[3];
        normalise(x, xn);
        sh.x[lane][0] = xn[0];
        sh.x[lane][1] = xn[1];
        sh.x[lane][2] = xn[2];
        layer1(sh, lane, xn);
        __syncthreads();
        layer2(sh, lane, wf);
        __syncthreads();
        float o = b3;
        for (int j = 0; j < H; ++j) o += sh.w3[j] * gelu(sh.f[lane][j] + sh.b2[j]);
        const float d = valid ? o - y : 0.f;
        loss += d * d * invB;
        const float dout = 2.f * d * invB;
        sh.dout[lane] = dout;
        for (int j = 0; j < H; ++j) {
            const float z2 = sh.f[lane][j] + sh.b2[j];
            const unsigned short q = f2bf(dout * sh.w3[j] * gelu_d(z2));
            sh.d[lane][j] = q;
            sh.dt[j][lane] = q;
            sh.f[lane][j] = dout * gelu(z2);
        }
        __syncthreads();
        {  // column sums (lane = unit j): b2, W3, b3 gradients
            float s1 = 0.f, s2 = 0.f, s3 = 0.f;
            for (int q = 0; q < 64; ++q) {
                s1 += bf2f(sh.dt[lane][q]);
                s2 += sh.f[q][lane];
                s3 += sh.dout[q];
            }
            gb2 += s1;
            gw3 += s2;
            gb3 += s3;
        }
        // dW2[j][i] += sum_b dz2[b][j] h1[b][i]
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int kt = 0; kt < 4; ++kt) {
                    const bf16x4 a =
                        *reinterpret_cast<const bf16x4 *>(&sh.dt[mt * 16 + (lane & 15)][kt * 16 + 4 * (lane >> 4)]);
                    const bf16x4 bb =
                        *reinterpret_cast<const bf16x4 *>(&sh.h1t[nt * 16 + (lane & 15)][kt * 16 + 4 * (lane >> 4)]);
                    gw2[mt][nt] = mfma16(a, bb, gw2[mt][nt]);
                }
        __syncthreads();
        // dh1[b][i] = sum_j dz2[b][j] W2[j][i]
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kt = 0; kt < 4; ++kt) {
                    const bf16x4 a =
                        *reinterpret_cast<const bf16x4 *>(&sh.d[mt * 16 + (lane & 15)][kt * 16 + 4 * (lane >> 4)]);
                    acc = mfma16(a, wb[kt][nt], acc);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) sh.f[mt * 16 + 4 * (lane >> 4) + v][nt * 16 + (lane & 15)] = acc[v];
            }
        __syncthreads();
        for (int i = 0; i < H; ++i) {  // dz1 for the lane's own sample
            const float z = sh.b1[i] + sh.w1[3 * i] * xn[0] + sh.w1[3 * i + 1] * xn[1] +
                            sh.w1[3 * i + 2] * xn[2];
            sh.f[lane][i] = sh.f[lane][i] * gelu_d(z);
        }
        __syncthreads();
        {  // column sums (lane = unit i): b1, W1 gradients
            float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
            for (int q = 0; q < 64; ++q) {
                const float dz = sh.f[q][lane];
                s0 += dz;
                s1 += dz * sh.x[q][0];
                s2 += dz * sh.x[q][1];
                s3 += dz * sh.x[q][2];
            }
            gb1 += s0;
            gw1[0] += s1;
            gw1[1] += s2;
            gw1[2] += s3;
        }
        __syncthreads();
    }
    float *g = partial + (long long)blockIdx.x * (NP + 1);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = mt * 16 + 4 * (lane >> 4) + v, i = nt * 16 + (lane & 15);
                g[OW2 + j * H + i] = gw2[mt][nt][v];
            }
    g[OB2 + lane] = gb2;
    g[OW3 + lane] = gw3;
    g[OB1 + lane] = gb1;
    g[OW1 + 3 * lane] = gw1[0];
    g[OW1 + 3 * lane + 1] = gw1[1];
    g[OW1 + 3 * lane + 2] = gw1[2];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
    if (lane == 0) {
        g[OB3] = gb3;
        g[NP] = loss;
    }
}

// Sums the block partials (fixed order: deterministic) and applies one Adam
// step to the FP32 master parameters; refreshes their bf16 copy.
__global__ void k_adam(const float *__restrict__ partial, int nblocks, float *__restrict__ theta,
                       unsigned short *__restrict__ thq, float *__restrict__ m,
                       float *__restrict__ v, int t, float lr, float *__restrict__ loss_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NP) return;
    float g = 0.f;
    for (int b = 0; b < nblocks; ++b) g += partial[(long long)b * (NP + 1) + p];
    if (p == NP) {
        if (loss_out) *loss_out = g;
        return;
    }
    if (!theta) {  // gradient only (grad check): leave it in m
        m[p] = g;
        return;
    }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.f - powf(b1, (float)t)), vh = vv / (1.f - powf(b2, (float)t));
    const float th = theta[p] - lr * mh / (sqrtf(vh) + eps);
    theta[p] = th;
    thq[p] = f2bf(th);
}

__global__ void k_quantize(const float *__restrict__ theta, unsigned short *__restrict__ thq) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NP) thq[p] = f2bf(theta[p]);
}

// Network output for n input triples (64 per wavefront).
__global__ __launch_bounds__(64) void k_mlp_eval(const unsigned short *__restrict__ thq,
                                                 const float *__restrict__ X, long long n,
                                                 float *__restrict__ out) {
    __shared__ MlpSmall sm;
    const int lane = threadIdx.x;
    load_small_f(sm, thq, lane, 64);
    const float b3 = bf2f(thq[OB3]);
    bf16x4 w2a[4][4];
    load_w2_fwd(thq, lane, w2a);
    __syncthreads();
    for (long long base = (long long)blockIdx.x * 64; base < n; base += (long long)gridDim.x * 64) {
        const long long s = base + lane;
        float x[3] = {0.f, 0.f, 0.f};
        if (s < n) {
            x[0] = X[3 * s];
            x[1] = X[3 * s + 1];
            x[2] = X[3 * s + 2];
        }
        float xn[3];
        normalise(x, xn);
        const float o = mlp_wave(sm, w2a, b3, lane, xn);
        if (s < n) out[s] = o;
    }
}

// ---------------------------------------------------------------------------
// Cooling of the fused step in modes 2 and 3.
//
// k_stage2 classifies every interior cell as it writes it.  Cells below T_min
// are left alone; cells with dt <= eps t_cool take one substep (all the
// converged integrator would take); their CFL maxima are formed there.  Every
// other cell goes to the queue segment of its stage-2 block (one LDS atomic
// per wavefront).
//
// k_cool_dense then runs the queued cells in full wavefronts: mode 3
// integrates them to convergence; mode 2 evaluates the surrogate and keeps
// its answer T1 when
//   * the inputs lie in the training box,
//   * 10 K <= T1 <= T0 (cooling only lowers T, never below the floor), and
//   * the exact cooling time from T0 to T1, t = k_B (Y(T0) - Y(T1)) / ((gamma-1) n),
//     is within 0.05 min(dt, t_cool(T1)) of dt, so that ln T1 is within 5% of
//     its change (at most 0.05) of the exact solution -- or, for T1 within 10%
//     of the gate T_min, the gas reaches T_min within that time, and the
//     answer is 0.995 T_min (the converged integrator stops within 1% below
//     the gate);
// other cells go to a fallback queue that a mode-3 dense pass integrates.
// Then the CFL maxima of the cells finished here.

// Exclusive prefix sums of the per-tile queue counts: off[b] = cnt[0] + ... +
// cnt[b-1], off[n] = the total.  Block q scans the q-th list of n counts into
// the q-th list of n+1 offsets (the network's and the integrator's queues, in
// one launch); block 0 also zeroes the fallback counter.
__global__ __launch_bounds__(1024) void k_queue_scan(const unsigned *__restrict__ cnts,
                                                     unsigned *__restrict__ offs, int n,
                                                     unsigned *__restrict__ fcount) {
    __shared__ unsigned part[1024];
    const unsigned *const cnt = cnts + (long long)blockIdx.x * n;
    unsigned *const off = offs + (long long)blockIdx.x * (n + 1);
    const int t = threadIdx.x, per = (n + 1023) / 1024;
    const int b0 = min(t * per, n), b1 = min(b0 + per, n);
    unsigned sum = 0;
    for (int b = b0; b < b1; ++b) sum += cnt[b];
    part[t] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // inclusive scan of the partial sums
        const unsigned v = t >= d ? part[t - d] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    unsigned run = part[t] - sum;
    for (int b = b0; b < b1; ++b) {
        off[b] = run;
        run += cnt[b];
    }
    if (t == 1023) off[n] = part[1023];
    if (t == 0 && blockIdx.x == 0) *fcount = 0;
}

// Segment b of the queue holding position p: off[b] <= p < off[b+1] (off
// non-decreasing; empty segments repeat a value).
__device__ __forceinline__ int queue_segment(const unsigned *__restrict__ off, int nseg,
                                             unsigned p) {
    int lo = 0, hi = nseg;  // off[lo] <= p < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= p) lo = mid;
        else hi = mid;
    }
    return lo;
}

// The dense passes: wavefronts take 64 consecutive queue positions at a
// time, grid-stride, so the work is spread evenly whatever the tiles' counts.
// MODE 2, the network: the cells k_stage2 put at the front of each tile's
// segment (offn: their scanned counts); rejected cells go to the fallback
// queue through a per-wavefront LDS buffer, flushed with one global atomic
// per FBUF_FLUSH cells (a global atomic per wavefront serialises on the
// counter when a tenth of the cells fall back).  MODE 3, the converged
// integrator: the cells at the back of each segment (offi), then the
// *fcount fallbacks.
constexpr int FBUF = 1024, FBUF_FLUSH = FBUF - 64;

// Copies a wavefront's n buffered fallback cells to the fallback queue.
__device__ __forceinline__ void flush_fallbacks(const unsigned *buf, unsigned n, int lane,
                                                unsigned *__restrict__ fqueue,
                                                unsigned *__restrict__ fcount) {
    __builtin_amdgcn_wave_barrier();  // the wavefront's LDS writes of buf come first
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(fcount, n);
    base = __shfl(base, 0);
    for (unsigned q = lane; q < n; q += 64) fqueue[base + q] = buf[q];
}

template <class R, int MODE>
__global__ __launch_bounds__(128) void k_cool_dense(const Problem *__restrict__ Pp,
                                                    R *__restrict__ U,
                                                    const unsigned short *__restrict__ thq,
                                                    StepState *__restrict__ st, int parity,
                                                    const unsigned *__restrict__ queue,
                                                    const unsigned *__restrict__ offn,
                                                    const unsigned *__restrict__ offi, int nseg,
                                                    long long C, unsigned *__restrict__ fqueue,
                                                    unsigned *__restrict__ fcount, int with_y) {
    __shared__ MlpSmall sm;
    __shared__ unsigned s_fbuf[MODE == 2 ? 2 : 1][MODE == 2 ? FBUF : 1];
    // The FP32 cooling table in LDS (cool_tab_bytes(ntable, with_y) bytes of
    // dynamic LDS): for the network's inputs and check (MODE 2, either state
    // precision) and the FP32 integrators; none in the FP64 integrator pass,
    // whose arithmetic reads the FP64 table.
    extern __shared__ __align__(16) unsigned char s_tab[];
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    TabRow *const trow = reinterpret_cast<TabRow *>(s_tab);
    YRow *const ty = with_y ? reinterpret_cast<YRow *>(s_tab + P.ntable * sizeof(TabRow)) : nullptr;
    const unsigned nseg_cells = MODE == 2 ? offn[nseg] : offi[nseg];
    const unsigned nq = MODE == 2 ? nseg_cells : nseg_cells + *fcount;
    // The grid is sized for the largest queue: a block past this one's end has
    // no cell (its first wavefront's first position is blockIdx.x * 128).
    if (blockIdx.x * 128u >= nq) return;
    if (sizeof(R) == 4 || MODE == 2) stage_cool_tab(P, trow, ty, threadIdx.x, blockDim.x);  // visible after the barrier below
    const CoolTab tab = cool_tab(P, trow, ty);
    const double dt = dt_from_maxima(P, st->maxima[parity]);
    unsigned nfb = 0;  // this wavefront's buffered fallbacks
    bf16x4 w2a[4][4];
    float b3 = 0.f;
    // T_min / Lambda(T_min): k_B T_min / ((gamma-1) n Lambda) is the cooling time at the gate
    if (MODE == 2) {
        load_small_f(sm, thq, threadIdx.x, blockDim.x);
        b3 = bf2f(thq[OB3]);
        load_w2_fwd(thq, lane, w2a);
    }
    __syncthreads();
    const float gate_tc = MODE == 2 ? P.f_tmin / lambda_f(tab, P.f_tmin) : 0.f;
    double mx = 0.0, my = 0.0, mz = 0.0;
    const unsigned stride = gridDim.x * 128u;
    for (unsigned base = (blockIdx.x * 2u + wv) * 64u; base < nq; base += stride) {
        const unsigned qi = base + lane;
        const bool valid = qi < nq;
        const unsigned pos = valid ? qi : base;
        long long c;
        if (MODE == 2) {
            const int b = queue_segment(offn, nseg, pos);
            c = queue[(long long)b * C + (pos - offn[b])];
        } else if (pos < nseg_cells) {
            const int b = queue_segment(offi, nseg, pos);
            c = queue[(long long)b * C + (C - 1 - (pos - offi[b]))];
        } else {
            c = fqueue[pos - nseg_cells];
        }
        double u[5];
        const CellState s = cell_state(P, U, g.nc, c, u);
        bool need_exact = valid;
        double Tn = s.T;
        if constexpr (MODE == 2) {
            const float T0 = (float)s.T;
            const float kn = (float)(KB / (P.gamma - 1.0)) / (float)s.nd;
            const float fdt = (float)dt;
            // A cell whose exact cooling time to the gate (closed form) fits in
            // dt ends at the gate, where the integrator stops: the answer the
            // check below accepts when the network predicts the gate, given
            // without the network.
            const bool gate =
                valid && kn * cool_dY(tab, T0, P.f_tmin) <= fdt + 0.05f * fminf(fdt, kn * gate_tc);
            const bool net = valid && !gate;
            bool ok = false, snap = false;
            float T1 = T0;
            if (__ballot(net)) {  // wavefront-uniform: the network runs unless every cell is gated
                const float lam = lambda_f(tab, T0);
                const float ce = lam * (float)s.nd * (float)s.nd * (float)(P.gamma - 1.0) *
                                 __builtin_amdgcn_rcpf((float)s.p);
                float x[3] = {lg10((float)s.nd), lg10(T0), lg10(ce * fdt)};
                float xn[3];
                normalise(x, xn);
                const float o = mlp_wave(sm, w2a, b3, lane, xn);
                const bool inbox = x[0] >= X0LO && x[0] <= X0HI && x[1] >= X1LO &&
                                   x[1] <= X1HI && x[2] <= X2HI;
                const float d = __builtin_amdgcn_exp2f(x[2] * 3.32192809488736f);
                const float rh = -0.4342944819032518f * d / (1.f + d) * o;  // log10(T1/T0)
                const float rmin = lg10(10.f * __builtin_amdgcn_rcpf(T0));
                ok = net && inbox && rh <= 1e-3f && rh >= rmin - 1e-3f;
                T1 = T0 * __builtin_amdgcn_exp2f(fminf(rh, 0.f) * 3.32192809488736f);
                if (ok) {
                    // a time error dt' shifts ln T1 by dt'/t_cool(T1): allow 5% of
                    // the change in ln T, at most 0.05
                    const float Te = fmaxf(T1, P.f_tmin);
                    const float tol = 0.05f * fminf(fdt, kn * Te / lambda_f(tab, Te));
                    if (T1 <= 1.1f * P.f_tmin) {
                        // the network says the gas reaches the gate: accepted when
                        // the exact cooling time to T_min fits in dt, and the answer
                        // is the gate itself (where the integrator stops)
                        ok = kn * cool_dY(tab, T0, P.f_tmin) <= fdt + tol;
                        snap = true;
                    } else {
                        ok = fabsf(kn * cool_dY(tab, T0, T1) - fdt) <= tol;
                    }
                }
            }
            if (gate || ok) {
                Tn = gate || snap ? (double)SNAP * P.tmin : (double)T1;
                need_exact = false;
            }
            // fallbacks go to the integrator's pass: a lane that needs the
            // integrator does not hold up the other 63
            const unsigned long long fm = __ballot(need_exact);
            if (fm) {
                if (need_exact) s_fbuf[wv][nfb + __popcll(fm & ((1ull << lane) - 1ull))] = (unsigned)c;
                nfb += (unsigned)__popcll(fm);
                if (nfb > FBUF_FLUSH) {  // the buffer is this wavefront's alone: no barrier
                    flush_fallbacks(s_fbuf[wv], nfb, lane, fqueue, fcount);
                    nfb = 0;
                }
            }
            if (valid && !need_exact) finish_cell(P, U, g.nc, c, u, s, Tn, mx, my, mz);
        } else {
            // mode 3: the substepped integrator; mode 4 and the surrogate's
            // integrator cells and fallbacks: the exact integration
            if (valid)
                finish_cell(P, U, g.nc, c, u, s,
                            P.cooling == 3 ? integrate<R>(P, tab, s, dt, COOL_MAXSUB)
                                           : cool_int<R>(P, tab, s.nd, s.T, dt),
                            mx, my, mz);
        }
    }
    if (MODE == 2 && nfb) flush_fallbacks(s_fbuf[wv], nfb, lane, fqueue, fcount);
    block_max3<128>(mx, my, mz, st->maxima[parity ^ 1]);
}

// ---------------------------------------------------------------------------
// launchers

hipError_t sur_generate(hipStream_t s, const Problem *P, long long n, unsigned long long seed,
                        float *X, float *Y, double eps, int max_sub) {
    long long nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_gen_samples, dim3((unsigned)nb), dim3(256), 0, s, P, n, seed, X, Y, eps,
                       max_sub);
    return hipGetLastError();
}

hipError_t sur_grad(hipStream_t s, const unsigned short *thq, const float *X, const float *Y,
                    long long off, int B, float *partial, int nblocks) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(64), 0, s, thq, X, Y, off, B, partial);
    return hipGetLastError();
}

hipError_t sur_adam(hipStream_t s, const float *partial, int nblocks, float *theta,
                    unsigned short *thq, float *m, float *v, int t, float lr, float *loss_out) {
    hipLaunchKernelGGL(k_adam, dim3((NP + 1 + 255) / 256), dim3(256), 0, s, partial, nblocks, theta,
                       thq, m, v, t, lr, loss_out);
    return hipGetLastError();
}

hipError_t sur_quantize(hipStream_t s, const float *theta, unsigned short *thq) {
    hipLaunchKernelGGL(k_quantize, dim3((NP + 255) / 256), dim3(256), 0, s, theta, thq);
    return hipGetLastError();
}

hipError_t sur_eval(hipStream_t s, const unsigned short *thq, const float *X, long long n,
                    float *out) {
    long long nb = (n + 63) / 64;
    if (nb > 16384) nb = 16384;
    hipLaunchKernelGGL(k_mlp_eval, dim3((unsigned)nb), dim3(64), 0, s, thq, X, n, out);
    return hipGetLastError();
}

template <class R>
hipError_t sur_cool_pass(hipStream_t s, const Problem *P, int nseg, long long C, R *U,
                         const unsigned short *thq, StepState *st, int parity, unsigned *queue,
                         const unsigned *seg, unsigned *off, unsigned *fqueue, unsigned *fcount,
                         int mode, int ntable) {
    constexpr unsigned nb = 4096;  // 8192 wavefronts, grid-stride over the queue
    unsigned *offn = off, *offi = off + nseg + 1;
    // The network pass reads the FP32 table and Y (inputs and check); in FP32
    // the integrator of mode 3 reads the rows, and the exact integration (mode
    // 4, and the surrogate's integrator cells and fallbacks) also Y.
    const int y3 = mode != 3;
    const unsigned lds2 = cool_tab_bytes(ntable, true);
    const unsigned lds3 = sizeof(R) == 4 ? cool_tab_bytes(ntable, y3) : 0u;
    hipLaunchKernelGGL(k_queue_scan, dim3(2), dim3(1024), 0, s, seg, offn, nseg, fcount);
    if (mode == 2)
        hipLaunchKernelGGL((k_cool_dense<R, 2>), dim3(nb), dim3(128), lds2, s, P, U, thq, st,
                           parity, queue, offn, offi, nseg, C, fqueue, fcount, 1);
    hipLaunchKernelGGL((k_cool_dense<R, 3>), dim3(nb), dim3(128), lds3, s, P, U, thq, st, parity,
                       queue, offn, offi, nseg, C, fqueue, fcount, y3);
    return hipGetLastError();
}

template hipError_t sur_cool_pass<float>(hipStream_t, const Problem *, int, long long, float *,
                                         const unsigned short *, StepState *, int, unsigned *,
                                         const unsigned *, unsigned *, unsigned *, unsigned *, int,
                                         int);
template hipError_t sur_cool_pass<double>(hipStream_t, const Problem *, int, long long, double *,
                                          const unsigned short *, StepState *, int, unsigned *,
                                          const unsigned *, unsigned *, unsigned *, unsigned *, int,
                                          int);

}  // namespace tlq
