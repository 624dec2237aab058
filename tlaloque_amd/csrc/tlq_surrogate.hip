// tlq_surrogate.hip -- kernels of the learned cooling surrogate (see
// tlq_surrogate.hpp): training data, MFMA training gradient, Adam, inference,
// and the cooling pass of the fused step.
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

// GELU, tanh form: 0.5 x (1 + tanh(u/2)) = x sigmoid(u), u = 2 sqrt(2/pi) (x + 0.044715 x^3).
__device__ __forceinline__ float gelu(float x) {
    const float u = 1.5957691216057308f * (x + 0.044715f * x * x * x);
    return x * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.4426950408889634f * u));
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
// Cooling integrated to convergence: apply_cooling's per-cell formula
// (cooling.f90:67-112) in substeps h <= eps t_cool with Lambda re-read at the
// current temperature, until dt is used or T < T_min (the reference's gate).
__device__ double cool_sub_d(const Problem &P, double n, double T, double dt, double eps,
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

__device__ float cool_sub_f(const Problem &P, float n, float T, float dt, float eps, int max_sub) {
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
struct MlpSharedInfer {
    unsigned short h1[64][72];
    float f[64][65];
    float w1[3 * H], b1[H], b2[H], w3[H];
};

// B-operand fragments of W2 for z2 = h1 W2^T: lane holds W2[j][i..i+3] with
// j = 16 nt + lane%16, i = 16 kt + 4 (lane/16).
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

// layer 1 for the lane's own sample -> h1 (bf16) in shared memory
template <class S>
__device__ __forceinline__ void layer1(S &sh, int lane, const float xn[3], bool transposed) {
    for (int i = 0; i < H; ++i) {
        const float z = sh.b1[i] + sh.w1[3 * i] * xn[0] + sh.w1[3 * i + 1] * xn[1] +
                        sh.w1[3 * i + 2] * xn[2];
        const unsigned short q = f2bf(gelu(z));
        sh.h1[lane][i] = q;
        if constexpr (sizeof(S) == sizeof(MlpShared)) {
            if (transposed) sh.h1t[i][lane] = q;
        }
    }
}

// layer 2 (MFMA): f[sample][j] = sum_i h1[sample][i] W2[j][i]
template <class S>
__device__ __forceinline__ void layer2(S &sh, int lane, const bf16x4 w[4][4]) {
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

template <class S>
__device__ __forceinline__ void load_small(S &sh, const unsigned short *__restrict__ thq, int lane) {
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
        layer1(sh, lane, xn, true);
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
    __shared__ MlpSharedInfer sh;
    const int lane = threadIdx.x;
    load_small(sh, thq, lane);
    const float b3 = bf2f(thq[OB3]);
    bf16x4 wf[4][4];
    load_w2_fwd(thq, lane, wf);
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
        layer1(sh, lane, xn, false);
        __syncthreads();
        layer2(sh, lane, wf);
        __syncthreads();
        float o = b3;
        for (int j = 0; j < H; ++j) o += sh.w3[j] * gelu(sh.f[lane][j] + sh.b2[j]);
        if (s < n) out[s] = o;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Cooling pass of the fused step (modes 2 and 3): one wavefront per 64 cells
// of an x row.  Cells at T >= T_min with dt <= eps t_cool take one substep
// (the converged integrator takes no more).  For the others mode 3 runs the
// converged integrator and mode 2 evaluates the surrogate (skipped by a
// wavefront with no such cell) and keeps its answer T1 when
//   * the inputs lie in the training box,
//   * 10 K <= T1 <= T0 (cooling only lowers T, never below the floor), and
//   * the exact cooling time from T0 to T1, t = k_B (Y(T0) - Y(T1)) / ((gamma-1) n),
//     is within 0.05 min(dt, t_cool(T1)) of dt, so that ln T1 is within 5% of
//     its change (at most 0.05) of the exact solution -- or, for T1 below the
//     gate T_min (where the converged integrator stops), the gas reaches T_min
//     within that time;
// other cells take the exact converged integrator (the fallback).  Then the
// CFL maxima of the cooled state (time_step of the next step).
constexpr float COOL_EPS = 0.01f;
constexpr int COOL_MAXSUB = 100000;

template <class R, int MODE>
__global__ __launch_bounds__(64) void k_cool_pass(const Problem *__restrict__ Pp, R *__restrict__ U,
                                                  const unsigned short *__restrict__ thq,
                                                  StepState *__restrict__ st, int parity,
                                                  unsigned long long *__restrict__ counts) {
    __shared__ MlpSharedInfer sh;
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const int lane = threadIdx.x;
    const int nxw = (g.nx + 63) / 64;  // wavefronts per row
    const double dt = dt_from_maxima(P, st->maxima[parity]);
    bf16x4 wf[4][4];
    float b3 = 0.f;
    if (MODE == 2) {
        load_small(sh, thq, lane);
        b3 = bf2f(thq[OB3]);
        load_w2_fwd(thq, lane, wf);
        __syncthreads();
    }
    double mxd = 0.0, myd = 0.0, mzd = 0.0;
    unsigned long long ncool = 0, nfall = 0;
    const long long nrows = (long long)g.ny * g.nz * nxw;
    for (long long w = blockIdx.x; w < nrows; w += gridDim.x) {
        const int wx = (int)(w % nxw);
        const long long r = w / nxw;
        const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
        const int i = wx * 64 + lane + 1;
        const bool in = i <= g.nx;
        const long long c = g.idx(in ? i : g.nx, j, k);
        double u[5];
#pragma unroll
        for (int e = 0; e < 5; ++e) u[e] = (double)U[e * g.nc + c];
        const double rho = u[0];
        const double ek = 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / rho;
        double p = (P.gamma - 1.0) * (u[4] - ek);
        if (!(p > 1e-20)) p = 1e-20;
        const double nd = rho / (P.mu0 * AMU);
        const double T = p * P.mu0 * AMU / (rho * KB);
        const bool hot = in && T >= P.tmin;
        double Tn = T;
        // dt <= eps t_cool: the converged integrator takes exactly one
        // substep, i.e. the reference's formula -- no network, no fallback
        bool big = hot;
        if (hot) {
            if constexpr (sizeof(R) == 8) {
                big = find_coef(P, T) * nd * nd / (p / (P.gamma - 1.0)) * dt > (double)COOL_EPS;
            } else {
                const float ce = lambda_f(P, (float)T) * (float)nd * (float)nd *
                                 (float)(P.gamma - 1.0) / (float)p;
                big = ce * (float)dt > COOL_EPS;
            }
        }
        bool need_exact = false;
        if (MODE == 2 && __any(big)) {
            float x[3] = {0.f, 0.f, 0.f};
            if (big) {
                const float lam = lambda_f(P, (float)T);
                const float ce = lam * (float)nd * (float)nd * (float)(P.gamma - 1.0) / (float)p;
                x[0] = log10f((float)nd);
                x[1] = log10f((float)T);
                x[2] = log10f(ce * (float)dt);
            }
            float xn[3];
            normalise(x, xn);
            layer1(sh, lane, xn, false);
            __syncthreads();
            layer2(sh, lane, wf);
            __syncthreads();
            float o = b3;
            for (int q = 0; q < H; ++q) o += sh.w3[q] * gelu(sh.f[lane][q] + sh.b2[q]);
            __syncthreads();
            if (big) {
                const bool inbox = x[0] >= X0LO && x[0] <= X0HI && x[1] >= X1LO && x[1] <= X1HI &&
                                   x[2] <= X2HI;
                const float d = __builtin_amdgcn_exp2f(x[2] * 3.32192809488736f);
                const float rh = -0.4342944819032518f * d / (1.f + d) * o;  // log10(T1/T0)
                const float rmin = log10f(10.f / (float)T);
                bool ok = inbox && rh <= 1e-3f && rh >= rmin - 1e-3f;
                const float T0 = (float)T;
                const float T1 = T0 * __builtin_amdgcn_exp2f(fminf(rh, 0.f) * 3.32192809488736f);
                if (ok) {
                    const float kn = (float)(KB / (P.gamma - 1.0)) / (float)nd;
                    const float fdt = (float)dt;
                    // a time error dt' shifts ln T1 by dt'/t_cool(T1): allow 5% of
                    // the change in ln T, at most 0.05
                    const float Te = fmaxf(T1, P.f_tmin);
                    const float tol = 0.05f * fminf(fdt, kn * Te / lambda_f(P, Te));
                    if (T1 <= P.f_tmin)
                        ok = T1 >= 0.97f * P.f_tmin && kn * cool_dY(P, T0, P.f_tmin) <= fdt + tol;
                    else
                        ok = fabsf(kn * cool_dY(P, T0, T1) - fdt) <= tol;
                }
                if (ok) Tn = (double)T1;
                else need_exact = true;
            }
        } else if (MODE != 2) {
            need_exact = big;
        }
        if (hot && !big) {  // one substep: apply_cooling's formula
            if constexpr (sizeof(R) == 8) Tn = cool_sub_d(P, nd, T, dt, COOL_EPS, 1);
            else Tn = cool_sub_f(P, (float)nd, (float)T, (float)dt, COOL_EPS, 1);
        }
        if (need_exact) {  // the converged integrator (mode 3, and mode 2's fallback)
            if constexpr (sizeof(R) == 8) Tn = cool_sub_d(P, nd, T, dt, COOL_EPS, COOL_MAXSUB);
            else Tn = cool_sub_f(P, (float)nd, (float)T, (float)dt, COOL_EPS, COOL_MAXSUB);
        }
        if (hot) {
            ncool += big;
            if (MODE == 2 && need_exact) ++nfall;
            const double pn = p * (Tn / T);
            const R e5 = (R)(ek + pn / (P.gamma - 1.0));
            U[4 * g.nc + c] = e5;
            p = pn;
        }
        if (in) {
            const double cs = sqrt(P.gamma * p / rho);
            mxd = fmax(mxd, fabs(u[1] / rho) + cs);
            myd = fmax(myd, fabs(u[2] / rho) + cs);
            mzd = fmax(mzd, fabs(u[3] / rho) + cs);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ncool += __shfl_xor(ncool, o);
        nfall += __shfl_xor(nfall, o);
    }
    if (lane == 0 && counts) {
        atomicAdd(counts, ncool);
        atomicAdd(counts + 1, nfall);
    }
    block_max3<64>(mxd, myd, mzd, st->maxima[parity ^ 1]);
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
hipError_t sur_cool_pass(hipStream_t s, const Problem *P, const Grid &g, R *U,
                         const unsigned short *thq, StepState *st, int parity,
                         unsigned long long *counts, int mode) {
    const long long nrows = (long long)g.ny * g.nz * ((g.nx + 63) / 64);
    long long nb = nrows < 256 * 64 ? nrows : 256 * 64;
    if (mode == 2)
        hipLaunchKernelGGL((k_cool_pass<R, 2>), dim3((unsigned)nb), dim3(64), 0, s, P, U, thq, st,
                           parity, counts);
    else
        hipLaunchKernelGGL((k_cool_pass<R, 3>), dim3((unsigned)nb), dim3(64), 0, s, P, U, thq, st,
                           parity, counts);
    return hipGetLastError();
}

template hipError_t sur_cool_pass<float>(hipStream_t, const Problem *, const Grid &, float *,
                                         const unsigned short *, StepState *, int,
                                         unsigned long long *, int);
template hipError_t sur_cool_pass<double>(hipStream_t, const Problem *, const Grid &, double *,
                                          const unsigned short *, StepState *, int,
                                          unsigned long long *, int);

}  // namespace tlq
