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
// Cooling of the fused step in modes 2 and 3, after stage 2, in two passes.
//
// k_cool_classify visits every interior cell.  Cells below T_min are left
// alone; cells with dt <= eps t_cool take one substep (all the converged
// integrator would take); their CFL maxima are formed here.  Every other
// cell goes to a queue (one atomic per wavefront).
//
// k_cool_dense runs the queued cells in full wavefronts: mode 3 integrates
// them to convergence; mode 2 evaluates the surrogate and keeps its answer T1
// when
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
//
constexpr float COOL_EPS = 0.01f;
constexpr int COOL_MAXSUB = 100000;
// A surrogate answer at the gate is set just below it, in the 1% band where
// the converged integrator's last substep ends (so the cell stays gated).
constexpr double SNAP = 0.995;

struct CellState {
    double rho, ek, p, nd, T;
    float rinv;  // 1/rho (FP32 state)
};

// Primitive state of one cell for cooling: FP64 arithmetic for an FP64
// state; for an FP32 state the guarded fast path of calc_prims (pressure
// recovery in FP64 with a refined reciprocal, the rest in FP32).
template <class R>
__device__ __forceinline__ CellState cell_state(const Problem &P, const R *__restrict__ U,
                                                long long nc, long long c, double u[5]) {
#pragma unroll
    for (int e = 0; e < 5; ++e) u[e] = (double)U[e * nc + c];
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

template <class R>
__device__ __forceinline__ void finish_cell(const Problem &P, R *__restrict__ U, long long nc,
                                            long long c, const double u[5], const CellState &s,
                                            double Tn, double &mx, double &my, double &mz) {
    if constexpr (sizeof(R) == 8) {
        const double pn = s.p * (Tn / s.T);
        U[4 * nc + c] = (R)(s.ek + pn / (P.gamma - 1.0));
        const double cs = sqrt(P.gamma * pn / s.rho);
        mx = fmax(mx, fabs(u[1] / s.rho) + cs);
        my = fmax(my, fabs(u[2] / s.rho) + cs);
        mz = fmax(mz, fabs(u[3] / s.rho) + cs);
    } else {
        const float pn = (float)s.p * ((float)Tn * __builtin_amdgcn_rcpf((float)s.T));
        U[4 * nc + c] = (R)(s.ek + (double)pn * (double)P.f_inv_gm1);
        const float cs = __builtin_amdgcn_sqrtf((float)P.gamma * pn * s.rinv);
        mx = fmax(mx, (double)(fabsf((float)u[1] * s.rinv) + cs));
        my = fmax(my, (double)(fabsf((float)u[2] * s.rinv) + cs));
        mz = fmax(mz, (double)(fabsf((float)u[3] * s.rinv) + cs));
    }
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

// Block b of the classifying pass owns the contiguous interior cells
// [b C, (b+1) C) (flat index, x fastest) and the queue segment of the same
// range; it fills the segment with an LDS counter and leaves its length in
// seg[b].  No global atomics besides the CFL maxima.
template <class R>
__global__ __launch_bounds__(256) void k_cool_classify(const Problem *__restrict__ Pp,
                                                       R *__restrict__ U,
                                                       StepState *__restrict__ st, int parity,
                                                       unsigned *__restrict__ queue,
                                                       unsigned *__restrict__ seg, long long C) {
    __shared__ unsigned s_n;
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const double dt = dt_from_maxima(P, st->maxima[parity]);
    const int lane = threadIdx.x & 63;
    const long long ncell = (long long)g.nx * g.ny * g.nz;
    const long long first = (long long)blockIdx.x * C;
    unsigned *qseg = queue + first;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    double mx = 0.0, my = 0.0, mz = 0.0;
    // C is a multiple of the block size: every wavefront runs whole
    // iterations, so its ballot sees all 64 lanes
    for (long long t = first + threadIdx.x; t < first + C; t += blockDim.x) {
        const bool in = t < ncell;
        const long long tt = in ? t : ncell - 1;
        const int i = (int)(tt % g.nx) + 1;
        const long long r = tt / g.nx;
        const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
        const long long c = g.idx(i, j, k);
        double u[5];
        const CellState s = cell_state(P, U, g.nc, c, u);
        const bool hot = in && s.T >= P.tmin;
        const bool big = hot && dt_over_tcool<R>(P, s, dt) > (double)COOL_EPS;
        if (in && !big) {
            if (hot) finish_cell(P, U, g.nc, c, u, s, integrate<R>(P, s, dt, 1), mx, my, mz);
            else cell_maxima<R>(P, u, s, mx, my, mz);
        }
        const unsigned long long mask = __ballot(big);
        if (mask) {
            const int leader = __ffsll((long long)mask) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&s_n, (unsigned)__popcll(mask));
            base = __shfl(base, leader);
            if (big) qseg[base + __popcll(mask & ((1ull << lane) - 1ull))] = (unsigned)c;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) seg[blockIdx.x] = s_n;
    block_max3<256>(mx, my, mz, st->maxima[parity ^ 1]);
}

// Block b of the dense pass runs queue segment b, 64 cells per wavefront
// step; fall[b] receives the segment's fallback count.
template <class R, int MODE>
__global__ __launch_bounds__(128) void k_cool_dense(const Problem *__restrict__ Pp,
                                                    R *__restrict__ U,
                                                    const unsigned short *__restrict__ thq,
                                                    StepState *__restrict__ st, int parity,
                                                    const unsigned *__restrict__ queue,
                                                    const unsigned *__restrict__ seg,
                                                    unsigned *__restrict__ fqueue,
                                                    unsigned *__restrict__ fall, long long C) {
    __shared__ MlpSmall sm;
    __shared__ unsigned s_fall;
    unsigned *fseg = MODE == 2 ? fqueue + (long long)blockIdx.x * C : nullptr;
    const Problem &P = *Pp;
    const Grid &g = P.g;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long long nq = seg[blockIdx.x];
    const unsigned *qseg = queue + (long long)blockIdx.x * C;
    const double dt = dt_from_maxima(P, st->maxima[parity]);
    bf16x4 w2a[4][4];
    float b3 = 0.f;
    if (threadIdx.x == 0) s_fall = 0;
    if (MODE == 2 && nq > 0) {
        load_small_f(sm, thq, threadIdx.x, blockDim.x);
        b3 = bf2f(thq[OB3]);
        load_w2_fwd(thq, lane, w2a);
    }
    __syncthreads();
    double mx = 0.0, my = 0.0, mz = 0.0;
    for (long long base = (long long)wv * 64; base < nq; base += 128) {
        const long long qi = base + lane;
        const bool valid = qi < nq;
        const long long c = (long long)qseg[valid ? qi : base];
        double u[5];
        const CellState s = cell_state(P, U, g.nc, c, u);
        bool need_exact = valid;
        double Tn = s.T;
        if constexpr (MODE == 2) {
            const float lam = lambda_f(P, (float)s.T);
            const float ce = lam * (float)s.nd * (float)s.nd * (float)(P.gamma - 1.0) *
                             __builtin_amdgcn_rcpf((float)s.p);
            float x[3] = {lg10((float)s.nd), lg10((float)s.T), lg10(ce * (float)dt)};
            float xn[3];
            normalise(x, xn);
            const float o = mlp_wave(sm, w2a, b3, lane, xn);
            const bool inbox = x[0] >= X0LO && x[0] <= X0HI && x[1] >= X1LO && x[1] <= X1HI &&
                               x[2] <= X2HI;
            const float d = __builtin_amdgcn_exp2f(x[2] * 3.32192809488736f);
            const float rh = -0.4342944819032518f * d / (1.f + d) * o;  // log10(T1/T0)
            const float rmin = lg10(10.f * __builtin_amdgcn_rcpf((float)s.T));
            bool ok = valid && inbox && rh <= 1e-3f && rh >= rmin - 1e-3f;
            bool snap = false;
            const float T0 = (float)s.T;
            const float T1 = T0 * __builtin_amdgcn_exp2f(fminf(rh, 0.f) * 3.32192809488736f);
            if (ok) {
                const float kn = (float)(KB / (P.gamma - 1.0)) / (float)s.nd;
                const float fdt = (float)dt;
                // a time error dt' shifts ln T1 by dt'/t_cool(T1): allow 5% of
                // the change in ln T, at most 0.05
                const float Te = fmaxf(T1, P.f_tmin);
                const float tol = 0.05f * fminf(fdt, kn * Te / lambda_f(P, Te));
                if (T1 <= 1.1f * P.f_tmin) {
                    // the network says the gas reaches the gate: accepted when
                    // the exact cooling time to T_min fits in dt, and the answer
                    // is the gate itself (where the integrator stops)
                    ok = kn * cool_dY(P, T0, P.f_tmin) <= fdt + tol;
                    snap = true;
                } else {
                    ok = fabsf(kn * cool_dY(P, T0, T1) - fdt) <= tol;
                }
            }
            if (ok) {
                Tn = snap ? (double)SNAP * P.tmin : (double)T1;
                need_exact = false;
            }
            // fallbacks go to a queue of their own, integrated by a dense
            // mode-3 pass: a lane that needs the integrator does not hold up
            // the other 63
            const unsigned long long fm = __ballot(need_exact);
            if (fm) {
                const int leader = __ffsll((long long)fm) - 1;
                unsigned fb = 0;
                if (lane == leader) fb = atomicAdd(&s_fall, (unsigned)__popcll(fm));
                fb = __shfl(fb, leader);
                if (need_exact) fseg[fb + __popcll(fm & ((1ull << lane) - 1ull))] = (unsigned)c;
            }
            if (valid && !need_exact) finish_cell(P, U, g.nc, c, u, s, Tn, mx, my, mz);
        } else {
            if (valid) finish_cell(P, U, g.nc, c, u, s, integrate<R>(P, s, dt, COOL_MAXSUB), mx,
                                   my, mz);
        }
    }
    __syncthreads();
    if (MODE == 2 && threadIdx.x == 0) fall[blockIdx.x] = s_fall;
    if (nq > 0) block_max3<128>(mx, my, mz, st->maxima[parity ^ 1]);
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
                         const unsigned short *thq, StepState *st, int parity, unsigned *queue,
                         unsigned *seg, unsigned *fqueue, unsigned *fall, int mode) {
    const long long ncell = (long long)g.nx * g.ny * g.nz;
    const long long C = cool_segment(ncell);
    const unsigned nb = (unsigned)((ncell + C - 1) / C);
    hipLaunchKernelGGL(k_cool_classify<R>, dim3(nb), dim3(256), 0, s, P, U, st, parity, queue,
                       seg, C);
    if (mode == 2) {
        hipLaunchKernelGGL((k_cool_dense<R, 2>), dim3(nb), dim3(128), 0, s, P, U, thq, st, parity,
                           queue, seg, fqueue, fall, C);
        hipLaunchKernelGGL((k_cool_dense<R, 3>), dim3(nb), dim3(128), 0, s, P, U, thq, st, parity,
                           fqueue, fall, nullptr, nullptr, C);
    } else {
        hipLaunchKernelGGL((k_cool_dense<R, 3>), dim3(nb), dim3(128), 0, s, P, U, thq, st, parity,
                           queue, seg, nullptr, nullptr, C);
    }
    return hipGetLastError();
}

template hipError_t sur_cool_pass<float>(hipStream_t, const Problem *, const Grid &, float *,
                                         const unsigned short *, StepState *, int, unsigned *,
                                         unsigned *, unsigned *, unsigned *, int);
template hipError_t sur_cool_pass<double>(hipStream_t, const Problem *, const Grid &, double *,
                                          const unsigned short *, StepState *, int, unsigned *,
                                          unsigned *, unsigned *, unsigned *, int);

}  // namespace tlq
