// mcgb_surrogate.hpp — the learned first stage of the delayed-acceptance sweep.
//
// A per-pair MLP 4 -> 64 -> 64 -> 1 with SiLU activations on the features
// (r, u_i.r^, u_j.r^, u_i.u_j) of a GB pair inside the cutoff (r^ from i to j), fitted to the
// exact pair energy clipped at `clip`.  The reference has no learned part.
//
// Parameters: one flat FP32 vector (kNParam floats), W[out][in] row-major:
//   W1[64][4] b1[64] W2[64][64] b2[64] W3[64] b3.
// Input normalisation: x0 = r/2 - 1 (r in [0, 4) -> [-1, 1)), x1..x3 = a, b, c.
//
// Inference in the sweep runs the two dense layers on the matrix cores
// (v_mfma_f32_32x32x16_bf16: BF16 operands, FP32 accumulation), 32 pairs per wavefront tile.
// Operand layouts (one wavefront, lane l, g = l / 32):
//   A (32 x 16): lane l holds row l % 32, k-slots of group g (8 values);
//   B (16 x 32): lane l holds column l % 32, k-slots of group g (8 values);
//   C (32 x 32, 16 values): lane l holds column l % 32, rows 8 (r / 4) + 4 g + r % 4.
// Pairs are the columns.  Layer 1 contracts the inputs split into BF16 high and low parts
// (xh0..xh3, xl0..xl3: ~16 significant bits) against (W1, W1) in group 0 and the constant 1
// against b1 in group 1; its C tile, after SiLU and BF16 rounding, is layer 2's B operand directly:
// the k-slot (g, idx) of K step m carries hidden unit hidx(m / 2, 8 (m % 2) + idx, g), and the
// host packs W2 into the matching A operand.  Only the k order changes, never the sum.
//
// Training (FP32): one single-wavefront workgroup of 64 samples computes forward and backward,
// the three 64 x 64 x 64 products on the matrix cores (v_mfma_f32_32x32x2_f32), the rest on the
// vector ALU with the per-sample intermediates in LDS; a second kernel sums the workgroup
// partials in a fixed order; a third applies Adam.
#pragma once

#include "mcgb_device.hpp"

namespace mcgb {

constexpr int kHid = 64;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = 256;
constexpr int kOffW2 = 320;
constexpr int kOffB2 = 4416;
constexpr int kOffW3 = 4480;
constexpr int kOffB3 = 4544;
constexpr int kNParam = 4545;
constexpr int kPairRec = 5;          // emitted record: r, a, b, c, U
constexpr int kDaBins = 64;          // DA error histogram: 64 bins over [-kDaRange, kDaRange] kT
constexpr float kDaRange = 4.0f;     //  + underflow + overflow

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline int mfma_row(int r, int g) { return 8 * (r >> 2) + 4 * g + (r & 3); }

// Packed MFMA operands, lane-minor so that a wavefront loads them coalesced.
struct MlpPacked {
    uint16_t w1[2][8][64];       // [hidden tile][elem][lane]   (BF16 bits)
    uint16_t w2[2][4][8][64];    // [out tile][K step][elem][lane]
    float b2[2][64][16];         // [out tile][lane][C row slot]: a lane's 16 values contiguous
    float w3[2][64][16];
    float b3;
    float pad[15];
};

__device__ inline float silu_f(float z) {
    const float e = __builtin_amdgcn_exp2f(-1.4426950408889634f * z);
    return z * __builtin_amdgcn_rcpf(1.0f + e);
}
// d silu / dz = s (1 + z (1 - s)), s = sigmoid(z)
__device__ inline float silu_grad(float z) {
    const float s = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
    return s * (1.0f + z * (1.0f - s));
}

// Per-lane MFMA operands of the network: the matrix operands held in registers by the kernels
// that use them; the output layer's bias and weights (64 more VGPRs) are read per tile from
// the packed record (L1/L2-resident), which keeps the delayed-acceptance sweep under 256
// registers.
struct MlpRegs {
    bf16x8 w1[2];
    bf16x8 w2[2][4];
    const floatx16* b2;          // this lane's [2] output-tile biases
    const floatx16* w3;          // this lane's [2] output-layer weights
    float b3;
};

__device__ inline bf16x8 load_bf16x8(const uint16_t* p, int lane) {
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = __builtin_bit_cast(__bf16, p[e * 64 + lane]);
    return v;
}

__device__ inline void mlp_load(const MlpPacked* __restrict__ M, int lane, MlpRegs& R) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        R.w1[t] = load_bf16x8(&M->w1[t][0][0], lane);
#pragma unroll
        for (int m = 0; m < 4; ++m) R.w2[t][m] = load_bf16x8(&M->w2[t][m][0][0], lane);
    }
    R.b2 = (const floatx16*)&M->b2[0][lane][0];        // tile t at [t * 64]
    R.w3 = (const floatx16*)&M->w3[0][lane][0];
    R.b3 = M->b3;
}

// Network input of one pair: lane group 0 supplies the inputs as BF16 high + low parts (zeros
// when !on), lane group 1 the constant 1 of the bias slot.
__device__ inline bf16x8 mlp_input(float r, float a, float b, float c, bool on, int lane) {
    bf16x8 x;
    if (lane < 32) {
        const float v[4] = {0.5f * r - 1.0f, a, b, c};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float f = on ? v[e] : 0.0f;
            const __bf16 h = (__bf16)f;
            x[e] = h;
            x[4 + e] = (__bf16)(f - (float)h);
        }
    } else {
        x[0] = (__bf16)1.0f;
#pragma unroll
        for (int e = 1; e < 8; ++e) x[e] = (__bf16)0.0f;
    }
    return x;
}

// Outputs of the 32 pairs whose inputs lanes 0..31 supplied: lanes j and j + 32 both return
// pair j's value.  Must be called by the whole wavefront (wavefront-uniform control flow).
__device__ inline float mlp_tile32(const MlpRegs& R, bf16x8 X) {
    // the output layer's operands, loaded first so that the loads overlap layer 1
    const floatx16 b2a = R.b2[0], b2b = R.b2[64], w3a = R.w3[0], w3b = R.w3[64];
    floatx16 z1a = {}, z1b = {};
    z1a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(R.w1[0], X, z1a, 0, 0, 0);
    z1b = __builtin_amdgcn_mfma_f32_32x32x16_bf16(R.w1[1], X, z1b, 0, 0, 0);
    bf16x8 h[4];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        h[0][e] = (__bf16)silu_f(z1a[e]);
        h[1][e] = (__bf16)silu_f(z1a[8 + e]);
        h[2][e] = (__bf16)silu_f(z1b[e]);
        h[3][e] = (__bf16)silu_f(z1b[8 + e]);
    }
    floatx16 z2a = b2a, z2b = b2b;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        z2a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(R.w2[0][m], h[m], z2a, 0, 0, 0);
        z2b = __builtin_amdgcn_mfma_f32_32x32x16_bf16(R.w2[1][m], h[m], z2b, 0, 0, 0);
    }
    float y = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) y += w3a[r] * silu_f(z2a[r]) + w3b[r] * silu_f(z2b[r]);
    y += __shfl_xor(y, 32);
    return y + R.b3;
}

// FP32 reference forward of one pair (VALU), P = flat parameter vector.
__device__ inline float mlp_fwd_f32(const float* __restrict__ P, const float x[4]) {
    float h1[kHid];
#pragma unroll
    for (int h = 0; h < kHid; ++h) {
        float z = P[kOffB1 + h];
#pragma unroll
        for (int k = 0; k < 4; ++k) z += P[kOffW1 + 4 * h + k] * x[k];
        h1[h] = silu_f(z);
    }
    float y = P[kOffB3];
    for (int o = 0; o < kHid; ++o) {
        float z = P[kOffB2 + o];
#pragma unroll
        for (int h = 0; h < kHid; ++h) z += P[kOffW2 + kHid * o + h] * h1[h];
        y += P[kOffW3 + o] * silu_f(z);
    }
    return y;
}

// ------------------------------------------------------------------ inference kernels (tests)

// y[i] of n feature rows x[i] = (r, a, b, c); mfma = 1: BF16 matrix-core tiles, else FP32.
__global__ __launch_bounds__(64) void k_mlp_eval(const float* __restrict__ P,
                                                 const MlpPacked* __restrict__ M, int n,
                                                 const float* __restrict__ x, int mfma,
                                                 float* __restrict__ y) {
    const int lane = threadIdx.x;
    if (mfma) {
        MlpRegs R;
        mlp_load(M, lane, R);
        const int j = blockIdx.x * 32 + (lane & 31);
        const bool on = lane < 32 && j < n;
        float f[4] = {0.f, 0.f, 0.f, 0.f};
        if (on)
            for (int k = 0; k < 4; ++k) f[k] = x[4 * (size_t)j + k];
        const float v = mlp_tile32(R, mlp_input(f[0], f[1], f[2], f[3], on, lane));
        if (on) y[j] = v;
    } else {
        const int j = blockIdx.x * 64 + lane;
        if (j >= n) return;
        float f[4];
        for (int k = 0; k < 4; ++k) f[k] = x[4 * (size_t)j + k];
        f[0] = 0.5f * f[0] - 1.0f;
        y[j] = mlp_fwd_f32(P, f);
    }
}

// ------------------------------------------------------------------ training kernels

// One 32 x 32 tile of a 64-deep product on the matrix cores, FP32 operands and accumulation
// (v_mfma_f32_32x32x2_f32, 32 K steps of 2): sum_k a(i, k) b(k, j) for tile rows i and columns
// j.  Operand layouts (lane l, g = l / 32): A (32 x 2) lane l holds a(l % 32, k0 + g); B (2 x 32)
// lane l holds b(k0 + g, l % 32); C (32 x 32) lane l, element r holds row mfma_row(r, g),
// column l % 32.
template <typename FA, typename FB>
__device__ inline floatx16 mm64_tile(int lane, FA a, FB b) {
    const int i = lane & 31, g = lane >> 5;
    floatx16 c = {};
#pragma unroll 8
    for (int k0 = 0; k0 < kHid; k0 += 2)
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(a(i, k0 + g), b(k0 + g, i), c, 0, 0, 0);
    return c;
}

// Gradient of (1/n) sum w (y - min(t, clip))^2 over one workgroup's 64 samples (one wavefront),
// with w = 1 (wfloor <= 0) or w = (wfloor + 1) / (wfloor + |min(t, clip)|) (relative weighting).
//   data != nullptr: sample s of step `step` is record floor(u * ndata), u a Philox draw
//   (key seed, counter {sample, step, 0, 0x5A5A0000}), mapped by a random element of the
//   energy's symmetry group (u_i -> -u_i: (a, c) -> (-a, -c); u_j -> -u_j: (b, c) -> (-b, -c);
//   i <-> j: (a, b) -> (-b, -a)); else samples are xs/ts[0, nexp) as given.
// The three 64 x 64 x 64 products run on the matrix cores in FP32:
//   Z2 = H1 W2^T (forward), gW2 = dZ2^T H1, dH1 = dZ2 W2 (backward);
// layer 1 (4 inputs), the output layer and the per-row and per-column sums on the vector ALU,
// thread t owning sample t for the row work and output row t of a gradient for the column work.
// Writes kNParam partial sums to partial[blockIdx.x] and the loss sum to lpart[blockIdx.x].
__global__ __launch_bounds__(64) void k_mlp_grad(const float* __restrict__ Pg,
                                                 const float* __restrict__ data, long long ndata,
                                                 const float* __restrict__ xs,
                                                 const float* __restrict__ ts, int nexp,
                                                 uint32_t step, uint32_t key0, uint32_t key1,
                                                 float clip, float wfloor, float inv_n,
                                                 float* __restrict__ partial,
                                                 double* __restrict__ lpart) {
    // rows padded to 65 floats: column walks and row walks both stay free of bank conflicts
    __shared__ float W2[kHid][kHid + 1];      // W2[o][h]
    __shared__ float Q[kOffW2 + 1];           // W1 (256) and b1 (64)
    __shared__ float R[kHid * 2 + 1];         // b2 (64), W3 (64), b3
    __shared__ float H1[kHid][kHid + 1];      // h1 rows [sample][h]; later dz1 rows
    __shared__ float Z2[kHid][kHid + 1];      // z2 rows [sample][o]; later dz2 rows
    __shared__ float X[kHid][4];
    __shared__ float DY[kHid];
    const int t = threadIdx.x;
    for (int p = t; p < kHid * kHid; p += 64) W2[p >> 6][p & 63] = Pg[kOffW2 + p];
    for (int p = t; p < kOffW2; p += 64) Q[p] = Pg[p];
    for (int p = t; p < 2 * kHid + 1; p += 64) R[p] = Pg[kOffB2 + p];
    const size_t s = (size_t)blockIdx.x * 64 + t;
    float x[4] = {0.f, 0.f, 0.f, 0.f}, tg = 0.f;
    bool valid;
    if (data) {
        uint32_t u[4];
        philox4x32_10((uint32_t)s, step, 0u, 0x5A5A0000u, key0, key1, u);
        const long long idx = (long long)(((uint64_t)u[0] * (uint64_t)ndata) >> 32);
        const float* rec = data + (size_t)idx * kPairRec;
        x[0] = rec[0]; x[1] = rec[1]; x[2] = rec[2]; x[3] = rec[3]; tg = rec[4];
        if (u[1] & 1u) { x[1] = -x[1]; x[3] = -x[3]; }
        if (u[1] & 2u) { x[2] = -x[2]; x[3] = -x[3]; }
        if (u[1] & 4u) { const float ta = x[1]; x[1] = -x[2]; x[2] = -ta; }
        valid = true;
    } else {
        valid = (long long)s < (long long)nexp;
        if (valid) {
            for (int k = 0; k < 4; ++k) x[k] = xs[4 * s + k];
            tg = ts[s];
        }
    }
    tg = tg < clip ? tg : clip;
    x[0] = 0.5f * x[0] - 1.0f;
    for (int k = 0; k < 4; ++k) X[t][k] = x[k];
    __syncthreads();

    // layer 1 (thread t: sample t)
#pragma unroll
    for (int h = 0; h < kHid; ++h) {
        float z = Q[kOffB1 + h];
#pragma unroll
        for (int k = 0; k < 4; ++k) z += Q[kOffW1 + 4 * h + k] * x[k];
        H1[t][h] = silu_f(z);
    }
    __syncthreads();
    const int g = t >> 5, j = t & 31;
    // layer 2 on the matrix cores: Z2[s][o] = sum_h H1[s][h] W2[o][h] + b2[o]
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const floatx16 c = mm64_tile(
                t, [&](int i, int k) { return H1[32 * mt + i][k]; },
                [&](int k, int i) { return W2[32 * nt + i][k]; });
#pragma unroll
            for (int r = 0; r < 16; ++r)
                Z2[32 * mt + mfma_row(r, g)][32 * nt + j] = c[r] + R[32 * nt + j];
        }
    __syncthreads();
    // output layer and loss (thread t: sample t)
    float y = R[2 * kHid];
#pragma unroll 8
    for (int o = 0; o < kHid; ++o) y += R[kHid + o] * silu_f(Z2[t][o]);
    const float diff = y - tg;
    const float wt = wfloor > 0.0f ? (wfloor + 1.0f) / (wfloor + fabsf(tg)) : 1.0f;
    const float dy = valid ? 2.0f * wt * diff * inv_n : 0.0f;
    DY[t] = dy;
    const double loss = valid ? (double)wt * (double)diff * (double)diff : 0.0;
    __syncthreads();
    // thread t: column o = t of Z2 -> gW3[o], gb2[o], and dz2 in place
    float* out = partial + (size_t)blockIdx.x * kNParam;
    {
        const float w3 = R[kHid + t];
        float gw3 = 0.f, gb2 = 0.f;
        for (int q = 0; q < 64; ++q) {
            const float z = Z2[q][t];
            gw3 += DY[q] * silu_f(z);
            const float d2 = DY[q] * w3 * silu_grad(z);
            gb2 += d2;
            Z2[q][t] = d2;
        }
        out[kOffW3 + t] = gw3;
        out[kOffB2 + t] = gb2;
    }
    __syncthreads();
    // gW2[o][h] = sum_s dZ2[s][o] H1[s][h] on the matrix cores, written from the C tiles
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const floatx16 c = mm64_tile(
                t, [&](int i, int k) { return Z2[k][32 * mt + i]; },
                [&](int k, int i) { return H1[k][32 * nt + i]; });
#pragma unroll
            for (int r = 0; r < 16; ++r)
                out[kOffW2 + kHid * (32 * mt + mfma_row(r, g)) + 32 * nt + j] = c[r];
        }
    __syncthreads();
    // dH1[s][h] = sum_o dZ2[s][o] W2[o][h] on the matrix cores; dz1 = dh1 silu'(z1) into H1
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const floatx16 c = mm64_tile(
                t, [&](int i, int k) { return Z2[32 * mt + i][k]; },
                [&](int k, int i) { return W2[k][32 * nt + i]; });
            const int h = 32 * nt + j;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int sr = 32 * mt + mfma_row(r, g);
                float z = Q[kOffB1 + h];
#pragma unroll
                for (int k = 0; k < 4; ++k) z += Q[kOffW1 + 4 * h + k] * X[sr][k];
                H1[sr][h] = c[r] * silu_grad(z);
            }
        }
    __syncthreads();
    // thread t: hidden unit h = t of layer 1 -> gb1[h], gW1[h][k]
    {
        float gb1 = 0.f, gw1[4] = {0.f, 0.f, 0.f, 0.f};
        for (int q = 0; q < 64; ++q) {
            const float d1 = H1[q][t];
            gb1 += d1;
#pragma unroll
            for (int k = 0; k < 4; ++k) gw1[k] += d1 * X[q][k];
        }
        out[kOffB1 + t] = gb1;
#pragma unroll
        for (int k = 0; k < 4; ++k) out[kOffW1 + 4 * t + k] = gw1[k];
    }
    // db3: fixed-order sum by thread 0
    if (t == 0) {
        float gb3 = 0.f;
        for (int q = 0; q < 64; ++q) gb3 += DY[q];
        out[kOffB3] = gb3;
    }
    // loss: wavefront tree (fixed order)
    double l = loss;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) l += __shfl_xor(l, off);
    if (t == 0) lpart[blockIdx.x] = l;
}

// grad[p] = sum over workgroups of partial[b][p], in workgroup order; loss likewise.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ partial, int nblk,
                                                    const double* __restrict__ lpart,
                                                    float* __restrict__ grad,
                                                    double* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kNParam) {
        float g = 0.f;
        for (int b = 0; b < nblk; ++b) g += partial[(size_t)b * kNParam + p];
        grad[p] = g;
    }
    if (p == 0 && loss) {
        double l = 0.0;
        for (int b = 0; b < nblk; ++b) l += lpart[b];
        *loss = l;
    }
}

// Adam (Kingma & Ba 2015) on the flat parameter vector; step = 1, 2, ...
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ P, const float* __restrict__ g,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  int step, float lr, float b1, float b2, float eps) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNParam) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.0f - b1) * gp;
    const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    const float mh = mp / (1.0f - powf(b1, (float)step));
    const float vh = vp / (1.0f - powf(b2, (float)step));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

// ------------------------------------------------------------------ exact pair emitter

// Partners within the cutoff of the sampled list positions i = offset + m * stride (m < M):
// pass 0 counts them into cnt[m]; pass 1 writes records (r, a, b, c, U) at out + 5*(off[m] + .),
// U the FP64 exact pair energy (1e30 for a capped pair).  Every ordered pair (i, j), j != i.
template <typename Real>
__global__ __launch_bounds__(128) void k_emit(const RepConst* __restrict__ RC,
                                              const RepDyn* __restrict__ RD, int r, int pass,
                                              const uint32_t* __restrict__ qx,
                                              const uint32_t* __restrict__ qy,
                                              const uint32_t* __restrict__ qz,
                                              const Real* __restrict__ ux, const Real* __restrict__ uy,
                                              const Real* __restrict__ uz,
                                              const int* __restrict__ cell_start, int stride,
                                              int offset, int M, int* __restrict__ cnt,
                                              const long long* __restrict__ off,
                                              float* __restrict__ out) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const RepConst& C = RC[r];
    const RepDyn& D = RD[r];
    const int i = offset + m * stride;
    if (i >= C.N) { if (pass == 0) cnt[m] = 0; return; }
    const int n0 = D.n[0], n1 = D.n[1], n2 = D.n[2];
    const double s0 = D.s[0], s1 = D.s[1], s2 = D.s[2];
    const int po = C.part_off;
    const int gi = po + i;
    const uint32_t xi = qx[gi], yi = qy[gi], zi = qz[gi];
    double uix = ux[gi], uiy = uy[gi], uiz = uz[gi];
    const double ni = 1.0 / sqrt(uix * uix + uiy * uiy + uiz * uiz);
    const int cx = (int)cell_coord(xi, D.g[0], n0), cy = (int)cell_coord(yi, D.g[1], n1),
              cz = (int)cell_coord(zi, D.g[2], n2);
    const int* cs = cell_start + C.cs_off;
    const GBd gd = make_gbd(C);
    int c = 0;
    long long base = pass ? off[m] : 0;
    for (int e = 0; e < 27; ++e) {
        int ex = cx + (e % 3) - 1, ey = cy + ((e / 3) % 3) - 1, ez = cz + e / 9 - 1;
        ex += ex < 0 ? n0 : 0; ex -= ex >= n0 ? n0 : 0;
        ey += ey < 0 ? n1 : 0; ey -= ey >= n1 ? n1 : 0;
        ez += ez < 0 ? n2 : 0; ez -= ez >= n2 ? n2 : 0;
        const int nb = ex + n0 * (ey + n1 * ez);
        for (int kk = cs[nb]; kk < cs[nb + 1]; ++kk) {
            if (kk == i) continue;
            const int gj = po + kk;
            const double dx = (double)(int32_t)(qx[gj] - xi) * s0;
            const double dy = (double)(int32_t)(qy[gj] - yi) * s1;
            const double dz = (double)(int32_t)(qz[gj] - zi) * s2;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 < C.rc2)) continue;
            if (pass == 1) {
                const double ujx = ux[gj], ujy = uy[gj], ujz = uz[gj];
                bool cap;
                const double u = gb_pair_f64(gd, uix, uiy, uiz, ujx, ujy, ujz, dx, dy, dz, r2, &cap);
                const double rr = sqrt(r2), ir = 1.0 / rr;
                const double nj = 1.0 / sqrt(ujx * ujx + ujy * ujy + ujz * ujz);
                float* o = out + (size_t)(base + c) * kPairRec;
                o[0] = (float)rr;
                o[1] = (float)((dx * uix + dy * uiy + dz * uiz) * ir * ni);
                o[2] = (float)((dx * ujx + dy * ujy + dz * ujz) * ir * nj);
                o[3] = (float)((uix * ujx + uiy * ujy + uiz * ujz) * ni * nj);
                o[4] = cap ? 1.0e30f : (float)u;
            }
            ++c;
        }
    }
    if (pass == 0) cnt[m] = c;
}

}  // namespace mcgb
