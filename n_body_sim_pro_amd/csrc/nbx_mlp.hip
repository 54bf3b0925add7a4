// nbx_mlp.hip - learned Barnes-Hut residual correction (include/nbx_learned.h).
//
// MLP 16 -> 64 -> 64 -> 3 (SiLU, SiLU, linear) on v_mfma_f32_16x16x16_bf16:
// BF16 weights and activations, FP32 accumulation, FP32 master parameters.
//
// Layout. Every layer is computed transposed, Z^T = W * X^T, with 16 samples
// per MFMA column block: the accumulator of a 16x16x16 MFMA holds, in lane l,
// sample (l % 16) and rows 4*(l/16) .. 4*(l/16)+3, which is exactly the
// B-operand layout (column = sample, k = 4*(l/16)+v) the next layer needs, so
// the forward pass never shuffles data. The weight-gradient products sum over
// samples (K = samples) and take their operands through a 16x16 LDS transpose.
//
// Rounding points (mirrored by the CPU oracle in oracle/mlp_oracle.py):
//   xb = bf16((f - mean) * inv_std), inv_std = float32(1 / std);
//   z1 = W1b xb + b1; h1b = bf16(silu(z1));
//   z2 = W2b h1b + b2; h2b = bf16(silu(z2)); y = W3b h2b + b3;
//   dy = 2 (y - t) / B; dyb = bf16(dy); gW3 = dyb h2b^T; dh2 = W3b^T dyb;
//   dz2 = dh2 silu'(z2); dz2b = bf16(dz2); gW2 = dz2b h1b^T; dh1 = W2b^T dz2b;
//   dz1 = dh1 silu'(z1); dz1b = bf16(dz1); gW1 = dz1b xb^T;
//   bias gradients sum the FP32 dy, dz2, dz1.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "nbx_internal.h"
#include "nbx_learned.h"

namespace nbx {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kF = NBX_MLP_FEATURES;  // 16
constexpr int kH = NBX_MLP_HIDDEN;    // 64
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + kH * kF;
constexpr int OFF_W2 = OFF_B1 + kH;
constexpr int OFF_B2 = OFF_W2 + kH * kH;
constexpr int OFF_W3 = OFF_B2 + kH;
constexpr int OFF_B3 = OFF_W3 + 16 * kH;
constexpr int kParams = OFF_B3 + 16;  // 6288
constexpr int kPartial = kParams + 1;  // + loss
constexpr int kMlpThreads = 256;
constexpr int kWavesPerBlock = kMlpThreads / kWave;

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ inline short f2bf(float x) {
    return __builtin_bit_cast(short, static_cast<__bf16>(x));
}

__device__ inline float bf2f(short h) {
    return __uint_as_float(static_cast<uint32_t>(static_cast<unsigned short>(h)) << 16);
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ inline float sigmoidf_(float z) {
    return __builtin_amdgcn_rcpf(1.f + __expf(-z));
}

__device__ inline bf16x4 pack(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = f2bf(a);
    r[1] = f2bf(b);
    r[2] = f2bf(c);
    r[3] = f2bf(d);
    return r;
}

// A fragment of a row-major matrix M (rows x cols, leading dim ld): rows
// r0 + (lane % 16), columns c0 + 4*(lane/16) .. +3.
__device__ inline bf16x4 frag_rows(const float* __restrict__ M, int ld, int r0, int c0, int lane) {
    const float* p = M + (r0 + (lane & 15)) * ld + c0 + 4 * (lane >> 4);
    return pack(p[0], p[1], p[2], p[3]);
}

// A fragment of the transpose of M: element (i, k) = M[k][i] for
// i = r0 + (lane % 16), k = c0 + 4*(lane/16) .. +3.
__device__ inline bf16x4 frag_cols(const float* __restrict__ M, int ld, int r0, int c0, int lane) {
    const int i = r0 + (lane & 15);
    const int k = c0 + 4 * (lane >> 4);
    return pack(M[k * ld + i], M[(k + 1) * ld + i], M[(k + 2) * ld + i], M[(k + 3) * ld + i]);
}


// Accumulator-layout bias: rows r0 + 4*(lane/16) + v.
__device__ inline f32x4 bias4(const float* __restrict__ b, int r0, int lane) {
    const float* p = b + r0 + 4 * (lane >> 4);
    f32x4 c = {p[0], p[1], p[2], p[3]};
    return c;
}

struct Weights {
    bf16x4 w1[4];       // A of layer 1: W1 rows (hidden tile mt), cols = features
    bf16x4 w2[4][4];    // A of layer 2: W2[mt][ks]
    bf16x4 w3[4];       // A of layer 3: W3 (outputs) x hidden k-step
    f32x4 b1[4], b2[4], b3;   // biases in accumulator layout
};

__device__ inline void load_weights(const float* __restrict__ P, int lane, Weights& w) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        w.w1[mt] = frag_rows(P + OFF_W1, kF, 16 * mt, 0, lane);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            w.w2[mt][ks] = frag_rows(P + OFF_W2, kH, 16 * mt, 16 * ks, lane);
        }
        w.w3[mt] = frag_rows(P + OFF_W3, kH, 0, 16 * mt, lane);
        w.b1[mt] = bias4(P + OFF_B1, 16 * mt, lane);
        w.b2[mt] = bias4(P + OFF_B2, 16 * mt, lane);
    }
    w.b3 = bias4(P + OFF_B3, 0, lane);
}

// Standardised BF16 input fragment: sample (lane % 16) of the group, features
// 4*(lane/16) .. +3; also returns the FP32 values for the LDS transpose.
__device__ inline bf16x4 load_x(const float* __restrict__ feats, const int* __restrict__ index,
                                int sample, bool valid, const float* __restrict__ mean,
                                const float* __restrict__ inv_std, int lane, float xs[4]) {
    const int f0 = 4 * (lane >> 4);
    const int s = index != nullptr && valid ? index[sample] : sample;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const float raw = valid ? feats[static_cast<size_t>(s) * kF + f0 + v] : 0.f;
        xs[v] = valid ? (raw - mean[f0 + v]) * inv_std[f0 + v] : 0.f;
    }
    return pack(xs[0], xs[1], xs[2], xs[3]);
}

struct Forward {
    f32x4 z1[4], z2[4];
    bf16x4 h1[4], h2[4];
    f32x4 y;
};

__device__ inline void forward(const Weights& w, bf16x4 xb, Forward& f) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f.z1[mt] = mfma(w.w1[mt], xb, w.b1[mt]);
        f.h1[mt] = pack(f.z1[mt][0] * sigmoidf_(f.z1[mt][0]), f.z1[mt][1] * sigmoidf_(f.z1[mt][1]),
                        f.z1[mt][2] * sigmoidf_(f.z1[mt][2]), f.z1[mt][3] * sigmoidf_(f.z1[mt][3]));
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f32x4 acc = w.b2[mt];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            acc = mfma(w.w2[mt][ks], f.h1[ks], acc);
        }
        f.z2[mt] = acc;
        f.h2[mt] = pack(acc[0] * sigmoidf_(acc[0]), acc[1] * sigmoidf_(acc[1]),
                        acc[2] * sigmoidf_(acc[2]), acc[3] * sigmoidf_(acc[3]));
    }
    f32x4 y = w.b3;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        y = mfma(w.w3[ks], f.h2[ks], y);
    }
    f.y = y;
}

// Inference: raw features (n x 16) -> y (n x 3).
__global__ __launch_bounds__(kMlpThreads) void mlp_forward_kernel(
    const float* __restrict__ P, const float* __restrict__ mean, const float* __restrict__ inv_std,
    const float* __restrict__ feats, int n, float* __restrict__ out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kMlpThreads + threadIdx.x) / kWave;
    const int waves = gridDim.x * kWavesPerBlock;
    Weights w;
    load_weights(P, lane, w);
    const int groups = (n + 15) / 16;
    for (int g = wave; g < groups; g += waves) {
        const int sample = 16 * g + (lane & 15);
        const bool valid = sample < n;
        float xs[4];
        const bf16x4 xb = load_x(feats, nullptr, sample, valid, mean, inv_std, lane, xs);
        Forward f;
        forward(w, xb, f);
        if (valid && lane < 16) {
            out[static_cast<size_t>(sample) * 3 + 0] = f.y[0];
            out[static_cast<size_t>(sample) * 3 + 1] = f.y[1];
            out[static_cast<size_t>(sample) * 3 + 2] = f.y[2];
        }
    }
}

// Orders one wave's LDS writes before its later LDS reads (lanes of a wave
// exchange data through LDS; no other wave takes part).
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Transpose through LDS. `vals` holds a 16x16 matrix M in accumulator layout
// (lane l: M[4*(l/16)+v][l % 16]); the result holds lane l: M[l % 16][4*(l/16)+v],
// which is both the A fragment of M and the B fragment of M^T.
__device__ inline bf16x4 transpose_frag(float* __restrict__ lds, const float vals[4], int lane) {
    const int r0 = 4 * (lane >> 4);
    const int c = lane & 15;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        lds[(r0 + v) * 17 + c] = vals[v];
    }
    wave_lds_sync();
    const float* p = lds + c * 17 + r0;
    const bf16x4 r = pack(p[0], p[1], p[2], p[3]);
    wave_lds_sync();
    return r;
}

__device__ inline void bf_vals(bf16x4 b, float out[4]) {
    out[0] = bf2f(b[0]);
    out[1] = bf2f(b[1]);
    out[2] = bf2f(b[2]);
    out[3] = bf2f(b[3]);
}

// Gradients over `n` samples (optionally through `index`), one FP32 partial
// per wave: [gradients (kParams), loss].
__global__ __launch_bounds__(kMlpThreads) void mlp_grad_kernel(
    const float* __restrict__ P, const float* __restrict__ mean, const float* __restrict__ inv_std,
    const float* __restrict__ feats, const float* __restrict__ targets,
    const int* __restrict__ index, int n, float inv_batch, float* __restrict__ partial) {
    __shared__ float lds_all[kWavesPerBlock][16 * 17];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    float* lds = lds_all[wid];
    const int wave = blockIdx.x * kWavesPerBlock + wid;
    const int waves = gridDim.x * kWavesPerBlock;
    Weights w;
    load_weights(P, lane, w);
    // transposed weights for the backward products
    bf16x4 w3t[4], w2t[4][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        w3t[mt] = frag_cols(P + OFF_W3, kH, 16 * mt, 0, lane);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            w2t[mt][ks] = frag_cols(P + OFF_W2, kH, 16 * mt, 16 * ks, lane);
        }
    }
    f32x4 gw1[4], gw2[4][4], gw3[4];
    float gb1[4][4], gb2[4][4], gb3[4];
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gw1[a] = zero;
        gw3[a] = zero;
        gb3[a] = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            gw2[a][b] = zero;
            gb1[a][b] = 0.f;
            gb2[a][b] = 0.f;
        }
    }
    float loss = 0.f;
    const int groups = (n + 15) / 16;
    for (int g = wave; g < groups; g += waves) {
        const int sample = 16 * g + (lane & 15);
        const bool valid = sample < n;
        float xs[4];
        const bf16x4 xb = load_x(feats, index, sample, valid, mean, inv_std, lane, xs);
        Forward f;
        forward(w, xb, f);
        // dL/dy for outputs o = 4*(lane/16)+v < 3 of sample lane % 16
        float dy[4];
        const int s = index != nullptr && valid ? index[sample] : sample;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int o = 4 * (lane >> 4) + v;
            float d = 0.f;
            if (valid && o < 3) {
                const float diff = f.y[v] - targets[static_cast<size_t>(s) * 3 + o];
                loss += diff * diff;
                d = 2.f * diff * inv_batch;
            }
            dy[v] = d;
            gb3[v] += d;
        }
        const bf16x4 dyb = pack(dy[0], dy[1], dy[2], dy[3]);
        float t4[4];
        // gW3[o][h] += sum_p dyb[o][p] h2b[p][h]
        bf_vals(dyb, t4);
        const bf16x4 dy_a = transpose_frag(lds, t4, lane);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            bf_vals(f.h2[nt], t4);
            gw3[nt] = mfma(dy_a, transpose_frag(lds, t4, lane), gw3[nt]);
        }
        // dz2 = (W3b^T dyb) * silu'(z2)
        bf16x4 dz2b[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const f32x4 dh = mfma(w3t[mt], dyb, zero);
            float dz[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float z = f.z2[mt][v];
                const float sg = sigmoidf_(z);
                dz[v] = valid ? dh[v] * (sg + z * sg * (1.f - sg)) : 0.f;
                gb2[mt][v] += dz[v];
            }
            dz2b[mt] = pack(dz[0], dz[1], dz[2], dz[3]);
        }
        // gW2[h2][h1] += sum_p dz2b[h2][p] h1b[p][h1]
        bf16x4 h1_b[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            bf_vals(f.h1[nt], t4);
            h1_b[nt] = transpose_frag(lds, t4, lane);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            bf_vals(dz2b[mt], t4);
            const bf16x4 dz_a = transpose_frag(lds, t4, lane);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                gw2[mt][nt] = mfma(dz_a, h1_b[nt], gw2[mt][nt]);
            }
        }
        // dz1 = (W2b^T dz2b) * silu'(z1);  gW1[h][f] += sum_p dz1b[h][p] xb[p][f]
        bf_vals(xb, t4);
        const bf16x4 x_b = transpose_frag(lds, t4, lane);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4 dh = zero;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                dh = mfma(w2t[mt][ks], dz2b[ks], dh);
            }
            float dz[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float z = f.z1[mt][v];
                const float sg = sigmoidf_(z);
                dz[v] = valid ? dh[v] * (sg + z * sg * (1.f - sg)) : 0.f;
                gb1[mt][v] += dz[v];
            }
            const bf16x4 dzb = pack(dz[0], dz[1], dz[2], dz[3]);
            bf_vals(dzb, t4);
            gw1[mt] = mfma(transpose_frag(lds, t4, lane), x_b, gw1[mt]);
        }
    }
    // Write this wave's partial: accumulator tiles (lane l: column l % 16,
    // rows 4*(l/16)+v) and bias sums reduced over the 16 samples of a column group.
    float* out = partial + static_cast<size_t>(wave) * kPartial;
    const int col = lane & 15;
    const int r0 = 4 * (lane >> 4);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            out[OFF_W1 + (16 * mt + r0 + v) * kF + col] = gw1[mt][v];
            out[OFF_W3 + (r0 + v) * kH + 16 * mt + col] = gw3[mt][v];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                out[OFF_W2 + (16 * mt + r0 + v) * kH + 16 * nt + col] = gw2[mt][nt][v];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float b1v = gb1[a][v], b2v = gb2[a][v];
            float b3v = a == 0 ? gb3[v] : 0.f;
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) {
                b1v += __shfl_xor(b1v, off);
                b2v += __shfl_xor(b2v, off);
                b3v += __shfl_xor(b3v, off);
            }
            if (col == 0) {
                out[OFF_B1 + 16 * a + r0 + v] = b1v;
                out[OFF_B2 + 16 * a + r0 + v] = b2v;
                if (a == 0) {
                    out[OFF_B3 + r0 + v] = b3v;
                }
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        loss += __shfl_xor(loss, off);
    }
    if (lane == 0) {
        out[kParams] = loss * inv_batch;
    }
}

// Fixed-order sum of the per-wave partials.
__global__ void mlp_reduce_kernel(const float* __restrict__ partial, int waves,
                                  float* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kPartial) {
        return;
    }
    float s = 0.f;
    for (int w = 0; w < waves; ++w) {
        s += partial[static_cast<size_t>(w) * kPartial + k];
    }
    out[k] = s;
}

__global__ void adam_kernel(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kParams) {
        return;
    }
    // rows 3..15 of W3 and entries 3..15 of b3 are padding and stay zero
    const bool pad = (k >= OFF_W3 && k < OFF_B3 && (k - OFF_W3) / kH >= 3) || (k >= OFF_B3 + 3);
    if (pad) {
        P[k] = 0.f;
        return;
    }
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    P[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

__global__ void probe_mfma_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                  float* __restrict__ d) {
    const int lane = threadIdx.x;
    const bf16x4 fa = frag_rows(a, 16, 0, 0, lane);
    const bf16x4 fb = frag_cols(b, 16, 0, 0, lane);   // B[k][j]: j = lane % 16, k = 4*(lane/16)+v
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    const f32x4 r = mfma(fa, fb, zero);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        d[(4 * (lane >> 4) + v) * 16 + (lane & 15)] = r[v];
    }
}

}  // namespace

struct Mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float* params = nullptr;   // kParams
    float* norm = nullptr;     // mean[16], std[16]
    float* grads = nullptr;    // kPartial
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    int adam_step = 0;
    float out_scale = 1.f;
    bool ready = false;
    Buf partial;
    Buf feats, targets, index, scratch_in, scratch_out;
    size_t dataset = 0;
    int cu_count = 256;
    int apply_blocks_per_cu = 1;   // resident learned_apply_kernel workgroups per CU
};

int grad_waves(const Mlp& m, size_t n) {
    const int groups = static_cast<int>((n + 15) / 16);
    const int cap = m.cu_count * 8 * kWavesPerBlock;
    int waves = groups < cap ? groups : cap;
    waves = (waves + kWavesPerBlock - 1) / kWavesPerBlock * kWavesPerBlock;
    return waves > 0 ? waves : kWavesPerBlock;
}

NBodySimProStatus mlp_gradients_device(Mlp& m, const float* feats, const float* targets,
                                       const int* index, size_t n, NBodySimProError* error) {
    const int waves = grad_waves(m, n);
    NBX_TRY(m.partial.need<float>(static_cast<size_t>(waves) * kPartial, error));
    mlp_grad_kernel<<<waves / kWavesPerBlock, kMlpThreads, 0, m.stream>>>(
        m.params, m.norm, m.norm + 2 * kF, feats, targets, index, static_cast<int>(n),
        1.f / static_cast<float>(n), m.partial.as<float>());
    mlp_reduce_kernel<<<grid_for(kPartial, 256), 256, 0, m.stream>>>(m.partial.as<float>(), waves,
                                                                     m.grads);
    NBX_HIP(error, hipGetLastError());
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus mlp_forward_device(Mlp& m, const float* feats, size_t n, float* out,
                                     NBodySimProError* error) {
    const int groups = static_cast<int>((n + 15) / 16);
    int blocks = (groups + kWavesPerBlock - 1) / kWavesPerBlock;
    blocks = blocks < m.cu_count * 8 ? blocks : m.cu_count * 8;
    mlp_forward_kernel<<<blocks > 0 ? blocks : 1, kMlpThreads, 0, m.stream>>>(
        m.params, m.norm, m.norm + 2 * kF, feats, static_cast<int>(n), out);
    NBX_HIP(error, hipGetLastError());
    return N_BODY_SIM_PRO_STATUS_OK;
}

}  // namespace nbx

struct NBXMlp {
    nbx::Mlp m;
};

namespace nbx {

// Walk record (kWalkRecord floats per tree position, G not applied):
//   [0..2] total sum, [3] accepted cells, [4..6] quadrupole force of the
//   accepted cells (sum of 5/2 (d.Q.d) d / r^7 - Q.d / r^5), [7] exact leaf
//   terms, [8] sum and [9] minimum of (distance / cell edge)^2 over accepted
//   cells, [10..15] zero.

// Local frame of one query: e1 along the walk's total acceleration, e2 along
// the quadrupole force perpendicular to e1, e3 = e1 x e2. Features are
// frame-free and the model's outputs are residual components in this frame.
struct Frame {
    float e1[3], e2[3], e3[3];
    float a;        // |total| (G not applied)
    float q_par;    // quadrupole force along e1
    float q_perp;   // |quadrupole force perpendicular to e1|
};

__device__ inline void local_frame(const float* __restrict__ wf, Frame& fr) {
    const float ax = wf[0], ay = wf[1], az = wf[2];
    const float a = sqrtf(ax * ax + ay * ay + az * az);
    fr.a = a;
    if (a > 0.f) {
        const float ia = __builtin_amdgcn_rcpf(a);
        fr.e1[0] = ax * ia;
        fr.e1[1] = ay * ia;
        fr.e1[2] = az * ia;
    } else {
        fr.e1[0] = 1.f;
        fr.e1[1] = 0.f;
        fr.e1[2] = 0.f;
    }
    const float cx = wf[4], cy = wf[5], cz = wf[6];
    const float cp = cx * fr.e1[0] + cy * fr.e1[1] + cz * fr.e1[2];
    float px = cx - cp * fr.e1[0], py = cy - cp * fr.e1[1], pz = cz - cp * fr.e1[2];
    float pn = sqrtf(px * px + py * py + pz * pz);
    fr.q_par = cp;
    fr.q_perp = pn;
    if (!(pn > 1e-7f * a)) {
        // no usable perpendicular quadrupole force: the axis least aligned with e1
        const float ux = fabsf(fr.e1[0]), uy = fabsf(fr.e1[1]), uz = fabsf(fr.e1[2]);
        const bool kx = ux <= uy && ux <= uz;
        const bool ky = !kx && uy <= uz;
        const bool kz = !kx && !ky;
        const float d = kx ? fr.e1[0] : (ky ? fr.e1[1] : fr.e1[2]);
        px = (kx ? 1.f : 0.f) - d * fr.e1[0];
        py = (ky ? 1.f : 0.f) - d * fr.e1[1];
        pz = (kz ? 1.f : 0.f) - d * fr.e1[2];
    }
    // second Gram-Schmidt pass: the first loses orthogonality in FP32 when
    // the perpendicular part is a small difference of large vectors
    const float q = px * fr.e1[0] + py * fr.e1[1] + pz * fr.e1[2];
    px -= q * fr.e1[0];
    py -= q * fr.e1[1];
    pz -= q * fr.e1[2];
    const float ipn = __builtin_amdgcn_rcpf(sqrtf(px * px + py * py + pz * pz));
    fr.e2[0] = px * ipn;
    fr.e2[1] = py * ipn;
    fr.e2[2] = pz * ipn;
    fr.e3[0] = fr.e1[1] * fr.e2[2] - fr.e1[2] * fr.e2[1];
    fr.e3[1] = fr.e1[2] * fr.e2[0] - fr.e1[0] * fr.e2[2];
    fr.e3[2] = fr.e1[0] * fr.e2[1] - fr.e1[1] * fr.e2[0];
}

// Raw features (before standardisation) from one walk record.
__device__ inline void raw_features(const float* __restrict__ wf, const Frame& fr, float m,
                                    int leaf_depth, float log2_root_edge, float f[kF]) {
    const float cells = wf[3], leaves = wf[7];
    const float inv = fr.a > 0.f ? __builtin_amdgcn_rcpf(fr.a) : 0.f;
    const float qn = sqrtf(fr.q_par * fr.q_par + fr.q_perp * fr.q_perp);
    const float iq = qn > 0.f ? __builtin_amdgcn_rcpf(qn) : 0.f;
    f[0] = cells > 0.f ? sqrtf(wf[8] * __builtin_amdgcn_rcpf(cells)) : 0.f;   // rms distance / edge
    f[1] = cells > 0.f ? sqrtf(wf[9]) : 0.f;            // minimum distance / cell edge
    // own mass over the leaf cell volume: edge = 2 h_root 2^-depth
    f[2] = log2f(m > 0.f ? m : 1e-30f) - 3.f * (log2_root_edge - static_cast<float>(leaf_depth));
    f[3] = log2f(1.f + cells);
    f[4] = log2f(1.f + leaves);
    f[5] = fr.q_par * inv;                               // quadrupole force along e1 / |a|
    f[6] = fr.q_perp * inv;                              // and perpendicular to it / |a|
    f[7] = fr.q_par * iq;                                // cosine and sine of the angle
    f[8] = fr.q_perp * iq;                               // between q and a
    f[9] = log2f(qn * inv + 0x1p-40f);                   // log2 |q| / |a|
#pragma unroll
    for (int k = 10; k < kF; ++k) {
        f[k] = 0.f;
    }
}

// Depth of the leaf at tree position t: one below the deeper of the cells it
// shares with its neighbours (delta[t] = common levels of t and t + 1, -1 at
// the ends).
__device__ inline int leaf_depth(const int32_t* __restrict__ delta, int t) {
    const int a = t > 0 ? delta[t - 1] : -1;
    const int b = delta[t];
    return 1 + (a > b ? a : b);
}

__global__ void features_kernel(const float* __restrict__ walkf, const float4* __restrict__ s4,
                                const int32_t* __restrict__ delta,
                                const uint32_t* __restrict__ tperm, int n, float log2_root_edge,
                                float G, double* __restrict__ ax, double* __restrict__ ay,
                                double* __restrict__ az, float* __restrict__ out,
                                float* __restrict__ frames) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) {
        return;
    }
    float f[kF];
    const float* wf = walkf + static_cast<size_t>(t) * kWalkRecord;
    Frame fr;
    local_frame(wf, fr);
    raw_features(wf, fr, s4[t].w, leaf_depth(delta, t), log2_root_edge, f);
    const size_t p = tperm[t];
    // the plain walk's accelerations, bit for bit
    ax[p] = static_cast<double>(G * wf[0]);
    ay[p] = static_cast<double>(G * wf[1]);
    az[p] = static_cast<double>(G * wf[2]);
    float* o = out + p * kF;
#pragma unroll
    for (int k = 0; k < kF; ++k) {
        o[k] = f[k];
    }
    if (frames != nullptr) {
        float* q = frames + p * 9;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            q[k] = fr.e1[k];
            q[3 + k] = fr.e2[k];
            q[6 + k] = fr.e3[k];
        }
    }
}

// Corrected accelerations: a += G |q| (y0 e1 + y1 e2 + y2 e3) / out_scale,
// with a the walk's result and q its quadrupole force (G not applied; no
// accepted cell, no correction). Each wave takes 64 queries per step: every lane
// assembles one query's features into LDS, four 16-query MFMA passes run the
// network, and every lane applies its own query's correction.
__global__ __launch_bounds__(kMlpThreads) void learned_apply_kernel(
    const float* __restrict__ P, const float* __restrict__ mean, const float* __restrict__ inv_std,
    float out_scale, const float* __restrict__ walkf, const float4* __restrict__ s4,
    const int32_t* __restrict__ delta, const uint32_t* __restrict__ tperm, int n,
    float log2_root_edge, float G, double* __restrict__ ax, double* __restrict__ ay,
    double* __restrict__ az) {
    __shared__ float xs_all[kWavesPerBlock][kWave][kF + 1];
    __shared__ float ys_all[kWavesPerBlock][kWave][4];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    float (*xs)[kF + 1] = xs_all[wid];
    float (*ys)[4] = ys_all[wid];
    const int wave = blockIdx.x * kWavesPerBlock + wid;
    const int waves = gridDim.x * kWavesPerBlock;
    Weights w;
    load_weights(P, lane, w);
    const int tiles = (n + kWave - 1) / kWave;
    for (int g = wave; g < tiles; g += waves) {
        const int t = kWave * g + lane;
        const bool valid = t < n;
        Frame fr;
        float f[kF];
        double bx = 0.0, by = 0.0, bz = 0.0;
        if (valid) {
            const float* wf = walkf + static_cast<size_t>(t) * kWalkRecord;
            local_frame(wf, fr);
            raw_features(wf, fr, s4[t].w, leaf_depth(delta, t), log2_root_edge, f);
            // the plain walk's accelerations, bit for bit
            bx = static_cast<double>(G * wf[0]);
            by = static_cast<double>(G * wf[1]);
            bz = static_cast<double>(G * wf[2]);
        } else {
#pragma unroll
            for (int k = 0; k < kF; ++k) f[k] = mean[k];
        }
#pragma unroll
        for (int k = 0; k < kF; ++k) {
            xs[lane][k] = (f[k] - mean[k]) * inv_std[k];
        }
        wave_lds_sync();
        const int f0 = 4 * (lane >> 4);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int p = 16 * q + (lane & 15);
            const bf16x4 xb = pack(xs[p][f0], xs[p][f0 + 1], xs[p][f0 + 2], xs[p][f0 + 3]);
            Forward fw;
            forward(w, xb, fw);
            if (lane < 16) {
                ys[p][0] = fw.y[0];
                ys[p][1] = fw.y[1];
                ys[p][2] = fw.y[2];
            }
        }
        wave_lds_sync();
        if (valid) {
            const float y0 = ys[lane][0], y1 = ys[lane][1], y2 = ys[lane][2];
            const uint32_t o = tperm[t];
            // the walk's quadrupole force q plus the network's estimate of
            // what q leaves, in units of |q|
            const float* wf = walkf + static_cast<size_t>(t) * kWalkRecord;
            const double qn = sqrt(static_cast<double>(fr.q_par) * fr.q_par +
                                   static_cast<double>(fr.q_perp) * fr.q_perp);
            const double k = static_cast<double>(G) * qn / out_scale;
            const double gq = static_cast<double>(G);
            ax[o] = bx + gq * wf[4] + k * (y0 * fr.e1[0] + y1 * fr.e2[0] + y2 * fr.e3[0]);
            ay[o] = by + gq * wf[5] + k * (y0 * fr.e1[1] + y1 * fr.e2[1] + y2 * fr.e3[1]);
            az[o] = bz + gq * wf[6] + k * (y0 * fr.e1[2] + y1 * fr.e2[2] + y2 * fr.e3[2]);
        }
        wave_lds_sync();
    }
}

NBodySimProStatus bh_learned(Engine* e, const NBodySimProGravity* g, NBodySimProError* error) {
    Mlp* m = static_cast<Mlp*>(e->mlp);
    if (m == nullptr || !m->ready) {
        return barnes_hut(e, g, false, error);   // plain BH fallback
    }
    NBX_TRY(bh_build(e, true, error, true));
    BHState& s = e->bh;
    NBX_TRY(e->scratch2.need<float>(s.n * kWalkRecord, error));
    float* walkf = e->scratch2.as<float>();
    NBX_TRY(bh_walk(e, g, false, walkf, error));
    NBX_HIP(error, hipEventRecord(e->ev[7], e->stream));
    // One resident round of workgroups that stride over the 64-query tiles:
    // a grid beyond what fits leaves a second, partial round on some CUs.
    // (Measured on MI355X at 10M: 4 workgroups per CU requested with 3
    // resident, 0.86 ms; the resident count, 0.74 ms. Forcing 4 resident
    // workgroups through a waves-per-EU bound spills 10 VGPRs: 0.81 ms.)
    const int tiles = static_cast<int>((s.n + kWave - 1) / kWave);
    int blocks = (tiles + kWavesPerBlock - 1) / kWavesPerBlock;
    const int resident = m->cu_count * m->apply_blocks_per_cu;
    blocks = blocks < resident ? blocks : resident;
    learned_apply_kernel<<<blocks, kMlpThreads, 0, e->stream>>>(
        m->params, m->norm, m->norm + 2 * kF, m->out_scale, walkf, s.s4.as<float4>(),
        s.delta.as<int32_t>(), s.tperm.as<uint32_t>(), static_cast<int>(s.n),
        static_cast<float>(std::log2(2.0 * s.root[3])),
        static_cast<float>(g->gravitational_constant), e->ax, e->ay, e->az);
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipEventRecord(e->ev[8], e->stream));
    NBX_HIP(error, hipEventSynchronize(e->ev[8]));
    float t[10] = {0};
    for (int k : {1, 3, 4, 6, 8, 9}) {
        NBX_HIP(error, hipEventElapsedTime(&t[k], e->ev[0], e->ev[k]));
    }
    float mlp_ms = 0.f;
    NBX_HIP(error, hipEventElapsedTime(&mlp_ms, e->ev[7], e->ev[8]));
    NBXTimings tm{};
    tm.bbox_keys_ms = t[1];
    tm.sort_ms = t[9] - t[1];
    tm.tree_ms = t[3] - t[9];
    tm.com_ms = t[4] - t[3];
    float walk_ms = 0.f;
    NBX_HIP(error, hipEventElapsedTime(&walk_ms, e->ev[5], e->ev[6]));
    tm.walk_ms = walk_ms;
    tm.force_ms = t[8];
    tm.kernel_ms = walk_ms;
    tm.mlp_ms = mlp_ms;
    const double interactions = static_cast<double>(e->stats.accepted_approximations +
                                                    e->stats.exact_interactions);
    tm.kernel_flops = 20.0 * interactions;
    e->timings = tm;
    e->stats.build_time_seconds = t[4] * 1e-3;
    e->stats.evaluation_time_seconds = (t[8] - t[4]) * 1e-3;
    return N_BODY_SIM_PRO_STATUS_OK;
}

}  // namespace nbx

using nbx::Mlp;

extern "C" {

size_t nbx_mlp_param_count(void) {
    return nbx::kParams;
}

void nbx_mlp_destroy(NBXMlp* p) {
    if (p == nullptr) {
        return;
    }
    Mlp& m = p->m;
    (void)hipSetDevice(m.device);
    if (m.stream != nullptr) (void)hipStreamSynchronize(m.stream);
    float* arrays[] = {m.params, m.norm, m.grads, m.adam_m, m.adam_v};
    for (float* a : arrays) {
        if (a != nullptr) (void)hipFree(a);
    }
    nbx::Buf* bufs[] = {&m.partial, &m.feats, &m.targets, &m.index, &m.scratch_in, &m.scratch_out};
    for (nbx::Buf* b : bufs) {
        if (b->p != nullptr) (void)hipFree(b->p);
    }
    if (m.stream != nullptr) (void)hipStreamDestroy(m.stream);
    delete p;
}

void nbx_mlp_release(void* mlp) {
    (void)mlp;   // the engine does not own the model; nbx_mlp_destroy frees it
}

NBXMlp* nbx_mlp_create(int device, NBodySimProError* error) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_UNSUPPORTED_PLATFORM, __FILE__, __LINE__,
                       "no such HIP device");
        return nullptr;
    }
    NBXMlp* p = new (std::nothrow) NBXMlp();
    if (p == nullptr) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY, __FILE__, __LINE__,
                       "model allocation failed");
        return nullptr;
    }
    Mlp& m = p->m;
    m.device = device;
    hipDeviceProp_t prop;
    bool ok = hipSetDevice(device) == hipSuccess &&
              hipGetDeviceProperties(&prop, device) == hipSuccess &&
              hipStreamCreateWithFlags(&m.stream, hipStreamNonBlocking) == hipSuccess;
    if (ok) m.cu_count = prop.multiProcessorCount;
    int per_cu = 0;
    ok = ok && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, nbx::learned_apply_kernel,
                                                            nbx::kMlpThreads, 0) == hipSuccess;
    if (ok) m.apply_blocks_per_cu = per_cu > 0 ? per_cu : 1;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m.params), nbx::kParams * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m.norm), 3 * nbx::kF * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m.grads), nbx::kPartial * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m.adam_m), nbx::kParams * sizeof(float)) == hipSuccess;
    ok = ok && hipMalloc(reinterpret_cast<void**>(&m.adam_v), nbx::kParams * sizeof(float)) == hipSuccess;
    ok = ok && hipMemset(m.params, 0, nbx::kParams * sizeof(float)) == hipSuccess &&
         hipMemset(m.adam_m, 0, nbx::kParams * sizeof(float)) == hipSuccess &&
         hipMemset(m.adam_v, 0, nbx::kParams * sizeof(float)) == hipSuccess;
    if (ok) {
        // mean, std, 1 / std
        std::vector<float> norm(3 * nbx::kF, 1.f);
        for (int k = 0; k < nbx::kF; ++k) norm[k] = 0.f;
        ok = hipMemcpy(m.norm, norm.data(), norm.size() * sizeof(float), hipMemcpyHostToDevice) ==
             hipSuccess;
    }
    if (!ok) {
        nbx::set_error(error, N_BODY_SIM_PRO_STATUS_OUT_OF_MEMORY, __FILE__, __LINE__,
                       "device allocation failed while creating the model");
        nbx_mlp_destroy(p);
        return nullptr;
    }
    return p;
}

NBodySimProStatus nbx_mlp_set_params(NBXMlp* p, const float* params, const float* mean,
                                     const float* stdv, float out_scale, NBodySimProError* error) {
    if (p == nullptr || params == nullptr || mean == nullptr || stdv == nullptr ||
        !(out_scale > 0.f)) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "invalid model parameters");
    }
    for (int k = 0; k < nbx::kF; ++k) {
        if (!(stdv[k] > 0.f)) {
            NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "std entries must be > 0");
        }
    }
    Mlp& m = p->m;
    NBX_HIP(error, hipSetDevice(m.device));
    NBX_HIP(error, hipMemcpy(m.params, params, nbx::kParams * sizeof(float), hipMemcpyHostToDevice));
    NBX_HIP(error, hipMemcpy(m.norm, mean, nbx::kF * sizeof(float), hipMemcpyHostToDevice));
    NBX_HIP(error, hipMemcpy(m.norm + nbx::kF, stdv, nbx::kF * sizeof(float), hipMemcpyHostToDevice));
    float inv[nbx::kF];
    for (int k = 0; k < nbx::kF; ++k) {
        inv[k] = 1.f / stdv[k];   // +inf -> 0: the feature is ignored
    }
    NBX_HIP(error, hipMemcpy(m.norm + 2 * nbx::kF, inv, sizeof(inv), hipMemcpyHostToDevice));
    NBX_HIP(error, hipMemset(m.adam_m, 0, nbx::kParams * sizeof(float)));
    NBX_HIP(error, hipMemset(m.adam_v, 0, nbx::kParams * sizeof(float)));
    m.adam_step = 0;
    m.out_scale = out_scale;
    m.ready = true;
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_mlp_get_params(NBXMlp* p, float* params, float* mean, float* stdv,
                                     float* out_scale, NBodySimProError* error) {
    if (p == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "model must not be null");
    }
    Mlp& m = p->m;
    NBX_HIP(error, hipSetDevice(m.device));
    NBX_HIP(error, hipStreamSynchronize(m.stream));
    if (params != nullptr) {
        NBX_HIP(error, hipMemcpy(params, m.params, nbx::kParams * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (mean != nullptr) {
        NBX_HIP(error, hipMemcpy(mean, m.norm, nbx::kF * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (stdv != nullptr) {
        NBX_HIP(error, hipMemcpy(stdv, m.norm + nbx::kF, nbx::kF * sizeof(float), hipMemcpyDeviceToHost));
    }
    if (out_scale != nullptr) {
        *out_scale = m.out_scale;
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_mlp_forward(NBXMlp* p, const float* features, size_t n, float* out,
                                  NBodySimProError* error) {
    if (p == nullptr || features == nullptr || out == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "model, features, out required");
    }
    Mlp& m = p->m;
    if (n == 0) {
        return N_BODY_SIM_PRO_STATUS_OK;
    }
    NBX_HIP(error, hipSetDevice(m.device));
    NBX_TRY(m.scratch_in.need<float>(n * nbx::kF, error));
    NBX_TRY(m.scratch_out.need<float>(n * 3, error));
    NBX_HIP(error, hipMemcpyAsync(m.scratch_in.p, features, n * nbx::kF * sizeof(float),
                                  hipMemcpyHostToDevice, m.stream));
    NBX_TRY(nbx::mlp_forward_device(m, m.scratch_in.as<float>(), n, m.scratch_out.as<float>(), error));
    NBX_HIP(error, hipMemcpyAsync(out, m.scratch_out.p, n * 3 * sizeof(float), hipMemcpyDeviceToHost,
                                  m.stream));
    NBX_HIP(error, hipStreamSynchronize(m.stream));
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_mlp_gradients(NBXMlp* p, const float* features, const float* targets,
                                    size_t n, float* grads, float* loss, NBodySimProError* error) {
    if (p == nullptr || features == nullptr || targets == nullptr || n == 0) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "model and data required");
    }
    Mlp& m = p->m;
    NBX_HIP(error, hipSetDevice(m.device));
    NBX_TRY(m.scratch_in.need<float>(n * nbx::kF, error));
    NBX_TRY(m.scratch_out.need<float>(n * 3, error));
    NBX_HIP(error, hipMemcpyAsync(m.scratch_in.p, features, n * nbx::kF * sizeof(float),
                                  hipMemcpyHostToDevice, m.stream));
    NBX_HIP(error, hipMemcpyAsync(m.scratch_out.p, targets, n * 3 * sizeof(float),
                                  hipMemcpyHostToDevice, m.stream));
    NBX_TRY(nbx::mlp_gradients_device(m, m.scratch_in.as<float>(), m.scratch_out.as<float>(), nullptr,
                                      n, error));
    std::vector<float> g(nbx::kPartial);
    NBX_HIP(error, hipMemcpyAsync(g.data(), m.grads, nbx::kPartial * sizeof(float),
                                  hipMemcpyDeviceToHost, m.stream));
    NBX_HIP(error, hipStreamSynchronize(m.stream));
    if (grads != nullptr) {
        std::copy(g.begin(), g.begin() + nbx::kParams, grads);
    }
    if (loss != nullptr) {
        *loss = g[nbx::kParams];
    }
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_mlp_load_dataset(NBXMlp* p, const float* features, const float* targets,
                                       size_t n, NBodySimProError* error) {
    if (p == nullptr || features == nullptr || targets == nullptr || n == 0 || n > 0x7FFFFFF0ull) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "invalid training set");
    }
    Mlp& m = p->m;
    NBX_HIP(error, hipSetDevice(m.device));
    NBX_TRY(m.feats.need<float>(n * nbx::kF, error));
    NBX_TRY(m.targets.need<float>(n * 3, error));
    NBX_TRY(m.index.need<int>(n, error));
    NBX_HIP(error, hipMemcpy(m.feats.p, features, n * nbx::kF * sizeof(float), hipMemcpyHostToDevice));
    NBX_HIP(error, hipMemcpy(m.targets.p, targets, n * 3 * sizeof(float), hipMemcpyHostToDevice));
    m.dataset = n;
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_mlp_train(NBXMlp* p, int epochs, int batch, float lr, uint64_t seed,
                                float* loss_history, NBodySimProError* error) {
    if (p == nullptr || epochs < 0 || batch <= 0 || !(lr > 0.f)) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "invalid training arguments");
    }
    Mlp& m = p->m;
    if (m.dataset == 0) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_STATE, "no training set loaded");
    }
    NBX_HIP(error, hipSetDevice(m.device));
    const size_t n = m.dataset;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        NBX_HIP(error, hipMemcpyAsync(m.index.p, order.data(), n * sizeof(int),
                                      hipMemcpyHostToDevice, m.stream));
        double epoch_loss = 0.0;
        int batches = 0;
        std::vector<float> losses;
        for (size_t start = 0; start < n; start += static_cast<size_t>(batch)) {
            const size_t cnt = std::min(static_cast<size_t>(batch), n - start);
            NBX_TRY(nbx::mlp_gradients_device(m, m.feats.as<float>(), m.targets.as<float>(),
                                              m.index.as<int>() + start, cnt, error));
            ++m.adam_step;
            const float c1 = 1.f - std::pow(b1, static_cast<float>(m.adam_step));
            const float c2 = 1.f - std::pow(b2, static_cast<float>(m.adam_step));
            nbx::adam_kernel<<<nbx::grid_for(nbx::kParams, 256), 256, 0, m.stream>>>(
                m.params, m.adam_m, m.adam_v, m.grads, lr, b1, b2, eps, c1, c2);
            NBX_HIP(error, hipGetLastError());
            if (loss_history != nullptr) {
                float l = 0.f;
                NBX_HIP(error, hipMemcpyAsync(&l, m.grads + nbx::kParams, sizeof(float),
                                              hipMemcpyDeviceToHost, m.stream));
                NBX_HIP(error, hipStreamSynchronize(m.stream));
                epoch_loss += l;
            }
            ++batches;
        }
        if (loss_history != nullptr) {
            loss_history[ep] = static_cast<float>(epoch_loss / batches);
        }
    }
    NBX_HIP(error, hipStreamSynchronize(m.stream));
    m.ready = true;
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_engine_set_mlp(NBXEngine* engine, NBXMlp* mlp, NBodySimProError* error) {
    if (engine == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "engine must not be null");
    }
    if (mlp != nullptr && mlp->m.device != engine->e.device) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT,
                 "model and engine must live on the same device");
    }
    engine->e.mlp = mlp == nullptr ? nullptr : &mlp->m;
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_engine_bh_features(NBXEngine* engine, const NBodySimProGravity* g,
                                         float* features, float* frames, NBodySimProError* error) {
    if (engine == nullptr || g == nullptr || features == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "engine, gravity, out required");
    }
    nbx::Engine* e = &engine->e;
    NBX_HIP(error, hipSetDevice(e->device));
    NBX_TRY(nbx::bh_build(e, true, error, true));
    nbx::BHState& s = e->bh;
    NBX_TRY(e->scratch2.need<float>(s.n * nbx::kWalkRecord, error));
    NBX_TRY(e->scratch.need<float>(s.n * (nbx::kF + 9), error));
    NBX_TRY(nbx::bh_walk(e, g, false, e->scratch2.as<float>(), error));
    float* dfeat = e->scratch.as<float>();
    float* dframe = dfeat + s.n * nbx::kF;
    nbx::features_kernel<<<nbx::grid_for(s.n, 256), 256, 0, e->stream>>>(
        e->scratch2.as<float>(), s.s4.as<float4>(), s.delta.as<int32_t>(),
        s.tperm.as<uint32_t>(), static_cast<int>(s.n),
        static_cast<float>(std::log2(2.0 * s.root[3])),
        static_cast<float>(g->gravitational_constant), e->ax, e->ay, e->az, dfeat,
        frames != nullptr ? dframe : nullptr);
    NBX_HIP(error, hipGetLastError());
    NBX_HIP(error, hipMemcpyAsync(features, dfeat, s.n * nbx::kF * sizeof(float),
                                  hipMemcpyDeviceToHost, e->stream));
    if (frames != nullptr) {
        NBX_HIP(error, hipMemcpyAsync(frames, dframe, s.n * 9 * sizeof(float),
                                      hipMemcpyDeviceToHost, e->stream));
    }
    NBX_HIP(error, hipStreamSynchronize(e->stream));
    return N_BODY_SIM_PRO_STATUS_OK;
}

NBodySimProStatus nbx_mlp_probe_mfma(int device, const float* a, const float* b, float* d,
                                     NBodySimProError* error) {
    if (a == nullptr || b == nullptr || d == nullptr) {
        NBX_FAIL(error, N_BODY_SIM_PRO_STATUS_INVALID_ARGUMENT, "a, b, d required");
    }
    NBX_HIP(error, hipSetDevice(device));
    float* buf = nullptr;
    NBX_HIP(error, hipMalloc(reinterpret_cast<void**>(&buf), 3 * 256 * sizeof(float)));
    NBX_HIP(error, hipMemcpy(buf, a, 256 * sizeof(float), hipMemcpyHostToDevice));
    NBX_HIP(error, hipMemcpy(buf + 256, b, 256 * sizeof(float), hipMemcpyHostToDevice));
    nbx::probe_mfma_kernel<<<1, nbx::kWave>>>(buf, buf + 256, buf + 512);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipMemcpy(d, buf + 512, 256 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    NBX_HIP(error, e1);
    NBX_HIP(error, e2);
    return N_BODY_SIM_PRO_STATUS_OK;
}

}  // extern "C"
