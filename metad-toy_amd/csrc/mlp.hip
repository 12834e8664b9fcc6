// Learned bias: a 3 -> 64 -> 64 -> 1 tanh network fitted on the GPU to the
// exact FP64 sum of the hills a run has deposited, so a 3-D CV space needs no
// dense grid.  No reference counterpart (the reference's 3-D uniform grid has
// no gradient, src/grids/GH_t0_uniformGrid.cu:375-382).
//
// Kernels
//   k_mlp_data     sample points around the deposited hills and label them with
//                  the exact FP64 sum of Gaussians (terms with exponent > 40,
//                  i.e. below 4e-18 of a hill height, are skipped)
//   k_mlp_stats    input normalisation and output scale of a training set
//   k_mlp_train    fused forward + backward of a 16-point tile per one-wave
//                  workgroup: hidden 64x64 layer and its two backward products
//                  on BF16 MFMA (v_mfma_f32_16x16x16_bf16, FP32 accumulation),
//                  the 3-wide and 1-wide layers on FP32 VALU; one gradient row
//                  per tile
//   k_mlp_adam     gradient rows summed in a fixed order + Adam update, BF16 copy of W2
//                  (the training loop of a fit is replayed as one captured graph)
//   k_mlp_forward  batched MFMA inference (held-out check, bias evaluation)
//   k_mlp_commit   held-out gate: adopt the fitted network or fall back to the
//                  exact hills sum
//
// MFMA 16x16x16 BF16 operand layout (lane l, r = l % 16, g = l / 16):
//   A[i][k]: lane holds A[r][4g + 0..3];  B[k][j]: lane holds B[4g + 0..3][r];
//   D[i][j]: lane holds D[4g + 0..3][r].
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "engine.h"
#include "mlp_tanh.h"

namespace mdt {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int H = kMlpH;

__device__ __forceinline__ short f2bf(float v) {
    // round to nearest even
    uint32_t u = __float_as_uint(v);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

// counter-based random numbers (SplitMix64 finaliser of seed and index)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t& s) { s = mix64(s); return (double)(s >> 11) * (1.0 / 9007199254740992.0); }
__device__ __forceinline__ double gauss(uint64_t& s) {
    double a = u01(s), b = u01(s);
    if (a < 1e-300) a = 1e-300;
    return sqrt(-2.0 * log(a)) * cos(6.283185307179586 * b);
}

__device__ __forceinline__ float tanh_f(float z) { return mlp_tanh(z); }

// ---------------------------------------------------------------------------
// Training and held-out points with exact FP64 labels (the sum of every hill).
//   mode 0: fresh points (90 % within spread sigma of a random hill, 10 % uniform over
//           the hill box widened by 4 sigma), labelled by every hill;
//   mode 1: the points already in s_out keep their place; the hills deposited since the
//           last fit (stage n_fit .. nhills) are added to their labels;
//   mode 2: fresh points around the hills deposited since the last fit, labelled by
//           every hill.
// ---------------------------------------------------------------------------
constexpr int kHillChunk = 256;
__global__ __launch_bounds__(256) void k_mlp_data(MlpDev m, int w, int64_t n, uint64_t seed, float4* __restrict__ s_out,
                                                   double* __restrict__ v_out, int mode) {
    __shared__ double hc[kHillChunk * 4];
    __shared__ double hsc[kHillChunk];
    const int64_t nh = m.nhills[w];
    const int64_t hprev = m.stage[w].n_fit < nh ? m.stage[w].n_fit : nh;
    const int64_t lab0 = mode == 1 ? hprev : 0;                  // first hill of the label sum
    const int64_t pick0 = mode == 2 && hprev < nh ? hprev : 0;   // first hill sampled around
    const double* Hs = m.hills + (int64_t)w * m.hills_cap * (m.D + 3);
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s[3] = {0, 0, 0};
    if (p < n && mode == 1) {
        const float4 q = s_out[p];
        s[0] = q.x; s[1] = q.y; s[2] = q.z;
    } else if (p < n && nh > 0) {
        uint64_t st = mix64(seed ^ (uint64_t)(p * 0x2545F4914F6CDD1Dull));
        const double pick = u01(st);
        if (pick < 0.9) {
            int64_t h = pick0 + (int64_t)(u01(st) * (double)(nh - pick0));
            if (h >= nh) h = nh - 1;
            for (int d = 0; d < 3; ++d) s[d] = Hs[h * (m.D + 3) + 1 + d] + m.spread * m.sigma[d] * gauss(st);
        } else {
            // uniform over the hills' box widened by 4 sigma
            for (int d = 0; d < 3; ++d) {
                const double lo = m.box_lo_out[3 * w + d], hi = m.box_hi_out[3 * w + d];
                s[d] = lo - 4 * m.sigma[d] + u01(st) * (hi - lo + 8 * m.sigma[d]);
            }
        }
    }
    const double i2s0 = 0.5 / (m.sigma[0] * m.sigma[0]), i2s1 = 0.5 / (m.sigma[1] * m.sigma[1]),
                 i2s2 = 0.5 / (m.sigma[2] * m.sigma[2]);
    double v = 0.0;
    for (int64_t h0 = lab0; h0 < nh; h0 += kHillChunk) {
        __syncthreads();
        for (int k = threadIdx.x; k < kHillChunk; k += blockDim.x) {
            const int64_t h = h0 + k;
            if (h < nh) {
                // a replayed hill keeps its own sigma: its widths are the bias widths times
                // sigma / sigma_0
                const double* rec = Hs + h * (m.D + 3);
                hc[4 * k] = rec[1]; hc[4 * k + 1] = rec[2]; hc[4 * k + 2] = rec[3]; hc[4 * k + 3] = rec[4];
                hsc[k] = rec[5] == m.sigma[0] ? 1.0 : (m.sigma[0] * m.sigma[0]) / (rec[5] * rec[5]);
            } else {
                hc[4 * k + 3] = 0.0; hc[4 * k] = hc[4 * k + 1] = hc[4 * k + 2] = 1e300;
                hsc[k] = 1.0;
            }
        }
        __syncthreads();
        const int lim = nh - h0 < kHillChunk ? (int)(nh - h0) : kHillChunk;
        for (int k = 0; k < lim; ++k) {
            const double a = s[0] - hc[4 * k], b = s[1] - hc[4 * k + 1], c = s[2] - hc[4 * k + 2];
            const double e = (a * a * i2s0 + b * b * i2s1 + c * c * i2s2) * hsc[k];
            if (e < 40.0) v += hc[4 * k + 3] * exp(-e);
        }
    }
    if (p < n) {
        if (mode == 1) {
            v_out[p] += v;
        } else {
            s_out[p] = make_float4((float)s[0], (float)s[1], (float)s[2], 0.0f);
            v_out[p] = v;
        }
    }
}

// box of the hill centres (for sampling) -- one block
__global__ __launch_bounds__(256) void k_mlp_box(MlpDev m, int w) {
    __shared__ double sh[6][256];
    const int64_t nfit = m.nhills[w];
    const double* Hs = m.hills + (int64_t)w * m.hills_cap * (m.D + 3);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t h = threadIdx.x; h < nfit; h += blockDim.x)
        for (int d = 0; d < 3; ++d) {
            const double v = Hs[h * (m.D + 3) + 1 + d];
            lo[d] = fmin(lo[d], v); hi[d] = fmax(hi[d], v);
        }
    for (int d = 0; d < 3; ++d) { sh[d][threadIdx.x] = lo[d]; sh[3 + d][threadIdx.x] = hi[d]; }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int d = 0; d < 3; ++d) {
                sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + o]);
                sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        MlpNet* net = m.stage + w;
        for (int d = 0; d < 3; ++d) {
            double a = nfit > 0 ? sh[d][0] : 0.0, b = nfit > 0 ? sh[3 + d][0] : 0.0;
            m.box_lo_out[3 * w + d] = a;
            m.box_hi_out[3 * w + d] = b;
            net->in_shift[d] = 0.5 * (a + b);
            net->in_scale[d] = 1.0 / (0.5 * (b - a) + 2.0 * m.sigma[d]);
        }
        net->n_fit = nfit;
    }
}

// output scale: max label of the training set -- one block
__global__ __launch_bounds__(256) void k_mlp_stats(MlpDev m, int w, int64_t n, const double* __restrict__ v) {
    __shared__ double sh[256];
    double mx = 0.0;
    for (int64_t p = threadIdx.x; p < n; p += blockDim.x) mx = fmax(mx, v[p]);
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        MlpNet* net = m.stage + w;
        net->out_scale = sh[0] > 1e-300 ? sh[0] : 1.0;
        net->out_shift = 0.0;
    }
}

// ---------------------------------------------------------------------------
// shared forward of one 16-point tile per wave (layer 1 in A layout, layer 2 on MFMA)
// ---------------------------------------------------------------------------
struct TileFwd {
    float h1a[4][4];     // h1[r][16 ks + 4g + t]  (A layout of layer 2)
    float h2[4][4];      // h2[4g + t][16 nb + r]  (D layout)
    float y[4];          // y[4g + t]
};

__device__ __forceinline__ void tile_forward(const float* sW1, const float* sb1, const short* sW2, const short* sW2l,
                                             const float* sb2, const float* sW3, float b3,
                                             const float* xr /* x of point r */,
                                             float* h1s /* [16][64] this wave */, TileFwd& T) {
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    // layer 1: lane computes h1[r][k] for k = 16 ks + 4 g + t
    #pragma unroll
    for (int ks = 0; ks < 4; ++ks)
        #pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int k = 16 * ks + 4 * g + t;
            float z = sb1[k];
            z = __builtin_fmaf(sW1[3 * k], xr[0], z);
            z = __builtin_fmaf(sW1[3 * k + 1], xr[1], z);
            z = __builtin_fmaf(sW1[3 * k + 2], xr[2], z);
            const float h = tanh_f(z);
            T.h1a[ks][t] = h;
            h1s[r * H + k] = h;
        }
    // layer 2: z2[p][n] = sum_k h1[p][k] W2[n][k]; B[k][n] = W2[n][k].  Both operands
    // split into BF16 hi + lo; hi*hi + hi*lo + lo*hi keeps ~16 mantissa bits.
    f32x4 acc[4];
    #pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        acc[nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        #pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            bf16x4 ah, al, bh, bl;
            #pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float v = T.h1a[ks][t];
                ah[t] = f2bf(v);
                al[t] = f2bf(v - bf2f(ah[t]));
                bh[t] = sW2[(16 * nb + r) * H + 16 * ks + 4 * g + t];
                bl[t] = sW2l[(16 * nb + r) * H + 16 * ks + 4 * g + t];
            }
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc[nb], 0, 0, 0);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc[nb], 0, 0, 0);
            acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc[nb], 0, 0, 0);
        }
    }
    // h2 and the output: lane holds p = 4g + t, n = 16 nb + r
    float part[4] = {0, 0, 0, 0};
    #pragma unroll
    for (int nb = 0; nb < 4; ++nb)
        #pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int n = 16 * nb + r;
            const float h = tanh_f(acc[nb][t] + sb2[n]);
            T.h2[nb][t] = h;
            part[t] = __builtin_fmaf(sW3[n], h, part[t]);
        }
    #pragma unroll
    for (int t = 0; t < 4; ++t) {
        float v = part[t];
        #pragma unroll
        for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
        T.y[t] = v + b3;
    }
}

// The network's parameters to LDS by an NT-thread workgroup, every load issued before
// the first LDS store (one memory round trip); W2's BF16 halves move as 32-bit words
// (sW2 and sW2l 4-byte aligned).
template <int NT>
__device__ __forceinline__ void load_net(const MlpNet* net, float* sW1, float* sb1, short* sW2, short* sW2l, float* sb2,
                                         float* sW3) {
    static_assert(NT >= H && (H * H / 2) % NT == 0, "layout assumed below");
    constexpr int NW2 = H * H / 2 / NT, NW1 = (H * 3 + NT - 1) / NT;
    const int t = threadIdx.x;
    const uint32_t* gh = reinterpret_cast<const uint32_t*>(net->w2bf);
    const uint32_t* gl = reinterpret_cast<const uint32_t*>(net->w2lo);
    uint32_t vh[NW2], vl[NW2];
    #pragma unroll
    for (int u = 0; u < NW2; ++u) { vh[u] = gh[t + u * NT]; vl[u] = gl[t + u * NT]; }
    float w1[NW1];
    #pragma unroll
    for (int u = 0; u < NW1; ++u) w1[u] = t + u * NT < H * 3 ? net->p[kW1 + t + u * NT] : 0.0f;
    float b1 = 0.0f, b2 = 0.0f, w3 = 0.0f;
    if (t < H) { b1 = net->p[kB1 + t]; b2 = net->p[kB2 + t]; w3 = net->p[kW3 + t]; }
    uint32_t* dh = reinterpret_cast<uint32_t*>(sW2);
    uint32_t* dl = reinterpret_cast<uint32_t*>(sW2l);
    #pragma unroll
    for (int u = 0; u < NW2; ++u) { dh[t + u * NT] = vh[u]; dl[t + u * NT] = vl[u]; }
    #pragma unroll
    for (int u = 0; u < NW1; ++u) if (t + u * NT < H * 3) sW1[t + u * NT] = w1[u];
    if (t < H) { sb1[t] = b1; sb2[t] = b2; sW3[t] = w3; }
}

// ---------------------------------------------------------------------------
// k_mlp_train: forward + backward of one 16-point tile per workgroup (one wave);
// the tile's dL/dparams for L = (1/B) sum (y - t)^2 is written as gradient row
// blockIdx.x (every parameter once), and the Adam kernel adds the rows in order.
// One wave per tile gives the step ceil(B / 16) workgroups to spread over the CUs.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_mlp_train(MlpDev m, int w, int64_t first, int64_t count, int64_t n_total,
                                                   const float4* __restrict__ s_in, const double* __restrict__ v_in,
                                                   float* __restrict__ gpart) {
    __shared__ float sW1[H * 3], sb1[H], sb2[H], sW3[H];
    __shared__ __attribute__((aligned(16))) short sW2[H * H];
    __shared__ __attribute__((aligned(16))) short sW2l[H * H];
    __shared__ float h1s[16 * H];
    __shared__ float d2s[16 * H];
    __shared__ float xs[16][4];
    const MlpNet* net = m.stage + w;
    // one training step per launch: its Adam update (the next kernel) reads the advanced
    // step count, so a captured epoch replays with the right step and learning rate
    if (blockIdx.x == 0 && threadIdx.x == 0 && m.adam_t) m.adam_t[w] += 1;
    const int l = threadIdx.x, r = l & 15, g = l >> 4;
    const int64_t p0 = first + (int64_t)blockIdx.x * kMlpTile;
    // the tile's points and the labels of the lane's four points are loaded together
    // with the network
    float4 sv = make_float4(0, 0, 0, 0);
    if (l < 16 && p0 + l < first + count) sv = s_in[(p0 + l) % n_total];
    double lab[4];
    #pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int64_t p = p0 + 4 * g + t;
        lab[t] = p < first + count ? v_in[p % n_total] : 0.0;
    }
    load_net<64>(net, sW1, sb1, sW2, sW2l, sb2, sW3);
    // normalised inputs of the tile
    if (l < 16) {
        const bool ok = p0 + l < first + count;
        xs[l][0] = ok ? (float)((sv.x - net->in_shift[0]) * net->in_scale[0]) : 0.0f;
        xs[l][1] = ok ? (float)((sv.y - net->in_shift[1]) * net->in_scale[1]) : 0.0f;
        xs[l][2] = ok ? (float)((sv.z - net->in_shift[2]) * net->in_scale[2]) : 0.0f;
        xs[l][3] = ok ? 1.0f : 0.0f;   // point weight (0 pads the last tile)
    }
    __syncthreads();
    TileFwd T;
    tile_forward(sW1, sb1, sW2, sW2l, sb2, sW3, net->p[kB3], xs[r], h1s, T);
    __syncthreads();   // h1s written by other lanes
    // dL/dy
    const float inv_b = 2.0f / (float)count;
    const float osc = (float)(1.0 / net->out_scale);
    float dy[4];
    #pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int pl = 4 * g + t;
        const float wt = xs[pl][3];
        const float tgt = wt > 0.0f ? (float)(lab[t] * (double)osc) : 0.0f;
        dy[t] = wt * inv_b * (T.y[t] - tgt);
    }
    // delta2[p][n] = dy[p] W3[n] (1 - h2^2); dW3, db2, db3
    float d2[4][4];
    float gW3[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3 = 0.0f;
    #pragma unroll
    for (int nb = 0; nb < 4; ++nb)
        #pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int n = 16 * nb + r;
            const float h = T.h2[nb][t];
            gW3[nb] = __builtin_fmaf(dy[t], h, gW3[nb]);
            d2[nb][t] = dy[t] * sW3[n] * (1.0f - h * h);
            gb2[nb] += d2[nb][t];
            d2s[(4 * g + t) * H + n] = d2[nb][t];
        }
    #pragma unroll
    for (int t = 0; t < 4; ++t) gb3 += dy[t];
    // sum over the 4 point groups (lanes r, r+16, r+32, r+48)
    #pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        gW3[nb] += __shfl_xor(gW3[nb], 16, 64); gW3[nb] += __shfl_xor(gW3[nb], 32, 64);
        gb2[nb] += __shfl_xor(gb2[nb], 16, 64); gb2[nb] += __shfl_xor(gb2[nb], 32, 64);
    }
    gb3 += __shfl_xor(gb3, 16, 64); gb3 += __shfl_xor(gb3, 32, 64);
    __syncthreads();   // d2s written by other lanes
    float* out = gpart + (int64_t)blockIdx.x * kMlpP;
    // dW2[n][k] = sum_p delta2[p][n] h1[p][k]: A[n][p] (lane: n = 16nb + r, p = 4g+t), B[p][k] = h1[p][k]
    #pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        bf16x4 a;
        #pragma unroll
        for (int t = 0; t < 4; ++t) a[t] = f2bf(d2[nb][t]);
        #pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            bf16x4 b;
            #pragma unroll
            for (int t = 0; t < 4; ++t) b[t] = f2bf(h1s[(4 * g + t) * H + 16 * kb + r]);
            const f32x4 gw = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, f32x4{0, 0, 0, 0}, 0, 0, 0);
            #pragma unroll
            for (int t = 0; t < 4; ++t) out[kW2 + (16 * nb + 4 * g + t) * H + 16 * kb + r] = gw[t];
        }
    }
    // delta1[p][k] = (sum_n delta2[p][n] W2[n][k]) (1 - h1^2): A[p][n] from d2s, B[n][k] = W2[n][k]
    #pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
        f32x4 acc = {0, 0, 0, 0};
        #pragma unroll
        for (int ns = 0; ns < 4; ++ns) {
            bf16x4 a, b;
            #pragma unroll
            for (int t = 0; t < 4; ++t) {
                a[t] = f2bf(d2s[r * H + 16 * ns + 4 * g + t]);
                b[t] = sW2[(16 * ns + 4 * g + t) * H + 16 * kb + r];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        // lane holds S[p = 4g + t][k = 16 kb + r]
        float gb1 = 0.0f, gW1[3] = {0.0f, 0.0f, 0.0f};
        #pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int pl = 4 * g + t;
            const float h = h1s[pl * H + 16 * kb + r];
            const float d1 = acc[t] * (1.0f - h * h);
            gb1 += d1;
            gW1[0] = __builtin_fmaf(d1, xs[pl][0], gW1[0]);
            gW1[1] = __builtin_fmaf(d1, xs[pl][1], gW1[1]);
            gW1[2] = __builtin_fmaf(d1, xs[pl][2], gW1[2]);
        }
        gb1 += __shfl_xor(gb1, 16, 64); gb1 += __shfl_xor(gb1, 32, 64);
        for (int d = 0; d < 3; ++d) { gW1[d] += __shfl_xor(gW1[d], 16, 64); gW1[d] += __shfl_xor(gW1[d], 32, 64); }
        if (g == 0) {
            const int k = 16 * kb + r;
            out[kB1 + k] = gb1;
            out[kW1 + 3 * k] = gW1[0];
            out[kW1 + 3 * k + 1] = gW1[1];
            out[kW1 + 3 * k + 2] = gW1[2];
        }
    }
    if (g == 0) {
        #pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            out[kW3 + 16 * nb + r] = gW3[nb];
            out[kB2 + 16 * nb + r] = gb2[nb];
        }
        if (r == 0) out[kB3] = gb3;
    }
}

// sum of the gradient rows (fixed order) + Adam.  A workgroup covers 64 parameters:
// thread t sums the rows b = q, q+16, ... (q = t / 64) of parameter t % 64 with
// independent loads (sixteen in flight), and the sixteen group sums are added in order.
// step: 1-based Adam step within the current fit (the moments restart with each fit);
// nit > 0: the step is the device counter m.adam_t[w] and the rate lr follows the cosine
// schedule lr (0.01 + 0.99 (1 + cos(pi (step - 1) / nit)) / 2) over the nit steps of the fit
constexpr int kAdamThreads = 1024;
__global__ __launch_bounds__(kAdamThreads) void k_mlp_adam(MlpDev m, int w, int nblocks, const float* __restrict__ gpart,
                                                   float lr, float* __restrict__ gsum_out, int64_t step, int64_t nit) {
    if (nit > 0) {
        step = m.adam_t[w];
        lr = lr * (float)(0.01 + 0.99 * 0.5 * (1.0 + cos(3.141592653589793 * (double)(step - 1) / (double)nit)));
    }
    constexpr int NQ = kAdamThreads / 64;
    __shared__ float part[NQ][64];
    MlpNet* net = m.stage + w;
    float* mo = m.adam_m + (int64_t)w * kMlpP;
    float* ve = m.adam_v + (int64_t)w * kMlpP;
    const int pl = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + pl;
    // the moments and the parameter are loaded together with the gradient rows
    float m0 = 0.0f, v0 = 0.0f, p0 = 0.0f;
    if (q == 0 && k < kMlpP && lr > 0.0f) { m0 = mo[k]; v0 = ve[k]; p0 = net->p[k]; }
    float acc = 0.0f;
    if (k < kMlpP) {
        constexpr int U = 16;
        for (int b0 = q; b0 < nblocks; b0 += NQ * U) {
            float x[U];
            #pragma unroll
            for (int u = 0; u < U; ++u) {
                const int b = b0 + NQ * u;
                x[u] = b < nblocks ? gpart[(int64_t)b * kMlpP + k] : 0.0f;
            }
            #pragma unroll
            for (int u = 0; u < U; ++u) acc += x[u];
        }
    }
    part[q][pl] = acc;
    __syncthreads();
    if (q == 0 && k < kMlpP) {
        float gs = part[0][pl];
        #pragma unroll
        for (int u = 1; u < NQ; ++u) gs += part[u][pl];
        if (gsum_out) gsum_out[k] = gs;
        if (lr > 0.0f) {
            const float b1 = 0.9f, b2 = 0.999f;
            float mm = b1 * m0 + (1.0f - b1) * gs;
            float vv = b2 * v0 + (1.0f - b2) * gs * gs;
            mo[k] = mm;
            ve[k] = vv;
            const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
            const float upd = lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
            const float pn = p0 - upd;
            net->p[k] = pn;
            if (k >= kW2 && k < kW2 + H * H) {
                const short hi = f2bf(pn);
                net->w2bf[k - kW2] = hi;
                net->w2lo[k - kW2] = f2bf(pn - bf2f(hi));
            }
        }
    }
}

// batched MFMA inference of a network on n raw points; optional error vs exact labels
__global__ __launch_bounds__(256) void k_mlp_forward(const MlpNet* __restrict__ net, int64_t n,
                                                      const float4* __restrict__ s_in, const double* __restrict__ v_ref,
                                                      double* __restrict__ v_out, double* __restrict__ blk_err) {
    __shared__ float sW1[H * 3], sb1[H], sb2[H], sW3[H];
    __shared__ __attribute__((aligned(16))) short sW2[H * H];
    __shared__ __attribute__((aligned(16))) short sW2l[H * H];
    __shared__ float h1s[4][16 * H];
    __shared__ float xs[4][16][4];
    __shared__ double red[256];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    const int64_t p0 = (int64_t)blockIdx.x * 64 + wv * 16;
    // the tile's points are loaded together with the network
    float4 sv = make_float4(0, 0, 0, 0);
    if (l < 16 && p0 + l < n) sv = s_in[p0 + l];
    load_net<256>(net, sW1, sb1, sW2, sW2l, sb2, sW3);
    if (l < 16) {
        xs[wv][l][0] = (float)((sv.x - net->in_shift[0]) * net->in_scale[0]);
        xs[wv][l][1] = (float)((sv.y - net->in_shift[1]) * net->in_scale[1]);
        xs[wv][l][2] = (float)((sv.z - net->in_shift[2]) * net->in_scale[2]);
    }
    __syncthreads();
    TileFwd T;
    tile_forward(sW1, sb1, sW2, sW2l, sb2, sW3, net->p[kB3], xs[wv][r], h1s[wv], T);
    double err = 0.0;
    if (r == 0) {
        #pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int64_t p = p0 + 4 * g + t;
            if (p < n) {
                const double V = (double)T.y[t] * net->out_scale + net->out_shift;
                if (v_out) v_out[p] = V;
                if (v_ref) err = fmax(err, fabs(V - v_ref[p]));
            }
        }
    }
    if (blk_err) {
        red[threadIdx.x] = err;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
            __syncthreads();
        }
        if (threadIdx.x == 0) blk_err[blockIdx.x] = red[0];
    }
}

// held-out gate: adopt the staged network when max |dV| <= tol
__global__ __launch_bounds__(256) void k_mlp_commit(MlpDev m, int w, int nblk, const double* __restrict__ blk_err,
                                                     double tol) {
    __shared__ double red[256];
    double e = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) e = fmax(e, blk_err[b]);
    red[threadIdx.x] = e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    const double err = red[0];
    const bool ok = err <= tol && m.stage[w].n_fit > 0;
    // copy the staged network to the live slot (words, all threads)
    if (ok) {
        const int nw = (int)(sizeof(MlpNet) / 4);
        const int* src = reinterpret_cast<const int*>(m.stage + w);
        int* dst = reinterpret_cast<int*>(m.live + w);
        for (int k = threadIdx.x; k < nw; k += blockDim.x) dst[k] = src[k];
    }
    if (threadIdx.x == 0) {
        m.ok[w] = ok ? 1 : 0;
        m.err[w] = err;
        m.fits[w] += 1;
        if (!ok) m.fallbacks[w] += 1;
    }
}

__global__ void k_mlp_copy_net(const MlpNet* __restrict__ src, MlpNet* __restrict__ dst) {
    const int nw = (int)(sizeof(MlpNet) / 4);
    const int* s = reinterpret_cast<const int*>(src);
    int* d = reinterpret_cast<int*>(dst);
    for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < nw; k += blockDim.x * gridDim.x) d[k] = s[k];
}

__global__ void k_mlp_init(MlpNet* net, uint64_t seed) {
    for (int k = threadIdx.x + blockIdx.x * blockDim.x; k < kMlpP; k += blockDim.x * gridDim.x) {
        uint64_t st = mix64(seed + (uint64_t)k * 0x9E3779B97F4A7C15ull);
        double z = gauss(st);
        float v;
        if (k < kB1) v = (float)(z * 1.0);                       // W1: inputs are O(1)
        else if (k < kW2) v = 0.0f;                              // b1
        else if (k < kB2) v = (float)(z / 8.0);                  // W2: 1/sqrt(64)
        else if (k < kW3) v = 0.0f;                              // b2
        else if (k < kB3) v = (float)(z / 8.0);                  // W3
        else v = 0.0f;                                           // b3
        net->p[k] = v;
        if (k >= kW2 && k < kB2) {
            const short hi = f2bf(v);
            net->w2bf[k - kW2] = hi;
            net->w2lo[k - kW2] = f2bf(v - bf2f(hi));
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t mlp_data(const MlpDev& m, int w, int64_t n, uint64_t seed, float4* s, double* v, int mode, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_data, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, m, w, n, seed, s, v, mode);
    return hipGetLastError();
}
hipError_t mlp_box(const MlpDev& m, int w, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_box, dim3(1), dim3(256), 0, st, m, w);
    return hipGetLastError();
}
hipError_t mlp_stats(const MlpDev& m, int w, int64_t n, const double* v, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_stats, dim3(1), dim3(256), 0, st, m, w, n, v);
    return hipGetLastError();
}
hipError_t mlp_train_step(const MlpDev& m, int w, int64_t first, int64_t count, int64_t n_total, const float4* s,
                          const double* v, float* gpart, float lr, float* gsum, int64_t step, int64_t nit,
                          hipStream_t st) {
    const int nb = (int)mlp_grad_rows(count);
    hipLaunchKernelGGL(k_mlp_train, dim3(nb), dim3(64), 0, st, m, w, first, count, n_total, s, v, gpart);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpP + 63) / 64), dim3(kAdamThreads), 0, st, m, w, nb, gpart, lr, gsum, step, nit);
    return hipGetLastError();
}
hipError_t mlp_forward(const MlpNet* net, int64_t n, const float4* s, const double* vref, double* vout,
                       double* blk_err, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st, net, n, s, vref, vout, blk_err);
    return hipGetLastError();
}
hipError_t mlp_commit(const MlpDev& m, int w, int nblk, const double* blk_err, double tol, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_commit, dim3(1), dim3(256), 0, st, m, w, nblk, blk_err, tol);
    return hipGetLastError();
}
hipError_t mlp_copy_net(const MlpNet* src, MlpNet* dst, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_copy_net, dim3(8), dim3(256), 0, st, src, dst);
    return hipGetLastError();
}
hipError_t mlp_init(MlpNet* net, uint64_t seed, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_init, dim3(18), dim3(256), 0, st, net, seed);
    return hipGetLastError();
}

}  // namespace mdt
