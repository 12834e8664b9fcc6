// gfx950 kernels of the metadynamics CV-and-bias path.
//
// Per MD step (one walker or many, atoms of all walkers concatenated):
//   k_env        N-nearest selection (bit-exact FP64 r^2, reference insertion
//                semantics when the cap binds) fused with FP32 Y_lm / q_lm /
//                q_l for every degree on the environment, FP64 per-tile sums.
//   k_ld         Lechner-Dellago local Q_l from neighbours' q_lm.
//   k_sk_amp     structure-factor amplitudes A(k) (FP64 phase, FP64 sums).
//   k_reduce     per walker: FP64 CV sums, DIM expressions with exact
//                derivatives, hill deposit (well-tempered), bias value and
//                gradient, chain-rule weights dV/dv.
//   k_ld_coeff   two-hop coefficients of the local-Q bias.
//   k_force      exact bias forces by a deterministic gather over each atom's
//                own selected pairs and the reverse pairs (no atomics).
//   k_sk_force   structure-factor bias forces.
//
// Reference counterparts (paths relative to the Metad-toy source tree):
//   selection  src/cvs/CVs_Steinhardt.cu:1324-1406
//   q_lm/q_l   src/cvs/CVs_Steinhardt.cu:1410-1532, include/CVsets/CV_Steinhardt_math.h:981-1215
//   gradient   src/cvs/CVs_Steinhardt.cu:1538-1698 (replaced: exact gradient)
//   grid       src/grids/GH_t0_uniformGrid.cu:194-461
//   chain rule src/grids/DimSet.cu:68-97
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <utility>

#include "engine.h"
#include "mlp_tanh.h"

namespace mdt {
namespace {

// ---------------------------------------------------------------------------
// Compile-time tables for Y_lm(n) = c_lm(n_z) (n_x + i n_y)^m with
// c_lm(u) = N_lm (-1)^m d^m P_l(u)/du^m (orthonormal, Condon-Shortley phase:
// the same functions as the reference closed forms).
// ---------------------------------------------------------------------------
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));   // a 128-bit buffer load
constexpr double kPi = 3.14159265358979323846;
constexpr double cfact(int n) { double f = 1; for (int i = 2; i <= n; ++i) f *= i; return f; }
constexpr double csqrt(double v) {
    double r = v > 1 ? v : 1.0;
    for (int it = 0; it < 200; ++it) r = 0.5 * (r + v / r);
    return r;
}
struct YTab { double c[7][7][8]; };
constexpr YTab make_ytab() {
    YTab t{};
    for (int l = 0; l <= 6; ++l) {
        double P[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; 2 * k <= l; ++k) {
            double sg = (k % 2) ? -1.0 : 1.0;
            double v = sg * cfact(l) / (cfact(k) * cfact(l - k)) *
                       cfact(2 * l - 2 * k) / (cfact(l) * cfact(l - 2 * k));
            double p2 = 1; for (int q = 0; q < l; ++q) p2 *= 2.0;
            P[l - 2 * k] += v / p2;
        }
        for (int m = 0; m <= l; ++m) {
            double D[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int k = 0; k <= l; ++k) D[k] = P[k];
            for (int d = 0; d < m; ++d) {
                for (int k = 0; k < 7; ++k) D[k] = D[k + 1] * (k + 1);
                D[7] = 0;
            }
            double N = csqrt((2 * l + 1) / (4 * kPi) * cfact(l - m) / cfact(l + m));
            double sg = (m % 2) ? -1.0 : 1.0;
            for (int k = 0; k <= l; ++k) t.c[l][m][k] = sg * N * D[k];
        }
    }
    return t;
}
constexpr YTab kY = make_ytab();

// coefficient of u^K of: W=0 c_lm, W=1 mu_m*m*c_lm, W=2 mu_m*c'_lm
template <int L, int M, int W, int K>
struct YC {
    static constexpr double mu = M ? 2.0 : 1.0;
    static constexpr double raw =
        W == 0 ? kY.c[L][M][K]
      : W == 1 ? mu * M * kY.c[L][M][K]
               : (K + 1 <= L ? mu * (K + 1) * kY.c[L][M][K + 1] : 0.0);
    static constexpr float v = (float)raw;
};
// Horner in v = u^2 over k = K, K+2, ..., TOP
template <int L, int M, int W, int K, int TOP>
struct HornerU2 {
    static __device__ __forceinline__ float run(float v) {
        if constexpr (K >= TOP) return YC<L, M, W, TOP>::v;
        else return __builtin_fmaf(v, HornerU2<L, M, W, K + 2, TOP>::run(v), YC<L, M, W, K>::v);
    }
};
// polynomial of degree DEG with parity DEG%2 evaluated at u (u2 = u*u)
template <int L, int M, int W>
__device__ __forceinline__ float ypoly(float u, float u2) {
    constexpr int DEG = (W == 2) ? (L - M - 1) : (L - M);
    if constexpr (DEG < 0) {
        return 0.0f;
    } else {
        float h = HornerU2<L, M, W, DEG % 2, DEG>::run(u2);
        if constexpr (DEG % 2) return h * u; else return h;
    }
}

template <int... Is, class F>
__device__ __forceinline__ void sfor_impl(std::integer_sequence<int, Is...>, F&& f) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void sfor(F&& f) { sfor_impl(std::make_integer_sequence<int, N>{}, f); }

// powers z^m = (nx + i ny)^m, m = 0..LMAX
template <int LMAX>
__device__ __forceinline__ void zpow(float nx, float ny, float* zr, float* zi) {
    zr[0] = 1.0f; zi[0] = 0.0f;
    zr[1] = nx; zi[1] = ny;
    #pragma unroll
    for (int m = 2; m <= LMAX; ++m) {
        zr[m] = __builtin_fmaf(zr[m - 1], nx, -zi[m - 1] * ny);
        zi[m] = __builtin_fmaf(zr[m - 1], ny, zi[m - 1] * nx);
    }
}

constexpr int lmax_of(int lmask) { return (lmask & kL6) ? 6 : (lmask & kL4) ? 4 : 3; }

// q[off..off+2L] += Y_l (packed: Re q_l0, then Re/Im for m = 1..L)
template <int L>
__device__ __forceinline__ void acc_ylm(float u, float u2, const float* zr, const float* zi, float* q) {
    q[0] += ypoly<L, 0, 0>(u, u2);
    sfor<L>([&](auto mm) {
        constexpr int M = decltype(mm)::value + 1;
        float c = ypoly<L, M, 0>(u, u2);
        q[2 * M - 1] = __builtin_fmaf(c, zr[M], q[2 * M - 1]);
        q[2 * M] = __builtin_fmaf(c, zi[M], q[2 * M]);
    });
}

// G += grad_n of sum_m mu_m Re(conj(K_m) c_lm(n_z) z^m)   (K packed as the record)
template <int L>
__device__ __forceinline__ void acc_grad(float u, float u2, const float* zr, const float* zi,
                                         const float* K, float& gx, float& gy, float& gz) {
    gz = __builtin_fmaf(ypoly<L, 0, 2>(u, u2), K[0], gz);
    sfor<L>([&](auto mm) {
        constexpr int M = decltype(mm)::value + 1;
        const float kr = K[2 * M - 1], ki = K[2 * M];
        // conj(K) z^(M-1)  (z^0 = 1)
        float tr, ti;
        if constexpr (M == 1) {
            tr = kr; ti = -ki;
        } else {
            tr = __builtin_fmaf(kr, zr[M - 1], ki * zi[M - 1]);
            ti = __builtin_fmaf(kr, zi[M - 1], -ki * zr[M - 1]);
        }
        float a = ypoly<L, M, 1>(u, u2);          // mu m c_lm
        gx = __builtin_fmaf(a, tr, gx);
        gy = __builtin_fmaf(-a, ti, gy);
        if constexpr (M < L) {                    // c'_lm = 0 for m = l
            float b = ypoly<L, M, 2>(u, u2);      // mu c'_lm
            gz = __builtin_fmaf(b, __builtin_fmaf(kr, zr[M], ki * zi[M]), gz);
        }
    });
}

__device__ __forceinline__ int owner_of(const Geo& g, int j) {
    return j < g.n_local_total ? j : g.owner[j - g.n_local_total];
}

// tile index with XCD-contiguous ranges: workgroup b runs on XCD b % 8
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    int per = nblocks >> 3;
    return (b & 7) * per + (b >> 3);
}
inline int xcd_grid(int ntiles) { return ((ntiles + 7) / 8) * 8; }

// FP64 lane move through DPP (both 32-bit halves); CTRL is the DPP control word.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)__double_as_longlong(v), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(__double_as_longlong(v) >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ double readlane_f64(double v, int l) {
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)__double_as_longlong(v), l);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)(__double_as_longlong(v) >> 32), l);
    return __longlong_as_double(((long long)(uint32_t)hi << 32) | (uint32_t)lo);
}
// Sum over the 64 lanes of a wave (all lanes active) in a fixed order: pairs in
// quads, quads in rows (mirrors), then the four row totals in order; every lane
// gets the same value.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);   // row_half_mirror
    v += dpp_f64<0x140>(v);   // row_mirror
    return ((readlane_f64(v, 15) + readlane_f64(v, 31)) + readlane_f64(v, 47)) + readlane_f64(v, 63);
}

// Block-wide FP64 sum of NV values per thread in a fixed order; result on thread 0.
// NW = waves per block (blockDim.x / 64); sh holds NV * NW doubles.
template <int NV, int NW = 4>
__device__ __forceinline__ void block_sum(double* v, double* sh) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    #pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = wave_sum(v[k]);
        if (lane == 0) sh[k * NW + wv] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        #pragma unroll
        for (int k = 0; k < NV; ++k) {
            double a = 0.0;
            #pragma unroll
            for (int q = 0; q < NW; ++q) a += sh[k * NW + q];
            v[k] = a;
        }
    }
}

// FP64 pair vector and r^2 exactly as the reference evaluates it without
// contraction: r2 = (dx*dx + dy*dy) + dz*dz.
// FP64 position of atom j for the selection's exact tests: from the engine buffer when
// the geometry carries one (x64), else from x4
__device__ __forceinline__ double3 pos64(const Geo& g, int j) {
    if (g.x64) {
        const double* p = g.x64 + 3 * (int64_t)j;
        return make_double3(p[0], p[1], p[2]);
    }
    const double4 v = g.x4[j];
    return make_double3(v.x, v.y, v.z);
}
__device__ __forceinline__ double pair_r2(const Geo& g, int j, double xi0, double xi1, double xi2) {
#pragma clang fp contract(off)
    const double3 p = pos64(g, j);
    const double d0 = p.x - xi0, d1 = p.y - xi1, d2 = p.z - xi2;
    return d0 * d0 + d1 * d1 + d2 * d2;
}

// ---------------------------------------------------------------------------
// List topology, once per neighbour-list binding: for entry e = (k -> j) the
// local owner i of j, and the reverse position: where in i's list the image
// of k sits at displacement -(x_j - x_k).  An atom sees one image of another
// inside the list cutoff when every box edge exceeds twice the cutoff, so the
// reverse entry is unique.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int list_len(const Geo& g, int i) {
    return g.cnt ? g.cnt[i] : g.off[i + 1] - g.off[i];
}

// Owner and reverse position of every list entry: for entry k -> j, the local owner i
// of j and the position in i's list of the image of k that sits at -(x_j - x_k) from
// i (NONE when i's list holds none).  With a near-list NearDev the build runs only when
// the current epoch's rebuild flag is up (or when forced).
template <class RT, int NONE>
__global__ __launch_bounds__(256) void k_topology(Geo g, int64_t n, int32_t* __restrict__ eown,
                                                   RT* __restrict__ revpos, const int32_t* __restrict__ flag,
                                                   const int64_t* __restrict__ epoch, int force) {
    if (flag && !force && flag[*epoch & 1] == 0) return;
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int beg = g.off[k], end = beg + list_len(g, (int)k);
    const double4 xk = g.x4[k];
    for (int e = beg; e < end; ++e) {
        const int j = g.idx[e];
        const int i = owner_of(g, j);
        const double4 xj = g.x4[j];
        const double d0 = xj.x - xk.x, d1 = xj.y - xk.y, d2 = xj.z - xk.z;
        const double4 xi = g.x4[i];
        int best = NONE;
        double bd = 1e-8;
        const int b2 = g.off[i], e2 = b2 + list_len(g, i);
        for (int q = b2; q < e2; ++q) {
            const int t = g.idx[q];
            if (owner_of(g, t) != (int)k) continue;
            const double4 xt = g.x4[t];
            double s0 = xt.x - xi.x + d0, s1 = xt.y - xi.y + d1, s2 = xt.z - xi.z + d2;
            double dd = s0 * s0 + s1 * s1 + s2 * s2;
            if (dd < bd) { bd = dd; best = q - b2; }
        }
        eown[e] = i;
        revpos[e] = (RT)best;
    }
}

// Near-list build: the entries of each bound list within Rp (FP64 r^2 <= Rp^2, the
// atom itself excluded), in list order, at the list's own offsets; the build positions
// of every atom (locals and ghosts) are kept for the displacement test.
__global__ __launch_bounds__(256) void k_prune(Geo g, NearDev nd, int force) {
    if (!force && nd.flag[*nd.epoch & 1] == 0) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nd.n_all) return;
    nd.xpr[i] = g.x4f[i];
    if (i >= nd.n_local) return;
    const int beg = nd.off[i], end = nd.off[i + 1];
    const double4 xi = g.x4[i];
    int k = 0;
    for (int e = beg; e < end; ++e) {
        const int j = nd.idx[e];
        const double4 xj = g.x4[j];
        const double d0 = xj.x - xi.x, d1 = xj.y - xi.y, d2 = xj.z - xi.z;
        const double r2 = (d0 * d0 + d1 * d1) + d2 * d2;
        if (r2 > nd.rp2 || r2 < 1e-12) continue;
        if (k < kMaxList) nd.nidx[beg + k] = j;
        ++k;
    }
    if (k > kMaxList) { *nd.overflow = 1; k = kMaxList; }
    nd.ncnt[i] = k;
}

// Packed FP32 position: xyz rounded to FP32, and in the fourth word the FP64
// rounding residual x - fl32(x) of each coordinate as a signed 10-bit multiple of
// ulp(fl32(x))/1024.  One 16-byte load then gives a pair vector to ~1e-7
// relative, the precision of the FP32 arithmetic that consumes it.
__device__ __forceinline__ uint32_t pack_res1(double x, float xf) {
    const uint32_t E = (__float_as_uint(xf) >> 23) & 0xFFu;
    if (E < 34u) return 0u;
    const double sc = (double)__uint_as_float((E - 33u) << 23);   // ulp(xf) / 1024, a power of two
    double q = rint((x - (double)xf) / sc);
    q = fmin(fmax(q, -511.0), 511.0);
    return (uint32_t)(int)q & 0x3FFu;
}
// The residual times 2^34: q ulp(xf)/1024 2^34 = q 2^(e) with e the frexp exponent
// of xf (q = 0 whenever the packing stored none).
__device__ __forceinline__ float unpack_res34(float xf, uint32_t w, int c) {
    const int q = ((int)(w << (22 - 10 * c))) >> 22;
    return __builtin_amdgcn_ldexpf((float)q, __builtin_amdgcn_frexp_expf(xf));
}
struct PPos { float x, y, z, rx, ry, rz; };   // r* = residual * 2^34
__device__ __forceinline__ PPos unpack_pos(const float4& p) {
    const uint32_t w = __float_as_uint(p.w);
    return PPos{p.x, p.y, p.z, unpack_res34(p.x, w, 0), unpack_res34(p.y, w, 1), unpack_res34(p.z, w, 2)};
}
// x_j - x_i from packed positions: (xf_j - xf_i) + (r_j - r_i), the residual
// difference rescaled exactly by the power of two
__device__ __forceinline__ void pdiff(const PPos& pj, const PPos& pi, float& d0, float& d1, float& d2) {
    constexpr float s = 0x1p-34f;
    d0 = __builtin_fmaf(s, pj.rx - pi.rx, pj.x - pi.x);
    d1 = __builtin_fmaf(s, pj.ry - pi.ry, pj.y - pi.y);
    d2 = __builtin_fmaf(s, pj.rz - pi.rz, pj.z - pi.z);
}

__global__ void k_pack_x4(const double* __restrict__ x, double4* __restrict__ x4, float4* __restrict__ x4f,
                          int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = x[3 * i], b = x[3 * i + 1], c = x[3 * i + 2];
        x4[i] = make_double4(a, b, c, 0.0);
        const float af = (float)a, bf = (float)b, cf = (float)c;
        const uint32_t w = pack_res1(a, af) | (pack_res1(b, bf) << 10) | (pack_res1(c, cf) << 20);
        x4f[i] = make_float4(af, bf, cf, __uint_as_float(w));
    }
}

// One walker's positions as the MD engine holds them (locals then ghost images,
// 3 FP64 each) packed into the context's concatenated layout: locals at x4[loc..],
// ghosts at x4[gh..].  Runs at the start of a step when the walker's positions are
// read from an attached engine buffer, and in mdt_update_positions.
// A workgroup takes 256 atoms at a time: their 768 coordinates are read with
// unit-stride loads (each load instruction touches consecutive lines) into LDS,
// then each thread packs its own atom.  With x4_loc and x4_gh NULL only the FP32
// packing is written (the step's FP64 reads go to the engine buffer itself).
__global__ __launch_bounds__(256) void k_pack_walker(const double* __restrict__ x, int64_t n_local, int64_t n_all,
                                                     double4* __restrict__ x4_loc, float4* __restrict__ x4f_loc,
                                                     double4* __restrict__ x4_gh, float4* __restrict__ x4f_gh) {
    __shared__ double xs[3 * 256];
    for (int64_t base = (int64_t)blockIdx.x * 256; base < n_all; base += (int64_t)gridDim.x * 256) {
        const int64_t nd = 3 * (n_all - base < 256 ? n_all - base : 256);
        __syncthreads();                               // the previous chunk's reads are complete
        #pragma unroll
        for (int u = 0; u < 3; ++u) {
            const int k = threadIdx.x + 256 * u;
            if (k < nd) xs[k] = x[3 * base + k];
        }
        __syncthreads();
        const int64_t i = base + threadIdx.x;
        if (i >= n_all) continue;
        const double a = xs[3 * threadIdx.x], b = xs[3 * threadIdx.x + 1], c = xs[3 * threadIdx.x + 2];
        const float af = (float)a, bf = (float)b, cf = (float)c;
        const uint32_t w = pack_res1(a, af) | (pack_res1(b, bf) << 10) | (pack_res1(c, cf) << 20);
        const bool loc = i < n_local;
        const int64_t k = loc ? i : i - n_local;
        if (x4_loc) (loc ? x4_loc : x4_gh)[k] = make_double4(a, b, c, 0.0);
        (loc ? x4f_loc : x4f_gh)[k] = make_float4(af, bf, cf, __uint_as_float(w));
    }
}

// The same, with the displacement test of the near list: any atom farther than
// sqrt(move2) from its position at the last build raises the current epoch's rebuild
// flag; the next epoch's flag is cleared here (nothing reads it during this epoch).
__global__ void k_pack_walker_near(const double* __restrict__ x, int64_t n_local, int64_t n_all, double4* __restrict__ x4_loc,
                                   float4* __restrict__ x4f_loc, double4* __restrict__ x4_gh, float4* __restrict__ x4f_gh,
                                   const float4* __restrict__ xpr_loc, const float4* __restrict__ xpr_gh, NearDev nd) {
    const int64_t e = *nd.epoch;
    if (blockIdx.x == 0 && threadIdx.x == 0) nd.flag[(e + 1) & 1] = 0;
    bool moved = false;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_all; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = x[3 * i], b = x[3 * i + 1], c = x[3 * i + 2];
        const float af = (float)a, bf = (float)b, cf = (float)c;
        const uint32_t w = pack_res1(a, af) | (pack_res1(b, bf) << 10) | (pack_res1(c, cf) << 20);
        const bool loc = i < n_local;
        const int64_t k = loc ? i : i - n_local;
        (loc ? x4_loc : x4_gh)[k] = make_double4(a, b, c, 0.0);
        (loc ? x4f_loc : x4f_gh)[k] = make_float4(af, bf, cf, __uint_as_float(w));
        const float4 p = (loc ? xpr_loc : xpr_gh)[k];
        const float d0 = af - p.x, d1 = bf - p.y, d2 = cf - p.z;
        moved |= d0 * d0 + d1 * d1 + d2 * d2 > nd.move2;
    }
    if (__ballot(moved) != 0ull && (threadIdx.x & 63) == 0) nd.flag[e & 1] = 1;
}

// ---------------------------------------------------------------------------
// 128-bit candidate masks and lane-group helpers.  The per-atom kernels run
// kLanes adjacent lanes per atom; lane l takes the set bits of rank l, l+kLanes,
// ...  Group sums use a fixed xor butterfly, so every lane of a group holds the
// bitwise identical result and the outcome does not depend on scheduling.
// ---------------------------------------------------------------------------
struct M128 { uint64_t lo, hi; };
__device__ __forceinline__ bool mbit(const M128& m, int p) {
    return p < 64 ? ((m.lo >> p) & 1ull) : ((m.hi >> (p - 64)) & 1ull);
}
__device__ __forceinline__ void mset(M128& m, int p) {
    if (p < 64) m.lo |= 1ull << p; else m.hi |= 1ull << (p - 64);
}
__device__ __forceinline__ bool mempty(const M128& m) { return (m.lo | m.hi) == 0ull; }
__device__ __forceinline__ int mpop_count(const M128& m) { return __popcll(m.lo) + __popcll(m.hi); }
__device__ __forceinline__ int mpop(M128& m) {
    if (m.lo) { int p = __builtin_ctzll(m.lo); m.lo &= m.lo - 1; return p; }
    int p = 64 + __builtin_ctzll(m.hi); m.hi &= m.hi - 1; return p;
}
__device__ __forceinline__ void mdrop(M128& m, int k) {
    for (int t = 0; t < k; ++t) { if (m.lo) m.lo &= m.lo - 1; else m.hi &= m.hi - 1; }
}
// next set bit of this lane's round-robin share (-1 when done); `it` must have
// had mdrop(it, lane) applied once before the first call
__device__ __forceinline__ int mnext(M128& it) {
    if (mempty(it)) return -1;
    const int p = mpop(it);
    mdrop(it, kLanes - 1);
    return p;
}
// Quad butterflies through DPP (quad_perm [1,0,3,2] then [2,3,0,1]); callers keep
// every lane of the wave active.
static_assert(kLanes == 4, "the group helpers are quad permutations");
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}
__device__ __forceinline__ uint64_t grp_or_u64(uint64_t v) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo |= dpp_u32<0xB1>(lo); hi |= dpp_u32<0xB1>(hi);
    lo |= dpp_u32<0x4E>(lo); hi |= dpp_u32<0x4E>(hi);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ M128 grp_or(M128 m) { return M128{grp_or_u64(m.lo), grp_or_u64(m.hi)}; }
__device__ __forceinline__ float grp_sum(float v) {
    v += __uint_as_float(dpp_u32<0xB1>(__float_as_uint(v)));
    v += __uint_as_float(dpp_u32<0x4E>(__float_as_uint(v)));
    return v;
}
// FP64 sum over an atom's lane group (quad butterfly through DPP; every lane gets it)
__device__ __forceinline__ double grp_sum_f64(double v) {
    v += dpp_f64<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);    // quad_perm [2,3,0,1]
    return v;
}

__device__ __forceinline__ M128 grp_bcast0(M128 m) {
    const int src = (int)(threadIdx.x & ~(unsigned)(kLanes - 1)) & 63;
    m.lo = (uint64_t)__shfl((long long)m.lo, src, 64);
    m.hi = (uint64_t)__shfl((long long)m.hi, src, 64);
    return m;
}

// Orders one wave's LDS writes and reads when lanes exchange data through LDS:
// the hardware completes a wave's LDS accesses in issue order, and the fences
// keep the compiler from moving a lane's accesses across the exchange point
// (it cannot see that other lanes read what this lane writes).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// ---------------------------------------------------------------------------
// k_env: selection + q_lm for every degree in LMASK, kLanes lanes per atom.
// Phase A scans the list with FP64 r^2; when the number within rc does not
// exceed N every such candidate is kept (what the reference insertion keeps
// then), otherwise the atom's lanes replay the reference insertion to find the
// kept set.  Phase B accumulates Y_lm in FP32 over the kept set.
// ---------------------------------------------------------------------------
// The reference insertion (get_environment_Steinhardt_Q) over one atom's list, with
// the sorted arrays in LDS (written by lane 0 of the atom's group), so that the
// cap-binding path does not set the register budget of the whole kernel: skip
// r^2 > rc^2 or r^2 < 1e-12, then walk the N slots with strict '>' carrying the
// displaced entry.  Entries outside rc never enter, so only the within-rc mask is
// walked, four entries at a time in list order: lane l computes the FP64 r^2 of the
// l-th (the four loads in flight together) and lane 0 inserts the four in order,
// entry by entry as the reference does, so the slots are the reference's.
// Every lane of the group calls it (the group's control flow is uniform).
template <int NMAX>
__device__ __forceinline__ void literal_insert_group(const Geo& g, int beg, const M128& inrc, double xi0, double xi1,
                                                     double xi2, double rc2, int N, double* sr2, int* sp, int lane) {
    if (lane == 0)
        for (int s = 0; s < N; ++s) { sr2[s] = __builtin_inf(); sp[s] = -1; }
    M128 it = inrc;
    mdrop(it, lane);
    const int nr = (mpop_count(inrc) + kLanes - 1) / kLanes;
    const int src0 = (int)(threadIdx.x & ~(unsigned)(kLanes - 1)) & 63;
    for (int rd = 0; rd < nr; ++rd) {
        const int p = mnext(it);
        double r2 = __builtin_inf();
        if (p >= 0) r2 = pair_r2(g, g.idx[beg + p], xi0, xi1, xi2);
        #pragma unroll
        for (int k = 0; k < kLanes; ++k) {
            const int q0 = __shfl(p, src0 + k, 64);
            const double r0 = __shfl(r2, src0 + k, 64);
            // the slots stay sorted (non-decreasing), so an entry no smaller than the last
            // slot changes nothing: it is passed over after one read
            if (lane == 0 && q0 >= 0 && !((r0 > rc2) || (r0 < 1e-12)) && sr2[N - 1] > r0) {
                double r = r0;
                int q = q0;
                for (int s = 0; s < N; ++s) {
                    const double tr = sr2[s];
                    if (tr > r) {
                        const int tq = sp[s];
                        sr2[s] = r; sp[s] = q;
                        r = tr; q = tq;
                    }
                }
            }
        }
    }
}


// EXPORT instance: phase A, then the insertion for every atom, whatever its count,
// writing the slots in the reference order (sel_out) and the count within rc; no
// q_lm, record or partial-sum writes.  Its kept set equals the step's by construction
// (the same scan and the same insertion).
template <int NMAX, int LMASK, bool EXPORT = false>
__global__ __launch_bounds__(kBlock) void k_env(Geo g, EnvDev e) {
    constexpr int NL = rec_nl(LMASK);
    constexpr int QF = rec_used(LMASK);
    constexpr int RS = rec_stride(LMASK);
    constexpr int LMAX = lmax_of(LMASK);
    __shared__ double sh[NL * 4];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const int a = threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
    const bool valid = a < tl.count;
    const int i = tl.start + (valid ? a : 0);
    double part[NL];
    #pragma unroll
    for (int k = 0; k < NL; ++k) part[k] = 0.0;

    const bool center = valid && (g.group == nullptr || g.group[i]);
    const int beg = g.off[i];
    const int c = center ? list_len(g, i) : 0;
    const double rc2 = e.rc2, rc2s = e.rc2 * (1.0 + kRcSlack);
    const int N = e.N;
    // Phase A: masks.  Each pass of the scan takes 16 consecutive entries of the
    // list, entry base + 4u + lane in slot u of the lane, with all list loads, then
    // all position loads, issued before any test.  The FP32 pair vector of the
    // rounded positions has |dd_k| <= u(2X + |d_k|) (u = 2^-24, |coordinates| <= X),
    // so its r^2 is within 7u X r + 5u r^2 of the FP64 r^2; every candidate farther
    // than twice that from the cutoffs is decided in FP32, the rest (and
    // near-coincident pairs) with the exact FP64 r^2, so the masks are the FP64
    // scan's.  A wave ballot per slot gives each group the 4 flags of its 4 lanes,
    // so the masks and counts are group-uniform without any exchange, and every
    // entry within rc is appended, packed position and all, to the atom's LDS list
    // in list order: when the cap does not bind that list is the selected set.
    constexpr int KC = NMAX < 32 ? NMAX : 32;       // LDS list length per atom
    constexpr int KB = (KC + kLanes - 1) / kLanes;
    // per atom: the kept list, or the sorted arrays of the literal insertion
    constexpr int KW = KC > (12 * NMAX + 15) / 16 ? KC : (12 * NMAX + 15) / 16;
    __shared__ float4 kept[kTile][KW];
    const int gsh = (threadIdx.x & 63) & ~(kLanes - 1);   // bit offset of the group in a ballot
    const float4 pif = g.x4f[i];
    float margin;
    {
        constexpr float u = 5.9604645e-8f;
        const float rcf = e.rcs;
        const float X = fmaxf(fmaxf(fabsf(pif.x), fabsf(pif.y)), fabsf(pif.z)) + rcf + 1.0f;
        margin = 2.0f * (5.0f * u * (float)rc2s + 7.0f * u * X * rcf) + 1e-6f;
    }
    const float r2lo = (float)rc2 - margin, r2hi = (float)rc2s + margin;
    M128 own{0, 0}, rcm{0, 0};
    int cnt = 0;                                     // group-uniform
    bool fbl = false;                                // this lane took the exact FP64 test
    // the list and the packed positions as raw buffers addressed by 32-bit offsets
    const __amdgpu_buffer_rsrc_t ridx = __builtin_amdgcn_make_buffer_rsrc((void*)g.idx, (short)0, 0x7FFFFFFF, 0x00020000);
    const __amdgpu_buffer_rsrc_t rx4f = __builtin_amdgcn_make_buffer_rsrc((void*)g.x4f, (short)0, 0x7FFFFFFF, 0x00020000);
    for (int base = 0; base < c; base += 4 * kLanes) {   // trip count uniform within the group
        int jj[4];
        #pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = base + u * kLanes + lane;
            jj[u] = (int)__builtin_amdgcn_raw_buffer_load_b32(ridx, 4u * (uint32_t)(beg + (p < c ? p : c - 1)), 0, 0);
        }
        float4 xx[4];
        #pragma unroll
        for (int u = 0; u < 4; ++u) {
            const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(rx4f, 16u * (uint32_t)jj[u], 0, 0);
            xx[u] = make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
        }
        uint32_t cin = 0, crc = 0;
        #pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = base + u * kLanes + lane;
            bool in = false, inrc = false;
            if (p < c) {
                const float e0 = xx[u].x - pif.x, e1 = xx[u].y - pif.y, e2 = xx[u].z - pif.z;
                const float r2f = __builtin_fmaf(e2, e2, __builtin_fmaf(e1, e1, e0 * e0));
                if (r2f < r2lo && r2f > 1e-2f) {
                    in = inrc = true;
                } else if (!(r2f > r2hi)) {
                    fbl = true;
                    const double3 xi = pos64(g, i);   // the FP64 position only on this rare path
                    const double r2 = pair_r2(g, jj[u], xi.x, xi.y, xi.z);
                    inrc = r2 >= 1e-12 && r2 <= rc2s;
                    in = !((r2 > rc2) || (r2 < 1e-12));
                }
            }
            const uint32_t ni = (uint32_t)(__ballot(in) >> gsh) & 0xFu;
            const uint32_t nr = (uint32_t)(__ballot(inrc) >> gsh) & 0xFu;
            if (in) {
                const int rk = cnt + __popc(ni & ((1u << lane) - 1u));
                if (rk < KC) kept[a][rk] = xx[u];
            }
            cnt += __popc(ni);
            cin |= ni << (4 * u);
            crc |= nr << (4 * u);
        }
        if (base < 64) {
            own.lo |= (uint64_t)cin << base;
            rcm.lo |= (uint64_t)crc << base;
        } else {
            own.hi |= (uint64_t)cin << (base - 64);
            rcm.hi |= (uint64_t)crc << (base - 64);
        }
    }
    const bool band = ((uint32_t)(__ballot(fbl) >> gsh) & 0xFu) != 0u;
    // Phase B: Y_lm of the kept set, round-robin over the group's lanes.
    const PPos pi = unpack_pos(pif);
    float q[QF];
    #pragma unroll
    for (int k = 0; k < QF; ++k) q[k] = 0.0f;
    auto add_ylm = [&](const float4& pj) {
        float d0, d1, d2;
        pdiff(unpack_pos(pj), pi, d0, d1, d2);
        const float ir = __builtin_amdgcn_rsqf(__builtin_fmaf(d2, d2, __builtin_fmaf(d1, d1, d0 * d0)));
        const float nx = d0 * ir, ny = d1 * ir, uu = d2 * ir;
        const float u2 = uu * uu;
        float zr[LMAX + 1], zi[LMAX + 1];
        zpow<LMAX>(nx, ny, zr, zi);
        if constexpr (LMASK & kL3) acc_ylm<3>(uu, u2, zr, zi, q + rec_off(LMASK, 3));
        if constexpr (LMASK & kL4) acc_ylm<4>(uu, u2, zr, zi, q + rec_off(LMASK, 4));
        if constexpr (LMASK & kL6) acc_ylm<6>(uu, u2, zr, zi, q + rec_off(LMASK, 6));
    };
    wave_lds_sync();
    if (!EXPORT && cnt <= N && cnt <= KC) {          // uniform within the group
        float4 pk[KB];
        #pragma unroll
        for (int s = 0; s < KB; ++s)
            if (lane + s * kLanes < cnt) pk[s] = kept[a][lane + s * kLanes];
        #pragma unroll
        for (int s = 0; s < KB; ++s)
            if (lane + s * kLanes < cnt) add_ylm(pk[s]);
    } else {
        if (EXPORT || cnt > N) {                     // the cap binds: the literal insertion
            double* sr2 = reinterpret_cast<double*>(&kept[a][0]);
            int* sp = reinterpret_cast<int*>(sr2 + NMAX);
            const double3 xi = pos64(g, i);
            literal_insert_group<NMAX>(g, beg, own, xi.x, xi.y, xi.z, rc2, N, sr2, sp, lane);
            if (lane == 0) {
                own = M128{0, 0};
                for (int s = 0; s < N; ++s)
                    if (sp[s] >= 0) mset(own, sp[s]);
                if constexpr (EXPORT) {
                    if (valid) {
                        for (int s = 0; s < N; ++s) e.sel_out[(int64_t)i * N + s] = sp[s] >= 0 ? g.idx[beg + sp[s]] : -1;
                        e.sel_cnt[i] = cnt;
                    }
                }
            }
            own = grp_bcast0(own);
        }
        if constexpr (EXPORT) return;
        M128 it = own;
        mdrop(it, lane);
        for (int p = mnext(it); p >= 0; p = mnext(it)) add_ylm(g.x4f[g.idx[beg + p]]);
    }
    const int n = cnt < N ? cnt : N;
    const float inv_n = n > 0 ? __builtin_amdgcn_rcpf((float)n) : 0.0f;
    #pragma unroll
    for (int k = kRecHdr; k < QF; ++k) q[k] = grp_sum(q[k]) * inv_n;
    float out[RS];
    #pragma unroll
    for (int k = 0; k < RS; ++k) out[k] = k < QF ? q[k] : 0.0f;
    // header: the packed position and the two flags the force pass reads
    out[0] = pif.x;
    out[1] = pif.y;
    out[2] = pif.z;
    out[3] = __uint_as_float((__float_as_uint(pif.w) & ~(kHdrCap | kHdrBand)) | (cnt > N ? kHdrCap : 0u) |
                             (band ? kHdrBand : 0u));
    // the atom's share of the CV sums (an atom outside the context's share keeps its q_lm
    // for its neighbours' use, but adds nothing to the sums nor to their gradient)
    const bool inshare = g.share == nullptr || g.share[i];
    auto finish = [&](auto ll) {
        constexpr int L = decltype(ll)::value;
        if constexpr ((L == 3 && (LMASK & kL3)) || (L == 4 && (LMASK & kL4)) || (L == 6 && (LMASK & kL6))) {
            constexpr int o = rec_off(LMASK, L);
            constexpr int li = rec_li(LMASK, L);
            float s2 = q[o] * q[o];
            #pragma unroll
            for (int m = 1; m <= L; ++m) s2 += 2.0f * (q[o + 2 * m - 1] * q[o + 2 * m - 1] + q[o + 2 * m] * q[o + 2 * m]);
            constexpr float P = (float)(4.0 * kPi / (2 * L + 1));
            const float ql = __builtin_amdgcn_sqrtf(s2 * P);
            out[rec_a(LMASK, L)] = (ql > 0.0f && n > 0 && inshare) ? P * __builtin_amdgcn_rcpf(ql * (float)n) : 0.0f;
            if (valid && lane == 0) e.ql[(int64_t)li * g.n_local_total + i] = ql;
            part[li] = (center && inshare && lane == 0) ? (double)ql : 0.0;
        }
    };
    finish(std::integral_constant<int, 3>{});
    finish(std::integral_constant<int, 4>{});
    finish(std::integral_constant<int, 6>{});
    if (valid) {
        float4* dst = reinterpret_cast<float4*>(e.rec + (int64_t)i * RS);
        #pragma unroll
        for (int k = 0; k < RS / 4; ++k)
            if (k % kLanes == lane) dst[k] = make_float4(out[4 * k], out[4 * k + 1], out[4 * k + 2], out[4 * k + 3]);
        if (lane == 0) {
            e.nsel[i] = n;
            e.cnt_rc[i] = cnt;
            reinterpret_cast<ulonglong2*>(e.selmask)[2 * i] = make_ulonglong2(own.lo, own.hi);
            reinterpret_cast<ulonglong2*>(e.selmask)[2 * i + 1] = make_ulonglong2(rcm.lo, rcm.hi);
        }
    }
    block_sum<NL>(part, sh);
    if (threadIdx.x == 0) {
        #pragma unroll
        for (int k = 0; k < NL; ++k) e.partial[(int64_t)e.slot_ql[k] * g.ntiles + t] = part[k];
    }
}

// Reverse-selected candidates of atom i: the candidate's owner selects i.
// Checked only inside the slack-widened rc mask, distributed over the lanes.
__device__ __forceinline__ M128 reverse_mask(const Geo& g, const EnvDev& e, int beg, const M128& rcm, int lane) {
    M128 rev{0, 0};
    M128 it = rcm;
    mdrop(it, lane);
    while (!mempty(it)) {
        const int p = mpop(it);
        const int rp = g.revpos[beg + p];
        if (rp != 0xFF) {
            const ulonglong2 mo = reinterpret_cast<const ulonglong2*>(e.selmask)[2 * g.eown[beg + p]];
            if (rp < 64 ? ((mo.x >> rp) & 1ull) : ((mo.y >> (rp - 64)) & 1ull)) mset(rev, p);
        }
        mdrop(it, kLanes - 1);
    }
    return grp_or(rev);
}

// ---------------------------------------------------------------------------
// Lechner-Dellago local q: qbar_i = (q_i + sum_{j in S(i)} q_owner(j)) / (n_i + 1).
// Writes lq, bq = P / (lq (n+1)) * qbar, and per-tile sums of lq (the 1/N_group of the
// average enters through the chain-rule weight).
// ---------------------------------------------------------------------------
template <int L, int LMASK>
__global__ __launch_bounds__(kBlock) void k_ld(Geo g, EnvDev e, LdDev d) {
    // The kLanes lanes of an atom walk its selected neighbours together; lane t
    // owns the float4 slice t of the degree-L block, so each neighbour's block is
    // one cache line touched by one load instruction of the group.
    constexpr int RS = rec_stride(LMASK);
    constexpr int O = rec_off(LMASK, L);
    constexpr int W = 2 * L + 1;
    constexpr int BS = (W + 3) / 4 * 4;
    constexpr int NB = BS / 4;                 // float4 slices of the block
    static_assert(NB <= kLanes, "a degree block must fit one float4 per lane");
    static_assert(O % 4 == 0 && O + BS <= RS, "degree block must be float4 aligned inside the record");
    __shared__ double sh[4];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const int a = threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
    const bool valid = a < tl.count;
    const int i = tl.start + (valid ? a : 0);
    double part[1] = {0.0};
    const bool center = valid && (g.group == nullptr || g.group[i]);
    const int sl = lane < NB ? lane : 0;       // lanes beyond the block repeat slice 0 and drop it
    const bool mine = lane < NB;
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const ulonglong2 mk = reinterpret_cast<const ulonglong2*>(e.selmask)[2 * i];
    M128 it{valid ? mk.x : 0ull, valid ? mk.y : 0ull};
    const int beg = g.off[i];
    // batches of kLdBatch neighbours: all owner indices, then all blocks in flight
    constexpr int kLdBatch = 8;
    while (!mempty(it)) {
        int oo[kLdBatch];
        #pragma unroll
        for (int u = 0; u < kLdBatch; ++u) {
            const int p = mempty(it) ? -1 : mpop(it);
            oo[u] = p >= 0 ? g.eown[beg + p] : -1;
        }
        float4 vv[kLdBatch];
        #pragma unroll
        for (int u = 0; u < kLdBatch; ++u)
            vv[u] = oo[u] >= 0 ? reinterpret_cast<const float4*>(e.rec + (int64_t)oo[u] * RS + O)[sl]
                               : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        #pragma unroll
        for (int u = 0; u < kLdBatch; ++u) { acc.x += vv[u].x; acc.y += vv[u].y; acc.z += vv[u].z; acc.w += vv[u].w; }
    }
    const float4 vi = reinterpret_cast<const float4*>(e.rec + (int64_t)i * RS + O)[sl];
    const int n = e.nsel[i];
    const float inv = 1.0f / (float)(n + 1);
    float qb[4] = {(acc.x + vi.x) * inv, (acc.y + vi.y) * inv, (acc.z + vi.z) * inv, (acc.w + vi.w) * inv};
    float s2 = 0.0f;
    #pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int comp = 4 * sl + k;
        const float wgt = !mine || comp >= W ? 0.0f : (comp == 0 ? 1.0f : 2.0f);
        s2 += wgt * qb[k] * qb[k];
    }
    s2 = grp_sum(s2);
    constexpr float P = (float)(4.0 * kPi / (2 * L + 1));
    const float lq = center ? __builtin_sqrtf(P * s2) : 0.0f;
    const bool inshare = g.share == nullptr || g.share[i];
    const float beta = lq > 0.0f && inshare ? P / (lq * (float)(n + 1)) : 0.0f;
    if (valid && mine && d.wslot >= 0) {      // gradient coefficients only when LQ.AVE is biased
        float o4[4];
        #pragma unroll
        for (int k = 0; k < 4; ++k) o4[k] = 4 * sl + k < W ? beta * qb[k] : 0.0f;
        reinterpret_cast<float4*>(d.bq + (int64_t)i * BS)[sl] = make_float4(o4[0], o4[1], o4[2], o4[3]);
    }
    if (valid && lane == 0) {
        d.lq[i] = lq;
        part[0] = inshare ? (double)lq : 0.0;
    }
    block_sum<1>(part, sh);
    if (threadIdx.x == 0) e.partial[(int64_t)d.slot * g.ntiles + t] = part[0];
}

// extra_i[l block] = w / n_i * (bq_i + sum_{i' : i in S(i')} bq_i')
template <int L, int LMASK>
__global__ __launch_bounds__(kBlock) void k_ld_coeff(Geo g, EnvDev e, LdDev d) {
    constexpr int RS = rec_stride(LMASK);
    constexpr int O = rec_off(LMASK, L);
    constexpr int W = 2 * L + 1;
    constexpr int BS = (W + 3) / 4 * 4;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const int a = threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
    const bool valid = a < tl.count;
    const int i = tl.start + (valid ? a : 0);
    const double w = e.wcv[(int64_t)tl.walker * kWcvStride + d.wslot];
    const ulonglong2 mr = reinterpret_cast<const ulonglong2*>(e.selmask)[2 * i + 1];
    const int beg = g.off[i];
    M128 rcm{valid ? mr.x : 0ull, valid ? mr.y : 0ull};
    M128 rev = reverse_mask(g, e, beg, rcm, lane);
    float C[BS];
    #pragma unroll
    for (int k = 0; k < BS; ++k) C[k] = 0.0f;
    mdrop(rev, lane);
    while (!mempty(rev)) {
        const int p = mpop(rev);
        const float4* bo = reinterpret_cast<const float4*>(d.bq + (int64_t)g.eown[beg + p] * BS);
        #pragma unroll
        for (int k = 0; k < BS / 4; ++k) { float4 v = bo[k]; C[4*k] += v.x; C[4*k+1] += v.y; C[4*k+2] += v.z; C[4*k+3] += v.w; }
        mdrop(rev, kLanes - 1);
    }
    const float4* bi = reinterpret_cast<const float4*>(d.bq + (int64_t)i * BS);
    float B[BS];
    #pragma unroll
    for (int k = 0; k < BS / 4; ++k) { float4 v = bi[k]; B[4*k] = v.x; B[4*k+1] = v.y; B[4*k+2] = v.z; B[4*k+3] = v.w; }
    const int n = e.nsel[i];
    const float s = n > 0 ? (float)w / (float)n : 0.0f;
    float* ex = e.extra + (int64_t)i * RS + O;
    #pragma unroll
    for (int k = 0; k < W; ++k) {
        const float v = s * (B[k] + grp_sum(C[k]));
        if (valid && lane == 0) ex[k] = v;
    }
}

// ---------------------------------------------------------------------------
// k_force: F_k = sum_{j in S(k)} g_kj - sum_{(i,j): owner(j)=k, j in S(i)} g_ij,
// g_ij = sum_lm mu_m Re(conj(K_ilm) grad Y_lm(x_j - x_i)),
// K_i = w_l a_il q_il (+ extra_i).  Gather over k's own list: every candidate
// that k selects (own) or whose owner selects k (rev) contributes once, with
// the two coefficient sets added before the contraction.  kLanes lanes per
// atom share the active pairs round-robin.
// ---------------------------------------------------------------------------
// K from a record already in registers (header quad ignored)
template <int LMASK, bool EXTRA>
__device__ __forceinline__ void coeff_from_record(const EnvDev& e, int a, const float* wl, const float* r, float* K) {
    constexpr int RS = rec_stride(LMASK);
    constexpr int NV = rec_used(LMASK) / 4;
    float x[NV * 4];
    if constexpr (EXTRA) {
        const float4* xp = reinterpret_cast<const float4*>(e.extra + (int64_t)a * RS);
        #pragma unroll
        for (int k = 0; k < kRecHdr; ++k) x[k] = 0.0f;
        #pragma unroll
        for (int k = 1; k < NV; ++k) { float4 v = xp[k]; x[4*k] = v.x; x[4*k+1] = v.y; x[4*k+2] = v.z; x[4*k+3] = v.w; }
    }
    #pragma unroll
    for (int k = 0; k < NV * 4; ++k) K[k] = 0.0f;
    auto one = [&](auto ll) {
        constexpr int L = decltype(ll)::value;
        if constexpr ((L == 3 && (LMASK & kL3)) || (L == 4 && (LMASK & kL4)) || (L == 6 && (LMASK & kL6))) {
            constexpr int o = rec_off(LMASK, L);
            constexpr int li = rec_li(LMASK, L);
            const float s = wl[li] * r[rec_a(LMASK, L)];
            #pragma unroll
            for (int k = 0; k < 2 * L + 1; ++k) {
                float v = s * r[o + k];
                if constexpr (EXTRA) v += x[o + k];
                K[o + k] = v;
            }
        }
    };
    one(std::integral_constant<int, 3>{});
    one(std::integral_constant<int, 4>{});
    one(std::integral_constant<int, 6>{});
}

template <int LMASK, bool EXTRA>
__device__ __forceinline__ void load_coeff(const EnvDev& e, int a, const float* wl, float* K) {
    constexpr int RS = rec_stride(LMASK);
    constexpr int NV = rec_used(LMASK) / 4;     // float4s actually used (the first is the header)
    const float4* rp = reinterpret_cast<const float4*>(e.rec + (int64_t)a * RS);
    float r[NV * 4];
    #pragma unroll
    for (int k = 0; k < kRecHdr; ++k) r[k] = 0.0f;
    #pragma unroll
    for (int k = 1; k < NV; ++k) { float4 v = rp[k]; r[4*k] = v.x; r[4*k+1] = v.y; r[4*k+2] = v.z; r[4*k+3] = v.w; }
    coeff_from_record<LMASK, EXTRA>(e, a, wl, r, K);
}

// Quad broadcast: every lane of an aligned group of four gets lane r's value.
template <int R>
__device__ __forceinline__ int quad_bcast(int v) {
    return __builtin_amdgcn_mov_dpp(v, R | (R << 2) | (R << 4) | (R << 6), 0xF, 0xF, true);
}
template <int R>
__device__ __forceinline__ float quad_bcastf(float v) { return __int_as_float(quad_bcast<R>(__float_as_int(v))); }
template <int R>
__device__ __forceinline__ uint64_t quad_bcast64(uint64_t v) {
    return ((uint64_t)(uint32_t)quad_bcast<R>((int)(v >> 32)) << 32) | (uint32_t)quad_bcast<R>((int)(uint32_t)v);
}

template <int LMASK, bool EXTRA, int LDL>
__global__ __launch_bounds__(kBlock) void k_force(Geo g, EnvDev e, double* __restrict__ f) {
    static_assert(LDL == 0 || (LDL == 3 && (LMASK & kL3)) || (LDL == 4 && (LMASK & kL4)) || (LDL == 6 && (LMASK & kL6)),
                  "the fused local-Q degree must be on the environment");
    constexpr int LW = LDL ? 2 * LDL + 1 : 1;           // fused local-Q accumulators
    constexpr int LO = LDL ? rec_off(LMASK, LDL) : 0;
    float qb[LW];
    #pragma unroll
    for (int k = 0; k < LW; ++k) qb[k] = 0.0f;
    constexpr int QF = rec_used(LMASK);
    constexpr int NL = rec_nl(LMASK);
    constexpr int LMAX = lmax_of(LMASK);
    constexpr int RS = rec_stride(LMASK);
    constexpr int NQ = RS / 16;                // 64-byte quads per record
    constexpr int NV = QF / 4;                 // float4 of the record that are used
    // per wave: for each quad and candidate slot, the 64 lanes' float4 of one global
    // to LDS load (32 KB per workgroup for two-quad records)
    __shared__ float4 dst[kBlock / 64][NQ][kLanes][64];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const int a = threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
    const bool valid = a < tl.count;
    const int i = tl.start + (valid ? a : 0);
    // the neighbour records as a raw buffer (32-bit offsets; loads past its end read 0)
    const __amdgpu_buffer_rsrc_t rrec = __builtin_amdgcn_make_buffer_rsrc(
        (void*)e.rec, (short)0, (int)((uint32_t)g.n_local_total * (uint32_t)(RS * 4)), 0x00020000);
    float wl[NL];
    #pragma unroll
    for (int k = 0; k < NL; ++k)
        wl[k] = e.wslot[k] >= 0 ? (float)e.wcv[(int64_t)tl.walker * kWcvStride + e.wslot[k]] : 0.0f;
    // The atom's own data is loaded once per group, each lane one share, and
    // broadcast through DPP: the four mask words (own lo/hi, rc lo/hi), the
    // record, the packed position.
    const uint64_t mw = valid ? reinterpret_cast<const uint64_t*>(e.selmask)[4 * (int64_t)i + lane] : 0ull;
    const M128 own{quad_bcast64<0>(mw), quad_bcast64<1>(mw)};
    const M128 rcm{quad_bcast64<2>(mw), quad_bcast64<3>(mw)};
    const int beg = g.off[i];
    float Ki[QF];
    uint32_t hwi;                                     // the atom's header flags word
    if constexpr (EXTRA) {
        load_coeff<LMASK, EXTRA>(e, i, wl, Ki);
        hwi = __float_as_uint(e.rec[(int64_t)i * RS + 3]);
    } else {
        constexpr int NCH = (NV + kLanes - 1) / kLanes;
        const float4* rp = reinterpret_cast<const float4*>(e.rec + (int64_t)i * RS);
        float4 ch[NCH];
        #pragma unroll
        for (int h = 0; h < NCH; ++h)
            ch[h] = h * kLanes + lane < NV ? rp[h * kLanes + lane] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float r[NV * 4];
        #pragma unroll
        for (int k = 0; k < kRecHdr; ++k) r[k] = 0.0f;
        hwi = __float_as_uint(quad_bcastf<0>(ch[0].w));
        sfor<NV>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            if constexpr (k >= 1) {
                const float4 v = ch[k / kLanes];
                r[4 * k] = quad_bcastf<k % kLanes>(v.x); r[4 * k + 1] = quad_bcastf<k % kLanes>(v.y);
                r[4 * k + 2] = quad_bcastf<k % kLanes>(v.z); r[4 * k + 3] = quad_bcastf<k % kLanes>(v.w);
            }
        });
        coeff_from_record<LMASK, EXTRA>(e, i, wl, r, Ki);
    }
    PPos pi;
    {
        const float pw = reinterpret_cast<const float*>(g.x4f + i)[lane];
        pi = unpack_pos(make_float4(quad_bcastf<0>(pw), quad_bcastf<1>(pw), quad_bcastf<2>(pw), quad_bcastf<3>(pw)));
    }
    float Fx = 0.0f, Fy = 0.0f, Fz = 0.0f;
    // Rounds over the slack-widened rc candidates, lane l taking the set bits of
    // rank l, l+4, ...  In each round the group's four candidate records are
    // fetched cooperatively (every load instruction of the group touches one line)
    // straight into LDS; each lane reads back its own candidate's record, whose
    // header carries the candidate's packed position (ghost images load theirs),
    // and evaluates its pair.
    const int nrounds = (mpop_count(rcm) + kLanes - 1) / kLanes;   // uniform within the group
    // wave (wave-uniform: the LDS destinations of the record loads are scalar), lane group within it
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), gq = (threadIdx.x & 63) >> 2;
    // lists of at most 64 entries (the usual case, uniform over the wave): the
    // lane's share of the low word is walked with plain 64-bit clears
    const bool wide = __ballot(rcm.hi != 0ull) != 0ull;
    // candidates all within the first 32 list positions (uniform over the wave): 32-bit clears
    const bool narrow = __ballot(rcm.hi != 0ull || (rcm.lo >> 32) != 0ull) == 0ull;
    M128 it = rcm;
    mdrop(it, lane);
    uint64_t it64 = it.lo;
    uint32_t it32 = (uint32_t)it.lo;
    auto next_p = [&]() {
        int q;
        if (narrow) {
            q = __ffs(it32) - 1;
            #pragma unroll
            for (int k = 0; k < kLanes; ++k) it32 &= it32 - 1u;
        } else if (!wide) {
            q = (int)__ffsll((unsigned long long)it64) - 1;
            #pragma unroll
            for (int k = 0; k < kLanes; ++k) it64 &= it64 - 1ull;
        } else {
            q = mnext(it);
        }
        return q;
    };
    // the list words (reverse position, owner, index) of the next round are loaded
    // while the current round runs
    int pn = next_p();
    // the list words as raw buffers addressed by 32-bit offsets
    const __amdgpu_buffer_rsrc_t rrev = __builtin_amdgcn_make_buffer_rsrc((void*)g.revpos, (short)0, 0x7FFFFFFF, 0x00020000);
    const __amdgpu_buffer_rsrc_t rown = __builtin_amdgcn_make_buffer_rsrc((void*)g.eown, (short)0, 0x7FFFFFFF, 0x00020000);
    const __amdgpu_buffer_rsrc_t ridx = __builtin_amdgcn_make_buffer_rsrc((void*)g.idx, (short)0, 0x7FFFFFFF, 0x00020000);
    auto words = [&](int p, int& rp, int& o, int& j) {
        const uint32_t ent = (uint32_t)(beg + p);
        rp = p >= 0 ? (int)__builtin_amdgcn_raw_buffer_load_b8(rrev, ent, 0, 0) : 0xFF;
        o = p >= 0 ? (int)__builtin_amdgcn_raw_buffer_load_b32(rown, 4u * ent, 0, 0) : i;
        j = p >= 0 ? (int)__builtin_amdgcn_raw_buffer_load_b32(ridx, 4u * ent, 0, 0) : i;
    };
    int rpn, on, jn;
    words(pn, rpn, on, jn);
    for (int rd = 0; rd < nrounds; ++rd) {
        const int p = pn, rp = rpn, o = on, j = jn;
        // The group's four candidate records go straight from global memory into
        // LDS, one quad of one candidate per load instruction, every instruction
        // touching one line per group.  In the load for candidate slot c, lane l of
        // a group fetches float4 (l + c) mod 4 of the quad, so that each lane later
        // reads its own candidate's quad from four different banks.
        wave_lds_sync();                             // last round's reads are complete
        {
            const int oc[kLanes] = {quad_bcast<0>(o), quad_bcast<1>(o), quad_bcast<2>(o), quad_bcast<3>(o)};
            sfor<NQ>([&](auto qq) {
                constexpr int q = decltype(qq)::value;
                #pragma unroll
                for (int c = 0; c < kLanes; ++c) {
                    const uint32_t vo = (uint32_t)oc[c] * (uint32_t)(RS * 4) + (uint32_t)(((lane + c) & 3) * 16);
                    // the quad's byte offset goes in the scalar offset, which moves the global
                    // address only (the immediate offset would move the LDS destination too)
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(rrec, (__attribute__((address_space(3))) void*)&dst[wv][q][c][0],
                                                             16, vo, 64 * q, 0, 0);
                }
            });
        }
        if (rd + 1 < nrounds) {                      // uniform within the group
            pn = next_p();
            words(pn, rpn, on, jn);
        }
        float rr[NV * 4];
        #pragma unroll
        for (int q = 0; q < NQ; ++q) {
            #pragma unroll
            for (int k = 0; k < kLanes; ++k) {
                if (4 * q + k < NV) {
                    const float4 u4 = dst[wv][q][lane][4 * gq + ((k - lane) & 3)];
                    rr[16 * q + 4 * k] = u4.x; rr[16 * q + 4 * k + 1] = u4.y;
                    rr[16 * q + 4 * k + 2] = u4.z; rr[16 * q + 4 * k + 3] = u4.w;
                }
            }
        }
        // the candidate's packed position: a local neighbour's is its record header,
        // a ghost image's is loaded
        float4 pj = make_float4(rr[0], rr[1], rr[2], rr[3]);
        if (j >= g.n_local_total) pj = g.x4f[j];
        // Reverse selection: the owner selects i.  When the owner's cap does not bind
        // and no entry of i lies near rc (where the two frames' FP64 distances might
        // decide differently), every candidate inside rc is selected by its owner;
        // otherwise the owner's selection mask decides.
        bool is_own = false, is_rev = false;
        if (p >= 0) {
            is_own = mbit(own, p);
            if (rp == 0xFF) {
                is_rev = false;
            } else if (((__float_as_uint(rr[3]) & kHdrCap) | (hwi & kHdrBand)) == 0u) {
                is_rev = true;
            } else {
                const uint64_t w = reinterpret_cast<const uint64_t*>(e.selmask)[4 * (int64_t)o + (rp >> 6)];
                is_rev = (w >> (rp & 63)) & 1ull;
            }
        }
        // every active pair of the wave selected from both sides (the usual case when
        // the cap does not bind): the combination needs no per-lane selects
        const bool mutual = __ballot(is_own != is_rev) == 0ull;
        {
            if constexpr (LDL != 0) {
                if (is_own) {
                    #pragma unroll
                    for (int k = 0; k < LW; ++k) qb[k] += rr[LO + k];
                }
            }
            if (is_own || is_rev) {
                // C = [own] K_i + [rev] (-1)^l K_o with K_o = w_l a_o q_o (+ extra_o): the
                // reverse flag and the sign fold into one scale per degree
                float C[QF];
                auto comb = [&](auto ll) {
                    constexpr int L = decltype(ll)::value;
                    if constexpr ((L == 3 && (LMASK & kL3)) || (L == 4 && (LMASK & kL4)) || (L == 6 && (LMASK & kL6))) {
                        constexpr int ob = rec_off(LMASK, L);
                        constexpr float sg = (L & 1) ? -1.0f : 1.0f;
                        const float s0 = sg * wl[rec_li(LMASK, L)] * rr[rec_a(LMASK, L)];
                        if (mutual) {
                            #pragma unroll
                            for (int k = 0; k < 2 * L + 1; ++k) C[ob + k] = __builtin_fmaf(s0, rr[ob + k], Ki[ob + k]);
                        } else {
                            const float sc = is_rev ? s0 : 0.0f;
                            #pragma unroll
                            for (int k = 0; k < 2 * L + 1; ++k)
                                C[ob + k] = __builtin_fmaf(sc, rr[ob + k], is_own ? Ki[ob + k] : 0.0f);
                        }
                    }
                };
                comb(std::integral_constant<int, 3>{});
                comb(std::integral_constant<int, 4>{});
                comb(std::integral_constant<int, 6>{});
                if constexpr (EXTRA) {
                    if (is_rev) {
                        const float4* xp = reinterpret_cast<const float4*>(e.extra + (int64_t)o * RS);
                        float x[NV * 4];
                        #pragma unroll
                        for (int k = 1; k < NV; ++k) {
                            const float4 v = xp[k];
                            x[4 * k] = v.x; x[4 * k + 1] = v.y; x[4 * k + 2] = v.z; x[4 * k + 3] = v.w;
                        }
                        auto addx = [&](auto ll) {
                            constexpr int L = decltype(ll)::value;
                            if constexpr ((L == 3 && (LMASK & kL3)) || (L == 4 && (LMASK & kL4)) || (L == 6 && (LMASK & kL6))) {
                                constexpr int ob = rec_off(LMASK, L);
                                constexpr float sg = (L & 1) ? -1.0f : 1.0f;
                                #pragma unroll
                                for (int k = 0; k < 2 * L + 1; ++k) C[ob + k] = __builtin_fmaf(sg, x[ob + k], C[ob + k]);
                            }
                        };
                        addx(std::integral_constant<int, 3>{});
                        addx(std::integral_constant<int, 4>{});
                        addx(std::integral_constant<int, 6>{});
                    }
                }
                float d0, d1, d2;
                pdiff(unpack_pos(pj), pi, d0, d1, d2);
                const float ir = __builtin_amdgcn_rsqf(__builtin_fmaf(d2, d2, __builtin_fmaf(d1, d1, d0 * d0)));
                const float nx = d0 * ir, ny = d1 * ir, u = d2 * ir;
                const float u2 = u * u;
                float zr[LMAX + 1], zi[LMAX + 1];
                zpow<LMAX>(nx, ny, zr, zi);
                float gx = 0.0f, gy = 0.0f, gz = 0.0f;
                if constexpr (LMASK & kL3) acc_grad<3>(u, u2, zr, zi, C + rec_off(LMASK, 3), gx, gy, gz);
                if constexpr (LMASK & kL4) acc_grad<4>(u, u2, zr, zi, C + rec_off(LMASK, 4), gx, gy, gz);
                if constexpr (LMASK & kL6) acc_grad<6>(u, u2, zr, zi, C + rec_off(LMASK, 6), gx, gy, gz);
                const float dot = gx * nx + gy * ny + gz * u;
                Fx = __builtin_fmaf(ir, __builtin_fmaf(-dot, nx, gx), Fx);
                Fy = __builtin_fmaf(ir, __builtin_fmaf(-dot, ny, gy), Fy);
                Fz = __builtin_fmaf(ir, __builtin_fmaf(-dot, u, gz), Fz);
            }
        }
    }
    Fx = grp_sum(Fx); Fy = grp_sum(Fy); Fz = grp_sum(Fz);
    if (valid && lane == 0) {
        f[3 * (int64_t)i] += (double)Fx;
        f[3 * (int64_t)i + 1] += (double)Fy;
        f[3 * (int64_t)i + 2] += (double)Fz;
    }
    if constexpr (LDL != 0) {
        // Lechner-Dellago lq(i) = |qbar_i|, qbar_i = (q_i + sum_{j in S(i)} q_j) / (n_i + 1),
        // from the neighbour records gathered above
        // group sums on every lane (DPP), then lane 0 alone loads q_i and finishes
        #pragma unroll
        for (int k = 0; k < LW; ++k) qb[k] = grp_sum(qb[k]);
        double part[1] = {0.0};
        if (valid && lane == 0) {
            const bool center = g.group == nullptr || g.group[i];
            const float* qi = e.rec + (int64_t)i * RS + LO;
            const float inv = __builtin_amdgcn_rcpf((float)(e.nsel[i] + 1));
            float s2 = 0.0f;
            #pragma unroll
            for (int k = 0; k < LW; ++k) {
                const float v = (qb[k] + qi[k]) * inv;
                s2 += (k == 0 ? 1.0f : 2.0f) * v * v;
            }
            constexpr float P = (float)(4.0 * kPi / (2 * LDL + 1));
            const float lq = center ? __builtin_amdgcn_sqrtf(P * s2) : 0.0f;
            e.fld_lq[i] = lq;
            part[0] = (g.share == nullptr || g.share[i]) ? (double)lq : 0.0;
        }
        // the staging buffer is free again: its first bytes hold the block-sum scratch
        __syncthreads();
        block_sum<1>(part, reinterpret_cast<double*>(&dst[0][0][0][0]));
        if (threadIdx.x == 0) e.partial[(int64_t)e.fld_slot * g.ntiles + t] = part[0];
    }
}

// ---------------------------------------------------------------------------
// Bias: uniform grid (Catmull-Rom at bin centres, clamp) and exact hills sum.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double cr_value(double p0, double p1, double p2, double p3, double x) {
    return p1 + 0.5 * x * ((p2 - p0) + x * (2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3) +
                           x * x * (-p0 + 3.0 * p1 - 3.0 * p2 + p3));
}
__device__ __forceinline__ double cr_deriv(double p0, double p1, double p2, double p3, double x) {
    return 0.5 * ((p2 - p0) + 2.0 * x * (2.0 * p0 - 5.0 * p1 + 4.0 * p2 - p3) +
                  3.0 * x * x * (-p0 + 3.0 * p1 - 3.0 * p2 + p3));
}
__device__ __forceinline__ void center_frac_clamp(double cv, double lo, double h, int nb, int& i, double& f) {
    double u = (cv - lo) / h - 0.5;
    i = (int)floor(u);
    f = u - i;
    if (i < 1) { i = 1; f = 0.0; }
    if (i > nb - 3) { i = nb - 3; f = 1.0; }
}

// 3-D grid (the sparse-hash bias of GH_t1_sparseHash.cu:283-407 as a dense grid):
// tensor Catmull-Rom taken from the innermost dimension outwards, as the
// reference's recursion does, each derivative divided by its bin width where it
// is taken; NaN -> 0 per component.
__device__ double grid_eval3(const ReduceDev& r, const double* grid, const double* s, double* dV) {
#pragma clang fp contract(off)
    double h[3], f[3];
    int b[3];
    #pragma unroll
    for (int d = 0; d < 3; ++d) {
        h[d] = (r.hi[d] - r.lo[d]) / r.nb[d];
        center_frac_clamp(s[d], r.lo[d], h[d], r.nb[d], b[d], f[d]);
    }
    const int64_t n1 = r.nb[1], n2 = r.nb[2];
    // along dimension 2 for each of the 4x4 (dim 0, dim 1) rows: value and derivative
    double v01[4], d1_0[4], d2_0[4];   // per dim-0 index: CR over dim 1 of (value, value, deriv/h2)
    #pragma unroll
    for (int a = 0; a < 4; ++a) {
        double cv2[4], cd2[4];
        #pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double* p = grid + ((int64_t)(b[0] - 1 + a) * n1 + (b[1] - 1 + c)) * n2 + (b[2] - 1);
            cv2[c] = cr_value(p[0], p[1], p[2], p[3], f[2]);
            cd2[c] = cr_deriv(p[0], p[1], p[2], p[3], f[2]) / h[2];
        }
        v01[a] = cr_value(cv2[0], cv2[1], cv2[2], cv2[3], f[1]);
        d1_0[a] = cr_deriv(cv2[0], cv2[1], cv2[2], cv2[3], f[1]) / h[1];
        d2_0[a] = cr_value(cd2[0], cd2[1], cd2[2], cd2[3], f[1]);
    }
    dV[0] = cr_deriv(v01[0], v01[1], v01[2], v01[3], f[0]) / h[0];
    dV[1] = cr_value(d1_0[0], d1_0[1], d1_0[2], d1_0[3], f[0]);
    dV[2] = cr_value(d2_0[0], d2_0[1], d2_0[2], d2_0[3], f[0]);
    #pragma unroll
    for (int d = 0; d < 3; ++d) if (isnan(dV[d])) dV[d] = 0.0;
    return cr_value(v01[0], v01[1], v01[2], v01[3], f[0]);
}

// grid value and gradient at s (one thread)
__device__ double grid_eval(const ReduceDev& r, const double* grid, const double* s, double* dV) {
#pragma clang fp contract(off)
    if (r.D == 3) return grid_eval3(r, grid, s, dV);
    if (r.D == 1) {
        double h = (r.hi[0] - r.lo[0]) / r.nb[0];
        int i; double x;
        center_frac_clamp(s[0], r.lo[0], h, r.nb[0], i, x);
        double p0 = grid[i - 1], p1 = grid[i], p2 = grid[i + 1], p3 = grid[i + 2];
        double d = cr_deriv(p0, p1, p2, p3, x) / h;
        if (isnan(d)) d = 0.0;
        dV[0] = d;
        return cr_value(p0, p1, p2, p3, x);
    }
    double h0 = (r.hi[0] - r.lo[0]) / r.nb[0], h1 = (r.hi[1] - r.lo[1]) / r.nb[1];
    int i0, j0; double dx, dy;
    center_frac_clamp(s[0], r.lo[0], h0, r.nb[0], i0, dx);
    center_frac_clamp(s[1], r.lo[1], h1, r.nb[1], j0, dy);
    const int n1 = r.nb[1];
    double col[4], dcol[4];
    for (int t = 0; t < 4; ++t) {
        int jj = j0 - 1 + t;
        double p0 = grid[(int64_t)(i0 - 1) * n1 + jj], p1 = grid[(int64_t)i0 * n1 + jj];
        double p2 = grid[(int64_t)(i0 + 1) * n1 + jj], p3 = grid[(int64_t)(i0 + 2) * n1 + jj];
        col[t] = cr_value(p0, p1, p2, p3, dx);
        dcol[t] = cr_deriv(p0, p1, p2, p3, dx);
    }
    dV[0] = cr_value(dcol[0], dcol[1], dcol[2], dcol[3], dy) / h0;
    dV[1] = cr_deriv(col[0], col[1], col[2], col[3], dy) / h1;
    if (isnan(dV[0]) || isnan(dV[1])) { dV[0] = 0.0; dV[1] = 0.0; }
    return cr_value(col[0], col[1], col[2], col[3], dy);
}

// Deposit one hill into a grid with the block's threads (add_to_grid).
__device__ void grid_deposit_block(const ReduceDev& r, double* grid, const double* s, double w, double sig) {
#pragma clang fp contract(off)
    int lower[kMaxDim] = {0, 0, 0}, upper[kMaxDim] = {0, 0, 0};
    for (int k = 0; k < r.D; ++k) {
        double dcv = (r.hi[k] - r.lo[k]) / r.nb[k];
        int R = (int)ceil(4.0 * sig / dcv) + 1;
        int c = (int)((s[k] - r.lo[k]) / dcv);
        lower[k] = c - R > 0 ? c - R : 0;
        upper[k] = c + R < r.nb[k] - 1 ? c + R : r.nb[k] - 1;
    }
    if (r.D == 1) {
        for (int g = lower[0] + threadIdx.x; g <= upper[0]; g += blockDim.x) {
            double ctr = r.lo[0] + (g + 0.5) * (r.hi[0] - r.lo[0]) / r.nb[0];
            double dx = s[0] - ctr;
            grid[g] += w * exp(-0.5 * (dx * dx) / (sig * sig));
        }
    } else if (r.D == 3) {
        // recursive_add2grid / gauss_calc<3> (GH_t1_sparseHash.cu:242-262,425-431):
        // bin centre lo + (i + 1/2) dcv, exp(-0.5 sum_d dx_d^2 / sigma^2)
        const int n0 = upper[0] - lower[0] + 1, n1 = upper[1] - lower[1] + 1, n2 = upper[2] - lower[2] + 1;
        const double h0 = (r.hi[0] - r.lo[0]) / r.nb[0], h1 = (r.hi[1] - r.lo[1]) / r.nb[1];
        const double h2 = (r.hi[2] - r.lo[2]) / r.nb[2];
        for (int q = threadIdx.x; q < n0 * n1 * n2; q += blockDim.x) {
            if (n0 <= 0 || n1 <= 0 || n2 <= 0) break;
            const int i = lower[0] + q / (n1 * n2), j = lower[1] + (q / n2) % n1, k = lower[2] + q % n2;
            const double d0 = s[0] - (r.lo[0] + (i + 0.5) * h0);
            const double d1 = s[1] - (r.lo[1] + (j + 0.5) * h1);
            const double d2 = s[2] - (r.lo[2] + (k + 0.5) * h2);
            const double paw = d0 * d0 + d1 * d1 + d2 * d2;
            grid[((int64_t)i * r.nb[1] + j) * r.nb[2] + k] += w * exp(-0.5 * (paw) / (sig * sig));
        }
    } else {
        const int n0 = upper[0] - lower[0] + 1, n1 = upper[1] - lower[1] + 1;
        for (int q = threadIdx.x; q < n0 * n1; q += blockDim.x) {
            if (n0 <= 0 || n1 <= 0) break;
            const int i = lower[0] + q / n1, j = lower[1] + q % n1;
            double d0 = s[0] - (r.lo[0] + (i + 0.5) * (r.hi[0] - r.lo[0]) / r.nb[0]);
            double d1 = s[1] - (r.lo[1] + (j + 0.5) * (r.hi[1] - r.lo[1]) / r.nb[1]);
            grid[(int64_t)i * r.nb[1] + j] += w * exp(-0.5 * (d0 * d0 + d1 * d1) / (sig * sig));
        }
    }
}

// A wave-uniform FP64 value moved to scalar registers (both halves from the first lane).
__device__ __forceinline__ double uniform_f64(double v) {
    const int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)__double_as_longlong(v));
    const int hi = __builtin_amdgcn_readfirstlane((int)(uint32_t)(__double_as_longlong(v) >> 32));
    return __longlong_as_double(((long long)(uint32_t)hi << 32) | (uint32_t)lo);
}

// One thread's share of the exact FP64 hills sum at s: hills first, first + G,
// first + 2G, ... below nh, in that order, added to acc = (V, dV/ds).
__device__ __forceinline__ void hills_thread_sum(const ReduceDev& r, int w, const double* s, int64_t first,
                                                 int64_t G, double* acc) {
    const double* H = r.hills + (int64_t)w * r.hills_cap * (r.D + 3);
    const int64_t nh = r.nhills[w];
    // four records loaded ahead of their arithmetic (the loop is latency-bound one
    // record at a time)
    constexpr int HU = 4;
    // 1/(2 sigma^2) per dimension, once (FP64 divisions cost more than the exp); twice
    // it is 1/sigma^2 exactly.  The values are uniform: held in scalar registers.
    double i2s2[kMaxDim] = {0, 0, 0};
    #pragma unroll
    for (int d = 0; d < kMaxDim; ++d) if (d < r.D) i2s2[d] = uniform_f64(0.5 / (r.sigma[d] * r.sigma[d]));
    // A hill's widths are the bias widths scaled by its own sigma over the first bias
    // width (a hill replayed from a HILLS file keeps the file's sigma); deposits carry
    // the first width, so their scale is exactly 1 and costs no division.
    const double sig0 = uniform_f64(r.sigma[0]);
    for (int64_t h = first; h < nh; h += HU * G) {
        double c[HU][kMaxDim], hgt[HU], hs[HU];
        #pragma unroll
        for (int u = 0; u < HU; ++u) {
            const int64_t hu = h + u * G < nh ? h + u * G : h;
            const double* rec = H + hu * (r.D + 3);
            #pragma unroll
            for (int d = 0; d < kMaxDim; ++d) c[u][d] = d < r.D ? rec[1 + d] : 0.0;
            hgt[u] = rec[1 + r.D];
            hs[u] = rec[2 + r.D];
        }
        #pragma unroll
        for (int u = 0; u < HU; ++u) {
            if (h + u * G >= nh) break;
            const double sc = hs[u] == sig0 ? 1.0 : (sig0 * sig0) / (hs[u] * hs[u]);
            double ex = 0, dd[kMaxDim] = {0, 0, 0};
            #pragma unroll
            for (int d = 0; d < kMaxDim; ++d) if (d < r.D) {
                dd[d] = s[d] - c[u][d];
                ex += dd[d] * dd[d] * i2s2[d];
            }
            double t = hgt[u] * exp(-ex * sc);
            acc[0] += t;
            #pragma unroll
            for (int d = 0; d < kMaxDim; ++d) if (d < r.D) acc[1 + d] -= t * dd[d] * (2.0 * i2s2[d] * sc);
        }
    }
}

// Exact FP64 hills sum at s, block-cooperative; V and dV valid on thread 0.
template <int NW>
__device__ double hills_eval_block(const ReduceDev& r, int w, const double* s, double* dV, double* sh,
                                   int64_t h_first = 0) {
    double acc[4] = {0, 0, 0, 0};
    hills_thread_sum(r, w, s, h_first + threadIdx.x, blockDim.x, acc);
    block_sum<4, NW>(acc, sh);
    if (threadIdx.x == 0) {
        #pragma unroll
        for (int d = 0; d < kMaxDim; ++d) if (d < r.D) dV[d] = acc[1 + d];
    }
    __syncthreads();
    return acc[0];
}

// Learned bias network at one point, block-cooperative (threads 0..63 are the
// hidden units), FP32 master weights (the MFMA kernels carry W2 as BF16 hi + lo,
// the same function to ~1e-5).  V and dV/ds valid on thread 0.
__device__ double mlp_eval_block(const MlpNet* net, const double* s, double* dV) {
    __shared__ float h1s[kMlpH], g2s[kMlpH], ys[kMlpH], gis[3][kMlpH];
    const int a = threadIdx.x;
    float x[3];
    #pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = (float)((s[d] - net->in_shift[d]) * net->in_scale[d]);
    if (a < kMlpH) {
        float z = net->p[kB1 + a];
        #pragma unroll
        for (int d = 0; d < 3; ++d) z = __builtin_fmaf(net->p[kW1 + 3 * a + d], x[d], z);
        h1s[a] = mlp_tanh(z);
    }
    __syncthreads();
    if (a < kMlpH) {
        float z = net->p[kB2 + a];
        for (int b = 0; b < kMlpH; ++b) z = __builtin_fmaf(net->p[kW2 + a * kMlpH + b], h1s[b], z);
        const float h = mlp_tanh(z);
        ys[a] = net->p[kW3 + a] * h;
        g2s[a] = net->p[kW3 + a] * (1.0f - h * h);
    }
    __syncthreads();
    if (a < kMlpH) {
        float t = 0.0f;
        for (int c = 0; c < kMlpH; ++c) t = __builtin_fmaf(g2s[c], net->p[kW2 + c * kMlpH + a], t);
        t *= 1.0f - h1s[a] * h1s[a];
        #pragma unroll
        for (int d = 0; d < 3; ++d) gis[d][a] = t * net->p[kW1 + 3 * a + d];
    }
    __syncthreads();
    double V = 0.0;
    if (a == 0) {
        float y = net->p[kB3];
        for (int c = 0; c < kMlpH; ++c) y += ys[c];
        for (int d = 0; d < 3; ++d) {
            float gsum = 0.0f;
            for (int c = 0; c < kMlpH; ++c) gsum += gis[d][c];
            dV[d] = (double)gsum * net->in_scale[d] * net->out_scale;
        }
        V = (double)y * net->out_scale + net->out_shift;
    }
    __syncthreads();
    return V;
}

// V(s) and dV/ds of walker w's bias, block-cooperative; valid on thread 0.
template <int NW>
__device__ double bias_eval_block(const ReduceDev& r, int w, const double* s, double* dV, double* sh) {
    if (r.kind == BIAS_GRID) {
        double v = 0.0;
        if (threadIdx.x == 0) v = grid_eval(r, r.grid + (int64_t)w * r.grid_size, s, dV);
        return v;
    }
    if (r.kind == BIAS_MLP && r.mlp_ok[w]) {
        // network for the fitted hills + exact sum of the hills deposited since
        const MlpNet* net = r.mlp + w;
        double dn[kMaxDim] = {0, 0, 0};
        const double vn = mlp_eval_block(net, s, dn);
        const double vh = hills_eval_block<NW>(r, w, s, dV, sh, net->n_fit);
        if (threadIdx.x == 0) {
            #pragma unroll
            for (int d = 0; d < kMaxDim; ++d) if (d < r.D) dV[d] += dn[d];
        }
        return vn + vh;
    }
    return hills_eval_block<NW>(r, w, s, dV, sh);
}

// ---------------------------------------------------------------------------
// Expression evaluation with forward-mode derivatives over the symbols.
// ---------------------------------------------------------------------------
struct Dual { double v; double d[kMaxSym]; };

// st: kStack entries of scratch (LDS in k_reduce: an interpreter stack is dynamically indexed).
// Kept out of line: its FP64 math constants would otherwise occupy the caller's registers.
__device__ __noinline__ void expr_eval(const ExprProg& p, int nsym, const double* sym, double* val, double* dval, Dual* st) {
    int sp = 0;
    for (int k = 0; k < p.n; ++k) {
        const int op = p.op[k];
        if (op == OP_CONST || op == OP_SYM) {
            Dual& a = st[sp++];
            for (int q = 0; q < kMaxSym; ++q) a.d[q] = 0.0;
            if (op == OP_CONST) a.v = p.c[k];
            else { int q = (int)p.c[k]; a.v = sym[q]; a.d[q] = 1.0; }
            continue;
        }
        if ((op >= OP_ADD && op <= OP_POW) || op == OP_MIN || op == OP_MAX) {
            const Dual& b = st[--sp];
            Dual& a = st[sp - 1];
            double av = a.v, bv = b.v, da = 0, db = 0, rv = 0;
            switch (op) {
                case OP_ADD: rv = av + bv; da = 1; db = 1; break;
                case OP_SUB: rv = av - bv; da = 1; db = -1; break;
                case OP_MUL: rv = av * bv; da = bv; db = av; break;
                case OP_DIV: rv = av / bv; da = 1.0 / bv; db = -av / (bv * bv); break;
                case OP_POW: rv = pow(av, bv); da = bv * pow(av, bv - 1.0);
                             db = av > 0 ? rv * log(av) : 0.0; break;
                case OP_MIN: rv = av <= bv ? av : bv; da = av <= bv; db = !(av <= bv); break;
                case OP_MAX: rv = av >= bv ? av : bv; da = av >= bv; db = !(av >= bv); break;
            }
            for (int q = 0; q < kMaxSym; ++q) a.d[q] = da * a.d[q] + db * b.d[q];
            a.v = rv;
            continue;
        }
        Dual& a = st[sp - 1];
        double x = a.v, rv = 0, dx = 0;
        switch (op) {
            case OP_NEG: rv = -x; dx = -1; break;
            case OP_SIN: rv = sin(x); dx = cos(x); break;
            case OP_COS: rv = cos(x); dx = -sin(x); break;
            case OP_TAN: rv = tan(x); dx = 1.0 + rv * rv; break;
            case OP_EXP: rv = exp(x); dx = rv; break;
            case OP_LOG: rv = log(x); dx = 1.0 / x; break;
            case OP_LOG10: rv = log10(x); dx = 1.0 / (x * 2.302585092994046); break;
            case OP_SQRT: rv = sqrt(x); dx = 0.5 / rv; break;
            case OP_ABS: rv = fabs(x); dx = x >= 0 ? 1.0 : -1.0; break;
            case OP_TANH: rv = tanh(x); dx = 1.0 - rv * rv; break;
            case OP_ASIN: rv = asin(x); dx = 1.0 / sqrt(1.0 - x * x); break;
            case OP_ACOS: rv = acos(x); dx = -1.0 / sqrt(1.0 - x * x); break;
            case OP_ATAN: rv = atan(x); dx = 1.0 / (1.0 + x * x); break;
            case OP_SINH: rv = sinh(x); dx = cosh(x); break;
            case OP_COSH: rv = cosh(x); dx = sinh(x); break;
        }
        for (int q = 0; q < kMaxSym; ++q) a.d[q] *= dx;
        a.v = rv;
    }
    *val = st[0].v;
    for (int q = 0; q < nsym; ++q) dval[q] = st[0].d[q];
}

// Wave 0 of the reduce kernel writes the walker's state and chain-rule weights:
// lane c stores CV slot c's value and its weight w_c = sum_d dV/ds_d ds_d/dv_c
// (ac = lane c's ds_d/dv_c); V, dV/ds and s are taken from lane 0.
__device__ __forceinline__ void reduce_finish(const ReduceDev& r, int w, int c, double cvv, const double* ac,
                                              const double* sv, double V, const double* dV, double hill,
                                              int64_t step, double mlp_err, int mlp_ok, double inv_c) {
    V = readlane_f64(V, 0);
    double dv[kMaxDim];
    #pragma unroll
    for (int d = 0; d < kMaxDim; ++d) dv[d] = readlane_f64(dV[d], 0);
    double* st = r.state + (int64_t)w * kStStride;
    if (c < r.ncv) st[kStCv + c] = cvv;
    double wc = 0.0;
    if (c < kMaxSlotCV) {
        #pragma unroll
        for (int d = 0; d < kMaxDim; ++d) if (d < r.D) wc += dv[d] * ac[d];
        r.wcv[(int64_t)w * kWcvStride + c] = wc;
    }
    // the weight sum of each representative slot over the slots aliased to it, in slot
    // order, each weight times its CV's 1/N_group (a slot that nothing aliases gets its
    // own): the force passes of per-group averages take it
    const double wn = c < kMaxSlotCV ? wc * inv_c : 0.0;
    double comb = 0.0;
    #pragma unroll
    for (int q = 0; q < kMaxSlotCV; ++q) {
        const double v = __shfl(wn, q, 64);
        if (r.walias[q] == c) comb += v;
    }
    if (c < kMaxSlotCV) r.wcv[(int64_t)w * kWcvStride + kMaxSlotCV + c] = comb;
    if (c == 0) {
        #pragma unroll
        for (int d = 0; d < kMaxDim; ++d) if (d < r.D) { st[kStS + d] = sv[d]; st[kStDV + d] = dv[d]; }
        st[kStV] = V;
        st[kStHill] = hill;
        st[kStStep] = (double)step;
        st[kStMlpErr] = mlp_err;
        st[kStMlpOk] = (double)mlp_ok;
        r.step[w] = step + 1;
    }
}

// ---------------------------------------------------------------------------
// k_reduce: one workgroup per walker.  Post-force order of
// FixMetadynamics::post_force (src/core/fix_crystallize.cu:409-450):
// CVs -> expression -> deposit on step % pace == 0 (WT height from the bias
// before the deposit) -> V and dV/ds after the deposit -> chain rule.  A hill-log
// bias (exact sum or learned) ends this kernel at s: k_hills_part and k_bias_tail
// do the rest of the step.
//
// The kernel is one chain of dependent memory round trips, so it is laid out
// to have few: the partial sums load in the first one (tile ranges come with
// the arguments), together with the per-walker scalars, the affine expression
// coefficients and a small grid (to LDS).  After the CV sums, wave 0 carries
// one CV slot per lane: an affine DIM value is a wave sum and the chain-rule
// weight of slot c is lane c's own product, so no single-thread chains of
// LDS reads remain on the usual path.
// ---------------------------------------------------------------------------
constexpr int kRedThreads = 1024;
// Partial-sum slot of a CV source: a slot >= 0 is summed from this step's tile partials
// (k_env, k_ld, k_lq_sum); -1 is a direct value; -2 - p is slot p of an unbiased local Q
// fused into the force pass, which runs after this kernel, so the sum taken here is
// the previous step's (the same fixed-order tree as k_cv_tail, which gives the current
// step's value when the state is read).
__device__ __forceinline__ int partial_slot(int src) { return src >= 0 ? src : (src <= -2 ? -2 - src : -1); }
constexpr int kRedGridLds = 4096;     // grids up to this many bins are read from an LDS copy
// AFFINE: every DIM program is affine (r.affine), so the expression interpreter, an
// out-of-line call with a stack frame, is left out of that instance.
template <bool AFFINE>
__global__ __launch_bounds__(kRedThreads) void k_reduce(ReduceDev r) {
    const int w = blockIdx.x;
    __shared__ double sh[kMaxSlotCV * (kRedThreads / 64)];
    __shared__ ExprProg progs[kMaxDim];
    __shared__ double cv[kMaxSlotCV];
    __shared__ double s[kMaxDim];
    __shared__ double dsdv[kMaxDim * kMaxSym];
    __shared__ double bc[8];
    __shared__ Dual stack[kStack];
    __shared__ double symv[kMaxSym];
    __shared__ double gl[kRedGridLds];
    constexpr int U = 4;
    constexpr int SG = 4;
    constexpr int NW = kRedThreads / 64;
    const int lane64 = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // one step per launch: the near-list epoch moves on (every kernel that reads it this
    // step, the packs, the near-list build and the selection, has completed)
    if (w == 0 && threadIdx.x == 0 && r.epoch) *r.epoch += 1;
    const int t0 = r.tile_args ? r.tile_range[w] : r.tile_begin[w];
    const int t1 = r.tile_args ? r.tile_range[w + 1] : r.tile_begin[w + 1];
    const int64_t step = r.step[w];
    const int64_t nhill0 = r.nhills[w];
    const double mlp_err = r.kind == BIAS_MLP ? r.mlp_err[w] : 0.0;
    const int mlp_ok = r.kind == BIAS_MLP ? r.mlp_ok[w] : 0;
    // wave 0, lane c: CV slot c's source, direct value and affine coefficients
    const bool slot_lane = wv == 0 && lane64 < kMaxSlotCV;
    int src_c = -1;
    #pragma unroll
    for (int c = 0; c < kMaxSlotCV; ++c) if (lane64 == c) src_c = r.cv_src[c];
    const double direct_c = slot_lane && lane64 < r.ncv && src_c == -1 ? r.direct[(int64_t)w * kMaxSlotCV + lane64] : 0.0;
    // lane c's 1/N, loaded with the first round trip
    const double inv_c = slot_lane ? r.cv_inv[(int64_t)w * kMaxSlotCV + lane64] : 0.0;
    double acoef[kMaxDim] = {0, 0, 0};
    if (AFFINE && slot_lane) {
        #pragma unroll
        for (int d = 0; d < kMaxDim; ++d) if (d < r.D) acoef[d] = r.aff[d * kMaxSlotCV + lane64];
    }
    if (!AFFINE) {
        // expression programs to LDS for the interpreter (thread 0 below)
        const int nwords = r.D * (int)(sizeof(ExprProg) / 4);
        const int* src = reinterpret_cast<const int*>(r.prog);
        int* dst = reinterpret_cast<int*>(progs);
        for (int k = threadIdx.x; k < nwords; k += blockDim.x) dst[k] = src[k];
    }
    // a small grid is loaded to registers here and stored to LDS after the partial
    // sums, so both sets of loads are in flight together
    constexpr int GU = kRedGridLds / kRedThreads;
    const bool grid_in_lds = r.kind == BIAS_GRID && r.grid_size <= kRedGridLds;
    double gr[GU];
    if (grid_in_lds) {
        const double* src = r.grid + (int64_t)w * r.grid_size;
        #pragma unroll
        for (int u = 0; u < GU; ++u) {
            const int k = threadIdx.x + u * kRedThreads;
            gr[u] = k < r.grid_size ? src[k] : 0.0;
        }
    }
    // 1. CV values: one deterministic FP64 tree over tiles per slot (each thread
    //    sums its tiles in a fixed order, then waves, then the wave totals in
    //    order).  Up to four slots at a time have all their loads in flight
    //    together; a 1024-thread block has 128 VGPRs per lane.
    #pragma unroll
    for (int c0 = 0; c0 < kMaxSlotCV; c0 += SG) {
        if (c0 >= r.ncv) break;
        double acc[SG];
        #pragma unroll
        for (int g = 0; g < SG; ++g) acc[g] = 0.0;
        for (int t = t0 + threadIdx.x; t < t1; t += U * kRedThreads) {
            double x[SG][U];
            #pragma unroll
            for (int g = 0; g < SG; ++g) {
                const int c = c0 + g;
                const int src = c < r.ncv ? partial_slot(r.cv_src[c]) : -1;
                const double* pp = r.partial + (int64_t)(src >= 0 ? src : 0) * r.ntiles;
                #pragma unroll
                for (int u = 0; u < U; ++u)
                    x[g][u] = (src >= 0 && t + u * kRedThreads < t1) ? pp[t + u * kRedThreads] : 0.0;
            }
            #pragma unroll
            for (int g = 0; g < SG; ++g) acc[g] += (x[g][0] + x[g][1]) + (x[g][2] + x[g][3]);
        }
        #pragma unroll
        for (int g = 0; g < SG; ++g) {
            const double v = wave_sum(acc[g]);
            if (lane64 == 0 && c0 + g < r.ncv) sh[(c0 + g) * NW + wv] = v;
        }
    }
    if (grid_in_lds) {
        #pragma unroll
        for (int u = 0; u < GU; ++u) {
            const int k = threadIdx.x + u * kRedThreads;
            if (k < r.grid_size) gl[k] = gr[u];
        }
    }
    __syncthreads();
    const bool dep = r.pace > 0 && (step % r.pace) == 0;
    if (wv == 0) {
        // wave 0: lane c holds CV slot c
        const int c = lane64;
        double cvv = 0.0;
        if (c < r.ncv) {
            if (partial_slot(src_c) >= 0) {
                double a = 0.0;
                #pragma unroll
                for (int q = 0; q < NW; ++q) a += sh[c * NW + q];
                cvv = a * inv_c;
            } else {
                cvv = direct_c;                  // a direct value (distance, S(k))
            }
            cv[c] = cvv;
        }
        // 2. dimensions
        double sv[kMaxDim] = {0, 0, 0};
        if constexpr (AFFINE) {
            #pragma unroll
            for (int d = 0; d < kMaxDim; ++d) if (d < r.D) sv[d] = r.aff0[d] + wave_sum(acoef[d] * cvv);
            if (c == 0) {
                #pragma unroll
                for (int d = 0; d < kMaxDim; ++d) if (d < r.D) s[d] = sv[d];
            }
        } else {
            wave_lds_sync();
            if (c == 0) {
                for (int q = 0; q < r.nsym; ++q) symv[q] = cv[r.sym_slot[q]];
                for (int d = 0; d < r.D; ++d) expr_eval(progs[d], r.nsym, symv, &s[d], &dsdv[d * kMaxSym], stack);
            }
            wave_lds_sync();
            #pragma unroll
            for (int d = 0; d < kMaxDim; ++d) {
                if (d >= r.D) continue;
                sv[d] = s[d];
                // d s_d / d v_c: the derivatives of every symbol bound to slot c
                double a = 0.0;
                #pragma unroll
                for (int q = 0; q < kMaxSym; ++q)
                    if (q < r.nsym && r.sym_slot[q] == c) a += dsdv[d * kMaxSym + q];
                acoef[d] = a;
            }
        }
        if (r.kind == BIAS_GRID && !dep) {
            // a step without a deposit on a grid: lane 0 evaluates, wave 0 finishes;
            // no further workgroup barrier (uniform over the block)
            double V = 0.0, dVg[kMaxDim] = {0, 0, 0};
            if (c == 0) V = grid_eval(r, grid_in_lds ? gl : r.grid + (int64_t)w * r.grid_size, sv, dVg);
            reduce_finish(r, w, c, cvv, acoef, sv, V, dVg, 0.0, step, mlp_err, mlp_ok, inv_c);
            return;
        }
        if (r.kind != BIAS_GRID) {
            // a hill log: its sum is spread over many workgroups (k_hills_part) and the
            // step ends in k_bias_tail, which takes the slot values, s and ds/dv from here
            double* st = r.state + (int64_t)w * kStStride;
            double* bw = r.bwork + (int64_t)w * kBwStride;
            if (c < r.ncv) st[kStCv + c] = cvv;
            if (c < kMaxSlotCV) {
                #pragma unroll
                for (int d = 0; d < kMaxDim; ++d) bw[kBwAc + d * kMaxSlotCV + c] = acoef[d];
            }
            if (c == 0) {
                #pragma unroll
                for (int d = 0; d < kMaxDim; ++d) if (d < r.D) st[kStS + d] = sv[d];
            }
            return;
        }
    } else if (r.kind != BIAS_GRID || !dep) {
        return;
    }
    // a grid bias on a deposit step from here on
    // keep the slot values and the chain-rule coefficients of wave 0 across the
    // block-cooperative deposit below
    __shared__ double acoef_s[kMaxDim][kMaxSlotCV];
    if (slot_lane) {
        #pragma unroll
        for (int d = 0; d < kMaxDim; ++d) acoef_s[d][lane64] = acoef[d];
    }
    __syncthreads();
    // 3. deposit
    double* gw = r.grid + (int64_t)w * r.grid_size;
    double dV[kMaxDim] = {0, 0, 0};
    if (threadIdx.x == 0) {
        const double Vb = r.wt ? grid_eval(r, gw, s, dV) : 0.0;
        bc[0] = r.h0 * exp(-(Vb) / (r.kT * (r.biasf - 1.0)));
    }
    __syncthreads();
    const double hgt = bc[0];
    grid_deposit_block(r, gw, s, hgt, r.sigma[0]);
    if (threadIdx.x == 0) {
        int64_t k = nhill0;
        if (k < r.hills_cap) {
            double* rec = r.hills + ((int64_t)w * r.hills_cap + k) * (r.D + 3);
            rec[0] = (double)step;
            for (int d = 0; d < r.D; ++d) rec[1 + d] = s[d];
            rec[1 + r.D] = hgt;
            rec[2 + r.D] = r.sigma[0];
            r.nhills[w] = k + 1;
        }
    }
    __syncthreads();
    // 4. bias value and gradient after the deposit (valid on thread 0)
    double V = 0.0;
    if (threadIdx.x == 0) V = grid_eval(r, gw, s, dV);
    if (wv == 0) {
        const int c = lane64;
        const double cvv = c < r.ncv ? cv[c] : 0.0;
        double ac[kMaxDim] = {0, 0, 0}, sv[kMaxDim] = {0, 0, 0};
        #pragma unroll
        for (int d = 0; d < kMaxDim; ++d) if (d < r.D) { ac[d] = c < kMaxSlotCV ? acoef_s[d][c] : 0.0; sv[d] = s[d]; }
        reduce_finish(r, w, c, cvv, ac, sv, V, dV, hgt, step, mlp_err, mlp_ok, inv_c);
    }
}

// Partial sums of a walker's hill log at the walker's s of this step: workgroup b of
// walker w (blockIdx.y) takes hills first + b*256 + t + k*kHillBlocks*256 and writes
// (V, dV/ds) to partial slot b.  A trusted learned bias takes only the hills deposited
// since its fit (first = n_fit); otherwise first = 0.
constexpr int kHillThreads = 256;
__global__ __launch_bounds__(kHillThreads) void k_hills_part(ReduceDev r) {
    __shared__ double sh[4 * (kHillThreads / 64)];
    const int w = blockIdx.y, b = blockIdx.x;
    const double* st = r.state + (int64_t)w * kStStride;
    double s[kMaxDim] = {0, 0, 0};
    #pragma unroll
    for (int d = 0; d < kMaxDim; ++d) if (d < r.D) s[d] = st[kStS + d];
    const int64_t first = (r.kind == BIAS_MLP && r.mlp_ok[w]) ? r.mlp[w].n_fit : 0;
    double acc[4] = {0, 0, 0, 0};
    hills_thread_sum(r, w, s, first + (int64_t)b * kHillThreads + threadIdx.x, (int64_t)kHillBlocks * kHillThreads, acc);
    block_sum<4, kHillThreads / 64>(acc, sh);
    if (threadIdx.x == 0) {
        double* part = r.bwork + (int64_t)w * kBwStride + 4 * b;
        #pragma unroll
        for (int k = 0; k < 4; ++k) part[k] = acc[k];
    }
}

// The rest of a step on a hill-log bias, one wave per walker: the partial sums of
// k_hills_part in slot order (a fixed tree), plus the network term of a trusted learned
// bias, is the bias before the deposit; a deposit takes its well-tempered height from
// it, and the new hill adds exactly its height to V at its own centre and nothing to
// dV/ds there; then the chain-rule weights and the walker's state.
__global__ __launch_bounds__(64) void k_bias_tail(ReduceDev r) {
    const int w = blockIdx.x, c = threadIdx.x;
    const int64_t step = r.step[w];
    const int64_t nhill0 = r.nhills[w];
    const double mlp_err = r.kind == BIAS_MLP ? r.mlp_err[w] : 0.0;
    const int mlp_ok = r.kind == BIAS_MLP ? r.mlp_ok[w] : 0;
    const double* st = r.state + (int64_t)w * kStStride;
    const double* bw = r.bwork + (int64_t)w * kBwStride;
    const double inv_c = c < kMaxSlotCV ? r.cv_inv[(int64_t)w * kMaxSlotCV + c] : 0.0;
    const double cvv = c < r.ncv ? st[kStCv + c] : 0.0;
    double sv[kMaxDim] = {0, 0, 0}, ac[kMaxDim] = {0, 0, 0};
    #pragma unroll
    for (int d = 0; d < kMaxDim; ++d) if (d < r.D) {
        sv[d] = st[kStS + d];
        ac[d] = c < kMaxSlotCV ? bw[kBwAc + d * kMaxSlotCV + c] : 0.0;
    }
    double a[4] = {0, 0, 0, 0};
    for (int b = c; b < kHillBlocks; b += 64) {
        #pragma unroll
        for (int k = 0; k < 4; ++k) a[k] += bw[4 * b + k];
    }
    double Vb = wave_sum(a[0]);
    double dV[kMaxDim] = {0, 0, 0};
    #pragma unroll
    for (int d = 0; d < kMaxDim; ++d) if (d < r.D) dV[d] = wave_sum(a[1 + d]);
    if (mlp_ok) {
        double dn[kMaxDim] = {0, 0, 0};
        const double vn = mlp_eval_block(r.mlp + w, sv, dn);
        if (c == 0) {
            Vb += vn;
            #pragma unroll
            for (int d = 0; d < kMaxDim; ++d) if (d < r.D) dV[d] += dn[d];
        }
        Vb = readlane_f64(Vb, 0);
    }
    const bool dep = r.pace > 0 && (step % r.pace) == 0;
    double hill = 0.0, V = Vb;
    if (dep) {
        hill = r.h0 * exp(-(r.wt ? Vb : 0.0) / (r.kT * (r.biasf - 1.0)));
        if (nhill0 < r.hills_cap) {
            if (c == 0) {
                double* rec = r.hills + ((int64_t)w * r.hills_cap + nhill0) * (r.D + 3);
                rec[0] = (double)step;
                for (int d = 0; d < r.D; ++d) rec[1 + d] = sv[d];
                rec[1 + r.D] = hill;
                rec[2 + r.D] = r.sigma[0];
                r.nhills[w] = nhill0 + 1;
            }
            V = Vb + hill;
        }
    }
    reduce_finish(r, w, c, cvv, ac, sv, V, dV, hill, step, mlp_err, mlp_ok, inv_c);
}

// Values of the CVs whose partial sums come from the force pass (unbiased local-Q
// fused into k_force): one 1024-thread workgroup per walker, four independent loads
// in flight per thread, the same fixed-order tree as the reduce kernel.
__global__ __launch_bounds__(1024) void k_cv_tail(ReduceDev r) {
    constexpr int NT = 1024, U = 4;
    __shared__ double sh[NT / 64];
    const int w = blockIdx.x;
    const int t0 = r.tile_begin[w], t1 = r.tile_begin[w + 1];
    for (int c = 0; c < r.ncv; ++c) {
        if (r.cv_src[c] > -2) continue;
        const double* pp = r.partial + (int64_t)(-2 - r.cv_src[c]) * r.ntiles;
        double v[1] = {0.0};
        for (int t = t0 + threadIdx.x; t < t1; t += U * NT) {
            double x[U];
            #pragma unroll
            for (int u = 0; u < U; ++u) x[u] = t + u * NT < t1 ? pp[t + u * NT] : 0.0;
            v[0] += (x[0] + x[1]) + (x[2] + x[3]);
        }
        block_sum<1, NT / 64>(v, sh);
        if (threadIdx.x == 0) r.state[(int64_t)w * kStStride + kStCv + c] = v[0] * r.cv_inv[(int64_t)w * kMaxSlotCV + c];
        __syncthreads();
    }
}

// Raw sums of the tile partials, one 1024-thread workgroup per walker, in the tree
// order of k_cv_tail: the context's share of every per-atom-sum CV, before the sum over
// contexts and the division by the group count.
__global__ __launch_bounds__(1024) void k_cv_sums(ReduceDev r, const int32_t* __restrict__ psum, double* __restrict__ out) {
    constexpr int NT = 1024, U = 4;
    __shared__ double sh[NT / 64];
    const int w = blockIdx.x;
    const int t0 = r.tile_begin[w], t1 = r.tile_begin[w + 1];
    for (int c = 0; c < r.ncv; ++c) {
        if (psum[c] < 0) continue;
        const double* pp = r.partial + (int64_t)psum[c] * r.ntiles;
        double v[1] = {0.0};
        for (int t = t0 + threadIdx.x; t < t1; t += U * NT) {
            double x[U];
            #pragma unroll
            for (int u = 0; u < U; ++u) x[u] = t + u * NT < t1 ? pp[t + u * NT] : 0.0;
            v[0] += (x[0] + x[1]) + (x[2] + x[3]);
        }
        block_sum<1, NT / 64>(v, sh);
        if (threadIdx.x == 0) out[(int64_t)w * kMaxSlotCV + c] = v[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_bias_eval(ReduceDev r, int w, int64_t np, const double* s,
                                                      double* V, double* dV) {
    __shared__ double sh[64];
    for (int64_t p = blockIdx.x; p < np; p += gridDim.x) {
        double sp[kMaxDim] = {0, 0, 0};
        #pragma unroll
        for (int d = 0; d < kMaxDim; ++d) if (d < r.D) sp[d] = s[p * r.D + d];
        double g[kMaxDim] = {0, 0, 0};
        double v = bias_eval_block<kBlock / 64>(r, w, sp, g, sh);
        if (threadIdx.x == 0) {
            if (V) V[p] = v;
            #pragma unroll
            for (int d = 0; d < kMaxDim; ++d) if (dV && d < r.D) dV[p * r.D + d] = g[d];
        }
        __syncthreads();
    }
}

__global__ void k_bias_add(ReduceDev r, int w, int64_t nh, const double* s, const double* h, double sig) {
    for (int64_t k = 0; k < nh; ++k) {
        double sp[kMaxDim];
        for (int d = 0; d < r.D; ++d) sp[d] = s[k * r.D + d];
        if (r.kind == BIAS_GRID) {
            grid_deposit_block(r, r.grid + (int64_t)w * r.grid_size, sp, h[k], sig);
        } else if (threadIdx.x == 0) {
            int64_t q = r.nhills[w];
            if (q < r.hills_cap) {
                double* rec = r.hills + ((int64_t)w * r.hills_cap + q) * (r.D + 3);
                rec[0] = -1.0;
                for (int d = 0; d < r.D; ++d) rec[1 + d] = sp[d];
                rec[1 + r.D] = h[k];
                rec[2 + r.D] = sig;
                r.nhills[w] = q + 1;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// FES reconstruction from a hills list (tools/plot_hills.py:158-235): one thread
// per grid point sums the hills in list order (the order of the tool's numpy
// reduction over hills).  1-D: every hill at every point, exp(-(dx^2) / (2 s^2));
// 2-D: a hill covers the bins of its window int((c -+ 4 s - lo) / d), truncated
// toward zero and clamped to the grid, exp(-(dx^2 + dy^2) / (2 s^2)).  hills rows:
// step, s..., height, sigma.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int fes_trunc(double v) {
    return (int)fmin(fmax(v, -2.0e9), 2.0e9);      // int() of the tool, range-limited
}
__global__ __launch_bounds__(kBlock) void k_fes_sum(int D, int64_t nh, const double* __restrict__ hills,
                                                    const double* __restrict__ x, int n0,
                                                    const double* __restrict__ y, int n1, double lo0,
                                                    double lo1, double d0, double d1, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nout = D == 2 ? (int64_t)n0 * n1 : n0;
    if (b >= nout) return;
    const int stride = D + 3;
    double F = 0.0;
    if (D == 1) {
        const double X = x[b];
        for (int64_t k = 0; k < nh; ++k) {
            const double* r = hills + k * stride;
            const double t = X - r[1];
            F += r[2] * exp(-(t * t) / (2.0 * (r[3] * r[3])));
        }
    } else {
        const int i = (int)(b / n1), j = (int)(b % n1);
        const double X = x[i], Y = y[j];
        for (int64_t k = 0; k < nh; ++k) {
            const double* r = hills + k * stride;
            const double c1 = r[1], c2 = r[2], h = r[3], sg = r[4];
            const double cut = 4.0 * sg;
            const int rmin = max(0, fes_trunc((c1 - cut - lo0) / d0));
            const int rmax = min(n0 - 1, fes_trunc((c1 + cut - lo0) / d0));
            const int cmin = max(0, fes_trunc((c2 - cut - lo1) / d1));
            const int cmax = min(n1 - 1, fes_trunc((c2 + cut - lo1) / d1));
            if (i < rmin || i > rmax || j < cmin || j > cmax) continue;
            const double a = X - c1, e = Y - c2;
            F += h * exp(-(a * a + e * e) / (2.0 * (sg * sg)));
        }
    }
    out[b] = F;
}

// Marginal of a 2-D FES (tools/plot_hills.py:236-262): -kT ln(sum exp(-F/kT) * bin)
// along one axis (axis 1 sums over the second CV), one thread per output entry.
__global__ __launch_bounds__(kBlock) void k_fes_marginal(const double* __restrict__ F, int n0, int n1, int axis,
                                                         double kT, double bin, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int nout = axis == 1 ? n0 : n1, nsum = axis == 1 ? n1 : n0;
    if (m >= nout) return;
    double acc = 0.0;
    for (int k = 0; k < nsum; ++k) {
        const double f = axis == 1 ? F[(int64_t)m * n1 + k] : F[(int64_t)k * n1 + m];
        acc += exp(-f / kT);
    }
    out[m] = -kT * log(acc * bin);
}

// v -= min(v) over n values, one workgroup (the minimum is exact, so the result
// does not depend on the reduction order)
__global__ __launch_bounds__(kBlock) void k_sub_min(double* v, int n) {
    __shared__ double sm[kBlock];
    double m = __builtin_inf();
    for (int k = threadIdx.x; k < n; k += blockDim.x) m = fmin(m, v[k]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    const double mn = sm[0];
    for (int k = threadIdx.x; k < n; k += blockDim.x) v[k] -= mn;
}

// ---------------------------------------------------------------------------
// STRU_FACTOR (Debye form, CV function AVE; src/cvs/CVs_Strufactor.cu):
// S_i = 1 + sum_j sin(q r)/(q r) s(r^2) over the list entries with
// 1e-12 <= r^2 <= rc^2 (get_environment_Strufactor :1014-1080), summed in list
// order before the 1 is added (structure_factor_cv_kernel :1082-1120), with the
// switch s = 1 - ((r^2 - r_on)/(rc^2 - r_on))^3 above r_on = 0.8 rc^2;
// CV = sum over group atoms of S_i / N_group (compute_cv_AVE :642-656).  FP64.
// The gradient is the exact one: a pair (k, j) enters S_k when k is in the
// group and S_j when j is, so atom k's force carries the weight
// [k in G] + [j in G] (the reference's dcv kernel, :1136-1200, weights a pair
// 1 + [j in G] and leaves atoms outside the group without the force).
// ---------------------------------------------------------------------------
// The reference's switching functions (include/utils/zqc_switch_function.h:216-337),
// f and df/dx, including its RATIONAL branches (1/(1+x^6) for n = 6, m = 12; 1/(1+x^4)
// for 4, 8; 1/(1+x^n) for m = 2n; 1 + x^m for n = 2m; else (1-x^n)/(1-x^m)).  One
// departure: the derivative of the general branch at x = 1 is its limit
// n (n - m) / (2 m r_0), where the reference returns the negative of it (:325-327).
__device__ double sw_f(const SwFn& p, double v) {
    switch (p.type) {
        case 1: return v < p.r0 ? 1.0 : 0.0;
        case 2: return 1.0 / (1.0 + exp(-p.alpha * (v - p.r0)));
        case 3: return 0.5 * (1.0 + tanh(p.alpha * (v - p.r0)));
        case 4: {
            const double x = (v - p.d0) / p.r0;
            if (p.n == 6 && p.m == 12) { const double x2 = x * x, x6 = x2 * x2 * x2; return 1.0 / (1.0 + x6); }
            if (p.n == 4 && p.m == 8) { const double x2 = x * x, x4 = x2 * x2; return 1.0 / (1.0 + x4); }
            if (2 * p.n == p.m) return 1.0 / (1.0 + pow(x, p.n));
            if (p.n == 2 * p.m) return 1.0 + pow(x, p.m);
            return (1 - pow(x, p.n)) / (1 - pow(x, p.m));
        }
        default: return 1.0;
    }
}
__device__ double sw_df(const SwFn& p, double v) {
    switch (p.type) {
        case 2: {
            const double e = exp(-p.alpha * (v - p.r0)), dn = 1.0 + e;
            return (p.alpha * e) / (dn * dn);
        }
        case 3: {
            const double t = tanh(p.alpha * (v - p.r0));
            return 0.5 * p.alpha * (1.0 - t * t);
        }
        case 4: {
            const double x = (v - p.d0) / p.r0;
            if (p.n == 6 && p.m == 12) {
                const double x2 = x * x, x5 = x2 * x2 * x, x6 = x2 * x2 * x2;
                return -p.n * x5 / ((1.0 + x6) * (1.0 + x6)) / p.r0;
            }
            if (p.n == 4 && p.m == 8) {
                const double x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
                return -(double)p.n * x3 / ((1.0 + x4) * (1.0 + x4)) / p.r0;
            }
            if (2 * p.n == p.m) {
                const double a = pow(x, p.n - 1), b = pow(x, p.n);
                return -(double)p.n * a / ((1.0 + b) * (1.0 + b)) / p.r0;
            }
            if (p.n == 2 * p.m) return (double)p.m * pow(x, p.m - 1) / p.r0;
            const double xm1 = pow(x, p.m - 1), xm = xm1 * x;
            // at x = 1 the limit n (n - m) / (2 m r_0); the reference's limit has the opposite sign
            if (fabs(1.0 - xm) < 1e-12) return 0.5 * (double)(p.n * (p.n - p.m)) / (double)p.m / p.r0;
            const double xn1 = pow(x, p.n - 1);
            return ((double)(p.n - p.m) * xn1 * xm + p.m * xm1 - p.n * xn1) / ((1 - xm) * (1 - xm)) / p.r0;
        }
        default: return 0.0;
    }
}

__device__ __forceinline__ double sf_switch(double r2, double rc2, double r, double* ds) {
    const double r_on = 0.8 * rc2;
    if (r2 > r_on) {
        const double u = (r2 - r_on) / (rc2 - r_on);
        *ds = -3 * (u * u) * (2.0 * r) / (rc2 - r_on);
        return 1.0 - u * u * u;
    }
    *ds = 0.0;
    return 1.0;
}

// Per-atom pair sums of STRU_FACTOR (g = sin(q r)/(q r) s, CVs_Strufactor.cu:1082-1120, and
// with the chemical weights cv_kernel_structure_factor_chem, CVs_Strufactor_chem.cu:151-191)
// and WEIGHT_CHEM (g = s, Weighted_chem_pair_cv_kernel,
// CV_Weighted_Chemical_Pair_Order.cu:647-694): S_i = 1 + w_i sum_j w_j g(r_ij) over list
// entries with 1e-12 <= r^2 <= rc^2; g and dg/dr.
__device__ __forceinline__ double sf_pair(const SfDev& d, double r2, double* dg) {
    const double r = sqrt(r2);
    double ds;
    const double sw = sf_switch(r2, d.rc2, r, &ds);
    if (d.pair) { *dg = ds; return sw; }
    const double th = d.q * r;
    const double sn = sin(th), cs = cos(th);
    *dg = (cs / th - sn / (th * th)) * sw * d.q + ds * sn / th;   // d/dr of sin(q r)/(q r) s
    return sn / th * sw;
}

__global__ __launch_bounds__(64) void k_sf_value(Geo g, SfDev d) {
#pragma clang fp contract(off)
    const int t = blockIdx.x;
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const int a = threadIdx.x;
    const bool valid = a < tl.count;
    const int i = tl.start + (valid ? a : 0);
    double S = 0.0;
    if (valid) {
        const double4 p = g.x4[i];
        double acc = 0.0;
        for (int e = g.off[i]; e < g.off[i + 1]; ++e) {
            const int j = g.idx[e];
            const double4 q = g.x4[j];
            const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if ((r2 > d.rc2) || (r2 < 1e-12)) continue;
            double dg;
            const double gv = sf_pair(d, r2, &dg);
            acc += d.aw ? d.aw[j] * gv : gv;
        }
        S = (d.aw ? d.aw[i] * acc : acc) + 1.0;
        d.S[i] = S;
    }
    const bool in_g = valid && (g.group == nullptr || g.group[i]);
    const double term = d.count ? sw_f(d.sw, S) : S;             // COUNT: compute_cv_COUNT :658-674
    const double part = wave_sum(in_g ? term : 0.0);
    if (threadIdx.x == 0) d.part[t] = part;
}

// per walker: the tile sums in tile order -> CV value (one 64-lane wave)
__global__ __launch_bounds__(64) void k_sf_sum(SfDev d) {
    const int w = blockIdx.x;
    const int t0 = d.tile_begin[w], t1 = d.tile_begin[w + 1];
    double acc = 0.0;
    for (int t = t0 + threadIdx.x; t < t1; t += 64) acc += d.part[t];
    const double v = wave_sum(acc);
    if (threadIdx.x == 0) d.direct[(int64_t)w * kMaxSlotCV + d.slot] = d.count ? v : v * d.inv_ng[w];
}

__global__ __launch_bounds__(64) void k_sf_force(Geo g, SfDev d, double* __restrict__ f) {
#pragma clang fp contract(off)
    const int t = blockIdx.x;
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const int a = threadIdx.x;
    if (a >= tl.count) return;
    const int k = tl.start + a;
    // AVE: pair weight ([k in G] + [j in G]) / N_group; COUNT: [k in G] sw'(S_k) + [j in G] sw'(S_j)
    // (structure_factor_dcv_COUNT_kernel :1203-1271 for k in G)
    const double wc = d.wcv[(int64_t)tl.walker * kWcvStride + d.slot];
    const double W = d.count ? wc : wc * d.inv_ng[tl.walker];
    const double gk = (g.group == nullptr || g.group[k]) ? (d.count ? sw_df(d.sw, d.S[k]) : 1.0) : 0.0;
    const double wk = d.aw ? d.aw[k] : 1.0;
    const double4 p = g.x4[k];
    double F[3] = {0.0, 0.0, 0.0};
    for (int e = g.off[k]; e < g.off[k + 1]; ++e) {
        const int j = g.idx[e];
        const double4 q = g.x4[j];
        const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        if ((r2 > d.rc2) || (r2 < 1e-12)) continue;
        const double r = sqrt(r2);
        double fp;
        (void)sf_pair(d, r2, &fp);
        if (d.aw) fp *= wk * d.aw[j];
        const int jo = owner_of(g, j);
        const double gj = (g.group == nullptr || g.group[jo]) ? (d.count ? sw_df(d.sw, d.S[jo]) : 1.0) : 0.0;
        const double c = (gk + gj) * fp / r;
        F[0] += c * dx; F[1] += c * dy; F[2] += c * dz;
    }
    #pragma unroll
    for (int q = 0; q < 3; ++q) f[3 * (int64_t)k + q] += W * F[q];
}

// ---------------------------------------------------------------------------
// STEINH L: switching-weighted local Q (src/cvs/CVs_Steinhardt_Local.cu).
// Every list entry of atom a with sigma(r) >= eps counts (get_environment_
// Steinhardt_LocalQ :1210-1285): S_a = sum sigma, q_lm(a) = sum sigma Y_lm / S_a
// (:1310-1452).  A group atom i has LQ_lm(i) = (q_lm(i) + sum sigma q_lm(j)) /
// (S_i + 1) and LQ_l(i) (:1490-1579); the CV is sum q f(q) / sum f(q)
// (MEAN_SOLID), sum f(q) (NSOLID) or sum f(q) / N_group (FRAC) (:694-745).
// FP64 throughout, Y_lm from the same orthonormal polynomial table as the FP32
// path.  The gradient is exact and all atoms carry it: per entry e = (a -> j)
// the kernels form g_e = d(W CV)/d(x_j - x_a); the force on a is the sum of g
// over its own entries minus g of the reverse entries (the images of a in its
// neighbours' lists), so no atomics are needed.  Per-entry words live in a per-tile
// ELL layout: entry k of atom a sits at apos[a] + 64 k, apos[a] = tile base + lane,
// so the 64 lanes of a tile touch 64 consecutive words on every round.
// ---------------------------------------------------------------------------
template <int L, int M, bool DER, int K>
struct YC64 {
    static constexpr double v = DER ? (K + 1 <= L ? (K + 1) * kY.c[L][M][K + 1] : 0.0) : kY.c[L][M][K];
};
template <int L, int M, bool DER, int K, int TOP>
struct H64 {
    static __device__ __forceinline__ double run(double v) {
        if constexpr (K >= TOP) return YC64<L, M, DER, TOP>::v;
        else return fma(v, H64<L, M, DER, K + 2, TOP>::run(v), YC64<L, M, DER, K>::v);
    }
};
// c_lm(u) (DER = false) or c'_lm(u) (DER = true) in FP64
template <int L, int M, bool DER>
__device__ __forceinline__ double ypoly64(double u, double u2) {
    constexpr int DEG = DER ? (L - M - 1) : (L - M);
    if constexpr (DEG < 0) {
        return 0.0;
    } else {
        const double h = H64<L, M, DER, DEG % 2, DEG>::run(u2);
        if constexpr (DEG % 2) return h * u; else return h;
    }
}

// acc[2m], acc[2m+1] += s Re/Im Y_lm(n), m = 0..L
template <int L>
__device__ __forceinline__ void ylm64_acc(double nx, double ny, double u, double s, double* acc) {
    const double u2 = u * u;
    double zr = 1.0, zi = 0.0;
    sfor<L + 1>([&](auto mm) {
        constexpr int M = decltype(mm)::value;
        if constexpr (M > 0) {
            const double t = zr * nx - zi * ny;
            zi = zr * ny + zi * nx;
            zr = t;
        }
        const double c = s * ypoly64<L, M, false>(u, u2);
        acc[2 * M] += c * zr;
        acc[2 * M + 1] += c * zi;
    });
}

// <H, Y(d)> and its gradient in d: grad_n of sum_m c_lm(u) (H_re Re z^m + H_im Im z^m),
// z = n_x + i n_y, projected: grad_d = (grad_n - (n . grad_n) n) / r
template <int L>
__device__ __forceinline__ double ylm64_dot_grad(double nx, double ny, double u, double r, const double* H,
                                                 double& gx, double& gy, double& gz) {
    const double u2 = u * u;
    double zr = 1.0, zi = 0.0, pr = 0.0, pi = 0.0;    // z^m and z^(m-1)
    double val = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
    sfor<L + 1>([&](auto mm) {
        constexpr int M = decltype(mm)::value;
        if constexpr (M > 0) {
            pr = zr; pi = zi;
            const double t = zr * nx - zi * ny;
            zi = zr * ny + zi * nx;
            zr = t;
        }
        const double hr = H[2 * M], hi = H[2 * M + 1];
        const double c = ypoly64<L, M, false>(u, u2), cd = ypoly64<L, M, true>(u, u2);
        const double proj = hr * zr + hi * zi;
        val += c * proj;
        az += cd * proj;
        if constexpr (M > 0) {
            ax += c * M * (hr * pr + hi * pi);
            ay += c * M * (hi * pr - hr * pi);
        }
    });
    const double dn = nx * ax + ny * ay + u * az, ir = 1.0 / r;
    gx = (ax - dn * nx) * ir;
    gy = (ay - dn * ny) * ir;
    gz = (az - dn * u) * ir;
    return val;
}

// acc += sum over entries [beg, end) of sigma_e rec[owner_e]; the per-entry words come
// first and two entries' records are loaded before either is used
template <int LS>
__device__ __forceinline__ void gather_weighted(const int32_t* __restrict__ eo, const double2* __restrict__ es,
                                                const double* __restrict__ rec, int beg, int end, int64_t ap,
                                                double* acc) {
    int e = beg;
    for (; e + 1 < end; e += 2) {
        const double s0 = es[ap + 64 * (int64_t)(e - beg)].x, s1 = es[ap + 64 * (int64_t)(e + 1 - beg)].x;
        const double* r0 = rec + (int64_t)eo[ap + 64 * (int64_t)(e - beg)] * LS;
        const double* r1 = rec + (int64_t)eo[ap + 64 * (int64_t)(e + 1 - beg)] * LS;
        double v0[LS], v1[LS];
        #pragma unroll
        for (int k = 0; k < LS; ++k) { v0[k] = r0[k]; v1[k] = r1[k]; }
        #pragma unroll
        for (int k = 0; k < LS; ++k) acc[k] += s0 * v0[k] + s1 * v1[k];
    }
    if (e < end) {
        const double s0 = es[ap + 64 * (int64_t)(e - beg)].x;
        const double* r0 = rec + (int64_t)eo[ap + 64 * (int64_t)(e - beg)] * LS;
        #pragma unroll
        for (int k = 0; k < LS; ++k) acc[k] += s0 * r0[k];
    }
}

// Store each lane's LS-double row of a tile's consecutive atoms: the rows pass through
// LDS so the wave writes the tile's contiguous block in order.  Every lane of the
// one-wave workgroup calls it (count <= 64 rows are written).
template <int LS>
__device__ __forceinline__ void tile_store_rows(double* __restrict__ dst, const double* v, int count, double* sh) {
    const int lane = threadIdx.x;
    #pragma unroll
    for (int k = 0; k < LS; ++k) sh[lane * LS + k] = v[k];
    __syncthreads();
    for (int t = lane; t < count * LS; t += 64) dst[t] = sh[t];
    __syncthreads();
}

template <int L>
__global__ __launch_bounds__(64) void k_lq_q(Geo g, LqDev d) {
    constexpr int LS = 2 * (L + 1);
    __shared__ double sh[64 * LS];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const bool valid = (int)threadIdx.x < tl.count;
    const int i = tl.start + (valid ? (int)threadIdx.x : 0);
    const double4 p = g.x4[i];
    double acc[LS];
    #pragma unroll
    for (int k = 0; k < LS; ++k) acc[k] = 0.0;
    double S = 0.0;
    int cnt = 0;
    const int beg = g.off[i];
    const int end = valid ? g.off[i + 1] : beg;
    const int64_t ap = d.apos[i];
    // software pipeline: the next entry's position and the list index after it are
    // loaded while the current entry is evaluated
    int jn = beg + 1 < end ? g.idx[beg + 1] : i;
    double4 qn = g.x4[beg < end ? g.idx[beg] : i];
    for (int e = beg; e < end; ++e) {
        const double4 q = qn;
        qn = g.x4[jn];
        jn = e + 2 < end ? g.idx[e + 2] : i;
        const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        double s = 0.0, ds = 0.0;
        if (r2 >= 1e-12) {
            const double r = sqrt(r2);
            s = sw_f(d.swr, r);
            if (s >= d.eps) {
                ds = sw_df(d.swr, r);
                S += s;
                ++cnt;
                const double ir = 1.0 / r;
                ylm64_acc<L>(dx * ir, dy * ir, dz * ir, s, acc);
            } else {
                s = 0.0;
            }
        }
        d.es[ap + 64 * (int64_t)(e - beg)] = make_double2(s, ds);
    }
    const bool live = cnt > 0 && S >= 1e-12;
    const double inv = live ? 1.0 / S : 0.0;
    #pragma unroll
    for (int k = 0; k < LS; ++k) acc[k] *= inv;
    tile_store_rows<LS>(d.q + (int64_t)tl.start * LS, acc, tl.count, sh);
    if (valid) d.S[i] = live ? S : 0.0;
}

// LQ_lm(i) = (q_lm(i) + sum_e sigma_e q_lm(owner(e))) / (S_i + 1), kLanes lanes per atom:
// lane l gathers entries l, l + 4, ... of the atom's list, the group adds the shares
// (four short gather chains per atom instead of one long one; splitting each row's
// components over the lanes instead was slower).
template <int L>
__global__ __launch_bounds__(kBlock) void k_lq_lq(Geo g, LqDev d) {
    constexpr int LS = 2 * (L + 1);
    __shared__ double sh[2 * (kBlock / 64)];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const int a = threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
    const bool valid = a < tl.count;
    const int i = tl.start + (valid ? a : 0);
    const bool in_g = valid && (g.group == nullptr || g.group[i]);
    const double Si = d.S[i];
    const bool live = in_g && Si > 0.0;          // uniform within the group
    const double* __restrict__ qa = d.q;
    double acc[LS];
    #pragma unroll
    for (int k = 0; k < LS; ++k) acc[k] = 0.0;
    if (live) {
        const int beg = g.off[i], end = g.off[i + 1];
        const int64_t ap = d.apos[i];
        int e = beg + lane;
        for (; e + kLanes < end; e += 2 * kLanes) {
            const int64_t e0 = ap + 64 * (int64_t)(e - beg), e1 = e0 + 64 * kLanes;
            const double s0 = d.es[e0].x, s1 = d.es[e1].x;
            const double* r0 = qa + (int64_t)d.eown[e0] * LS;
            const double* r1 = qa + (int64_t)d.eown[e1] * LS;
            double v0[LS], v1[LS];
            #pragma unroll
            for (int k = 0; k < LS; ++k) { v0[k] = r0[k]; v1[k] = r1[k]; }
            #pragma unroll
            for (int k = 0; k < LS; ++k) acc[k] += s0 * v0[k] + s1 * v1[k];
        }
        if (e < end) {
            const int64_t e0 = ap + 64 * (int64_t)(e - beg);
            const double s0 = d.es[e0].x;
            const double* r0 = qa + (int64_t)d.eown[e0] * LS;
            #pragma unroll
            for (int k = 0; k < LS; ++k) acc[k] += s0 * r0[k];
        }
    }
    #pragma unroll
    for (int k = 0; k < LS; ++k) acc[k] = grp_sum_f64(acc[k]);
    double v = 0.0;
    if (live) {
        const double inv = 1.0 / (Si + 1.0);
        double sq = 0.0;
        #pragma unroll
        for (int k = 0; k < LS; ++k) {
            const double lq = (acc[k] + qa[(int64_t)i * LS + k]) * inv;
            if (k % kLanes == lane) d.LQ[(int64_t)i * LS + k] = lq;
            sq += (k < 2 ? 1.0 : 2.0) * lq * lq;
        }
        v = sqrt(sq * (4.0 * kPi) / (double)(2 * L + 1));
    } else if (valid) {
        #pragma unroll
        for (int k = 0; k < LS; ++k)
            if (k % kLanes == lane) d.LQ[(int64_t)i * LS + k] = 0.0;
    }
    if (valid && lane == 0) d.ql[i] = v;
    const bool one = in_g && lane == 0;            // each atom counted once
    const double f = one ? sw_f(d.swcv, v) : 0.0;
    double part[2] = {one ? v * f : 0.0, f};
    block_sum<2>(part, sh);
    if (threadIdx.x == 0) { d.part[2 * t] = part[0]; d.part[2 * t + 1] = part[1]; }
}

// per walker: sum q f and sum f over the tiles in a fixed order -> CV (one workgroup)
__global__ __launch_bounds__(kBlock) void k_lq_sum(LqDev d) {
    __shared__ double sh[2 * (kBlock / 64)];
    const int w = blockIdx.x;
    const int t0 = d.tile_begin[w], t1 = d.tile_begin[w + 1];
    double v[2] = {0.0, 0.0};
    for (int t = t0 + threadIdx.x; t < t1; t += kBlock) { v[0] += d.part[2 * t]; v[1] += d.part[2 * t + 1]; }
    block_sum<2>(v, sh);
    const double sq = v[0], sf = v[1];
    if (threadIdx.x != 0) return;
    const double N = d.ng[w];
    const double cv = d.mode == 0 ? (sf != 0.0 ? sq / sf : 0.0) : d.mode == 1 ? sf : (N > 0.0 ? sf / N : 0.0);
    d.direct[(int64_t)w * kMaxSlotCV + d.slot] = cv;
    d.sums[2 * w] = sf;
    d.sums[2 * w + 1] = cv;
}

// K_i = W dCV/dLQ_lm(i) / (S_i + 1): w_i = dCV/dLQ_l(i) per mode, then
// dLQ_l/dLQ_lm = 4 pi/(2l+1) c_m LQ_lm / LQ_l (c_0 = 1, c_m = 2 for m >= 1)
template <int L>
__global__ __launch_bounds__(64) void k_lq_k(Geo g, LqDev d) {
    constexpr int LS = 2 * (L + 1);
    __shared__ double sh[64 * LS];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const bool valid = (int)threadIdx.x < tl.count;
    const int i = tl.start + (valid ? (int)threadIdx.x : 0);
    const int w = tl.walker;
    const double ql = d.ql[i];
    double c = 0.0;
    if (valid && (g.group == nullptr || g.group[i]) && d.S[i] > 0.0 && ql > 0.0) {
        const double W = d.wcv[(int64_t)w * kWcvStride + d.slot];
        const double sf = d.sums[2 * w], cv = d.sums[2 * w + 1], N = d.ng[w];
        const double f = sw_f(d.swcv, ql), fp = sw_df(d.swcv, ql);
        const double wi = d.mode == 0 ? (sf != 0.0 ? (f + fp * (ql - cv)) / sf : 0.0)
                        : d.mode == 1 ? fp : (N > 0.0 ? fp / N : 0.0);
        c = W * wi * (4.0 * kPi / (double)(2 * L + 1)) / ql / (d.S[i] + 1.0);
    }
    double kv[LS];
    #pragma unroll
    for (int k = 0; k < LS; ++k) kv[k] = c != 0.0 ? (k < 2 ? c : 2.0 * c) * d.LQ[(int64_t)i * LS + k] : 0.0;
    tile_store_rows<LS>(d.K + (int64_t)tl.start * LS, kv, tl.count, sh);
}

// H_a = K_a + sum over counted entries of s K_owner(j), with <H_a, q_a> and <K_a, LQ_a>; for
// each counted entry e = (a -> j) also u_e = <K_owner(j), q_a>, which is the <K_i, q_j> the
// reverse entry of e needs (i = owner(j), j = a): the pair pass reads 8 bytes there instead
// of gathering a neighbour's q row
template <int L>
__global__ __launch_bounds__(kBlock) void k_lq_h(Geo g, LqDev d) {
    // kLanes lanes per atom: lane l takes entries l, l + 4, ...; the group adds the shares of H
    constexpr int LS = 2 * (L + 1);
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const int a = threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
    const bool valid = a < tl.count;
    const int i = tl.start + (valid ? a : 0);
    const double* __restrict__ Ka = d.K;
    const double* __restrict__ qa = d.q;
    double H[LS], qi[LS];
    #pragma unroll
    for (int k = 0; k < LS; ++k) { H[k] = 0.0; qi[k] = qa[(int64_t)i * LS + k]; }
    if (valid) {
        // u on every entry, counted here or not: the reverse entry may count where this one
        // falls below cutoff_eps by rounding, and u does not depend on sigma
        const bool live = d.S[i] > 0.0;
        const int beg = g.off[i], end = g.off[i + 1];
        const int64_t ap = d.apos[i];
        for (int e = beg + lane; e < end; e += kLanes) {
            const int64_t el = ap + 64 * (int64_t)(e - beg);
            const double s = live ? d.es[el].x : 0.0;
            const double* ko = Ka + (int64_t)d.eown[el] * LS;
            double v[LS];
            #pragma unroll
            for (int k = 0; k < LS; ++k) v[k] = ko[k];
            double u = 0.0;
            #pragma unroll
            for (int k = 0; k < LS; ++k) { H[k] += s * v[k]; u += v[k] * qi[k]; }
            d.u[el] = u;
        }
    }
    double hq = 0.0, kl = 0.0;
    #pragma unroll
    for (int k = 0; k < LS; ++k) {
        const double ki = Ka[(int64_t)i * LS + k];
        H[k] = grp_sum_f64(H[k]) + ki;
        kl += ki * d.LQ[(int64_t)i * LS + k];
        hq += H[k] * qi[k];
    }
    if (valid) {
        if (lane == 0) d.hk[i] = make_double2(hq, kl);
        #pragma unroll
        for (int k = 0; k < LS; ++k)
            if (k % kLanes == lane) d.H[(int64_t)i * LS + k] = H[k];
    }
}

// per entry of a: g_e = (1/S_a) [s' <H_a, Y - q_a> n + s grad_d <H_a, Y>] + s' (<K_a, q_j> - <K_a, LQ_a>) n,
// <K_a, q_j> read at the reverse entry (u, formed by k_lq_h)
template <int L>
__global__ __launch_bounds__(kBlock) void k_lq_pair(Geo g, LqDev d) {
    // kLanes lanes per atom: lane l takes entries l, l + 4, ... (each entry's g is its own)
    constexpr int LS = 2 * (L + 1);
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const int a = threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
    if (a >= tl.count) return;
    const int i = tl.start + a;
    const int beg = g.off[i], end = g.off[i + 1];
    const int64_t ap = d.apos[i];
    const double Si = d.S[i];
    double* __restrict__ go = d.g;
    if (!(Si > 0.0)) {
        for (int e = beg + lane; e < end; e += kLanes) {
            const int64_t o = 3 * (ap + 64 * (int64_t)(e - beg));
            go[o] = 0.0; go[o + 1] = 0.0; go[o + 2] = 0.0;
        }
        return;
    }
    const double4 p = g.x4[i];
    double H[LS];
    #pragma unroll
    for (int k = 0; k < LS; ++k) H[k] = d.H[(int64_t)i * LS + k];
    const double2 hk = d.hk[i];
    const double iS = 1.0 / Si;
    // Two-stage software pipeline over the lane's entries: while entry e is evaluated, the
    // neighbour position and reverse-entry u of e + 4 are in flight, and the (sigma, sigma'),
    // list index and reverse index of e + 8.
    auto words = [&](int e, double2& sd, int& j, int64_t& er) {
        if (e < end) {
            const int64_t el = ap + 64 * (int64_t)(e - beg);
            sd = d.es[el]; j = d.eidx[el]; er = d.erev[el];
        } else {
            sd = make_double2(0.0, 0.0); j = i; er = -1;
        }
    };
    double2 sd1, sd2;
    int j1, j2;
    int64_t er1, er2;
    words(beg + lane, sd1, j1, er1);
    words(beg + lane + kLanes, sd2, j2, er2);
    double4 q1 = g.x4[j1];
    double u1 = er1 >= 0 ? d.u[er1] : 0.0;
    for (int e = beg + lane; e < end; e += kLanes) {
        const int64_t el = ap + 64 * (int64_t)(e - beg);
        const double2 sd = sd1;
        const double4 q = q1;
        const double ue = u1;
        sd1 = sd2;
        q1 = g.x4[j2];
        u1 = er2 >= 0 ? d.u[er2] : 0.0;
        words(e + 2 * kLanes, sd2, j2, er2);
        double G0 = 0.0, G1 = 0.0, G2 = 0.0;
        if (sd.x != 0.0) {
            const double kq = ue - hk.y;
            const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
            const double r = sqrt(dx * dx + dy * dy + dz * dz), ir = 1.0 / r;
            const double nx = dx * ir, ny = dy * ir, nz = dz * ir;
            double gx, gy, gz;
            const double hy = ylm64_dot_grad<L>(nx, ny, nz, r, H, gx, gy, gz) - hk.x;
            const double cr = sd.y * hy * iS + sd.y * kq, ct = sd.x * iS;
            G0 = cr * nx + ct * gx;
            G1 = cr * ny + ct * gy;
            G2 = cr * nz + ct * gz;
        }
        go[3 * el] = G0; go[3 * el + 1] = G1; go[3 * el + 2] = G2;
    }
}

// per entry: the ELL index of its reverse entry (-1 = none), from the list topology
__global__ void k_lq_erev(Geo g, int64_t n, const int32_t* __restrict__ reown, const int32_t* __restrict__ rrev,
                          const int64_t* __restrict__ apos, int64_t* __restrict__ erev, int32_t* __restrict__ eidx,
                          int32_t* __restrict__ eown) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int beg = g.off[a];
    for (int e = beg; e < g.off[a + 1]; ++e) {
        const int rp = rrev[e];
        const int64_t el = apos[a] + 64 * (int64_t)(e - beg);
        erev[el] = rp < 0 ? -1 : apos[reown[e]] + 64 * (int64_t)rp;
        eidx[el] = g.idx[e];
        eown[el] = reown[e];
    }
}

// F_a += sum over a's entries of g_e - sum of g at their reverse entries, kLanes lanes
// per atom (lane l takes entries l, l + 4, ...; the group adds the shares)
__global__ __launch_bounds__(kBlock) void k_lq_force(Geo g, LqDev d, double* __restrict__ f) {
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= g.ntiles) return;
    const Tile tl = g.tiles[t];
    const int a = threadIdx.x / kLanes, lane = threadIdx.x % kLanes;
    const bool valid = a < tl.count;
    const int i = tl.start + (valid ? a : 0);
    const double* __restrict__ gg = d.g;
    const int len = valid ? g.off[i + 1] - g.off[i] : 0;
    const int64_t ap = d.apos[i];
    double F0 = 0.0, F1 = 0.0, F2 = 0.0;
    for (int k = lane; k < len; k += kLanes) {
        const int64_t el = ap + 64 * (int64_t)k;
        const int64_t er = d.erev[el];
        F0 += gg[3 * el]; F1 += gg[3 * el + 1]; F2 += gg[3 * el + 2];
        if (er < 0) continue;
        F0 -= gg[3 * er]; F1 -= gg[3 * er + 1]; F2 -= gg[3 * er + 2];
    }
    F0 = grp_sum_f64(F0); F1 = grp_sum_f64(F1); F2 = grp_sum_f64(F2);
    if (valid && lane == 0) {
        f[3 * (int64_t)i] += F0; f[3 * (int64_t)i + 1] += F1; f[3 * (int64_t)i + 2] += F2;
    }
}

// ---------------------------------------------------------------------------
// DISTANCE CV (src/cvs/CVs_DISTANCE.cu:55-123): r = |x_a2 - x_a1| with each
// component folded to the minimum image by the half-box rule of delta_x
// (:55-85); the bias force is -dV/dr on a1 along -d/r and on a2 along +d/r
// (bias_force :98-112 with get_dcvdx :114-121).  One thread per walker.
// ---------------------------------------------------------------------------
__global__ void k_dist_value(const double4* __restrict__ x4, DistDev d) {
#pragma clang fp contract(off)
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= d.n_walkers) return;
    const double4 p1 = x4[d.atom_off[w] + d.a1], p2 = x4[d.atom_off[w] + d.a2];
    double v[3] = {p2.x - p1.x, p2.y - p1.y, p2.z - p1.z};
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double L = d.box[3 * w + k];
        if (v[k] > L / 2) v[k] -= L;
        else if (v[k] < -L / 2) v[k] += L;
    }
    const double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double* o = d.dvec + 4 * w;
    o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = r;
    d.direct[(int64_t)w * kMaxSlotCV + d.slot] = r;
}

__global__ void k_dist_force(DistDev d, double* __restrict__ f) {
#pragma clang fp contract(off)
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= d.n_walkers) return;
    const double* o = d.dvec + 4 * w;
    const double dV = d.wcv[(int64_t)w * kWcvStride + d.slot];
    double* f1 = f + 3 * (d.atom_off[w] + d.a1);
    double* f2 = f + 3 * (d.atom_off[w] + d.a2);
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double g = -o[k] / o[3];
        f1[k] -= dV * g;
        f2[k] += dV * g;
    }
}

// Virial partials of a force buffer: sum_i F_i (x) unwrap(x_i), components
// xx yy zz xy xz yz (the tally of FixMetadynamics::post_force).
__global__ __launch_bounds__(kBlock) void k_virial(const double4* __restrict__ x4, const double* __restrict__ f,
                                                   const int32_t* __restrict__ image, int first, int n,
                                                   double L0, double L1, double L2, double* __restrict__ out) {
    __shared__ double sh[6 * 4];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (a < n) {
        const int i = first + a;
        const double4 p = x4[i];
        double u0 = p.x, u1 = p.y, u2 = p.z;
        if (image) { u0 += image[3 * a] * L0; u1 += image[3 * a + 1] * L1; u2 += image[3 * a + 2] * L2; }
        const double fx = f[3 * (int64_t)i], fy = f[3 * (int64_t)i + 1], fz = f[3 * (int64_t)i + 2];
        v[0] = fx * u0; v[1] = fy * u1; v[2] = fz * u2; v[3] = fx * u1; v[4] = fx * u2; v[5] = fy * u2;
    }
    block_sum<6>(v, sh);
    if (threadIdx.x == 0) for (int k = 0; k < 6; ++k) out[6 * blockIdx.x + k] = v[k];
}

__global__ void k_fill(double* p, int64_t n, double v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f(float* p, int64_t n, float v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// Structure factor.  k is stored in revolutions per Angstrom (k / 2pi) so the
// FP64 phase reduces to its fractional part before the FP32 sin/cos.
// ---------------------------------------------------------------------------
constexpr int kSkTile = 8;    // k vectors per workgroup in the amplitude pass
constexpr int kSkSlots = 256; // slots of a phase tile (one reference position each)
constexpr int kSkMaxTiles = 64;   // phase tiles of an amplitude chunk (chunks of <= 16,384 slots)

// sin and cos of 2 pi ph (ph in revolutions): FP64 reduction to [-0.5, 0.5], then the
// hardware sine/cosine, which take their argument in revolutions
__device__ __forceinline__ void sincos_rev(double ph, float& s, float& c) {
    const float fr = (float)(ph - rint(ph));
    s = __builtin_amdgcn_sinf(fr);
    c = __builtin_amdgcn_cosf(fr);
}
__device__ __forceinline__ float sin_rev(double ph) { return __builtin_amdgcn_sinf((float)(ph - rint(ph))); }

// Tile-relative phases.  Each walker's atoms are ordered along a Hilbert curve (sk.perm),
// so a run of kSkSlots consecutive slots is a compact set of atoms (within ~65 A of its
// first atom on the benchmark systems).  The phase of atom j in tile t is
//   k.r_j / 2pi = theta_t(k) + k.(r_j - r_t) / 2pi,   theta_t(k) = frac(k.r_t / 2pi),
// theta from the first atom's FP64 position once per (tile, k), the rest in FP32 from
// the FP32 offset r_j - r_t: a few revolutions, so the FP32 phase carries ~1e-5 rad.
// The hardware sine and cosine take that argument as it is (on gfx950 they hold 1.2e-7
// over |x| <= 300 revolutions, measured); when the tiles' extent at the list binding
// times the largest |k| / 2pi exceeds 128 revolutions (sk.wide), frac() first brings the
// argument to [0, 1).
__device__ __forceinline__ float frac_rev(double ph) { return (float)(ph - floor(ph)); }

// A(k) partial sums with the phase in FP64 (at most kSkTiledK k vectors): workgroup (k
// tile, atom chunk); k vectors are uniform per workgroup and live in scalar registers.
__global__ __launch_bounds__(kBlock) void k_sk_amp(Geo g, SkDev sk) {
    __shared__ double sh[kSkTile * 2 * 4];
    const int kb = blockIdx.x * kSkTile;
    const int ch = blockIdx.y;
    const int start = sk.chunk[3 * ch], count = sk.chunk[3 * ch + 1];
    double kx[kSkTile], ky[kSkTile], kz[kSkTile];
    #pragma unroll
    for (int q = 0; q < kSkTile; ++q) {
        const int qq = kb + q < sk.nk ? kb + q : sk.nk - 1;
        kx[q] = sk.k[3 * qq]; ky[q] = sk.k[3 * qq + 1]; kz[q] = sk.k[3 * qq + 2];
    }
    // FP32 sums over runs of kSkRun atoms of a thread, FP64 across runs and threads:
    // a run's rounding error (<= kSkRun^2 * 2^-24 in units of |e^{ik.r}| = 1) stays
    // below 1e-6 of |A| even at a Bragg peak
    constexpr int kSkRun = 8;
    double acc[kSkTile * 2];
    #pragma unroll
    for (int q = 0; q < kSkTile * 2; ++q) acc[q] = 0.0;
    for (int a0 = threadIdx.x; a0 < count; a0 += kSkRun * blockDim.x) {
        float run[kSkTile * 2];
        #pragma unroll
        for (int q = 0; q < kSkTile * 2; ++q) run[q] = 0.0f;
        // the run's positions and group flags are loaded together, ahead of the arithmetic
        double4 xv[kSkRun];
        bool on[kSkRun];
        #pragma unroll
        for (int r = 0; r < kSkRun; ++r) {
            const int a = a0 + r * blockDim.x;
            on[r] = a < count;
            const int i = start + (on[r] ? a : 0);
            xv[r] = g.x4[i];
            if (g.group) on[r] = on[r] && g.group[i];
        }
        #pragma unroll
        for (int r = 0; r < kSkRun; ++r) {
            if (!on[r]) continue;
            #pragma unroll
            for (int q = 0; q < kSkTile; ++q) {
                const double ph = xv[r].x * kx[q] + xv[r].y * ky[q] + xv[r].z * kz[q];
                float sn, cs;
                sincos_rev(ph, sn, cs);
                run[2 * q] += cs;
                run[2 * q + 1] += sn;
            }
        }
        #pragma unroll
        for (int q = 0; q < kSkTile * 2; ++q) acc[q] += (double)run[q];
    }
    // the block tree of block_sum, finished in parallel: wave sums, then thread q adds
    // value q's four wave sums in wave order and writes it
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    #pragma unroll
    for (int q = 0; q < kSkTile * 2; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == 0) sh[q * (kBlock / 64) + wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < kSkTile * 2) {
        const int q = threadIdx.x;
        double a = 0.0;
        #pragma unroll
        for (int k = 0; k < kBlock / 64; ++k) a += sh[q * (kBlock / 64) + k];
        if (kb + q / 2 < sk.nk) sk.apart[((int64_t)ch * sk.nk + kb) * 2 + q] = a;
    }
}

// Per-step slot data: workgroup (b, w) takes phase tile b of walker w: the tile's
// reference (its first atom's FP64 position) and every slot's FP32 offset from it, in
// slot order, so that the passes over (atom, k) read contiguous 16-byte records.
__global__ __launch_bounds__(kBlock) void k_sk_prep(Geo g, SkDev sk) {
    static_assert(kBlock == kSkSlots, "a workgroup is one phase tile");
    const int w = blockIdx.y, b = blockIdx.x;
    const int n = sk.n_atoms[w];
    const int a0 = b * kSkSlots;
    if (a0 >= n) return;
    const int sb = sk.atom_begin[w];
    const double4 r = g.x4[sk.perm[sb + a0]];
    if (threadIdx.x == 0) sk.tref[sk.tbeg[w] + b] = r;
    const int a = a0 + (int)threadIdx.x;
    if (a < n) {
        const int i = sk.perm[sb + a];
        const double4 x = g.x4[i];
        const float on = (g.group == nullptr || g.group[i]) ? 1.0f : 0.0f;
        sk.dr[sb + a] = make_float4((float)(x.x - r.x), (float)(x.y - r.y), (float)(x.z - r.z), on);
    }
}

// A(k) partial sums with tile-relative phases (more than kSkTiledK k vectors):
// workgroup (k tile, atom chunk); a chunk starts on a phase tile of its walker.
template <bool FRACT>
__global__ __launch_bounds__(kBlock) void k_sk_amp_tiled(SkDev sk) {
    static_assert(kBlock == kSkSlots, "a thread's r-th atom of a run lies in the chunk's tile r");
    __shared__ double sh[kSkTile * 2 * 4];
    __shared__ float th[kSkMaxTiles][kSkTile];
    const int kb = blockIdx.x * kSkTile;
    const int ch = blockIdx.y;
    const int start = sk.chunk[3 * ch], count = sk.chunk[3 * ch + 1], w = sk.chunk[3 * ch + 2];
    const int ntl = (count + kSkSlots - 1) / kSkSlots;
    const double4* tref = sk.tref + sk.tbeg[w] + (start - sk.atom_begin[w]) / kSkSlots;
    double kx[kSkTile], ky[kSkTile], kz[kSkTile];
    #pragma unroll
    for (int q = 0; q < kSkTile; ++q) {
        const int qq = kb + q < sk.nk ? kb + q : sk.nk - 1;
        kx[q] = sk.k[3 * qq]; ky[q] = sk.k[3 * qq + 1]; kz[q] = sk.k[3 * qq + 2];
    }
    // theta per (tile, k) from the tiles' reference positions
    for (int e = threadIdx.x; e < ntl * kSkTile; e += blockDim.x) {
        const int r = e / kSkTile, q = e % kSkTile;
        const double4 p = tref[r];
        th[r][q] = frac_rev(p.x * kx[q] + p.y * ky[q] + p.z * kz[q]);
    }
    __syncthreads();
    float kxf[kSkTile], kyf[kSkTile], kzf[kSkTile];
    #pragma unroll
    for (int q = 0; q < kSkTile; ++q) { kxf[q] = (float)kx[q]; kyf[q] = (float)ky[q]; kzf[q] = (float)kz[q]; }
    // FP32 sums over runs of kSkRun atoms of a thread, FP64 across runs and threads:
    // a run's rounding error (<= kSkRun^2 * 2^-24 in units of |e^{ik.r}| = 1) stays
    // below 1e-6 of |A| even at a Bragg peak
    constexpr int kSkRun = 8;
    double acc[kSkTile * 2];
    #pragma unroll
    for (int q = 0; q < kSkTile * 2; ++q) acc[q] = 0.0;
    for (int a0 = threadIdx.x; a0 < count; a0 += kSkRun * blockDim.x) {
        float run[kSkTile * 2];
        #pragma unroll
        for (int q = 0; q < kSkTile * 2; ++q) run[q] = 0.0f;
        // the run's offsets and group flags are loaded together, ahead of the arithmetic
        float4 d[kSkRun];
        #pragma unroll
        for (int r = 0; r < kSkRun; ++r) {
            const int a = a0 + r * blockDim.x;
            d[r] = a < count ? sk.dr[start + a] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        // slot a0 + r*256 lies in tile a0/256 + r, the same for the whole workgroup
        const int t0 = a0 / kSkSlots;
        #pragma unroll
        for (int r = 0; r < kSkRun; ++r) {
            if (d[r].w == 0.0f) continue;
            const int t = t0 + r;
            #pragma unroll
            for (int q = 0; q < kSkTile; ++q) {
                const float ph = __builtin_fmaf(d[r].z, kzf[q], __builtin_fmaf(d[r].y, kyf[q],
                                                __builtin_fmaf(d[r].x, kxf[q], th[t][q])));
                const float fr = FRACT ? __builtin_amdgcn_fractf(ph) : ph;
                run[2 * q] += __builtin_amdgcn_cosf(fr);
                run[2 * q + 1] += __builtin_amdgcn_sinf(fr);
            }
        }
        #pragma unroll
        for (int q = 0; q < kSkTile * 2; ++q) acc[q] += (double)run[q];
    }
    // the block tree of block_sum, finished in parallel: wave sums, then thread q adds
    // value q's four wave sums in wave order and writes it
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    #pragma unroll
    for (int q = 0; q < kSkTile * 2; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == 0) sh[q * (kBlock / 64) + wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < kSkTile * 2) {
        const int q = threadIdx.x;
        double a = 0.0;
        #pragma unroll
        for (int k = 0; k < kBlock / 64; ++k) a += sh[q * (kBlock / 64) + k];
        if (kb + q / 2 < sk.nk) sk.apart[((int64_t)ch * sk.nk + kb) * 2 + q] = a;
    }
}

// The gradient record of k vector q with amplitude A = re + i im: the force pass takes
// Im(conj(A) e^{i k.r}) = |A| sin(k.r - arg A), one sine per atom and k.
__device__ __forceinline__ SkK sk_record(const SkDev& sk, int q, double re, double im) {
    SkK kp;
    for (int d = 0; d < 3; ++d) {
        kp.kr[d] = sk.k[3 * q + d];
        kp.krf[d] = (float)sk.k[3 * q + d];
        kp.kv[d] = (float)sk.kvec[3 * q + d];
    }
    kp.pad = 0.0f;
    kp.phi = atan2(im, re) * 0.15915494309189535;     // 1 / 2pi
    kp.amod = (float)sqrt(re * re + im * im);
    return kp;
}

// per walker: A(k) = sum over chunks, S = |A|^2/N, value = mean S, packed k data
// Amplitudes of walker w = blockIdx.y, k = blockIdx.x: a workgroup sums that k's
// chunk partials, thread t taking chunks t, t+256, ... in order, then the
// fixed-order block sum; thread 0 writes A(k), the packed k record of the gradient
// pass and the term |A(k)|^2 / N.
__global__ __launch_bounds__(kBlock) void k_sk_fold(SkDev sk) {
    __shared__ double sh[2 * (kBlock / 64)];
    const int w = blockIdx.y;
    const int q = blockIdx.x;
    const int c0 = sk.chunk_begin[w], c1 = sk.chunk_begin[w + 1];
    double v2[2] = {0.0, 0.0};
    for (int c = c0 + (int)threadIdx.x; c < c1; c += kBlock) {
        const double2 v = reinterpret_cast<const double2*>(sk.apart)[(int64_t)c * sk.nk + q];
        v2[0] += v.x; v2[1] += v.y;
    }
    block_sum<2>(v2, sh);
    if (threadIdx.x != 0) return;
    const double re = v2[0], im = v2[1];
    const double N = sk.n_norm ? sk.n_norm[w] : (double)sk.n_atoms[w];
    sk.amp[((int64_t)w * sk.nk + q) * 2] = re;
    sk.amp[((int64_t)w * sk.nk + q) * 2 + 1] = im;
    sk.kpack[(int64_t)w * sk.nk + q] = sk_record(sk, q, re, im);
    sk.sterm[(int64_t)w * sk.nk + q] = (re * re + im * im) / N;
}

// k_sk_fold and k_sk_cv in one workgroup per walker (blockIdx.x) when there are at most
// kSkFoldAll k vectors: every k's chunk partials in the order and block tree of
// k_sk_fold, then SK.AVE in the order of k_sk_cv, so the results are bitwise those
// of the two kernels.
constexpr int kSkFoldAll = 8;
__global__ __launch_bounds__(kBlock) void k_sk_fold_cv(SkDev sk) {
    constexpr int NW = kBlock / 64;
    __shared__ double sh[2 * kSkFoldAll * NW];
    __shared__ double st[kSkFoldAll];
    const int w = blockIdx.x;
    const int nk = sk.nk;
    const int c0 = sk.chunk_begin[w], c1 = sk.chunk_begin[w + 1];
    double v2[2 * kSkFoldAll];
    #pragma unroll
    for (int q = 0; q < 2 * kSkFoldAll; ++q) v2[q] = 0.0;
    // a thread's chunks c0 + t, c0 + t + 256, ... in that order, four chunks' loads in
    // flight together
    constexpr int U = 4;
    for (int c = c0 + (int)threadIdx.x; c < c1; c += U * kBlock) {
        double2 v[U][kSkFoldAll];
        #pragma unroll
        for (int u = 0; u < U; ++u) {
            const double2* pc = reinterpret_cast<const double2*>(sk.apart) + (int64_t)(c + u * kBlock) * nk;
            #pragma unroll
            for (int q = 0; q < kSkFoldAll; ++q) v[u][q] = (q < nk && c + u * kBlock < c1) ? pc[q] : make_double2(0.0, 0.0);
        }
        #pragma unroll
        for (int u = 0; u < U; ++u) {
            if (c + u * kBlock < c1) {
                #pragma unroll
                for (int q = 0; q < kSkFoldAll; ++q) { v2[2 * q] += v[u][q].x; v2[2 * q + 1] += v[u][q].y; }
            }
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    #pragma unroll
    for (int q = 0; q < 2 * kSkFoldAll; ++q) {
        if (q < 2 * nk) {
            const double v = wave_sum(v2[q]);
            if (lane == 0) sh[q * NW + wv] = v;
        }
    }
    __syncthreads();
    const double N = sk.n_norm ? sk.n_norm[w] : (double)sk.n_atoms[w];
    if ((int)threadIdx.x < nk) {
        const int q = threadIdx.x;
        double re = 0.0, im = 0.0;
        #pragma unroll
        for (int k = 0; k < NW; ++k) re += sh[2 * q * NW + k];
        #pragma unroll
        for (int k = 0; k < NW; ++k) im += sh[(2 * q + 1) * NW + k];
        sk.amp[((int64_t)w * nk + q) * 2] = re;
        sk.amp[((int64_t)w * nk + q) * 2 + 1] = im;
        sk.kpack[(int64_t)w * nk + q] = sk_record(sk, q, re, im);
        const double term = (re * re + im * im) / N;
        sk.sterm[(int64_t)w * nk + q] = term;
        st[q] = term;
    }
    __syncthreads();
    if (wv == 0) {
        // k_sk_cv's tree: lane q holds term q (the other waves of k_sk_cv add zeros)
        const double v = wave_sum(lane < nk ? st[lane] : 0.0);
        if (lane == 0) sk.direct[(int64_t)w * kMaxSlotCV + sk.slot] = v / nk;
    }
}

// The gradient record and |A|^2 / N of amplitudes already in sk.amp (summed over the
// contexts that each hold a share of the atoms): walker blockIdx.y, k = the thread.
__global__ __launch_bounds__(kBlock) void k_sk_pack(SkDev sk) {
    const int w = blockIdx.y;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= sk.nk) return;
    const double re = sk.amp[((int64_t)w * sk.nk + q) * 2], im = sk.amp[((int64_t)w * sk.nk + q) * 2 + 1];
    const double N = sk.n_norm ? sk.n_norm[w] : (double)sk.n_atoms[w];
    sk.kpack[(int64_t)w * sk.nk + q] = sk_record(sk, q, re, im);
    sk.sterm[(int64_t)w * sk.nk + q] = (re * re + im * im) / N;
}

// SK.AVE of walker w = blockIdx.x: the mean over k of |A(k)|^2 / N, a fixed-order
// FP64 sum (thread-strided, then the block tree)
__global__ __launch_bounds__(kBlock) void k_sk_cv(SkDev sk) {
    __shared__ double sh[4];
    const int w = blockIdx.x;
    double v[1] = {0.0};
    for (int q = threadIdx.x; q < sk.nk; q += blockDim.x) v[0] += sk.sterm[(int64_t)w * sk.nk + q];
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) sk.direct[(int64_t)w * kMaxSlotCV + sk.slot] = v[0] / sk.nk;
}

// F_j = w (2/(N K)) sum_k k |A_k| sin(k.r_j - arg A_k) (= k Im(conj(A_k) e^{i k.r_j})).
// The k loop index is uniform, so the packed k records load through the scalar path.
// A thread takes AT atoms
// (strided by the workgroup size): each k record feeds AT terms per lane, which hides
// the load of the next record behind the arithmetic.  With few k vectors (at most
// kSkFewK) a thread takes one atom, so that the grid has four times the workgroups to
// cover the memory latency, and the k records are staged in LDS in one round trip
// instead of one dependent load per record; an atom's terms are added in the same
// order either way.
constexpr int kSkAtoms = 4;
constexpr int kSkFewK = 64;
template <int AT>
__global__ __launch_bounds__(kBlock) void k_sk_force(Geo g, SkDev sk, double* f) {
    const int w = blockIdx.y;
    const int a0 = blockIdx.x * blockDim.x * AT + threadIdx.x;
    const int n = sk.n_atoms[w];
    if ((int)(blockIdx.x * blockDim.x * AT) >= n) return;
    const double wt = sk.wcv[(int64_t)w * kWcvStride + sk.wslot];
    const double nn = sk.n_norm ? sk.n_norm[w] : (double)n;
    const float scale = (float)(wt * 2.0 / (nn * sk.nk));
    int ii[AT];
    double px[AT], py[AT], pz[AT];
    float gx[AT], gy[AT], gz[AT];
    #pragma unroll
    for (int u = 0; u < AT; ++u) {
        const int a = a0 + u * (int)blockDim.x;
        ii[u] = sk.atom_begin[w] + (a < n ? a : 0);
        const double4 xv = g.x4[ii[u]];
        px[u] = xv.x; py[u] = xv.y; pz[u] = xv.z;
        gx[u] = gy[u] = gz[u] = 0.0f;
    }
    const SkK* kp = sk.kpack + (int64_t)w * sk.nk;
    if constexpr (AT == 1) {
        // few k vectors: all records to LDS with one round trip, read back as broadcasts
        __shared__ SkK ks[kSkFewK];
        if ((int)threadIdx.x < sk.nk) ks[threadIdx.x] = kp[threadIdx.x];
        __syncthreads();
        kp = ks;
    }
    #pragma unroll 2
    for (int q = 0; q < sk.nk; ++q) {
        const SkK k = kp[q];
        #pragma unroll
        for (int u = 0; u < AT; ++u) {
            const double ph = px[u] * k.kr[0] + py[u] * k.kr[1] + pz[u] * k.kr[2] - k.phi;
            const float imv = k.amod * sin_rev(ph);
            gx[u] = __builtin_fmaf(k.kv[0], imv, gx[u]);
            gy[u] = __builtin_fmaf(k.kv[1], imv, gy[u]);
            gz[u] = __builtin_fmaf(k.kv[2], imv, gz[u]);
        }
    }
    #pragma unroll
    for (int u = 0; u < AT; ++u) {
        const int a = a0 + u * (int)blockDim.x;
        const int i = ii[u];
        if (a < n && (g.group == nullptr || g.group[i])) {
            f[3 * (int64_t)i] += (double)(scale * gx[u]);
            f[3 * (int64_t)i + 1] += (double)(scale * gy[u]);
            f[3 * (int64_t)i + 2] += (double)(scale * gz[u]);
        }
    }
}

// The same force with tile-relative phases (more than kSkTiledK k vectors).
// A workgroup takes AT phase tiles of kSkSlots slots (thread t the t-th atom of each),
// so the phase of atom j is theta'_u(k) + k.(r_j - r_u) / 2pi with
// theta'_u(k) = frac(k.r_u / 2pi - arg A_k / 2pi) from the tile's first atom in FP64,
// staged in LDS for a block of k vectors at a time; the rest is FP32 (see k_sk_amp).
// The k loop index is uniform, so the packed k records load through the scalar path,
// and each record feeds AT terms per lane, added in k order.
constexpr int kSkForceKB = 1024;   // k vectors whose tile phases are staged at a time
template <bool FRACT>
__global__ __launch_bounds__(kBlock) void k_sk_force_tiled(SkDev sk, double* f) {
    static_assert(kBlock == kSkSlots, "thread t takes slot t of each of the workgroup's tiles");
    constexpr int AT = kSkAtoms, KB = kSkForceKB;
    __shared__ float th[AT][KB];
    const int w = blockIdx.y;
    const int n = sk.n_atoms[w];
    const int b0 = blockIdx.x * blockDim.x * AT;       // the workgroup's first slot in the walker
    if (b0 >= n) return;
    const int sb = sk.atom_begin[w];
    const double wt = sk.wcv[(int64_t)w * kWcvStride + sk.wslot];
    const double nn = sk.n_norm ? sk.n_norm[w] : (double)n;
    const float scale = (float)(wt * 2.0 / (nn * sk.nk));
    // the workgroup's tiles: b0 / kSkSlots + u of the walker (those past its end are empty)
    const double4* tref = sk.tref + sk.tbeg[w] + b0 / kSkSlots;
    const int ntl = (n - b0 + kSkSlots - 1) / kSkSlots < AT ? (n - b0 + kSkSlots - 1) / kSkSlots : AT;
    int ii[AT];
    bool ok[AT];
    float dx[AT], dy[AT], dz[AT];
    float gx[AT], gy[AT], gz[AT];
    #pragma unroll
    for (int u = 0; u < AT; ++u) {
        const int a = b0 + u * kSkSlots + (int)threadIdx.x;
        const float4 d = a < n ? sk.dr[sb + a] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        ok[u] = d.w != 0.0f;
        ii[u] = a < n ? sk.perm[sb + a] : 0;
        dx[u] = d.x; dy[u] = d.y; dz[u] = d.z;
        gx[u] = gy[u] = gz[u] = 0.0f;
    }
    const SkK* kp = sk.kpack + (int64_t)w * sk.nk;
    for (int q0 = 0; q0 < sk.nk; q0 += KB) {
        const int nq = sk.nk - q0 < KB ? sk.nk - q0 : KB;
        __syncthreads();                               // the last block's phases are read
        for (int e = threadIdx.x; e < ntl * nq; e += blockDim.x) {
            const int u = e / nq, qq = e % nq;
            const SkK& k = kp[q0 + qq];
            const double4 p = tref[u];
            th[u][qq] = frac_rev(p.x * k.kr[0] + p.y * k.kr[1] + p.z * k.kr[2] - k.phi);
        }
        __syncthreads();
        #pragma unroll 2
        for (int qq = 0; qq < nq; ++qq) {
            const SkK k = kp[q0 + qq];
            #pragma unroll
            for (int u = 0; u < AT; ++u) {
                const float ph = __builtin_fmaf(dz[u], k.krf[2], __builtin_fmaf(dy[u], k.krf[1],
                                                __builtin_fmaf(dx[u], k.krf[0], th[u][qq])));
                const float imv = k.amod * __builtin_amdgcn_sinf(FRACT ? __builtin_amdgcn_fractf(ph) : ph);
                gx[u] = __builtin_fmaf(k.kv[0], imv, gx[u]);
                gy[u] = __builtin_fmaf(k.kv[1], imv, gy[u]);
                gz[u] = __builtin_fmaf(k.kv[2], imv, gz[u]);
            }
        }
    }
    #pragma unroll
    for (int u = 0; u < AT; ++u) {
        const int i = ii[u];
        if (ok[u]) {
            f[3 * (int64_t)i] += (double)(scale * gx[u]);
            f[3 * (int64_t)i + 1] += (double)(scale * gy[u]);
            f[3 * (int64_t)i + 2] += (double)(scale * gz[u]);
        }
    }
}

template <int NMAX>
hipError_t env_dispatch_l(const Geo& g, const EnvDev& e, dim3 grid, hipStream_t st) {
    switch (e.lmask) {
        case kL6:             hipLaunchKernelGGL((k_env<NMAX, kL6>), grid, dim3(kBlock), 0, st, g, e); break;
        case kL4:             hipLaunchKernelGGL((k_env<NMAX, kL4>), grid, dim3(kBlock), 0, st, g, e); break;
        case kL4 | kL6:       hipLaunchKernelGGL((k_env<NMAX, kL4 | kL6>), grid, dim3(kBlock), 0, st, g, e); break;
        case kL3:             hipLaunchKernelGGL((k_env<NMAX, kL3>), grid, dim3(kBlock), 0, st, g, e); break;
        case kL3 | kL4:       hipLaunchKernelGGL((k_env<NMAX, kL3 | kL4>), grid, dim3(kBlock), 0, st, g, e); break;
        case kL3 | kL6:       hipLaunchKernelGGL((k_env<NMAX, kL3 | kL6>), grid, dim3(kBlock), 0, st, g, e); break;
        case kL3 | kL4 | kL6: hipLaunchKernelGGL((k_env<NMAX, kL3 | kL4 | kL6>), grid, dim3(kBlock), 0, st, g, e); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

hipError_t launch_topology_near(const Geo& g, const NearDev& nd, int32_t* eown, uint8_t* revpos, int force,
                                hipStream_t st) {
    if (nd.n_local == 0) return hipSuccess;
    hipLaunchKernelGGL((k_topology<uint8_t, 0xFF>), dim3((unsigned)((nd.n_local + 255) / 256)), dim3(256), 0, st, g,
                       (int64_t)nd.n_local, eown, revpos, (const int32_t*)nd.flag, (const int64_t*)nd.epoch, force);
    return hipGetLastError();
}

hipError_t launch_topology(const Geo& g, int64_t n, int32_t* eown, uint8_t* revpos, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_topology<uint8_t, 0xFF>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, n, eown, revpos,
                       (const int32_t*)nullptr, (const int64_t*)nullptr, 1);
    return hipGetLastError();
}

hipError_t launch_topology_raw(const Geo& g, int64_t n, int32_t* eown, int32_t* revpos, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_topology<int32_t, -1>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, n, eown, revpos,
                       (const int32_t*)nullptr, (const int64_t*)nullptr, 1);
    return hipGetLastError();
}

hipError_t launch_prune(const Geo& g, const NearDev& nd, int force, hipStream_t st) {
    if (nd.n_all == 0) return hipSuccess;
    hipLaunchKernelGGL(k_prune, dim3((unsigned)((nd.n_all + 255) / 256)), dim3(256), 0, st, g, nd, force);
    return hipGetLastError();
}

hipError_t launch_pack_x4(const double* x, double4* x4, float4* x4f, int64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    int64_t b = (n + 255) / 256; if (b > 8192) b = 8192;
    hipLaunchKernelGGL(k_pack_x4, dim3((unsigned)b), dim3(256), 0, st, x, x4, x4f, n);
    return hipGetLastError();
}

hipError_t launch_pack_walker(const double* x, int64_t n_local, int64_t n_all, double4* x4_loc, float4* x4f_loc,
                              double4* x4_gh, float4* x4f_gh, hipStream_t st) {
    if (n_all == 0) return hipSuccess;
    int64_t b = (n_all + 255) / 256; if (b > 8192) b = 8192;
    hipLaunchKernelGGL(k_pack_walker, dim3((unsigned)b), dim3(256), 0, st, x, n_local, n_all, x4_loc, x4f_loc, x4_gh, x4f_gh);
    return hipGetLastError();
}

hipError_t launch_pack_walker_near(const double* x, int64_t n_local, int64_t n_all, double4* x4_loc, float4* x4f_loc,
                                   double4* x4_gh, float4* x4f_gh, const float4* xpr_loc, const float4* xpr_gh,
                                   const NearDev& nd, hipStream_t st) {
    int64_t b = (n_all + 255) / 256; if (b > 8192) b = 8192; if (b < 1) b = 1;
    hipLaunchKernelGGL(k_pack_walker_near, dim3((unsigned)b), dim3(256), 0, st, x, n_local, n_all, x4_loc, x4f_loc,
                       x4_gh, x4f_gh, xpr_loc, xpr_gh, nd);
    return hipGetLastError();
}

__global__ void k_fill_u8(uint8_t* p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}
hipError_t launch_fill_u8(uint8_t* p, int64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    int64_t b = (n + 255) / 256; if (b > 8192) b = 8192;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)b), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_env(const Geo& g, const EnvDev& e, hipStream_t st) {
    if (g.ntiles == 0) return hipSuccess;
    dim3 grid(xcd_grid(g.ntiles));
    if (e.N <= 12) return env_dispatch_l<12>(g, e, grid, st);
    if (e.N <= 16) return env_dispatch_l<16>(g, e, grid, st);
    if (e.N <= 18) return env_dispatch_l<18>(g, e, grid, st);
    if (e.N <= 24) return env_dispatch_l<24>(g, e, grid, st);
    if (e.N <= 32) return env_dispatch_l<32>(g, e, grid, st);
    if (e.N <= 64) return env_dispatch_l<64>(g, e, grid, st);
    return hipErrorInvalidValue;
}

hipError_t launch_select_export(const Geo& g, const EnvDev& e, hipStream_t st) {
    if (g.ntiles == 0) return hipSuccess;
    if (!e.sel_out || !e.sel_cnt) return hipErrorInvalidValue;
    dim3 grid(xcd_grid(g.ntiles));
#define MDT_SEL(NM) hipLaunchKernelGGL((k_env<NM, kL6, true>), grid, dim3(kBlock), 0, st, g, e)
    if (e.N <= 12) MDT_SEL(12);
    else if (e.N <= 16) MDT_SEL(16);
    else if (e.N <= 18) MDT_SEL(18);
    else if (e.N <= 24) MDT_SEL(24);
    else if (e.N <= 32) MDT_SEL(32);
    else if (e.N <= 64) MDT_SEL(64);
    else return hipErrorInvalidValue;
#undef MDT_SEL
    return hipGetLastError();
}

template <int L>
static hipError_t ld_dispatch(const Geo& g, const EnvDev& e, const LdDev& d, dim3 grid, hipStream_t st, bool coeff) {
#define MDT_LD(LM) do { if (coeff) hipLaunchKernelGGL((k_ld_coeff<L, LM>), grid, dim3(kBlock), 0, st, g, e, d); \
                        else hipLaunchKernelGGL((k_ld<L, LM>), grid, dim3(kBlock), 0, st, g, e, d); } while (0)
    switch (e.lmask) {
        case kL6: if constexpr (L == 6) MDT_LD(kL6); else return hipErrorInvalidValue; break;
        case kL4: if constexpr (L == 4) MDT_LD(kL4); else return hipErrorInvalidValue; break;
        case kL3: if constexpr (L == 3) MDT_LD(kL3); else return hipErrorInvalidValue; break;
        case kL4 | kL6: if constexpr (L != 3) MDT_LD(kL4 | kL6); else return hipErrorInvalidValue; break;
        case kL3 | kL4: if constexpr (L != 6) MDT_LD(kL3 | kL4); else return hipErrorInvalidValue; break;
        case kL3 | kL6: if constexpr (L != 4) MDT_LD(kL3 | kL6); else return hipErrorInvalidValue; break;
        case kL3 | kL4 | kL6: MDT_LD(kL3 | kL4 | kL6); break;
        default: return hipErrorInvalidValue;
    }
#undef MDT_LD
    return hipGetLastError();
}

hipError_t launch_ld(const Geo& g, const EnvDev& e, const LdDev& d, hipStream_t st) {
    if (g.ntiles == 0) return hipSuccess;
    dim3 grid(xcd_grid(g.ntiles));
    if (d.l == 6) return ld_dispatch<6>(g, e, d, grid, st, false);
    if (d.l == 4) return ld_dispatch<4>(g, e, d, grid, st, false);
    if (d.l == 3) return ld_dispatch<3>(g, e, d, grid, st, false);
    return hipErrorInvalidValue;
}

hipError_t launch_ld_coeff(const Geo& g, const EnvDev& e, const LdDev& d, hipStream_t st) {
    if (g.ntiles == 0) return hipSuccess;
    dim3 grid(xcd_grid(g.ntiles));
    if (d.l == 6) return ld_dispatch<6>(g, e, d, grid, st, true);
    if (d.l == 4) return ld_dispatch<4>(g, e, d, grid, st, true);
    if (d.l == 3) return ld_dispatch<3>(g, e, d, grid, st, true);
    return hipErrorInvalidValue;
}

template <int LM>
static void force_dispatch(const Geo& g, const EnvDev& e, double* f, dim3 grid, hipStream_t st) {
    if (e.extra != nullptr) { hipLaunchKernelGGL((k_force<LM, true, 0>), grid, dim3(kBlock), 0, st, g, e, f); return; }
    if constexpr ((LM & kL3) != 0) if (e.fld_l == 3) { hipLaunchKernelGGL((k_force<LM, false, 3>), grid, dim3(kBlock), 0, st, g, e, f); return; }
    if constexpr ((LM & kL4) != 0) if (e.fld_l == 4) { hipLaunchKernelGGL((k_force<LM, false, 4>), grid, dim3(kBlock), 0, st, g, e, f); return; }
    if constexpr ((LM & kL6) != 0) if (e.fld_l == 6) { hipLaunchKernelGGL((k_force<LM, false, 6>), grid, dim3(kBlock), 0, st, g, e, f); return; }
    hipLaunchKernelGGL((k_force<LM, false, 0>), grid, dim3(kBlock), 0, st, g, e, f);
}

hipError_t launch_force(const Geo& g, const EnvDev& e, double* f, hipStream_t st) {
    if (g.ntiles == 0) return hipSuccess;
    if (e.fld_l != 0 && e.extra != nullptr) return hipErrorInvalidValue;
    dim3 grid(xcd_grid(g.ntiles));
    switch (e.lmask) {
        case kL6: force_dispatch<kL6>(g, e, f, grid, st); break;
        case kL4: force_dispatch<kL4>(g, e, f, grid, st); break;
        case kL4 | kL6: force_dispatch<kL4 | kL6>(g, e, f, grid, st); break;
        case kL3: force_dispatch<kL3>(g, e, f, grid, st); break;
        case kL3 | kL4: force_dispatch<kL3 | kL4>(g, e, f, grid, st); break;
        case kL3 | kL6: force_dispatch<kL3 | kL6>(g, e, f, grid, st); break;
        case kL3 | kL4 | kL6: force_dispatch<kL3 | kL4 | kL6>(g, e, f, grid, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_cv_tail(const ReduceDev& r, hipStream_t st) {
    hipLaunchKernelGGL(k_cv_tail, dim3(r.n_walkers), dim3(1024), 0, st, r);
    return hipGetLastError();
}

hipError_t launch_reduce(const ReduceDev& r, hipStream_t st) {
    if (r.n_walkers == 0) return hipSuccess;
    if (r.affine) hipLaunchKernelGGL(k_reduce<true>, dim3(r.n_walkers), dim3(kRedThreads), 0, st, r);
    else hipLaunchKernelGGL(k_reduce<false>, dim3(r.n_walkers), dim3(kRedThreads), 0, st, r);
    if (r.kind != BIAS_GRID) {
        hipLaunchKernelGGL(k_hills_part, dim3(kHillBlocks, r.n_walkers), dim3(kHillThreads), 0, st, r);
        hipLaunchKernelGGL(k_bias_tail, dim3(r.n_walkers), dim3(64), 0, st, r);
    }
    return hipGetLastError();
}

hipError_t launch_virial(const double4* x4, const double* f, const int32_t* image, int first, int n,
                         const double L[3], double* out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_virial, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, x4, f, image, first, n,
                       L[0], L[1], L[2], out);
    return hipGetLastError();
}

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    int64_t b = (n + 255) / 256; if (b > 4096) b = 4096;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)b), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_f(float* p, int64_t n, float v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    int64_t b = (n + 255) / 256; if (b > 4096) b = 4096;
    hipLaunchKernelGGL(k_fill_f, dim3((unsigned)b), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_bias_eval(const ReduceDev& r, int w, int64_t np, const double* s, double* V,
                            double* dV, hipStream_t st) {
    if (np == 0) return hipSuccess;
    int64_t b = np < 1024 ? np : 1024;
    hipLaunchKernelGGL(k_bias_eval, dim3((unsigned)b), dim3(kBlock), 0, st, r, w, np, s, V, dV);
    return hipGetLastError();
}

hipError_t launch_bias_add(const ReduceDev& r, int w, int64_t nh, const double* s, const double* h,
                           double sigma, hipStream_t st) {
    if (nh == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bias_add, dim3(1), dim3(kBlock), 0, st, r, w, nh, s, h, sigma);
    return hipGetLastError();
}

hipError_t launch_sk_amp(const Geo& g, const SkDev& s, hipStream_t st) {
    dim3 grid((s.nk + kSkTile - 1) / kSkTile, s.nchunk);
    if (s.nk > kSkTiledK) {
        if (s.max_atoms > 0) {
            hipLaunchKernelGGL(k_sk_prep, dim3((s.max_atoms + kSkSlots - 1) / kSkSlots, s.n_walkers), dim3(kBlock), 0,
                               st, g, s);
            hipError_t e0 = hipGetLastError();
            if (e0 != hipSuccess) return e0;
        }
        if (s.wide) hipLaunchKernelGGL(k_sk_amp_tiled<true>, grid, dim3(kBlock), 0, st, s);
        else hipLaunchKernelGGL(k_sk_amp_tiled<false>, grid, dim3(kBlock), 0, st, s);
    } else {
        hipLaunchKernelGGL(k_sk_amp, grid, dim3(kBlock), 0, st, g, s);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (s.nk <= kSkFoldAll) {
        hipLaunchKernelGGL(k_sk_fold_cv, dim3(s.n_walkers), dim3(kBlock), 0, st, s);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_sk_fold, dim3(s.nk, s.n_walkers), dim3(kBlock), 0, st, s);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sk_cv, dim3(s.n_walkers), dim3(kBlock), 0, st, s);
    return hipGetLastError();
}

hipError_t launch_sk_pack(const SkDev& s, hipStream_t st) {
    hipLaunchKernelGGL(k_sk_pack, dim3((s.nk + kBlock - 1) / kBlock, s.n_walkers), dim3(kBlock), 0, st, s);
    return hipGetLastError();
}

hipError_t launch_cv_sums(const ReduceDev& r, const int32_t* psum, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_cv_sums, dim3(r.n_walkers), dim3(1024), 0, st, r, psum, out);
    return hipGetLastError();
}

hipError_t launch_sk_force(const Geo& g, const SkDev& s, double* f, hipStream_t st) {
    // grid.y = walkers, grid.x covers the largest walker
    if (s.nk > kSkTiledK) {
        dim3 grid((s.max_atoms + kBlock * kSkAtoms - 1) / (kBlock * kSkAtoms), s.n_walkers);
        if (s.wide) hipLaunchKernelGGL(k_sk_force_tiled<true>, grid, dim3(kBlock), 0, st, s, f);
        else hipLaunchKernelGGL(k_sk_force_tiled<false>, grid, dim3(kBlock), 0, st, s, f);
    } else if (s.nk <= kSkFewK) {
        dim3 grid((s.max_atoms + kBlock - 1) / kBlock, s.n_walkers);
        hipLaunchKernelGGL(k_sk_force<1>, grid, dim3(kBlock), 0, st, g, s, f);
    } else {
        dim3 grid((s.max_atoms + kBlock * kSkAtoms - 1) / (kBlock * kSkAtoms), s.n_walkers);
        hipLaunchKernelGGL(k_sk_force<kSkAtoms>, grid, dim3(kBlock), 0, st, g, s, f);
    }
    return hipGetLastError();
}

hipError_t launch_fes_sum(int D, int64_t nh, const double* hills, const double* x, int n0, const double* y,
                          int n1, double lo0, double lo1, double d0, double d1, double* out, hipStream_t st) {
    const int64_t nout = D == 2 ? (int64_t)n0 * n1 : n0;
    if (nout <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fes_sum, dim3((unsigned)((nout + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, D, nh, hills,
                       x, n0, y, n1, lo0, lo1, d0, d1, out);
    return hipGetLastError();
}

hipError_t launch_fes_marginal(const double* F, int n0, int n1, int axis, double kT, double bin, double* out,
                               hipStream_t st) {
    const int nout = axis == 1 ? n0 : n1;
    if (nout <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fes_marginal, dim3((unsigned)((nout + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, F, n0, n1,
                       axis, kT, bin, out);
    hipLaunchKernelGGL(k_sub_min, dim3(1), dim3(kBlock), 0, st, out, nout);
    return hipGetLastError();
}

hipError_t launch_dist_value(const double4* x4, const DistDev& d, hipStream_t st) {
    hipLaunchKernelGGL(k_dist_value, dim3((d.n_walkers + 63) / 64), dim3(64), 0, st, x4, d);
    return hipGetLastError();
}

hipError_t launch_dist_force(const DistDev& d, double* f, hipStream_t st) {
    hipLaunchKernelGGL(k_dist_force, dim3((d.n_walkers + 63) / 64), dim3(64), 0, st, d, f);
    return hipGetLastError();
}

hipError_t launch_sf_value(const Geo& g, const SfDev& d, hipStream_t st) {
    if (g.ntiles > 0) hipLaunchKernelGGL(k_sf_value, dim3(g.ntiles), dim3(64), 0, st, g, d);
    hipLaunchKernelGGL(k_sf_sum, dim3(d.n_walkers), dim3(64), 0, st, d);
    return hipGetLastError();
}

hipError_t launch_sf_force(const Geo& g, const SfDev& d, double* f, hipStream_t st) {
    if (g.ntiles > 0) hipLaunchKernelGGL(k_sf_force, dim3(g.ntiles), dim3(64), 0, st, g, d, f);
    return hipGetLastError();
}

// one STEINH L kernel: 0 k_lq_q, 1 k_lq_lq, 2 k_lq_sum (values, before the reduce);
// 3 k_lq_k, 4 k_lq_h, 5 k_lq_pair, 6 k_lq_force (forces, after it)
template <int L>
hipError_t launch_lq_l(int stage, const Geo& g, const LqDev& d, double* f, hipStream_t st) {
    const dim3 gr(xcd_grid(g.ntiles)), bl(64);
    if (stage == 2) { hipLaunchKernelGGL(k_lq_sum, dim3(d.n_walkers), dim3(kBlock), 0, st, d); return hipGetLastError(); }
    if (g.ntiles == 0) return hipGetLastError();
    switch (stage) {
        case 0: hipLaunchKernelGGL(k_lq_q<L>, gr, bl, 0, st, g, d); break;
        case 1: hipLaunchKernelGGL(k_lq_lq<L>, gr, dim3(kBlock), 0, st, g, d); break;
        case 3: hipLaunchKernelGGL(k_lq_k<L>, gr, bl, 0, st, g, d); break;
        case 4: hipLaunchKernelGGL(k_lq_h<L>, gr, dim3(kBlock), 0, st, g, d); break;
        case 5: hipLaunchKernelGGL(k_lq_pair<L>, gr, dim3(kBlock), 0, st, g, d); break;
        case 6: hipLaunchKernelGGL(k_lq_force, gr, dim3(kBlock), 0, st, g, d, f); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_lq(int stage, const Geo& g, const LqDev& d, double* f, hipStream_t st) {
    switch (d.l) {
        case 3: return launch_lq_l<3>(stage, g, d, f, st);
        case 4: return launch_lq_l<4>(stage, g, d, f, st);
        case 6: return launch_lq_l<6>(stage, g, d, f, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_lq_erev(const Geo& g, int64_t n, const int32_t* reown, const int32_t* rrev, const int64_t* apos,
                          int64_t* erev, int32_t* eidx, int32_t* eown,
                          hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_lq_erev, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, n, reown, rrev, apos,
                                  erev, eidx, eown);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free virial (mdt_fix_virial_noalloc): the per-block
// partial sums of k_virial, 6 doubles per 256-atom block, for walkers of up to
// kScratchBlocks * kBlock local atoms.  One copy per device, kept until the process
// ends; its device address is looked up once per device (scratch_setup) and only
// that address is handed to kernels.  Every call on a device shares it, so calls that
// overlap on one device race; every block's partials are written before they are
// read, so nothing in it has to start at zero.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[6 * kScratchBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the six sums of a walker's block partials, each in block order (the order of the
// host sum of mdt_fix_virial, so both give the same bits)
__global__ void k_virial_sum(const double* __restrict__ part, int nb, double* __restrict__ v) {
    const int k = threadIdx.x;
    if (k >= 6) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[6 * b + k];
    v[k] = s;
}
}  // namespace

hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<double*>(p);
    return hipSuccess;
}

double* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

hipError_t launch_virial_sum(const double* part, int nb, double* v, hipStream_t st) {
    hipLaunchKernelGGL(k_virial_sum, dim3(1), dim3(64), 0, st, part, nb, v);
    return hipGetLastError();
}


}  // namespace mdt
