// Device build of ghost images and full neighbour lists for a periodic
// orthorhombic box: the input side of the CV path that LAMMPS provides in a
// run (NeighHub::ensure/list, src/core/neigh_hub.cu:191-231, delivers Atom::x
// with ghosts and a full NeighList to the reference), built here on the GPU
// so a standalone run needs no host list.
//
// The order is part of the contract (the N-nearest insertion breaks ties by
// list order) and equals the host generator's:
//   * ghosts: image shifts (sz, sy, sx) each over -1..1 with sz outermost,
//     (0,0,0) skipped; within a shift the local atoms in ascending index;
//     an image is kept when every coordinate lies in [lo - rg, hi + rg);
//   * bins over the extended box, nb_d = max(1, floor(ext_d / rg)) of width
//     ext_d / nb_d; an atom's bin is floor((x - (lo - rg)) / width), clamped;
//   * a list visits the 27 bins of the stencil (dz outer, dy, dx inner, each
//     -1..1) and the atoms of a bin in ascending index, keeping r^2 <= rg^2
//     (FP64, r^2 = (dx^2 + dy^2) + dz^2 without contraction).
// Ghost compaction and list offsets are exclusive scans; bins are a stable
// radix sort of (bin, index) pairs, so each bin holds its atoms in ascending
// index order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <cmath>
#include <cstdint>

#include "engine.h"

namespace mdt {
namespace {

constexpr int kNlBlock = 256;

__device__ __forceinline__ bool image_in(const NlBox& b, const double* x, int s, double* p) {
    const int sh[3] = {s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1};   // sx fastest, sz slowest
    bool keep = true;
    #pragma unroll
    for (int d = 0; d < 3; ++d) {
        p[d] = x[d] + sh[d] * b.L[d];
        keep = keep && !(p[d] < b.lo[d] - b.rg || p[d] >= b.hi[d] + b.rg);
    }
    return keep;
}

// flag[k * n + i] = image of i under the k-th nonzero shift is a ghost (k = 0..25)
__global__ __launch_bounds__(kNlBlock) void k_ghost_flags(const double* __restrict__ x, int64_t n, NlBox b,
                                                          int32_t* __restrict__ flag) {
#pragma clang fp contract(off)
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 26 * n) return;
    const int k = (int)(q / n);
    const int64_t i = q - (int64_t)k * n;
    const int s = k < 13 ? k : k + 1;                               // skip shift (0,0,0) = 13
    double p[3];
    flag[q] = image_in(b, x + 3 * i, s, p) ? 1 : 0;
}

__global__ __launch_bounds__(kNlBlock) void k_ghost_scatter(const double* __restrict__ x, int64_t n, NlBox b,
                                                            const int32_t* __restrict__ flag,
                                                            const int32_t* __restrict__ pos,
                                                            double* __restrict__ xall, int32_t* __restrict__ owner) {
#pragma clang fp contract(off)
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 26 * n) return;
    if (!flag[q]) return;
    const int k = (int)(q / n);
    const int64_t i = q - (int64_t)k * n;
    const int s = k < 13 ? k : k + 1;
    double p[3];
    image_in(b, x + 3 * i, s, p);
    const int64_t g = pos[q];
    xall[3 * (n + g)] = p[0];
    xall[3 * (n + g) + 1] = p[1];
    xall[3 * (n + g) + 2] = p[2];
    owner[g] = (int32_t)i;
}

__global__ __launch_bounds__(kNlBlock) void k_copy_f64(const double* __restrict__ a, double* __restrict__ out,
                                                       int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = a[q];
}

__device__ __forceinline__ int bin_coord(double v, double elo, double bw, int nb) {
    int c = (int)floor((v - elo) / bw);
    return c < 0 ? 0 : (c > nb - 1 ? nb - 1 : c);
}

__global__ __launch_bounds__(kNlBlock) void k_bin(const double* __restrict__ xall, int64_t na, NlBox b,
                                                  int32_t* __restrict__ bin_of, int32_t* __restrict__ val,
                                                  int32_t* __restrict__ bin_count) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= na) return;
    int c[3];
    #pragma unroll
    for (int d = 0; d < 3; ++d) c[d] = bin_coord(xall[3 * j + d], b.elo[d], b.bw[d], b.nb[d]);
    const int32_t bb = (int32_t)(((int64_t)c[2] * b.nb[1] + c[1]) * b.nb[0] + c[0]);
    bin_of[j] = bb;
    val[j] = (int32_t)j;
    atomicAdd(bin_count + bb, 1);
}

// COUNT: write the list length of each local atom; else write the list at off[i]
template <bool COUNT>
__global__ __launch_bounds__(kNlBlock) void k_list(const double* __restrict__ xall, int64_t n, NlBox b,
                                                   const int32_t* __restrict__ bin_of,
                                                   const int32_t* __restrict__ bin_start,
                                                   const int32_t* __restrict__ bin_atoms,
                                                   int64_t* __restrict__ count, const int64_t* __restrict__ off,
                                                   int32_t* __restrict__ idx) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t bi = bin_of[i];
    const int bx = bi % b.nb[0], by = (bi / b.nb[0]) % b.nb[1], bz = bi / (b.nb[0] * b.nb[1]);
    const double x0 = xall[3 * i], x1 = xall[3 * i + 1], x2 = xall[3 * i + 2];
    const double rg2 = b.rg * b.rg;
    int64_t w = COUNT ? 0 : off[i];
    int64_t c = 0;
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = bz + dz;
        if (z < 0 || z >= b.nb[2]) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = by + dy;
            if (y < 0 || y >= b.nb[1]) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int xx = bx + dx;
                if (xx < 0 || xx >= b.nb[0]) continue;
                const int64_t bb = ((int64_t)z * b.nb[1] + y) * b.nb[0] + xx;
                for (int32_t t = bin_start[bb]; t < bin_start[bb + 1]; ++t) {
                    const int32_t j = bin_atoms[t];
                    if (j == i) continue;
                    const double d0 = xall[3 * (int64_t)j] - x0, d1 = xall[3 * (int64_t)j + 1] - x1;
                    const double d2 = xall[3 * (int64_t)j + 2] - x2;
                    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                    if (r2 <= rg2) {
                        if (COUNT) ++c;
                        else idx[w++] = j;
                    }
                }
            }
        }
    }
    if (COUNT) count[i] = c;
}

__global__ __launch_bounds__(kNlBlock) void k_off32(const int64_t* __restrict__ o64, int32_t* __restrict__ o32,
                                                    int64_t n) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) o32[q] = (int32_t)o64[q];
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kNlBlock - 1) / kNlBlock); }

}  // namespace

// ---- launchers (fix.cpp owns the buffers and the order of the phases) ----
hipError_t nl_ghost_flags(const double* x, int64_t n, const NlBox& b, int32_t* flag, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ghost_flags, dim3(nblk(26 * n)), dim3(kNlBlock), 0, st, x, n, b, flag);
    return hipGetLastError();
}

hipError_t nl_ghost_scatter(const double* x, int64_t n, const NlBox& b, const int32_t* flag, const int32_t* pos,
                            double* xall, int32_t* owner, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_f64, dim3(nblk(3 * n)), dim3(kNlBlock), 0, st, x, xall, 3 * n);
    hipLaunchKernelGGL(k_ghost_scatter, dim3(nblk(26 * n)), dim3(kNlBlock), 0, st, x, n, b, flag, pos, xall, owner);
    return hipGetLastError();
}

hipError_t nl_bin(const double* xall, int64_t na, const NlBox& b, int32_t* bin_of, int32_t* val,
                  int32_t* bin_count, hipStream_t st) {
    if (na == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bin, dim3(nblk(na)), dim3(kNlBlock), 0, st, xall, na, b, bin_of, val, bin_count);
    return hipGetLastError();
}

hipError_t nl_list_count(const double* xall, int64_t n, const NlBox& b, const int32_t* bin_of,
                         const int32_t* bin_start, const int32_t* bin_atoms, int64_t* count, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_list<true>), dim3(nblk(n)), dim3(kNlBlock), 0, st, xall, n, b, bin_of, bin_start,
                       bin_atoms, count, (const int64_t*)nullptr, (int32_t*)nullptr);
    return hipGetLastError();
}

hipError_t nl_list_fill(const double* xall, int64_t n, const NlBox& b, const int32_t* bin_of,
                        const int32_t* bin_start, const int32_t* bin_atoms, const int64_t* off, int32_t* idx,
                        int32_t* off32, hipStream_t st) {
    if (n > 0)
        hipLaunchKernelGGL((k_list<false>), dim3(nblk(n)), dim3(kNlBlock), 0, st, xall, n, b, bin_of, bin_start,
                           bin_atoms, (int64_t*)nullptr, off, idx);
    hipLaunchKernelGGL(k_off32, dim3(nblk(n + 1)), dim3(kNlBlock), 0, st, off, off32, n + 1);
    return hipGetLastError();
}

// exclusive prefix sums and the stable pair sort (temporary storage queried with tmp == nullptr)
hipError_t nl_scan_i32(void* tmp, size_t& bytes, const int32_t* in, int32_t* out, int64_t n, hipStream_t st) {
    return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, st);
}
hipError_t nl_scan_i64(void* tmp, size_t& bytes, const int64_t* in, int64_t* out, int64_t n, hipStream_t st) {
    return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, (int)n, st);
}
hipError_t nl_sort_pairs(void* tmp, size_t& bytes, const int32_t* kin, int32_t* kout, const int32_t* vin,
                         int32_t* vout, int64_t n, int end_bit, hipStream_t st) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, (int)n, 0, end_bit, st);
}

}  // namespace mdt
