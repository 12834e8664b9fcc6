// Synthetic atomistic systems and LAMMPS-style full neighbour lists.
//
// This is the input side of the CV path: in a LAMMPS run, Atom::x (locals
// followed by ghost images) and a full NeighList come from the MD engine
// (src/core/neigh_hub.cu:191-231 delivers them to the reference).  Here the
// same data are produced on the host for the standalone driver, the tests and
// the benchmark:
//   * seeded FCC lattices with per-coordinate Gaussian jitter, split into a
//     lower and an upper slab with different jitter widths,
//   * uniform random ("liquid") atoms in a z range,
//   * ghost images of a periodic orthorhombic box out to the list cutoff,
//     with a ghost -> owner map,
//   * a full list per local atom over [0, n_all) within the list cutoff.
// List order is part of the input contract because the N-nearest insertion
// breaks ties by list order: bins are visited in the 27-bin stencil order
// (dz outer, dy, dx inner, each -1..1) and atoms inside a bin in ascending
// index order.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <thread>

#include "mdt.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0,1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    // standard normal by Box-Muller (one value per call, the sine branch dropped)
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// static-chunk parallel loop over [0, n) on the host cores
template <class F>
void parallel_for(int64_t n, F&& f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        const int64_t a = t * chunk, b = std::min<int64_t>(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([&f, a, b] { for (int64_t i = a; i < b; ++i) f(i); });
    }
    for (auto& x : th) x.join();
}

inline double wrap(double v, double lo, double L) {
    double t = v - lo;
    t -= L * std::floor(t / L);
    if (t >= L) t -= L;     // floor rounding at the upper edge
    return lo + t;
}

}  // namespace

struct mdt_system {
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    std::vector<double> x;        // local atoms, 3 per atom
    std::vector<double> xall;     // locals then ghosts
    std::vector<int32_t> owner;   // ghost -> local owner
    std::vector<int32_t> offsets; // n_local + 1
    std::vector<int32_t> idx;
    double rg = 0;
    int64_t n_local = 0, n_all = 0;
    std::string err;
};

extern "C" {

int mdt_system_create(mdt_system** out) {
    if (!out) return -1;
    *out = new mdt_system();
    return 0;
}

void mdt_system_destroy(mdt_system* s) { delete s; }

const char* mdt_system_error(const mdt_system* s) { return s ? s->err.c_str() : "null system"; }

int mdt_system_set_box(mdt_system* s, const double lo[3], const double hi[3]) {
    for (int d = 0; d < 3; ++d) {
        if (!(hi[d] > lo[d])) { s->err = "box hi must exceed lo"; return -1; }
        s->lo[d] = lo[d]; s->hi[d] = hi[d];
    }
    return 0;
}

void mdt_system_get_box(const mdt_system* s, double lo[3], double hi[3]) {
    for (int d = 0; d < 3; ++d) { lo[d] = s->lo[d]; hi[d] = s->hi[d]; }
}

// FCC lattice of nx*ny*nz cubic cells of edge a (4 atoms per cell, cells in
// C order x-slowest, basis (0,0,0),(0,.5,.5),(.5,0,.5),(.5,.5,0)), appended to
// the system.  The box becomes [0,nx a)x[0,ny a)x[0,nz a) if it is unset.
// Every coordinate gets N(0, sigma^2) jitter, sigma = sigma_below for lattice
// sites with z < z_split and sigma_above otherwise; positions are wrapped.
int mdt_system_add_fcc(mdt_system* s, int nx, int ny, int nz, double a, double sigma_below,
                       double sigma_above, double z_split, uint64_t seed) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(a > 0)) { s->err = "bad FCC shape"; return -1; }
    if (!(s->hi[0] > s->lo[0])) {
        double lo[3] = {0, 0, 0}, hi[3] = {nx * a, ny * a, nz * a};
        mdt_system_set_box(s, lo, hi);
    }
    static const double basis[4][3] = {{0, 0, 0}, {0, .5, .5}, {.5, 0, .5}, {.5, .5, 0}};
    SplitMix64 rng(seed);
    double L[3] = {s->hi[0] - s->lo[0], s->hi[1] - s->lo[1], s->hi[2] - s->lo[2]};
    s->x.reserve(s->x.size() + (size_t)12 * nx * ny * nz);
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz)
                for (int b = 0; b < 4; ++b) {
                    double p[3] = {(ix + basis[b][0]) * a, (iy + basis[b][1]) * a, (iz + basis[b][2]) * a};
                    double sig = p[2] < z_split ? sigma_below : sigma_above;
                    for (int d = 0; d < 3; ++d) {
                        double v = s->lo[d] + p[d];
                        if (sig > 0) v += sig * rng.normal();
                        s->x.push_back(wrap(v, s->lo[d], L[d]));
                    }
                }
    return 0;
}

// n atoms uniform in the box (x, y) and in [z0, z1) along z, appended.
int mdt_system_add_uniform(mdt_system* s, int64_t n, double z0, double z1, uint64_t seed) {
    if (!(s->hi[0] > s->lo[0])) { s->err = "set the box before adding uniform atoms"; return -1; }
    SplitMix64 rng(seed);
    for (int64_t i = 0; i < n; ++i) {
        double ux = rng.uniform(), uy = rng.uniform(), uz = rng.uniform();
        s->x.push_back(s->lo[0] + ux * (s->hi[0] - s->lo[0]));
        s->x.push_back(s->lo[1] + uy * (s->hi[1] - s->lo[1]));
        s->x.push_back(z0 + uz * (z1 - z0));
    }
    return 0;
}

// Replace the local positions (n atoms, 3 doubles each).
int mdt_system_set_positions(mdt_system* s, const double* x, int64_t n) {
    s->x.assign(x, x + 3 * n);
    return 0;
}

// Ghost images and the full neighbour list within rg (r^2 <= rg^2, j != i).
// The box must exceed 2*rg along every axis so that an atom sees at most one
// image of any other atom inside the list cutoff.
int mdt_system_build(mdt_system* s, double rg) {
    const int64_t n = (int64_t)(s->x.size() / 3);
    double L[3];
    for (int d = 0; d < 3; ++d) {
        L[d] = s->hi[d] - s->lo[d];
        if (!(L[d] > 2.0 * rg)) { s->err = "box edge must exceed twice the list cutoff"; return -1; }
    }
    s->rg = rg;
    s->n_local = n;
    // ghosts: every nonzero image shift whose image lies within rg of the box
    s->xall.assign(s->x.begin(), s->x.end());
    s->owner.clear();
    for (int sz = -1; sz <= 1; ++sz)
        for (int sy = -1; sy <= 1; ++sy)
            for (int sx = -1; sx <= 1; ++sx) {
                if (!sx && !sy && !sz) continue;
                int sh[3] = {sx, sy, sz};
                for (int64_t i = 0; i < n; ++i) {
                    double p[3];
                    bool keep = true;
                    for (int d = 0; d < 3; ++d) {
                        p[d] = s->x[3 * i + d] + sh[d] * L[d];
                        if (p[d] < s->lo[d] - rg || p[d] >= s->hi[d] + rg) { keep = false; break; }
                    }
                    if (!keep) continue;
                    s->xall.push_back(p[0]); s->xall.push_back(p[1]); s->xall.push_back(p[2]);
                    s->owner.push_back((int32_t)i);
                }
            }
    const int64_t na = (int64_t)(s->xall.size() / 3);
    s->n_all = na;
    // bins over the extended box
    double elo[3], bw[3];
    int nb[3];
    for (int d = 0; d < 3; ++d) {
        elo[d] = s->lo[d] - rg;
        double ext = L[d] + 2 * rg;
        nb[d] = std::max(1, (int)std::floor(ext / rg));
        bw[d] = ext / nb[d];
    }
    const int64_t nbins = (int64_t)nb[0] * nb[1] * nb[2];
    std::vector<int32_t> bin_of(na), bin_start(nbins + 1, 0), bin_atoms(na);
    for (int64_t j = 0; j < na; ++j) {
        int b[3];
        for (int d = 0; d < 3; ++d) {
            int v = (int)std::floor((s->xall[3 * j + d] - elo[d]) / bw[d]);
            b[d] = std::min(std::max(v, 0), nb[d] - 1);
        }
        bin_of[j] = (int32_t)(((int64_t)b[2] * nb[1] + b[1]) * nb[0] + b[0]);
        bin_start[bin_of[j] + 1]++;
    }
    for (int64_t b = 0; b < nbins; ++b) bin_start[b + 1] += bin_start[b];
    {
        std::vector<int32_t> fill(bin_start.begin(), bin_start.end() - 1);
        for (int64_t j = 0; j < na; ++j) bin_atoms[fill[bin_of[j]]++] = (int32_t)j;  // ascending index
    }
    const double rg2 = rg * rg;
    std::vector<int32_t> count(n, 0);
    auto visit = [&](int64_t i, auto&& emit) {
        int b = bin_of[i];
        int bx = b % nb[0], by = (b / nb[0]) % nb[1], bz = b / (nb[0] * nb[1]);
        const double* xi = &s->xall[3 * i];
        for (int dz = -1; dz <= 1; ++dz) {
            int z = bz + dz; if (z < 0 || z >= nb[2]) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                int y = by + dy; if (y < 0 || y >= nb[1]) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    int xx = bx + dx; if (xx < 0 || xx >= nb[0]) continue;
                    int64_t bb = ((int64_t)z * nb[1] + y) * nb[0] + xx;
                    for (int32_t t = bin_start[bb]; t < bin_start[bb + 1]; ++t) {
                        int32_t j = bin_atoms[t];
                        if (j == i) continue;
                        const double* xj = &s->xall[3 * (int64_t)j];
                        double ddx = xj[0] - xi[0], ddy = xj[1] - xi[1], ddz = xj[2] - xi[2];
                        double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (r2 <= rg2) emit(j);
                    }
                }
            }
        }
    };
    parallel_for(n, [&](int64_t i) {
        int32_t c = 0;
        visit(i, [&](int32_t) { ++c; });
        count[i] = c;
    });
    s->offsets.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        int64_t v = (int64_t)s->offsets[i] + count[i];
        if (v > INT32_MAX) { s->err = "neighbour list exceeds 2^31 entries"; return -1; }
        s->offsets[i + 1] = (int32_t)v;
    }
    s->idx.assign((size_t)s->offsets[n], 0);
    parallel_for(n, [&](int64_t i) {
        int32_t w = s->offsets[i];
        visit(i, [&](int32_t j) { s->idx[w++] = j; });
    });
    return 0;
}

int64_t mdt_system_n_local(const mdt_system* s) { return s->n_local ? s->n_local : (int64_t)(s->x.size() / 3); }
int64_t mdt_system_n_all(const mdt_system* s) { return s->n_all; }
int64_t mdt_system_nnz(const mdt_system* s) { return s->offsets.empty() ? 0 : s->offsets.back(); }

// Copy out any of: x_all (n_all*3), offsets (n_local+1), idx (nnz), ghost_owner
// (n_all - n_local).  Before build, x_all receives the local positions.
int mdt_system_copy(const mdt_system* s, double* x_all, int32_t* offsets, int32_t* idx,
                    int32_t* ghost_owner) {
    if (x_all) {
        const std::vector<double>& src = s->xall.empty() ? s->x : s->xall;
        std::memcpy(x_all, src.data(), src.size() * sizeof(double));
    }
    if (offsets) std::memcpy(offsets, s->offsets.data(), s->offsets.size() * sizeof(int32_t));
    if (idx) std::memcpy(idx, s->idx.data(), s->idx.size() * sizeof(int32_t));
    if (ghost_owner) std::memcpy(ghost_owner, s->owner.data(), s->owner.size() * sizeof(int32_t));
    return 0;
}

}  // extern "C"
