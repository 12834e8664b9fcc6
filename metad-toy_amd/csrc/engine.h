// Internal interface between the host orchestration (fix.cpp) and the gfx950
// kernels (kernels.hip).  Not part of the public ABI.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace mdt {

constexpr int kBlock = 256;        // threads per workgroup for per-atom kernels
constexpr int kMaxList = 128;      // max neighbour-list length per atom (2 x 64-bit masks)
constexpr int kMaxSlots = 64;      // max cutoff_Natoms
constexpr int kMaxSlotCV = 16;     // CV value slots per walker
constexpr int kMaxSym = 8;         // symbols usable in DIM expressions
constexpr int kMaxDim = 3;
constexpr int kMaxProg = 96;       // expression bytecode length per dimension
constexpr int kStack = 16;

// Bit set of Steinhardt degrees an environment computes.
constexpr int kL3 = 1, kL4 = 2, kL6 = 4;

// Per-atom record: for each l present (3, 4, 6 in that order) 2l+1 floats
// (Re q_l0, then Re/Im q_lm for m = 1..l), then one scale a_l per l,
// padded to a multiple of 4 floats.
__host__ __device__ constexpr int rec_nl(int lmask) {
    return ((lmask & kL3) ? 1 : 0) + ((lmask & kL4) ? 1 : 0) + ((lmask & kL6) ? 1 : 0);
}
__host__ __device__ constexpr int rec_qfloats(int lmask) {
    return ((lmask & kL3) ? 7 : 0) + ((lmask & kL4) ? 9 : 0) + ((lmask & kL6) ? 13 : 0);
}
__host__ __device__ constexpr int rec_stride(int lmask) {
    return (rec_qfloats(lmask) + rec_nl(lmask) + 3) / 4 * 4;
}
// offset of degree l's block inside the record
__host__ __device__ constexpr int rec_off(int lmask, int l) {
    return l == 3 ? 0
         : l == 4 ? ((lmask & kL3) ? 7 : 0)
                  : ((lmask & kL3) ? 7 : 0) + ((lmask & kL4) ? 9 : 0);
}
// index (0..nl-1) of degree l among the present ones
__host__ __device__ constexpr int rec_li(int lmask, int l) {
    return l == 3 ? 0 : l == 4 ? ((lmask & kL3) ? 1 : 0)
                               : ((lmask & kL3) ? 1 : 0) + ((lmask & kL4) ? 1 : 0);
}

// A tile is a run of <= kBlock consecutive local atoms of one walker.
struct Tile {
    int32_t start;
    int32_t count;
    int32_t walker;
    int32_t pad;
};

// Geometry shared by all per-atom kernels (device pointers).
struct Geo {
    const double* x;          // (n_local_total + n_ghost_total) * 3
    const int32_t* off;       // n_local_total + 1 (global CSR offsets)
    const int32_t* idx;       // global all-space indices
    const int32_t* owner;     // ghost -> local index (indexed by j - n_local_total)
    const uint8_t* revpos;    // per CSR entry: position of the reverse entry, 0xFF = none
    const uint8_t* group;     // per local atom, NULL = all
    const Tile* tiles;
    int32_t ntiles;
    int32_t n_local_total;
};

// One N-nearest environment (cutoff, cap) and the Steinhardt degrees on it.
struct EnvDev {
    double rc2;
    int32_t N;
    int32_t lmask;
    float* rec;               // n_local_total * rec_stride(lmask)
    float* ql;                // nl * n_local_total (SoA per degree)
    int32_t* nsel;            // n_local_total
    int32_t* cnt_rc;          // n_local_total
    uint64_t* selmask;        // n_local_total * 2
    float* extra;             // optional n_local_total * rec_stride: extra coefficients
    double* partial;          // [slot][tile]
    int32_t slot_ql[3];       // partial slot per present degree (order 3,4,6)
    const double* inv_ng;     // per walker 1/N_group
    const double* wcv;        // per walker [kMaxSlotCV] weights dV/dv (from the reduce kernel)
    int32_t wslot[3];         // wcv slot of Q_l.AVE per present degree (-1 = not biased)
};

// Lechner-Dellago local Q_l on an environment.
struct LdDev {
    int32_t l;
    float* lq;                // n_local_total
    float* bq;                // n_local_total * (2l+1): P/(Ng lq (n+1)) * qbar
    int32_t slot;             // partial slot of the lq sum
    int32_t wslot;            // wcv slot of LQ.AVE (-1 = not biased)
};

// Expression program (reverse Polish) for one dimension.
enum ExprOp : int32_t {
    OP_CONST = 0, OP_SYM, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_POW, OP_NEG,
    OP_SIN, OP_COS, OP_TAN, OP_EXP, OP_LOG, OP_SQRT, OP_ABS, OP_TANH, OP_MIN, OP_MAX,
    OP_ASIN, OP_ACOS, OP_ATAN, OP_SINH, OP_COSH, OP_LOG10
};
struct ExprProg {
    int32_t n;
    int32_t op[kMaxProg];
    double c[kMaxProg];       // constant (OP_CONST) or symbol index (OP_SYM)
};

// Bias representations.
enum BiasKind : int32_t { BIAS_GRID = 0, BIAS_HILLS = 2, BIAS_MLP = 3 };

// Everything the per-walker reduce/bias kernel needs (device pointers).
struct ReduceDev {
    int32_t n_walkers;
    int32_t nslot_sum;        // partial-sum slots
    int32_t ntiles;
    const Tile* tiles;
    const int32_t* tile_begin;  // per walker first tile, n_walkers + 1
    const double* partial;    // [slot][tile]
    const double* inv_ng;     // per walker
    const double* direct;     // per walker [kMaxSlotCV] values written by other kernels
    int32_t ncv;              // value slots
    int32_t cv_src[kMaxSlotCV];  // >= 0: partial slot; -1: direct
    int32_t nsym;
    int32_t sym_slot[kMaxSym];
    int32_t D;
    const ExprProg* prog;     // D programs
    // bias
    int32_t kind;
    int32_t pace;
    double lo[kMaxDim], hi[kMaxDim];
    int32_t nb[kMaxDim];
    int32_t radius[kMaxDim];
    int64_t grid_size;
    double* grid;             // per walker grid_size
    double sigma[kMaxDim];    // hill widths per dimension (grid: isotropic sigma[0])
    double h0, biasf, kT;
    int32_t wt;
    double* hills;            // per walker [cap][D+3]: step, s..., h, sigma
    int64_t* nhills;          // per walker
    int64_t hills_cap;
    // per-walker step counters and outputs
    int64_t* step;            // per walker
    double* state;            // per walker [kStateStride]
    double* wcv;              // per walker [kMaxSlotCV]
    // learned bias
    const float* mlp;         // packed network (see kernels.hip), NULL unless kind == BIAS_MLP
    const int32_t* mlp_ok;    // per walker: 1 = network trusted, 0 = use exact sum
    double mlp_scale_in[kMaxDim], mlp_shift_in[kMaxDim], mlp_scale_out, mlp_shift_out;
};

// state layout per walker
constexpr int kStCv = 0;                       // [kMaxSlotCV] CV values
constexpr int kStS = kStCv + kMaxSlotCV;       // [kMaxDim] s
constexpr int kStV = kStS + kMaxDim;           // V after deposit
constexpr int kStDV = kStV + 1;                // [kMaxDim] dV/ds
constexpr int kStHill = kStDV + kMaxDim;       // height deposited this step (0 if none)
constexpr int kStStep = kStHill + 1;           // step of this state
constexpr int kStVexact = kStStep + 1;         // exact hills-sum V (learned-bias check)
constexpr int kStStride = 32;

// ---- launchers (kernels.hip) ----
hipError_t launch_revpos(const Geo& g, int64_t n_local_total, uint8_t* revpos, hipStream_t st);
hipError_t launch_env(const Geo& g, const EnvDev& e, hipStream_t st);
hipError_t launch_select_literal(const Geo& g, int32_t first, int32_t count, double rc2, int N,
                                 int32_t* sel_idx, int32_t* cnt_rc, int32_t* nsel, hipStream_t st);
hipError_t launch_ld(const Geo& g, const EnvDev& e, const LdDev& d, hipStream_t st);
hipError_t launch_ld_coeff(const Geo& g, const EnvDev& e, const LdDev& d, hipStream_t st);
hipError_t launch_force(const Geo& g, const EnvDev& e, double* f, hipStream_t st);
hipError_t launch_reduce(const ReduceDev& r, hipStream_t st);
hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t st);
hipError_t launch_fill_f(float* p, int64_t n, float v, hipStream_t st);
hipError_t launch_bias_eval(const ReduceDev& r, int walker, int64_t np, const double* s,
                            double* V, double* dV, hipStream_t st);
hipError_t launch_bias_add(const ReduceDev& r, int walker, int64_t nh, const double* s,
                           const double* h, double sigma, hipStream_t st);

// Structure factor
struct SkDev {
    int32_t nk;
    const double* k;          // nk * 3 (k / 2pi, revolutions per length)
    const double* kvec;       // nk * 3 (k itself)
    int32_t nchunk;
    const int32_t* chunk;     // per chunk: start, count, walker (3 ints)
    double* apart;            // [chunk][nk][2]
    double* amp;              // per walker [nk][2]
    float* ampf;              // per walker [nk][2] scaled for the gradient
    const int32_t* chunk_begin;  // per walker first chunk (n_walkers + 1)
    double* direct;           // per walker [kMaxSlotCV]
    int32_t slot;             // direct slot of SK.AVE
    int32_t wslot;            // wcv slot
    const double* wcv;
    const int32_t* n_atoms;   // per walker
    const int32_t* atom_begin;// per walker
    int32_t n_walkers;
    int32_t max_atoms;
};
hipError_t launch_sk_amp(const Geo& g, const SkDev& s, hipStream_t st);
hipError_t launch_sk_force(const Geo& g, const SkDev& s, double* f, hipStream_t st);

}  // namespace mdt
