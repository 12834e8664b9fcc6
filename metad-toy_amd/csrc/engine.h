// Internal interface between the host orchestration (fix.cpp) and the gfx950
// kernels (kernels.hip).  Not part of the public ABI.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace mdt {

constexpr int kBlock = 256;        // threads per workgroup for per-atom kernels
constexpr int kLanes = 4;          // lanes per atom in the per-atom kernels
constexpr int kTile = kBlock / kLanes;   // atoms per tile (one workgroup)
constexpr int kMaxList = 128;      // max near-list length per atom (2 x 64-bit masks)
// candidates with r^2 <= rc^2 (1 + kRcSlack) in an atom's frame: a superset of every
// pair that either side of a pair can select (the two frames differ by rounding)
constexpr double kRcSlack = 1e-9;
constexpr int kMaxSlots = 64;      // max cutoff_Natoms
constexpr int kMaxSlotCV = 16;     // CV value slots per walker
// per-walker chain-rule weights: [0, kMaxSlotCV) dV/dv of each CV slot; from kMaxSlotCV
// on, per representative slot, the sum over the slots aliased to it (CALs with identical
// STEINH Q definitions share one environment degree and its force pass)
constexpr int kWcvStride = 2 * kMaxSlotCV;
constexpr int kMaxSym = 8;         // symbols usable in DIM expressions
constexpr int kMaxDim = 3;
constexpr int kMaxProg = 96;       // expression bytecode length per dimension
constexpr int kStack = 16;
constexpr int kRedArgWalkers = 64; // walkers whose tile ranges travel in the reduce kernel's arguments
constexpr int64_t kMaxGridBins = 1ll << 28;   // bins per walker grid (2 GiB of FP64)
// Sums over a hill log are spread over kHillBlocks workgroups per walker.  Per-walker
// work buffer of that path: [kHillBlocks][4] partial sums (V, dV/ds), then the
// chain-rule coefficients ds_d/dv_c as [kMaxDim][kMaxSlotCV].
constexpr int kHillBlocks = 64;
constexpr int kBwAc = 4 * kHillBlocks;
constexpr int kBwStride = kBwAc + kMaxDim * kMaxSlotCV;

// Bit set of Steinhardt degrees an environment computes.
constexpr int kL3 = 1, kL4 = 2, kL6 = 4;

// Per-atom record: a header of four floats, the atom's packed position (the
// x4f layout: xyz in FP32 and three 10-bit residuals in bits 0-29 of the fourth
// word) with two flags in bits 30-31 of that word; then for each l present (3, 4,
// 6 in that order) a 16-byte aligned block of 2l+1 floats (Re q_l0, then Re/Im
// q_lm for m = 1..l) followed by the scale a_l = 4pi/(2l+1) / (N_group q_l n).
// Blocks are padded to 4 floats and the record to 64 bytes, so a neighbour's
// record is whole 64-byte quads that four lanes fetch together, one float4 each.
constexpr int kRecHdr = 4;
constexpr uint32_t kHdrCap = 1u << 30;    // the cap binds: the selection is not the whole in-rc set
constexpr uint32_t kHdrBand = 1u << 31;   // some list entry needed the exact FP64 distance test
__host__ __device__ constexpr int rec_nl(int lmask) {
    return ((lmask & kL3) ? 1 : 0) + ((lmask & kL4) ? 1 : 0) + ((lmask & kL6) ? 1 : 0);
}
__host__ __device__ constexpr int rec_blk(int l) { return (2 * l + 2 + 3) / 4 * 4; }
__host__ __device__ constexpr int rec_used(int lmask) {
    return kRecHdr + ((lmask & kL3) ? rec_blk(3) : 0) + ((lmask & kL4) ? rec_blk(4) : 0) +
           ((lmask & kL6) ? rec_blk(6) : 0);
}
__host__ __device__ constexpr int rec_stride(int lmask) { return (rec_used(lmask) + 15) / 16 * 16; }
// offset of degree l's block inside the record
__host__ __device__ constexpr int rec_off(int lmask, int l) {
    return kRecHdr + (l == 3 ? 0
                     : l == 4 ? ((lmask & kL3) ? rec_blk(3) : 0)
                              : ((lmask & kL3) ? rec_blk(3) : 0) + ((lmask & kL4) ? rec_blk(4) : 0));
}
// slot of the scale a_l
__host__ __device__ constexpr int rec_a(int lmask, int l) { return rec_off(lmask, l) + 2 * l + 1; }
// index (0..nl-1) of degree l among the present ones
__host__ __device__ constexpr int rec_li(int lmask, int l) {
    return l == 3 ? 0 : l == 4 ? ((lmask & kL3) ? 1 : 0)
                               : ((lmask & kL3) ? 1 : 0) + ((lmask & kL4) ? 1 : 0);
}

// A tile is a run of <= kTile consecutive local atoms of one walker.
struct Tile {
    int32_t start;
    int32_t count;
    int32_t walker;
    int32_t pad;
};

// Geometry shared by all per-atom kernels (device pointers).
struct Geo {
    const double4* x4;        // (n_local_total + n_ghost_total), xyz + pad, FP64
    // the same FP64 positions as 3 doubles per atom in the x4 order, when the context's
    // one walker reads them from an attached engine buffer (locals then ghosts, which is
    // that order); NULL otherwise.  The selection's exact FP64 tests read it, so a step
    // that has no other FP64 reader need not copy the positions into x4.
    const double* x64;
    const float4* x4f;        // the same positions rounded to FP32 (list-scan prefilter)
    const int32_t* off;       // n_local_total + 1 (global CSR offsets)
    const int32_t* idx;       // global all-space indices
    const int32_t* owner;     // ghost -> local index (indexed by j - n_local_total)
    const int32_t* eown;      // per list entry: local owner of idx[e]
    const uint8_t* revpos;    // per list entry: position of the reverse entry (the image of this
                              // center in the owner's list), 0xFF = none
    const int32_t* cnt;       // per local atom: list length (the near list); NULL: off[i+1] - off[i]
    const uint8_t* group;     // per local atom, NULL = all
    const uint8_t* share;     // per local atom: adds to this context's CV sums (NULL = all)
    const Tile* tiles;
    int32_t ntiles;
    int32_t n_local_total;
};

// The near list: every bound list cut to the entries within a radius Rp above the largest
// Steinhardt cutoff, in list order (so the N-nearest insertion, which never keeps an entry
// beyond rc, selects the same entries in the same order), stored in place at the bound
// list's offsets with a length per atom.  It is rebuilt inside the step whenever some atom
// has moved more than a threshold since the last build (k_pack_walker raises the flag of
// the current epoch; the reduce kernel advances the epoch once per step).
struct NearDev {
    const int32_t* off;       // bound CSR offsets
    const int32_t* idx;       // bound entries
    int32_t* nidx;            // near entries at off[i] .. off[i] + ncnt[i]
    int32_t* ncnt;            // per local atom
    float4* xpr;              // positions at the last build (packed layout: locals, ghosts)
    int32_t* flag;            // [2] rebuild flags, by epoch parity
    int64_t* epoch;           // step epoch counter
    int32_t* overflow;        // set when an atom has more than kMaxList near entries
    double rp2;               // Rp^2
    float move2;              // squared displacement that forces a rebuild
    int32_t n_local, n_all;
};

// One N-nearest environment (cutoff, cap) and the Steinhardt degrees on it.
struct EnvDev {
    double rc2;
    float rcs;                // sqrt(rc2 (1 + kRcSlack)) rounded to FP32 (the scan's error bound)
    int32_t N;
    int32_t lmask;
    float* rec;               // n_local_total * rec_stride(lmask)
    float* ql;                // nl * n_local_total (SoA per degree)
    int32_t* nsel;            // n_local_total
    int32_t* cnt_rc;          // n_local_total
    uint64_t* selmask;        // n_local_total * 4: selected mask, within-rc (slack) mask
    float* extra;             // optional n_local_total * rec_stride: extra coefficients
    double* partial;          // [slot][tile]
    int32_t slot_ql[3];       // partial slot per present degree (order 3,4,6)
    const double* wcv;        // per walker [kWcvStride] weights dV/dv (from the reduce kernel)
    int32_t wslot[3];         // wcv slot of Q_l.AVE per present degree (-1 = not biased)
    // an unbiased Lechner-Dellago CV on this (biased) environment, accumulated by
    // the force kernel from the neighbour records it loads anyway (0 = none)
    int32_t fld_l;
    int32_t fld_slot;         // its partial-sum slot
    float* fld_lq;            // its per-atom lq
    // selection export (the k_env export instance only): slot-ordered neighbour
    // indices [n_local_total * N] and count within rc [n_local_total]
    int32_t* sel_out;
    int32_t* sel_cnt;
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
    int32_t linear;           // 1: s = lin0 + sum_q lin[q] v_q (evaluated without the interpreter)
    int32_t op[kMaxProg];
    double c[kMaxProg];       // constant (OP_CONST) or symbol index (OP_SYM)
    double lin0;
    double lin[kMaxSym];
};

// Bias representations.
enum BiasKind : int32_t { BIAS_GRID = 0, BIAS_HILLS = 2, BIAS_MLP = 3 };

// Learned bias network 3 -> kMlpH -> kMlpH -> 1 (tanh), FP32 master parameters
// (row-major W1[a][d], b1, W2[a][b] with z2[a] = sum_b W2[a][b] h1[b], b2, W3, b3),
// a BF16 copy of W2 for the MFMA products, and the input/output normalisation.
constexpr int kMlpH = 64;
constexpr int kW1 = 0, kB1 = kW1 + kMlpH * 3, kW2 = kB1 + kMlpH, kB2 = kW2 + kMlpH * kMlpH,
              kW3 = kB2 + kMlpH, kB3 = kW3 + kMlpH, kMlpP = kB3 + 1;
// training points per workgroup (one wave) of a training step; a step of `count` points
// writes one gradient row of kMlpP floats per tile
constexpr int kMlpTile = 16;
inline int64_t mlp_grad_rows(int64_t count) { return (count + kMlpTile - 1) / kMlpTile; }
struct MlpNet {
    float p[kMlpP + 3];
    short w2bf[kMlpH * kMlpH];      // BF16(W2)
    short w2lo[kMlpH * kMlpH];      // BF16(W2 - BF16(W2)): hi + lo carries 16 mantissa bits
    double in_shift[3], in_scale[3], out_scale, out_shift;
    int64_t n_fit;            // hills represented by the network; later hills are summed exactly
};
struct MlpDev {
    int32_t D;
    int64_t hills_cap;
    const double* hills;
    const int64_t* nhills;
    double sigma[3];
    double spread;            // sampling width around a hill, in sigmas
    MlpNet* live;             // per walker, used by the step
    MlpNet* stage;            // per walker, being fitted
    float* adam_m;
    float* adam_v;
    int32_t* ok;              // per walker: 1 = live network trusted
    double* err;              // per walker: last held-out max |dV|
    int64_t* fits;
    int64_t* fallbacks;
    double* box_lo_out;       // per walker [3]: box of the hill centres
    double* box_hi_out;
    int64_t* adam_t;          // per walker Adam step of the current fit (advanced by k_mlp_train)
};

// Everything the per-walker reduce/bias kernel needs (device pointers).
struct ReduceDev {
    int32_t n_walkers;
    int32_t nslot_sum;        // partial-sum slots
    int32_t ntiles;
    const Tile* tiles;
    const int32_t* tile_begin;  // per walker first tile, n_walkers + 1
    const double* partial;    // [slot][tile]
    const double* cv_inv;     // per walker [kMaxSlotCV]: 1 / group count of each partial-sum CV
    const double* direct;     // per walker [kMaxSlotCV] values written by other kernels
    int32_t ncv;              // value slots
    int32_t walias[kMaxSlotCV];  // representative slot of each slot's weight sum (-1: none)
    int32_t cv_src[kMaxSlotCV];  // >= 0: partial slot; -1: direct; -2 - s: partial slot s summed after the force pass
    int32_t nsym;
    int32_t sym_slot[kMaxSym];
    int32_t D;
    const ExprProg* prog;     // D programs
    // every program affine: s_d = aff0[d] + sum_c aff[d][c] v_c over value slots c
    int32_t affine;
    double aff0[kMaxDim];
    const double* aff;        // [kMaxDim][kMaxSlotCV]
    // per-walker tile ranges in the arguments (n_walkers <= kRedArgWalkers), so the
    // partial sums load without waiting on tile_begin
    int32_t tile_args;
    int32_t tile_range[kRedArgWalkers + 1];
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
    double* wcv;              // per walker [kWcvStride]
    int64_t* epoch;           // near-list epoch, advanced once per step (NULL: no near list)
    // learned bias
    const MlpNet* mlp;        // per walker live network, NULL unless kind == BIAS_MLP
    const int32_t* mlp_ok;    // per walker: 1 = network trusted, 0 = use exact sum
    const double* mlp_err;    // per walker: held-out max |dV| of the last fit
    double* bwork;            // per walker [kBwStride]: hill-sum partials and ds/dv (hills kinds)
};

// state layout per walker
constexpr int kStCv = 0;                       // [kMaxSlotCV] CV values
constexpr int kStS = kStCv + kMaxSlotCV;       // [kMaxDim] s
constexpr int kStV = kStS + kMaxDim;           // V after deposit
constexpr int kStDV = kStV + 1;                // [kMaxDim] dV/ds
constexpr int kStHill = kStDV + kMaxDim;       // height deposited this step (0 if none)
constexpr int kStStep = kStHill + 1;           // step of this state
constexpr int kStMlpErr = kStStep + 1;         // learned bias: held-out max |dV| of the live fit
constexpr int kStMlpOk = kStMlpErr + 1;        // learned bias: 1 if the network is in use
constexpr int kStStride = 32;

// Switching function of the reference (LINE, STEP, FERMI, TANH, RATIONAL).
struct SwFn {
    int32_t type;             // 0 LINE, 1 STEP, 2 FERMI, 3 TANH, 4 RATIONAL
    int32_t n, m;
    double r0, d0, alpha;
};

// STRU_FACTOR (Debye) CV: per-atom S_i over each atom's list within rc.
struct SfDev {
    int32_t count;            // 0: AVE (sum S_i / N_group); 1: COUNT (sum sw(S_i))
    int32_t pair;             // pair term g(r): 0 sin(q r)/(q r) s(r^2) (STRU_FACTOR); 1 s(r^2) (WEIGHT_CHEM)
    SwFn sw;
    double rc2, q;
    const double* aw;         // per atom (locals then ghosts) chemical weight w_type; NULL = 1
    double* S;                // per local atom S_i
    double* part;             // per tile FP64 sum of S_i over group atoms
    const int32_t* tile_begin;  // per walker first tile (n_walkers + 1)
    const double* inv_ng;     // per walker 1 / N_group
    double* direct;           // per walker [kMaxSlotCV] CV values
    const double* wcv;        // per walker [kWcvStride] chain-rule weights
    int32_t slot;
    int32_t n_walkers;
};
hipError_t launch_sf_value(const Geo& g, const SfDev& d, hipStream_t st);
hipError_t launch_sf_force(const Geo& g, const SfDev& d, double* f, hipStream_t st);

// STEINH L (switching-weighted local Q of degree l): per-atom state, per-tile sums
// and per-entry pair gradients.  Arrays of 2(l+1) doubles hold Re/Im for m = 0..l.
struct LqDev {
    int32_t l;
    int32_t mode;             // 0 MEAN_SOLID, 1 NSOLID, 2 FRAC
    SwFn swr, swcv;           // sigma(r) on list entries; f on LQ_l
    double eps;               // entries with sigma < eps do not count
    double* S;                // per local atom sum of sigma (0: no entry counts)
    double* q;                // per local atom q_lm
    double* LQ;               // per local atom LQ_lm (0 outside the group)
    double* ql;               // per local atom LQ_l
    double* K;                // per local atom W dCV/dLQ_lm / (S + 1)
    double* H;                // per local atom dCV/dq_lm (K plus the weighted K of its entries)
    double2* hk;              // per local atom (<H, q>, <K, LQ>)
    double* u;                // per ELL entry a -> j: <K_owner(j), q_a>
    const int64_t* apos;      // per local atom: ELL position of its first entry (tile base + lane);
                              // entry k at apos + 64 k
    double* g;                // per ELL entry: d(W CV)/d(pair vector), 3 doubles
    double2* es;              // per ELL entry: (sigma, dsigma/dr), sigma = 0 where the entry does not count
    const int64_t* erev;      // per ELL entry: ELL index of the reverse entry, -1 = none
    const int32_t* eidx;      // per ELL entry: the list index (all-space atom)
    const int32_t* eown;      // per ELL entry: its local owner
    double* part;             // per tile (sum q f, sum f)
    double* sums;             // per walker (sum f, CV)
    const int32_t* tile_begin;  // per walker first tile (n_walkers + 1)
    const double* ng;         // per walker N_group
    double* direct;           // per walker [kMaxSlotCV] CV values
    const double* wcv;        // per walker [kWcvStride] chain-rule weights
    int32_t slot;
    int32_t n_walkers;
};
// stage 0 k_lq_q, 1 k_lq_lq, 2 k_lq_sum (values); 3 k_lq_k, 4 k_lq_h, 5 k_lq_pair, 6 k_lq_force (forces)
hipError_t launch_lq(int stage, const Geo& g, const LqDev& d, double* f, hipStream_t st);
hipError_t launch_lq_erev(const Geo& g, int64_t n, const int32_t* reown, const int32_t* rrev, const int64_t* apos,
                          int64_t* erev, int32_t* eidx, int32_t* eown, hipStream_t st);

// DISTANCE CV between two atoms of each walker (one thread per walker).
struct DistDev {
    int32_t n_walkers;
    int32_t a1, a2;           // local indices within a walker
    const int64_t* atom_off;  // per walker first atom in the concatenated arrays
    const double* box;        // per walker box lengths (3)
    double* dvec;             // per walker (dx, dy, dz, r) of the last value pass
    double* direct;           // per walker [kMaxSlotCV] CV values
    const double* wcv;        // per walker [kWcvStride] chain-rule weights
    int32_t slot;
};
hipError_t launch_dist_value(const double4* x4, const DistDev& d, hipStream_t st);
hipError_t launch_dist_force(const DistDev& d, double* f, hipStream_t st);

// Periodic box and bins of a device neighbour build (nlist.hip).
struct NlBox {
    double lo[3], hi[3], L[3];
    double rg;                // list cutoff
    double elo[3], bw[3];     // extended box lower corner (lo - rg) and bin widths
    int32_t nb[3];            // bins per dimension
};
hipError_t nl_ghost_flags(const double* x, int64_t n, const NlBox& b, int32_t* flag, hipStream_t st);
hipError_t nl_ghost_scatter(const double* x, int64_t n, const NlBox& b, const int32_t* flag, const int32_t* pos,
                            double* xall, int32_t* owner, hipStream_t st);
hipError_t nl_bin(const double* xall, int64_t na, const NlBox& b, int32_t* bin_of, int32_t* val,
                  int32_t* bin_count, hipStream_t st);
hipError_t nl_list_count(const double* xall, int64_t n, const NlBox& b, const int32_t* bin_of,
                         const int32_t* bin_start, const int32_t* bin_atoms, int64_t* count, hipStream_t st);
hipError_t nl_list_fill(const double* xall, int64_t n, const NlBox& b, const int32_t* bin_of,
                        const int32_t* bin_start, const int32_t* bin_atoms, const int64_t* off, int32_t* idx,
                        int32_t* off32, hipStream_t st);
hipError_t nl_scan_i32(void* tmp, size_t& bytes, const int32_t* in, int32_t* out, int64_t n, hipStream_t st);
hipError_t nl_scan_i64(void* tmp, size_t& bytes, const int64_t* in, int64_t* out, int64_t n, hipStream_t st);
hipError_t nl_sort_pairs(void* tmp, size_t& bytes, const int32_t* kin, int32_t* kout, const int32_t* vin,
                         int32_t* vout, int64_t n, int end_bit, hipStream_t st);

// ---- launchers (kernels.hip) ----
hipError_t launch_prune(const Geo& g, const NearDev& nd, int force, hipStream_t st);
// owners and reverse positions of the near lists (built when the epoch's flag is up, or forced)
hipError_t launch_topology_near(const Geo& g, const NearDev& nd, int32_t* eown, uint8_t* revpos, int force,
                                hipStream_t st);
// owners and reverse positions of the bound lists (at most kMaxList entries each)
hipError_t launch_topology(const Geo& g, int64_t n, int32_t* eown, uint8_t* revpos, hipStream_t st);
// the topology of the bound lists with int32 reverse positions (-1 = none), for the
// kernels that run on the bound lists (STEINH L)
hipError_t launch_topology_raw(const Geo& g, int64_t n, int32_t* eown, int32_t* revpos, hipStream_t st);

hipError_t launch_pack_x4(const double* x, double4* x4, float4* x4f, int64_t n, hipStream_t st);
hipError_t launch_pack_walker(const double* x, int64_t n_local, int64_t n_all, double4* x4_loc, float4* x4f_loc,
                              double4* x4_gh, float4* x4f_gh, hipStream_t st);
// the same, comparing with the positions of the last near-list build (xpr at the same
// offsets as x4f) and raising the rebuild flag of the current epoch
hipError_t launch_pack_walker_near(const double* x, int64_t n_local, int64_t n_all, double4* x4_loc, float4* x4f_loc,
                                   double4* x4_gh, float4* x4f_gh, const float4* xpr_loc, const float4* xpr_gh,
                                   const NearDev& nd, hipStream_t st);
hipError_t launch_fill_u8(uint8_t* p, int64_t n, hipStream_t st);
hipError_t launch_env(const Geo& g, const EnvDev& e, hipStream_t st);
// the selection export instance of k_env (e.sel_out, e.sel_cnt set)
hipError_t launch_select_export(const Geo& g, const EnvDev& e, hipStream_t st);
hipError_t launch_ld(const Geo& g, const EnvDev& e, const LdDev& d, hipStream_t st);
hipError_t launch_ld_coeff(const Geo& g, const EnvDev& e, const LdDev& d, hipStream_t st);
hipError_t launch_force(const Geo& g, const EnvDev& e, double* f, hipStream_t st);
hipError_t launch_reduce(const ReduceDev& r, hipStream_t st);
// values of CVs whose partial sums are produced after the reduce (unbiased, fused into the force pass)
hipError_t launch_cv_tail(const ReduceDev& r, hipStream_t st);
hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t st);
// allocation-free virial: the per-device scratch of block partials (looked up once per
// device by scratch_setup, which must run with no graph capture active) and the
// fixed-order sum of a walker's partials into v[6]
constexpr int kScratchBlocks = 8192;
hipError_t scratch_setup(int device);
double* scratch_ptr(int device);
hipError_t launch_virial_sum(const double* part, int nb, double* v, hipStream_t st);
hipError_t launch_virial(const double4* x4, const double* f, const int32_t* image, int first, int n,
                         const double L[3], double* out, hipStream_t st);
hipError_t launch_fill_f(float* p, int64_t n, float v, hipStream_t st);
hipError_t launch_bias_eval(const ReduceDev& r, int walker, int64_t np, const double* s,
                            double* V, double* dV, hipStream_t st);
hipError_t launch_fes_sum(int D, int64_t nh, const double* hills, const double* x, int n0, const double* y,
                          int n1, double lo0, double lo1, double d0, double d1, double* out, hipStream_t st);
hipError_t launch_fes_marginal(const double* F, int n0, int n1, int axis, double kT, double bin, double* out,
                               hipStream_t st);
hipError_t launch_bias_add(const ReduceDev& r, int walker, int64_t nh, const double* s,
                           const double* h, double sigma, hipStream_t st);

// learned bias (mlp.hip)
// mode 0: fresh points and labels; 1: labels of the points in s raised by the hills since
// the last fit; 2: fresh points around the hills since the last fit (k_mlp_data)
hipError_t mlp_data(const MlpDev& m, int w, int64_t n, uint64_t seed, float4* s, double* v, int mode, hipStream_t st);
hipError_t mlp_box(const MlpDev& m, int w, hipStream_t st);
hipError_t mlp_stats(const MlpDev& m, int w, int64_t n, const double* v, hipStream_t st);
hipError_t mlp_train_step(const MlpDev& m, int w, int64_t first, int64_t count, int64_t n_total, const float4* s,
                          const double* v, float* gpart, float lr, float* gsum, int64_t step, int64_t nit,
                          hipStream_t st);
hipError_t mlp_forward(const MlpNet* net, int64_t n, const float4* s, const double* vref, double* vout,
                       double* blk_err, hipStream_t st);
hipError_t mlp_commit(const MlpDev& m, int w, int nblk, const double* blk_err, double tol, hipStream_t st);
hipError_t mlp_copy_net(const MlpNet* src, MlpNet* dst, hipStream_t st);
hipError_t mlp_init(MlpNet* net, uint64_t seed, hipStream_t st);

// Structure factor
// Above this many k vectors the structure-factor passes over (atom, k) take
// tile-relative FP32 phases (SkDev.perm, tbeg, tref, dr): their per-step preparation
// (one pass over the atoms in Hilbert order, and the force written through the
// permutation) costs less than the FP64 phase work it saves.
constexpr int kSkTiledK = 256;
struct SkK {                  // per walker and k vector, rebuilt every step
    double kr[3];             // k / 2pi (revolutions per length)
    double phi;               // arg A(k) / 2pi (revolutions)
    float kv[3];              // k
    float amod;               // |A(k)|
    float krf[3];             // k / 2pi rounded to FP32 (the tile-relative phase)
    float pad;
};
struct SkDev {
    int32_t nk;
    const double* k;          // nk * 3 (k / 2pi, revolutions per length)
    const double* kvec;       // nk * 3 (k itself)
    int32_t nchunk;
    const int32_t* chunk;     // per chunk: start, count, walker (3 ints), in slots
    // slot -> local atom index: each walker's atoms ordered along a Hilbert curve, so
    // that every run of 256 slots (a phase tile) is a compact set of atoms
    const int32_t* perm;
    const int32_t* tbeg;      // per walker: its first phase tile (n_walkers + 1)
    double4* tref;            // per phase tile: the FP64 position of its first atom (this step)
    float4* dr;               // per slot: FP32 offset of the atom from its tile's reference;
                              // w = 1 when the atom adds to the sums (its group), else 0
    int32_t wide;             // tile phases may exceed 128 revolutions: reduce them first
    double* apart;            // [chunk][nk][2]
    double* amp;              // per walker [nk][2]
    SkK* kpack;               // per walker [nk]
    float* ampf;              // per walker [nk][2] scaled for the gradient
    double* sterm;            // per walker [nk] |A(k)|^2 / N
    const int32_t* chunk_begin;  // per walker first chunk (n_walkers + 1)
    double* direct;           // per walker [kMaxSlotCV]
    int32_t slot;             // direct slot of SK.AVE
    int32_t wslot;            // wcv slot
    const double* wcv;
    const int32_t* n_atoms;   // per walker
    const int32_t* atom_begin;// per walker
    const double* n_norm;     // per walker N of S = |A|^2 / N (NULL: n_atoms)
    int32_t n_walkers;
    int32_t max_atoms;
};
// per-step slot data of the structure-factor passes (tile references and offsets), then
// the amplitudes
hipError_t launch_sk_amp(const Geo& g, const SkDev& s, hipStream_t st);
hipError_t launch_sk_force(const Geo& g, const SkDev& s, double* f, hipStream_t st);
// the gradient records and S terms of amplitudes written into s.amp (reduced over contexts)
hipError_t launch_sk_pack(const SkDev& s, hipStream_t st);
// raw FP64 sums of the tile partials of every CV slot with psum[c] >= 0 (a partial slot),
// per walker: out[w * kMaxSlotCV + c], in a fixed order
hipError_t launch_cv_sums(const ReduceDev& r, const int32_t* psum, double* out, hipStream_t st);

}  // namespace mdt
