/*
 * mdt.h - C ABI of the MI355X-native metadynamics CV-and-bias path.
 *
 * Drop-in boundary for the Metad-toy LAMMPS plugin's hot path.  The plugin's
 * FixMetadynamics / CV / Gaussian_Hill_Base C++ interfaces are replaced by
 * this plain-pointer C ABI; a LAMMPS-side fix (or any MD engine) binds
 * positions and full neighbour lists, hands over the `fix metad` keyword list
 * unchanged, and calls mdt_fix_post_force once per step.  Positions,
 * neighbour lists, selections, per-atom order parameters, bias grids and
 * forces stay in device memory; per step the host sees only a few doubles.
 *
 * Conventions
 *   - Every function returning int returns MDT_OK (0) or a negative status;
 *     the message is available from mdt_last_error(ctx).  Nothing here calls
 *     exit(), MPI or any collective.
 *   - One HIP stream per context.  A context is not thread-safe.
 *   - Positions are FP64, 3 per atom, locals first then ghost images (the
 *     LAMMPS Atom::x layout).  Neighbour lists are CSR over local atoms with
 *     entries indexing [0, n_all); list order is preserved (it defines
 *     N-nearest tie-breaking).
 *   - "walker" indexes an independent system; a context holds n_walkers of
 *     them, all advanced together by one batched launch sequence per step.
 *
 * Reference paths below are relative to the Metad-toy source tree.
 */
#ifndef MDT_H
#define MDT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDT_OK 0
#define MDT_EINVAL (-22)   /* bad argument or bad fix keyword list        */
#define MDT_ENOMEM (-12)   /* device or host allocation failed            */
#define MDT_EHIP (-5)      /* a HIP runtime call failed                   */
#define MDT_ENOTSUP (-95)  /* a feature the reference has but this build
                              does not provide (message names it)        */
#define MDT_ESTATE (-71)   /* call out of order (e.g. step before bind)   */

typedef struct mdt_ctx mdt_ctx;
typedef struct mdt_fix mdt_fix;
typedef struct mdt_system mdt_system;

/* ------------------------------------------------------------------------
 * Context
 * ---------------------------------------------------------------------- */

/* Create a context on HIP device `device` holding `n_walkers` independent
 * systems.  Replaces the per-rank GPU state the reference keeps in CV/env
 * objects (src/cvs/CVs_Steinhardt.cu:189-228, include/utils/zqc_CVs_tools.h:33-229). */
int mdt_create(mdt_ctx** ctx, int device, int n_walkers);
void mdt_destroy(mdt_ctx* ctx);
const char* mdt_last_error(const mdt_ctx* ctx);
/* Library build tag ("mdt <version> gfx950"). */
const char* mdt_version(void);

/* ------------------------------------------------------------------------
 * Inputs from the MD engine (replaces the host flattening in
 * Steinhardt_env::refresh_lmpbox/get_env, src/cvs/CVs_Steinhardt.cu:391-566,
 * and the NeighHub list hand-off, src/core/neigh_hub.cu:191-231).
 * ---------------------------------------------------------------------- */

/* Orthorhombic periodic box of one walker.  Each edge must exceed twice the
 * list cutoff so an atom sees at most one image of another inside it. */
int mdt_set_box(mdt_ctx* ctx, int walker, const double lo[3], const double hi[3],
                const int periodic[3]);

/* Positions x[n_all*3] (locals then ghosts).  on_device != 0 means `x` is a
 * device pointer.  The data are copied; call again (or mdt_update_positions)
 * whenever they change. */
int mdt_bind_positions(mdt_ctx* ctx, int walker, const double* x, int64_t n_local,
                       int64_t n_all, int on_device);

/* Full neighbour list of every local atom: offsets[n_center+1], idx[offsets[n]],
 * entries in [0, n_all); ghost_owner[n_all-n_local] maps each ghost image to
 * its local atom (LAMMPS: atom->map(tag[j])).  LAMMPS NEIGHMASK bits must
 * already be stripped.  Lists may be of any length: when one is longer than 128
 * entries, the N-nearest selection walks near lists cut just above its largest
 * cutoff (list order kept, rebuilt on the device when atoms move), and at most
 * 128 entries may lie within that cutoff. */
int mdt_bind_neighbors(mdt_ctx* ctx, int walker, const int32_t* offsets, const int32_t* idx,
                       int64_t n_center, const int32_t* ghost_owner, int on_device);

/* Optional group membership of local atoms (1 = in group; default: all).
 * Replaces the mask/groupbit compaction of refresh_lmpbox (:391-425). */
int mdt_bind_group(mdt_ctx* ctx, int walker, const uint8_t* in_group);

/* Membership of a named group (the group word of a CAL: STEINH ... <group>,
 * STRU_FACTOR <group>), 1 = in group, for local atoms; in_group == NULL unbinds
 * it.  CALs may name different groups, each averaged over its own atoms, as
 * the reference pools environments per group (Steinhardt_env::get_or_create,
 * src/cvs/CVs_Steinhardt.cu:191-220, over LAMMPS atom->mask group bits).  A
 * group name without a binding of its own uses the mdt_bind_group membership. */
int mdt_bind_group_named(mdt_ctx* ctx, int walker, const char* name, const uint8_t* in_group);

/* Atom types of a walker's local atoms (LAMMPS atom->type, 1-based), for the
 * chemically weighted CVs: STRU_FACTOR ... Chemical (Stru_fact_chem_env,
 * src/cvs/CVs_Strufactor_chem.cu:29-66,95-123) and WEIGHT_CHEM
 * (src/cvs/CV_Weighted_Chemical_Pair_Order.cu:32-248).  Ghost images take their
 * owner's type.  type == NULL unbinds them. */
int mdt_bind_types(mdt_ctx* ctx, int walker, const int32_t* type);

/* This context's share of a walker's atoms (1 = mine): the atoms whose terms it adds
 * to the CV sums, as one MPI rank adds its local atoms; every atom still takes part
 * (neighbours, q_lm) and receives forces.  mine == NULL: every atom (the default). */
int mdt_bind_share(mdt_ctx* ctx, int walker, const uint8_t* mine);

/* Convenience: box, positions and lists of a host-built mdt_system. */
int mdt_bind_system(mdt_ctx* ctx, int walker, const mdt_system* sys);

/* New positions for an already bound walker (same n_all, same lists).  A device
 * buffer is packed straight into the context's layout (one kernel, no copy). */
int mdt_update_positions(mdt_ctx* ctx, int walker, const double* x, int on_device);

/* Read a walker's positions from the MD engine's device buffer x (n_all*3
 * doubles, locals then ghosts, the layout bound last) at the start of every step,
 * inside the step's launches and graphs: the per-step upload of
 * Steinhardt_env::get_env (src/cvs/CVs_Steinhardt.cu:552-566) without a copy or a
 * host call.  The buffer must stay valid and keep its layout until the next
 * attach; x == NULL detaches. */
int mdt_attach_positions(mdt_ctx* ctx, int walker, const double* x);

/* ------------------------------------------------------------------------
 * fix metad
 * ---------------------------------------------------------------------- */

/* Parse the `fix ID group metad ...` keyword list (argv = the words after
 * "metad") and build CVs, symbols, dimensions and the bias for every walker.
 * Keywords: GAUSSIAN PACE RECORD CV_dim TEMP CAL SYMBOL|SIMBOL DIM
 * Gaussian_Hill_type METAD_RESTART WT, plus BOLTZ <kB> and HILLS_FILE <path>.
 * Replaces FixMetadynamics::FixMetadynamics (src/core/fix_crystallize.cu:32-302),
 * CVFactory::create (src/core/CVFactory.cu:8-30) and Steinhardt::create
 * (src/cvs/CVs_Steinhardt.cu:33-161). */
int mdt_fix_create(mdt_ctx* ctx, int argc, const char* const* argv, mdt_fix** fix);
void mdt_fix_destroy(mdt_fix* fix);

/* Parsed value of a fix keyword, for tools that read a `fix metad` line without
 * running it (e.g. on a parse-only context, device -1): key "GAUSSIAN" gives
 * (sigma, height, biasf); "PACE", "CV_dim", "TEMP", "BOLTZ", "WT",
 * "METAD_RESTART", "Gaussian_Hill_type" one value; "SIGMAS" one per dimension;
 * "DIM" (lower, upper, bins) per dimension; "<CAL>.KVEC" the k vectors of an
 * SK CAL, (kx, ky, kz) each.  *n receives the count; at most cap
 * values are written to out.  Replaces the keyword scan of the post-processing
 * script tools/plot_hills.py:41-119 with the fix's own parser
 * (FixMetadynamics::FixMetadynamics, src/core/fix_crystallize.cu:32-302). */
int mdt_fix_param(mdt_fix* fix, const char* key, double* out, int cap, int* n);

/* First-run set-up: allocate the bias, open/replay HILLS (METAD_RESTART 1).
 * Replaces FixMetadynamics::init (:343-378), GH_t0_uniformGrid::init_set_mode
 * and init_hills (src/grids/GH_t0_uniformGrid.cu:73-192).  Called implicitly
 * by the first step if omitted. */
int mdt_fix_init(mdt_fix* fix);

/* One MD step for every walker: CVs, hill deposit on step % PACE == 0 (before
 * the force, as the reference does), bias value/gradient, chain rule and
 * f += F_bias for local atoms.  `f` holds n_local*3 doubles per walker,
 * walkers concatenated in order; on_device selects a device or host pointer;
 * f == NULL accumulates into the context's own device force buffer.
 * Replaces FixMetadynamics::post_force (src/core/fix_crystallize.cu:386-472). */
int mdt_fix_post_force(mdt_fix* fix, int64_t step, double* f, int on_device);

/* One step split at the global reduction, for callers that spread a system over
 * several contexts (MPI ranks, each holding a share of the atoms: mdt_bind_share).
 * mdt_fix_partials runs the step up to the CV sums and writes, per walker, the
 * context's partial sums: per CAL in order, STEINH Q/LD (sum of q_l over the share's
 * group atoms, their count) and SK (Re A(k), Im A(k) per k vector, atom count);
 * *n = their total (out == NULL: only the count).  The caller sums these vectors
 * element-wise over the contexts (MPI_Allreduce) and hands the sums to
 * mdt_fix_resume, which finishes the step: CV values, the hill deposit, bias and
 * chain rule, and f += the forces from this context's share of the CVs (f as in
 * mdt_fix_post_force; over all contexts these add up to the forces of the whole
 * system).  Replaces the per-CV MPI_Allreduce of STEIN_QL::compute_cv_AVE
 * (src/cvs/CVs_Steinhardt.cu:824-839) and the bias broadcast of
 * FixMetadynamics::post_force (src/core/fix_crystallize.cu:437-442).  A fix that
 * calls mdt_fix_partials steps only this way from then on; DISTANCE, STRU_FACTOR
 * and STEINH L CALs are refused in it. */
int mdt_fix_partials(mdt_fix* fix, int64_t step, double* out, int64_t cap, int64_t* n);
int mdt_fix_resume(mdt_fix* fix, int64_t step, const double* reduced, double* f, int on_device);

/* `nsteps` consecutive steps step0..step0+nsteps-1 into the context's device
 * force buffer with no host synchronisation (standalone/batched driver); every
 * step is the whole path of mdt_fix_post_force.  Results are complete after
 * mdt_fix_sync.  mdt_fix_run_steps replays graphs of `steps_per_launch` (1..64)
 * captured steps, or with 0 launches each step's kernels one by one (no graph);
 * mdt_fix_run uses 8.  With a learned bias the launches stop at
 * every refit step.  No reference counterpart: the reference has no driver
 * outside the LAMMPS step loop. */
int mdt_fix_run(mdt_fix* fix, int64_t step0, int64_t nsteps);
int mdt_fix_run_steps(mdt_fix* fix, int64_t step0, int64_t nsteps, int steps_per_launch);
int mdt_fix_sync(mdt_fix* fix);

/* Run `nsteps` steps with plain launches and a HIP event pair around every
 * kernel on the context stream; ms[k] = mean duration of kernel k (names
 * comma-separated in `names`).  Measurement aid with no reference
 * counterpart (the reference times nothing below the LAMMPS Modify total). */
int mdt_fix_profile(mdt_fix* fix, int64_t step0, int64_t nsteps, char* names, int names_cap,
                    double* ms, int ms_cap, int* nkern);

/* Bias energy V_b of a walker at the last step (compute_scalar, :484-487). */
int mdt_fix_scalar(mdt_fix* fix, int walker, double* value);
/* Dimension value s_n (compute_vector, :489-494). */
int mdt_fix_vector(mdt_fix* fix, int walker, int n, double* value);
/* Value of a CV function, e.g. "Q6.AVE" (STEIN_QL::compute_cv_AVE,
 * src/cvs/CVs_Steinhardt.cu:824-839), at the last step. */
int mdt_fix_cv_value(mdt_fix* fix, int walker, const char* cal_func, double* value);
/* dV/dv for a CV function at the last step (the weight the chain rule gives
 * STEIN_QL::bias_force_AVE, src/grids/DimSet.cu:83-97). */
int mdt_fix_cv_weight(mdt_fix* fix, int walker, const char* cal_func, double* value);

/* Per-atom export "CAL.prop" for n_local atoms of a walker (Fix::extract
 * "colvar_peratom:<CAL>.<prop>", :592-613; props: stein_q, local_q, nsel,
 * count_rc; qlm (2l+1 per atom); kept: the neighbours the last step selected,
 * N per atom in list order, -1 after the last, decoded from the step's own
 * selection mask). */
int mdt_fix_extract_peratom(mdt_fix* fix, int walker, const char* key, double* out);

/* N-nearest selection of a CAL's environment at the current positions, in the
 * reference slot order (get_environment_Steinhardt_Q,
 * src/cvs/CVs_Steinhardt.cu:1324-1406): sel_idx[n_local*N] (-1 = empty),
 * count_rc[n_local], n_sel[n_local].  Computed by an instance of the step's
 * own selection kernel that runs the insertion for every atom and writes its
 * slots.  Any pointer may be NULL. */
int mdt_fix_selection(mdt_fix* fix, int walker, const char* cal_name, int32_t* sel_idx,
                      int32_t* count_rc, int32_t* n_sel);

/* Copy the context's device force buffer of a walker to host (n_local*3). */
int mdt_fix_get_forces(mdt_fix* fix, int walker, double* f_host);
/* Virial of the bias forces held in the context's device force buffer
 * (f == NULL or host-f steps), sum_i F_i (x) unwrap(x_i) over a walker's local
 * atoms as v[6] = xx yy zz xy xz yz, the tally of FixMetadynamics::post_force
 * (src/core/fix_crystallize.cu:453-469).  image: LAMMPS image flags, 3 ints
 * per local atom, NULL = all zero. */
int mdt_fix_virial(mdt_fix* fix, int walker, const int32_t* image, double v[6]);
/* The virial of mdt_fix_virial with nothing allocated: no reference counterpart;
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  image: device pointer to
 * 3 ints per local atom, or NULL; v: device pointer to 6 doubles, written on the
 * context's stream (complete after mdt_fix_sync or any later synchronising call),
 * the same bits as mdt_fix_virial.  Block partial sums go to a fixed per-device
 * scratch array, so a walker may hold at most 8192 * 256 = 2,097,152 local atoms
 * (larger: MDT_EINVAL), and calls that overlap on one device race. */
int mdt_fix_virial_noalloc(mdt_fix* fix, int walker, const int32_t* image, double* v);
/* Zero the context's device force buffer. */
int mdt_fix_zero_forces(mdt_fix* fix);

/* Bias of a walker evaluated at np points s[np*D] on the device:
 * V[np] and dV/ds[np*D] (either may be NULL).  Grid: Catmull-Rom as
 * get_bias_energy/get_dVdcv (src/grids/GH_t0_uniformGrid.cu:312-461);
 * hills sum: exact FP64 sum of Gaussians. */
int mdt_fix_bias_eval(mdt_fix* fix, int walker, int64_t np, const double* s, double* V,
                      double* dVds);
/* Deposit nh hills (centres s[nh*D], heights h[nh], isotropic sigma) into a
 * walker's bias as add_to_grid does (:226-282). */
int mdt_fix_bias_add_hills(mdt_fix* fix, int walker, int64_t nh, const double* s,
                           const double* h, double sigma);
/* Copy a walker's dense grid (row-major, last dimension fastest). */
int mdt_fix_grid(mdt_fix* fix, int walker, double* out, int64_t n);
/* Number of hills deposited so far by a walker and a copy of them:
 * rec[nh*(D+3)] = (step, s..., height, sigma) per hill. */
int mdt_fix_hills(mdt_fix* fix, int walker, int64_t* nh, double* rec, int64_t cap);

/* ------------------------------------------------------------------------
 * Learned bias (Gaussian_Hill_type 3): no reference counterpart.  The
 * network represents the hills deposited up to its last fit; hills deposited
 * since are summed exactly, and the exact FP64 sum replaces the network
 * whenever its held-out max |dV| exceeds MLP_TOL kT.
 * ---------------------------------------------------------------------- */

/* Fit the network of a walker to the hills deposited so far now
 * (asynchronous; runs by itself every MLP_REFIT steps). */
int mdt_fix_mlp_refit(mdt_fix* fix, int walker);
/* Held-out max |dV| of the last fit, whether the network is in use, number of
 * fits and of fallbacks, hills the last fit was trained on. */
int mdt_fix_mlp_status(mdt_fix* fix, int walker, double* err, int32_t* ok, int64_t* fits,
                       int64_t* fallbacks, int64_t* n_fit);
/* Parameters (live != 0: network in use; 0: last fitted) as 4481 floats
 * W1[64][3] b1[64] W2[64][64] b2[64] W3[64] b3, and norm[8] = input shift[3],
 * input scale[3], output scale, output shift. */
int mdt_fix_mlp_params(mdt_fix* fix, int walker, int live, float* params, double* norm);

/* Standalone network on a context's device (no fix needed): params and norm
 * as returned by mdt_fix_mlp_params.  mdt_mlp_forward: V at np points s[np*3]
 * on the BF16 MFMA inference kernel.  mdt_mlp_loss_grad: gradient of
 * L = (1/np) sum_p (y(x_p) - v_p / out_scale)^2 with respect to the 4481
 * parameters (y = network output before the output scale), one pass of the
 * MFMA training kernel without an update. */
int mdt_mlp_forward(mdt_ctx* ctx, const float* params, const double* norm, int64_t np, const double* s,
                    double* V);
int mdt_mlp_loss_grad(mdt_ctx* ctx, const float* params, const double* norm, int64_t np, const double* s,
                      const double* v, float* grad);

/* Ghost images and full neighbour lists built on a context's device, for n
 * local atoms x_local[n*3] (host, or device memory when on_device) in the
 * periodic box [lo, hi) with list cutoff rg.  Replaces the lists LAMMPS hands
 * the reference through NeighHub::ensure/list (src/core/neigh_hub.cu:191-231)
 * with the same data the host generator mdt_system_build produces, in the same
 * order: ghosts shift-major (sz, sy, sx over -1..1, sz outermost, zero shift
 * skipped), local atoms ascending within a shift; each list visits the 27-bin
 * stencil (dz outer, dx inner) and a bin's atoms in ascending index, keeping
 * FP64 r^2 <= rg^2.  The result stays on the device (n_all atoms, nnz list
 * entries): mdt_neighbor_fetch copies any of x_all[n_all*3],
 * offsets[n+1], idx[nnz], ghost_owner[n_all-n] to the host, and
 * mdt_neighbor_bind binds box, positions and lists to a walker as
 * mdt_set_box + mdt_bind_positions + mdt_bind_neighbors would. */
int mdt_neighbor_build(mdt_ctx* ctx, const double* x_local, int64_t n, const double lo[3], const double hi[3],
                       double rg, int on_device, int64_t* n_all, int64_t* nnz);
int mdt_neighbor_fetch(mdt_ctx* ctx, double* x_all, int32_t* offsets, int32_t* idx, int32_t* ghost_owner);
int mdt_neighbor_bind(mdt_ctx* ctx, int walker);

/* FES reconstruction from a HILLS list on a context's device.  Replaces the
 * hill sum of tools/plot_hills.py:158-223 (the reference's post-processing
 * script, numpy on the host).  hills: nh rows of (step, s_1..s_D, height,
 * sigma), the HILLS line layout (src/grids/GH_t0_uniformGrid.cu:210-220).
 * D = 1: out[n0] = sum over every hill of h exp(-(x - c)^2 / (2 sigma^2)) at
 * the points x[n0] (:158-172).  D = 2: out[n0 * n1] (row-major, x fastest
 * varying along the second index) sums each hill over the bins of its window
 * int((c -+ 4 sigma - lo) / d), clamped to the grid (:174-223); lo[2] and d[2]
 * are the grid's lower bounds and bin widths.  Hills are summed in list order
 * in FP64, as the script's numpy reduction does. */
int mdt_fes_sum(mdt_ctx* ctx, int D, int64_t nh, const double* hills, const double* x, int n0,
                const double* y, int n1, const double* lo, const double* d, double* out);
/* Marginal free energy of a 2-D FES F[n0 * n1] (tools/plot_hills.py:236-262):
 * out = -kT ln(sum_k exp(-F/kT) * bin) along axis (1: over the second CV,
 * giving n0 values F(CV1); 0: over the first, n1 values F(CV2)), shifted so
 * that its minimum is 0. */
int mdt_fes_marginal(mdt_ctx* ctx, const double* F, int n0, int n1, int axis, double kT, double bin,
                     double* out);

/* ------------------------------------------------------------------------
 * Synthetic systems (host; the standalone driver's stand-in for LAMMPS
 * create_atoms + neighbor build)
 * ---------------------------------------------------------------------- */
int mdt_system_create(mdt_system** out);
void mdt_system_destroy(mdt_system* s);
const char* mdt_system_error(const mdt_system* s);
int mdt_system_set_box(mdt_system* s, const double lo[3], const double hi[3]);
void mdt_system_get_box(const mdt_system* s, double lo[3], double hi[3]);
int mdt_system_add_fcc(mdt_system* s, int nx, int ny, int nz, double a, double sigma_below,
                       double sigma_above, double z_split, uint64_t seed);
int mdt_system_add_uniform(mdt_system* s, int64_t n, double z0, double z1, uint64_t seed);
int mdt_system_set_positions(mdt_system* s, const double* x, int64_t n);
int mdt_system_build(mdt_system* s, double list_cutoff);
int64_t mdt_system_n_local(const mdt_system* s);
int64_t mdt_system_n_all(const mdt_system* s);
int64_t mdt_system_nnz(const mdt_system* s);
int mdt_system_copy(const mdt_system* s, double* x_all, int32_t* offsets, int32_t* idx,
                    int32_t* ghost_owner);

#ifdef __cplusplus
}
#endif

#endif /* MDT_H */
