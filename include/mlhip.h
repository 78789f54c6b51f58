/*
 * mlhip.h — C ABI of libmlhip.so, the MI355X (gfx950) implementation of the
 * single-level implicit solve of SOFTBALL (jrgparkinson/mushy-layer):
 *   - the nonlinear FAS multigrid for enthalpy H and bulk concentration
 *     Theta of AMRNonLinearMultiCompOp, with the phase-diagram closure fused
 *     into red/black Gauss-Seidel smoothers;
 *   - the variable-coefficient Helmholtz / Poisson operators behind
 *     DarcyBrinkmanOp and AMRProjectionOp.
 *
 * Each operator entry point replaces one method of the reference's Chombo
 * AMRLevelOp<LevelData<FArrayBox>> implementation; the citation names the
 * reference method (file:line).  All field arguments are DEVICE pointers in
 * Chombo FAB layout (one ghost cell, component outermost):
 *     idx(i,j,k,c) = (i+1) + (n0+2)*((j+1) + (n1+2)*(k+gz)) + c*ncell
 * gz = 1 in 3D and 0 in 2D; ncell = ml_cell_count().  Face coefficients of
 * direction d live on surroundingNodes(valid box, d) without ghosts
 * (ml_face_count()), component outermost.  Only single-box levels whose box
 * starts at the origin are accepted (lo == 0): red/black colouring uses the
 * absolute parity of i+j+k, as the reference's Fortran does.
 *
 * Return codes: 0 success; negative values are errors (ML_E*), with a
 * message from ml_last_error().  Whole solves report the multigrid exit
 * status separately (ml_solve_stats.exit_status, bit set 1/2/4/8, as the
 * reference inspects at srcSubcycle/AMRLevelMushyLayer.cpp:715-749).
 *
 * Threading: a context owns one HIP stream; calls on a context are
 * asynchronous on that stream and must not be issued concurrently from
 * several host threads.  ml_buf_download and ml_ctx_synchronize block.
 */
#ifndef MLHIP_H
#define MLHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MLHIP_ABI_VERSION 1

enum {
  ML_OK = 0,
  ML_EINVAL = -1,   /* invalid argument */
  ML_EHIP = -2,     /* HIP runtime error */
  ML_ENOTSUP = -3,  /* unsupported configuration */
  ML_ENOMEM = -4    /* device allocation failed */
};

enum { ML_BC_DIRICHLET = 0, ML_BC_NEUMANN = 1 };
enum { ML_PRECISION_FP64 = 0, ML_PRECISION_FP32_REFINED = 1 };

typedef struct ml_ctx_s* ml_ctx;
typedef struct ml_hc_op_s* ml_hc_op;
typedef struct ml_vc_op_s* ml_vc_op;

/* Physical parameters (MushyLayerParams, src/MushyLayerParams.h;
 * nondimensional defaults src/MushyLayerParams.cpp:266-270,:340-365,:476-484,
 * :736-753). */
typedef struct {
  double stefan;
  double compositionRatio;
  double waterDistributionCoeff;
  double specificHeatRatio;
  double heatConductivityRatio;
  double thetaEutectic;
  double ThetaEutectic;
  double heatDiffusionCoeff;   /* D_H */
  double saltDiffusionCoeff;   /* D_Theta = 1/Le */
} ml_params;

/* Boundary condition per [dir][side lo=0/hi=1][component]: type
 * ML_BC_DIRICHLET (order 1, BCutil/BCFuncF.ChF:14-56) or ML_BC_NEUMANN
 * (BCutil/BCFuncF.ChF:63-96); ignored in periodic directions.  For the HC
 * operator it is the (theta, Theta_l) derived-variable BC
 * (PhysBCUtil::temperatureLiquidSalinityBC, BCutil/PhysBCUtil.cpp:3741-3787)
 * and `apply` is m_apply_bcs_to_diagnostic_var
 * (src/AMRNonLinearMultiCompOp.cpp:92-96). */
typedef struct {
  int type[3][2][2];
  double val[3][2][2];
  int apply;
} ml_bc;

/* Multigrid options (defaults srcSubcycle/MushyLayerSubcycleUtils.cpp:664-676:
 * nu1 = 1, nu2 = 4, bottom = 40, max_iter = 10, eps = hang = norm_thresh =
 * 1e-10, precision FP64).  flags: ML_SOLVE_HOST_DRIVER runs the solve through
 * the per-operator entry points with the host C++ driver (host/FASMultiGrid.H)
 * instead of the fused engine; both give identical FP64 results.
 * ML_SOLVE_NO_LOOKAHEAD enqueues one iteration at a time (the host reads the
 * stopping rule after each) instead of batches that the device cuts short:
 * same results, no launches that return at once (for per-launch
 * measurements).  ML_SOLVE_NO_GRAPH launches every kernel from the host, each
 * step of an iteration as its own kernel, instead of replaying the
 * iterations captured as a HIP graph with the residual fused into the last
 * relaxation: same results.
 * bottom_solver (linear operator solves only): ML_BOTTOM_RELAX relaxes the
 * coarsest MG level `bottom` times; ML_BOTTOM_BICGSTAB follows those
 * relaxations with BiCGStab of at most bottom_max_iter iterations, the
 * reference's bottom solver of the projection (Projector::makeBottomSolvers,
 * srcSubcycle/Projector.cpp:3004-3023, m_imax 20) and of the U* solves
 * (s_botSolverUStar, srcSubcycle/AMRLevelMushyLayerInit.cpp:739, m_imax 80).
 * BiCGStab needs a coarsest level of at most 1024 cells (it runs inside the
 * single-workgroup V-cycle; ML_ENOTSUP otherwise).  Defaults: relax, 80. */
typedef struct {
  int nu1, nu2, bottom, max_iter;
  double eps, hang, norm_thresh;
  int precision;
  int flags;
  int bottom_solver, bottom_max_iter;
} ml_solve_opts;
enum { ML_BOTTOM_RELAX = 0, ML_BOTTOM_BICGSTAB = 1 };

enum { ML_SOLVE_HOST_DRIVER = 1, ML_SOLVE_NO_LOOKAHEAD = 2, ML_SOLVE_NO_GRAPH = 4 };

typedef struct {
  int exit_status;   /* 1 |r|<=eps|r0|, 2 iteration max, 4 hang, 8 |r|<=norm_thresh */
  int iterations;    /* V-cycles */
  double initial_residual;
  double final_residual;
} ml_solve_stats;

/* ---- context, buffers, layout ---------------------------------------- */
int ml_ctx_create(int device, unsigned flags, ml_ctx* out);
int ml_ctx_destroy(ml_ctx ctx);
const char* ml_last_error(ml_ctx ctx);
int ml_ctx_synchronize(ml_ctx ctx);
/* the HIP stream (hipStream_t) the context launches on */
void* ml_ctx_stream(ml_ctx ctx);
/* HIP events on the context stream (slots 0..15) for timing a sequence of
 * calls: record a slot, then read the milliseconds between two recorded
 * slots (blocks until slot_b has completed). */
int ml_ctx_event_record(ml_ctx ctx, int slot);
int ml_ctx_event_elapsed(ml_ctx ctx, int slot_a, int slot_b, float* ms);
/* Launch timing of one kernel family (measurement; no counterpart in the
 * reference).  While started, every launch of `kernel` on the context is
 * bracketed by HIP events on the context stream (the first max_launches
 * launches) and its algorithmic bytes (every input read once, every output
 * written once) are summed.  ml_ctx_timing_stop synchronises the stream,
 * returns the timed launches, their summed duration in ms and bytes, and the
 * launches beyond max_launches, and ends the timing. */
enum {
  ML_KT_VC_RELAX = 1, /* tiled red/black relaxation of the linear VC operator */
  ML_KT_HC_RELAX = 2, /* tiled red/black relaxation of the (H, Theta) operator with the closure */
  ML_KT_BATCH_RELAX = 3 /* batched form: FP32 block relaxation of the linearised (H, Theta) correction
                         * (every instance's level in one launch; launches are timed when the step runs
                         * with ML_SOLVE_NO_GRAPH) */
};
int ml_ctx_timing_start(ml_ctx ctx, int kernel, int max_launches);
int ml_ctx_timing_stop(ml_ctx ctx, long long* launches, double* total_ms, double* bytes, long long* missed);
int ml_buf_alloc(ml_ctx ctx, size_t bytes, void** dptr);
int ml_buf_free(ml_ctx ctx, void* dptr);
int ml_buf_upload(ml_ctx ctx, void* dst_dev, const void* src_host, size_t bytes);
int ml_buf_download(ml_ctx ctx, void* dst_host, const void* src_dev, size_t bytes);
int ml_buf_copy(ml_ctx ctx, void* dst_dev, const void* src_dev, size_t bytes);
size_t ml_cell_count(int dim, const int n[]);
size_t ml_face_count(int dim, const int n[], int dir);
/* number of MG levels: coarsen while every length divides by 2^(l+1)
 * (src/AMRNonLinearMultiCompOp.cpp:1316-1325 with s_maxCoarse = 2) */
int ml_mg_depth(int dim, const int n[]);
void ml_default_opts(ml_solve_opts* o);

/* ---- nonlinear (H, Theta) operator: AMRNonLinearMultiCompOp ------------ */

/* AMRNonLinearMultiCompOpFactory::define + MGnewOp for every depth
 * (src/AMRNonLinearMultiCompOp.cpp:1201-1274, :1285-1501). */
int ml_hc_op_create(ml_ctx ctx, int dim, const int lo[], const int hi[], double dx,
                    const int periodic[], const ml_params* params, const ml_bc* bc,
                    ml_hc_op* out);
int ml_hc_op_destroy(ml_hc_op op);
/* n and dx of MG level `depth` (0 = finest) */
int ml_hc_level_info(ml_hc_op op, int depth, int n_out[3], double* dx_out);
int ml_hc_num_levels(ml_hc_op op);

/* setCoefs (src/AMRNonLinearMultiCompOp.cpp:525-538) on the finest level and
 * the MG coarsening of aCoef/bCoef to every depth with arithmetic averaging
 * (MGnewOp, src/AMRNonLinearMultiCompOp.cpp:1423-1494).  aCoef may be NULL
 * (aCoef == 1).  Coefficients are copied. */
int ml_hc_set_coefs(ml_hc_op op, const double* aCoef, const double* const bCoef[],
                    double alpha, double beta);
/* setAlphaAndBeta (src/AMRNonLinearMultiCompOp.cpp:515-523) */
int ml_hc_set_alpha_beta(ml_hc_op op, double alpha, double beta);
/* diagonalScale (src/AMRNonLinearMultiCompOp.H:75-83): rhs *= aCoef of level
 * `depth` (identity when aCoef was NULL) */
int ml_hc_diagonal_scale(ml_hc_op op, int depth, double* rhs);

/* computeDiffusedVar (src/AMRNonLinearMultiCompOp.cpp:36-98): (theta, Theta_l)
 * on valid cells and wall ghosts of level `depth`. */
int ml_hc_compute_diffused_var(ml_hc_op op, int depth, const double* phi, double* out,
                               int homogeneous);
/* relax -> levelGSRB x iters (src/AMRNonLinearMultiCompOp.cpp:815-949) */
int ml_hc_relax(ml_hc_op op, int depth, double* phi, const double* rhs, int iters);
/* No counterpart in the reference: ml_hc_relax for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and same results as
 * ml_hc_relax, but the smoother's ping-pong partner array (2 components of
 * the level's cells, FP64) is taken from a static device array of
 * ML_NOALLOC_SCRATCH_BYTES per GPU instead of being allocated; no
 * allocating call is made.  A level whose state exceeds that size is
 * rejected (ML_ENOTSUP).  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (other contexts or host threads) race. */
#define ML_NOALLOC_SCRATCH_BYTES (8u << 20)
int ml_hc_relax_noalloc(ml_hc_op op, int depth, double* phi, const double* rhs, int iters);
/* residualI (src/AMRNonLinearMultiCompOp.cpp:100-166): res = rhs - L(phi) */
int ml_hc_residual(ml_hc_op op, int depth, double* res, double* phi, const double* rhs,
                   int homogeneous);
/* applyOpI (src/AMRNonLinearMultiCompOp.cpp:228-300): lhs = L(phi) */
int ml_hc_apply_op(ml_hc_op op, int depth, double* lhs, double* phi, int homogeneous);
/* restrictResidual (src/AMRNonLinearMultiCompOp.cpp:373-451):
 * resCoarse (level depth+1) = R(rhs - L(phi)) */
int ml_hc_restrict_residual(ml_hc_op op, int depth, double* resCoarse, double* phiFine,
                            const double* rhsFine);
/* restrictR (src/AMRNonLinearMultiCompOp.cpp:492-514): phiCoarse = R phiFine */
int ml_hc_restrict_r(ml_hc_op op, int depth, double* phiCoarse, const double* phiFine);
/* prolongIncrement (src/AMRNonLinearMultiCompOp.cpp:454-482):
 * phiFine += P corrCoarse (linear, ratio 2) */
int ml_hc_prolong_increment(ml_hc_op op, int depth, double* phiFine, const double* corrCoarse);

/* levelJacobi (src/AMRNonLinearMultiCompOp.cpp:1077-1101), `iters` sweeps
 * (relax mode 4): res = rhs - L(phi); phi += 0.5 * lambda * res */
int ml_hc_jacobi(ml_hc_op op, int depth, double* phi, const double* rhs, int iters);
/* preCond (src/AMRNonLinearMultiCompOp.cpp:172-206): phi = lambda * rhs, then
 * two levelGSRB iterations */
int ml_hc_precond(ml_hc_op op, int depth, double* phi, const double* rhs);
/* getFlux (src/AMRNonLinearMultiCompOp.cpp:1103-1162) on every face of level
 * `depth`: flux[d] (2 components, face layout) = -bCoef * beta *
 * (phi_l(i) - phi_l(i - e_d)) / dx, phi_l the diffused variable with the
 * derived-variable BC.  (reflux, :601-693, needs a finer AMR level; a
 * single-level solve has none.) */
int ml_hc_get_flux(ml_hc_op op, int depth, const double* phi, double* const flux[]);

/* AMRFASMultiGrid::solve with this operator (Chombo driver; build-owned
 * specification).  phi is the initial guess and the result. */
int ml_hc_solve(ml_hc_op op, double* phi, const double* rhs, const ml_solve_opts* opts,
                ml_solve_stats* stats);
/* LevelBackwardEuler::updateSoln as driven by multiCompAdvectDiffuse
 * (srcSubcycle/AMRLevelMushyLayer.cpp:1645-1831): alpha = 1, beta = -dt,
 * rhs = aCoef*phi_old + dt*src (src may be NULL), solve from phi_new. */
int ml_hc_solve_be(ml_hc_op op, double* phi_new, const double* phi_old, const double* src,
                   double dt, const ml_solve_opts* opts, ml_solve_stats* stats);

/* ---- linear variable-coefficient operator: DarcyBrinkmanOp and
 * AMRProjectionOp --------------------------------------------------------
 * L u = alpha*aCoef*u - beta*cCoef*u - beta*div(bCoef grad u), one
 * component.  Darcy-Brinkman U* solve (one call per velocity component):
 * aCoef = 1, bCoef = -Pr, cCoef = (Pr/Da) chi/Pi, backward Euler alpha = 1,
 * beta = -dt (srcSubcycle/AMRLevelMushyLayerInit.cpp:579-747).  Projection:
 * cCoef = NULL, alpha = 0, beta = 1, bCoef = -S_face (srcSubcycle/
 * Projector.cpp:2850-2991).  The BC uses component 0 of `bc` and is always
 * applied; MG prolongation is piecewise constant. */
/* DarcyBrinkmanOpFactory::define + MGnewOp for every depth
 * (src/DarcyBrinkmanOp.cpp:801-1017) */
int ml_vc_op_create(ml_ctx ctx, int dim, const int lo[], const int hi[], double dx,
                    const int periodic[], const ml_bc* bc, ml_vc_op* out);
int ml_vc_op_destroy(ml_vc_op op);
int ml_vc_num_levels(ml_vc_op op);
/* DarcyBrinkmanOp::setCoefs (src/DarcyBrinkmanOp.cpp:250-263) plus the MG
 * coarsening of aCoef, bCoef, cCoef to every depth with arithmetic averaging
 * (src/DarcyBrinkmanOp.cpp:977-1008).  aCoef NULL: aCoef == 1; cCoef NULL:
 * no c term (the projection's VCAMRPoissonOp2 form).  Copied. */
int ml_vc_set_coefs(ml_vc_op op, const double* aCoef, const double* const bCoef[],
                    const double* cCoef, double alpha, double beta);
/* setAlphaAndBeta (src/DarcyBrinkmanOp.cpp:242-248) */
int ml_vc_set_alpha_beta(ml_vc_op op, double alpha, double beta);
/* relax -> levelGSRB x iters, homogeneous BCs (src/DarcyBrinkmanOp.cpp:503-576) */
int ml_vc_relax(ml_vc_op op, int depth, double* phi, const double* rhs, int iters);
/* residualI (src/DarcyBrinkmanOp.cpp:31-83): res = rhs - L(phi) */
int ml_vc_residual(ml_vc_op op, int depth, double* res, double* phi, const double* rhs,
                   int homogeneous);
/* applyOpI (src/DarcyBrinkmanOp.cpp:123-137): lhs = L(phi) */
int ml_vc_apply_op(ml_vc_op op, int depth, double* lhs, double* phi, int homogeneous);
/* preCond (src/DarcyBrinkmanOp.cpp:89-121): phi = lambda*rhs on the valid
 * cells, then two red/black iterations (homogeneous BCs) */
int ml_vc_precond(ml_vc_op op, int depth, double* phi, const double* rhs);
/* dotProduct (AMRPoissonOp, inherited by DarcyBrinkmanOp, src/DarcyBrinkmanOp.H:27):
 * *out = (x, y) over the valid cells of level `depth`, in the fixed order of
 * the BiCGStab bottom (64 interleaved partial sums, then a pairwise tree over
 * the smallest power of two of them covering min(cells, 64));
 * blocks until the value is on the host */
int ml_vc_dot(ml_vc_op op, int depth, const double* x, const double* y, double* out);
/* restrictResidual (src/DarcyBrinkmanOp.cpp:183-240), homogeneous BCs */
int ml_vc_restrict_residual(ml_vc_op op, int depth, double* resCoarse, double* phiFine,
                            const double* rhsFine);
/* prolongIncrement, piecewise constant (AMRProjectionOp::prolongIncrement,
 * src/AMRProjectionOp.cpp:67-98) */
int ml_vc_prolong_increment(ml_vc_op op, int depth, double* phiFine, const double* corrCoarse);
/* AMRMultiGrid::solve on one level with this operator (Chombo driver;
 * build-owned specification); FP64 only */
int ml_vc_solve(ml_vc_op op, double* phi, const double* rhs, const ml_solve_opts* opts,
                ml_solve_stats* stats);
/* backward-Euler U* step (srcSubcycle/AMRLevelMushyLayerVel.cpp:1326-1403):
 * alpha = 1, beta = -dt, rhs = aCoef*phi_old + dt*src (src may be NULL) */
int ml_vc_solve_be(ml_vc_op op, double* phi_new, const double* phi_old, const double* src,
                   double dt, const ml_solve_opts* opts, ml_solve_stats* stats);
/* n (<= ML_VC_MAX_SOLVES) independent backward-Euler solves with the same
 * coefficients, as the reference runs them one velocity component at a time
 * (UstarBE[comp]->updateSoln, srcSubcycle/AMRLevelMushyLayerVel.cpp:1368-1373);
 * here they run concurrently on separate streams (each after the work
 * already enqueued on the context, the context continuing after all of
 * them).  Solve i gives exactly the result of ml_vc_solve_be on
 * (phi_new[i], phi_old[i], src ? src[i] : NULL); stats[i] its statistics. */
enum { ML_VC_MAX_SOLVES = 4 };

/* ---- batched linear solves (no counterpart in the reference: its runs
 * advance one mushy layer per process) ----------------------------------
 * nops 2D operators (distinct, one context; every level at most 512x512
 * cells and at least one level above the single-workgroup coarse level)
 * with per_op (<= ML_VC_MAX_SOLVES) solves each: unit u is solve
 * u % per_op of ops[u / per_op].  One call iterates every unit together,
 * each launch covering all units (2K + 1 launches per iteration, K the
 * deepest tiled-level count), and gives every unit exactly the states and
 * statistics of ml_vc_solve_many / ml_vc_solve_be_many on its operator
 * (the U* solves of many Darcy-Brinkman instances, or their projections).
 * Needs nu2 >= 1 and 2*nu + 1 passes within one relaxation launch. */
typedef struct ml_vc_batch_s* ml_vc_batch;
int ml_vc_batch_create(ml_ctx ctx, int nops, const ml_vc_op ops[], int per_op, ml_vc_batch* out);
int ml_vc_batch_destroy(ml_vc_batch b);
/* phi[u] updated in place from its value, rhs[u]; stats[u] may be NULL */
int ml_vc_batch_solve(ml_vc_batch b, double* const phi[], const double* const rhs[], const ml_solve_opts* opts,
                      ml_solve_stats stats[]);
/* backward-Euler form: alpha = 1, beta = -dt on every operator, rhs_u =
 * phi_old[u] + dt*src[u] (src or src[u] may be NULL), solved from phi_new[u] */
int ml_vc_batch_solve_be(ml_vc_batch b, double* const phi_new[], const double* const phi_old[],
                         const double* const src[], double dt, const ml_solve_opts* opts, ml_solve_stats stats[]);
int ml_vc_solve_be_many(ml_vc_op op, int n, double* const phi_new[], const double* const phi_old[],
                        const double* const src[], double dt, const ml_solve_opts* opts,
                        ml_solve_stats* stats);
/* Coefficient set-up of n independent operators of one context and
 * dimension, equal to n calls of ml_vc_set_coefs / ml_hc_set_coefs
 * (src/DarcyBrinkmanOp.cpp:250-263,977-1008; src/AMRNonLinearMultiCompOp.H
 * :75-93 with the hierarchy of src/AMRNonLinearMultiCompOp.cpp:1423-1494)
 * but in three launches for all of them: the level-0 copies, then the
 * hierarchy averages (one thread, or one workgroup, per coarse item).
 * bCoef holds n*dim pointers (operator i's direction d at i*dim + d);
 * aCoef and cCoef are n pointers (entries may be NULL: the operator then
 * has no a, or no c, term) or NULL for all.  No counterpart in the
 * reference, which sets its one operator up at a time; the batched
 * Darcy-Brinkman steps use it. */
int ml_vc_set_coefs_many(ml_ctx ctx, int n, const ml_vc_op ops[], const double* const aCoef[],
                         const double* const bCoef[], const double* const cCoef[], const double alpha[],
                         const double beta[]);
int ml_hc_set_coefs_many(ml_ctx ctx, int n, const ml_hc_op ops[], const double* const aCoef[],
                         const double* const bCoef[], const double alpha[], const double beta[]);
/* U* Darcy coefficient on valid cells: cCoef = (chi / Pi(chi)) * darcy_coeff,
 * Kozeny-Carman Pi = chi^3/(1-chi)^2 (srcSubcycle/AMRLevelMushyLayerInit.cpp
 * :665-689, src/MushyLayerParams.cpp:1046-1049); darcy_coeff = Pr/Da */
int ml_vc_darcy_ccoef(ml_vc_op op, const double* chi, double darcy_coeff, double* cCoef);
/* bCoef[d] = scale * sqrt(|chi_lo chi_hi|) on every face (geometric face
 * porosity, src/CellToEdge2F.ChF:20-46); scale = -1 gives the projection's
 * -S_face with S = chi (srcSubcycle/AMRLevelMushyLayerInit.cpp:1141) */
int ml_vc_porosity_faces(ml_vc_op op, const double* chi, double scale, double* const bCoef[]);
/* ---- batched form (no counterpart in the reference: its runs advance one
 * mushy layer per process) ------------------------------------------------
 * `count` independent (H, Theta) instances, each an ml_hc_op of its own grid
 * size, parameters and BCs (2D; finest level at most 512x512 cells, aCoef
 * == 1), advance one implicit step together.  For every instance i the step
 * is exactly the single-instance sequence
 *   ml_hc_porosity(op_i, phi[i], chi) -> ml_hc_assemble_bcoef(op_i, chi, b)
 *   -> ml_hc_set_coefs(op_i, NULL, b, 1, -dt)
 *   -> ml_hc_solve_be(op_i, phi[i], copy of phi[i], NULL, dt, opts)
 * with opts.precision == ML_PRECISION_FP32_REFINED, and gives bit-identical
 * states and statistics; but every kernel launch covers all instances (2K + 2
 * launches per outer iteration for the whole batch, K the deepest tiled-level
 * count, replayed from a HIP graph unless ML_SOLVE_NO_GRAPH).  The operators
 * are borrowed and must outlive the batch; their coefficient arrays are
 * overwritten by each step.  Requires nu1 <= 3, nu2 <= 4 and max_iter <= 127. */
typedef struct ml_hc_batch_s* ml_hc_batch;
int ml_hc_batch_create(ml_ctx ctx, int count, const ml_hc_op ops[], ml_hc_batch* out);
int ml_hc_batch_destroy(ml_hc_batch b);
int ml_hc_batch_size(ml_hc_batch b);
/* Device memory of the batch's set-up: the bytes of its one device
 * allocation, and how many device allocations and fill launches set-up made
 * (one each: every buffer of the batch is carved out of that allocation and
 * zeroed by one fill). */
int ml_hc_batch_memory(ml_hc_batch b, size_t* bytes, int* allocs, int* fills);
/* Tile shape of the batched FP32 relaxation (results do not depend on it).
 * By default the first step times every compiled shape on the batch's own
 * launches and keeps the fastest (the autotuner); set_tile fixes shape
 * `variant` (-1: tune again at the next step).  ml_hc_batch_tile reports the
 * shape in use and, when tune_ms is not NULL, the tuning time of each shape
 * (up to max_shapes); it returns the number of shapes. */
int ml_hc_batch_set_tile(ml_hc_batch b, int variant);
int ml_hc_batch_tile(ml_hc_batch b, int* variant, int* tile_x, int* tile_y, int* threads, float* tune_ms,
                     int max_shapes);
/* phi[i]: instance i's state (device, FAB layout, 2 components), advanced
 * in place; stats[i] (may be NULL) its solve statistics */
int ml_hc_batch_step(ml_hc_batch b, double* const phi[], double dt, const ml_solve_opts* opts,
                     ml_solve_stats stats[]);
/* the same step with a source per instance (src[i] or src may be NULL): the
 * single-instance sequence with ml_hc_solve_be(op_i, phi[i], copy, src[i],
 * dt, opts), e.g. the advective source of a Darcy-Brinkman step */
int ml_hc_batch_step_src(ml_hc_batch b, double* const phi[], const double* const src[], double dt,
                         const ml_solve_opts* opts, ml_solve_stats stats[]);

/* ---- adjoint of the implicit step (no counterpart in the reference) ----
 * Differentiable time stepping of the no-flow (H, Theta) problem with a
 * bottom heat-flux profile F (one value per bottom face, n[0] values; 2D,
 * conductivity ratio 1, aCoef == 1).  The flux enters the forward step as
 * the source that is exactly the Neumann ghost theta_g = theta_1 - dx F
 * (BCutil/BCFuncF.ChF:80-88) on theta, the bottom BC of theta itself being
 * homogeneous Neumann:
 *   ml_hc_flux_source: src = 0 except src_H = b_H,bottom face F / dx in the
 *   bottom cells, from the operator's current level-0 bCoef; then
 *   ml_hc_solve_be(op, phi, phi_old, src, dt, ...) is the forward step.
 * ml_hc_adjoint_step: given g_new = dL/dphi^n of a loss L of the state phi^n
 * reached by the forward step from phi^{n-1} = phi_old (with bCoef from the
 * porosity of phi_old, as the step assembled it), writes g_old = dL/dphi^{n-1}
 * through this step (valid cells; ghosts untouched), adds dL/dF to g_flux
 * (may be NULL) and dL/dSt to *g_stefan (may be NULL).  Solves J^T lambda =
 * g_new (J the step's Jacobian with the linearised closure) by BiCGStab to
 * opts->eps relative (max_iter iterations), preconditioned by opts->nu1
 * sweeps of the transposed smoother (2x2-block red/black Gauss-Seidel on J^T,
 * black first: the transpose of the forward red-then-black smoother), or by
 * the 2x2 block diagonal when nu1 == 0; stats: status 1 converged / 2 iteration limit / 4 breakdown,
 * iterations, |g_new|_2, final residual.  Overwrites the operator's level-0
 * bCoef.
 * ml_hc_porosity_vjp: g_phi += g_chi dchi/dphi (valid cells) and, when
 * g_stefan is not NULL, *g_stefan += sum g_chi dchi/dSt (the porosity's own
 * dependence on St), for losses on the porosity. */
int ml_hc_flux_source(ml_hc_op op, const double* flux, double* src);
int ml_hc_adjoint_step(ml_hc_op op, const double* phi_new, const double* phi_old, double dt,
                       const double* g_new, double* g_old, double* g_flux, double* g_stefan,
                       const ml_solve_opts* opts, ml_solve_stats* stats);
int ml_hc_porosity_vjp(ml_hc_op op, const double* phi, const double* g_chi, double* g_phi, double* g_stefan);

/* ---- state update and coefficient assembly ----------------------------- */
/* porosity of every valid cell (CALCULATEPOROSITY_MULTICOMP,
 * src/EnthalpyVariablesF.ChF:56-94, as in updateEnthalpyVariables
 * src/phaseDiagram.cpp:92-177) */
int ml_hc_porosity(ml_hc_op op, const double* phi, double* chi);
/* HC bCoef from cell porosity (srcSubcycle/AMRLevelMushyLayerInit.cpp:914-953,
 * geometric faces src/CellToEdge2F.ChF:20-46) */
int ml_hc_assemble_bcoef(ml_hc_op op, const double* chi, double* const bCoef[]);

/* updateEnthalpyVariables (src/phaseDiagram.cpp:92-177): from (H, Theta)
 * the 4-component cell field vars = (theta, Theta_l, Theta_s, porosity) on
 * the valid cells of level 0 (CALCULATET, CALCULATECL, CALCULATECS,
 * CALCULATEPOROSITY, src/EnthalpyVariablesF.ChF:13-287) */
int ml_hc_enthalpy_vars(ml_hc_op op, const double* phi, double* vars);

/* ---- cell-centred velocity of a Darcy-Brinkman step --------------------
 * (computeCCvelocity, srcSubcycle/AMRLevelMushyLayerVel.cpp:673-917), on
 * level 0 of a VC operator of the grid.  vars_old / vars_new are the
 * enthalpy variables (ml_hc_enthalpy_vars) at t^n and t^{n+1}; fields at
 * the half time are 0.5*x^n + 0.5*x^{n+1}.  pi is the pressure of the last
 * projection (1 component; ghosts are not read: beyond a wall the pressure
 * is extrapolated to second order, BCutil/PhysBCUtil.cpp:2524-2536, 2628-2633).
 * Velocity components, sources and the projection rhs are separate
 * 1-component cell arrays; only valid cells are written.  Velocity walls
 * are no-slip (BCutil/PhysBCUtil.cpp:1577-1611); periodic sides wrap. */
typedef struct {
  double buoyancyTCoeff;  /* m_buoyancyTCoeff: Pr*Ra_T (diffusive time, advective velocity) */
  double buoyancySCoeff;  /* m_buoyancySCoeff: Pr*Ra_C */
  double porosityLimit;   /* velocity set to zero where the new porosity <= this
                           * (ccVelPorosityLimit = solidPorosity = 0.05) */
  int scalePressure;      /* scaleP_CC: pressure gradients multiplied by the porosity (default 1) */
} ml_flow_params;
/* computeUstarSrc (srcSubcycle/AMRLevelMushyLayerVel.cpp:1724-1853):
 * src_d = buoyancy_d - porosity*grad(pi)_d, buoyancy along the last direction
 * porosity*(buoyancyTCoeff*theta - buoyancySCoeff*Theta_l) (fillBuoyancy,
 * srcSubcycle/AMRLevelMushyLayerFill.cpp:870-885; fillPressureSrcTerm :787-807) */
int ml_vc_ustar_src(ml_vc_op op, const ml_flow_params* fp, const double* vars_old, const double* vars_new,
                    const double* pi, double* const src[]);
/* u = ustar + dt*porosity*grad(pi) (removeGradPressure,
 * srcSubcycle/AMRLevelMushyLayerVel.cpp:654-671), u = 0 where the new
 * porosity <= limit (setVelZero), then rhs = div(u)/dt, the cell-centred
 * divergence through arithmetic face values (Projector::levelProject,
 * srcSubcycle/Projector.cpp:1379-1397; util/Divergence.cpp:349-390) */
int ml_vc_cc_project_rhs(ml_vc_op op, const ml_flow_params* fp, double dt, const double* vars_old,
                         const double* vars_new, const double* pi, const double* const ustar[], double* const u[],
                         double* rhs);
/* the projection's face coefficient bCoef[d] = -(porosity at the half time,
 * arithmetic face mean; srcSubcycle/AMRLevelMushyLayerVel.cpp:717-731),
 * wall faces the adjacent cell's value */
int ml_vc_pressure_faces(ml_vc_op op, const double* vars_old, const double* vars_new, double* const bCoef[]);
/* applyCCcorrection (srcSubcycle/Projector.cpp:1632-1683):
 * u -= (grad(pi)*porosity)*(1/(1/dt)), then u = 0 where the new porosity <= limit */
int ml_vc_cc_correct(ml_vc_op op, const ml_flow_params* fp, double dt, const double* vars_old,
                     const double* vars_new, const double* pi, double* const u[]);
/* Permeability laws (PermeabilityFunctions, src/MushyLayerParams.h:108-130;
 * MushyLayerParams::calculatePermeability, src/MushyLayerParams.cpp:1031-1090) */
enum { ML_PERM_PURE_FLUID = 0, ML_PERM_CUBIC = 1, ML_PERM_KOZENY_CARMAN = 2, ML_PERM_LOG = 3, ML_PERM_POROSITY = 4 };
/* ml_vc_darcy_ccoef with any permeability law; hele_shaw > 0 combines the law
 * harmonically with the Hele-Shaw permeability 1/(1/hele_shaw + 1/Pi) */
int ml_vc_darcy_ccoef_law(ml_vc_op op, const double* chi, double darcy_coeff, int law, double hele_shaw,
                          double* cCoef);

/* ---- the advective source of the (H, Theta) equations ---------------------
 * multiCompAdvectDiffuse -> computeScalarAdvectiveSrcHC
 * (srcSubcycle/AMRLevelMushyLayer.cpp:1645-1672, 2074-2104): face values of
 * W = (theta, Theta_l) at t^n + dt/2 by the unsplit PLM Godunov predictor
 * (PatchGodunov with AdvectionPhysics, srcSubcycle/AdvectionPhysics.cpp;
 * HC options srcSubcycle/AMRLevelMushyLayerInit.cpp:497-537), fluxes
 * W_half*u_face*advectionCoeff (computeScalarAdvectiveFlux, :1165-1198),
 * src = src - div(flux) (util/Divergence.cpp:19-60).
 * Boundary types per [dir][side][comp] in AdvectIBC's numbering
 * (BCutil/AdvectIBC.H:137-144; primBC BCutil/AdvectIBC.cpp:130-260); the
 * plume inflow is not supported (ML_ENOTSUP). */
enum { ML_ADV_DIRICHLET = 0, ML_ADV_EXTRAP = 1, ML_ADV_INFLOW_OUTFLOW = 2, ML_ADV_PLUME_INFLOW = 3,
       ML_ADV_NEUMANN = 4 };
typedef struct {
  double advectionCoeff;   /* m_advectionCoeff (src/MushyLayerParams.cpp:714-801) */
  int type[3][2][2];
  double val[3][2][2];
} ml_advect_params;
/* computeScalDiffusion (srcSubcycle/AMRLevelMushyLayer.cpp:1280-1340): the
 * operator applied with alpha = 0, beta = 1 and inhomogeneous BCs, i.e. the
 * diffusion terms of (H, Theta) at the state phi (2 components; the
 * operator's current face coefficients) */
int ml_hc_diffusive_src(ml_hc_op op, double* phi, double* out);
/* W: 2-component cell array (theta, Theta_l), e.g. the first two components
 * of ml_hc_enthalpy_vars' output; u_face[d]: face velocities; S: the
 * diffusive source (2 components) or NULL (diffusiveSrcForAdvection false);
 * src: 2 components, updated in place on the valid cells; flux[d]: optional
 * 2-component face outputs (edgeScalTotal), NULL to skip.  Ghosts of W and S
 * are not read (periodic sides wrap, the slopes next to a wall are
 * one-sided).  The operator keeps a device workspace for the stages. */
int ml_hc_advect_src(ml_hc_op op, const ml_advect_params* ap, double dt, const double* W,
                     const double* const u_face[], const double* S, double* src, double* const flux[]);
/* CellToEdge of a no-slip cell-centred velocity (1-component arrays per
 * direction): interior and periodic faces 0.5*(u_lo + u_hi), wall faces 0.
 * The advection velocity a caller without the reference's MAC-projected
 * predictor passes to ml_hc_advect_src. */
int ml_vc_face_velocity(ml_vc_op op, const double* const u[], double* const u_face[]);
/* The advective source of a Darcy-Brinkman step for n independent
 * operators of one context and dimension, equal for every i to
 *   ml_hc_diffusive_src(ops[i], phi[i], dsrc[i]);
 *   ml_vc_face_velocity(<ops[i]'s grid>, u_i, u_face_i);
 *   src[i] = 0;
 *   ml_hc_advect_src(ops[i], ap[i], dt, W[i], u_face_i, dsrc[i], src[i], NULL);
 * (srcSubcycle/AMRLevelMushyLayer.cpp:1645-1672,2074-2104) but in one launch
 * per stage for all of them; u and u_face hold n*dim arrays (operator i's
 * direction d at i*dim + d).  No counterpart in the reference, which
 * advances one mushy layer per process; the batched Darcy-Brinkman steps
 * use it. */
int ml_hc_db_advect_src_many(ml_ctx ctx, int n, const ml_hc_op ops[], const ml_advect_params* const ap[], double dt,
                             double* const phi[], const double* const W[], const double* const u[],
                             double* const dsrc[], double* const u_face[], double* const src[]);
/* The other per-instance stages of a batched Darcy-Brinkman step, each
 * equal for every i to the single-operator calls it names, in one launch
 * per kernel for all operators (u, u_old, src, ustar and bCoef hold n*dim
 * arrays, operator i's direction d at i*dim + d; no counterpart in the
 * reference, which advances one mushy layer per process):
 *   ml_hc_db_coefs_many:   ml_hc_enthalpy_vars(ops[i], phi[i], vars[i]);
 *                          ml_hc_assemble_bcoef(ops[i], vars[i] + 3*ncell, bCoef_i)
 *   ml_vc_db_ustar_many:   ml_hc_enthalpy_vars(hops[i], phi[i], vars_new[i]);
 *                          ml_vc_darcy_ccoef_law(uops[i], vars_old[i] + 3*ncell,
 *                              darcy_coeff[i], law[i], hele_shaw[i], cCoef[i]);
 *                          ml_vc_ustar_src(uops[i], fp[i], vars_old[i], vars_new[i], pi[i], src_i);
 *                          u_old_i = u_i
 *   ml_vc_db_project_many: ml_vc_cc_project_rhs(uops[i], fp[i], dt, vars_old[i], vars_new[i],
 *                              pi[i], ustar_i, u_i, rhs[i]);
 *                          ml_vc_pressure_faces(uops[i], vars_old[i], vars_new[i], bCoef_i);
 *                          pi[i] = 0
 *   ml_vc_db_correct_many: ml_vc_cc_correct(uops[i], fp[i], dt, vars_old[i], vars_new[i], pi[i], u_i) */
int ml_hc_db_coefs_many(ml_ctx ctx, int n, const ml_hc_op ops[], const double* const phi[], double* const vars[],
                        double* const bCoef[]);
int ml_vc_db_ustar_many(ml_ctx ctx, int n, const ml_hc_op hops[], const ml_vc_op uops[],
                        const ml_flow_params* const fp[], const double darcy_coeff[], const int law[],
                        const double hele_shaw[], const double* const phi[], const double* const vars_old[],
                        double* const vars_new[], const double* const pi[], double* const cCoef[],
                        double* const src[], const double* const u[], double* const u_old[]);
int ml_vc_db_project_many(ml_ctx ctx, int n, const ml_vc_op uops[], const ml_flow_params* const fp[], double dt,
                          const double* const vars_old[], const double* const vars_new[], double* const pi[],
                          const double* const ustar[], double* const u[], double* const rhs[],
                          double* const bCoef[]);
int ml_vc_db_correct_many(ml_ctx ctx, int n, const ml_vc_op uops[], const ml_flow_params* const fp[], double dt,
                          const double* const vars_old[], const double* const vars_new[], const double* const pi[],
                          double* const u[]);

/* ---- multi-box levels ----------------------------------------------------
 * A level of the reference is a DisjointBoxLayout (the default
 * max_grid_size 128, srcSubcycle/MushyLayerSubcycleUtils.cpp:1011-1012,
 * cuts a 256^2 domain into four boxes), each box with its own FAB of one
 * ghost layer (LevelData<FArrayBox>, component outermost).  The solves run
 * on the whole domain as one array (the layout described at the top); these
 * copy between the boxes and that array on the device, and synchronise.
 * box_lo / box_hi: nbox*3 ints, the inclusive valid-cell bounds of each box
 * inside the domain [0, n-1] (the boxes must be disjoint and cover it);
 * fabs[b]: box b's FAB.
 *   gather:  the domain array's valid cells from the boxes' valid cells;
 *   scatter: the boxes' valid cells from the domain array, and their ghost
 *            cells that lie inside the domain or across a periodic side
 *            (what LevelData::exchange fills); ghosts on a physical wall are
 *            left to the caller's boundary conditions;
 *   gather_faces: the domain's face array of direction dir (FluxBox
 *            component dir, surroundingNodes(box, dir) per box); a face
 *            shared by two boxes takes the later box's value. */
int ml_level_gather(ml_ctx ctx, int dim, const int n[], const int periodic[], int ncomp, int nbox,
                    const int box_lo[], const int box_hi[], const double* const fabs[], double* level);
int ml_level_scatter(ml_ctx ctx, int dim, const int n[], const int periodic[], int ncomp, int nbox,
                     const int box_lo[], const int box_hi[], const double* level, double* const fabs[]);
int ml_level_gather_faces(ml_ctx ctx, int dim, const int n[], int ncomp, int nbox, const int box_lo[],
                          const int box_hi[], int dir, const double* const fabs[], double* level_faces);

/* ---- norms and LevelData arithmetic ------------------------------------ */
/* max |x| over valid cells and ncomp components (AMRPoissonOp::norm, type 0) */
int ml_norm_max(ml_ctx ctx, int dim, const int n[], int ncomp, const double* x, double* out);
/* Elementwise over n doubles (the LinearOp vector functions the reference's
 * operators inherit from AMRPoissonOp):
 *   axby: lhs = a*x + b*y  (y may be NULL: lhs = a*x)
 *   incr: lhs = lhs + scale*x
 *   mult: lhs = lhs*x      (diagonalScale with aCoef, src/AMRNonLinearMultiCompOp.H:75-83)
 *   set:  lhs = v */
int ml_vec_axby(ml_ctx ctx, size_t n, double* lhs, const double* x, const double* y, double a,
                double b);
int ml_vec_incr(ml_ctx ctx, size_t n, double* lhs, const double* x, double scale);
int ml_vec_mult(ml_ctx ctx, size_t n, double* lhs, const double* x);
int ml_vec_set(ml_ctx ctx, size_t n, double* lhs, double v);

#ifdef __cplusplus
}
#endif
#endif
