/*
 * adseismic.h -- C-ABI of the MI355X (gfx950) wave-propagation library.
 *
 * One shared library, libadseismic_hip.so, is the drop-in boundary for ADSeismic.jl's hot
 * path.  Plain pointers and sizes only.  Device pointers are HIP device memory; "host" marks
 * the few small host arrays (geometry).  Every call is stream-ordered on the hipStream_t
 * passed in (NULL = the null stream); nothing synchronises the device except where noted.
 * Functions return ADS_OK (0) or an ADS_ERR_* code; ads_last_error() gives the text
 * (thread-local).
 *
 * Index conventions follow the reference at this boundary: source/receiver positions are
 * 1-based (i, j) on the padded (NX+2) x (NY+2) grid (src/Core.jl:600, 727) and are converted
 * once to 0-based flat indices (i-1)*(NY+2) + (j-1) inside the library.  Fields are row-major
 * (NX+2) x (NY+2), j (depth) fastest.
 */
#ifndef ADSEISMIC_H
#define ADSEISMIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *ads_stream_t; /* hipStream_t */

enum {
    ADS_OK = 0,
    ADS_ERR_ARG = 1,       /* null pointer / bad enum / bad parameter value */
    ADS_ERR_SHAPE = 2,     /* sizes inconsistent */
    ADS_ERR_INDEX = 3,     /* source / receiver index outside the padded grid */
    ADS_ERR_CFL = 4,       /* time step violates the stability bound */
    ADS_ERR_HIP = 5,       /* a HIP runtime call or kernel launch failed */
    ADS_ERR_WORKSPACE = 6, /* workspace too small */
    ADS_ERR_NONFINITE = 7  /* NaN/Inf in the misfit (blow-up) */
};

enum { ADS_F32 = 0, ADS_F64 = 1 };

/* Thread-local text of the last error (empty string if none). */
const char *ads_last_error(void);
/* Library version: major*10000 + minor*100 + patch. */
int ads_version(void);

/* ======================================================================================
 * Drop-in per-step operators.  Each replaces one TensorFlow custom op of the reference
 * (under deps/CustomOps) with the same argument meaning; fields are device arrays of
 * N = (NX+2)*(NY+2) elements; scalar arguments are host values.
 * ==================================================================================== */

/* AcousticOneStep forward: deps/CustomOps/AcousticOneStep/AcousticOneStep.cpp:14-61 and
 * AcousticOneStepCpu.h:1-48 (phi/psi from w).  Ring cells of u, phiout, psiout are set to 0. */
int ads_acoustic_one_step_f64(const double *w, const double *wold, const double *phi,
                              const double *psi, const double *sigma, const double *tau,
                              const double *c, double dt, double hx, double hy, int64_t nx,
                              int64_t ny, double *u, double *phiout, double *psiout,
                              ads_stream_t stream);
int ads_acoustic_one_step_f32(const float *w, const float *wold, const float *phi,
                              const float *psi, const float *sigma, const float *tau,
                              const float *c, double dt, double hx, double hy, int64_t nx,
                              int64_t ny, float *u, float *phiout, float *psiout,
                              ads_stream_t stream);

/* AcousticOneStepGrad: AcousticOneStep.cpp:63-93 (signature), AcousticOneStepCpu.h:51-125
 * (numerics).  Gather form; outputs are overwritten on every cell, ring cells included, and
 * equal the CPU op's scatter result.  The reference's grad_sigma/tau/dt/hx/hy are never
 * written by it and are not produced here. */
int ads_acoustic_one_step_grad_f64(const double *grad_u, const double *grad_phiout,
                                   const double *grad_psiout, const double *w, const double *wold,
                                   const double *phi, const double *psi, const double *sigma,
                                   const double *tau, const double *c, double dt, double hx,
                                   double hy, int64_t nx, int64_t ny, double *grad_w,
                                   double *grad_wold, double *grad_phi, double *grad_psi,
                                   double *grad_c, ads_stream_t stream);
int ads_acoustic_one_step_grad_f32(const float *grad_u, const float *grad_phiout,
                                   const float *grad_psiout, const float *w, const float *wold,
                                   const float *phi, const float *psi, const float *sigma,
                                   const float *tau, const float *c, double dt, double hx,
                                   double hy, int64_t nx, int64_t ny, float *grad_w,
                                   float *grad_wold, float *grad_phi, float *grad_psi,
                                   float *grad_c, ads_stream_t stream);

/* MpiAcousticOneStep forward on one subdomain:
 * deps/CustomOps/MPIAcousticOneStepCpu/MpiAcousticOneStep.h:1-42.  w, wold, phi, psi, sigma,
 * tau are (n+2)^2 halo'd blocks (the caller fills the halo); c, u, phiout, psiout are n*n. */
int ads_mpi_acoustic_one_step_f64(const double *w, const double *wold, const double *phi,
                                  const double *psi, const double *sigma, const double *tau,
                                  const double *c, double dt, double hx, double hy, int64_t n,
                                  double *u, double *phiout, double *psiout, ads_stream_t stream);
int ads_mpi_acoustic_one_step_f32(const float *w, const float *wold, const float *phi,
                                  const float *psi, const float *sigma, const float *tau,
                                  const float *c, double dt, double hx, double hy, int64_t n,
                                  float *u, float *phiout, float *psiout, ads_stream_t stream);

/* GatherOps (GatherOps.h:1-6): out[p] = v[ii[p]-1]; ii is a 1-based device int64 array. */
int ads_gather_f64(const double *v, int64_t nv, const int64_t *ii, int64_t n, double *out,
                   ads_stream_t stream);
int ads_gather_f32(const float *v, int64_t nv, const int64_t *ii, int64_t n, float *out,
                   ads_stream_t stream);
/* ScatterAddOps (ScatterAddOps.h:1-7): out = ipt; out[ii[p]-1] += upd[p] (duplicates add,
 * in index order). */
int ads_scatter_add_f64(const double *ipt, int64_t d, const int64_t *ii, const double *upd,
                        int64_t n, double *out, ads_stream_t stream);
int ads_scatter_add_f32(const float *ipt, int64_t d, const int64_t *ii, const float *upd,
                        int64_t n, float *out, ads_stream_t stream);
/* ScatterNdOps (ScatterNdOps.h:1-7): out = zeros(m); out[ii[p]-1] = upd[p] (last one wins). */
int ads_scatter_nd_f64(const int64_t *ii, const double *upd, int64_t n, int64_t m, double *out,
                       ads_stream_t stream);
int ads_scatter_nd_f32(const int64_t *ii, const float *upd, int64_t n, int64_t m, float *out,
                       ads_stream_t stream);

/* AddSource forward (deps/CustomOps/SourceOps/AddSource.cpp:33-87, CPU semantics: +=):
 * out_k = in_k for the five fields (order sxx, syy, sxy, vx, vy as the op's inputs,
 * AddSource.cc:75-123), then field_type[srci-1, srcj-1] += srcv[s] in source order.
 * srci/srcj/srctype are HOST int64 arrays (validated); srcv is a device array of nsrc. */
int ads_add_source_f64(const double *sxx, const double *syy, const double *sxy, const double *vx,
                       const double *vy, const int64_t *srci, const int64_t *srcj,
                       const int64_t *srctype, int64_t nsrc, int64_t nx, int64_t ny,
                       const double *srcv, double *sxx_out, double *syy_out, double *sxy_out,
                       double *vx_out, double *vy_out, ads_stream_t stream);
/* AddSource forward that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, semantics and results as
 * ads_add_source_f64, but the source table lives in a fixed per-device array of
 * ADS_NOALLOC_MAX_SOURCES entries held until the process ends; nsrc above that is
 * ADS_ERR_SHAPE.  Sources are validated before any output is written.  Every call on one
 * device shares that array: calls that overlap on one device (other streams, other threads)
 * race. */
#define ADS_NOALLOC_MAX_SOURCES 65536
int ads_add_source_noalloc_f64(const double *sxx, const double *syy, const double *sxy,
                               const double *vx, const double *vy, const int64_t *srci,
                               const int64_t *srcj, const int64_t *srctype, int64_t nsrc,
                               int64_t nx, int64_t ny, const double *srcv, double *sxx_out,
                               double *syy_out, double *sxy_out, double *vx_out, double *vy_out,
                               ads_stream_t stream);
/* AddSource backward (AddSource.cpp:91-159): grads pass through; grad_srcv[s] = cotangent of
 * the source's field at its cell. */
int ads_add_source_grad_f64(const double *g_sxx, const double *g_syy, const double *g_sxy,
                            const double *g_vx, const double *g_vy, const int64_t *srci,
                            const int64_t *srcj, const int64_t *srctype, int64_t nsrc, int64_t nx,
                            int64_t ny, double *grad_srcv, ads_stream_t stream);
/* GetReceive forward (deps/CustomOps/ReceiveOps/GetReceive.cpp:10-46):
 * out[r*nt + k] = field_type_r[k][(rcvi-1)(NY+2) + rcvj-1] for nt = NSTEP+1 history levels.
 * Histories are device arrays (nt, N); receivers HOST int64 arrays. */
int ads_get_receive_f64(const double *vx, const double *vy, const double *sxx, const double *syy,
                        const double *sxy, int64_t nt, const int64_t *rcvi, const int64_t *rcvj,
                        const int64_t *rcvtype, int64_t nrcv, int64_t nx, int64_t ny, double *out,
                        ads_stream_t stream);

/* ======================================================================================
 * Parameters (mirror src/Struct.jl) and host-side coefficient helpers.
 * ==================================================================================== */

/* AcousticPropagatorParams, src/Struct.jl:82-121 (+ MPIAcoustic.jl:3-49 origin fields). */
typedef struct {
    int64_t NX, NY, NSTEP;
    double DELTAX, DELTAY, DELTAT;
    int32_t USE_PML_XMIN, USE_PML_XMAX, USE_PML_YMIN, USE_PML_YMAX;
    int64_t NPOINTS_PML;
    double Rcoef, vp_ref;
    /* 0: phi/psi advanced from the new u (default, src/Core.jl:528-549);
     * 1 or 2: phi/psi advanced from w (AcousticOneStep op, Core.jl:475-525). */
    int32_t PropagatorKernel;
    /* 1: the model argument is the velocity and is squared (Core.jl:564), gradients are
     *    returned w.r.t. the velocity; 0: the model is used as is (MPIAcoustic.jl:336). */
    int32_t square_velocity;
    /* x/y coordinate of padded cell 0 (0 for the full grid; subdomain offset otherwise,
     * MPIAcoustic.jl:177-178) and the global interior sizes for the far PML edge. */
    double origin_x, origin_y;
    int64_t NX_global, NY_global; /* 0 = same as NX / NY */
} ads_acoustic_params;

/* compute_PML_Params!(::AcousticPropagatorParams), src/Core.jl:622-655 + pml_helper
 * :738-761.  sigma depends on x only and tau on y only: writes the 1-D host arrays
 * sigma_x[NX+2] and tau_y[NY+2]; the reference's full grids are Sx[i,j] = sigma_x[i],
 * Sy[i,j] = tau_y[j]. */
int ads_acoustic_pml_profiles(const ads_acoustic_params *p, double *sigma_x, double *tau_y);

/* Stability bound of the 2nd-order scheme: c_max * dt * sqrt(1/hx^2 + 1/hy^2) <= 1.
 * Returns ADS_OK or ADS_ERR_CFL. */
int ads_acoustic_check_cfl(const ads_acoustic_params *p, double c_max);

/* ElasticPropagatorParams, src/Struct.jl:4-32 (+ MPIElastic.jl subdomain fields). */
typedef struct {
    int64_t NX, NY, NSTEP;
    double DELTAX, DELTAY, DELTAT;
    double f0, vp_ref;
    int32_t USE_PML_XMIN, USE_PML_XMAX, USE_PML_YMIN, USE_PML_YMAX;
    int64_t NPOINTS_PML;
    double NPOWER, K_MAX_PML, ALPHA_MAX_PML, Rcoef;
    /* 0: ElasticPropagatorSolver semantics (parameter averaging, reduced cell ranges,
     *    src/Core.jl:96-213), arrays of (NX+2) x (NY+2).
     * 1: one MPIElastic block stepped with zero halos (MPIElasticPropagatorSolver,
     *    MPIElastic.jl:374-620): NX x NY is the block (n x n there), arrays are
     *    (NX+4) x (NY+4) with the block at 1-based [3, NX+2] x [3, NY+2]; every fw updates the
     *    whole block, parameters are not averaged, the CPML profiles have NX (NY) entries
     *    indexed by the block cell.  Sources are 1-based into the (NX+4) x (NY+4) arrays (the
     *    MPIElasticSource local index, MPIElastic.jl:85-86); receivers are 1-based into the
     *    block (MPIElasticReceiver, MPIElastic.jl:110-111). */
    int32_t subdomain;
    /* physical x/y of profile index 0 (MPIElastic.jl:231, 306: (II-1) n DELTAX and, as the
     * reference writes it, (JJ-1) n DELTAX) and the global grid size whose far edge places the
     * right/top CPML (0: NX, NY) */
    double origin_x, origin_y;
    int64_t NX_global, NY_global;
} ads_elastic_params;

/* compute_PML_Params(::ElasticPropagatorParams), src/Core.jl:231-407: writes host arrays
 * ax, bx, kx [2*NX] and ay, by, ky [2*NY], row 0 = full grid, row 1 = half grid
 * (Core.jl:397-402). */
int ads_elastic_pml_params(const ads_elastic_params *p, double *ax, double *bx, double *kx,
                           double *ay, double *by, double *ky);

/* ======================================================================================
 * Shot batches (the solver-level hot path).
 * ==================================================================================== */

/* Geometry of a batch of shots; all HOST arrays, copied at plan creation.
 *   shot s has sources src_ptr[s] .. src_ptr[s+1]-1 and receivers rcv_ptr[s] .. rcv_ptr[s+1]-1.
 *   nstep[s] <= NSTEP is the record length of shot s (NULL: NSTEP for every shot).
 * Device data that goes with it:
 *   srcv   : concatenation over shots of an (nstep[s] x nsrc_s) row-major block (the
 *            reference's srcv, Struct.jl:126/54); level n takes row n-1.
 *   traces : acoustic: per shot (nstep[s]+1) x nrcv_s (Core.jl:728);
 *            elastic : per shot nrcv_s x (nstep[s]+1) (GetReceive.cc:65).
 *   offsets: ads_*_plan_offsets() returns the element offset of each shot's block. */
typedef struct {
    int64_t nshots;
    const int64_t *nstep;
    const int64_t *src_ptr;
    const int64_t *srci, *srcj;
    const int64_t *srctype; /* elastic only (Struct.jl:44-48); ignored for acoustic */
    const int64_t *rcv_ptr;
    const int64_t *rcvi, *rcvj;
    const int64_t *rcvtype; /* elastic only */
} ads_shot_geometry;

typedef struct ads_acoustic_plan ads_acoustic_plan;

/* Creates a plan: validates parameters and geometry (ADS_ERR_INDEX for a position outside the
 * padded grid), computes PML profiles and uploads index tables.  dtype: ADS_F32 or ADS_F64. */
int ads_acoustic_plan_create(const ads_acoustic_params *p, const ads_shot_geometry *g, int dtype,
                             ads_acoustic_plan **out);
int ads_acoustic_plan_destroy(ads_acoustic_plan *plan);
/* Sizes in elements of the srcv and trace arrays, and per-shot element offsets (host arrays
 * of nshots entries, may be NULL). */
int ads_acoustic_plan_sizes(const ads_acoustic_plan *plan, int64_t *srcv_elems,
                            int64_t *trace_elems, int64_t *srcv_off, int64_t *trace_off);
/* Workspace bytes needed for a checkpoint interval K (levels between checkpoints; K >= 1).
 * K >= NSTEP-1 stores every level (no recomputation). */
size_t ads_acoustic_plan_workspace_size(const ads_acoustic_plan *plan, int64_t K);
/* Checkpoint interval for a memory budget (0 if none fits): the fewest segments M whose
 * interval fits in budget_bytes, and among the intervals with M segments the shortest,
 * K = ceil((NSTEP-1)/M), which minimises the recomputed levels ((M-1) K). */
int64_t ads_acoustic_plan_choose_interval(const ads_acoustic_plan *plan, size_t budget_bytes);
/* Checkpoint schedule: binomial / sqrt(N) checkpointing with recomputation, in place of the
 * store-every-level tape of Core.jl:570-572, 614.  The NSTEP-1 levels are cut
 * into M = ceil((NSTEP-1)/K) segments; a snapshot holds u^a, u^{a-1} on the full grid and
 * phi^a, psi^a outside the box where both PML profiles vanish, for the shots active at a.
 *   S < 0 (the default): one snapshot per segment start (one-level sqrt(N) scheme, each
 *         segment but the last rebuilt once);
 *   0 <= S < M-1: S snapshot slots reused by the binomial (Revolve-type) reversal schedule.
 * ads_acoustic_plan_workspace_size(K) follows the plan's setting, and a backward must follow a
 * forward made with the same K and setting. */
int ads_acoustic_plan_set_snapshots(ads_acoustic_plan *plan, int64_t snapshots);
int64_t ads_acoustic_plan_snapshots(const ads_acoustic_plan *plan);
/* The checkpoint schedule over M segments with S snapshot slots (S < 0: one per segment
 * start), as rows of three int64 (kind, a, b) in execution order:
 *   4 (m, q): the forward pass keeps state m (the state at the start of segment m) in slot q;
 *   then the adjoint sweep: 0 (m, q) restore state m from slot q (q < 0: the zero state);
 *   1 (m0, m1) advance from state m0 through segments m0 .. m1-1; 2 (m, q) save state m in
 *   slot q; 3 (m, -) adjoint sweep of segment m.  Host arrays; ADS_ERR_SHAPE (with *nops set)
 *   when cap rows are too few. */
int ads_checkpoint_schedule(int64_t M, int64_t S, int64_t *ops, int64_t cap, int64_t *nops);
/* The (K, S) with the least modelled time (forward + adjoint + schedule rebuilds, weighted by
 * the shots active at each level) whose workspace fits budget_bytes; sets the plan's snapshot
 * count to S.  ADS_ERR_WORKSPACE if nothing fits. */
int ads_acoustic_plan_choose_schedule(ads_acoustic_plan *plan, size_t budget_bytes, int64_t *K,
                                      int64_t *S);

/* Forward: traces of every shot (model = velocity or C as square_velocity says).  Keeps the
 * checkpoints in the workspace for a following ads_acoustic_plan_backward. */
int ads_acoustic_plan_forward(ads_acoustic_plan *plan, const void *model, const void *srcv,
                              void *traces, void *workspace, size_t ws_bytes, int64_t K,
                              ads_stream_t stream);
/* Adjoint: given d loss / d traces, accumulates nothing else and OVERWRITES grad_model with
 * d loss / d model (summed over shots).  Needs the workspace of the preceding forward with
 * the same model, srcv and K. */
int ads_acoustic_plan_backward(ads_acoustic_plan *plan, const void *model, const void *srcv,
                               const void *grad_traces, void *grad_model, void *workspace,
                               size_t ws_bytes, int64_t K, ads_stream_t stream);
/* FWI gradient in one call: forward, misfit L = sum (traces - observed)^2 (src/Utils.jl:308),
 * residual, adjoint.  loss: device double[1] (overwritten); traces may be NULL.  Returns
 * ADS_ERR_NONFINITE (after a device sync) if the misfit is not finite. */
int ads_acoustic_plan_gradient(ads_acoustic_plan *plan, const void *model, const void *srcv,
                               const void *observed, double *loss, void *grad_model,
                               void *traces, void *workspace, size_t ws_bytes, int64_t K,
                               ads_stream_t stream);

/* Kernel-shape autotuner: times every candidate shape (rows per wavefront 1/2/4 x shots per
 * workgroup 1..32 x shots per batch 1/2, batch 2 only with 1-2 rows; the forward also 8 rows,
 * the streaming form of its two-level rectangle pass, one shot per wavefront) of the forward
 * and adjoint stencil kernels on this plan's own grid and shots for `reps` steps each, keeps
 * the fastest in the plan, and writes it to shape_out[6] = {fwd rows, fwd shots, fwd batch,
 * adj rows, adj shots, adj batch} (may be NULL).  Synchronises the stream; leaves the
 * workspace contents undefined.  ads_acoustic_plan_set_shape() sets a shape directly (same
 * 6-entry layout; forward rows 8 or 16 select the streaming rectangle where the two-level pass
 * applies and 2 rows elsewhere). */
int ads_acoustic_plan_autotune(ads_acoustic_plan *plan, const void *model, const void *srcv,
                               void *workspace, size_t ws_bytes, int64_t K, int reps,
                               int *shape_out, ads_stream_t stream);
int ads_acoustic_plan_set_shape(ads_acoustic_plan *plan, const int *shape);

/* hipGraph replay (no counterpart in the reference): with enable = 1, each checkpoint segment
 * of the forward sweep, of the recomputation and of the adjoint sweep is captured once into a
 * graph (on a private stream) and replayed on the caller's stream afterwards.  A graph is keyed
 * by the segment, the workspace, srcv, traces and residual addresses; calls with other buffers
 * capture new graphs.  Graphs are not used while ads_profile_begin() timing is active, and are
 * dropped by set_shape/autotune and at plan destruction. */
int ads_acoustic_plan_set_graph(ads_acoustic_plan *plan, int enable);

/* Temporal blocking of the forward sweeps (no counterpart in the reference): with levels = 2
 * (the default) the interior rectangle advances two time levels per pass (u^n and u^{n+1}
 * from u^{n-1}, u^{n-2}), which cuts its HBM traffic by a third; with PropagatorKernel 1 the
 * PML frame does too, in the same launch.  Results are bit-identical to levels = 1.  It
 * applies only when no source lies in the interior rectangle and to pairs of levels with the
 * same number of active shots; otherwise levels are stepped one at a time.
 * ads_acoustic_plan_blocking returns the levels per pass in effect (1 or 2). */
int ads_acoustic_plan_set_blocking(ads_acoustic_plan *plan, int levels);
int ads_acoustic_plan_blocking(const ads_acoustic_plan *plan);

/* Rows per item of the two-level PML-frame pass (no counterpart in the reference): 4, 8, 16,
 * 32 or 64 force that height; 0 (the default) chooses per launch the tallest height not above
 * the number of active shots whose per-wavefront walk over the item's rows, shot after shot,
 * stays well inside the launch's streaming time.  Results are bit-identical for every value. */
int ads_acoustic_plan_set_frame_rows(ads_acoustic_plan *plan, int rows);

/* Plain drop-in of AcousticPropagatorSolver (Core.jl:562-620) on the plan's shots (every shot
 * must have nstep == NSTEP): writes the full histories u, phi, psi, each laid out
 * [shot][level 0..NSTEP][N] (phi/psi may be NULL).  scratch: device buffer of
 * ads_acoustic_plan_history_scratch_size() bytes. */
size_t ads_acoustic_plan_history_scratch_size(const ads_acoustic_plan *plan);
int ads_acoustic_plan_forward_history(ads_acoustic_plan *plan, const void *model,
                                      const void *srcv, void *u_hist, void *phi_hist,
                                      void *psi_hist, void *scratch, ads_stream_t stream);
/* Its pullback: given the dense cotangent of the u history (same layout), OVERWRITES
 * grad_model with d loss / d model.  Needs u_hist from the forward; phi/psi cotangents are
 * taken as zero (no caller of the reference differentiates through them). */
int ads_acoustic_plan_backward_history(ads_acoustic_plan *plan, const void *model,
                                       const void *srcv, const void *u_hist,
                                       const void *grad_u_hist, void *grad_model, void *scratch,
                                       ads_stream_t stream);

/* ---------------------------------- elastic ---------------------------------------- */
typedef struct ads_elastic_plan ads_elastic_plan;

int ads_elastic_plan_create(const ads_elastic_params *p, const ads_shot_geometry *g, int dtype,
                            ads_elastic_plan **out);
int ads_elastic_plan_destroy(ads_elastic_plan *plan);
int ads_elastic_plan_sizes(const ads_elastic_plan *plan, int64_t *srcv_elems,
                           int64_t *trace_elems, int64_t *srcv_off, int64_t *trace_off);
size_t ads_elastic_plan_workspace_size(const ads_elastic_plan *plan, int64_t K);
/* Adjoint kernels: full-grid streaming (2, the default: one wavefront per shot walking rows
 * over the whole grid, x-direction stencils from registers, y-direction from neighbouring
 * lanes, the CPML memory cotangents formed on the fly), streaming over the interior rectangle
 * with separate frame passes (1) or LDS-tiled over the rectangle (0); rows per streaming
 * workgroup (1..32, default 8); shots per adjoint group (1..8, default 4; the parameter terms
 * are accumulated per group); rows x shots <= 64.  The streaming kernels need 3 x shots x rows
 * x 64 values of LDS; above 64 KiB (FP64 past 32 rows x shots) the LDS-tiled kernels run.
 * Changes the workspace layout (mode 2 keeps a second set of memory-variable cotangents):
 * call before ads_elastic_plan_workspace_size and the forward. */
int ads_elastic_plan_set_adjoint_kernels(ads_elastic_plan *plan, int streaming, int rows,
                                         int shots);
/* Forward kernels (no counterpart in the reference): fused = 1 (the default) steps a level in
 * one launch over the whole grid, the stress update two rows ahead of the velocity update in
 * registers, each wavefront walking `rows` rows of one shot (0, the default: 8, 16 or 32 rows,
 * the tallest that leaves enough wavefronts to fill the GPU); fused = 0 runs the
 * stress and the velocity phase as two launches (interior rectangle tiles and the CPML frame
 * as streaming items).  Sources and receivers are applied after either. */
int ads_elastic_plan_set_forward_kernels(ads_elastic_plan *plan, int fused, int rows);
/* As ads_acoustic_plan_choose_interval, over the NSTEP elastic levels. */
int64_t ads_elastic_plan_choose_interval(const ads_elastic_plan *plan, size_t budget_bytes);
/* As the acoustic snapshot calls, for the elastic plan: the NSTEP levels are cut into
 * M = ceil(NSTEP/K) segments; a snapshot slot holds a whole level (13 arrays) of every shot. */
int ads_elastic_plan_set_snapshots(ads_elastic_plan *plan, int64_t snapshots);
int64_t ads_elastic_plan_snapshots(const ads_elastic_plan *plan);
int ads_elastic_plan_choose_schedule(ads_elastic_plan *plan, size_t budget_bytes, int64_t *K,
                                     int64_t *S);
/* Forward: traces of every shot from rho, lambda, mu (device, N each). */
int ads_elastic_plan_forward(ads_elastic_plan *plan, const void *rho, const void *lam,
                             const void *mu, const void *srcv, void *traces, void *workspace,
                             size_t ws_bytes, int64_t K, ads_stream_t stream);
/* Adjoint: OVERWRITES grad_rho, grad_lam, grad_mu (device, N each) with d loss / d(.). */
int ads_elastic_plan_backward(ads_elastic_plan *plan, const void *rho, const void *lam,
                              const void *mu, const void *srcv, const void *grad_traces,
                              void *grad_rho, void *grad_lam, void *grad_mu, void *workspace,
                              size_t ws_bytes, int64_t K, ads_stream_t stream);
int ads_elastic_plan_gradient(ads_elastic_plan *plan, const void *rho, const void *lam,
                              const void *mu, const void *srcv, const void *observed,
                              double *loss, void *grad_rho, void *grad_lam, void *grad_mu,
                              void *traces, void *workspace, size_t ws_bytes, int64_t K,
                              ads_stream_t stream);
/* Pullback of the full-history drop-in: given dense cotangents of the 5 field histories
 * (layout below), OVERWRITES the three gradients.  Reruns the checkpointed forward internally. */
int ads_elastic_plan_backward_dense(ads_elastic_plan *plan, const void *rho, const void *lam,
                                    const void *mu, const void *srcv, const void *grad_hist,
                                    void *grad_rho, void *grad_lam, void *grad_mu, void *workspace,
                                    size_t ws_bytes, int64_t K, ads_stream_t stream);
/* Plain drop-in of ElasticPropagatorSolver (Core.jl:31-94): the full histories of every shot
 * (nstep == NSTEP), dense layout [shot][field vx, vy, sxx, syy, sxy][level 0..NSTEP][N].
 * scratch: ads_elastic_plan_history_scratch_size() bytes (CPML memory variables). */
size_t ads_elastic_plan_history_scratch_size(const ads_elastic_plan *plan);
int ads_elastic_forward_history(ads_elastic_plan *plan, const void *rho, const void *lam,
                                const void *mu, const void *srcv, void *hist, void *scratch,
                                ads_stream_t stream);

/* ======================================================================================
 * Kernel timing for benchmarks.  Between ads_profile_begin() and ads_profile_end() every
 * launch of the main stencil kernels made by the plan entry points is bracketed by HIP events
 * on its launch stream; ads_profile_end() waits for them and accumulates, per kernel class,
 * the summed duration, the launch count and the algorithmic HBM bytes of those launches
 * (compulsory per-cell traffic: acoustic forward w(4 + 4 f_pml), adjoint w(7 + 4 f_pml) for
 * PropagatorKernel 1; w(4 + 2 f_pml) and w(7 + 2 f_pml) for kernel 0 whose phi/psi updates
 * run in a separate pass; elastic forward w(8 + 4 f_x + 4 f_y) stress and w(5 + 4 f_x + 4 f_y)
 * velocity phase, adjoint w(12 + 4 f_x + 4 f_y) per phase: the whole-step counts
 * w(13 + 8 f_x + 8 f_y) and w(24 + 8 f_x + 8 f_y) split by the fields each phase updates; the
 * fused elastic forward step, the default, is timed in ADS_PROF_EL_FWD_STRESS with the
 * whole-step count and leaves ADS_PROF_EL_FWD_VEL empty).  A launch that advances two time
 * levels (the acoustic forward's two-level pass) counts as two launches.
 * ads_profile_begin_sampled(every) brackets only launches 0, every, 2*every, ... of each class
 * (counted from the begin call), so that a long timed run carries few events; the queried
 * duration, launch count and bytes then cover those launches alone.  ads_profile_begin() is
 * ads_profile_begin_sampled(1).  Not thread-safe across concurrent callers; meant for one
 * benchmarking thread.
 * ==================================================================================== */
enum {
    ADS_PROF_AC_FWD = 0,       /* acoustic forward step kernel */
    ADS_PROF_AC_ADJ = 1,       /* acoustic adjoint step kernel */
    ADS_PROF_EL_FWD_STRESS = 2,
    ADS_PROF_EL_FWD_VEL = 3,
    ADS_PROF_EL_ADJ_VEL = 4,
    ADS_PROF_EL_ADJ_STRESS = 5,
    ADS_PROF_COUNT = 6
};
int ads_profile_begin(void);
int ads_profile_begin_sampled(int every);
int ads_profile_end(void);
int ads_profile_query(int kernel_class, double *total_ms, int64_t *launches, double *alg_bytes);

#ifdef __cplusplus
}
#endif
#endif /* ADSEISMIC_H */
