/*
 * openfea_amd.h -- C ABI of the MI355X-native OpenFEA element hot path.
 *
 * libofea_hip.so exports exactly these symbols.  All indices are int32,
 * node ids in element connectivity are 1-based (as in the reference's
 * node_topos_), every other index is 0-based.  Functions return OFEA_OK (0)
 * or a negative status; they never exit().  ofea_last_error() gives the
 * message of the last failure on the calling thread.  One context per host
 * thread; work on a context is ordered on the context's own HIP stream.
 * Host buffers are caller-owned and only read/written during the call.
 *
 * Reference call surface replaced (citations relative to the reference tree):
 *   ofea_inp_*            ReadInfo::read_ele_node_num/read_geo_mesh/
 *                         read_load_bcs/read_dis_bcs/read_mat/read_time
 *                         (src/data2cae.cpp:18-179, :315-492)
 *   ofea_build_free_index set_BCs::build_free_index (src/set_bc.cpp:20-43)
 *   ofea_build_single_load set_BCs::build_single_load (src/set_bc.cpp:46-77)
 *   ofea_set_mesh         data_management fields coords_/node_topos_/
 *                         ele_list_idx_/resort_free_nodes_
 *                         (src/include/data_management.h:45-104)
 *   ofea_pattern          assamble_stiffness::build_CSR (src/assemble.cpp:19-77)
 *   ofea_assemble         assamble_stiffness::fill_CSR_sparse_mat
 *                         (src/assemble.cpp:95-151)
 *   ofea_solve            solution_api(..., solver_type) (src/linear_solution.cpp:17-64)
 *                         with GPU PCG methods beside "Pardiso_*"/"SuperLU_*"
 *   ofea_nl_assemble      assamble_nl_stiffness::fill_CSR_sparse_mat
 *                         (src/nl_assemble.cpp:94-166)
 *   ofea_nl_solve         Newton loop of src/cae.cpp:209-253 with load increments
 *   ofea_nl_solve_ref     the same loop with the reference's stopping rule
 *   ofea_lumped_mass      ele_base::build_ele_mass over all elements (src/cae.cpp:439-448)
 *   ofea_explicit_run     CAE_process::explicit_analysis time loop (src/cae.cpp:420-574),
 *                         fixed step count
 *   ofea_explicit_analysis the same loop with its time control (auto step, CheckTime, outputs)
 *   ofea_build_explicit_load OutFroce_ of src/cae.cpp:449-456 (int-truncated value)
 *   ofea_reset_displacement simulation_post::reset_displacement (src/post_process.cpp:18-40)
 *   ofea_ca_info_*        ReadInfo::CA_read_change (src/data2cae.cpp:519-665)
 *   ofea_ca_delta         ca_get_delt_stiffness / ca_fill_CSR_sparse_mat (src/ca_analysis.cpp:18-143)
 *   ofea_ca_solve         ca_build_rom + ca_SVD + ca_solve (src/ca_analysis.cpp:146-309),
 *                         driven as in CAE_process::CA_ReAnalysis (src/cae.cpp:317-376)
 *   ofea_reset_ca_displacement simulation_post::reset_ca_displacement (src/post_process.cpp:42-76)
 *   ofea_post_nodal       simulation_post::get_cauchy_stress_3d (src/post_process.cpp:79-168)
 *   ofea_write_vtk        data_process::export_dis_2_vtk / CA_export_dis_2_vtk
 *                         (src/data2vtk.cpp:47-264, :266-427)
 *   ofea_sfem_setup       SFEM3D::build + assamble_stiffness::SFEM_build_CSR
 *                         (src/SFEM3D.cpp:8-134, src/assemble.cpp:291-350)
 *   ofea_sfem_assemble    SFEM_fill_CSR_sparse_mat + SFEM_C3D4::build_ele_stiff_mat
 *                         (src/assemble.cpp:369-414, src/Elements/SFEM_C3D4.cpp:53-427)
 */
#ifndef OPENFEA_AMD_H
#define OPENFEA_AMD_H

#ifdef __cplusplus
extern "C" {
#endif
#pragma GCC visibility push(default)

#define OFEA_OK 0
#define OFEA_NOT_CONVERGED 1   /* solver hit maxit; x holds the last iterate */
#define OFEA_ERR_ARG (-1)      /* invalid argument or size                    */
#define OFEA_ERR_HIP (-2)      /* HIP runtime failure (no device, OOM, ...)   */
#define OFEA_ERR_STATE (-3)    /* call out of order (e.g. assemble before mesh) */
#define OFEA_ERR_IO (-4)       /* file could not be read / parsed             */
#define OFEA_ERR_LIMIT (-5)    /* mesh exceeds a documented kernel limit      */

/* element type ids, as ELE_TYPES in src/include/data_management.h:74 */
#define OFEA_C3D4 0
#define OFEA_C3D8 1

/* solver methods for ofea_solve */
#define OFEA_PCG_JACOBI 0        /* FP64 values, point-Jacobi PCG                       */
#define OFEA_PCG_BLOCK_JACOBI 1  /* FP64 values, 3x3 node-block Jacobi PCG              */
#define OFEA_PCG_MIXED 2         /* FP32 matrix values, FP64 vectors, block Jacobi, FP64 reliable updates */
#define OFEA_PCG_LEARNED 3       /* FP64 PCG with the learned block-Jacobi correction            */
#define OFEA_PCG_MIXED32 4       /* FP32 values and FP32 vectors, block Jacobi, FP64 reliable updates */
#define OFEA_PCG_LEARNED_MIXED 5 /* as MIXED with the learned block-Jacobi correction            */
/* two-level methods: the local part (3x3 block Jacobi or the learned blocks)
 * plus an additive coarse correction Z A_c^-1 Z^T r over rigid-body modes of
 * node aggregates (FP32 coarse inverse, set up from the resident K) */
#define OFEA_PCG_TWO_LEVEL 6                /* FP64 values, block Jacobi + coarse             */
#define OFEA_PCG_TWO_LEVEL_MIXED 7          /* FP32 values, FP64 vectors, reliable updates    */
#define OFEA_PCG_LEARNED_TWO_LEVEL 8        /* FP64 values, learned blocks + coarse           */
#define OFEA_PCG_LEARNED_TWO_LEVEL_MIXED 9  /* FP32 values, learned blocks + coarse           */
#define OFEA_PCG_TWO_LEVEL_MIXED32 10       /* FP32 values and vectors, block Jacobi + coarse */
#define OFEA_PCG_LEARNED_TWO_LEVEL_MIXED32 11 /* FP32 values and vectors, learned blocks + coarse */

/* quantities of ofea_post_nodal */
#define OFEA_POST_STRAIN 0  /* B u_e and its "Mises" value: what the reference writes as S_* */
#define OFEA_POST_STRESS 1  /* Cauchy stress D B u_e and its von Mises value                */

/* explicit precision for ofea_explicit_run */
#define OFEA_FP32 0
#define OFEA_FP64 1
#define OFEA_BF16 2            /* learned-network inference only (ofea_learned_precision) */

typedef struct ofea_ctx ofea_ctx;
typedef struct ofea_deck ofea_deck;
typedef struct ofea_ca_info ofea_ca_info;

const char *ofea_last_error(void);
/* library version string, e.g. "openfea_amd 0.1 gfx950" */
const char *ofea_version(void);

/* ---- host drop-in: reference-dialect .inp reader (no GPU needed) ---- */
/* Reads the deck the way src/data2cae.cpp does (line order defines node and
 * element numbering, counts from the first *Nset and *Elset follow-on lines, set
 * keywords "Set-load"/"Set-fix", first "*Cload" line, "*Elastic",
 * "*Density", "*Dynamic, Explicit" time line). */
int ofea_inp_open(const char *path, ofea_deck **deck);
int ofea_inp_sizes(const ofea_deck *deck, int *nd, int *ne, int *nfix, int *nload);
/* xyz nd*3, conn ne*8 (1-based, unused slots 0), etype ne, fix_set nfix,
 * load_set nload; scalars may be NULL.  A missing *Elastic leaves E/nu as
 * NaN (the reference keeps the constructor's material). */
int ofea_inp_get(const ofea_deck *deck, double *xyz, int *conn, int *etype, int *fix_set, int *load_set,
                 int *load_dof, double *load_value, double *E, double *nu, double *rho, double *dt,
                 double *t_total);
void ofea_inp_close(ofea_deck *deck);

/* ---- .inp dialect extensions (no reference counterpart; no GPU needed) ----
 * ofea_inp_open also reads, at keyword level, what the reference's reader
 * leaves out: *Nset and *Elset (ids, set names, generate), *Surface type=ELEMENT
 * ("elset_or_element, S#", Abaqus face numbering), every *Cload line (loads
 * add), *Dsload (P: pressure, traction -p n_out; TRVEC: magnitude and
 * direction) as consistent nodal forces (tri faces exact, quad faces 2x2
 * Gauss), *Boundary (ENCASTRE, PINNED, XSYMM/YSYMM/ZSYMM or "first[, last[,
 * value]]"), *Initial Conditions type=VELOCITY and *Material, *Elastic,
 * *Density assigned by *Solid Section.  Ids map to line-order indices id-1
 * as in the reference.  A deck whose extension blocks cannot be read still
 * opens; these getters then return OFEA_ERR_IO with the reason. */
/* f: nd*3 nodal forces of all *Cload and *Dsload lines */
int ofea_inp_nodal_loads(const ofea_deck *deck, double *f);
/* nodes with all three DOFs constrained (1-based ids into fixed, capacity nd;
 * ascending; NULL to count): the clamp list of ofea_build_free_index.
 * npartial counts constrained DOFs of nodes not fully fixed and nnonzero the
 * entries with a prescribed nonzero value; neither fits the reference's
 * node elimination (callers reject or treat them otherwise). */
int ofea_inp_boundary(const ofea_deck *deck, int *nfixed, int *fixed, int *npartial, int *nnonzero);
/* v0: nd*3 initial velocity (0 where none is given) */
int ofea_inp_initial_velocity(const ofea_deck *deck, double *v0);
/* per-element E, nu, rho from the section assignments (ne each, nullable;
 * NaN for an element without a section) */
int ofea_inp_materials(const ofea_deck *deck, double *E_elem, double *nu_elem, double *rho_elem);

/* ---- host drop-in: boundary conditions (no GPU needed) ---- */
/* set_BCs::build_free_index: sorts fix_set in place, fills resort (nd) with
 * the free-node rank or -1; returns the number of free nodes (>= 0) or < 0. */
int ofea_build_free_index(int nd, int *fix_set, int nfix, int *resort);
/* set_BCs::build_single_load: F (n = 3*free nodes) zeroed, then F[3r+dof-1]
 * = value for each load-set node (assignment, not sum). */
int ofea_build_single_load(int n, const int *resort, const int *load_set, int nload, int dof, double value,
                           double *F);
/* build_CSR on the host, without a device context: pass col_ptr/row_idx NULL
 * to get n and nnz, then arrays of n+1 and nnz entries for the CSC pattern
 * (src/assemble.cpp:19-77: rows ascending per column, clamped DOFs dropped). */
int ofea_build_csr(int nd, int ne, const int *conn, const int *etype, const int *resort, int *n, long long *nnz,
                   int *col_ptr, int *row_idx);
/* reset_displacement: u_full (3*nd) from the free-DOF u (3*free). */
int ofea_reset_displacement(int nd, const int *resort, const double *u_free, double *u_full);

/* reset_ca_displacement: u_new (3*nd_new, zeroed first) receives, for every
 * listed node i < nmap with idref2new[i] != -1, the displacement of node i of
 * the original model (free-DOF u of that model, 0 at clamped nodes). */
int ofea_reset_ca_displacement(int nd_all, const int *resort, const double *u_free, int nmap, const int *idref2new,
                               int nd_new, double *u_new);

/* ---- host drop-in: CA change file (no GPU needed) ---- */
/* "*nodes_all, N" + "x, y, z" lines; "*eles_change, M, type=T" + "kind, n1..nk"
 * lines (kind 0 delete, 1 move, 2 add; node ids 0-based); "*idref2new, N" +
 * one index per line (-1 = node removed) up to "*End File". */
int ofea_ca_info_open(const char *path, ofea_ca_info **info);
int ofea_ca_info_sizes(const ofea_ca_info *info, int *nd_all, int *nmod, int *nmap);
/* xyz_all nd_all*3, kind nmod, conn_mod nmod*8 (0-based, unused slots 0),
 * etype_mod nmod, idref2new nmap; any may be NULL */
int ofea_ca_info_get(const ofea_ca_info *info, double *xyz_all, int *kind, int *conn_mod, int *etype_mod,
                     int *idref2new);
void ofea_ca_info_close(ofea_ca_info *info);

/* ---- host drop-in: VTK output (no GPU needed) ---- */
/* Legacy ASCII VTK in the reference's dialect: points xyz + scale*u, tets in
 * node order (0,2,1,3), hexes in voxel order, cell types 10/11, fields u_x,
 * u_y, u_z, u_magnitude and, if node_fields (7*nd, component-major, as
 * ofea_post_nodal writes) is not NULL, S_xx S_yy S_zz S_xy S_yz S_xz S_Mises.
 * conn is ne*8, 1-based. */
int ofea_write_vtk(const char *path, int nd, const double *xyz, int ne, const int *conn, const int *etype,
                   const double *u_full, double scale, const double *node_fields);

/* ---- context ---- */
int ofea_create(ofea_ctx **ctx, int device);
int ofea_destroy(ofea_ctx *ctx);
/* Copies the mesh, builds the CSC pattern and the gather plans on the host,
 * uploads everything to HBM.  conn: ne*8 1-based ids (C3D4 uses slots 0..3),
 * etype: 0 C3D4 / 1 C3D8, resort: nd free-node rank or -1. */
int ofea_set_mesh(ofea_ctx *ctx, int nd, const double *xyz, int ne, const int *conn, const int *etype,
                  const int *resort);
/* build_CSR: n free DOFs and nnz; when col_ptr/row_idx are non-NULL they
 * receive the CSC arrays (n+1, nnz), identical to the reference's. */
int ofea_pattern(ofea_ctx *ctx, int *n, long long *nnz, int *col_ptr, int *row_idx);

/* Per-element Young's modulus (ne), kept resident for later assemblies;
 * NULL clears it (then the global E of ofea_assemble is used). */
int ofea_set_elem_E(ofea_ctx *ctx, const double *E_elem);
/* fill_CSR_sparse_mat on the GPU (FP64).  The values stay resident in HBM for
 * ofea_solve; nz_val (nnz) receives a copy when non-NULL (then the call
 * synchronises), otherwise the call is asynchronous on the context stream. */
int ofea_assemble(ofea_ctx *ctx, double E, double nu, double *nz_val);

/* y = K x with the resident K (FP64), host vectors of length n. */
int ofea_spmv(ofea_ctx *ctx, const double *x, double *y);

/* RHS upload (n) kept resident; ofea_solve with b == NULL uses it. */
int ofea_set_rhs(ofea_ctx *ctx, const double *b);
/* Solve K x = b with the resident K.  Stops when the true FP64 residual
 * ||b - Kx||_2 <= rtol ||b||_2 (checked in FP64 for every method) or after
 * maxit inner iterations.  b/x may be NULL (resident RHS / keep x on device).
 * iters/relres may be NULL.  Returns OFEA_OK or OFEA_NOT_CONVERGED. */
int ofea_solve(ofea_ctx *ctx, int method, const double *b, double *x, double rtol, int maxit, int *iters,
               double *relres);
int ofea_get_solution(ofea_ctx *ctx, double *x);

/* Largest system (free DOFs) that ofea_solve_noalloc takes. */
#define OFEA_NOALLOC_MAX_DOFS 262144
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ofea_solve's OFEA_PCG_JACOBI / OFEA_PCG_BLOCK_JACOBI on
 * caller-owned DEVICE vectors d_b (read) and d_x (written), n = the free DOFs
 * of the resident K (ofea_set_mesh + ofea_assemble beforehand).  The call
 * allocates nothing: the run's vectors, preconditioner, partial sums and
 * scalars live in one static device array of 23.1 MB that every GPU the
 * library runs on holds from ofea_create until the process ends.  It returns
 * the same x (bit for bit) and iteration count as ofea_solve with the same
 * method; systems above OFEA_NOALLOC_MAX_DOFS give OFEA_ERR_LIMIT.  All calls
 * on one GPU share that array: calls that overlap in time on one GPU (other
 * threads, other contexts of the same device) race and must not be issued.
 * The call returns after the solve has finished (it synchronises the context
 * stream); the caller must have finished writing d_b beforehand. */
int ofea_solve_noalloc(ofea_ctx *ctx, int method, const double *d_b, double *d_x, double rtol, int maxit,
                       int *iters, double *relres);

/* ---- explicit dynamics ---- */
/* Lumped nodal mass (nd) of all elements at density rho. */
int ofea_lumped_mass(ofea_ctx *ctx, double rho, double *mass);
/* Half-step central-difference loop with fixed dt for nsteps steps.
 * mass nd, f_ext 3*nd, v0 3*nd (nullable), fixed_mask nd (nonzero = clamped),
 * precision OFEA_FP32 / OFEA_FP64, u_out 3*nd (nullable: keep on device). */
int ofea_explicit_run(ofea_ctx *ctx, double E, double nu, const double *mass, const double *f_ext,
                      const double *v0, const int *fixed_mask, double dt, int nsteps, int precision,
                      double *u_out);

/* CAE_process::explicit_analysis (src/cae.cpp:420-574) with the reference's
 * time control.  time_step 0 selects the automatic step, recomputed every
 * step as 0.9 x the smallest 2 / sqrt(max row sum of |Ke|) over the elements
 * at the current coordinates (UpdateTimeStep, src/explicit_tools.cpp:21-31;
 * update_timestep, src/Elements/C3D8.cpp:195-219, src/Elements/C3D4.cpp:140-164);
 * a given time_step is used as it is and, as in the reference, only lowered
 * by CheckTime (src/explicit_tools.cpp:48-56), which clamps the step to land
 * on each multiple of time_total / output_gap (the reference's output_gap_
 * is 10) and on time_total.  The displacement (3nd) is saved at t = 0 and
 * whenever the current time equals time_total or the current output time;
 * the loop ends after the save at time_total or when the time passes it.
 * Saves go to frames (max_frames x 3nd), with the output index of each
 * (frame_index, the number of the reference's "<path>_<k>.vtk") and its time
 * (frame_time); *n_frames counts the saves, *n_steps the loop iterations.
 * Time control is FP64; precision selects the arithmetic of the element and
 * node phases.  Inputs as ofea_explicit_run (f_ext from
 * ofea_build_explicit_load for the reference's single load).  Returns
 * OFEA_ERR_LIMIT when max_steps iterations end before time_total or the saves
 * exceed max_frames (the first max_frames are returned). */
int ofea_explicit_analysis(ofea_ctx *ctx, double E, double nu, const double *mass, const double *f_ext,
                           const double *v0, const int *fixed_mask, double time_total, double time_step,
                           int output_gap, int precision, int max_steps, int max_frames, double *frames,
                           int *frame_index, double *frame_time, int *n_frames, int *n_steps);
/* External force of the explicit loop (src/cae.cpp:449-456): f_ext (3nd)
 * zero except f_ext[3 (node - 1) + dof - 1] = (int)value for every node of
 * load_set (1-based) -- the reference stores the load value in an int, so
 * its fraction is dropped. */
int ofea_build_explicit_load(int nd, const int *load_set, int nload, int dof, double value, double *f_ext);

/* ---- geometric nonlinear (total Lagrangian St.Venant-Kirchhoff) ---- */
/* Tangent K_T (resident, copied to nz_val when non-NULL) and internal force
 * f_int (n, free DOFs) at the free-DOF displacement u (n). */
int ofea_nl_assemble(ofea_ctx *ctx, double E, double nu, const double *u, double *nz_val, double *f_int);
/* Newton loop of CAE_process::implict_analysis (src/cae.cpp:209-253) on the GPU,
 * extended with load stepping: the load F (n) is ramped in n_incr equal
 * increments; each increment iterates K_T du = lam F - f_int, u += du until
 * ||lam F - f_int||_2 <= tol ||lam F||_2 (FP64) or max_newton iterations.
 * Linear solves: ofea_solve method with inner_rtol; a negative inner_rtol
 * asks for inexact Newton instead: the Eisenstat-Walker forcing terms
 * (eta_k = 0.9 (|r_k| / |r_k-1|)^2, safeguarded by 0.9 eta_k-1^2 when that
 * exceeds 0.1, at most 0.1) bounded below by -inner_rtol and by half of what
 * the outer tolerance needs.  A learned method is
 * judged by the same solver with plain block Jacobi: on the first linear
 * solve the plain solver runs, then the learned one limited to the plain
 * iteration count, and the learned method serves the rest of the run only if
 * it converged in fewer iterations; the plain solver also takes over whenever
 * the learned one fails (linear_total counts every PCG iteration).  u_out
 * (n), counters nullable.
 * Returns OFEA_OK or OFEA_NOT_CONVERGED. */
int ofea_nl_solve(ofea_ctx *ctx, double E, double nu, const double *F, int n_incr, double tol, int max_newton,
                  int method, double inner_rtol, double *u_out, int *newton_iters, int *linear_iters,
                  double *relres);
/* The Newton loop of CAE_process::implict_analysis exactly as the reference
 * runs it (src/cae.cpp:209-253): res_norm = 1; while res_norm > res_limit
 * (data_management::res_lmit, 1e-6 by default) and loop < max_loop (100 in
 * the reference): K_T and f_int at u, r = F - f_int, K_T du = r, u += du,
 * res_norm = ||r||_2 / n of that r (n = 3 x free nodes), loop++.  The full
 * load at once; the du of the last residual test is applied.  The linear
 * solves are GPU PCG (method, inner_rtol) in place of PARDISO.  u_out (n),
 * counters nullable.  Returns OFEA_OK, or OFEA_NOT_CONVERGED when max_loop
 * loops end with res_norm above res_limit. */
int ofea_nl_solve_ref(ofea_ctx *ctx, double E, double nu, const double *F, double res_limit, int max_loop,
                      int method, double inner_rtol, double *u_out, int *loops, int *linear_iters,
                      double *res_norm);

/* ---- learned factorized correction of block Jacobi (no counterpart in the
 * reference) ----
 * M = G^T G with G block lower-triangular on K's node-block pattern:
 *   G_ii = L_i^-1 (D_i = L_i L_i^T, the 3x3 diagonal block of free node i),
 *   G_is = -alpha_is L_i^-1 K_is D_s^-1 for each neighbour s < i,
 *   alpha_is = 2 sigmoid(u_i + v_s) in (0, 2),
 * where (u_i, v_i) are the two outputs of a 2-layer node-graph network
 * (hidden 64, mean aggregation) over 9 features per node of the resident K.
 * alpha = 0 is block Jacobi; M is SPD for any parameters.  OFEA_PCG_LEARNED*
 * keep G in BF16 and apply z = G^T y with y = G r carried by its own
 * recursion (one sweep over each half of G per iteration). */
int ofea_learned_nparam(void);
/* features of the resident K, deterministic initial parameters (head zero:
 * alpha = 1, the first-order rule), forward pass */
int ofea_learned_init(ofea_ctx *ctx, unsigned long long seed);
int ofea_learned_set_params(ofea_ctx *ctx, const float *params, int nparam);
int ofea_learned_get_params(ofea_ctx *ctx, float *params, int nparam);
/* loss tr(G K G^T) / n for the resident K (n = free DOFs; 1 at alpha = 0;
 * det(G K G^T) does not depend on alpha, so this is Kaporin's condition
 * number of G K G^T up to a constant factor) and its gradient (nparam) */
int ofea_learned_loss_grad(ofea_ctx *ctx, double *loss, float *grad);
/* `steps` Adam steps at learning rate lr on the resident K; loss_hist (steps, nullable) */
int ofea_learned_train(ofea_ctx *ctx, int steps, float lr, double *loss_hist);
/* the same on the mean loss of nctx contexts (one resident system each, all on
 * one device, each after ofea_learned_init): the parameters start from
 * ctxs[0]'s and end in every context */
int ofea_learned_train_multi(ofea_ctx *const *ctxs, int nctx, int steps, float lr, double *loss_hist);
/* the factor G for the resident K as nnz values on K's layout (per free node
 * i and block slot k a row-major 3x3: G_is for s < i, G_ii, (G_si)^T for
 * s > i; stored BF16, returned widened to FP64) and the features (9 per free
 * node); either pointer nullable */
int ofea_learned_blocks(ofea_ctx *ctx, double *blocks, float *features);
/* the network's outputs (u_i, v_i) per free node (2 nfree), as they serve the solver */
int ofea_learned_outputs(ofea_ctx *ctx, float *uv);
/* recompute features and the factor for a newly assembled K */
int ofea_learned_refresh(ofea_ctx *ctx);
/* inference precision of the network's hidden layer when it produces the
 * solver's factor: OFEA_FP32 (default) or OFEA_BF16 (BF16 operands, FP32
 * accumulation on the matrix cores, v_mfma_f32_16x16x16_bf16); training stays
 * FP32.  Re-runs the network for the resident K. */
int ofea_learned_precision(ofea_ctx *ctx, int precision);
/* out = Z A_c^-1 Z^T r (n), the coarse correction of the two-level methods for
 * the resident K (built on first use). */
int ofea_coarse_apply(ofea_ctx *ctx, const double *r, double *out);
/* The aggregates of that coarse space (no counterpart in the reference):
 * agg (nfree = n/3) receives the aggregate of every free node; returns the
 * number of aggregates (>= 0) or an error.  Aggregates are compact parts of
 * the free nodes (slabs across the longest axis, cut by recursive coordinate
 * bisection), all of the same node count except the last. */
int ofea_coarse_aggregates(ofea_ctx *ctx, int *agg);
/* Coarse inverses A_c^-1 built since the mesh was set (no counterpart in the
 * reference): every two-level solve after a new assembly builds one, the
 * Newton drivers only when a tangent's solve slows down (>= 0, or an error). */
int ofea_coarse_builds(ofea_ctx *ctx);

/* ---- combined-approximation (CA) reanalysis on the resident K0 ----
 * ofea_ca_delta: Delta K on the pattern of the resident K (nnz values, copied
 * to dK_val when not NULL) from nmod modified elements: kind 0 (deleted) adds
 * -K_e, kind 2 (added) +K_e, kind 1 (moved) K_e(new) - K_e(old); conn_mod is
 * nmod*8, 0-based node ids of the context's mesh; the element stiffness uses
 * xyz_m (nd*3, the modified coordinates; NULL = the mesh's), the old side of a
 * moved element the mesh's coordinates.  Summation per CSC slot is in list
 * order.  Every modified element must couple only node pairs already in the
 * pattern (OFEA_ERR_ARG otherwise; the reference leaves new couplings as a
 * TODO, src/cae.cpp:324-329). */
int ofea_ca_delta(ofea_ctx *ctx, double E, double nu, int nmod, const int *kind, const int *conn_mod,
                  const int *etype_mod, const double *xyz_m, double *dK_val);
/* ofea_ca_solve: u_out (n) = U y with U an orthonormal basis of
 * span{r_0 = u0, r_i = -K0^-1 Delta K r_{i-1}, i < n_basis} and
 * (U^T (K0 + Delta K) U) y = U^T F; the K0 solves run the device PCG `method`
 * on the resident K0 to rtol.  Series vectors numerically inside the span of
 * the earlier ones are dropped; *n_used (nullable) is the basis size kept,
 * *pcg_iters (nullable) the PCG iterations over all K0 solves, reduced
 * (nullable, >= n_basis^2 + 2 n_basis doubles) receives K_r (row-major),
 * F_r and y.  n_basis in [1, 32].  Overwrites the context's right-hand side.
 * Returns OFEA_NOT_CONVERGED if a K0 solve missed rtol. */
int ofea_ca_solve(ofea_ctx *ctx, int n_basis, const double *u0, const double *F, int method, double rtol, int maxit,
                  double *u_out, int *n_used, int *pcg_iters, double *reduced);

/* ---- nodal strain / stress recovery ----
 * From the full displacement u_full (3*nd, as ofea_reset_displacement gives):
 * per element 7 values [xx, yy, zz, xy, yz, xz, Mises] (elem_out 7*ne,
 * component-major, nullable), and their averages over each node's element
 * incidences (node_out 7*nd, component-major, nullable; 0 for a node in no
 * element).  quantity OFEA_POST_STRAIN gives what the reference writes
 * (engineering strain B u_e with its formula sqrt(0.5((e0-e1)^2 + (e1-e2)^2 +
 * (e2-e0)^2 + 6(e3^2+e4^2+e5^2)))); OFEA_POST_STRESS the Cauchy stress
 * D B u_e (E, or the moduli of ofea_set_elem_E) and its von Mises value.
 * C3D8 (no routine in the reference) uses the detJ-weighted mean of its 8
 * Gauss-point values.  u_full NULL uses the context's resident solution of
 * the last ofea_solve (expanded on the device, 0 at clamped nodes); with
 * elem_out and node_out NULL the results stay on the device and the call
 * returns without synchronising (time it with ofea_timer_*). */
int ofea_post_nodal(ofea_ctx *ctx, double E, double nu, const double *u_full, int quantity, double *elem_out,
                    double *node_out);

/* ---- node-based smoothed FEM, C3D4 only (CAE_process::implict_SFEManalysis,
 * src/cae.cpp:379-416) ----
 * ofea_sfem_setup builds the smoothing domains of the context's mesh (one per
 * node: its cell in the surrounding tets, closed by the surface faces through
 * it) and replaces the context's pattern by the smoothed one: column j holds
 * row i when some node's ring contains both (SFEM_build_CSR), so ofea_pattern,
 * ofea_solve, ofea_spmv and ofea_post_nodal work on the smoothed system.
 * ofea_assemble, ofea_nl_*, ofea_assemble_batch and ofea_ca_delta return
 * OFEA_ERR_STATE on such a context until ofea_set_mesh is called again.
 * ofea_sfem_assemble fills K = sum over domains k of V_k B(G)^T D B(G) +
 * (V_k/3) sum_in B(a_in)^T D B(a_in), G the smoothed gradients and a_in the
 * columns of the gradient-smoothing term scaled by (3 V_k / 4 pi)^(1/3) (the
 * reference's six-term form, expanded), into the CSC values (resident, and
 * copied to nz_val when not NULL).  Blocks sum their domains in ascending
 * order; no atomics. */
int ofea_sfem_setup(ofea_ctx *ctx);
int ofea_sfem_assemble(ofea_ctx *ctx, double E, double nu, double *nz_val);

/* ---- batched reanalysis (no counterpart in the reference) ----
 * Many design variants of one mesh: variant v has per-element moduli
 * E_var[v*ne + e] and shares nu, pattern and geometry.  Each variant's K_v is
 * what ofea_set_elem_E + ofea_assemble (fill_CSR_sparse_mat,
 * src/assemble.cpp:95-151, with a per-element E) would give, all nvar of them
 * from one kernel launch.  Single-type meshes (all C3D4 or all C3D8),
 * nvar <= 64.  E_var NULL reuses the resident moduli of the previous call
 * (same nvar).  nz_val (nvar x nnz, nullable) receives the values. */
int ofea_assemble_batch(ofea_ctx *ctx, int nvar, const double *E_var, double nu, double *nz_val);
/* nrhs load cases (1, 2, 4 or 8) per variant, solved together by FP64 PCG
 * with an FP32 3x3 block-Jacobi preconditioner; every read of K_v serves all
 * nrhs systems.  B and X are [variant][case][n]; iters and relres
 * (nvar x nrhs, nullable) get per-system iteration counts and true FP64
 * relative residuals.  Each solve is the one solution_api
 * (src/linear_solution.cpp:17-64) would do for (K_v, b).  Returns OFEA_OK, or
 * OFEA_NOT_CONVERGED when a system misses rtol within maxit. */
int ofea_solve_batch(ofea_ctx *ctx, int nvar, int nrhs, const double *B, double *X, double rtol, int maxit,
                     int *iters, double *relres);
/* As ofea_solve_batch with the preconditioner chosen: OFEA_PCG_BLOCK_JACOBI
 * (ofea_solve_batch's), or OFEA_PCG_TWO_LEVEL: block Jacobi plus the coarse
 * correction Z A_c^-1 Z^T r of ofea_solve's two-level methods, A_c = Z^T K_v Z
 * built on the device for every variant and shared by its nrhs systems (the
 * systems run in the two-level node order; B and X keep the caller's). */
int ofea_solve_batch_pc(ofea_ctx *ctx, int nvar, int nrhs, const double *B, double *X, double rtol, int maxit,
                        int precond, int *iters, double *relres);
/* Times the FP64 SpMV of the resident K with 4, 8, 16, 32 and 64 lanes per
 * node (times[5] in ms, nullable) and makes the fastest the lane-group size
 * of every later solve on this context. */
int ofea_autotune(ofea_ctx *ctx, int *best_group, float *times);
/* The multi-RHS solver picks its own lane-group size (16, 32 or 64 lanes per
 * node) per nrhs by timing K_v X on its first call; this reports the choice
 * and the time of one multi-RHS SpMV (ms).  OFEA_ERR_STATE before that call. */
int ofea_batch_tiles(ofea_ctx *ctx, int nrhs, int *group, float *ms);

/* ---- timing on the context stream (HIP events) ---- */
int ofea_timer_start(ofea_ctx *ctx);
int ofea_timer_stop(ofea_ctx *ctx, float *ms);
int ofea_sync(ofea_ctx *ctx);
/* HBM calibration: copies `bytes` between two device buffers `reps` times with
 * a plain vector-load/store kernel on the context stream; *ms is the average
 * per copy.  Used to confirm the achievable bandwidth and to scale the
 * profiler's FETCH_SIZE / WRITE_SIZE counters. */
int ofea_stream_copy(ofea_ctx *ctx, long long bytes, int reps, float *ms);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif
#endif
