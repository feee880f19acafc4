/* nuto_amd: MI355X (gfx950) element assembly for NuTo's SimpleAssembler path.
 *
 * Plain C ABI: no exceptions cross it, every call returns an int status (NUTO_OK == 0) and the
 * message of the last failure on a context is available from nuto_last_error().
 *
 * Units of data (all host arrays are row-major, all indices int32):
 *   cell group     n_cells coordinate elements (quad4 / hex8 / tet4) + per dof type the cell dof
 *                  numbering exactly as Cell::DofNumbering returns it (node-major, component-minor)
 *   dof vectors    one contiguous FP64 array per dof type, indexed by global dof number
 *   history        kappa(cellId, ip) at kappa[cellId + ip * history_ld] (Eigen column-major MatrixXd)
 *   matrix         monolithic CSR, rows [type0 dofs ; type1 dofs], cols [type0 , type1]
 *                  (ToEigen(K, {type0, type1}), DofMatrixSparseConvertEigen.h:40-67)
 *
 * Device pointers: dof vectors, history and outputs passed to the nuto_build_* calls live in
 * device memory (e.g. torch tensors on cuda:0); kernels run on the context's stream.
 */
#ifndef NUTO_AMD_H
#define NUTO_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define NUTO_OK 0
#define NUTO_ERR_ARGUMENT 1    /* bad argument (NuTo::Exception in the reference) */
#define NUTO_ERR_STATE 2       /* call out of order, e.g. no dof numbering (SimpleAssembler.cpp:16-23) */
#define NUTO_ERR_DEVICE 3      /* HIP runtime error */
#define NUTO_ERR_MEMORY 4      /* device or host allocation failed */
#define NUTO_ERR_UNSUPPORTED 5 /* combination not provided by these kernels */
#define NUTO_ERR_NO_CONVERGENCE 6 /* NewtonRaphson::NoConvergence (NewtonRaphson.h:12-20) */

/* element types (interpolations): InterpolationQuadLinear, InterpolationBrickLinear,
 * InterpolationTetrahedronLinear, InterpolationTetrahedronQuadratic, InterpolationBrickQuadratic */
#define NUTO_QUAD4 0
#define NUTO_HEX8 1
#define NUTO_TET4 2
#define NUTO_TET10 3
#define NUTO_HEX20 4

/* integrands */
#define NUTO_MOMENTUM_LINEAR_ELASTIC 0 /* Integrands::MomentumBalance<D> + Laws::LinearElastic<D> */
#define NUTO_GRADIENT_DAMAGE 1         /* Integrands::GradientDamage<3> (Constant interaction, exp. damage) */
#define NUTO_MOMENTUM_LOCAL_DAMAGE 2   /* MomentumBalance<3> + Laws::LocalIsotropicDamage<3> */

typedef struct nuto_ctx nuto_ctx;

/* Integrand parameters. Mirrors the constructor arguments of
 * Laws::LinearElastic(E, nu, planeState)                       LinearElastic.h:16
 * Laws::LinearElasticDamage(E, nu, damageApplication)          LinearElasticDamage.h:32
 * Constitutive::ModifiedMisesStrainNorm(nu, k)                 ModifiedMisesStrainNorm.h:38
 * Constitutive::DamageLawExponential(kappa0, beta, alpha)      DamageLawExponential.h:26
 * Integrands::GradientDamage(disp, eeq, c, ...)                GradientDamage.h:36 */
typedef struct nuto_params
{
    int kind;            /* NUTO_MOMENTUM_LINEAR_ELASTIC | NUTO_GRADIENT_DAMAGE | NUTO_MOMENTUM_LOCAL_DAMAGE */
    int dim;             /* 2 or 3 (2 only for linear elasticity) */
    double E, nu;
    int plane_state;     /* 0 PLANE_STRESS (reference default for LinearElastic<2>), 1 PLANE_STRAIN */
    double k, kappa0, beta, alpha, c;
    int damage_application; /* 0 FULL, 1 UNILATERAL (LinearElasticDamage.h:14-18) */
    /* gradient damage only: NonlocalInteraction (GradientDamage.h:150-184), the factor g(omega) of
     * the nonlocal term c g grad(eeq). 0 Constant (g = 1, the integrand's default), 1 Decreasing
     * with R = interaction_r, eta = interaction_eta (the reference's defaults are 0.005 and 5).
     * The node kernel, the adjoint and batches take Constant only; Decreasing runs on the
     * generic row-owner kernel. */
    int interaction;
    double interaction_r, interaction_eta;
} nuto_params;

/* ---- set-up helpers (host only, no device needed) ------------------------------------------- */

/* Structured mesh of [0,lx]x[0,ly](x[0,lz]): UnitMeshFem::CreateQuads / CreateBricks followed by
 * UnitMeshFem::Transform(x * L) (UnitMeshFem.cpp:67-125); coord_type NUTO_TET4 splits every brick
 * into six tetrahedra. jitter > 0 moves each node by a seeded uniform offset of up to
 * jitter * spacing per direction. Pass coords == conn == NULL to query the sizes. */
int nuto_mesh_structured(int coord_type, int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                         unsigned long long seed, int* n_nodes, int* n_cells, double* coords, int* conn);

/* First-touch dof layer: AddDofInterpolation (MeshFemDofConvert.cpp:164-210) numbered by
 * DofNumbering::Build (DofNumbering.cpp:32-58). Writes per-cell dof node ranks
 * (n_cells x nodes(dof_type)); the dof numbers are rank * ncomp + component.
 * Coordinate elements: quad4, hex8, tet4 under any dof element of their shape, or tet10 under
 * tet10 (isoparametric: the dof nodes are the coordinate nodes).
 * Pass cell_ranks == NULL to query n_dof_nodes only. */
int nuto_dof_layer(int coord_type, int n_cells, const int* coord_conn, int dof_type, int* cell_ranks,
                   int* n_dof_nodes);

/* ---- context ------------------------------------------------------------------------------------ */

/* Replaces constructing SimpleAssembler (SimpleAssembler.h:14-15); one context per device. */
int nuto_create(nuto_ctx** ctx, int device);
int nuto_destroy(nuto_ctx* ctx);
const char* nuto_last_error(const nuto_ctx* ctx);
/* Run subsequent kernels on this HIP stream (hipStream_t); NULL is the device's default (null)
 * stream. A new context uses a stream of its own until this is called. */
int nuto_set_stream(nuto_ctx* ctx, void* stream);
int nuto_synchronize(nuto_ctx* ctx);

/* The cell group: Group<CellInterface> of Cells built by CellStorage::AddCells over a MeshFem
 * (CellStorage.cpp:6-16, Cell.h:14-30). coords: n_coord_nodes x dim (node-major),
 * coord_conn: n_cells x nodes(coord_type), cell_ids: Cell::Id per cell (NULL: 0..n_cells-1).
 * coord_type: NUTO_QUAD4, NUTO_HEX8, NUTO_TET4 or NUTO_TET10. */
int nuto_set_cells(nuto_ctx* ctx, int coord_type, int n_cells, int n_coord_nodes, const double* coords,
                   const int* coord_conn, const int* cell_ids);

/* Dof type slot 0 (displacements) or 1 (nonlocal equivalent strains): its interpolation, number
 * of components, total dofs (numIndependent + numDependent, SimpleAssembler.cpp:158-166) and the
 * cell dof numbering n_cells x (nodes(dof_type) * ncomp) as ElementFem::GetDofNumbering returns it
 * (ElementFem.h:62-75). Replaces SimpleAssembler::SetDofInfo (SimpleAssembler.cpp:11-14). */
int nuto_set_dof_type(nuto_ctx* ctx, int slot, int dof_type, int ncomp, int n_dofs, const int* cell_dofs);

/* The integrand bound into the VectorFunction / MatrixFunction (CellInterface.h:19-20, Bind.h:11-22). */
int nuto_set_integrand(nuto_ctx* ctx, const nuto_params* params);

/* IntegrationTypeBase of the cells: kind 0 = IntegrationTypeTensorProduct<dim>(n, GAUSS), n = 1..10
 * (Legendre roots by Newton iteration, Legendre.cpp:23-55), kind 1 = IntegrationTypeTetrahedron
 * (order n = 1, 2) (IntegrationCompanion.cpp:18-52). Kernels exist for quad4 with 1, 2, 3, 4, 5 and
 * hex8 with 2, 3, 4 points per direction (hex20: 3); other valid rules give NUTO_ERR_UNSUPPORTED at
 * the first assembly. */
int nuto_set_rule(nuto_ctx* ctx, int kind, int n);

/* Leading dimension of the history matrix (rows of GradientDamage::mKappas, GradientDamage.h:137). */
int nuto_set_history_ld(nuto_ctx* ctx, int history_ld);

/* Per-cell material fields (E, f_t, G_f) for gradient damage: E replaces E, kappa0 = f_t / E and
 * beta = f_t / G_f (the DamageLawExponential(Material::Softening) mapping, DamageLawExponential.h:32-37).
 * Host arrays of n_cells, or NULL to return to the scalar parameters. */
int nuto_set_material_fields(nuto_ctx* ctx, const double* E, const double* ft, const double* gf);

/* The same fields from device memory (n_cells doubles each), e.g. tensors that carry gradients. */
int nuto_set_material_fields_device(nuto_ctx* ctx, const double* E, const double* ft, const double* gf);

/* Sparsity of BuildMatrix's result (setFromTriplets pattern: every dof pair of every cell, explicit
 * zeros kept; SimpleAssembler.cpp:100-156), built once and kept on the device. The host part runs on
 * up to 16 threads for meshes of 16,384 dof nodes or more (the calling thread waits for them). */
int nuto_build_pattern(nuto_ctx* ctx, long long* nnz, int* n_rows);
int nuto_copy_pattern(nuto_ctx* ctx, int* row_ptr, int* col);
int nuto_pattern_device(nuto_ctx* ctx, const int** row_ptr, const int** col);

/* Block (slot_i, slot_j) of the matrix: K(dofI, dofJ) of NuTo's DofMatrixSparse, a map of one sparse
 * matrix per dof type pair (DofMatrixContainer.h:16-24; BuildMatrix fills it per (dofI, dofJ),
 * SimpleAssembler.cpp:105-155), whose monolithic export is ToEigen (DofMatrixSparseConvertEigen.h:
 * 40-67). The block is CSR over the rows of slot_i (dof numbers of that type), columns numbered
 * within slot_j. In every row the block's entries are one run of the monolithic row (slot 0's
 * columns come before slot 1's): block entry k of row r is K[run_begin[r] + k - row_ptr[r]].
 * nnz: the block's nonzeros; row_ptr (rows of slot_i + 1), col (nnz) and run_begin (rows of slot_i):
 * host arrays, each may be NULL. After nuto_build_pattern. */
int nuto_block_pattern(nuto_ctx* ctx, int slot_i, int slot_j, long long* nnz, int* row_ptr, int* col,
                       long long* run_begin);
/* The values of block (slot_i, slot_j) gathered from K (as nuto_build_matrix writes it) into
 * `values` in the block's CSR order; device arrays, values: the block's nonzeros. */
int nuto_block_values(nuto_ctx* ctx, int slot_i, int slot_j, const double* K, double* values);

/* SimpleAssembler::BuildVector (SimpleAssembler.cpp:38-65): R (length n_rows) from the dof vectors
 * u (slot 0) and e (slot 1, NULL for single-field integrands) and history kappa (NULL if unused). */
int nuto_build_vector(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, double* R);

/* SimpleAssembler::BuildMatrix (SimpleAssembler.cpp:100-156): K values (length nnz) in the pattern.
 * 3D momentum-balance groups (linear elasticity, local damage) assemble in two phases through a
 * staging buffer the context keeps (8 * n_cells * ndof_per_cell^2 bytes: 28.8 KB per hex20 cell,
 * 4.6 KB per hex8 cell); NUTO_AMD_TWO_PHASE=0 in the environment selects the single-pass kernel. */
int nuto_build_matrix(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, double* K);

/* BuildVector and BuildMatrix of the same state in one pass over the cells. */
int nuto_build_vector_matrix(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, double* R,
                             double* K);

/* No counterpart in the reference: nuto_build_vector_matrix for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc). It calls no
 * allocating routine: the hex8 gradient-damage node kernel's pre-pass records go to a static 32 MiB
 * device scratch array (one per device), so it takes at most 29127 cells (8 IPs each; more:
 * NUTO_ERR_ARGUMENT). Needs the node tables from set-up: the integrand and the rule set before
 * nuto_build_pattern, or one nuto_build_vector_matrix call (NUTO_ERR_STATE otherwise); other
 * integrands and elements: NUTO_ERR_UNSUPPORTED. Every context on one device shares that device's
 * scratch: calls that overlap in time on one device (other streams, other host threads) race. */
int nuto_build_vector_matrix_noalloc(nuto_ctx* ctx, const double* u, const double* e, const double* kappa,
                                     double* R, double* K);

/* TimeDependentProblem::UpdateHistory -> GradientDamage::Update (TimeDependentProblem.cpp:78-85,
 * GradientDamage.h:122-125) or EvolutionImplicit::Update (LocalIsotropicDamage.h:130-133):
 * kappa(cell, ip) = max(kappa(cell, ip), eeq(ip)), in place. */
int nuto_update_history(nuto_ctx* ctx, const double* u, const double* e, double* kappa);

/* SimpleAssembler::BuildDiagonallyLumpedMatrix (SimpleAssembler.cpp:67-98) with the integrand
 * DynamicMomentumBalance::Hessian2 = N^T N rho (DynamicMomentumBalance.h:22-30) on dof type slot 0
 * (the displacements, dim components): per cell the consistent mass integrated with the context's
 * rule (Cell.h:105-119), its diagonal scaled by (sum of all entries / sum of the diagonal) /
 * number of components (DofType::GetNum), summed over the cells into M (device, length
 * n_dofs of slot 0, every component of a node holding the same value). Needs the pattern. */
int nuto_build_lumped_mass(nuto_ctx* ctx, double rho, double* M);

/* Fields at the integration points, Cell::Eval (Cell.h:80-93) of the context's integrand, for
 * every cell (group order) and IP (rule order): out[(cell * n_ip + ip) * n_comp + comp], device.
 *   NUTO_IP_STRAIN     B u, engineering Voigt [xx,yy,xy] / [xx,yy,zz,yz,zx,xy]
 *                      (Nabla::Strain, DifferentialOperators.h:24-86); n_comp 3 (2D) or 6 (3D)
 *   NUTO_IP_STRESS     the law's stress: LinearElastic (LinearElastic.h:76-95), LinearElasticDamage
 *                      at omega(kappa) (LinearElasticDamage.h:65-72); n_comp as the strain
 *   NUTO_IP_KAPPA      GradientDamage::Kappa = max(history, eeq) (GradientDamage.h:126-129), local
 *                      damage max(history, eps_eq) (LocalIsotropicDamage.h:52-56); 1
 *   NUTO_IP_DAMAGE     DamageLawExponential::Damage(kappa) (DamageLawExponential.h:39-45); 1
 *   NUTO_IP_EQ_STRAIN  ModifiedMisesStrainNorm::Value(strain) (ModifiedMisesStrainNorm.h:70-84); 1
 * The damage fields need a damage integrand. Inputs as nuto_build_vector. */
#define NUTO_IP_STRAIN 0
#define NUTO_IP_STRESS 1
#define NUTO_IP_KAPPA 2
#define NUTO_IP_DAMAGE 3
#define NUTO_IP_EQ_STRAIN 4
int nuto_eval_ip(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, int field, double* out);

/* ---- constraints and the constrained linear solve ------------------------------------------------
 * Constraints::Add (Constraints.cpp:7-21) for dof type slot `slot`: n_eq equations (Equation.h,
 * Term.h), equation i reading  1 * x[dep[i]] + sum_t coef[t] * x[term_dof[t]] = rhs_i(time)  over
 * t in [term_ptr[i], term_ptr[i + 1]), dof numbers of that slot's DofType. Replaces the slot's
 * equations (n_eq = 0 removes them; setting the slot's dof type also does). The checks of
 * Constraints::TermChecker (Constraints.h:112-129) and GetJKNumbering (Constraints.cpp:62-71)
 * apply: every dependent dof is numbered and in range, appears in no other equation, and no term
 * is a dependent dof. Host arrays. */
int nuto_set_constraints(nuto_ctx* ctx, int slot, int n_eq, const int* dep, const int* term_ptr, const int* term_dof,
                         const double* coef);

#define NUTO_SOLVER_CG 0       /* EigenConjugateGradient (EigenSparseSolve.cpp:44), symmetric K */
#define NUTO_SOLVER_BICGSTAB 1 /* EigenBiCGSTAB (EigenSparseSolve.cpp:48), general K */
typedef struct nuto_solver_options
{
    int method;           /* NUTO_SOLVER_CG or NUTO_SOLVER_BICGSTAB; diagonal preconditioner, x0 = 0 */
    double tolerance;     /* stop at |C^T f - A u| <= tolerance |C^T f|; <= 0: machine epsilon (Eigen's) */
    int max_iterations;   /* <= 0: 2 x the independent dofs (Eigen's default) */
} nuto_solver_options;
typedef struct nuto_solve_info
{
    int iterations;            /* Eigen's count: the iterations before the converging one */
    double relative_residual;  /* |C^T f - A u| / |C^T f| at the end */
    int independent_dofs;      /* size of the reduced system */
} nuto_solve_info;

/* ConstrainedSystemSolver::Solve -> NuTo::Solve (Solve.cpp:8-32): with C the unit constraint
 * matrix of all slots (Constraints::BuildUnitConstraintMatrix, Constraints.cpp:89-137, block
 * diagonal), x = C u with (C^T K C) u = C^T f. K: values in the context's pattern (as
 * nuto_build_matrix writes them), f and x: length n_rows; device arrays. A = C^T K C is applied
 * as three sparse products, never formed; K stays on the device. options NULL: BiCGSTAB,
 * machine-epsilon tolerance, 2 n iterations. info may be NULL. The solver's vectors, about
 * 4 n_rows + 11 n_independent doubles of device memory, are allocated by the context's first solve
 * and kept for the next ones (grown when needed) until nuto_destroy. */
int nuto_solve(nuto_ctx* ctx, const double* K, const double* f, double* x, const nuto_solver_options* options,
               nuto_solve_info* info);

/* ConstrainedSystemSolver::SolveTrialState -> NuTo::SolveTrialState (Solve.cpp:34-78): with
 * delta = rhs(newTime) - rhs(oldTime) at the dependent dofs (GetSparseGlobalRhs, zero elsewhere),
 * x = C (C^T K C)^-1 C^T (f + K delta) - delta (the negative increment). rhs_old / rhs_new: the
 * equations' right-hand sides at the two times, slot 0's equations then slot 1's (device arrays;
 * rhs_old NULL means zeros). */
int nuto_solve_trial_state(nuto_ctx* ctx, const double* K, const double* f, const double* rhs_old,
                           const double* rhs_new, double* x, const nuto_solver_options* options,
                           nuto_solve_info* info);

/* One load step of QuasistaticSolver::DoStep (QuasistaticSolver.cpp:116-147) on the device: the
 * trial state x - SolveTrialState(K(x), R(x), rhs_old, rhs_new) from the last converged state
 * (TrialState, :51-66), then NewtonRaphson::Solve (NewtonRaphson.h:62-94) with LineSearch
 * (LineSearch.h:39-63: x = x0 - alpha dx, alpha halved up to max_line_search_steps times, accepted
 * once |R| < (1 - alpha) |R0|) or, with max_line_search_steps == 0, NoLineSearch; the residual norm
 * is |C^T R| (QuasistaticSolver::Norm, :84-90). On convergence the history is updated
 * (UpdateHistory, :79-82) and x is overwritten; otherwise x and kappa are left unchanged and the
 * call returns NUTO_ERR_NO_CONVERGENCE, which AdaptiveSolve (AdaptiveSolve.cpp:18-67) answers by
 * halving the time step. x: the monolithic dof vector [u; e] (n_rows, device); kappa: the history
 * (device, NULL for linear elasticity); rhs_old / rhs_new as for nuto_solve_trial_state. */
typedef struct nuto_newton_options
{
    double tolerance;          /* |C^T R| below which the step has converged (QuasistaticSolver::mTolerance, 1e-10 if <= 0) */
    int max_iterations;        /* Newton iterations (DoStep passes 6; 6 if <= 0) */
    int max_line_search_steps; /* LineSearch steps (6 by default in DoStep); 0: NoLineSearch; < 0: 6 */
    nuto_solver_options linear;/* the solver of every linear system */
} nuto_newton_options;
typedef struct nuto_newton_info
{
    int iterations;            /* DoStep's return value: Newton iterations to convergence */
    int linear_iterations;     /* Krylov iterations over all linear solves of the step */
    int residual_evaluations;  /* assemblies of R (trial state excluded) */
    double residual_norm;      /* |C^T R| at the returned state */
} nuto_newton_info;
int nuto_quasistatic_step(nuto_ctx* ctx, double* x, double* kappa, const double* rhs_old, const double* rhs_new,
                          const nuto_newton_options* options, nuto_newton_info* info);

/* Adjoint of BuildVector for the gradient-damage integrand (no counterpart in the reference, which
 * has no derivatives w.r.t. material parameters; its K^T lambda is the transpose of BuildMatrix,
 * SimpleAssembler.cpp:100-156 with GradientDamage::Hessian0, GradientDamage.h:92-120).
 * Given the cotangent lambda (n_rows, layout of R):
 *   d_state (n_rows)       = K^T lambda = (dR / d(u, eeq))^T lambda, computed without forming K;
 *   d_theta (3 x n_cells)  = (dR / d theta)^T lambda for theta = (E, f_t, G_f) per cell, stored
 *                            [E row; f_t row; G_f row]; NULL skips it. Without material fields the
 *                            scalar parameters are the point (f_t = kappa0 E, G_f = f_t / beta).
 * kappa (the history) is a constant. All pointers are device pointers. Hex8 cells and dofs, 2x2x2
 * Gauss points, any number of cells and neighbours per node. */
int nuto_vjp(nuto_ctx* ctx, const double* u, const double* e, const double* kappa, const double* lambda,
             double* d_state, double* d_theta);

/* ---- batches of independent meshes -------------------------------------------------------------
 * No counterpart in the reference, which assembles one cell group per SimpleAssembler call
 * (SimpleAssembler.cpp:38-156); a batch runs that call for many meshes at once. Every mesh is one
 * cell group with its own dof numbering (as nuto_set_cells + nuto_set_dof_type would take it) and
 * the batch's integrand; its rule is the default of its element pair (tet4/tet4: tetrahedron order
 * 1, tet4/tet10: order 2, hex8/hex8: 2x2x2 Gauss, hex8/hex20: 3x3x3 Gauss).
 *
 * Layout: the batch's u, eeq, R, K, history and d_theta arrays hold all meshes; nuto_batch_mesh_info
 * gives each mesh's offsets:
 *   info[0] group       info[1] dof nodes   info[2] cells          info[3] IPs per cell
 *   info[4] u offset    info[5] eeq offset (-1)                    info[6] R offset (displacement rows)
 *   info[7] R offset (eeq rows, -1)         info[8] rows of the mesh
 *   info[9] K offset (displacement rows)    info[10] their nonzeros
 *   info[11] K offset (eeq rows, -1)        info[12] their nonzeros
 *   info[13] history offset of (cell 0, ip 0); (cell c, ip i) at info[13] + c + i * info[14]
 *   info[15] d_theta offset of cell 0 (E); f_t and G_f follow at strides of info[16]
 * Within its slices a mesh's R and K are exactly nuto_build_vector / nuto_build_matrix of that mesh
 * (monolithic CSR of nuto_batch_copy_pattern). */
typedef struct nuto_batch nuto_batch;
int nuto_batch_create(nuto_batch** batch, int device);
int nuto_batch_destroy(nuto_batch* batch);
const char* nuto_batch_last_error(const nuto_batch* batch);
int nuto_batch_set_stream(nuto_batch* batch, void* stream); /* non-NULL for graph replay (mode -1) */
int nuto_batch_synchronize(nuto_batch* batch);
int nuto_batch_set_integrand(nuto_batch* batch, const nuto_params* params); /* dim 3 integrands */
/* cell_ranks: n_cells x nodes(dof_type) dof node ranks (nuto_dof_layer); mesh receives its index */
int nuto_batch_add_mesh(nuto_batch* batch, int coord_type, int dof_type, int n_cells, int n_coord_nodes,
                        const double* coords, const int* coord_conn, int n_dof_nodes, const int* cell_ranks,
                        int* mesh);
/* The meshes of one element pair, in the order added, fill groups (super-meshes with one CSR, one
 * launch each) of at most max_dof_nodes dof nodes; a larger mesh forms a group of its own. 0 (the
 * default): one group per element pair, except gradient damage, whose groups hold at most
 * (2^31 - 1) / (16 * 32) nodes (the node kernel takes at most 32 neighbours per node). In every
 * case a group closes before its CSR would exceed 2^31 - 1 nonzeros (int32 offsets). Takes effect
 * at the next nuto_batch_build. */
int nuto_batch_set_group_limit(nuto_batch* batch, long long max_dof_nodes);
/* Staging of the two-phase launches (mode 5; 3D momentum balance; mode 6 stages one mesh at a
 * time): each group stages the cell
 * matrices of consecutive meshes up to `bytes` at a time (8 (nd^2 + nd) bytes per cell, nd = 3 x
 * nodes per cell; a larger mesh alone), so that any batch size runs in bounded memory. 0: the whole
 * group at once. Default: the whole group when it can be allocated, else 1 GiB; nuto_batch_autotune
 * chooses per group among the whole group, 1 GiB, 256 MiB and 64 MiB. After nuto_batch_build. */
int nuto_batch_set_stage_bytes(nuto_batch* batch, long long bytes);
/* patterns of all meshes, once (host work on up to 16 threads, as nuto_build_pattern) */
int nuto_batch_build(nuto_batch* batch);
/* sizes[6]: lengths of u, eeq, R, K, history and d_theta */
int nuto_batch_sizes(nuto_batch* batch, long long* sizes);
int nuto_batch_mesh_info(nuto_batch* batch, int mesh, long long* info /* 17 */);
int nuto_batch_copy_pattern(nuto_batch* batch, int mesh, int* row_ptr, int* col);
/* info[7] of a group: dof element type, meshes, dof nodes, cells, cell-matrix kernel of its
 * two-phase launch (0 by columns, 1 / 2 Gram form with one / two row nodes per lane, 3 Gram form
 * on the FP64 matrix cores, -1 the element's default until nuto_batch_autotune times each),
 * number of groups, cells per staging
 * chunk (0: the whole group, -1: not chosen yet) */
int nuto_batch_group_info(nuto_batch* batch, int group, long long* info /* 7 */);
/* per-cell (E, f_t, G_f) over all cells, meshes in the order they were added (host arrays) */
int nuto_batch_set_material_fields(nuto_batch* batch, const double* E, const double* ft, const double* gf);
/* R and K of every mesh. mode -1: the tuned launch (mode 3 until tuned) replayed from a HIP graph; 0..2: one fused launch
 * over all meshes with 64 / 128 / 256 nodes per workgroup; 3: one launch per group (the meshes of
 * one element type); 4: one launch per mesh (the row-owner kernel); 5: two-phase (cell matrices,
 * then the row gather) per group; 6: two-phase per mesh (what separate
 * nuto_build_vector_matrix calls do). Gradient damage maps 0..2 and 5 to 3, and 6 to 4. */
int nuto_batch_assemble(nuto_batch* batch, const double* u, const double* e, const double* kappa, double* R,
                        double* K, int mode);
/* Times modes 0..6 (ms[7], -1 where not applicable) and makes the fastest of 0..3 and 5 (the batched
 * forms) the tuned mode. */
int nuto_batch_autotune(nuto_batch* batch, const double* u, const double* e, const double* kappa, double* R,
                        double* K, double* ms, int* best);
/* Gradient damage: nuto_vjp for every mesh (d_state layout of R, d_theta layout of info[15]). */
int nuto_batch_vjp(nuto_batch* batch, const double* u, const double* e, const double* kappa, const double* lambda,
                   double* d_state, double* d_theta);
/* Gradient damage, loss L = 0.5 ||R - r_obs||^2 of every mesh in one graph replay: R and K, then
 * lambda = R - r_obs, then d_state = K^T lambda and d_theta = dR/dtheta^T lambda. */
int nuto_batch_residual_adjoint(nuto_batch* batch, const double* u, const double* e, const double* kappa,
                                const double* r_obs, double* R, double* K, double* d_state, double* d_theta);

/* ---- device memory helpers for hosts without a HIP toolchain -------------------------------- */
int nuto_device_alloc(nuto_ctx* ctx, void** ptr, unsigned long long bytes);
int nuto_device_free(nuto_ctx* ctx, void* ptr);
int nuto_copy_to_device(nuto_ctx* ctx, void* dst, const void* src, unsigned long long bytes);
int nuto_copy_to_host(nuto_ctx* ctx, void* dst, const void* src, unsigned long long bytes);

/* ---- mesh input: gmsh MSH 2 files (host only) ------------------------------------------------
 * MeshGmsh (MeshGmsh.h:17-60, MeshGmsh.cpp:440-550): ASCII and binary files, $PhysicalNames,
 * $Nodes and $Elements; nodes and elements in file order, point elements dropped, node orders of
 * gmsh types 8, 10 and 12 changed to the interpolations' (MeshGmsh.cpp:398-436), elements with
 * det J <= 0 at the natural origin rejected (MeshGmsh.cpp:85-91). Mesh dimension: 3 if any
 * z != 0, else 2 if any y != 0, else 1 (MeshGmsh.cpp:130-147). */
typedef struct nuto_gmsh nuto_gmsh;
/* Reads path into *mesh. *mesh is set also on failure: read the message with
 * nuto_gmsh_last_error, then free it. */
int nuto_gmsh_read(const char* path, nuto_gmsh** mesh);
void nuto_gmsh_free(nuto_gmsh* mesh);
const char* nuto_gmsh_last_error(const nuto_gmsh* mesh);
int nuto_gmsh_sizes(const nuto_gmsh* mesh, int* dim, int* n_nodes, int* n_elements, int* conn_size);
/* coords: n_nodes x dim */
int nuto_gmsh_nodes(const nuto_gmsh* mesh, double* coords);
/* gmsh element type, CSR offsets (n_elements + 1) into conn (0-based node indices), first tag;
 * any pointer may be NULL */
int nuto_gmsh_elements(const nuto_gmsh* mesh, int* gmsh_type, int* offsets, int* conn, int* physical);
/* MeshGmsh::GetPhysicalGroup(int) (MeshGmsh.cpp:26-32): element indices of a physical group;
 * elements may be NULL to query n. An unknown id is an error. */
int nuto_gmsh_group(nuto_gmsh* mesh, int id, int* n, int* elements);
/* MeshGmsh::GetPhysicalGroup(std::string) (MeshGmsh.cpp:34-43): id of a named group, name
 * matched case-insensitively. */
int nuto_gmsh_group_id(nuto_gmsh* mesh, const char* name, int* id);
/* Connectivity (n_cells x nodes per cell) of the elements of a group (group < 0: all elements)
 * that are of the gmsh type of coord_type (quad4: 3, hex8: 5, tet4: 4, tet10: 11), ready for
 * nuto_set_mesh's conn; conn may be NULL to query n_cells. */
int nuto_gmsh_cells(nuto_gmsh* mesh, int group, int coord_type, int* n_cells, int* conn);

/* ---- Morton (Z-order) codes (NuTo::MortonOrder, MortonOrder.h:11-107) ------------------------
 * 3D: 10 bits per axis, x in bit 0, y in bit 1, z in bit 2 (EncodeMorton3D, DecodeMorton3X/Y/Z);
 * 2D: 16 bits per axis, x in bit 0 (Part1By1, DecodeMorton2X/Y). Inputs are cut to those bits. */
unsigned nuto_morton_encode3d(unsigned x, unsigned y, unsigned z);
unsigned nuto_morton_encode2d(unsigned x, unsigned y);
void nuto_morton_decode3d(unsigned code, unsigned* x, unsigned* y, unsigned* z);
void nuto_morton_decode2d(unsigned code, unsigned* x, unsigned* y);
/* Neighbor3D / NegNeighbor3D: the code of the cell shifted by (dx, dy, dz); 0xffffffff when a
 * coordinate would turn negative. */
unsigned nuto_morton_neighbor3d(unsigned key, int dx, int dy, int dz);
/* The permutation that sorts n points (n x dim, dim 1..3) along a Morton curve over their bounding
 * box, quantised to 2^10 steps per axis in 3D and 2^16 in 1D and 2D; ties keep the input order. */
int nuto_morton_order(int n, int dim, const double* coords, int* order);

#ifdef __cplusplus
}
#endif

#endif /* NUTO_AMD_H */
