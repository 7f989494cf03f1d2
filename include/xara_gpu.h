/*
 * xara_gpu.h -- C-ABI of the MI355X state-determination library.
 *
 * One handle (xg_domain) holds one finite-element model resident on one GPU:
 * nodes, fixities, materials, sections, elements, nodal reference loads, the
 * equation numbering, the fixed-pattern CSR system and all trial/committed
 * material state in structure-of-arrays form.  Every call returns an int
 * status: 0 = success, negative = failure (the reference convention of
 * Domain::update / LinearSOE::solve, Domain.cpp:1996-2003), and the last error
 * message is available from xg_last_error().
 *
 * Arrays are plain host pointers with explicit sizes.  Node-indexed arrays are
 * in ascending node-tag order (the order of the reference's node map and of
 * PlainHandler's DOF groups, PlainHandler.cpp:60-120); element-indexed arrays
 * are in ascending element-tag order (FE_Element order, PlainHandler.cpp:212).
 *
 * Each declaration names the reference interface it replaces.  The path is
 * xara's nonlinear state determination: Element::update / getTangentStiff /
 * getResistingForce of stdBrick, FourNodeQuad, DispBeamColumn2d/3d over
 * J2Plasticity, Steel02, Concrete02 and fiber sections, assembled into a CSR
 * LinearSOE and solved by preconditioned CG on the GPU.
 */
#ifndef XARA_GPU_H
#define XARA_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xg_domain xg_domain;

/* ---- status codes (SolutionAlgorithm codes, NewtonRaphson.cpp:139-219) ---- */
#define XG_OK                 0
#define XG_ERR_ARG           -1   /* bad argument / unknown tag                */
#define XG_ERR_STATE         -2   /* call out of order (e.g. before setup)     */
#define XG_ERR_DEVICE        -3   /* HIP runtime failure                       */
#define XG_ERR_MATERIAL      -4   /* return map failed (J2Plasticity.cpp:309)  */
#define XG_BAD_LINEAR_SOLVE  -5   /* LinearSOE::solve < 0                      */
#define XG_TEST_FAILED       -6   /* ConvergenceTest::Failure                  */
#define XG_BAD_STEP          -7   /* integrator newStep/update failed          */

/* ---- enumerations --------------------------------------------------------- */
/* system (soe_table, SRC/runtime/commands/analysis/solver.hpp:113) */
#define XG_SYSTEM_PCG_JACOBI       1  /* CSR, Jacobi-preconditioned CG           */
#define XG_SYSTEM_PCG_BLOCKJACOBI  2  /* CSR, nodal block-Jacobi CG             */
#define XG_SYSTEM_DENSE            3  /* dense LU (small models, testing)       */
/* numberer (SRC/runtime/commands/analysis/numberer.cpp) */
#define XG_NUMBERER_RCM    1          /* DOF_Numberer(RCM(false)), the default  */
#define XG_NUMBERER_PLAIN  2          /* PlainNumberer                          */
/* test (SRC/analysis/criteria) */
#define XG_TEST_NORM_UNBALANCE  1     /* CTestNormUnbalance.cpp:81              */
#define XG_TEST_NORM_DISP_INCR  2     /* CTestNormDispIncr.cpp:78               */
#define XG_TEST_ENERGY_INCR     3     /* CTestEnergyIncr.cpp:85                 */
/* per-Gauss-point / per-section outputs for xg_get_point_data */
#define XG_OUT_STRESS      1          /* NDMaterial::getStress / section s      */
#define XG_OUT_TANGENT     2          /* NDMaterial::getTangent / section ks    */
#define XG_OUT_STRAIN      3          /* NDMaterial::getStrain / section e      */
#define XG_OUT_PLASTIC     4          /* J2: eps_p(6) + xi trial                */
/* operations for xg_time_op */
#define XG_OP_UPDATE       1
#define XG_OP_TANGENT      2
#define XG_OP_UNBALANCE    3
#define XG_OP_SPMV         4
#define XG_OP_COPY         5          /* 256 MiB device copy: HBM counter calibration */

typedef struct xg_counters {
  long long updates;          /* Domain::update calls                         */
  long long material_points;  /* Gauss points + fibers updated, cumulative    */
  long long newton_iters;     /* NewtonRaphson iterations                     */
  long long solves;           /* LinearSOE::solve calls                        */
  long long cg_iters;         /* cumulative PCG iterations                    */
  double    last_cg_relres;   /* last PCG relative residual                   */
  double    last_norm;        /* last convergence-test norm                   */
  double    time_update_ms;   /* device time, cumulative                      */
  double    time_tangent_ms;
  double    time_unbalance_ms;
  double    time_solve_ms;
  long long band_bw;          /* half bandwidth of the band-Cholesky
                                 preconditioner, -1 when it is not used    */
  long long band_fallbacks;   /* matrices whose tiled band factor met a pivot
                                 that is not positive and were solved with
                                 block Jacobi instead                      */
  long long band_reuses;      /* assembled matrices equal bit for bit to the
                                 last one factored, whose band factor was
                                 kept instead of formed again              */
} xg_counters;

/* ---- lifetime -------------------------------------------------------------- */
/* `model basic -ndm ndm -ndf ndf` (BasicModelBuilder, SRC/runtime/runtime/BasicModelBuilder.h) */
int  xg_create(int ndm, int ndf, int device, xg_domain** out);
int  xg_destroy(xg_domain* d);
const char* xg_last_error(void);
int  xg_device_name(char* buf, int len);

/* ---- model definition (Tcl commands; Domain::addNode/addElement) ---------- */
/* `node tag x y <z>`  (SRC/domain/node/Node.cpp constructor)                  */
int  xg_nodes(xg_domain* d, int n, const int* tags, const double* crds /* n*ndm */);
/* `fix tag f1 .. fndf` -> SP_Constraints seen by PlainHandler.cpp:80-110     */
int  xg_fix(xg_domain* d, int n, const int* tags, const int* flags /* n*ndf */);
/* `mass tag m1 .. mndf` (Node::setMass)                                       */
int  xg_mass(xg_domain* d, int n, const int* tags, const double* m /* n*ndf */);
/* `nDMaterial J2Plasticity tag K G sig0 sigInf delta H <eta>`
   (plastic.cpp:616 -> J2Plasticity.cpp:83-110)                                */
int  xg_nd_j2plasticity(xg_domain* d, int tag, double K, double G, double sig0,
                        double sigInf, double delta, double H, double eta);
/* `uniaxialMaterial Steel02 tag fy E b R0 cR1 cR2 a1 a2 a3 a4 sigInit`
   (fedeas.cpp:578 -> Steel02.cpp:44-54)                                      */
int  xg_uniaxial_steel02(xg_domain* d, int tag, double fy, double E0, double b,
                         double R0, double cR1, double cR2, double a1, double a2,
                         double a3, double a4, double sigInit);
/* `uniaxialMaterial Concrete02 tag fpc epsc0 fpcu epscu rat ft Ets`
   (Concrete02.cpp:52-112)                                                    */
int  xg_uniaxial_concrete02(xg_domain* d, int tag, double fc, double epsc0,
                            double fcu, double epscu, double rat, double ft,
                            double Ets);
/* `uniaxialMaterial Elastic tag E` (ElasticMaterial), used for -GJ torsion    */
int  xg_uniaxial_elastic(xg_domain* d, int tag, double E);
/* `section Fiber tag <-GJ v> {fiber y z A mat ...}`
   (section.cpp:443/497 -> FiberSection2d.cpp:99 / FiberSection3d.cpp:294).
   For a 2D section z is ignored (may be NULL).  GJ < 0: no torsion.
   compute_centroid = 0 is `-noCentroid`.                                     */
int  xg_section_fiber(xg_domain* d, int tag, int nfib, const int* mat_tags,
                      const double* y, const double* z, const double* A,
                      double GJ, int compute_centroid);
/* `geomTransf Linear tag <vecxz>` (geomTransf.cpp:350/468 ->
   LinearCrdTransf2d/3d).  vecxz is ignored for ndm = 2.                      */
int  xg_geom_transf_linear(xg_domain* d, int tag, const double* vecxz);
/* `element stdBrick tag n1..n8 mat <b1 b2 b3>` (brick.cpp:25,126 -> Brick.cpp:77) */
int  xg_elements_brick(xg_domain* d, int n, const int* tags, const int* conn /* n*8 */,
                       int mat_tag, double b1, double b2, double b3);
/* `element quad tag n1..n4 thk PlaneStrain mat <pressure rho b1 b2>`
   (plane.cpp:99,249,395 -> FourNodeQuad.cpp:36)                              */
int  xg_elements_quad(xg_domain* d, int n, const int* tags, const int* conn /* n*4 */,
                      double thk, int mat_tag, double pressure, double rho,
                      double b1, double b2);
/* `element dispBeamColumn tag i j transf "Legendre" sec nip <-mass rho>`
   (frames.cpp:449,906 -> DispBeamColumn2d.cpp:48 / DispBeamColumn3d.cpp) */
int  xg_elements_dispbeam(xg_domain* d, int n, const int* tags, const int* conn /* n*2 */,
                          int transf_tag, int sec_tag, int nip, double rho);
/* `element forceBeamColumn tag i j transf "Lobatto" sec nip <-iter maxIters tol>`
   (frames.cpp:200-205,544-545,884-889 -> ForceBeamColumn2d.cpp:113 /
   ForceBeamColumn3d.cpp; LobattoBeamIntegration, 2 <= nip <= 10, no mass,
   no element loads; in 3D the section needs -GJ)                             */
int  xg_elements_forcebeam(xg_domain* d, int n, const int* tags, const int* conn /* n*2 */,
                           int transf_tag, int sec_tag, int nip, int max_iter, double tol);
/* `pattern Plain 1 Linear { load tag P1..Pndf }` (NodalLoad; LoadPattern
   with a Linear series, applied as lambda*P by Domain::applyLoad,
   Domain.cpp:1765)                                                           */
int  xg_nodal_loads(xg_domain* d, int n, const int* tags, const double* P /* n*ndf */);
/* `pattern UniformExcitation tag dir -accel {Path -dt dt -values ...}`
   (UniformExcitation.cpp:132, PathSeries.cpp:334).  dir is 0-based.         */
int  xg_uniform_excitation(xg_domain* d, int dir, int npts, double dt,
                           const double* accel, double factor);
/* `rayleigh alphaM betaK betaKinit betaKcomm` (Domain::setRayleighDampingFactors,
   Domain.cpp:1858-1873; Element::getDamp / getRayleighDampingForces,
   Element.cpp:181-200,283-323): betaK on the current tangent, betaKcomm on
   the tangent of the last commit; betaKinit must be 0 (XG_ERR_ARG).  The
   ensemble and the adjoint take alphaM only. */
int  xg_rayleigh(xg_domain* d, double alphaM, double betaK, double betaKinit,
                 double betaKcomm);

/* ---- analysis objects (BasicAnalysisBuilder::set, BasicAnalysisBuilder.cpp:781-822) */
/* `system ...` (solver.hpp:113 soe_table)                                    */
int  xg_system(xg_domain* d, int kind, double rtol, int maxit);
/* SpMV layout of the PCG systems.  No counterpart in the reference (its
   sparse solvers take the CSR arrays as they are): a tuning knob.
   XG_SPMV_AUTO uses node-blocked column indices (one index per neighbour
   node instead of one per nonzero, same values and pattern) when every node
   has all its DOFs free or none, and plain CSR otherwise; lanes = threads per
   matrix row (4, 8, 16, 32 or 64; 0 = chosen from the mean row length).
   XG_SPMV_SYM (2 or 3 DOFs per node, node-blocked pattern): the products
   read a block-symmetric copy of the matrix (upper node blocks only,
   refreshed after each assembly), for symmetric tangents; lanes = threads
   per node block row.  Under XG_SPMV_AUTO the PCG's own products use it too
   (the PCG assumes a symmetric matrix) while xg_spmv keeps the node-blocked
   CSR product.  Falls back to the node-blocked or plain CSR product where
   the layout does not apply.  xg_time_op(XG_OP_SPMV) times the PCG's.      */
#define XG_SPMV_AUTO 0
#define XG_SPMV_CSR  1
#define XG_SPMV_SYM  2
int  xg_spmv_config(xg_domain* d, int format, int lanes);
/* PCG issue pattern.  No counterpart in the reference: a tuning knob.  The
   host issues `chunk` iterations (even, 2..1024; default 16) before its
   first read of the convergence flag, then twice as many before each next
   read, up to 8 x chunk (iterations after convergence return at once).
   XG_CG_GRAPH: a full chunk is replayed from a captured hipGraph of its
   kernel launches.  XG_CG_FUSE: on systems small enough (at most 1024
   partial sums per reduction) an iteration is two launches instead of
   five: the product forms the new direction on the fly (beta reduced by
   every workgroup) and the update reduces alpha itself; the iterates are
   the same bits with and without either.
   XG_CG_SMALL: systems of at most 1,024 equations and 32,768 nonzeros are
   solved whole by one workgroup in one launch (same iteration, its sums in
   another order: the same result to rounding).  Default all of them.     */
#define XG_CG_GRAPH 1
#define XG_CG_FUSE  2
#define XG_CG_SMALL 4
#define XG_CG_WARM  8  /* under LoadControl the first Newton iteration of a step
                          starts from the previous step's first solution,
                          DisplacementControl's reference-load solves from the
                          previous one (the chunked PCG and the band-factor
                          preconditioned solves; the default; the stop test
                          stays |r| <= rtol |b|)                               */
#define XG_CG_WIDE  16 /* with XG_CG_SMALL, under LoadControl and
                          DisplacementControl: systems of more than 1,024
                          equations whose half bandwidth in the band order is
                          at most 1,024 are preconditioned by their exact band
                          Cholesky factor, formed in 64 x 64 tiles (one or two
                          PCG iterations per solve); on by default           */
int  xg_cg_config(xg_domain* d, int chunk, int flags);
/* `numberer RCM|Plain`                                                       */
int  xg_numberer(xg_domain* d, int kind);
/* `test NormUnbalance tol maxIter` etc.                                      */
int  xg_test(xg_domain* d, int kind, double tol, int max_iter);
/* `integrator LoadControl dlambda` (LoadControl.cpp:70-120)                  */
int  xg_integrator_load_control(xg_domain* d, double dlambda);
/* `integrator DisplacementControl node dof incr` (DisplacementControl.cpp:121-283);
   dof is 1-based as in the Tcl command.                                      */
int  xg_integrator_disp_control(xg_domain* d, int node_tag, int dof, double incr);
/* `integrator Newmark gamma beta` (Newmark.cpp:105-458)                      */
int  xg_integrator_newmark(xg_domain* d, double gamma, double beta);
/* `analyze n <dt>`: n steps of NewtonRaphson (NewtonRaphson.cpp:125-221)
   under the chosen integrator.  Returns 0 or the failing status.            */
int  xg_analyze(xg_domain* d, int nsteps, double dt);

/* ---- domainChanged: numbering and CSR pattern (bit-exact) ----------------- */
/* PlainHandler::handle + DOF_Numberer::numberDOF (RCM.cpp:67-259) +
   AnalysisModel::getDOFGraph + SparseGenRowLinSOE::setSize
   (SparseGenRowLinSOE.cpp:129-234).  Also builds the assembly scatter maps. */
int  xg_setup(xg_domain* d);
int  xg_num_nodes(xg_domain* d);
int  xg_num_elements(xg_domain* d);
int  xg_num_eqn(xg_domain* d);
long long xg_csr_nnz(xg_domain* d);
/* equation number of every node DOF (node-tag order, n*ndf), -1 for fixed
   (DOF_Group::getID after numberDOF, DOF_Numberer.cpp:120-133)               */
int  xg_node_eqns(xg_domain* d, int* eqn);
/* CSR pattern: rowptr (neq+1), colind (nnz)                                  */
int  xg_csr_pattern(xg_domain* d, int* rowptr, int* colind);

/* ---- state determination: the hot path ------------------------------------ */
/* Node::setTrialDisp for every node (n*ndf, node-tag order)                   */
int  xg_set_trial_disp(xg_domain* d, const double* U);
int  xg_get_trial_disp(xg_domain* d, double* U);
/* AnalysisModel::incrDisp (AnalysisModel.cpp:452): U += dU (dU per equation) */
int  xg_incr_disp(xg_domain* d, const double* dU);
/* Domain::update (Domain.cpp:1988-2006): Element::update of every element,
   every Gauss point / fiber trial state.                                     */
int  xg_update(xg_domain* d);
/* IncrementalIntegrator::formTangent (IncrementalIntegrator.cpp:74-98) with
   SparseGenRowLinSOE::addA (SparseGenRowLinSOE.cpp:237-291): K into the CSR
   values; copied to vals (nnz) when vals != NULL.                            */
int  xg_form_tangent(xg_domain* d, double* vals);
/* LinearSOE::addA (SparseGenRowLinSOE::addA, SparseGenRowLinSOE.cpp:237-291)
   for contributions other than the element tangents -- the DOF_Group
   blocks of TransientIntegrator::formTangent (TransientIntegrator.cpp:88-95)
   -- after xg_form_tangent: nblk blocks of bs x bs values (row-major, the
   factor already applied) added at the CSR entries of their equation
   numbers id (nblk*bs, negative ids skipped).  An entry outside the
   pattern is XG_ERR_ARG.                                                    */
int  xg_add_a(xg_domain* d, int nblk, int bs, const double* m, const int* id);
/* StaticIntegrator::formUnbalance (StaticIntegrator.cpp:51-70): B = -sum R_e
   + lambda P; copied to B (neq) when B != NULL.                              */
int  xg_form_unbalance(xg_domain* d, double* B);
/* LinearSOE::solve (LinearSOE.h:46-89) of the assembled system with right-hand
   side B (neq) into X (neq).                                                 */
int  xg_solve(xg_domain* d, const double* B, double* X);
/* xg_solve without allocating: no counterpart in the reference; for callers
   that cannot allocate device memory (code around HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc).  Same result as
   xg_solve; every PCG work buffer (staging of B and X, the six iteration
   vectors, the block-Jacobi inverses or the one-workgroup band factor,
   partial sums, scalars) comes from a fixed 128 MiB device array per GPU,
   looked up at xg_create; the tiled band factor (XG_CG_WIDE) uses the
   domain's own buffers, allocated at xg_setup; the kernels are launched
   one by one (no graph).  A system whose buffers do not fit (about 1.1 M
   equations with nodal 3x3 blocks) is XG_ERR_ARG.  Calls that overlap on
   one GPU share that array and race.                                       */
int  xg_solve_noalloc(xg_domain* d, const double* B, double* X);
/* y = A x with the assembled CSR values (for tests)                          */
int  xg_spmv(xg_domain* d, const double* x, double* y);
/* Domain::commit (Domain.cpp:1893-1923) / revertToLastCommit / revertToStart */
int  xg_commit(xg_domain* d);
int  xg_revert_to_last_commit(xg_domain* d);
int  xg_revert_to_start(xg_domain* d);
/* Load factor lambda (Domain::applyLoad) and pseudo-time.                    */
int  xg_set_load_factor(xg_domain* d, double lambda);
double xg_get_load_factor(xg_domain* d);

/* ---- element / material point read-back (Element::getResponse, print) ----- */
/* Per-point data in element-tag order, points in element quadrature order.
   what = XG_OUT_*; returns the number of doubles written per point in *width. */
int  xg_get_point_data(xg_domain* d, int what, double* out, int* width);
/* Element::getTangentStiff / getResistingForce of one element (by tag).      */
int  xg_element_tangent(xg_domain* d, int ele_tag, double* K, int* ndof);
int  xg_element_resisting_force(xg_domain* d, int ele_tag, double* R, int* ndof);
int  xg_get_counters(xg_domain* d, xg_counters* c);
/* Device time of `reps` back-to-back launches of op (XG_OP_*), measured with
   HIP events on the stream the kernels run on.  *ms = average per launch.    */
int  xg_time_op(xg_domain* d, int op, int reps, double* ms);


/* ---- batched ensemble of independent 2D frame models ----------------------
   No counterpart in the reference, which analyses one model per interpreter:
   each member is a model defined through the entry points above (2D,
   dispBeamColumn over fiber sections, nodal masses, UniformExcitation,
   Newmark, NormDispIncr/NormUnbalance test) whose `analyze n dt` would run
   BasicAnalysisBuilder::analyzeTransient (BasicAnalysisBuilder.cpp:422-520).
   All members advance together: one persistent launch per Newton iteration,
   one captured hipGraph per time step.  A member whose step fails is
   reverted to its last committed state and stops, as its analyze would. */
typedef struct xg_ensemble xg_ensemble;
int  xg_ensemble_create(int device, xg_ensemble** out);
int  xg_ensemble_destroy(xg_ensemble* e);
/* snapshot of a defined (not necessarily set up) model; the handle may be
   destroyed afterwards */
int  xg_ensemble_add(xg_ensemble* e, xg_domain* model);
int  xg_ensemble_setup(xg_ensemble* e);
int  xg_ensemble_num_models(xg_ensemble* e);
/* autotuner: times every (block, grid, model order) mapping and keeps the
   fastest; writes up to `cap` rows of (block, grid, order, ms) into `table` */
int  xg_ensemble_tune(xg_ensemble* e, double* table, int cap, int* nrows);
int  xg_ensemble_set_mapping(xg_ensemble* e, int block, int grid, int order, int use_graph);
/* nsteps Newmark steps of size dt for every running member; *nstopped =
   members stopped by a failed step so far */
int  xg_ensemble_analyze(xg_ensemble* e, int nsteps, double dt, int* nstopped);
/* trial displacements of member `model` (its nodes in ascending tag order) */
int  xg_ensemble_model_disp(xg_ensemble* e, int model, double* U, int* nnode);
/* per-member status (0/1 running, -2 stopped) and Newton iteration totals */
int  xg_ensemble_status(xg_ensemble* e, int* status, long long* newton_iters);

/* ---- Committed state (checkpoint / restart, hand-over between models) -----
   The committed state in the reference's sendSelf layouts, which define it
   (FE_Datastore save/restore of MovableObjects):
     XG_STATE_NODES       per node (ascending tag): committed displacement,
                          velocity, acceleration (ndf each; Node.cpp:767-791)
     XG_STATE_J2          per Gauss point (elements ascending, points in
                          element order): xi_n, epsilon_p_n row by row (9)
                          (J2Plasticity::sendSelf, J2Plasticity.cpp:597-614)
     XG_STATE_STEEL02     per Steel02 fiber (elements ascending, sections,
                          fibers in definition order): epsminP, epsmaxP,
                          epsplP, epss0P, sigs0P, epssrP, sigsrP, konP, epsP,
                          sigP, eP (Steel02::sendSelf, Steel02.cpp:287-308)
     XG_STATE_CONCRETE02  per Concrete02 fiber: ecminP, deptP, epsP, sigP, eP
                          (Concrete02::sendSelf, Concrete02.cpp:339-354)
     XG_STATE_DOMAIN      committed time and load factor (2 values)
   xg_set_state writes the committed state and sets the trial state equal
   to it, as after revertToLastCommit; stresses and tangents follow at the
   next update. */
#define XG_STATE_NODES      1
#define XG_STATE_J2         2
#define XG_STATE_STEEL02    3
#define XG_STATE_CONCRETE02 4
#define XG_STATE_DOMAIN     5
long long xg_state_size(xg_domain* d, int kind);   /* doubles; < 0 on error */
int  xg_get_state(xg_domain* d, int kind, double* out);
int  xg_set_state(xg_domain* d, int kind, const double* in);

/* ---- Adjoint sensitivity of a frame transient ----------------------------
   Gradients of
       loss = sum over the steps n of a forward run of sum_i U_n(node_i, dof)^2
   with respect to every Steel02 material's (Fy, E0, b) and every
   Concrete02 material's fc (as stored, compression negative), by a discrete
   adjoint of the converged Newmark + Newton steps.  It stands in for the
   reference's direct-differentiation path for such a gradient
   (sensitivityAlgorithm / computeGradients, SRC/runtime/commands/analysis/
   sensitivity.cpp; Newmark::computeSensitivities, Newmark.cpp:904-950),
   which has no Steel02/Concrete02 material partials
   (UniaxialMaterial.cpp:340-379).  Models: dispBeamColumn2d/3d over fiber
   sections only, Newmark, nodal masses, mass-proportional damping.  The
   state at the start of the forward run is treated as fixed. */
/* loss nodes (tags) and DOF (1-based) */
int  xg_adjoint_loss(xg_domain* d, int n, const int* node_tags, int dof);
/* nsteps Newmark steps of size dt from the current committed state,
   checkpointing the committed state every `ckpt` steps; *loss = loss */
int  xg_adjoint_forward(xg_domain* d, int nsteps, double dt, int ckpt, double* loss);
/* reverse sweep over the last forward run, recomputing each checkpoint
   segment; gradients in ascending material-tag order (3 per Steel02,
   1 per Concrete02).  The domain is left at the end of the forward run. */
int  xg_adjoint_backward(xg_domain* d, double* grad_steel, double* grad_conc);
/* wall times (ms) of the last forward and backward runs and the number of
   CG iterations of the backward linear solves */
int  xg_adjoint_stats(xg_domain* d, double* forward_ms, double* backward_ms, long long* backward_cg);
/* Host evaluation of the per-fiber adjoint the kernels run: for a Steel02
   (kind 1, prm = Fy E0 b R0 cR1 cR2 a1 a2 a3 a4 sigini) or Concrete02
   (kind 2, prm = fc epsc0 fcu epscu rat ft Ets) trial update from the
   committed history h (pool order: 10 / 5 values; kon for Steel02) at
   `strain`, pull the adjoint hbar of the trial history back to the
   committed history (cbar), the parameters (thbar: Fy E0 b / fc) and the
   strain (*sbar).  The adjoint of the tangent entry (e) is not propagated. */
int  xg_material_vjp(int kind, const double* prm, const double* h, int kon, double strain, const double* hbar,
                     double* cbar, double* thbar, double* sbar);

#ifdef __cplusplus
}
#endif
#endif /* XARA_GPU_H */
