/*
 * dvide.h -- C ABI of libdvide_hip.so, the MI355X (gfx950) contact time step.
 *
 * The reference exposes this path as C++ classes only (System.cuh:279-326,
 * Solver.cuh:17-46, APGD.cuh:62-67, Body.cuh:28-99, CollisionDetector.cuh:48-103);
 * it has no C ABI (DanVIApi.h:7-20 defines export macros that are applied to
 * nothing).  Each entry point below names the reference member it replaces.
 * The C++ drop-in surface (include/dvide/System.hpp: System, Body, Solver,
 * APGD, CollisionDetector) is a header-only layer over these functions.
 *
 * Conventions
 *   - every function returns int: 0 on success, a negative DV_ERR_* code on
 *     failure; dv_last_error() gives the message for the last failure on a
 *     scene;
 *   - all buffers are caller-owned host memory, plain pointers and sizes;
 *   - state vectors use the reference's flat layout: p, v, f_contact have
 *     3 entries per body (DOF 3*j .. 3*j+2 of body j, System.cu:342-424);
 *     gamma has the bilateral rows first, then 3 rows (n,u,v) per contact
 *     (System.cu:1380, APGD.cu:166);
 *   - one HIP stream per scene handle; calls on distinct handles may come from
 *     distinct host threads, calls on one handle must not overlap.
 */
#ifndef DVIDE_H_
#define DVIDE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dv_scene dv_scene;

/* error codes */
#define DV_OK 0
#define DV_ERR_ARG (-1)         /* invalid argument */
#define DV_ERR_STATE (-2)       /* call not valid in the scene's state */
#define DV_ERR_HIP (-3)         /* a HIP runtime call failed */
#define DV_ERR_NOMEM (-4)       /* device allocation failed or budget exceeded */
#define DV_ERR_UNSUPPORTED (-5) /* option not supported by the selected solver */
#define DV_ERR_DIVERGED (-6)    /* solver produced a non-finite value */

/* System(int solverType) (System.cu:77-101): 1 = APGD; 2..7 name PDIP, TPAS,
 * JKIP, PGJ, PGS and PJKIP, which are not on this path, so dv_scene_create
 * rejects them with DV_ERR_UNSUPPORTED; any other value is the reference's
 * default case, APGD. */
#define DV_SOLVER_APGD 1

/* APGD builds (DV_PARAM_APGD_BUILD); each runs APGD::solve (APGD.cu:164-361) */
#define DV_APGD_FP64 0  /* device-resident state machine in hipGraph chunks, FP64 (default) */
#define DV_APGD_MIXED 1 /* same, FP32 D/M^-1/products, FP64 iterates, tests and update */
#define DV_APGD_PLAIN 2 /* host-driven FP64, one kernel per CUSP operation; the build
                           that runs anti-relaxation */

/* double parameters (dv_set_param / dv_get_param) */
#define DV_PARAM_H 0          /* time step              System::setTimeStep, System.cu:135 */
#define DV_PARAM_MU 1         /* friction coefficient   System::setFrictionCoefficient, System.cu:140 */
#define DV_PARAM_TOL 2        /* APGD tolerance         Solver::tolerance, Solver.cuh:23 */
#define DV_PARAM_MAX_ITER 3   /* APGD iteration cap     Solver::maxIterations, Solver.cuh:24 */
#define DV_PARAM_GX 4         /* gravity                System::gravity, System.cu:16 */
#define DV_PARAM_GY 5
#define DV_PARAM_GZ 6
#define DV_PARAM_ENVELOPE 7   /* collision envelope     CollisionDetector::setEnvelope, CollisionDetector.cu:525 */
#define DV_PARAM_WARM_START 8 /* 0/1                    APGD::setWarmStarting, APGD.cuh:62 */
#define DV_PARAM_APGD_BUILD 9 /* DV_APGD_*: FP64, mixed or plain (no counterpart: the
                                 reference has one APGD, in FP64) */
#define DV_PARAM_GEOM_MODE 10 /* 1: lengths/normals rounded through binary32 as the
                                 reference's double3 helpers compile under nvcc
                                 (helper_math_double.h:576-598); 0: plain FP64 */
#define DV_PARAM_TIME 11      /* System::time (read/write) */
#define DV_PARAM_VERBOSE 12   /* Solver::verbose */
#define DV_PARAM_GRID 13      /* unused (accepted for compatibility) */
#define DV_PARAM_PROFILE 14   /* time one kernel of the graph-driven solvers with HIP
                                 events: 0 off, 1 P1 (contacts; for one scene it also
                                 takes the previous pass's decision), 2 P2 (bodies),
                                 4 the decide kernel (3 is refused: a pass has two
                                 kernels, P1 and P2).  One chunk of 16
                                 passes per solve (the first after the first chunk whose
                                 entry state has the momentum term on) is then launched
                                 on the scene's stream, and in its first pass (a deciding
                                 P1: its second) that kernel runs as a captured graph of
                                 16 back-to-back launches between two events (the decide
                                 kernel once, alone, where it runs: after the chunk's
                                 last pass for one scene); the rest run in the captured
                                 chunk graph */
#define DV_PARAM_ROWS_PER_THREAD 15 /* contact kernels of the graph-driven solvers: 0 the
                                       tuned choice (default), 1 or 2 contacts per thread
                                       (results agree to rounding, each choice is
                                       deterministic) */
#define DV_PARAM_ANTI_RELAX 16 /* 0/1                  APGD::setAntiRelaxation, APGD.cuh:63
                                  (APGD.cu:104-114, 335-345); runs on DV_APGD_PLAIN, and a
                                  DV_APGD_FP64 scene with it on is solved by that build */

/* state vectors (dv_get_vector / dv_set_vector) */
#define DV_VEC_P 0         /* System::p_d        3*nb */
#define DV_VEC_V 1         /* System::v_d        3*nb */
#define DV_VEC_GAMMA 2     /* System::gamma_d    nbil + 3*nc */
#define DV_VEC_F_CONTACT 3 /* System::f_contact_d 3*nb */
#define DV_VEC_R 4         /* System::r_d        nbil + 3*nc */
#define DV_VEC_K 5         /* System::k_d        3*nb */
#define DV_VEC_B 6         /* System::b_d        nbil + 3*nc */

/* per-step statistics (dv_get_stats) */
typedef struct dv_stats {
  int num_contacts;       /* CollisionDetector::numCollisions */
  int num_bilaterals;     /* constraintsBilateralDOF_d.size() */
  int iterations;         /* Solver::iterations (APGD.cu:357) */
  int backtracks;         /* passes of the Lipschitz backtracking loop */
  double residual;        /* best APGD residual (APGD.cu:293-301) */
  double objective;       /* System::objectiveCCP (APGD.cu:353-354) */
  double ms_total;        /* System::elapsedTime (System.cu:586-590) */
  double ms_collision;    /* per-phase device time */
  double ms_assembly;
  double ms_solve;
  double ms_update;
  double time;            /* System::time after the step */
  double prof_kernel_ms;  /* DV_PARAM_PROFILE: summed duration of the timed launches */
  int prof_kernel_count;  /* ... and their number */
  int passes;             /* solver passes (Schur products) of the step */
  /* launch schedule of the graph-driven builds' last chunk of passes (none
     changes a result): contact rows per thread, XCD-contiguous tile order,
     and, for several slots, workgroups per CU of the contact and body
     kernels; schedule_tuning is 1 when the solve measured candidate
     schedules (see DVIDE_SCHEDULE_FILE in dvide_amd/csrc/dv_apgd_fused.hip) */
  int rows_per_thread;
  int xcd_order;
  int wg_per_cu_contacts;
  int wg_per_cu_bodies;
  int schedule_tuning;
} dv_stats;

/* --- scene lifetime ------------------------------------------------------ */

/* System::System(int solverType), System.cu:62-133.  solverType 2..7 (PDIP,
 * TPAS, JKIP, PGJ, PGS, PJKIP) fail with DV_ERR_UNSUPPORTED; 1 and the
 * reference's default case give APGD in the DV_APGD_FP64 build.  Fails with
 * DV_ERR_HIP when no HIP device is visible (there is no CPU fallback);
 * dv_last_error(NULL) then names the cause. */
int dv_scene_create(dv_scene** out, int solverType);
int dv_scene_destroy(dv_scene* s);
const char* dv_last_error(const dv_scene* s);

/* System::add(Body*), System.cu:145-150, with the Body fields of Body.cuh:45-98:
 * pos[3*n], vel[3*n], mass[n], geom[3*n] ((r,0,0) sphere, (hx,hy,hz) box),
 * fixed[n] (Body::setBodyFixed), family[n] (PhysicsItem::setCollisionFamily).
 * Bodies get consecutive identifiers starting at the current body count. */
int dv_add_bodies(dv_scene* s, int n, const double* pos, const double* vel, const double* mass,
                  const double* geom, const int* fixed, const int* family);

/* System::addBilateralConstraintDOF(int,int,double,double), System.cu:510-532:
 * row p[dofA] - p[dofB] (or p[dofA] - p0 when dofB < 0) driven at `velocity`
 * from `start_time` on. */
int dv_add_bilateral_dof(dv_scene* s, int dofA, int dofB, double velocity, double start_time);

/* System::importMesh(filename, stiffness, numContactPointsPerElement,
 * pressure), System.cu:1932-2060: an ANCF shell mesh in the reference's text
 * format (a header line "nodes, shells, M entries, M^-1 entries", then the
 * nodes' position and two gradient rows, connectivity, materials, geometry
 * (a, b, thickness, -), the element-to-DOF map, external forces and the M and
 * M^-1 triplets, 1-based).  Its DOFs follow the bodies' (9 per node) in every
 * DOF vector and in D's columns.  Each element carries (points - 2)^2 contact
 * points (System.cu:494-499), spheres of half the thickness in collision
 * family -2, whose D blocks are 3x36 (System.cu:904-961).  Only fixed shells
 * are supported: a mesh with a nonzero M^-1 entry (a shell free to move,
 * which needs the shell's elastic forces) fails with DV_ERR_UNSUPPORTED, and
 * stiffness and pressure, which act on free DOFs only, are not used.  One
 * mesh per scene, after the bodies and before dv_initialize; at most 2^24
 * nodes or elements and 2^26 contact points. */
int dv_import_mesh(dv_scene* s, const char* path, double stiffness, int points, double pressure);

/* The same fixed shell mesh from arrays (no counterpart in the reference;
 * the arrays importMesh fills): nodes[9*n_nodes] (position, x-gradient,
 * y-gradient per node), connectivity[4*n_elements] (0-based),
 * geometry[4*n_elements] (a, b, thickness, ignored), `points` contact points
 * per element side (as importMesh's numContactPointsPerElement), and fext[9*
 * n_nodes] (may be NULL: zeros), which enters the potential energy only.
 * At most 2^26 contact points per mesh (DV_ERR_ARG beyond). */
int dv_add_shell_mesh(dv_scene* s, int n_nodes, const double* nodes, int n_elements, const int* connectivity,
                      const double* geometry, int points, const double* fext);

/* CollisionDetector::setBinsPerAxis, CollisionDetector.cu:6 (accepted for API
 * parity; the spatial-hash broadphase sizes its cells from the geometry). */
int dv_set_bins_per_axis(dv_scene* s, int bx, int by, int bz);

int dv_set_param(dv_scene* s, int param, double value);
int dv_get_param(const dv_scene* s, int param, double* value);

/* System::initializeSystem, System.cu:342-508: packs and uploads the bodies.
 * A scene holds at most 2^28 bodies, shell contact points included
 * (DV_ERR_UNSUPPORTED beyond). */
int dv_initialize(dv_scene* s);

/* System::DoTimeStep, System.cu:544-602: collision detection, D and r
 * assembly, APGD, and the update v = M^-1 (k + D^T gamma), p += h v.  A step
 * whose collision detection finds more than 2^28 contacts fails with
 * DV_ERR_UNSUPPORTED. */
int dv_step(dv_scene* s);

/* --- data access (System's public device vectors) ------------------------- */

int dv_num_bodies(const dv_scene* s, int* n);
/* nodes of the scene's shell mesh (System::nodes_h / 3); 9 DOFs each */
int dv_num_shell_nodes(const dv_scene* s, int* n);
/* System::bodies (Body::getMass, getGeometry, isFixed, getCollisionFamily):
 * copies of the first min(n, count) bodies' fields; any pointer may be NULL */
int dv_get_bodies(const dv_scene* s, double* mass, double* geom, int* fixed, int* family, size_t n);
int dv_num_contacts(const dv_scene* s, int* nc);
int dv_get_stats(const dv_scene* s, dv_stats* out);

/* copies min(n, length) entries of the vector into out; *len gets its length.
 * DOF vectors (p, v, f_contact, k) hold the bodies' 3 DOFs each, then a shell
 * mesh's 9 per node (fixed: its nodal coordinates, zero velocity, and the
 * contact force D^T gamma / h on its DOFs); k of a scene with a shell fails
 * with DV_ERR_UNSUPPORTED (it would need the shell's elastic forces). */
int dv_get_vector(dv_scene* s, int which, double* out, size_t n, size_t* len);
/* overwrites p or v (drivers mutate p_d/v_d between steps, models/test_draft.cu:438-439);
 * a fixed shell's DOFs must keep their values */
int dv_set_vector(dv_scene* s, int which, const double* in, size_t n);

/* CollisionDetector::collisionIdentifierA_d/B_d and normalsAndPenetrations_d
 * (CollisionDetector.cu:495-498): idA < idB geometry indices, normal from B to A
 * and the gap (= -penetration); loc (may be NULL) = collisionLocations_d.
 * Arrays sized for nc entries (nap: 4*nc, loc: 3*nc), canonical (idA, idB) order. */
int dv_get_contacts(dv_scene* s, uint32_t* idA, uint32_t* idB, double* nap, double* loc,
                    size_t nc);

/* System::getCCPViolation, System.cu:1605-1643: min(gamma_n,0), min((Dv+b)_n,0),
 * gamma_n . (Dv+b)_n and the 2-norm of min(mu gamma_n - |gamma_t|, 0). */
int dv_ccp_violation(dv_scene* s, double out[4]);

/* System::D (System.cu:1351-1386) as row-ordered COO: the bilateral rows
 * first (+1 at dofA, -1 at dofB), then per contact the rows n, u, v, each with
 * body A's three columns (+) before body B's (-), or for a shell contact
 * point its element's 36 columns weighted by the shape functions.  *nnz,
 * *rows, *cols get the sizes; I, J, V (cap entries, may be NULL to query) get
 * the entries.  Feeds System::exportMatrices (System.cu:1912-1930). */
int dv_get_jacobian(dv_scene* s, int* I, int* J, double* V, size_t cap, size_t* nnz, int* rows, int* cols);

/* System::calculateContactForcePerCollision (System.cu:1259-1289): per contact
 * gamma_n n / h and (gamma_u u + gamma_v v) / h, 3 values each (nc entries).
 * Like the reference it reads gamma at 3 i without the bilateral offset. */
int dv_contact_forces(dv_scene* s, double* normal, double* friction, size_t nc);

/* System::getPotentialEnergy, getKineticEnergy, getStrainEnergy
 * (System.cu:1645-1666): -f.p, v.(M v)/2 and 0 (no elastic bodies here). */
int dv_energies(dv_scene* s, double out[3]);

/* --- hot-path pieces, exposed for parity tests ---------------------------- */

/* Collision detection only (CollisionDetector::generateAxisAlignedBoundingBoxes,
 * detectPossibleCollisions_spatialSubdivision, detectCollisions,
 * CollisionDetector.cu:547-665) plus assembly of k, D and r
 * (System::buildAppliedImpulseVector / buildContactJacobian / buildSchurVector,
 * System.cu:1351-1409, 1502-1580).  *nc gets the contact count. */
int dv_assemble(dv_scene* s, int* nc);

/* y = N x = D M^-1 D^T x with the assembled D (APGD::performSchurComplementProduct,
 * APGD.cu:81-87); x, y have nbil + 3*nc entries.  Terms are summed in D^T's
 * row order, as a row-sorted COO SpMV sums them. */
int dv_schur_product(dv_scene* s, const double* x, double* y, size_t n);

/* The same product N x (APGD::performSchurComplementProduct, APGD.cu:81-87)
 * by the kernels one pass of the graph-driven solver runs (build DV_APGD_FP64
 * or DV_APGD_MIXED): J = F^T x stored per contact, the per-body gather
 * tmp = M^-1 (sum_A J - sum_B J + bilateral rows) and the contact rows
 * F (tmp_A - tmp_B), in that build's arithmetic (FP32 products for the mixed
 * build).  Summed in another order than dv_schur_product, so it agrees with
 * it to rounding.  The scene's state (gamma, v, p) is left unchanged. */
int dv_pass_product(dv_scene* s, int build, const double* x, double* y, size_t n);

/* dv_schur_product without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  x and y are staged through a fixed device array of the
 * library (16 MiB per GPU, held until the process ends) instead of a buffer
 * allocated per call, so n is at most 1,048,576 rows (about 349k contacts);
 * larger n fails with DV_ERR_ARG.  The array is shared by every call on a
 * GPU and keeps the last call's contents: calls that overlap on one GPU
 * (from several host threads or scenes) race and must be serialised by the
 * caller.  Results are those of dv_schur_product, bit for bit. */
int dv_schur_product_noalloc(dv_scene* s, const double* x, double* y, size_t n);

/* Coulomb-cone projection of n 3-vectors in place with friction mu
 * (APGD.cu:47-78). */
int dv_project_cone(dv_scene* s, double* g, size_t n, double mu);

/* APGD::solve, APGD.cu:164-361, on the assembled system (no update). */
int dv_solve(dv_scene* s);

/* --- batches and serving ---------------------------------------------------
 * No counterpart in the reference, which steps one System at a time.  Both
 * step initialized scenes that stay resident on the device (bodies, contacts,
 * warm-start gamma) and remain usable alone.  A scene's result is
 * bit-identical whether it is stepped alone (dv_step with the same build) or
 * in a batch or a server mix: its partial sums are formed per tile of its
 * own rows and reduced in a fixed order. */

typedef struct dv_batch dv_batch;
typedef struct dv_server dv_server;

typedef struct dv_batch_stats {
  int num_scenes;
  int max_iterations;       /* over the scenes */
  long total_contacts;
  long total_iterations;
  double ms_total;          /* host wall time of the step */
  double ms_assembly;       /* collision detection and assembly of all scenes */
  double ms_solve;          /* the shared APGD launches */
  double ms_update;
  int rows_per_thread;      /* the solve's launch schedule, as in dv_stats */
  int xcd_order;
  int wg_per_cu_contacts;
  int wg_per_cu_bodies;
  int schedule_tuning;
} dv_batch_stats;

/* build: DV_APGD_FP64 or DV_APGD_MIXED (the graph-driven builds) */
int dv_batch_create(dv_batch** out, int build);
int dv_batch_destroy(dv_batch* b);
const char* dv_batch_last_error(const dv_batch* b);
/* adds an initialized scene; the batch does not own it */
int dv_batch_add(dv_batch* b, dv_scene* s);
/* one DoTimeStep of every scene: per-scene collision detection and
 * assembly, one APGD solve of all scenes in the same graph launches, per-scene
 * update; each scene's dv_get_stats is filled as by dv_step */
int dv_batch_step(dv_batch* b);
int dv_batch_get_stats(const dv_batch* b, dv_batch_stats* out);

/* A continuous-batching server of time-step requests.  Requests wait in
 * submission order; a waiting request is admitted when a slot is free and
 * the solver memory of the admitted scenes stays within hbm_budget bytes
 * (<= 0: 80% of the device memory free at creation).  Its collision
 * detection and assembly run at admission, its solve joins the slots
 * already iterating at the next chunk of passes, and it completes when its
 * APGD has converged (or hit maxIterations) and its update has run. */
int dv_server_create(dv_server** out, int build, int max_slots, double hbm_budget);
int dv_server_destroy(dv_server* srv);
const char* dv_server_last_error(const dv_server* srv);
/* queues one DoTimeStep of `scene` (one request per scene in flight);
 * *ticket identifies it in dv_server_poll */
int dv_server_submit(dv_server* srv, dv_scene* scene, long* ticket);
/* admits waiting requests, runs one chunk of solver passes, completes the
 * requests whose solves are done; their tickets go to done[0..*n_done) (at
 * most max_done; the rest are reported by later polls).  A request whose
 * solve diverged (non-finite residual) frees its slot without an update and
 * makes the poll return DV_ERR_DIVERGED (done[] is still filled); its ticket
 * is then reported by dv_server_failed.  If assembling admitted requests
 * fails, they wait again in their order and the poll returns the error. */
int dv_server_poll(dv_server* srv, long* done, int max_done, int* n_done);
/* tickets of diverged requests not yet reported (at most max_n) */
int dv_server_failed(dv_server* srv, long* tickets, int max_n, int* n);
/* number of requests waiting, in flight or completed and not yet reported */
int dv_server_pending(const dv_server* srv, int* n);

#ifdef __cplusplus
}
#endif

#endif /* DVIDE_H_ */
