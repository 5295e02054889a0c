/*
 * svlg.h -- C ABI of the MI355X-native SeismoVLAB dynamic-FEM hot path (libsvlg.so).
 *
 * Plain pointers and sizes only. Host arrays are copied to HBM at model creation;
 * every state vector afterwards lives on the device and is only copied back when a
 * caller asks for it. Return value: 0 = ok, nonzero = stop (the meaning of SVL's
 * `bool stop`); svlg_last_error() gives the message. One HIP stream per model; a model
 * handle is not thread-safe.
 *
 * The entry points replace, method for method, the SVL C++ interfaces that sit on the
 * hot path (citations are relative to the SVL source tree):
 *   Assembler    02-Run_Process/07-Assembler/Assembler.hpp:50-181
 *   Integrator   02-Run_Process/10-Integrators/Integrator.hpp:53-144
 *                (CentralDifference.hpp:63, ExtendedNewmarkBeta.hpp)
 *   LinearSystem 02-Run_Process/11-Solvers/LinearSystem.hpp:44-66
 */
#ifndef SVLG_H
#define SVLG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVLG_ELEM_QUAD4 0 /* lin2DQuad4  (04-Elements/06-Quadrilateral/lin2DQuad4.cpp) */
#define SVLG_ELEM_HEXA8 1 /* lin3DHexa8  (04-Elements/10-Hexahedron/lin3DHexa8.cpp) */

#define SVLG_MAT_ELASTIC 0 /* Elastic2DPlaneStrain / Elastic3DLinear */
#define SVLG_MAT_J2 1      /* PlasticPlaneStrainJ2 / Plastic3DJ2 */

#define SVLG_PRECOND_JACOBI 0
#define SVLG_PRECOND_LEARNED 1

/* Model description, in SVL's ascending-tag order (mesh nodes, then continuum
 * elements). All arrays are host memory, read only during svlg_model_create. */
typedef struct svlg_mesh_desc {
    int32_t ndim;             /* 2 (Quad4, plane strain) or 3 (Hexa8) */
    int32_t elem_type;        /* SVLG_ELEM_QUAD4 or SVLG_ELEM_HEXA8 */
    int32_t nnodes;           /* mesh nodes */
    int32_t nelem;            /* continuum elements */
    const double *coords;     /* [nnodes*ndim] */
    const int32_t *free_dof;  /* [nnodes*ndim] Plain-scheme free index, -1 restrained */
    int32_t nfree;
    const int32_t *conn;      /* [nelem*nen] 0-based mesh node index, SVL node order */
    const int32_t *mat_type;  /* [nelem] SVLG_MAT_* (NULL: all elastic) */
    const double *E;          /* [nelem] Young's modulus */
    const double *nu;         /* [nelem] Poisson ratio */
    const double *rho;        /* [nelem] density */
    const double *j2;         /* [nelem*5] K, G, sigma_y, H, beta (NULL if no J2) */
    double thickness;         /* lin2DQuad4 `th` */
    int32_t lumped;           /* Global massform: 1 LUMPED, 0 CONSISTENT */
    const double *ray_am;     /* [nelem] Rayleigh am (NULL: Free damping) */
    const double *ray_ak;     /* [nelem] Rayleigh ak (NULL: Free damping) */
    /* ZeroLength1D + Viscous1DLinear dashpots to a fully restrained partner node
     * (01-Pre_Process/Method/Builder.py:1086-1131), in element-tag order. */
    int32_t ndash;
    const int32_t *dash_node; /* [ndash] mesh node */
    const int32_t *dash_dir;  /* [ndash] 0-based direction */
    const double *dash_eta;   /* [ndash] viscous coefficient eta */
    /* POINTLOAD_CONCENTRATED_DYNAMIC loads of the active load combination
     * (Assembler.cpp:334-350 + :485; Load.cpp:69-71). */
    int32_t nload;
    int32_t load_nt;          /* samples per series */
    const int32_t *load_node; /* [nload] */
    const double *load_dir;   /* [nload*ndim] direction vector */
    const double *load_factor;/* [nload] combination factor */
    const double *load_series;/* [nload*load_nt] amplitude time series */
    /* Domain reduction (ElementLoad GENERALWAVE: ELEMENTLOAD_DOMAIN_REDUCTION, Assembler.cpp:
     * 460-478; lin*::ComputeDomainReductionForces, lin2DQuad4.cpp:564-614; node signals as
     * Driver.hpp:1669-1720 reads them). ndrm_elem = 0: none. */
    int32_t ndrm_elem;
    const int32_t *drm_elem;  /* [ndrm_elem] continuum element index */
    int32_t ndrm_node;
    const int32_t *drm_node;  /* [ndrm_node] mesh node; every node of a DRM element is one */
    const int32_t *drm_cond;  /* [ndrm_node] 1 = exterior node (its signal enters negated) */
    int32_t drm_nt;           /* rows of every node signal (row k at step k) */
    const double *drm_signal; /* [ndrm_node*drm_nt*3*ndim] rows (U, V, A components) */
    double drm_factor;        /* combination factor of the DRM load */
    /* Support motion: Supports (TIMESERIES / CONSTANT per DOF) activated by the combination's
     * POINTLOAD_SUPPORT_MOTION loads (Driver.hpp:505-563, 1725-1734; Assembler::
     * ComputeSupportMotionIncrement, Assembler.cpp:493-532). Entry j prescribes the displacement
     * g_j of a restrained DOF (node, dir) with the load's factor; with i = k - sm_offset, step k
     * moves the DOF by the sum over its entries of factor (g_j(i) - g_j(i-1)) (factor g_j(0) at
     * i = 0), g_j(i) = values[i] for i below its length, values[0] past it (Node::
     * GetSupportMotion, Node.cpp:228-249). sm_offset 0: CentralDifference and
     * ExtendedNewmarkBeta (CentralDifference.cpp:195, ExtendedNewmarkBeta.cpp:197); 1:
     * NewmarkBeta (NewmarkBeta.cpp:187). The internal force of step k sees the support DOFs
     * one increment late (SVL updates the strains before adding the increment, Algorithm.cpp:
     * 18-56). Elastic models with the Linear algorithm only. nsm = 0: none. */
    int32_t nsm;
    const int32_t *sm_node;   /* [nsm] mesh node (restrained in sm_dir) */
    const int32_t *sm_dir;    /* [nsm] 0-based direction */
    const double *sm_factor;  /* [nsm] combination factor */
    const int32_t *sm_ptr;    /* [nsm+1] offsets of each entry's values */
    const double *sm_values;  /* [sm_ptr[nsm]] displacement series */
    int32_t sm_offset;        /* 0 or 1, see above */
} svlg_mesh_desc;

typedef struct svlg_model svlg_model;

/* Execution options (no counterpart in SVL): node-tile size of the fused explicit
 * kernel, workgroup size, and the explicit kernel variant. Zero fields take defaults. */
typedef struct svlg_exec_opts {
    int32_t tile_nodes;   /* max nodes owned by one workgroup (default 160 in 3D, 448 in 2D) */
    int32_t block;        /* threads per workgroup of the colour-phase fused kernel (default 512) */
    int32_t variant;      /* 0 auto: 3 when the mesh is corner-unique and a tile fits one round,
                           * else 1; 1 fused node tiles with colour phases, 2 element-colored
                           * kernels, 3 fused node tiles with corner slots */
    int32_t geo_table;    /* parallelepipeds: 0 = share (J^-1, |J|) through a table of distinct
                           * geometries when it holds at most 65535 entries, -1 = never (stage
                           * node coordinates in LDS instead) */
    int32_t elem_precision; /* explicit element step: 0 = FP64 (the parity path), 1 = BF16 MFMA
                           * with FP32 accumulation (Hexa8 meshes with one element geometry and
                           * corner-slot tiles), 2 = FP32 modal element arithmetic (axis-aligned
                           * Hexa8 bricks, corner-slot tiles); displacements, the corner-slot
                           * sums and the updates stay FP64 */
    int32_t cg_launch;    /* CG solves of the Newmark / Newton steps: 0 = iterations replayed from
                           * captured hipGraphs (runs of 8, pairs, one), 1 = eager launches
                           * (the same kernels and iteration order: bitwise the same results) */
    int32_t reserved[2];
} svlg_exec_opts;

/* Error string of the last failing call on this thread. */
const char *svlg_last_error(void);
/* Library version string (also forces the HIP runtime to load). */
const char *svlg_version(void);
/* Number of HIP devices visible (0 on a machine without a GPU). */
int svlg_device_count(void);

/* Uploads a model (Mesh::Initialize + Element::SetDomain, 06-Mesh/Mesh.cpp:18-49),
 * builds the element coloring and node tiles, and assembles the lumped mass and the
 * dashpot damping diagonals on the device (Assembler.cpp:48-158). */
int svlg_model_create(const svlg_mesh_desc *desc, svlg_model **out);
/* Same with execution options (NULL = defaults). */
int svlg_model_create_ex(const svlg_mesh_desc *desc, const svlg_exec_opts *opts, svlg_model **out);
int svlg_model_destroy(svlg_model *m);
/* Uses `stream` (a hipStream_t) for every later launch of this model; NULL restores
 * the model's own stream. svlg_model_stream returns the stream in use. */
int svlg_model_set_stream(svlg_model *m, void *stream);
void *svlg_model_stream(svlg_model *m);
/* Host-only planning of the node tiles (no device needed): builds the tiles the fused
 * kernel would use for `tile_nodes` (0 = default), checks their invariants and returns
 * stats[0]=ntiles [1]=max owned nodes [2]=max staged nodes [3]=max elements per tile
 * [4]=sum of tile elements (elements are recomputed by every tile they touch)
 * [5]=sum of halo nodes [6]=LDS bytes per workgroup [7]=element colors. */
int svlg_plan_tiles(const svlg_mesh_desc *desc, int32_t tile_nodes, int64_t *stats8);
/* info[0]=ncolors info[1]=nelem info[2]=nnodes info[3]=nfree info[4]=explicit kernel
 * variant (1 fused + colour phases, 2 element-colored, 3 fused + corner slots) info[5]=ntiles info[6]=element path
 * of the fused kernel (0 Gauss points, 1 modal with staged coordinates, 2 modal with
 * the geometry table, 3 modal with a table of diagonal J^-1: axis-aligned elements) info[7]=dynamic LDS bytes of the fused kernel
 * info[8]=explicit element precision (0 FP64, 1 BF16 MFMA, 2 FP32) info[9]=distinct element geometries */
int svlg_model_info(svlg_model *m, int64_t *info10);
/* Element color of every element (ascending-tag greedy coloring; no two elements of
 * one color share a node). */
int svlg_element_colors(svlg_model *m, int32_t *colors /*[nelem]*/);

/* ---- Assembler (Assembler.hpp:50-181) ---------------------------------------- */
/* Lumped mass diagonal per free DOF (Assembler::ComputeMassMatrix, row-sum lumping
 * lin2DQuad4.cpp:272-282 / lin3DHexa8.cpp:272-282). */
int svlg_mass_diagonal(svlg_model *m, double *mdiag /*[nfree]*/);
/* Dashpot damping diagonal per free DOF (Assembler::ComputeDampingMatrix restricted
 * to ZeroLength1D dashpots, ZeroLength1D.cpp:211-234). */
int svlg_damping_diagonal(svlg_model *m, double *cdiag /*[nfree]*/);
/* Fint(U) in free DOFs: Assembler::ComputeInternalForceVector (Assembler.cpp:240-269)
 * with the element stress state taken from U (lin*::UpdateState then
 * ComputeInternalForces). `ftol` is the |Fint_e(j)| > ftol filter of Assembler.cpp:262. */
int svlg_internal_force(svlg_model *m, const double *U /*[nfree]*/, double ftol,
                        double *F /*[nfree]*/);

/* ---- CentralDifference (10-Integrators/02-CentralDifference/CentralDifference.cpp) */
/* Initialize (:36-71): zero state, Keff = M/dt^2 + C/(2dt) (diagonal), ftol filter. */
int svlg_cd_init(svlg_model *m, double dt, double ftol);
/* Advances nsteps steps starting at step index k0 (ComputeNewStep :124-152 via
 * Linear::ComputeNewIncrement, with DynamicAnalysis' index convention: step k uses
 * load sample k). If nrec > 0, U at the DOFs rec_dofs[] after every step is written to
 * hist[s*nrec + r] (the NODE recorder's row k-1); rec_dofs[r] >= 0 is a free DOF index,
 * -1 - t the total DOF t = node*ndim + c of a mesh node (restrained: 0 or its support motion). */
int svlg_cd_steps(svlg_model *m, int32_t k0, int32_t nsteps, int32_t nrec,
                  const int32_t *rec_dofs, double *hist);
/* Same as svlg_cd_steps but leaves results on the device and returns immediately
 * after enqueueing (no recording, no host synchronisation). */
int svlg_cd_steps_async(svlg_model *m, int32_t k0, int32_t nsteps);
/* Current U, V, A in free DOFs (CentralDifference::GetDisplacements/Velocities/
 * Accelerations). Any pointer may be NULL. */
int svlg_cd_get_state(svlg_model *m, double *U, double *V, double *A);
/* The pieces of the step for callers that keep SVL's Algorithm and LinearSystem (Linear::
 * ComputeNewIncrement, Linear.cpp:22-56, calls them in this order); svlg_cd_steps fuses them.
 * CentralDifference::ComputeEffectiveForce (:206-221) for step k at the current state:
 * Feff = T^T [Fext(k) - Fint(U) + (M/dt^2 - C/2dt)(U - Up)] over the free DOFs (Fbar, the force
 * left by a previous analysis stage, is 0: one stage). `factor` is unused, as in SVL. */
int svlg_cd_effective_force(svlg_model *m, int32_t k, double factor, double *Feff /*[nfree]*/);
/* CentralDifference::ComputeEffectiveStiffness (:225-231): Keff = T^T (M/dt^2 + C/2dt) T, diagonal
 * for the lumped mass and dashpots; its diagonal over the free DOFs. */
int svlg_cd_effective_stiffness(svlg_model *m, double *Kdiag /*[nfree]*/);
/* The state update of CentralDifference::ComputeNewStep (:137-148) for step k with the
 * algorithm's increment dU (free DOFs): Up = U, U += T dU + SupportMotion (a_k at the support-
 * motion DOFs); V and A follow from the kept U history. */
int svlg_cd_apply_increment(svlg_model *m, int32_t k, const double *dU /*[nfree]*/);
/* Integrator::ComputeProgressiveForce (CentralDifference.cpp:174-186; Fbar = 0): Fext(k) at
 * the mesh nodes, [nnodes*ndim] in description order (Assembler::ComputeExternalForceVector,
 * Assembler.cpp:291-489: point loads and domain reduction forces). */
int svlg_external_force(svlg_model *m, int32_t k, double *F);
/* Blocks until all work queued on the model's stream is done. */
int svlg_sync(svlg_model *m);

/* Reactions at the mesh nodes nodes[0..n) (description order) for the current state of the
 * integrator initialised last, after step k (DynamicAnalysis::UpdateDomain,
 * DynamicAnalysis.cpp:130-150 -> Integrator::ComputeReactionForce, CentralDifference.cpp:156-171):
 * R = Fint + C V + M A - Fext(k) assembled over the node's elements
 * (Assembler::ComputeDynamicInternalForceVector, Assembler.cpp:273-287), R [n*ndim]. The
 * restrained partner node of a dashpot carries -eta V of its mesh node. */
int svlg_reactions(svlg_model *m, int32_t n, const int32_t *nodes, int32_t k, double *R);

/* ---- Independent meshes advanced together (no SVL counterpart: SVL runs one model per
 * process) --------------------------------------------------------------------------
 * A model whose description is the disjoint union of nparts meshes (node ranges
 * part_node_ptr[p]..part_node_ptr[p+1] in description order, sharing no element) is stepped
 * by one fused launch per step over all of them; each part keeps its own dt (CentralDifference
 * coefficients per DOF, CentralDifference.cpp:64-70) and reads sample k of its own load series. */
int svlg_cd_init_batch(svlg_model *m, int32_t nparts, const int32_t *part_node_ptr /*[nparts+1]*/,
                       const double *dts /*[nparts]*/, double ftol);
/* Advances nsteps steps from step index k0 by replaying a captured hipGraph of graph_steps
 * fused launches (a multiple of 3; captured on first use, reused while the size and the U ring
 * slot match), the remainder eagerly. Asynchronous like svlg_cd_steps_async. */
int svlg_cd_graph_steps(svlg_model *m, int32_t k0, int32_t nsteps, int32_t graph_steps);
/* Autotuner of the node-tile size: builds the model once per candidate tile_nodes (other
 * options from base, may be NULL), times `steps` CentralDifference steps with HIP events and
 * returns the fastest candidate in best_tile_nodes and every candidate's time in us_per_step
 * (may be NULL; -1 where the candidate could not be built). nparts/part_node_ptr/dts as
 * svlg_cd_init_batch (nparts = 1: one mesh, dts[0]). */
int svlg_autotune_tiles(const svlg_mesh_desc *desc, const svlg_exec_opts *base, const int32_t *cand, int32_t ncand,
                        int32_t nparts, const int32_t *part_node_ptr, const double *dts, int32_t steps,
                        int32_t *best_tile_nodes, double *us_per_step);

/* ---- ExtendedNewmarkBeta + Linear + LinearSystem ---------------------------------
 * (10-Integrators/03-Newmark/ExtendedNewmarkBeta.cpp, gamma = 1/2, beta = 1/4, G = 0;
 * 09-Algorithms/01-Linear/Linear.cpp:22-56; 11-Solvers/LinearSystem.hpp:44-66.)
 * Needs a corner-unique mesh (explicit variant 3). The mass is the lumped diagonal or SVL's
 * consistent element mass (massform CONSISTENT, its default: Driver.hpp:1998,
 * lin3DHexa8.cpp:243-285), applied element by element; models with a domain reduction load
 * need the lumped mass. Keff = K + 4/dt^2 M + 2/dt C is applied matrix-free; the linear system
 * is solved by preconditioned conjugate gradients to ||b - Keff x|| <= rtol ||b||. */
/* Initialize (:36-63 + ComputeEffectiveStiffness :233-245): zero state, Keff terms and
 * the preconditioner (SVLG_PRECOND_JACOBI). */
int svlg_nm_init(svlg_model *m, double dt, double rtol, int32_t maxit, int32_t precond);
/* Advances nsteps steps from step index k0 (ComputeNewStep :123-154, Feff :208-229 with
 * the Linear algorithm). Recording as svlg_cd_steps; iters[s] and relres[s] (may be
 * NULL) receive the CG iteration count and the final relative residual of step s. */
int svlg_nm_steps(svlg_model *m, int32_t k0, int32_t nsteps, int32_t nrec, const int32_t *rec_dofs,
                  double *hist, int32_t *iters, double *relres);
/* ExtendedNewmarkBeta::ComputeEffectiveForce (:208-229) for step k at the current state, the
 * Linear algorithm's dU = 0 (Fbar = 0, G = 0): Feff = T^T [Fext(k) - Fint(U) + C V + M (4/dt V + A)]
 * over the free DOFs. `factor` is unused, as in SVL. Models with J2 elements form their Newton
 * residual in svlg_nl_steps. */
int svlg_nm_effective_force(svlg_model *m, int32_t k, double factor, double *Feff /*[nfree]*/);
/* The state update of ExtendedNewmarkBeta::ComputeNewStep (:123-154) for step k with the
 * algorithm's increment dU (free DOFs): U += T dU + SupportMotion, A = 4/dt^2 dU - 4/dt V - A,
 * V = 2/dt dU - V (V and A from the free increment alone, as in SVL). */
int svlg_nm_apply_increment(svlg_model *m, int32_t k, const double *dU /*[nfree]*/);
/* Integrator::ComputeSupportMotionVector for step k of the initialised integrator: Feff -=
 * T^T (K a_k) with a_k the step's support increments and K the total effective stiffness
 * (ExtendedNewmarkBeta.cpp:191-204: K + 4/dt^2 M + 2/dt C; CentralDifference.cpp:190-203:
 * M/dt^2 + C/2dt, diagonal on the device, so Feff is unchanged). `factor` is unused, as in
 * SVL. Feff [nfree] is read and written. */
int svlg_support_motion_vector(svlg_model *m, int32_t k, double factor, double *Feff);
/* Integrator::GetDisplacements: the total displacement vector of the initialised integrator
 * over the mesh nodes, [nnodes*ndim] in description order (restrained DOFs: 0, or their
 * support motion). */
int svlg_get_displacements(svlg_model *m, double *U);
/* Current U, V, A in free DOFs (ExtendedNewmarkBeta::GetDisplacements/...). */
int svlg_nm_get_state(svlg_model *m, double *U, double *V, double *A);
/* Switches the preconditioner of later solves (SVLG_PRECOND_LEARNED needs svlg_gnn_init;
 * it runs flexible CG). No SVL counterpart: SVL selects a LinearSystem at start-up
 * (Driver.hpp:1768-1783). */
int svlg_nm_set_precond(svlg_model *m, int32_t precond);
/* y = Keff x over the free DOFs (the matrix of LinearSystem::SetMatrix, applied). */
int svlg_keff_apply(svlg_model *m, const double *x /*[nfree]*/, double *y /*[nfree]*/);
/* y = Keff x as svlg_keff_apply, allocating nothing: no counterpart in SVL. For callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). The work vectors live in a fixed device array of 3 x 2^20 doubles (24 MiB,
 * held by every GPU the library runs on until the process ends); models with more than 2^20
 * node DOFs (nodes x ndim) are refused. Calls that overlap on one GPU share that array and race.
 * Same kernels, so the same bits as svlg_keff_apply. */
int svlg_keff_apply_noalloc(svlg_model *m, const double *x /*[nfree]*/, double *y /*[nfree]*/);
/* Solves Keff x = b from x = 0 (LinearSystem::SetVector + SolveSystem + GetSolution). */
int svlg_solve(svlg_model *m, const double *b /*[nfree]*/, double *x /*[nfree]*/, double rtol,
               int32_t maxit, int32_t *iters, double *relres);

/* ---- NewmarkBeta + NewtonRaphson (models with J2 elements) --------------------------
 * (10-Integrators/03-Newmark/NewmarkBeta.cpp:115-144, 198-235, gamma = 1/2, beta = 1/4;
 * 09-Algorithms/02-Newton/NewtonRaphson.cpp:24-75; Algorithm.cpp:18-56, 122-188;
 * 02-Materials/02-NonLinear/Plastic3DJ2.cpp, PlasticPlaneStrainJ2.cpp.)
 * Each Newton iteration forms Feff = Fext - Fint + M (4/dt V + A - 4/dt^2 dU) + C (V - 2/dt dU)
 * from the Gauss-point stresses, solves (K_tangent + 4/dt^2 M + 2/dt C) du = Feff by
 * Jacobi-preconditioned CG to cg_rtol, sets dU += du, evaluates ComputeConvergence(cnvgtest)
 * and updates the Gauss-point states at U + dU, while residual > cnvgtol and iterations < nmax.
 * C = am M + ak K_initial + dashpots. Elastic elements take part unchanged. */
/* Initialize: zero state, J2 Gauss-point states at rest, Keff terms. cnvgtest = 1..8 as
 * SVL's flags (1 UnbalanceForce ... 5 RelativeIncrementalDisplacement ... 8 none). */
int svlg_nl_init(svlg_model *m, double dt, double cg_rtol, int32_t cg_maxit, double cnvgtol, int32_t nmax,
                 int32_t cnvgtest);
/* Advances nsteps steps from step index k0; recording as svlg_cd_steps. Optional outputs
 * (NULL to skip): newton_iters[s] = Newton iterations of step s; newton_res[s * nmax + i] =
 * convergence-test value of its iteration i (entries past the last iteration are left as
 * they are); cg_iters[s] = CG iterations summed over its Newton solves. The preconditioner
 * is the one svlg_nm_set_precond selected (the learned one approximates the initial
 * effective stiffness). State: svlg_nm_get_state. */
int svlg_nl_steps(svlg_model *m, int32_t k0, int32_t nsteps, int32_t nrec, const int32_t *rec_dofs,
                  double *hist, int32_t *newton_iters, double *newton_res, int32_t *cg_iters);
/* Stresses of the Gauss points of J2 element elem (lin*::GetStress): [ngp][3] (xx, yy, xy)
 * for Quad4 plane strain, [ngp][6] (xx, yy, zz, xy, yz, zx) for Hexa8. */
int svlg_gp_stress(svlg_model *m, int32_t elem, double *stress);

/* ---- Learned preconditioner (no SVL counterpart; it replaces the factorisation inside
 * LinearSystem::SolveSystem with a trained preconditioner for flexible CG) ------------
 * A linear graph convolution network over Ehat = D^-1/2 Keff D^-1/2 (D = diag Keff):
 *   x0 = s r, H^1 = x0 c_1^T, H^l = H^(l-1) A_l + (Ehat H^(l-1)) B_l + x0 c_l^T, z = s (H^L w)
 * with s = D^-1/2, C channels and L layers. theta = for l = 1..L [A_l (CxC) | B_l (CxC) |
 * c_l (C)], then w (C); size L (2C^2 + C) + C. Needs svlg_nm_init (its dt fixes Keff). */
/* channels in [1,4], layers in [2,16]; mfma: 1 = BF16 MFMA element step (Hexa8, one element
 * geometry), 0 = FP64 element routine, -1 = MFMA when possible; theta NULL = Chebyshev
 * initialisation on power-iteration bounds of the spectrum of Ehat. */
int svlg_gnn_init(svlg_model *m, int32_t channels, int32_t layers, int32_t mfma, const double *theta);
/* info[0]=channels [1]=layers [2]=MFMA in use [3]=parameter count [4], [5]=validation
 * flexible-CG iterations before / after the last training; bounds = Lanczos (lmin, lmax)
 * of Ehat. */
int svlg_gnn_info(svlg_model *m, int32_t *info6, double *bounds2);
int svlg_gnn_get_params(svlg_model *m, double *theta);
int svlg_gnn_set_params(svlg_model *m, const double *theta);
/* z = P r over the free DOFs (mfma selects the element step of this call). */
int svlg_gnn_apply(svlg_model *m, int32_t mfma, const double *r /*[nfree]*/, double *z /*[nfree]*/);
/* Mean energy-norm loss ||P r_p - e_p||^2_Keff / ||e_p||^2_Keff over npairs pairs (r_p, e_p),
 * [npairs][nfree] each with Keff e_p = r_p, and its gradient w.r.t. theta (FP64 element
 * routine, hand-written backward kernels). grad may be NULL. */
int svlg_gnn_loss_grad(svlg_model *m, int32_t npairs, const double *r, const double *e, double *loss,
                       double *grad);
/* Trains theta with Adam (learning rate lr) for `epochs` full-batch epochs on pairs from
 * the model's own Newmark solves: steps 1..nsteps from rest, every `every`-th step gives
 * (Feff, dU), `sweeps` Jacobi-smoothed pairs and one white-noise pair (Keff e, e). The
 * trained theta is kept only if it lowers the flexible-CG iterations of a validation solve
 * (*accepted = 1), else the previous theta is restored. The state is back at rest afterwards.
 * loss_hist[epochs] and accepted may be NULL. */
int svlg_gnn_train(svlg_model *m, int32_t nsteps, int32_t every, int32_t sweeps, int32_t epochs, double lr,
                   double *loss_hist, int32_t *accepted);

/* Device-to-device copy of n doubles with a plain coalesced kernel on `stream` (a
 * hipStream_t, NULL = default stream). No counterpart in SVL: it gives the bytes that
 * HBM traffic counters read for a transfer of known size and the copy bandwidth the
 * device sustains. `dst` and `src` are device pointers. */
int svlg_stream_copy(double *dst, const double *src, int64_t n, void *stream);

/* ---- Perfectly matched layers (PML2DQuad4 / PML3DHexa8 + regular elements + EQUAL) ------
 * A mixed model: regular nodes carry ndim unknowns, PML nodes 5 in 2D (ux, uy and the
 * stress-like xx, yy, xy; 04-Elements/06-Quadrilateral/PML2DQuad4.cpp:11-12, 233-466) or 9 in
 * 3D (ux, uy, uz, xx, yy, zz, xy, yz, xz; 04-Elements/10-Hexahedron/PML3DHexa8.cpp:214-715),
 * the regions tied by EQUAL constraints (06-Mesh/Mesh.cpp:328-383). Integrated with
 * ExtendedNewmarkBeta + Linear (10-Integrators/ExtendedNewmarkBeta.cpp:123-154, 208-245): in 3D
 * the history matrix G and the integrated displacement Ubar enter Keff (+ dt/3 G), Feff
 * (- G (Ubar + dt U + dt^2/6 V)) and the update; the 2D element has no G. The effective matrix
 * is symmetric and indefinite (negative stress blocks): the device assembles it once, scales it
 * symmetrically by |diag|^-1/2 and solves with MINRES where SVL factorises it with Eigen
 * (11-Solvers/EigenSolver.cpp). */
#define SVLG_CLASS_LIN2DQUAD4 0
#define SVLG_CLASS_PML2DQUAD4 1
#define SVLG_CLASS_LIN3DHEXA8 2
#define SVLG_CLASS_PML3DHEXA8 3

typedef struct svlg_mixed_desc {
    int32_t ndim;              /* 2 (classes 0, 1) or 3 (classes 2, 3) */
    int32_t nnodes;
    const double *coords;      /* [nnodes][ndim] */
    const int32_t *dof_ptr;    /* [nnodes+1] first total DOF of each node */
    const int32_t *dof_free;   /* [ntotal] free DOF of each total DOF, -1 when restrained */
    const double *dof_factor;  /* [ntotal] its entry of T: 1, or the factor of an EQUAL slave */
    int32_t nelem;
    const int32_t *elem_class; /* [nelem] SVLG_CLASS_* */
    const int32_t *conn;       /* [nelem][4 in 2D, 8 in 3D] node indices in SVL's corner order */
    const double *E, *nu, *rho, *th; /* [nelem] material (plane strain in 2D), thickness (2D) */
    const double *pml;         /* [nelem][9] m, L, R, x0, y0, z0, nx, ny, nz (PML elements;
                                  z0, nz unused in 2D) */
    int32_t lumped;            /* regular-element mass: 1 row-sum lumped, 0 consistent; the
                                  PML mass is always consistent (PML2DQuad4.cpp:286) */
    int32_t nload, load_nt;    /* POINTLOAD_CONCENTRATED_DYNAMIC on total DOFs */
    const int32_t *load_dof;   /* [nload] */
    const double *load_factor; /* [nload] */
    const double *load_series; /* [nload][load_nt] value of step k at column k */
} svlg_mixed_desc;

typedef struct svlg_pml svlg_pml;

/* Computes the element matrices, assembles M, C, K, G and Keff = K + 4/dt^2 M + 2/dt C
 * + dt/3 G over the free DOFs (T^T A T) on the device. Replaces Assembler::ComputeMassMatrix / DampingMatrix /
 * StiffnessMatrix (Assembler.hpp:50-181) and ExtendedNewmarkBeta::Initialize for this model. */
int svlg_pml_create(const svlg_mixed_desc *desc, double dt, double rtol, int32_t maxit, svlg_pml **out);
int svlg_pml_destroy(svlg_pml *p);
/* info4 = nfree, ntotal, nonzeros of the free-space matrices, element-matrix contributions. */
int svlg_pml_info(svlg_pml *p, int64_t *info4);
/* Steps k0 .. k0+nsteps-1 (ExtendedNewmarkBeta::ComputeNewStep, load sample k of step k).
 * Records U, V, A (each pointer may be NULL) of the total DOFs rec_dofs after every step, row
 * s for step k0+s, and the MINRES iterations per step. Stop when a solve misses rtol within
 * maxit iterations. */
int svlg_pml_steps(svlg_pml *p, int32_t k0, int32_t nsteps, int32_t nrec, const int32_t *rec_dofs,
                   double *histU, double *histV, double *histA, int32_t *iters);
int svlg_pml_get_state(svlg_pml *p, double *U, double *V, double *A /* [nfree] each */);
/* Execution option (no SVL counterpart): on = 1 (default) replays the MINRES iterations of a
 * step from one captured hipGraph per handle, on = 0 launches them eagerly. */
int svlg_pml_set_graph(svlg_pml *p, int32_t on);
/* ExtendedNewmarkBeta::GetPMLHistoryVector (Integrator.hpp:90): Ubar over the free DOFs (0
 * without PML3DHexa8 elements, which alone carry the history matrix G). */
int svlg_pml_get_history(svlg_pml *p, double *Ubar /* [nfree] */);

#ifdef __cplusplus
}
#endif

#endif /* SVLG_H */
