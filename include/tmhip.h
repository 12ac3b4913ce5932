/*
 * tmhip.h - C ABI of the MI355X cyclic-symmetry modal solve engine.
 *
 * This is the drop-in boundary below turbomodal's C++ cyclic solver.  The
 * reference runs the whole pipeline of CyclicSymmetrySolver::solve_at_rpm
 * (src/cyclic_solver.cpp:725-1357) on the CPU with Eigen/Spectra; here the
 * host object model (Mesh, Material, CyclicSymmetrySolver, ModalResult) calls
 * these entry points and every numeric stage runs as gfx950 HIP kernels:
 *
 *   reference operator                              replaced by
 *   ---------------------------------------------   -------------------------
 *   GlobalAssembler::assemble / assemble_rotating_  tm_operator_create
 *     effects / assemble_centrifugal_load /           (element kernels, fixed-
 *     assemble_stress_stiffening                      order segmented assembly,
 *     (include/turbomodal/assembler.hpp:9-41)         static prestress solve)
 *   CyclicSymmetrySolver::static_centrifugal         tm_operator_create
 *     (src/cyclic_solver.cpp:456-558)
 *   precompute_cyclic_projections                    tm_mesh_create (symbolic)
 *     (src/cyclic_solver.cpp:560-723)                  + tm_operator_create
 *   SymShiftInvertLDLT, HermitianLanczosEigen-       tm_solve_batch
 *     Solver::solve, ModalSolver::solve_real /
 *     solve_complex_hermitian, solve_harmonic
 *     (include/turbomodal/sym_shift_invert_ldlt.hpp:25-112,
 *      src/hermitian_lanczos.cpp:12-460,
 *      src/modal_solver.cpp:12-147,
 *      src/cyclic_solver.cpp:937-1224)
 *
 * Conventions: plain C types only, row-major host arrays, every call returns
 * TM_OK (0) or a negative status; tm_last_error() holds the message.  No C++
 * exception crosses this boundary.  A context owns one HIP stream on one
 * device; calls on one context are serialised by the caller.
 */
#ifndef TMHIP_H
#define TMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TM_OK 0
#define TM_ERR_ARG -1        /* invalid argument (reference: std::invalid_argument) */
#define TM_ERR_RUNTIME -2    /* topology/runtime error (reference: std::runtime_error) */
#define TM_ERR_HIP -3        /* HIP runtime failure */
#define TM_ERR_NUMERIC -4    /* zero pivot / breakdown that could not be recovered */
#define TM_ERR_MEMORY -5     /* device memory exhausted or admission refused */

/* node classes, as CyclicSymmetrySolver::classify_dofs (src/cyclic_solver.cpp:214-238) */
#define TM_NODE_INTERIOR 0
#define TM_NODE_LEFT 1
#define TM_NODE_RIGHT 2

/* precision of the shift-invert factor */
#define TM_PREC_FP64 0       /* complex FP64 factor, FP64 everywhere */
#define TM_PREC_MIXED 1      /* complex FP32 factor (FP32 MFMA Schur updates, factor stored in
                                FP32), every shift-invert solve refined in FP64 against FP64
                                K(k), M(k) until it stalls; FP64 Rayleigh-Ritz and residuals.
                                Block method on the (K, M) pencil only (not the Lanczos
                                method, TM_CONV_REFERENCE, the Coriolis pencil or the server):
                                those requests return TM_ERR_ARG */

/* eigen-solver convergence policy */
#define TM_CONV_TIGHT 0      /* relative Ritz residual <= tol_rel (default 1e-12) */
#define TM_CONV_REFERENCE 1  /* |beta s_i| < 1e-8 max(1,|theta_i|) (src/hermitian_lanczos.cpp:414-424) */

/* Eigensolver (tm_problem.method) */
#define TM_METHOD_LANCZOS 0  /* single-vector shift-invert Lanczos, exact-shift implicit restarts
                                (src/hermitian_lanczos.cpp), conv_policy applies */
#define TM_METHOD_BLOCK 1    /* block shift-invert Krylov (8 or 16 vectors per triangular sweep),
                                FP64 Rayleigh-Ritz with K(k), M(k); converged when
                                ||K x - lambda M x|| / ||K x|| <= max(5e-9, 4 x its FP64 floor),
                                or below 1e-8 and no longer improving.
                                Problems with fewer than 256 free DOFs, and TM_CONV_REFERENCE,
                                run the Lanczos method */

typedef struct tm_context tm_context;
typedef struct tm_mesh tm_mesh;
typedef struct tm_operator tm_operator;

/* Context: one device, one stream.  device < 0 selects the current device. */
int tm_context_create(int device, tm_context** out);
void tm_context_destroy(tm_context* ctx);
const char* tm_last_error(const tm_context* ctx);
/* Per-kernel timing with HIP events on the context stream (0 = off). */
int tm_set_profiling(tm_context* ctx, int enabled);

typedef struct tm_kernel_stat {
  char name[48];
  int64_t launches;
  double total_ms;         /* sum of event-timed durations */
  double alg_bytes;        /* algorithmic HBM bytes, summed over launches */
  double alg_flops;        /* algorithmic flops, summed over launches */
} tm_kernel_stat;
/* Copies up to max_stats entries; returns the number of entries available. */
int tm_kernel_stats(tm_context* ctx, tm_kernel_stat* stats, int max_stats);
int tm_reset_kernel_stats(tm_context* ctx);

/*
 * Mesh + cyclic topology upload and symbolic analysis (integer work, once per
 * mesh): reduced/free node map, node-block CSR pattern of K(k), segmented
 * assembly slot maps, nested-dissection ordering, assembly tree, supernodal
 * fronts.  Mirrors Mesh (include/turbomodal/mesh.hpp:40-105) after
 * match_boundary_nodes and CyclicSymmetrySolver::classify_dofs.
 *   nodes           n_nodes x 3 float64
 *   elements        n_elems x 10 int32, gmsh TET10 order
 *   node_class      n_nodes int32, TM_NODE_*
 *   right_partner   n_nodes int32: left partner of each right node, else -1
 *   constrained     3*n_nodes uint8: 1 = DOF removed (build_constrained_dofs)
 */
int tm_mesh_create(tm_context* ctx, const double* nodes, int32_t n_nodes,
                   const int32_t* elements, int32_t n_elems,
                   const int32_t* node_class, const int32_t* right_partner,
                   const uint8_t* constrained, int32_t num_sectors,
                   int32_t rotation_axis, tm_mesh** out);
void tm_mesh_destroy(tm_mesh* mesh);

typedef struct tm_mesh_info {
  int32_t n_nodes, n_elems, n_reduced_nodes, n_free_nodes;
  int32_t n_free;            /* free DOFs (the reference's n_free_) */
  int32_t n_dof_system;      /* 3 * n_free_nodes: order of the factored system */
  int64_t nnz_blocks;        /* 3x3 blocks in the K(k) pattern (both triangles) */
  int64_t nnz_L;             /* complex entries stored in the supernodal factor */
  double factor_flops;       /* real flops of one complex LDL^H factorization */
  int32_t n_supernodes, tree_levels, max_front;
  int64_t front_bytes_fp64;  /* device bytes of one problem's factor workspace (FP64) */
} tm_mesh_info;
int tm_mesh_get_info(const tm_mesh* mesh, tm_mesh_info* info);

/*
 * Operator = mesh + material: element kernels (Ke, Me, Kw at unit omega, G,
 * centrifugal load at unit omega), fixed-order segmented assembly straight
 * into the phase-split cyclic blocks (K_c, K_p, M_c, M_p, G_c, G_p), and -
 * when prestress != 0 - the k = 0 static solve K u = F at unit omega followed
 * by the stress-stiffening kernel.  K_eff(omega) = K + omega^2 (Ks1 - Kw1)
 * exactly (Ks is linear in u, u in F, F in omega^2), so one operator serves
 * every rpm.  Material validation follows src/material.cpp:15-22.
 */
int tm_operator_create(tm_context* ctx, tm_mesh* mesh, double E, double nu,
                       double rho, int32_t prestress, tm_operator** out);
void tm_operator_destroy(tm_operator* op);

/* One eigenproblem: (operator, rpm, harmonic k, num_modes). */
typedef struct tm_problem {
  tm_operator* op;
  double rpm;
  int32_t harmonic;
  int32_t num_modes;
  int32_t precision;         /* TM_PREC_* */
  int32_t conv_policy;       /* TM_CONV_* */
  double shift;              /* sigma; the reference uses 1.0 */
  double tol_rel;            /* TM_CONV_TIGHT threshold (<= 0: 1e-12) */
  int32_t max_restarts;      /* <= 0: 1000 as SolverConfig (modal_solver.hpp:39) */
  int32_t method;            /* TM_METHOD_* */
  double min_frequency;      /* Hz; > 0: solve 6 extra modes and drop f <= min_frequency
                                (src/cyclic_solver.cpp:1011, 1058-1084) */
  int32_t coriolis;          /* != 0 with rpm != 0 and harmonic > 0: the gyroscopic QEP
                                (K + w D - w^2 M) phi = 0, D = i 2 Omega G(k), through its
                                Lancaster pencil (src/cyclic_solver.cpp:986-1005,
                                src/modal_solver.cpp:149-312); forward and backward modes
                                of smallest |w|, whirl = sign(w); block method, >= 64 free DOFs */
  int32_t warm_start;        /* != 0: rpm continuation in a sweep.  A block-method problem
                                (not the Coriolis pencil) may start its Krylov basis from the
                                converged Ritz vectors of the problem of the same operator,
                                harmonic and mode count at the adjacent rpm of the same call,
                                solved in an earlier wave of its slot group, and then checks
                                convergence from its fifth block on.  Results match the
                                oracle to the same bars; they differ from a cold start at
                                rounding level.  tm_solve_batch only (the server ignores it) */
} tm_problem;

/* Result of one problem, in the reference's ModalResult layout
 * (include/turbomodal/modal_solver.hpp:8-31).  Caller-owned buffers sized
 * from num_modes and 3*n_nodes; any pointer may be NULL to skip. */
typedef struct tm_result {
  double* frequencies;       /* [num_modes] Hz, after degenerate duplication */
  double* mode_shapes;       /* [3*n_nodes][n_modes] complex128 (re,im pairs), row-major,
                                row stride n_modes (<= num_modes) */
  int32_t* whirl;            /* [num_modes] */
  double* eigenvalues;       /* [num_modes] lambda of the reduced problem (before duplication) */
  double* residuals;         /* [num_modes] ||K x - lambda M x|| / ||K x|| (free space, FP64) */
  double* modal_mass;        /* [num_modes] N * Re(phi^H M_sector phi), kg (python/turbomodal/solver.py:91-103) */
  int32_t n_modes;           /* out: modes written */
  int32_t n_unique;          /* out: eigenpairs solved (nev_target) */
  int32_t converged;         /* out */
  int32_t restarts;          /* out: implicit restarts used */
  int32_t status;            /* out: TM_OK or failure (harmonic dropped by caller) */
  double seconds;            /* out: wall time of the batch this problem ran in */
  int32_t factor_retries;    /* out: solves repeated alone after a failure (tm_solve_batch) */
  int32_t pivot_flags;       /* out: status of the last factorization of the problem: bit 0
                                a zero or non-finite pivot, bit 1 a near-singular pivot
                                (|d_j| <= 1e-10 |a_jj| and <= 1e-3 |shift M_jj|: the shift
                                sits on an eigenvalue); either fails that attempt */
  double shift_used;         /* out: shift of the factorization behind the result: the
                                problem's shift, moved by 1 % of max(|shift|, 1) per retry
                                of a problem whose factor was flagged; 0 for the Coriolis
                                pencil (it factors K_eff(k)) */
} tm_result;

/* Solve a batch: problems sharing an operator run together (one symbolic
 * analysis, batched numeric factorizations and Lanczos).  Blocking.
 * A problem that fails is solved again alone, as CyclicSymmetrySolver::solve_at_rpm
 * retries failed harmonics (src/cyclic_solver.cpp:1334-1342); where its LDL^H met a
 * zero, non-finite or near-singular pivot (pivot_flags != 0), the retry moves the
 * shift (up to three times), where the reference falls back to LU on a zero pivot
 * (src/hermitian_lanczos.cpp:27-55).  Still failing: status != TM_OK. */
int tm_solve_batch(tm_context* ctx, const tm_problem* problems, int32_t n,
                   tm_result* results);

/* Continuous-batching server (serving requests of uneven size as they arrive;
 * no counterpart in the reference, whose solve_at_rpm is a blocking call per
 * rpm, src/cyclic_solver.cpp:725-1357).  A scheduler thread owns the context:
 * submitted problems are admitted into slot groups (one per operator and block
 * size, capacity from tm_problem_footprint against hbm_fraction of the free
 * device memory) as slots and memory free up; their factorizations run
 * batched, masked to the new slots, and their block-Krylov steps share the
 * launches of the problems already in flight.  Results land in the caller's
 * tm_result buffers, which must stay valid until tm_server_wait reports done.
 * A problem's result does not depend on what it shares the device with.
 * While a server exists, tm_solve_batch on its context fails. */
typedef struct tm_server tm_server;
int tm_server_create(tm_context* ctx, int32_t max_slots /* per group, <= 64 */,
                     double hbm_fraction /* of free device memory, e.g. 0.85 */,
                     tm_server** out);
int tm_server_submit(tm_server* s, const tm_problem* problem, tm_result* result,
                     uint64_t* ticket);
/* Blocks until the problem is done (*done = 1, the ticket is released) or
 * timeout_s passes (*done = 0; timeout_s < 0 waits without limit). */
int tm_server_wait(tm_server* s, uint64_t ticket, double timeout_s, int32_t* done);
/* Finishes every submitted problem, then stops the scheduler. */
void tm_server_destroy(tm_server* s);

/* Device bytes one problem holds while in flight (factor + Krylov basis);
 * the serving loop admits requests against this. */
int64_t tm_problem_footprint(const tm_operator* op, int32_t num_modes,
                             int32_t precision, int32_t harmonic, int32_t method);
int64_t tm_device_free_bytes(tm_context* ctx);
/* Fraction of the device memory that tm_solve_batch keeps out of its batch
 * sizing (default 0.15): batches hold at most (1 - reserve) x (free memory +
 * the context's cached blocks) / tm_problem_footprint problems, and at least
 * one.  The device counterpart of solve_at_rpm's memory_reserve_fraction
 * (src/python_bindings.cpp:360), which caps concurrent harmonics against
 * (1 - reserve) x host RAM (src/cyclic_solver.cpp:1257-1272).  Values outside
 * [0, 1] return TM_ERR_ARG.  Applies to later tm_solve_batch calls on ctx;
 * results do not depend on it, only how many problems share a batch. */
int tm_set_memory_reserve(tm_context* ctx, double reserve_fraction);

/* Diagnostics used by the parity tests: the assembled free-space matrices
 * K(k) (at rpm) and M(k) in COO form over free DOFs in the reference's free
 * ordering (interior DOFs, then left DOFs, node-ascending; constrained DOFs
 * removed).  Call with rows == NULL to get the entry count. */
int tm_operator_export(tm_operator* op, double rpm, int32_t harmonic,
                       int32_t which /* 0 = K_eff(k), 1 = M(k), 2 = G(k) = G_c + P G_p - conj(P) G_p^T,
                                        the skew projection behind D(k) = i 2 Omega G(k) */,
                       int64_t* nnz, int32_t* rows, int32_t* cols, double* vals_c128);
/* Static prestress displacement at unit omega (3*n_nodes), the reference's
 * static_centrifugal(1.0). */
int tm_operator_static_u(tm_operator* op, double* u_full);

/* Grounded springs of ELASTIC_SUPPORT constraint groups: spring_full holds 3*n_nodes
 * stiffnesses (N/m, >= 0) added to the diagonal of the sector stiffness K before the
 * cyclic projection, as CyclicSymmetrySolver::build_spring_stiffness and solve_at_rpm
 * step 1b (src/cyclic_solver.cpp:193-212, 747-751); a right-boundary node's spring acts
 * on its left partner through the rotation R(alpha).  The static prestress solve keeps
 * the base K, as the reference's static_centrifugal does (src/cyclic_solver.cpp:456-558).
 * Springs accumulate over calls; call before solving with the operator. */
int tm_operator_add_springs(tm_operator* op, const double* spring_full);

/* Inertia of K(k) - shift M(k) at rpm (Sylvester): the number of negative pivots of
 * its LDL^H, which is the number of eigenvalues lambda < shift of
 * K(k) phi = lambda M(k) phi over the free DOFs (a completeness check of a solve:
 * with shift just above the largest eigenvalue returned, it equals n_unique).  Pinned
 * DOFs count as positive.  min_pivot_ratio (may be NULL): the smallest |d_j| / |a_jj|;
 * pivot_status (may be NULL): bits as tm_result.pivot_flags.  No counterpart in the
 * reference (a diagnostic of this engine). */
int tm_operator_inertia(tm_operator* op, double rpm, int32_t harmonic, double shift,
                        int64_t* n_negative, double* min_pivot_ratio, int32_t* pivot_status);

/* tm_operator_inertia without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops or harnesses that forbid hipMalloc).  The factor workspace comes from a static
 * 64 MiB device array of the library (one per device, held until the process ends); a mesh
 * whose one-problem workspace exceeds it is rejected with TM_ERR_ARG (the 24-sector
 * fixtures and the 4.9k-DOF flat disk fit, the 45k-DOF bladed sector does not).  Calls that
 * overlap on one device share that array and race: serialise them. */
int tm_operator_inertia_noalloc(tm_operator* op, double rpm, int32_t harmonic, double shift,
                                int64_t* n_negative, double* min_pivot_ratio, int32_t* pivot_status);

#ifdef __cplusplus
}
#endif
#endif /* TMHIP_H */
