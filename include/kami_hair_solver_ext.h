/*
 * kami_hair_solver_ext.h - additive entry points of the MI355X hair solver.
 *
 * None of these has a counterpart in the reference ABI v5; KHS_ABI_VERSION
 * stays 5 so existing callers load the library unchanged. They serve three
 * kinds of caller:
 *  - parity tests, which read and write the solver state and evaluate the
 *    incremental potential at a given state (khsx* "state" and "evaluate");
 *  - benchmarks, which select the convergence rule and arithmetic mode;
 *  - the bake server and batched runs (khsxJob*, in kami_hair_server.h).
 * The CPU oracle library exports the same khsx functions, so one test
 * drives both.
 */
#ifndef KAMI_HAIR_SOLVER_EXT_H
#define KAMI_HAIR_SOLVER_EXT_H

#include "kami_hair_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Number of doubles in q (and in v): 6 per internal node. */
KHS_API uint64_t khsxGetDofCount(const KhsSolver *solver);

/* Copies the generalized coordinates q and velocities v (6 doubles per node:
   x, y, z, rotation vector). Either pointer may be NULL. */
KHS_API KhsResult khsxGetState(const KhsSolver *solver, double *q, double *v,
                               uint64_t count);

/* Sets q, the previous-substep coordinates q_old and v. q_old may be NULL,
   in which case q_old = q. */
KHS_API KhsResult khsxSetState(KhsSolver *solver, const double *q,
                               const double *q_old, const double *v,
                               uint64_t count);

/* Rest model built by khsBuild. Per node (count = internal node count):
   reference frame, row-major 9 doubles; mass; rotational mass; flags
   (bit 0 fixed, bit 1 virtual extension). Per element (count = element
   count): node pair (2 uint32), strand, contact flag; rest length, rest
   shear (3), rest curvature (3). Any pointer may be NULL. */
KHS_API KhsResult khsxCopyNodeModel(const KhsSolver *solver, double *frames9,
                                    double *mass, double *rotational_mass,
                                    uint32_t *flags, uint32_t capacity);
KHS_API KhsResult khsxCopyElementModel(const KhsSolver *solver, uint32_t *nodes2,
                                       uint32_t *strand, uint32_t *contact,
                                       double *rest7, uint32_t capacity);

typedef struct KhsxEvaluation {
    uint32_t struct_size;
    uint32_t feasible;
    uint32_t solve_failed;
    uint32_t reserved;
    uint64_t candidate_pairs;
    uint64_t active_contacts;
    double objective;
    double elastic;
    double contact;
    double kinetic;
    double minimum_gap;
} KhsxEvaluation;

/* Evaluates the incremental potential of one substep of length dt at the
   current state (q, q_old, v, current collider): energies, the gradient
   (6 per node, zero on fixed nodes) and, if direction is not NULL, the
   block-tridiagonal Gauss-Newton direction -H^-1 g before trust-radius
   scaling. gradient may be NULL. */
KHS_API KhsResult khsxEvaluate(KhsSolver *solver, double dt, double *gradient,
                               double *direction, KhsxEvaluation *out);

/* Gauss-Newton blocks of the last khsxEvaluate: per node the 6x6 diagonal
   block and the 6x6 block coupling it to the next node of its strand
   (row-major, 36 doubles each; zero for fixed nodes). */
KHS_API KhsResult khsxCopyBlocks(const KhsSolver *solver, double *diagonal, double *upper,
                                 uint32_t capacity);

/* Broad phase at the current state: every (element, triangle) pair whose
   axis-aligned boxes overlap with padding radius + offset + barrier
   distance, for contact-enabled elements with a free left node. Pairs are
   written as (element, triangle) uint32 pairs, sorted; *count receives the
   total even when capacity is too small. */
KHS_API KhsResult khsxBroadphasePairs(KhsSolver *solver, uint32_t *pairs,
                                      uint64_t capacity, uint64_t *count);

/* khsxBroadphasePairs without allocating device memory. No counterpart
   in the reference: it serves callers that cannot allocate (HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc). The GPU
   library stages the pairs in a fixed device array of at most
   KHSX_NO_ALLOC_PAIRS pairs (8 MiB, held by each GPU until the process
   ends) and sorts them in the caller's buffer; capacity above that limit is
   refused with KHS_ERROR_INVALID_ARGUMENT. Calls that overlap on one GPU
   share the staging array and race; each call clears what it uses. */
#define KHSX_NO_ALLOC_PAIRS 1048576u
KHS_API KhsResult khsxBroadphasePairsNoAlloc(KhsSolver *solver, uint32_t *pairs,
                                             uint64_t capacity, uint64_t *count);

/* Self-check of the device broad-phase structure; violations[6]:
   0 triangle boxes not enclosed by an ancestor's stored box, 1 broken parent
   chains, 2 out-of-order sort keys, 3 leaves whose stored box is not their
   triangle's box, 4 stored boxes with lo > hi, 5 depth of the deepest leaf.
   Entries 0-4 are zero for a sound structure; the oracle reports zeros. */
KHS_API KhsResult khsxCheckBroadphase(KhsSolver *solver, uint64_t *violations);

/* Convergence rule. 0 (default): the reference rule, with the relative and
   absolute tolerances floored at 1e-4 and 1e-5 sqrt(dof). 1: floors lifted,
   so the desc tolerances apply as given (used to run to a 1e-6 residual). */
KHS_API KhsResult khsxSetConvergenceMode(KhsSolver *solver, uint32_t mode);

/* Arithmetic mode of the step. 0 (default): FP64 throughout. 1: mixed
   precision: the Gauss-Newton blocks and the gradient are rounded to FP32
   as the block solve loads them, and every Schur complement, inverse and
   Thomas factor is formed and stored in FP32; the direction is widened to
   FP64. Assembly, energies, barrier, step bound, line search and the
   convergence test stay FP64, so each Newton iteration corrects the FP32
   direction from an FP64 residual. The per-strand systems have condition
   numbers near 3e12: the FP32 direction leaves relative linear residuals
   near 1e-3 and the Newton loop converges far more slowly (measured in
   tests/test_mixed_precision.py). */
KHS_API KhsResult khsxSetPrecisionMode(KhsSolver *solver, uint32_t mode);

/* Device-side timing of the step, from HIP events recorded on the
   library's stream around each launch of the named kernel classes while
   timing is enabled (off by default). Substep latencies of the last frame
   (event to event) are kept for the last frame stepped. */
typedef struct KhsxKernelTiming {
    uint32_t struct_size;
    uint32_t substep_count;          /* substeps of the last frame */
    uint64_t assembly_launches;      /* gradient + Gauss-Newton block assembly */
    double assembly_ms;
    uint64_t energy_launches;        /* line-search energy evaluations */
    double energy_ms;
    uint64_t solve_launches;         /* block-tridiagonal solves */
    double solve_ms;
    uint64_t collision_launches;     /* BVH rebuilds, sweeps and CCD */
    double collision_ms;
    uint64_t nodes;                  /* internal nodes of the model */
} KhsxKernelTiming;

KHS_API KhsResult khsxSetKernelTiming(KhsSolver *solver, uint32_t enable);
KHS_API KhsResult khsxGetKernelTiming(const KhsSolver *solver, KhsxKernelTiming *timing);
/* Milliseconds of each substep of the last frame (at most capacity). */
KHS_API KhsResult khsxGetSubstepTimes(const KhsSolver *solver, double *milliseconds,
                                      uint32_t capacity, uint32_t *count);

/* Per-element contact lists of the device step (collider triangles whose
   box overlaps a cached box around each hair element), as left by the last
   substep: how many elements hold one, their lengths and overflows. A
   diagnostic that waits for the device; the CPU oracle keeps no lists and
   reports zeros. */
typedef struct KhsxContactListStats {
    uint32_t struct_size;
    uint32_t capacity;                /* entries an element's list holds */
    uint64_t elements;                /* internal nodes (one element each) */
    uint64_t listed;                  /* elements queued for the list passes */
    uint64_t nonempty;                /* elements whose list is not empty */
    uint64_t overflowed;              /* lists over capacity (queries walk the tree) */
    uint64_t entries;                 /* triangles over all lists within capacity */
    uint64_t longest;                 /* longest list within capacity */
    uint64_t length_histogram[9];     /* lengths 1, 2-3, 4-7, ..., 128-255, 256 and up */
    uint64_t root_nonempty;           /* of nonempty: elements whose left node is fixed */
    uint64_t root_entries;            /* of entries: theirs */
} KhsxContactListStats;

KHS_API KhsResult khsxGetContactListStats(KhsSolver *solver, KhsxContactListStats *stats);

/* Rollback ring of animation states kept on the device: with `frames` > 0
   the state (q and v) after each stepped animation frame is kept in slot
   frame % frames (0 frees the ring). khsxRollback(frame) restores the state
   kept after `frame`, as khsRestoreAnimationCheckpoint restores a checkpoint
   written after that frame, without any host copy; the next frame to step
   is frame + 1. It fails with KHS_ERROR_INVALID_ARGUMENT when the slot no
   longer holds that frame. A rebuild drops the ring; a new animation
   empties it. It replaces the host ring of checkpoints the Blender
   add-on keeps for resuming a bake (blender_extension/kami_hair_solver/
   addon.py:100-103, 642-701), whose every save and restore crosses the host.
   The CPU oracle keeps its ring in host memory. */
KHS_API KhsResult khsxSetRollbackRing(KhsSolver *solver, uint32_t frames);
KHS_API KhsResult khsxRollback(KhsSolver *solver, uint32_t frame);

/* Appends the current positions of the original points to an open .khc hair
   cache: 3 * khsGetOriginalPointCount little-endian doubles, the frame
   record the Blender add-on writes after khsCopyOriginalPositions
   (blender_extension/kami_hair_solver/addon.py:580-584, header :23-24,
   :704-724). The gather runs on the device and the bytes go to `fd` from
   pinned staging, with no per-point host conversion. A failed write returns
   KHS_ERROR_INTERNAL with the system's message in khsGetLastError. */
KHS_API KhsResult khsxWriteCacheFrame(KhsSolver *solver, int fd);

/* Worker threads the CPU oracle uses over strands (oracle only; the GPU
   library ignores it and returns KHS_OK). */
KHS_API KhsResult khsxSetThreadCount(KhsSolver *solver, uint32_t threads);

/* Per-strand Young's modulus (Pa) and hair radius (m), one value each per
   hair strand (strand_count = the strands of khsSetHairCurves). They replace
   the material's young_modulus and radius for each strand everywhere the
   step uses them: masses, stretch/shear/bend/twist stiffness, the Jacobi
   preconditioner, and the contact radius (barrier gap, query padding,
   step bound and sweep targets); the trust radius uses the largest radius.
   Every other material field stays solver-wide. Both pointers NULL returns
   to the material's values. On a built solver it applies at once (masses
   are recomputed); new hair curves clear it. The reference's material is
   global per solver (include/kami_hair_solver.h:80-93, 214-219); this is
   the additive input of the uneven-material groom. */
KHS_API KhsResult khsxSetStrandMaterials(KhsSolver *solver, const double *young_modulus, const double *radius,
                                         uint32_t strand_count);

/* Stage hooks: each runs one stage of the substep on the current state, so
   that a test can compare the stage between libraries and against closed
   forms. None of them changes q.

   khsxSetColliderState sets the collider's current and previous positions
   (3 doubles per vertex; previous NULL means previous = current), as the
   collider interpolation leaves them inside a substep
   (reference src/cuda_backend.cu:770-774). */
KHS_API KhsResult khsxSetColliderState(KhsSolver *solver, const double *current, const double *previous,
                                       uint32_t vertex_count);

/* The post-substep Coulomb friction impulse at the current q, v and
   collider (current and previous), added to v (reference
   src/cuda_backend.cu:941-964, 615-621). */
KHS_API KhsResult khsxApplyFriction(KhsSolver *solver, double dt);

/* Largest step alpha in [0, 1] along `direction` (6 doubles per node) that
   conservative advancement of every contact-enabled free segment against
   the current collider allows, before the line search's 0.9 factor
   (reference src/cuda_backend.cu:989-1000). */
KHS_API KhsResult khsxStepBound(KhsSolver *solver, const double *direction, uint64_t count, double *alpha);

/* The moving-collider sweep at substep start: every candidate triangle
   advanced from its previous to its current position against the
   stationary segments (reference src/cuda_backend.cu:966-987).
   *feasible = 1 when no triangle reaches the target clearance; otherwise 0
   and `diagnostics` (may be NULL) receives the reported pair as
   khsGetFailureDiagnostics would. */
KHS_API KhsResult khsxSweepCheck(KhsSolver *solver, uint32_t *feasible, KhsFailureDiagnostics *diagnostics);

/* Newton direction rule. 0 (default): the block direction when it solves
   and descends, else Jacobi-preconditioned steepest descent
   (reference src/cuda_backend.cu:623-648). 1: steepest descent always, the
   fallback forced; khsxEvaluate then returns -P g, P the Jacobi
   preconditioner (reference src/cuda_backend.cu:815-846), with zero rows
   on fixed nodes and no trust-radius scaling. */
KHS_API KhsResult khsxSetDirectionMode(KhsSolver *solver, uint32_t mode);

/* Block-tridiagonal solve algorithm. 0: block Thomas from the root to the
   tip with the explicit Gauss-Jordan inverse and partial pivoting, the
   reference's (src/cuda_backend.cu:855-881). 1: twisted factorization,
   eliminating from the root and from the tip towards the middle node at
   once. 2: fused: block Thomas whose factors C' = S^-1 U and d' come out of
   one Gauss-Jordan elimination of [S | U | rhs] without row exchanges (the
   blocks are symmetric positive definite), with the reference's diagonal
   shift; no inverse is formed. All three give the same direction in exact
   arithmetic, rounded differently. Both libraries implement each with the
   same operations (the device's reciprocal in 2 is within an ulp of a
   division), so parity holds in any. Algorithm 2 on a solver whose D, U
   and gradient arrays reach 4 GiB together (about 6.8M nodes) runs 0.
   Precision mode 1 always runs the FP32 form of 0. */
KHS_API KhsResult khsxSetSolveAlgorithm(KhsSolver *solver, uint32_t algorithm);

/* Reuse of contact-list work inside the Newton loop. 1 (default): the
   gradient evaluation at an accepted line-search trial takes the closest
   triangles the trial's list pass found at the same positions, and a trial
   whose query boxes the step bound's swept pass already made every list
   cover runs no contact check of its own. 0: every evaluation runs its own
   check and list pass (the result is the same; this mode exists so that
   tests can show it). The CPU oracle keeps no lists and accepts both. */
KHS_API KhsResult khsxSetContactReuse(KhsSolver *solver, uint32_t mode);

/* Directions the Newton loop used in the last stepped frame (its final
   adaptive attempt): block-tridiagonal directions and steepest-descent
   directions (forced or fallback). */
KHS_API KhsResult khsxGetDirectionCounts(const KhsSolver *solver, uint64_t *block_directions,
                                         uint64_t *steepest_directions);

/* Work enqueued on the solver's stream since it was built: kernel launches
   and device-to-device copies (each a GPU dispatch), and host waits for the
   stream (read-backs of reductions and synchronizations). Callers take
   differences around the frames they measure. The CPU oracle reports zero
   for both. */
KHS_API KhsResult khsxGetLaunchCounts(const KhsSolver *solver, uint64_t *dispatches, uint64_t *host_waits);

#ifdef __cplusplus
}
#endif

#endif
