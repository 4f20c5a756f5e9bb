// Device kernels of the hair step and their launch helpers.
#pragma once

#include "lbvh.hpp"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

#include <cstdint>

namespace khs {

constexpr int kWave = 64;
constexpr int kTile = 63;   // nodes owned by one wavefront in tiled kernels
constexpr int kBlock = 256;  // four wavefronts
constexpr int kTilesPerBlock = kBlock / kWave;
constexpr int kNodesPerBlock = kTile * kTilesPerBlock;
// The evaluation's tiles own one node fewer: its lane 63 lends its node's
// rotation to lane 62 (step_kernels.hpp, k_evaluate).
constexpr int kEvalTile = 62;
constexpr int kEvalNodesPerBlock = kEvalTile * kTilesPerBlock;

// One block's share of every reduction the host needs after a pass; the
// reduce kernel folds them in a fixed order, so results are reproducible.
struct Partial {
    double sum[8];
    double min_gap;
    double alpha;
    double max_translation;
    unsigned long long candidates, active;
    unsigned int infeasible, failed;
};
enum PartialSum { kObjective = 0, kElastic, kContact, kKinetic, kGradNorm2, kDirNorm2, kSlope, kSpare };

// Read-only rod model and collider views passed by value to kernels.
struct RodView {
    uint32_t node_count, strand_count;
    const double *frame;  // 9 per node, row-major
    const double *rest;   // 8 per node: L0, shear0[3], curvature0[3], 0
    const double *mass, *rotational_mass;
    const uint32_t *flags;
    const uint32_t *strand_of;
    const uint32_t *strand_first, *strand_count_of, *strand_fixed;
};

struct ColliderView {
    uint32_t vertex_count, triangle_count;
    const double *current;  // 3 per vertex
    const double *previous;
    const uint32_t *triangles;
    const BvhNode *bvh;
    const double *leaf_box;  // triangle records at `current` (kTriRecord each), or null: leaves use the tree's boxes
};

// Per-element lists of the collider triangles whose leaf box overlaps a
// cached box around the element. Before each query pass the cached box is
// checked to contain the pass's query box (grown by the pass's pad plus a
// small rounding margin); only elements that left it walk the hierarchy
// again, gathering for their query box grown by a skin. A covered query thus
// finds in the list every triangle the hierarchy would report. Lists longer
// than kCacheCapacity are marked overflowed and their queries walk the tree.
constexpr uint32_t kCacheCapacity = 512;
constexpr uint32_t kCacheOverflow = 0xffffffffu;
constexpr uint32_t kNoTriangle = 0xffffffffu;
constexpr uint32_t kPairBlocks = 2048;  // grid of the per-element narrow-phase passes
struct ContactCache {
    double skin;
    uint32_t *count;        // per element: list length, or kCacheOverflow
    uint32_t *tris;         // kCacheCapacity per element, contiguous
    double *box;            // 6 per element: lo, hi of the covered box
    uint32_t *listed;       // per element: 1 once appended to the list below
    uint32_t *list;         // elements whose list became non-empty
    uint32_t *list_count;
    uint32_t *pending;      // elements to gather (the swept pass)
    uint32_t *pending_count;
    const double *tri_box;  // triangle records (kTriRecord each): box and vertices at the current position
    // The swept check marks every element it queues with the pass's epoch
    // (a count of swept passes, never 0) and records the listed count; the
    // CCD kernel after it bounds the queued elements where it gathers them,
    // and the rest from their lists.
    uint32_t *pending_epoch;    // per element
    uint32_t epoch;
    uint32_t *listed_snapshot;  // list_count when the swept check ran
};

// The Newton direction's choice on the device: the block direction if it
// was solved and descends, else steepest descent; and the factor that
// brings its largest free-node translation within the trust radius.
struct DirectionChoice {
    int used;  // 1: block direction, 0: steepest descent
    int pad;
    double scale;
};

// Scalars of one substep.
struct StepParams {
    double dt, inv_dt, inv_dt2;
    double gravity[3];
    double radius_offset;  // radius + collider offset
    double barrier, stiffness, damping, friction, smoothing, minimum_gap;
    double weights[6];     // kGA, kGA, EA, EI, EI, 2GI
    double contact_pad;    // radius + offset + barrier
    double sweep_target;   // radius + offset + minimum gap
    int has_collider;
    // Per-strand materials: 2 doubles per node, the Young's modulus and
    // radius of the node's strand, or null when the material's values
    // (above) hold for every strand.
    const double *material;
    double collider_offset, poisson, shear_correction;
};

struct SweepRecord {
    uint32_t claimed, strand, element, triangle;
    double distance, required, clearance, displacement;
    double hair_start[3], hair_end[3], collider_point[3];
};

}  // namespace khs
