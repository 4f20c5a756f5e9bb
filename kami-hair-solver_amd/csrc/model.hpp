// Host construction of the Cosserat rod model and the collider mesh that the
// GPU step consumes: strand discretization, parallel-transported frames,
// rest strains, lumped masses, root targets, collider validation.
#pragma once

#include "host_math.hpp"
#include "kami_hair_solver.h"

#include <cstdint>
#include <string>
#include <vector>

namespace khs {

constexpr uint32_t kNodeFixed = 1u;      // Dirichlet root node
constexpr uint32_t kNodeVirtual = 2u;    // invisible extension node
constexpr uint32_t kNodeHasElement = 4u; // left node of an element
constexpr uint32_t kNodeContact = 8u;    // its element takes collider contact

// Nodes of a strand are contiguous; element k of the model joins nodes
// (k + s, k + s + 1) of strand s, and is stored at its left node.
struct RodModel {
    uint32_t node_count = 0, element_count = 0, strand_count = 0;
    std::vector<Vec3> position;              // rest / initial node positions
    std::vector<Mat3> frame;                 // reference frames
    std::vector<double> mass, rotational_mass;
    std::vector<uint32_t> flags;             // kNode* bits
    std::vector<uint32_t> strand_of;         // node -> strand
    std::vector<uint32_t> bind_left, bind_right;
    std::vector<double> bind_t;
    std::vector<double> rest;                // per node, 8 doubles: L0, shear0[3], curvature0[3], 0
    std::vector<uint32_t> strand_first, strand_count_of, strand_fixed;
    std::vector<uint32_t> original_to_internal;
    uint32_t merged_segments = 0, extension_strands = 0, extension_nodes = 0;
    double extension_length = 0.0, maximum_rest_length = 0.0;
    uint64_t free_dofs = 0;

    uint32_t element_of_node(uint32_t n) const { return n - strand_of[n]; }
};

// Discretizes the strands (reference src/solver.cpp:227-318), builds frames
// and rest strains (:320-356). Returns false with `error` on a degenerate
// strand.
bool build_rod_model(const std::vector<Vec3> &points, const std::vector<uint32_t> &offsets,
                     const KhsSolverDesc &desc, RodModel &model, std::string &error);

// Lumped translational and rotational masses (reference src/solver.cpp:358-373).
void assign_masses(const KhsHairMaterial &material, RodModel &model);

// Root target positions and rotation vectors of every node for one set of
// original point positions (reference src/solver.cpp:416-447). Runs over
// strands on `threads` host threads.
bool root_targets(const RodModel &model, const std::vector<Vec3> &points, std::vector<Vec3> &positions,
                  std::vector<Vec3> &rotations, unsigned threads);

struct ColliderMesh {
    std::vector<Vec3> vertices;
    std::vector<uint32_t> triangles;  // 3 per triangle
    uint32_t excluded = 0, boundary_edges = 0, nonmanifold_edges = 0, inconsistent_edges = 0, inverted = 0;
    bool closed = false;
    uint32_t triangle_count() const { return uint32_t(triangles.size() / 3); }
};

// Drops invalid or degenerate triangles (reference src/solver.cpp:179-225).
KhsResult filter_collider(const KhsVec3 *vertices, uint32_t vertex_count, const KhsTriangle *triangles,
                          uint32_t triangle_count, ColliderMesh &mesh, std::string &error);

// Edge-use diagnostics and closed-mesh orientation (reference
// src/solver.cpp:381-414).
void validate_collider(ColliderMesh &mesh);

}  // namespace khs
