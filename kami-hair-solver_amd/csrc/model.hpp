// What the host keeps of the Cosserat rod model (built on the device: strand
// layout and build statistics), and the host construction of the collider
// mesh: filtering and validation.
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
// (k + s, k + s + 1) of strand s, and is stored at its left node. The
// per-node arrays live on the device only.
struct RodModel {
    uint32_t node_count = 0, element_count = 0, strand_count = 0;
    std::vector<uint32_t> strand_first, strand_count_of, strand_fixed;
    std::vector<uint32_t> strand_real;       // nodes of the strand that are not virtual
    uint32_t merged_segments = 0, extension_strands = 0, extension_nodes = 0, fixed_nodes = 0;
    double extension_length = 0.0, maximum_rest_length = 0.0;
    uint64_t free_dofs = 0;
};

// Lumped masses per unit rest length of an element, split evenly between
// its two nodes (reference src/solver.cpp:358-373): translational
// rho pi r^2, rotational rho (2 I + J) / 3 with I = pi r^4 / 4, J = 2 I.
struct MassDensity {
    double translational = 0.0, rotational = 0.0;
};
MassDensity mass_density(const KhsHairMaterial &material);

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
