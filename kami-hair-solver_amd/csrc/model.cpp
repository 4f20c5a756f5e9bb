// Mass densities of the rod model and the host collider build.
#include "model.hpp"

#include <unordered_map>

namespace khs {

MassDensity mass_density(const KhsHairMaterial &mat)
{
    const double pi = 3.14159265358979323846;
    const double area = pi * mat.radius * mat.radius;
    const double second_moment = pi * std::pow(mat.radius, 4.0) / 4.0;
    const double polar = 2.0 * second_moment;
    MassDensity r;
    r.translational = mat.density * area;
    r.rotational = mat.density * (2.0 * second_moment + polar) / 3.0;
    return r;
}

KhsResult filter_collider(const KhsVec3 *vertices, uint32_t vertex_count, const KhsTriangle *triangles,
                          uint32_t triangle_count, ColliderMesh &mesh, std::string &error)
{
    mesh = ColliderMesh{};
    mesh.vertices.resize(vertex_count);
    for (uint32_t i = 0; i < vertex_count; ++i) mesh.vertices[i] = {vertices[i].x, vertices[i].y, vertices[i].z};
    mesh.triangles.reserve(size_t(triangle_count) * 3);
    for (uint32_t k = 0; k < triangle_count; ++k) {
        const KhsTriangle t = triangles[k];
        const bool bad_index = t.i0 >= vertex_count || t.i1 >= vertex_count || t.i2 >= vertex_count ||
                               t.i0 == t.i1 || t.i1 == t.i2 || t.i2 == t.i0;
        if (bad_index || !is_finite(mesh.vertices[t.i0]) || !is_finite(mesh.vertices[t.i1]) ||
            !is_finite(mesh.vertices[t.i2])) {
            ++mesh.excluded;
            continue;
        }
        const Vec3 n = cross(mesh.vertices[t.i1] - mesh.vertices[t.i0], mesh.vertices[t.i2] - mesh.vertices[t.i0]);
        if (length2(n) <= 1.0e-24) {
            ++mesh.excluded;
            continue;
        }
        mesh.triangles.insert(mesh.triangles.end(), {t.i0, t.i1, t.i2});
    }
    if (mesh.triangles.empty()) {
        error = "The collider has no valid triangle.";
        return KHS_ERROR_INVALID_COLLIDER;
    }
    return KHS_OK;
}

void validate_collider(ColliderMesh &mesh)
{
    struct Use { uint32_t count = 0; int direction = 0; };
    std::unordered_map<uint64_t, Use> uses;
    uses.reserve(mesh.triangles.size());
    for (size_t k = 0; k < mesh.triangles.size(); k += 3) {
        for (int e = 0; e < 3; ++e) {
            const uint32_t a = mesh.triangles[k + e], b = mesh.triangles[k + (e + 1) % 3];
            const uint64_t key = a < b ? (uint64_t(a) << 32) | b : (uint64_t(b) << 32) | a;
            Use &u = uses[key];
            ++u.count;
            u.direction += a < b ? 1 : -1;
        }
    }
    mesh.boundary_edges = mesh.nonmanifold_edges = mesh.inconsistent_edges = mesh.inverted = 0;
    for (const auto &kv : uses) {
        if (kv.second.count == 1) ++mesh.boundary_edges;
        else if (kv.second.count > 2) ++mesh.nonmanifold_edges;
        else if (kv.second.direction != 0) ++mesh.inconsistent_edges;
    }
    mesh.closed = mesh.boundary_edges == 0 && mesh.nonmanifold_edges == 0 && mesh.inconsistent_edges == 0;
    if (mesh.closed) {
        double six_volume = 0.0;
        for (size_t k = 0; k < mesh.triangles.size(); k += 3)
            six_volume += dot(mesh.vertices[mesh.triangles[k]],
                              cross(mesh.vertices[mesh.triangles[k + 1]], mesh.vertices[mesh.triangles[k + 2]]));
        if (six_volume < 0.0) mesh.inverted = 1;
    }
}

}  // namespace khs
