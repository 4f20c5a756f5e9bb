// Host rod-model build: see model.hpp for what each function restates.
#include "model.hpp"

#include <thread>
#include <unordered_map>

namespace khs {

bool build_rod_model(const std::vector<Vec3> &points, const std::vector<uint32_t> &offsets,
                     const KhsSolverDesc &desc, RodModel &m, std::string &error)
{
    m = RodModel{};
    const uint32_t strands = uint32_t(offsets.size() - 1);
    const double hmax = desc.maximum_element_length, zero = 1.0e-12;
    m.original_to_internal.assign(points.size(), 0);
    m.strand_first.resize(strands);
    m.strand_count_of.resize(strands);
    m.strand_fixed.resize(strands);
    auto push_node = [&](Vec3 x, uint32_t s, uint32_t left, uint32_t right, double t, bool is_virtual) {
        m.position.push_back(x);
        m.strand_of.push_back(s);
        m.bind_left.push_back(left);
        m.bind_right.push_back(right);
        m.bind_t.push_back(t);
        m.flags.push_back(is_virtual ? kNodeVirtual : 0u);
        return uint32_t(m.position.size() - 1);
    };
    for (uint32_t s = 0; s < strands; ++s) {
        const uint32_t begin = offsets[s], end = offsets[s + 1];
        const uint32_t first = push_node(points[begin], s, begin, begin, 0.0, false);
        m.strand_first[s] = first;
        m.original_to_internal[begin] = first;
        Vec3 previous = points[begin], tip_direction;
        uint32_t segments = 0, tip_left = begin, tip_right = begin;
        double strand_length = 0.0, tip_length = 0.0;
        for (uint32_t p = begin + 1; p < end; ++p) {
            const Vec3 target = points[p];
            const double len = length(target - previous);
            if (len <= zero) {  // zero-length segment: share the node
                m.original_to_internal[p] = uint32_t(m.position.size() - 1);
                ++m.merged_segments;
                previous = target;
                continue;
            }
            const uint32_t pieces = std::max(1u, uint32_t(std::ceil(len / hmax)));
            for (uint32_t d = 1; d <= pieces; ++d) {
                const double t = double(d) / pieces;
                push_node((1.0 - t) * previous + t * target, s, p - 1, p, t, false);
            }
            m.original_to_internal[p] = uint32_t(m.position.size() - 1);
            previous = target;
            strand_length += len;
            tip_left = p - 1;
            tip_right = p;
            tip_length = len;
            tip_direction = over(target - points[p - 1], len);
            ++segments;
        }
        uint32_t count = uint32_t(m.position.size()) - first;
        if (segments == 0 || count < 2) {
            error = "Strand " + std::to_string(s) + " has all its points at one position.";
            return false;
        }
        if (desc.minimum_dynamic_length > strand_length + zero) {
            // Invisible straight extension along the last segment up to the
            // minimum dynamic length.
            const double extra = desc.minimum_dynamic_length - strand_length;
            const uint32_t pieces = std::max(1u, uint32_t(std::ceil(extra / hmax)));
            const Vec3 tip = points[end - 1];
            for (uint32_t d = 1; d <= pieces; ++d) {
                const double dist = extra * double(d) / pieces;
                push_node(tip + dist * tip_direction, s, tip_left, tip_right, 1.0 + dist / tip_length, true);
            }
            ++m.extension_strands;
            m.extension_nodes += pieces;
            m.extension_length += extra;
            count += pieces;
        }
        const uint32_t fixed = std::min<uint32_t>(desc.fixed_root_nodes, count);
        for (uint32_t k = 0; k < fixed; ++k) m.flags[first + k] |= kNodeFixed;
        m.strand_count_of[s] = count;
        m.strand_fixed[s] = fixed;
    }
    m.node_count = uint32_t(m.position.size());
    m.strand_count = strands;
    m.element_count = m.node_count - strands;
    m.frame.assign(m.node_count, Mat3{});
    m.rest.assign(size_t(m.node_count) * 8, 0.0);
    m.mass.assign(m.node_count, 0.0);
    m.rotational_mass.assign(m.node_count, 0.0);

    // Frames by parallel transport along central-difference tangents, then
    // per element rest length, shear 1/2 (Ri^T + Rj^T) t and curvature
    // cayley(Ri^T Rj) / L0.
    for (uint32_t s = 0; s < strands; ++s) {
        const uint32_t f = m.strand_first[s], c = m.strand_count_of[s];
        std::vector<Vec3> tangent(c);
        for (uint32_t k = 0; k < c; ++k) {
            const uint32_t lo = k == 0 ? 0 : k - 1, hi = k + 1 == c ? k : k + 1;
            tangent[k] = unit(m.position[f + hi] - m.position[f + lo]);
        }
        Mat3 frame = tangent_frame(tangent[0], Vec3{1, 0, 0});
        m.frame[f] = frame;
        for (uint32_t k = 1; k < c; ++k) {
            frame = tangent_frame(tangent[k], transport(frame.column(0), tangent[k - 1], tangent[k]));
            m.frame[f + k] = frame;
        }
        for (uint32_t k = 0; k + 1 < c; ++k) {
            const uint32_t i = f + k, j = i + 1;
            m.flags[i] |= kNodeHasElement;
            if (!(m.flags[j] & kNodeVirtual)) m.flags[i] |= kNodeContact;
            const double L0 = length(m.position[j] - m.position[i]);
            const Vec3 t = over(m.position[j] - m.position[i], L0);
            const Mat3 rit = transposed(m.frame[i]), rjt = transposed(m.frame[j]);
            Mat3 sum;
            for (int e = 0; e < 9; ++e) sum.m[e] = rit.m[e] + rjt.m[e];
            const Vec3 shear = 0.5 * apply(sum, t);
            const Vec3 curvature = over(cayley_vector(product(rit, m.frame[j])), L0);
            double *r = &m.rest[size_t(i) * 8];
            r[0] = L0;
            r[1] = shear.x; r[2] = shear.y; r[3] = shear.z;
            r[4] = curvature.x; r[5] = curvature.y; r[6] = curvature.z;
            m.maximum_rest_length = std::max(m.maximum_rest_length, L0);
        }
    }
    m.free_dofs = 0;
    for (uint32_t n = 0; n < m.node_count; ++n) if (!(m.flags[n] & kNodeFixed)) m.free_dofs += 6;
    return true;
}

void assign_masses(const KhsHairMaterial &mat, RodModel &m)
{
    const double pi = 3.14159265358979323846;
    const double area = pi * mat.radius * mat.radius;
    const double second_moment = pi * std::pow(mat.radius, 4.0) / 4.0;
    const double polar = 2.0 * second_moment;
    std::fill(m.mass.begin(), m.mass.end(), 0.0);
    std::fill(m.rotational_mass.begin(), m.rotational_mass.end(), 0.0);
    for (uint32_t i = 0; i < m.node_count; ++i) {
        if (!(m.flags[i] & kNodeHasElement)) continue;
        const double L0 = m.rest[size_t(i) * 8];
        const double mass = mat.density * area * L0;
        const double rot = mat.density * (2.0 * second_moment + polar) / 3.0 * L0;
        m.mass[i] += 0.5 * mass;
        m.mass[i + 1] += 0.5 * mass;
        m.rotational_mass[i] += 0.5 * rot;
        m.rotational_mass[i + 1] += 0.5 * rot;
    }
}

bool root_targets(const RodModel &m, const std::vector<Vec3> &points, std::vector<Vec3> &pos,
                  std::vector<Vec3> &rot, unsigned threads)
{
    pos.resize(m.node_count);
    rot.assign(m.node_count, Vec3{});
    for (uint32_t n = 0; n < m.node_count; ++n)
        pos[n] = (1.0 - m.bind_t[n]) * points[m.bind_left[n]] + m.bind_t[n] * points[m.bind_right[n]];
    std::vector<uint8_t> ok(m.strand_count, 1);
    auto work = [&](uint32_t s0, uint32_t s1) {
        std::vector<Vec3> tangent;
        for (uint32_t s = s0; s < s1; ++s) {
            const uint32_t f = m.strand_first[s], c = m.strand_count_of[s];
            tangent.resize(c);
            for (uint32_t k = 0; k < c; ++k) {
                const uint32_t lo = k == 0 ? 0 : k - 1, hi = k + 1 == c ? k : k + 1;
                const Vec3 t = pos[f + hi] - pos[f + lo];
                if (length2(t) < 1.0e-24) { ok[s] = 0; break; }
                tangent[k] = unit(t);
            }
            if (!ok[s]) continue;
            Mat3 frame = tangent_frame(tangent[0], m.frame[f].column(0));
            rot[f] = axis_angle(product(frame, transposed(m.frame[f])));
            for (uint32_t k = 1; k < c; ++k) {
                frame = tangent_frame(tangent[k], transport(frame.column(0), tangent[k - 1], tangent[k]));
                rot[f + k] = axis_angle(product(frame, transposed(m.frame[f + k])));
            }
        }
    };
    const unsigned T = std::max(1u, std::min<unsigned>(threads, m.strand_count / 256 + 1));
    if (T == 1) {
        work(0, m.strand_count);
    } else {
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < T; ++t)
            pool.emplace_back(work, uint32_t(uint64_t(m.strand_count) * t / T), uint32_t(uint64_t(m.strand_count) * (t + 1) / T));
        for (auto &th : pool) th.join();
    }
    for (uint8_t v : ok) if (!v) return false;
    return true;
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
