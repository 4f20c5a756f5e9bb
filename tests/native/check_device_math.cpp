// CPU check of the HIP library's analytic element Jacobians and geometry
// (kami-hair-solver_amd/csrc/device_math.hpp, compiled for the host) against
// the oracle's dual-number differentiation and closest-point routines.
// Prints one JSON object with the largest relative errors found.
#include "../../oracle/oracle_math.hpp"
#include "../../oracle/oracle_geometry.hpp"
#include "../../kami-hair-solver_amd/csrc/device_math.hpp"

#include <cstdio>
#include <random>

using namespace oracle;

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t s) : g(s) {}
    double u(double a, double b) { return std::uniform_real_distribution<double>(a, b)(g); }
};

static M3 random_rotation(Rng &r)
{
    TV3<Dual<1>> w{Dual<1>(r.u(-3, 3)), Dual<1>(r.u(-3, 3)), Dual<1>(r.u(-3, 3))};
    TM3<Dual<1>> e = exp_so3(w);
    M3 out;
    for (int k = 0; k < 9; ++k) out.m[k] = e.m[k].v;
    return out;
}

int main()
{
    Rng rng(12345);
    double worst_strain = 0, worst_jac = 0;
    for (int trial = 0; trial < 4000; ++trial) {
        // Element with random frames, rest state and a perturbed state;
        // rotation magnitudes span the Taylor branch and large angles.
        const double scale = trial % 4 == 0 ? 1e-7 : (trial % 4 == 1 ? 1e-3 : (trial % 4 == 2 ? 0.1 : 1.5));
        M3 fi = random_rotation(rng), fj = random_rotation(rng);
        double q[12];
        const double L0 = rng.u(1e-3, 1e-2);
        for (int k = 0; k < 3; ++k) { q[k] = rng.u(-0.1, 0.1); q[6 + k] = q[k] + rng.u(-L0, L0); }
        for (int k = 3; k < 6; ++k) { q[k] = rng.u(-scale, scale); q[6 + k] = rng.u(-scale, scale); }
        double rest[8] = {L0, rng.u(-1, 1), rng.u(-1, 1), rng.u(-1, 1), rng.u(-50, 50), rng.u(-50, 50), rng.u(-50, 50), 0};
        // oracle (dual numbers)
        struct E { double rest_length; V3 rest_shear, rest_curvature; } dummy;
        (void)dummy;
        Dual<12> dq[12];
        for (int k = 0; k < 12; ++k) dq[k] = Dual<12>::var(q[k], k);
        typedef Dual<12> D;
        TV3<D> xi{dq[0], dq[1], dq[2]}, ui{dq[3], dq[4], dq[5]}, xj{dq[6], dq[7], dq[8]}, uj{dq[9], dq[10], dq[11]};
        TM3<D> ri0, rj0;
        for (int k = 0; k < 9; ++k) { ri0.m[k] = D(fi.m[k]); rj0.m[k] = D(fj.m[k]); }
        const TM3<D> ri = tmmul(exp_so3(ui), ri0), rj = tmmul(exp_so3(uj), rj0);
        const TV3<D> tangent = tscale(tsub(xj, xi), D(1.0 / L0));
        const TV3<D> shear = tscale(tmvec(tmadd(ttranspose(ri), ttranspose(rj)), tangent), D(0.5));
        const TV3<D> curv = tscale(cayley(tmmul(ttranspose(ri), rj)), D(1.0 / L0));
        D strain[6] = {shear.x - D(rest[1]), shear.y - D(rest[2]), shear.z - D(rest[3]),
                       curv.x - D(rest[4]), curv.y - D(rest[5]), curv.z - D(rest[6])};
        // device routine
        khs::dev::ElementJacobian J;
        khs::dev::element_strains<true>(q, q + 6, fi.m, fj.m, rest, J);
        for (int s = 0; s < 6; ++s) {
            const double e = std::fabs(J.strain[s] - strain[s].v) / std::fmax(1.0, std::fabs(strain[s].v));
            worst_strain = std::fmax(worst_strain, e);
        }
        // Jacobian rows: [dxi (= -Sx for shear), dui, dxj, duj]
        double jac[6][12] = {};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                jac[r][c] = -J.Sx[3 * r + c];
                jac[r][6 + c] = J.Sx[3 * r + c];
                jac[r][3 + c] = J.Si[3 * r + c];
                jac[r][9 + c] = J.Sj[3 * r + c];
                jac[3 + r][3 + c] = J.Ki[3 * r + c];
                jac[3 + r][9 + c] = J.Kj[3 * r + c];
            }
        for (int s = 0; s < 6; ++s) {
            double rowmax = 0;
            for (int a = 0; a < 12; ++a) rowmax = std::fmax(rowmax, std::fabs(strain[s].d[a]));
            for (int a = 0; a < 12; ++a) {
                const double e = std::fabs(jac[s][a] - strain[s].d[a]) / std::fmax(rowmax, 1e-300);
                worst_jac = std::fmax(worst_jac, e);
            }
        }
    }
    // Closest-point routines: device vs oracle on random configurations.
    double worst_geo = 0;
    for (int trial = 0; trial < 20000; ++trial) {
        V3 p0{rng.u(-1, 1), rng.u(-1, 1), rng.u(-1, 1)}, p1{rng.u(-1, 1), rng.u(-1, 1), rng.u(-1, 1)};
        V3 a{rng.u(-1, 1), rng.u(-1, 1), rng.u(-1, 1)}, b{rng.u(-1, 1), rng.u(-1, 1), rng.u(-1, 1)}, c{rng.u(-1, 1), rng.u(-1, 1), rng.u(-1, 1)};
        const bool hit = trial % 2;
        Closest o = closest_segment_triangle(p0, p1, a, b, c, hit);
        khs::dev::Closest d = khs::dev::closest_segment_triangle({p0.x, p0.y, p0.z}, {p1.x, p1.y, p1.z}, {a.x, a.y, a.z},
                                                                  {b.x, b.y, b.z}, {c.x, c.y, c.z}, hit);
        const double e = std::fabs(o.distance - d.distance) + std::fabs(o.t - d.t) + std::fabs(o.bary.x - d.bary.x) +
                         std::fabs(o.bary.y - d.bary.y) + std::fabs(o.bary.z - d.bary.z);
        worst_geo = std::fmax(worst_geo, e);
    }
    std::printf("{\"strain_rel\": %.3e, \"jacobian_rel\": %.3e, \"geometry_abs\": %.3e}\n", worst_strain, worst_jac, worst_geo);
    return 0;
}
