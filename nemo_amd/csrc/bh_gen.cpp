// Synthetic initial conditions with a counter-based generator: every draw is
// a hash of (seed, particle, draw number), so results do not depend on the
// number of threads.
//
// Plummer sphere: mkplummer's recipe (src/nbody/init/mkplummer.c:314-419):
// radius from the inverted cumulative mass drawn in [0, mfrac] with
// mfrac = min(0.999, m(rfrac)), rfrac = 22.8042468 virial units; isotropic
// direction; speed q*sqrt(2)*(1+r^2)^(-1/4) with q drawn by rejection under
// g(q) = q^2 (1-q^2)^3.5 < 0.1; structural units scaled to virial units by
// 16/(3 pi); equal masses normalised to 1; centre of mass and mean velocity
// subtracted.  The reference draws with Numerical Recipes' ran3; this file
// uses its own generator, so the samples differ, the distribution does not.
#include "../../include/bh.h"

#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct Stream {
    uint64_t key;
    uint64_t k = 0;
    Stream(uint64_t seed, uint64_t i) : key(mix64(seed * 0x9e3779b97f4a7c15ULL + 0x632be59bd9b4e019ULL) ^ mix64(i + 1)) {}
    double u01()   // [0, 1) with 53 random bits
    {
        uint64_t r = mix64(key + 0x9e3779b97f4a7c15ULL * (++k));
        return (double)(r >> 11) * (1.0 / 9007199254740992.0);
    }
    double range(double a, double b) { return a + (b - a) * u01(); }
};

template <class F>
void parallel_for(int64_t n, F &&f)
{
    unsigned nt = std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > 16) nt = 16;
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int64_t i = t; i < n; i += nt) f(i);
        });
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" void bh_mkplummer(int64_t n, uint64_t seed, double *mass, double *pos, double *vel)
{
    if (n < 1 || !mass || !pos) return;
    const double PI = 3.14159265358979323846;
    const double scalefactor = 16.0 / (3.0 * PI);
    const double inv_scale = 1.0 / scalefactor;
    const double sqrt_scale = std::sqrt(scalefactor);
    double mfrac = 0.999;
    const double rfrac = 22.8042468 * scalefactor;
    const double mrfrac = rfrac * rfrac * rfrac / std::pow(1.0 + rfrac * rfrac, 1.5);
    if (mrfrac < mfrac) mfrac = mrfrac;
    parallel_for(n, [&](int64_t i) {
        Stream s(seed, (uint64_t)i);
        double radius = 1.0 / std::sqrt(std::pow(s.range(0.0, mfrac), -2.0 / 3.0) - 1.0);
        double theta = std::acos(s.range(-1.0, 1.0));
        double phi = s.range(0.0, 2.0 * PI);
        pos[3 * i] = radius * std::sin(theta) * std::cos(phi) * inv_scale;
        pos[3 * i + 1] = radius * std::sin(theta) * std::sin(phi) * inv_scale;
        pos[3 * i + 2] = radius * std::cos(theta) * inv_scale;
        double x = 0.0, y = 0.1;
        while (y > x * x * std::pow(1.0 - x * x, 3.5)) {
            x = s.range(0.0, 1.0);
            y = s.range(0.0, 0.1);
        }
        double v = x * std::sqrt(2.0) * std::pow(1.0 + radius * radius, -0.25);
        theta = std::acos(s.range(-1.0, 1.0));
        phi = s.range(0.0, 2.0 * PI);
        if (vel) {
            vel[3 * i] = v * std::sin(theta) * std::cos(phi) * sqrt_scale;
            vel[3 * i + 1] = v * std::sin(theta) * std::sin(phi) * sqrt_scale;
            vel[3 * i + 2] = v * std::cos(theta) * sqrt_scale;
        }
    });
    double mtot = 0.0;
    for (int64_t i = 0; i < n; i++) mtot += 1.0 / (double)n;
    double cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; i++) {
        mass[i] = (1.0 / (double)n) / mtot;
        for (int k = 0; k < 3; k++) {
            cp[k] += pos[3 * i + k];
            if (vel) cv[k] += vel[3 * i + k];
        }
    }
    for (int k = 0; k < 3; k++) {
        cp[k] /= (double)n;
        cv[k] /= (double)n;
    }
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            pos[3 * i + k] -= cp[k];
            if (vel) vel[3 * i + k] -= cv[k];
        }
}

// Exponential disk in a Hernquist halo (G = 1), a synthetic galaxy for the
// time-integration workload; an approximate equilibrium built from the
// standard recipes, not a reference routine:
//  * halo: Hernquist (1990) profile of mass mhalo and scale radius ahalo,
//    radius from the inverted cumulative mass M(r)/M = r^2/(r+a)^2 drawn in
//    [0, 0.95] (truncated at 95 % of the mass, r < 38.5 a); isotropic
//    Gaussian velocities with the isotropic Jeans dispersion
//    sigma^2(r) = (1/rho) int_r^inf rho M(r')/r'^2 dr' in the potential of
//    halo + disk (disk mass taken as spherical), tabulated on a log grid,
//    redrawn above the escape speed;
//  * disk: surface density exp(-R/rdisk) (radius from the inverted
//    cumulative mass 1 - (1+x) e^-x, x <= 9.23), vertical sech^2(z/zdisk);
//    dispersions from Toomre's Q (sigma_R = 3.36 Q Sigma / kappa, capped at
//    the circular speed), the epicyclic ratio (sigma_phi^2 = sigma_R^2
//    kappa^2 / 4 Omega^2) and the isothermal sheet (sigma_z^2 = pi Sigma
//    zdisk); mean rotation from the asymmetric-drift relation; the circular
//    speed and kappa use the enclosed mass of halo + disk taken as spherical;
//  * equal masses within each component; centre of mass and mean velocity
//    subtracted.  Bodies 0..ndisk-1 are the disk, then the halo.
extern "C" void bh_mkdiskhalo(int64_t ndisk, int64_t nhalo, uint64_t seed, double mdisk, double rdisk, double zdisk,
                              double q, double mhalo, double ahalo, double *mass, double *pos, double *vel)
{
    if (ndisk < 0 || nhalo < 0 || ndisk + nhalo < 1 || !mass || !pos || !vel) return;
    const double PI = 3.14159265358979323846;
    auto gauss = [PI](Stream &s) {
        double u1 = s.u01(), u2 = s.u01();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
    };
    // enclosed mass (spherical) and its radial derivative
    auto menc = [&](double r, double *dm) {
        const double h = mhalo * r * r / ((r + ahalo) * (r + ahalo));
        const double x = r / rdisk, ex = std::exp(-x);
        const double d = mdisk * (1.0 - (1.0 + x) * ex);
        *dm = mhalo * 2.0 * ahalo * r / ((r + ahalo) * (r + ahalo) * (r + ahalo)) + mdisk * (r / (rdisk * rdisk)) * ex;
        return h + d;
    };
    // isotropic Jeans dispersion of the halo in the total (spherical) mass:
    // cumulative integral from outside in, trapezoid rule in ln r
    const int NT = 4096;
    const double lr0 = std::log(1e-5 * ahalo), lr1 = std::log(1e5 * ahalo), dl = (lr1 - lr0) / (NT - 1);
    std::vector<double> s2tab(NT);
    {
        auto rho = [&](double r) { return 1.0 / (r * std::pow(r + ahalo, 3)); };
        std::vector<double> f(NT);
        for (int k = 0; k < NT; k++) {
            const double r = std::exp(lr0 + dl * k);
            double dm;
            f[k] = rho(r) * menc(r, &dm) / (r * r) * r;      // d/d ln r
        }
        double acc = 0.0;
        for (int k = NT - 1; k >= 0; k--) {
            if (k < NT - 1) acc += 0.5 * (f[k] + f[k + 1]) * dl;
            const double r = std::exp(lr0 + dl * k);
            s2tab[k] = acc / rho(r);
        }
    }
    auto sigma2 = [&](double r) {
        double t = (std::log(std::max(r, 1e-5 * ahalo)) - lr0) / dl;
        t = std::min(std::max(t, 0.0), (double)(NT - 1) - 1e-9);
        const int k = (int)t;
        return s2tab[k] + (t - k) * (s2tab[k + 1] - s2tab[k]);
    };
    parallel_for(ndisk + nhalo, [&](int64_t i) {
        Stream s(seed, (uint64_t)i);
        double *x = pos + 3 * i, *v = vel + 3 * i;
        if (i < ndisk) {
            mass[i] = mdisk / (double)ndisk;
            const double u = s.range(0.0, 0.999);
            double t = 1.0;                                // solve 1 - (1+t) e^-t = u
            for (int it = 0; it < 60; it++) {
                const double f = 1.0 - (1.0 + t) * std::exp(-t) - u, fp = t * std::exp(-t);
                double tn = t - f / std::max(fp, 1e-300);
                if (tn <= 0.0) tn = 0.5 * t;
                if (std::fabs(tn - t) < 1e-14 * t) {
                    t = tn;
                    break;
                }
                t = tn;
            }
            const double R = rdisk * t, ph = s.range(0.0, 2.0 * PI);
            const double uz = s.range(1e-12, 1.0 - 1e-12);
            const double z = zdisk * std::atanh(2.0 * uz - 1.0);
            x[0] = R * std::cos(ph);
            x[1] = R * std::sin(ph);
            x[2] = z;
            const double Rs = std::max(R, 1e-6 * rdisk);
            double dm;
            const double M = menc(Rs, &dm);
            const double vc2 = M / Rs, om2 = M / (Rs * Rs * Rs);
            const double kap2 = M / (Rs * Rs * Rs) + dm / (Rs * Rs);
            const double sig = mdisk / (2.0 * PI * rdisk * rdisk) * std::exp(-R / rdisk);
            double sR = 3.36 * q * sig / std::sqrt(kap2);
            sR = std::min(sR, std::sqrt(vc2));
            const double sP = sR * std::sqrt(kap2 / (4.0 * om2));
            const double sZ = std::sqrt(PI * sig * zdisk);
            const double vm2 = vc2 + sR * sR * (1.0 - kap2 / (4.0 * om2) - 2.0 * R / rdisk);
            const double vR = sR * gauss(s), vP = std::sqrt(std::max(vm2, 0.0)) + sP * gauss(s), vZ = sZ * gauss(s);
            v[0] = vR * std::cos(ph) - vP * std::sin(ph);
            v[1] = vR * std::sin(ph) + vP * std::cos(ph);
            v[2] = vZ;
        } else {
            mass[i] = mhalo / (double)nhalo;
            const double su = std::sqrt(s.range(0.0, 0.95));
            const double r = ahalo * su / (1.0 - su);
            const double th = std::acos(s.range(-1.0, 1.0)), ph = s.range(0.0, 2.0 * PI);
            x[0] = r * std::sin(th) * std::cos(ph);
            x[1] = r * std::sin(th) * std::sin(ph);
            x[2] = r * std::cos(th);
            const double s2 = std::max(sigma2(r), 0.0);
            const double vesc2 = 2.0 * (mhalo / (r + ahalo) + mdisk / std::max(r, rdisk)), sd = std::sqrt(s2);
            for (int tries = 0; tries < 64; tries++) {
                v[0] = sd * gauss(s);
                v[1] = sd * gauss(s);
                v[2] = sd * gauss(s);
                if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < vesc2) break;
            }
        }
    });
    const int64_t n = ndisk + nhalo;
    double mt = 0.0, cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; i++) {
        mt += mass[i];
        for (int k = 0; k < 3; k++) {
            cp[k] += mass[i] * pos[3 * i + k];
            cv[k] += mass[i] * vel[3 * i + k];
        }
    }
    for (int64_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            pos[3 * i + k] -= cp[k] / mt;
            vel[3 * i + k] -= cv[k] / mt;
        }
}
