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
