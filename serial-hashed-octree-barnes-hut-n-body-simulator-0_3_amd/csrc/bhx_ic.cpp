// bhx_ic.cpp — seeded initial conditions (host).  No reference counterpart:
// the reference's plummerModelInitialConditions (ofApp.h:134-168) draws with
// ofRandom, does not virialise and forces every mass to 1.
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11), a
// counter-based generator, so body i's draws depend only on (seed, i, draw).
// Plummer sampling follows Aarseth, Henon & Wielen (1974): radius from the
// inverse cumulative mass, isotropic directions, speed by von Neumann
// rejection from g(q) = q^2 (1 - q^2)^{7/2} times the local escape speed.
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/bhx.h"

namespace {

struct Philox {
    uint32_t c[4];
    uint32_t k[2];
};

inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
}

// Four 32-bit outputs for counter (a, b, d, e) under a 64-bit key.
inline void philox4x32_10(uint32_t out[4], uint64_t key, uint32_t a, uint32_t b, uint32_t d, uint32_t e) {
    uint32_t c[4] = {a, b, d, e};
    uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    for (int r = 0; r < 10; ++r) {
        philox_round(c, k);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

// Uniform double in [0, 1) from 53 random bits.
inline double u01(uint32_t hi, uint32_t lo) {
    const uint64_t v = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6);
    return (double)v * (1.0 / 9007199254740992.0);
}

struct Draws {
    uint64_t seed;
    uint32_t body;
    uint32_t ctr = 0;
    uint32_t buf[4];
    int left = 0;
    double next() {
        if (left < 2) {
            philox4x32_10(buf, seed, body, ctr++, 0x504c554du /* "PLUM" */, 0);
            left = 4;
        }
        const double u = u01(buf[4 - left], buf[5 - left]);
        left -= 2;
        return u;
    }
};

}  // namespace

extern "C" int bhx_ic_plummer(size_t n, uint64_t seed, double total_mass, double a, double* pos, double* vel,
                              double* mass) {
    if (!pos || !vel || !mass) return BHX_E_NULL;
    if (!(total_mass > 0.0) || !(a > 0.0)) return BHX_E_ARG;
    const double two_pi = 6.283185307179586476925286766559;
    const double mcut = 0.999;
    const double mi = total_mass / (double)n;
    for (size_t i = 0; i < n; ++i) {
        Draws d{seed, (uint32_t)i};
        double X = d.next();
        X = 1e-10 + X * (mcut - 1e-10);                  // mass fraction in (0, 0.999]
        const double r = a / std::sqrt(std::pow(X, -2.0 / 3.0) - 1.0);
        const double ct = 1.0 - 2.0 * d.next();
        const double ph = two_pi * d.next();
        const double st = std::sqrt(std::fmax(0.0, 1.0 - ct * ct));
        pos[3 * i] = r * st * std::cos(ph);
        pos[3 * i + 1] = r * st * std::sin(ph);
        pos[3 * i + 2] = r * ct;
        double q = 0.0;
        for (int tries = 0; tries < 1000; ++tries) {
            const double x4 = d.next();
            const double x5 = d.next();
            if (0.1 * x5 < x4 * x4 * std::pow(1.0 - x4 * x4, 3.5)) {
                q = x4;
                break;
            }
        }
        const double ve = std::sqrt(2.0 * total_mass / std::sqrt(r * r + a * a));
        const double v = q * ve;
        const double vct = 1.0 - 2.0 * d.next();
        const double vph = two_pi * d.next();
        const double vst = std::sqrt(std::fmax(0.0, 1.0 - vct * vct));
        vel[3 * i] = v * vst * std::cos(vph);
        vel[3 * i + 1] = v * vst * std::sin(vph);
        vel[3 * i + 2] = v * vct;
        mass[i] = mi;
    }
    // centre-of-mass frame
    double cm[6] = {0, 0, 0, 0, 0, 0}, msum = 0.0;
    for (size_t i = 0; i < n; ++i) {
        msum += mass[i];
        for (int k = 0; k < 3; ++k) {
            cm[k] += mass[i] * pos[3 * i + k];
            cm[3 + k] += mass[i] * vel[3 * i + k];
        }
    }
    if (msum > 0.0)
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) {
                pos[3 * i + k] -= cm[k] / msum;
                vel[3 * i + k] -= cm[3 + k] / msum;
            }
    return BHX_OK;
}

// Sizes, masses and scale radii of an ensemble of Plummer systems: system k
// has N_k = round(exp(U[ln nmin, ln nmax])) bodies (log-uniform), total mass
// and scale radius 1 + jitter * U[-1, 1], each draw from Philox with counter
// (k, draw, "ENSM").
extern "C" int bhx_ic_ensemble_sizes(size_t nsys, uint64_t seed, uint64_t nmin, uint64_t nmax, double jitter,
                                     uint64_t* sizes, double* masses, double* radii) {
    if (!sizes) return BHX_E_NULL;
    if (nmin == 0 || nmax < nmin || !(jitter >= 0.0) || !(jitter < 1.0)) return BHX_E_ARG;
    const double l0 = std::log((double)nmin), l1 = std::log((double)nmax);
    for (size_t k = 0; k < nsys; ++k) {
        uint32_t r[4];
        philox4x32_10(r, seed, (uint32_t)k, 0u, 0x454e534du /* "ENSM" */, 0u);
        const double u = u01(r[0], r[1]);
        uint32_t r2[4];
        philox4x32_10(r2, seed, (uint32_t)k, 1u, 0x454e534du, 0u);
        double nk = std::nearbyint(std::exp(l0 + u * (l1 - l0)));
        if (nk < (double)nmin) nk = (double)nmin;
        if (nk > (double)nmax) nk = (double)nmax;
        sizes[k] = (uint64_t)nk;
        if (masses) masses[k] = 1.0 + jitter * (2.0 * u01(r[2], r[3]) - 1.0);
        if (radii) radii[k] = 1.0 + jitter * (2.0 * u01(r2[0], r2[1]) - 1.0);
    }
    return BHX_OK;
}

// The bodies of that ensemble, system after system: system k is
// bhx_ic_plummer(N_k, seed + 1 + k, M_k, a_k) at offset sum_{j<k} N_j.
// Systems are generated on up to `threads` host threads (0: all cores).
extern "C" int bhx_ic_ensemble(size_t nsys, uint64_t seed, const uint64_t* sizes, const double* masses,
                               const double* radii, int threads, double* pos, double* vel, double* mass) {
    if (!sizes || !masses || !radii || !pos || !vel || !mass) return BHX_E_NULL;
    std::vector<size_t> off(nsys + 1, 0);
    for (size_t k = 0; k < nsys; ++k) off[k + 1] = off[k] + (size_t)sizes[k];
    unsigned nt = threads > 0 ? (unsigned)threads : std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > nsys) nt = (unsigned)(nsys ? nsys : 1);
    std::atomic<size_t> next{0};
    std::atomic<int> status{BHX_OK};
    auto work = [&]() {
        for (size_t k; (k = next.fetch_add(1)) < nsys;) {
            const int rc = bhx_ic_plummer((size_t)sizes[k], seed + 1 + k, masses[k], radii[k], pos + 3 * off[k],
                                          vel + 3 * off[k], mass + off[k]);
            if (rc != BHX_OK) status = rc;
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    return status.load();
}
