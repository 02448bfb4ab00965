// bhx_ic.cpp — seeded initial conditions (host).  No reference counterpart:
// the reference's plummerModelInitialConditions (ofApp.h:134-168) draws with
// ofRandom, does not virialise and forces every mass to 1.
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11), a
// counter-based generator, so body i's draws depend only on (seed, i, draw).
// Plummer sampling follows Aarseth, Henon & Wielen (1974): radius from the
// inverse cumulative mass, isotropic directions, speed by von Neumann
// rejection from g(q) = q^2 (1 - q^2)^{7/2} times the local escape speed.
#include <algorithm>
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
    uint32_t tag = 0x504c554du;   // "PLUM"; the galaxy generator uses "GALX"
    double next() {
        if (left < 2) {
            philox4x32_10(buf, seed, body, ctr++, tag, 0);
            left = 4;
        }
        const double u = u01(buf[4 - left], buf[5 - left]);
        left -= 2;
        return u;
    }
    // standard normal (Box-Muller, one of the pair)
    double normal() {
        const double u1 = 1.0 - next();   // (0, 1]
        const double u2 = next();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925286766559 * u2);
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

// ---------------------------------------------------------------- galaxy pair

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

// Isotropic distribution function of the Hernquist (1990) sphere, eq. 17, up
// to its constant factor, as a function of q^2 = E a / (G M) in (0, 1) with E
// the (positive) binding energy per unit mass.
double hernquist_df(double q2) {
    if (!(q2 > 0.0)) return 0.0;
    if (q2 > 1.0 - 1e-15) q2 = 1.0 - 1e-15;
    const double q = std::sqrt(q2), s = std::sqrt(1.0 - q2);
    const double g = 3.0 * std::asin(q) + q * s * (1.0 - 2.0 * q2) * (8.0 * q2 * q2 - 8.0 * q2 - 3.0);
    return g > 0.0 ? g / std::pow(1.0 - q2, 2.5) : 0.0;
}

// Speed at radius r in a Hernquist halo of total mass M and scale a: density
// v^2 f(Psi(r) - v^2/2) on [0, v_esc), Psi = M / (r + a), by rejection under
// a bound from a grid dense near v = 0 (where the peak sits for small r),
// raised whenever a draw exceeds it.
double hernquist_speed(Draws& d, double r, double M, double a) {
    const double psi = M / (r + a);
    const double vesc = std::sqrt(2.0 * psi);
    auto g = [&](double v) { return v * v * hernquist_df((psi - 0.5 * v * v) * a / M); };
    double bound = 0.0, vbest = 0.0;
    for (int k = 1; k <= 128; ++k) {
        const double t = k / 128.0;
        const double v = vesc * t * t * t;
        const double gv = g(v);
        if (gv > bound) { bound = gv; vbest = v; }
    }
    bound *= 1.5;
    if (!(bound > 0.0)) return 0.0;
    for (int tries = 0; tries < 1000000; ++tries) {
        const double v = vesc * d.next();
        const double y = bound * d.next();
        const double gv = g(v);
        if (gv > bound) {       // the grid missed the peak: raise the bound, draw again
            bound = 1.5 * gv;
            continue;
        }
        if (y < gv) return v;
    }
    return vbest;
}

// Circular speed^2 in the plane of a galaxy: Hernquist halo (untruncated mass
// Mh, scale a) plus an exponential disk (mass Md, scale Rd; Freeman 1970:
// v^2 = 2 (Md / Rd) y^2 [I0(y) K0(y) - I1(y) K1(y)], y = R / (2 Rd)).
double vcirc2(double R, double Mh, double a, double Md, double Rd) {
    if (!(R > 1e-12)) return 0.0;
    const double vh = Mh * R / ((R + a) * (R + a));
    double vd = 0.0;
    if (Md > 0.0) {
        const double y = R / (2.0 * Rd);
        vd = 2.0 * (Md / Rd) * y * y *
             (std::cyl_bessel_i(0.0, y) * std::cyl_bessel_k(0.0, y) - std::cyl_bessel_i(1.0, y) * std::cyl_bessel_k(1.0, y));
        if (!(vd > 0.0)) vd = 0.0;
    }
    return vh + vd;
}

struct Galaxy {
    size_t first, nh, nd;          // body range: halo [first, first+nh), disk after it
    double mh_inf, a, rcut_x;      // untruncated halo mass, scale, mass fraction at the cut
    double md, rd, zd, rdcut;      // disk mass, scale length, scale height, cut radius
    double sigma, incl;
};

void galaxy_body(const Galaxy& g, size_t i, uint64_t seed, double* pos, double* vel) {
    Draws d{seed, (uint32_t)i};
    d.tag = 0x47414c58u;   // "GALX"
    double x[3], v[3];
    if (i < g.first + g.nh) {
        // halo: radius from the cumulative mass M(<r)/M = r^2 / (r + a)^2 up to the cut
        const double X = 1e-12 + d.next() * (g.rcut_x - 1e-12);
        const double sx = std::sqrt(X);
        const double r = g.a * sx / (1.0 - sx);
        const double ct = 1.0 - 2.0 * d.next(), ph = kTwoPi * d.next();
        const double st = std::sqrt(std::fmax(0.0, 1.0 - ct * ct));
        x[0] = r * st * std::cos(ph);
        x[1] = r * st * std::sin(ph);
        x[2] = r * ct;
        const double sp = hernquist_speed(d, r, g.mh_inf, g.a);
        const double vct = 1.0 - 2.0 * d.next(), vph = kTwoPi * d.next();
        const double vst = std::sqrt(std::fmax(0.0, 1.0 - vct * vct));
        v[0] = sp * vst * std::cos(vph);
        v[1] = sp * vst * std::sin(vph);
        v[2] = sp * vct;
    } else {
        // disk: R ~ Gamma(2) R_d below the cut, z from the sech^2 profile
        double R;
        do {
            R = g.rd * (-std::log(1.0 - d.next()) - std::log(1.0 - d.next()));
        } while (R > g.rdcut);
        const double ph = kTwoPi * d.next();
        double u = d.next();
        u = std::min(std::max(u, 1e-12), 1.0 - 1e-12);
        const double z = g.zd * std::atanh(2.0 * u - 1.0);
        const double vc = std::sqrt(vcirc2(R, g.mh_inf, g.a, g.md, g.rd));
        const double sg = g.sigma * vc;
        const double vR = sg * d.normal(), vp = vc + sg * d.normal(), vz = sg * d.normal();
        const double c = std::cos(ph), s = std::sin(ph);
        const double lx = R * c, ly = R * s, lvx = vR * c - vp * s, lvy = vR * s + vp * c;
        // inclination about the x axis
        const double ci = std::cos(g.incl), si = std::sin(g.incl);
        x[0] = lx;
        x[1] = ly * ci - z * si;
        x[2] = ly * si + z * ci;
        v[0] = lvx;
        v[1] = lvy * ci - vz * si;
        v[2] = lvy * si + vz * ci;
    }
    for (int k = 0; k < 3; ++k) {
        pos[3 * i + k] = x[k];
        vel[3 * i + k] = v[k];
    }
}

}  // namespace

extern "C" int bhx_galaxy_defaults(bhx_galaxy_params* p) {
    if (!p) return BHX_E_NULL;
    p->total_mass = 1.0;
    p->mass_ratio = 3.0;
    p->disk_fraction = 0.2;
    p->disk_scale = 1.0;
    p->disk_height = 0.1;
    p->disk_cut = 10.0;
    p->halo_scale = 3.0;
    p->halo_cut = 10.0;
    p->disk_sigma = 0.1;
    p->separation = 30.0;
    p->pericentre = 5.0;
    p->incl1 = 0.0;
    p->incl2 = kTwoPi / 6.0;
    return BHX_OK;
}

extern "C" int bhx_ic_galaxy_pair(size_t n, uint64_t seed, const bhx_galaxy_params* p, int threads, double* pos,
                                  double* vel, double* mass) {
    if (!p || !pos || !vel || !mass) return BHX_E_NULL;
    if (n < 4 || n > 0xFFFFFFFFull) return BHX_E_ARG;
    if (!(p->total_mass > 0.0) || !(p->mass_ratio > 0.0) || !(p->disk_fraction >= 0.0) ||
        !(p->disk_fraction <= 1.0) || !(p->disk_scale > 0.0) || !(p->disk_height > 0.0) || !(p->disk_cut > 0.0) ||
        !(p->halo_scale > 0.0) || !(p->halo_cut > 0.0) || !(p->disk_sigma >= 0.0) || !(p->separation > 0.0) ||
        !(p->pericentre > 0.0) || !(p->pericentre <= p->separation) || !std::isfinite(p->incl1) ||
        !std::isfinite(p->incl2))
        return BHX_E_ARG;
    const double mi = p->total_mass / (double)n;
    size_t n1 = (size_t)std::llround((double)n * p->mass_ratio / (1.0 + p->mass_ratio));
    n1 = std::min(std::max(n1, (size_t)1), n - 1);
    const size_t counts[2] = {n1, n - n1};
    Galaxy gal[2];
    size_t first = 0;
    for (int k = 0; k < 2; ++k) {
        Galaxy& g = gal[k];
        const double scale = k == 0 ? 1.0 : std::cbrt((double)counts[1] / (double)counts[0]);
        g.first = first;
        g.nd = (size_t)std::llround((double)counts[k] * p->disk_fraction);
        g.nh = counts[k] - g.nd;
        const double c = p->halo_cut;
        g.rcut_x = c * c / ((1.0 + c) * (1.0 + c));
        g.a = p->halo_scale * scale;
        g.mh_inf = (double)g.nh * mi / g.rcut_x;
        g.md = (double)g.nd * mi;
        g.rd = p->disk_scale * scale;
        g.zd = p->disk_height * scale;
        g.rdcut = p->disk_cut * g.rd;
        g.sigma = p->disk_sigma;
        g.incl = k == 0 ? p->incl1 : p->incl2;
        first += counts[k];
    }
    unsigned nt = threads > 0 ? (unsigned)threads : std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    const size_t block = 4096;
    const size_t nblocks = (n + block - 1) / block;
    if (nt > nblocks) nt = (unsigned)nblocks;
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t b; (b = next.fetch_add(1)) < nblocks;) {
            const size_t e = std::min(n, (b + 1) * block);
            for (size_t i = b * block; i < e; ++i) {
                galaxy_body(i < gal[1].first ? gal[0] : gal[1], i, seed, pos, vel);
                mass[i] = mi;
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    // each galaxy to its own centre-of-mass frame (sequential sums: the result
    // does not depend on the thread count), then onto the orbit
    const double mtot = p->total_mass;
    double gm[2];
    for (int k = 0; k < 2; ++k) {
        const Galaxy& g = gal[k];
        const size_t e = g.first + g.nh + g.nd;
        double cm[6] = {0, 0, 0, 0, 0, 0};
        for (size_t i = g.first; i < e; ++i)
            for (int a = 0; a < 3; ++a) {
                cm[a] += pos[3 * i + a];
                cm[3 + a] += vel[3 * i + a];
            }
        const double cnt = (double)(e - g.first);
        for (size_t i = g.first; i < e; ++i)
            for (int a = 0; a < 3; ++a) {
                pos[3 * i + a] -= cm[a] / cnt;
                vel[3 * i + a] -= cm[3 + a] / cnt;
            }
        gm[k] = cnt * mi;
    }
    // parabolic relative orbit: |v| = sqrt(2 M / d), angular momentum sqrt(2 M r_p)
    const double d0 = p->separation;
    const double v2 = 2.0 * mtot / d0;
    const double vt = std::sqrt(2.0 * mtot * p->pericentre) / d0;
    const double vr = -std::sqrt(std::fmax(0.0, v2 - vt * vt));
    const double rel[3] = {d0, 0.0, 0.0}, relv[3] = {vr, vt, 0.0};
    for (int k = 0; k < 2; ++k) {
        const Galaxy& g = gal[k];
        const double f = k == 0 ? -gm[1] / mtot : gm[0] / mtot;
        const size_t e = g.first + g.nh + g.nd;
        for (size_t i = g.first; i < e; ++i)
            for (int a = 0; a < 3; ++a) {
                pos[3 * i + a] += f * rel[a];
                vel[3 * i + a] += f * relv[a];
            }
    }
    return BHX_OK;
}
