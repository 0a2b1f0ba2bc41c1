// Host-side pieces: Philox RNG and lattice builders.
#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

#include "hmc.hpp"

namespace nslx {

// ---------------------------------------------------------------- Philox
// Philox-4x32 with 10 rounds (Salmon, Moraes, Dror, Shaw, SC'11).
static inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint64_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = M0 * c[0], p1 = M1 * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        if (r < 9) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    }
}

void host_philox(uint64_t seed, uint64_t offset, int64_t nblocks, uint32_t* out4) {
    for (int64_t i = 0; i < nblocks; ++i) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)offset, (uint32_t)(offset >> 32), 0u};
        philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        out4[4 * i + 0] = c[0]; out4[4 * i + 1] = c[1]; out4[4 * i + 2] = c[2]; out4[4 * i + 3] = c[3];
    }
}

static inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

void host_uniform(uint64_t seed, uint64_t offset, int64_t n, double* out) {
    const int64_t nb = (n + 1) / 2;
    for (int64_t i = 0; i < nb; ++i) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)offset, (uint32_t)(offset >> 32), 0u};
        philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        out[2 * i] = u53(c[0], c[1]);
        if (2 * i + 1 < n) out[2 * i + 1] = u53(c[2], c[3]);
    }
}

void host_normal(uint64_t seed, uint64_t offset, int64_t n, double* out) {
    const int64_t nb = (n + 1) / 2;
    for (int64_t i = 0; i < nb; ++i) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)offset, (uint32_t)(offset >> 32), 0u};
        philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double u0 = u53(c[0], c[1]), u1 = u53(c[2], c[3]);
        const double r = std::sqrt(-2.0 * std::log(1.0 - u0));
        out[2 * i] = r * std::cos(2.0 * M_PI * u1);
        if (2 * i + 1 < n) out[2 * i + 1] = r * std::sin(2.0 * M_PI * u1);
    }
}

// ---------------------------------------------------------------- lattices
static const double kS32 = 0.8660254037844386;   // sqrt(3)/2, as in honeycomb.tpp:76-80

// Periodic honeycomb: site/cell order and the min-image distance test of
// src/NSL/Lattice/Implementations/honeycomb.tpp:47-117.
void lattice_honeycomb(int L0, int L1, double kappa, double* K, double* sites) {
    NSLX_REQUIRE(L0 >= 1 && L1 >= 1 && K, "honeycomb: bad arguments");
    const int n = 2 * L0 * L1;
    const double a0[2] = {1.5, kS32}, a1[2] = {1.5, -kS32}, r[2] = {1.0, 0.0};
    std::vector<double> xy(2 * n);
    for (int l1 = 0; l1 < L0; ++l1)
        for (int l2 = 0; l2 < L1; ++l2) {
            const int i = l1 + L0 * l2;
            for (int d = 0; d < 2; ++d) {
                xy[2 * (2 * i) + d] = l1 * a0[d] + l2 * a1[d] + r[d] / 2;
                xy[2 * (2 * i + 1) + d] = l1 * a0[d] + l2 * a1[d] - r[d] / 2;
            }
        }
    const double P0[2] = {L0 * a0[0], L0 * a0[1]}, P1[2] = {L1 * a1[0], L1 * a1[1]};
    const double far = (P0[0] + P1[0]) * (P0[0] + P1[0]) + (P0[1] + P1[1]) * (P0[1] + P1[1]);
    std::fill(K, K + (size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            double best = far;
            for (int s0 = -1; s0 <= 1; ++s0)
                for (int s1 = -1; s1 <= 1; ++s1) {
                    const double vx = xy[2 * i] - xy[2 * j] + s0 * P0[0] + s1 * P1[0];
                    const double vy = xy[2 * i + 1] - xy[2 * j + 1] + s0 * P0[1] + s1 * P1[1];
                    best = std::min(best, vx * vx + vy * vy);
                }
            if (std::fabs(best - 1.0) < 1e-4) { K[(size_t)i * n + j] = kappa; K[(size_t)j * n + i] = kappa; }
        }
    if (sites) std::copy(xy.begin(), xy.end(), sites);
}

// Ring: src/NSL/Lattice/Implementations/ring.tpp:8-42 (real kappa).
void lattice_ring(int n, double kappa, double* K, double* sites) {
    NSLX_REQUIRE(n >= 1 && K, "ring: bad arguments");
    std::fill(K, K + (size_t)n * n, 0.0);
    for (int i = 0; i < n - 1; ++i) K[(size_t)i * n + i + 1] = kappa;
    for (int i = 1; i < n; ++i) K[(size_t)i * n + i - 1] = kappa;
    if (n > 1) { K[(size_t)0 * n + n - 1] = kappa; K[(size_t)(n - 1) * n + 0] = kappa; }
    else K[0] = 0.0;
    if (sites) {
        const double th = 2 * M_PI / n;
        for (int i = 0; i < n; ++i) { sites[3 * i] = std::cos(i * th); sites[3 * i + 1] = std::sin(i * th); sites[3 * i + 2] = 0.0; }
    }
}

// (n,n) armchair nanotube, open ends: graphene (honeycomb basis above)
// periodic in x with period 3n, cut to 0 <= y < cells*sqrt(3); sites sorted
// by the integer keys (round(1e6*y), round(1e6*x)); hop kappa at distance 1.
void lattice_nanotube(int n, int cells, double kappa, double* K, double* sites) {
    NSLX_REQUIRE(n >= 2 && cells >= 1 && K, "nanotube: bad arguments");
    const double a0[2] = {1.5, kS32}, a1[2] = {1.5, -kS32}, r[2] = {1.0, 0.0};
    const double circ = 3.0 * n, height = cells * 2 * kS32;
    std::map<std::pair<long long, long long>, std::pair<double, double>> pts;
    const int R = 2 * (n + cells) + 4;
    for (int l0 = -R; l0 <= R; ++l0)
        for (int l1 = -R; l1 <= R; ++l1)
            for (int sg = 1; sg >= -1; sg -= 2) {
                const double px = l0 * a0[0] + l1 * a1[0] + sg * r[0] / 2;
                const double py = l0 * a0[1] + l1 * a1[1] + sg * r[1] / 2;
                if (py < -1e-9 || py >= height - 1e-9) continue;
                double x = std::fmod(px, circ);
                if (x < 0) x += circ;
                if (x > circ - 1e-9) x -= circ;
                const auto key = std::make_pair(std::llround(py * 1e6), std::llround(x * 1e6));
                pts[key] = {x, py};
            }
    const int N = (int)pts.size();
    NSLX_REQUIRE(N == 4 * n * cells, "nanotube: unexpected site count");
    std::vector<std::pair<double, double>> v;
    v.reserve(N);
    for (auto& kv : pts) v.push_back(kv.second);
    std::fill(K, K + (size_t)N * N, 0.0);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double dx = v[i].first - v[j].first;
            dx -= circ * std::nearbyint(dx / circ);
            const double dy = v[i].second - v[j].second;
            if (std::fabs(dx * dx + dy * dy - 1.0) < 1e-4) K[(size_t)i * N + j] = kappa;
        }
    if (sites) {
        for (int i = 0; i < N; ++i) {
            const double th = 2 * M_PI * v[i].first / circ, rad = circ / (2 * M_PI);
            sites[3 * i] = rad * std::cos(th);
            sites[3 * i + 1] = rad * std::sin(th);
            sites[3 * i + 2] = v[i].second;
        }
    }
}

}  // namespace nslx
