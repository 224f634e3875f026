// Host-side builders of the Chebyshev store run under AddressSanitizer and UndefinedBehavior-
// Sanitizer (test infrastructure: built by tests/sanitize/Makefile, run by
// tests/test_sanitizers.py; no GPU call is made).  The propagator source is compiled into this
// program so its file-local builders can be called directly: random trees of 1..5,000 rows,
// binary and with rows of more neighbours, through the Sylvester-count spectral radius, the
// series coefficients and the preorder slot layout, whose invariants are checked.
#include "../../neuwon_amd/csrc/nw_propagator.hip"

#include <cstdio>
#include <random>

namespace {

int fails = 0;
#define CHECK(c, ...)                              \
    do {                                           \
        if (!(c)) {                                \
            std::fprintf(stderr, __VA_ARGS__);     \
            std::fprintf(stderr, "\n");            \
            ++fails;                               \
            return;                                \
        }                                          \
    } while (0)

// A random tree of n rows, parents before children, at most max_children children per row.
std::vector<int32_t> random_tree(int n, int max_children, std::mt19937& rng) {
    std::vector<int32_t> parent(n, -1), kids(n, 0);
    for (int i = 1; i < n; ++i) {
        int p;
        do p = (int)(rng() % (uint32_t)i);
        while (kids[p] >= max_children);
        parent[i] = p;
        ++kids[p];
    }
    return parent;
}

void one_cell(int n, int max_children, std::mt19937& rng) {
    std::vector<int32_t> parent = random_tree(n, max_children, rng);
    std::uniform_real_distribution<double> u(0.5, 2.0);
    std::vector<double> R(n), C(n);
    for (int i = 0; i < n; ++i) {
        R[i] = 1e6 * u(rng);
        C[i] = 1e-11 * u(rng);
    }
    const double dt_s = 0.05e-3;
    std::vector<std::vector<std::pair<int, double>>> adj(n);
    std::vector<double> diag(n, 0.0), off(n, 0.0);
    for (int c = 0; c < n; ++c) {
        const int q = parent[c];
        if (q < 0) continue;
        const double bc = dt_s / (R[c] * C[c]), bq = dt_s / (R[c] * C[q]);
        adj[c].push_back({q, bc});
        adj[q].push_back({c, bq});
        diag[c] += bc;
        diag[q] += bq;
        off[c] = -dt_s / (R[c] * std::sqrt(C[c] * C[q]));
    }
    std::vector<double> work;
    const double rho = nw::tree_spectral_radius(n, parent.data(), diag.data(), off.data(), work);
    double dmax = 0.0;
    for (double d : diag) dmax = std::max(dmax, d);
    CHECK(n == 1 || (rho >= dmax && rho <= 2.0 * dmax * (1.0 + 1e-7)),
          "n %d: spectral radius %g outside [max diag %g, 2 max diag]", n, rho, dmax);
    std::vector<double> coef;
    CHECK(nw::exp_chebyshev(rho, nw::kChebTol, coef) == 0, "n %d: no series for rho %g", n, rho);
    double sum = 0.0;
    for (double c : coef) sum += c;
    CHECK(std::fabs(sum - 1.0) < 1e-13, "n %d: series coefficients sum to %.17g", n, sum);
    int deg = 0;
    for (const auto& a : adj) deg = std::max<int>(deg, (int)a.size());
    if (deg > 3) return;  // such cells keep the natural order (no slot layout)
    const double scale = rho > 0.0 ? 2.0 / rho : 0.0;
    std::vector<uint16_t> pos(n);
    std::vector<uint32_t> far;
    std::vector<float4> cf;
    int J = 0;
    const int size = nw::preorder_layout(n, adj, scale, J, pos, far, cf);
    if (size < 0) {
        CHECK(n > kChebRows * kChebBlock - 64, "n %d: no layout", n);
        return;
    }
    const uint32_t places = (uint32_t)J * kChebBlock, zero = places - 1;
    CHECK(far.size() == places && cf.size() == places, "n %d: %zu far words for %u places", n,
          far.size(), places);
    std::vector<int> owner(places, -1);
    for (int i = 0; i < n; ++i) {
        CHECK(pos[i] < zero, "n %d: row %d at place %u (zero place %u)", n, i, pos[i], zero);
        CHECK(owner[pos[i]] < 0, "n %d: rows %d and %d share place %u", n, owner[pos[i]], i, pos[i]);
        owner[pos[i]] = i;
    }
    // every neighbour of a row is reached once: its slot neighbour (coefficient 0 or 1 of the
    // place) or one of its two far places, with the scaled coefficient
    for (int i = 0; i < n; ++i) {
        const uint32_t w = far[pos[i]];
        CHECK(w != kChebHole, "n %d: row %d has a hole word", n, i);
        const uint32_t f0 = w & 0xFFFFu, f1 = w >> 16;
        CHECK(f0 < places && f1 < places, "n %d: row %d far places %u %u", n, i, f0, f1);
        const float4 a = cf[pos[i]];
        double nsum = 0.0;
        for (const auto& e : adj[i]) nsum += (double)(float)(scale * e.second);
        const double csum = (double)a.x + (double)a.y + (double)a.z + (double)a.w;
        CHECK(std::fabs(csum - nsum) <= 1e-6 * std::max(1.0, std::fabs(nsum)),
              "n %d: row %d coefficients %g, neighbours %g", n, i, csum, nsum);
        for (uint32_t fp : {f0, f1})
            CHECK(fp == zero || owner[fp] >= 0, "n %d: row %d reads empty place %u", n, i, fp);
    }
}

}  // namespace

int main() {
    std::mt19937 rng(12345);
    const int sizes[] = {1, 2, 3, 63, 64, 65, 127, 1023, 1024, 1025, 2047, 3000, 4095, 4096, 5000};
    for (int n : sizes)
        for (int k : {2, 3})
            one_cell(n, k, rng);
    for (int t = 0; t < 60; ++t) one_cell(1 + (int)(rng() % 5000u), 2 + (int)(rng() % 2u), rng);
    std::printf("layout harness: %d failures\n", fails);
    return fails ? 1 : 0;
}
