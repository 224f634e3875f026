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

int fails = 0, emulated = 0, series_checked = 0;
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

// Invariants of a slot layout (pairs: the one-row-group layout, two reads per row; otherwise
// the port layout) and the product as the kernel forms it (lanes, DPP shifts, reads, term places)
// against the series of the same FP32 coefficients in natural order.
void layout_check(int n, const std::vector<std::vector<std::pair<int, double>>>& adj, double scale,
                  const std::vector<double>& coef, int J, int size, const std::vector<uint16_t>& pos,
                  const std::vector<uint32_t>& far, const std::vector<float4>& cf, bool pairs,
                  std::mt19937& rng) {
    const uint32_t places = (uint32_t)J * kChebBlock;
    const uint32_t zero_slot = places - 1;
    const uint32_t zero = (uint32_t)cheb_term_place(kChebBlock - 1, J - 1, J);
    CHECK(far.size() == places && cf.size() == places, "n %d: %zu far words for %u places", n,
          far.size(), places);
    CHECK((uint32_t)size <= zero_slot, "n %d: %d slots reach the zero slot", n, size);
    // slot <-> contract place <-> term place
    auto slot_of = [&](uint32_t pl) { return (pl % kChebBlock) * (uint32_t)J + pl / kChebBlock; };
    auto term_of = [&](uint32_t sl) { return (uint32_t)cheb_term_place(sl / J, sl % J, J); };
    std::vector<int> term_owner(places, -1);
    for (uint32_t sl = 0; sl < places; ++sl) {
        const uint32_t tp = term_of(sl);
        CHECK(tp < places && term_owner[tp] < 0, "n %d: term place %u of slot %u", n, tp, sl);
        term_owner[tp] = (int)sl;
    }
    std::vector<int> owner(places, -1);  // slot -> row
    for (int i = 0; i < n; ++i) {
        CHECK(pos[i] < places, "n %d: row %d at place %u", n, i, pos[i]);
        const uint32_t sl = slot_of(pos[i]);
        CHECK(sl < zero_slot, "n %d: row %d in slot %u (zero slot %u)", n, i, sl, zero_slot);
        CHECK(owner[sl] < 0, "n %d: rows %d and %d share slot %u", n, owner[sl], i, sl);
        owner[sl] = i;
    }
    for (uint32_t pl = 0; pl < places; ++pl) {
        const uint32_t sl = slot_of(pl);
        CHECK(((far[pl] & kChebNoRow) != 0) == (owner[sl] < 0), "n %d: slot %u row flag", n, sl);
        for (uint32_t rd : {far[pl] & 0xFFFFu, (far[pl] >> 16) & 0x7FFFu}) {
            if (!pairs && rd != (far[pl] & 0xFFFFu)) continue;
            CHECK(rd < places, "n %d: slot %u reads place %u", n, sl, rd);
            CHECK(rd == zero || owner[term_owner[rd]] >= 0, "n %d: slot %u reads an empty slot", n, sl);
        }
    }
    // every neighbour of a row is reached once, with its scaled coefficient: through slot s - 1
    // or s + 1 (same wave), the row's own port, or the port of slot s - 1 (same wave)
    const uint32_t W = (uint32_t)kWave * J;
    for (int i = 0; i < n; ++i) {
        const uint32_t sl = slot_of(pos[i]);
        const float4 a = cf[pos[i]];
        std::vector<std::pair<int, float>> got;
        if (a.x != 0.0f) {
            CHECK(sl % W != 0, "n %d: row %d left neighbour across a wave", n, i);
            got.push_back({owner[sl - 1], a.x});
        }
        if (a.y != 0.0f) {
            CHECK((sl + 1) % W != 0, "n %d: row %d right neighbour across a wave", n, i);
            got.push_back({owner[sl + 1], a.y});
        }
        if (a.z != 0.0f) got.push_back({owner[term_owner[far[pos[i]] & 0xFFFFu]], a.z});
        if (a.w != 0.0f && pairs) {
            got.push_back({owner[term_owner[(far[pos[i]] >> 16) & 0x7FFFu]], a.w});
        } else if (a.w != 0.0f) {
            CHECK(sl % W != 0, "n %d: row %d previous port across a wave", n, i);
            const uint32_t pp = ((sl - 1) % J) * kChebBlock + (sl - 1) / J;
            got.push_back({owner[term_owner[far[pp] & 0xFFFFu]], a.w});
        }
        CHECK(got.size() == adj[i].size(), "n %d: row %d reaches %zu of %zu neighbours", n, i,
              got.size(), adj[i].size());
        for (const auto& e : adj[i]) {
            int hits = 0;
            for (const auto& g : got)
                if (g.first == e.first && g.second == (float)(scale * e.second)) ++hits;
            CHECK(hits == 1, "n %d: row %d neighbour %d reached %d times", n, i, e.first, hits);
        }
    }
    // the product as the kernel forms it (lanes, DPP shifts, ports, term places), against the
    // series of the same FP32 coefficients in natural order
    std::vector<double> x(n);
    std::uniform_real_distribution<double> ux(-80.0, 40.0);
    for (auto& v : x) v = ux(rng);
    const int K = (int)coef.size() - 1;
    std::vector<double> ref(n, 0.0), tm2(n), tm1(x), tk(n);
    for (int i = 0; i < n; ++i) ref[i] = coef[0] * x[i];
    auto apply = [&](const std::vector<double>& v, std::vector<double>& w) {  // w = A~ v
        for (int i = 0; i < n; ++i) {
            double acc = v[i];
            for (const auto& e : adj[i]) acc += (double)(float)(scale * e.second) * (v[e.first] - v[i]);
            w[i] = acc;
        }
    };
    for (int k = 1; k <= K; ++k) {
        apply(tm1, tk);
        for (int i = 0; i < n; ++i) {
            if (k > 1) tk[i] = 2.0 * tk[i] - tm2[i];
            ref[i] += coef[k] * tk[i];
        }
        tm2.swap(tm1);
        tm1.swap(tk);
    }
    const int T = kChebBlock;
    std::vector<double> reg_t((size_t)T * J), reg_u((size_t)T * J, 0.0), out((size_t)T * J);
    std::vector<double> vec0(places, 0.0), vec1(places, 0.0);
    for (int t = 0; t < T; ++t)
        for (int j = 0; j < J; ++j) {
            const uint32_t pl = (uint32_t)(j * kChebBlock + t);
            const int r = owner[slot_of(pl)];
            reg_t[(size_t)t * J + j] = r >= 0 ? x[r] : 0.0;
            vec0[term_of(slot_of(pl))] = reg_t[(size_t)t * J + j];
            out[(size_t)t * J + j] = coef[0] * reg_t[(size_t)t * J + j];
        }
    auto lane_prev = [&](const std::vector<double>& v, int t, int j) {  // lane t-1's value, 0 at lane 0
        return t % kWave == 0 ? 0.0 : v[(size_t)(t - 1) * J + j];
    };
    auto lane_next = [&](const std::vector<double>& v, int t, int j) {
        return t % kWave == kWave - 1 ? 0.0 : v[(size_t)(t + 1) * J + j];
    };
    std::vector<double> port((size_t)T * J), second((size_t)T * J);
    for (int k = 1; k <= K; ++k) {
        const std::vector<double>& rd = (k % 2 == 1) ? vec0 : vec1;
        std::vector<double>& wr = (k % 2 == 1) ? vec1 : vec0;
        for (int t = 0; t < T; ++t)
            for (int j = 0; j < J; ++j) {
                port[(size_t)t * J + j] = rd[far[j * kChebBlock + t] & 0xFFFFu];
                second[(size_t)t * J + j] = rd[(far[j * kChebBlock + t] >> 16) & 0x7FFFu];
            }
        std::vector<double> nt((size_t)T * J);
        for (int t = 0; t < T; ++t)
            for (int j = 0; j < J; ++j) {
                const size_t q = (size_t)t * J + j;
                const float4 a = cf[j * kChebBlock + t];
                const bool row = !(far[j * kChebBlock + t] & kChebNoRow);
                const double cl = row ? 2.0 * a.x : 0.0, cr = row ? 2.0 * a.y : 0.0;
                const double co = row ? 2.0 * a.z : 0.0, cp = row ? 2.0 * a.w : 0.0;
                const double d2 = row ? 2.0 * (1.0 - (((double)a.x + (double)a.y) + ((double)a.z + (double)a.w))) : 0.0;
                const double tl = j > 0 ? reg_t[q - 1] : lane_prev(reg_t, t, J - 1);
                const double tr = j < J - 1 ? reg_t[q + 1] : lane_next(reg_t, t, 0);
                const double vp = pairs ? second[q] : j > 0 ? port[q - 1] : lane_prev(port, t, J - 1);
                double acc = k == 1 ? d2 * reg_t[q] : std::fma(d2, reg_t[q], -reg_u[q]);
                acc = std::fma(cl, tl, acc);
                acc = std::fma(cr, tr, acc);
                acc = std::fma(co, port[q], acc);
                acc = std::fma(cp, vp, acc);
                nt[q] = k == 1 ? 0.5 * acc : acc;
                out[q] = std::fma(coef[k], nt[q], out[q]);
            }
        for (int t = 0; t < T; ++t)
            for (int j = 0; j < J; ++j) {
                const size_t q = (size_t)t * J + j;
                reg_u[q] = reg_t[q];
                reg_t[q] = nt[q];
                wr[term_of((uint32_t)(t * J + j))] = nt[q];
            }
        CHECK(wr[zero] == 0.0, "n %d: the zero place holds %g after term %d", n, wr[zero], k);
    }
    double err = 0.0, mag = 0.0;
    for (int i = 0; i < n; ++i) {
        const uint32_t sl = slot_of(pos[i]);
        err = std::max(err, std::fabs(out[(size_t)(sl / J) * J + sl % J] - ref[i]));
        mag = std::max(mag, std::fabs(ref[i]));
    }
    CHECK(err <= 1e-11 * std::max(1.0, mag), "n %d: emulated product differs by %g (|ref| %g)", n, err, mag);
    ++emulated;
}

void one_cell(int n, int max_children, std::mt19937& rng) {
    std::vector<int32_t> parent = random_tree(n, max_children, rng);
    std::uniform_real_distribution<double> u(0.5, 2.0);
    std::vector<double> R(n), C(n);
    for (int i = 0; i < n; ++i) {
        R[i] = 1e6 * u(rng);
        C[i] = 1e-11 * u(rng);
    }
    // time steps of about 20 and about 100 series terms
    const double dt_s = (n % 2 ? 0.05e-3 : 1e-3);
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
    {  // the port layout (two or more row groups)
        std::vector<uint16_t> pos(n);
        std::vector<uint32_t> far;
        std::vector<float4> cf;
        int J = 0;
        const int size = nw::preorder_layout(n, adj, scale, J, pos, far, cf);
        if (size < 0) CHECK(n > kChebRows * kChebBlock - 64, "n %d: no port layout", n);
        else layout_check(n, adj, scale, coef, J, size, pos, far, cf, false, rng);
    }
    if (n <= kChebBlock - 64) {  // the one-row-group layout (two reads per row)
        std::vector<uint16_t> pos(n);
        std::vector<uint32_t> far;
        std::vector<float4> cf;
        const int size = nw::preorder_layout_pairs(n, adj, scale, pos, far, cf);
        CHECK(size > 0, "n %d: no one-row-group layout", n);
        layout_check(n, adj, scale, coef, 1, size, pos, far, cf, true, rng);
    }
}

// Poisson weights of the uniformization build: they sum to one, their mean is sigma, and each
// kept weight equals e^-sigma sigma^k / k! (from lgamma) where that is above 1e-200.
void poisson_checks() {
    for (double sigma : {0.0, 1e-3, 0.7, 5.0, 40.0, 300.0, 1500.0}) {
        std::vector<double> w;
        const int K = nw::poisson_weights(sigma, nw::kSeriesTol, w);
        CHECK(K >= 0 && (int)w.size() == K + 1, "sigma %g: %d weights for K %d", sigma, (int)w.size(), K);
        double sum = 0.0, mean = 0.0, worst = 0.0;
        for (int k = 0; k <= K; ++k) {
            sum += w[k];
            mean += k * w[k];
            if (sigma > 0.0) {
                const double ref = std::exp(-sigma + k * std::log(sigma) - std::lgamma(k + 1.0));
                if (ref > 1e-200 && w[k] > 0.0) worst = std::max(worst, std::fabs(w[k] / ref - 1.0));
            }
        }
        CHECK(std::fabs(sum - 1.0) < 1e-14, "sigma %g: weights sum to %.17g", sigma, sum);
        CHECK(std::fabs(mean - sigma) <= 1e-10 * std::max(1.0, sigma), "sigma %g: mean %.17g", sigma, mean);
        CHECK(worst < 1e-10, "sigma %g: weight off the Poisson formula by %g", sigma, worst);
    }
}

// The squaring choice: none for a soft cell, more as sigma grows, never past 8.
void squaring_choice_checks() {
    CHECK(nw::series_squarings(5000, 10.0) == 0, "squarings for sigma 10");
    const int s8 = nw::series_squarings(5000, 8000.0);
    CHECK(s8 >= 2 && s8 <= 8, "squarings for sigma 8000: %d", s8);
}

// The uniformization build of one cell, its kernel's arithmetic emulated on the host (column j,
// row j from the transposition identity), against a dense FP64 Taylor expm with scaling and
// squaring of the nonnegative shifted matrix: relative 1e-11 on every entry above 1e-14 and
// 1e-15 absolute elsewhere; rows sum to one.
void series_build_check(int n, std::mt19937& rng, int squarings = 0) {
    std::vector<int32_t> parent = random_tree(n, 2, rng);
    std::uniform_real_distribution<double> u(0.5, 2.0);
    std::vector<double> R(n), C(n);
    for (int i = 0; i < n; ++i) {
        R[i] = 1e7 * u(rng);
        C[i] = 1e-12 * u(rng);
    }
    const double dt_ms = 0.05;
    nw::SeriesOperator op;
    const int K = nw::series_operator(n, parent.data(), R.data(), C.data(),
                                      std::ldexp(dt_ms, -squarings), op);
    CHECK(K >= 0, "n %d: no series operator", n);
    std::vector<double> P((size_t)n * n);
    std::vector<double> v(n), t(n), nv(n), out(n);
    for (int col = 0; col < n; ++col) {
        for (int r = 0; r < n; ++r) {
            t[r] = v[r] = r == col ? 1.0 : 0.0;
            out[r] = op.w[0] * t[r];
        }
        for (int k = 1; k <= K; ++k) {
            for (int r = 0; r < n; ++r) {
                double sacc = op.dg[r] * t[r];
                sacc = std::fma(op.nc[3 * r], v[op.nb[3 * r]], sacc);
                sacc = std::fma(op.nc[3 * r + 1], v[op.nb[3 * r + 1]], sacc);
                if (op.nb[3 * r + 2] >= 0) {
                    sacc = std::fma(op.nc[3 * r + 2], v[op.nb[3 * r + 2]], sacc);
                } else {
                    for (int e = -op.nb[3 * r + 2] - 1; op.ovi[e] >= 0; ++e)
                        sacc = std::fma(op.ovc[e], v[op.ovi[e]], sacc);
                }
                nv[r] = sacc;
                out[r] = std::fma(op.w[k], sacc, out[r]);
            }
            t = nv;
            v = nv;
        }
        for (int r = 0; r < n; ++r) P[(size_t)col * n + r] = out[r] * (C[r] * (1.0 / C[col]));
    }
    // dense reference: B = A dt; X = (B + lambda I) / 2^s >= 0, exp(B) = e^-lambda (T20(X))^(2^s)
    const double dt = dt_ms * 1e-3;
    std::vector<double> B((size_t)n * n, 0.0);
    for (int c = 0; c < n; ++c) {
        const int q = parent[c];
        if (q < 0) continue;
        B[(size_t)c * n + q] += dt / (R[c] * C[c]);
        B[(size_t)c * n + c] -= dt / (R[c] * C[c]);
        B[(size_t)q * n + c] += dt / (R[c] * C[q]);
        B[(size_t)q * n + q] -= dt / (R[c] * C[q]);
    }
    double lambda = 0.0;
    for (int i = 0; i < n; ++i) lambda = std::max(lambda, -B[(size_t)i * n + i]);
    int sq = 0;
    while (lambda * std::ldexp(1.0, -sq) > 0.25) ++sq;
    std::vector<double> X((size_t)n * n), E((size_t)n * n, 0.0), T((size_t)n * n), Y((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            X[(size_t)i * n + j] = (B[(size_t)i * n + j] + (i == j ? lambda : 0.0)) * std::ldexp(1.0, -sq);
    auto mul = [&](const std::vector<double>& A1, const std::vector<double>& A2, std::vector<double>& O) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double acc = 0.0;
                for (int l = 0; l < n; ++l) acc += A1[(size_t)i * n + l] * A2[(size_t)l * n + j];
                O[(size_t)i * n + j] = acc;
            }
    };
    // Horner: E = I + X (I + X/2 (I + X/3 ...)), terms to X^20 / 20!
    for (int i = 0; i < n; ++i) E[(size_t)i * n + i] = 1.0;
    for (int k = 20; k >= 1; --k) {
        mul(X, E, T);
        for (size_t q = 0; q < T.size(); ++q) E[q] = T[q] / k;
        for (int i = 0; i < n; ++i) E[(size_t)i * n + i] += 1.0;
    }
    const double shrink = std::exp(-lambda * std::ldexp(1.0, -sq));
    for (auto& x : E) x *= shrink;
    for (int k = 0; k < sq; ++k) {
        mul(E, E, Y);
        E.swap(Y);
    }
    for (int q = 0; q < squarings; ++q) {  // the build's squarings of exp(B / 2^s)
        mul(P, P, Y);
        P.swap(Y);
    }
    double worst_rel = 0.0, worst_abs = 0.0, worst_sum = 0.0;
    for (int i = 0; i < n; ++i) {
        double rs = 0.0;
        for (int j = 0; j < n; ++j) {
            const double a = P[(size_t)i * n + j], b = E[(size_t)i * n + j];
            rs += a;
            if (std::fabs(b) > 1e-14) worst_rel = std::max(worst_rel, std::fabs(a - b) / std::fabs(b));
            else worst_abs = std::max(worst_abs, std::fabs(a - b));
        }
        worst_sum = std::max(worst_sum, std::fabs(rs - 1.0));
    }
    CHECK(worst_rel < 1e-11 && worst_abs < 1e-15 && worst_sum < 1e-13,
          "n %d, %d squarings: uniformization vs dense expm: rel %g, abs %g, row sums %g", n,
          squarings, worst_rel, worst_abs, worst_sum);
    ++series_checked;
}

}  // namespace

int main() {
    std::mt19937 rng(12345);
    const int sizes[] = {1, 2, 3, 63, 64, 65, 127, 1023, 1024, 1025, 2047, 3000, 4095, 4096, 5000};
    for (int n : sizes)
        for (int k : {2, 3})
            one_cell(n, k, rng);
    for (int t = 0; t < 60; ++t) one_cell(1 + (int)(rng() % 5000u), 2 + (int)(rng() % 2u), rng);
    poisson_checks();
    for (int n : {1, 2, 17, 120, 260})
        series_build_check(n, rng);
    for (int n : {17, 120, 260})
        series_build_check(n, rng, 3);
    squaring_choice_checks();
    std::printf("layout harness: %d cells emulated, %d series builds checked, %d failures\n",
                emulated, series_checked, fails);
    return fails ? 1 : 0;
}
