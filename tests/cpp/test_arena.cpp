// CPU check of the serving arena's offset allocator (trident_amd/csrc/arena.h)
// against a brute-force occupancy map: blocks stay in range and disjoint, an
// allocation fails only when no free run holds it, the chosen run is the
// smallest that does (best fit), and freeing everything leaves one range.
#include "../../trident_amd/csrc/arena.h"

#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

static int fails = 0;
#define CHECK(c, ...)                                  \
    do {                                               \
        if (!(c)) {                                    \
            std::fprintf(stderr, "FAIL: " __VA_ARGS__); \
            std::fprintf(stderr, "\n");                \
            if (++fails > 10) std::exit(1);            \
        }                                              \
    } while (0)

// free runs of the occupancy map: (offset, length)
static std::vector<std::pair<size_t, size_t>> runs(const std::vector<char> &occ) {
    std::vector<std::pair<size_t, size_t>> r;
    size_t i = 0;
    while (i < occ.size()) {
        if (occ[i]) {
            ++i;
            continue;
        }
        size_t j = i;
        while (j < occ.size() && !occ[j]) ++j;
        r.push_back({i, j - i});
        i = j;
    }
    return r;
}

int main() {
    std::mt19937_64 rng(5);
    for (int round = 0; round < 20; ++round) {
        const size_t N = 4096 + (rng() % 8192);
        tr::Arena a;
        a.init(N);
        std::vector<char> occ(N, 0);
        std::vector<std::pair<size_t, size_t>> live;
        for (int op = 0; op < 4000; ++op) {
            const bool do_alloc = live.empty() || (rng() % 100) < 55;
            if (do_alloc) {
                const size_t want = 1 + rng() % (round % 2 ? 64 : 700);
                size_t off = 0;
                const auto fr = runs(occ);
                size_t best = SIZE_MAX;
                for (auto &r : fr)
                    if (r.second >= want && r.second < best) best = r.second;
                const bool ok = a.alloc(want, off);
                CHECK(ok == (best != SIZE_MAX), "alloc(%zu) returned %d, brute force %d", want, ok, best != SIZE_MAX);
                if (!ok) continue;
                CHECK(off + want <= N, "block [%zu, %zu) past %zu", off, off + want, N);
                size_t run_len = 0;
                for (auto &r : fr)
                    if (off >= r.first && off < r.first + r.second) run_len = r.second;
                CHECK(run_len == best, "alloc(%zu) took a run of %zu, best fit is %zu", want, run_len, best);
                for (size_t k = off; k < off + want && k < N; ++k) {
                    CHECK(!occ[k], "byte %zu handed out twice", k);
                    occ[k] = 1;
                }
                live.push_back({off, want});
            } else {
                const size_t i = rng() % live.size();
                const auto b = live[i];
                live[i] = live.back();
                live.pop_back();
                a.free(b.first, b.second);
                for (size_t k = b.first; k < b.first + b.second; ++k) occ[k] = 0;
            }
            size_t freeb = 0;
            for (char c : occ) freeb += c ? 0 : 1;
            CHECK(a.free_bytes() == freeb, "free bytes %zu, brute force %zu", a.free_bytes(), freeb);
            CHECK(a.free_ranges() == runs(occ).size(), "free ranges %zu, brute force %zu (not coalesced)",
                  a.free_ranges(), runs(occ).size());
        }
        for (auto &b : live) a.free(b.first, b.second);
        CHECK(a.free_ranges() == 1 && a.free_bytes() == N, "after freeing all: %zu ranges, %zu bytes",
              a.free_ranges(), a.free_bytes());
    }
    size_t off;
    tr::Arena z;
    z.init(0);
    CHECK(!z.alloc(1, off), "an empty arena allocated");
    if (fails) return 1;
    std::printf("arena ok\n");
    return 0;
}
