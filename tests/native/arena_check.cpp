// Randomised check of the serving loop's range allocator (blitzen_amd/csrc/arena.h) against a brute-force model:
// every live block is inside the arena, aligned, and disjoint from every other; used() equals the sum of live
// rounded sizes; a block that the model says fits (some free gap of the rounded size) is granted (first fit);
// freeing everything coalesces back to one range. Prints "ok" or the first failure.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "../../blitzen_amd/csrc/arena.h"

struct Block { uint64_t off, n; };

static int fail(const char* what, uint64_t a = 0, uint64_t b = 0)
{
    std::printf("FAIL %s %llu %llu\n", what, (unsigned long long)a, (unsigned long long)b);
    return 1;
}

int main()
{
    std::mt19937_64 rng(12345);
    for (int round = 0; round < 50; ++round) {
        const uint64_t align = 1ull << (rng() % 9);          // 1 .. 256
        const uint64_t size = 1000 + rng() % 1000000;
        bz::Arena a(size, align);
        const uint64_t cap = size - size % align;
        std::vector<Block> live;
        for (int op = 0; op < 4000; ++op) {
            if (live.empty() || rng() % 3) {
                const uint64_t n = 1 + rng() % (size / 8);
                const uint64_t r = bz::Arena::round_up(n, align);
                // model: lowest-offset gap that holds r
                std::map<uint64_t, uint64_t> used;
                for (auto& b : live) used[b.off] = b.n;
                uint64_t cur = 0, first = ~0ull;
                for (auto& kv : used) {
                    if (kv.first - cur >= r) { first = cur; break; }
                    cur = kv.first + kv.second;
                }
                if (first == ~0ull && cap - cur >= r) first = cur;
                const uint64_t off = a.alloc(n);
                if (off != first) return fail("first-fit offset", off, first);
                if (off == bz::Arena::kFail) continue;
                if (off % align) return fail("alignment", off, align);
                live.push_back({ off, r });
            } else {
                const size_t k = rng() % live.size();
                a.release(live[k].off, live[k].n);
                live.erase(live.begin() + k);
            }
            uint64_t sum = 0;
            for (auto& b : live) {
                sum += b.n;
                if (b.off + b.n > cap) return fail("out of range", b.off, b.n);
            }
            if (sum != a.used()) return fail("used", sum, a.used());
        }
        for (auto& b : live) a.release(b.off, b.n);
        if (a.used() != 0 || a.free_ranges() != 1 || a.largest_free() != cap) return fail("coalesce", a.free_ranges());
        if (!a.fits_empty(cap) || a.fits_empty(cap + 1)) return fail("fits_empty", cap);
    }
    std::printf("ok\n");
    return 0;
}
