// Randomised check of the request server's block allocator (fastpcc_amd/csrc/arena.hpp)
// against a byte-map model: blocks never overlap, never leave the range, the used
// count is exact, and releasing everything coalesces back to one free block.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "arena.hpp"

using fpcc::Arena;

static int fail(const char *what, long a, long b)
{
    std::printf("FAIL %s %ld %ld\n", what, a, b);
    return 1;
}

int main()
{
    std::mt19937_64 rng(12345);
    for (int trial = 0; trial < 200; ++trial) {
        const size_t size = (size_t)(1 + rng() % 4096) * Arena::kAlign + rng() % Arena::kAlign;
        Arena a(size);
        if (a.size() != size / Arena::kAlign * Arena::kAlign) return fail("size", (long)a.size(), (long)size);
        std::vector<char> owner(a.size() / Arena::kAlign, 0);
        struct Blk { size_t off, bytes; };
        std::vector<Blk> live;
        size_t used = 0;
        for (int step = 0; step < 2000; ++step) {
            if (!live.empty() && (rng() % 3 == 0 || step > 1500)) {
                const size_t k = rng() % live.size();
                const Blk b = live[k];
                live[k] = live.back();
                live.pop_back();
                const size_t n = Arena::round(b.bytes ? b.bytes : 1) / Arena::kAlign;
                for (size_t i = 0; i < n; ++i) owner[b.off / Arena::kAlign + i] = 0;
                a.release(b.off, b.bytes);
                used -= n * Arena::kAlign;
            } else {
                const size_t bytes = rng() % 4 == 0 ? 0 : (size_t)(rng() % (size / 8 + 1));
                size_t off = 0;
                const size_t largest = a.largest_free();
                const bool ok = a.alloc(bytes, off);
                const size_t need = Arena::round(bytes ? bytes : 1);
                if (ok != (need <= largest)) return fail("first fit", (long)need, (long)largest);
                if (!ok) continue;
                if (off % Arena::kAlign || off + need > a.size()) return fail("range", (long)off, (long)need);
                for (size_t i = 0; i < need / Arena::kAlign; ++i) {
                    if (owner[off / Arena::kAlign + i]) return fail("overlap", (long)off, (long)i);
                    owner[off / Arena::kAlign + i] = 1;
                }
                live.push_back({off, bytes});
                used += need;
            }
            if (a.used() != used) return fail("used", (long)a.used(), (long)used);
            if (a.free_bytes() + a.used() != a.size()) return fail("free", (long)a.free_bytes(), (long)a.size());
        }
        for (const Blk &b : live) a.release(b.off, b.bytes);
        if (a.used() != 0 || a.fragments() != (a.size() ? 1u : 0u) || a.largest_free() != a.size())
            return fail("coalesce", (long)a.fragments(), (long)a.largest_free());
        // fits() agrees with a sequence of allocations and leaves the arena untouched
        const size_t nb = a.size() / Arena::kAlign;
        const size_t req[3] = {nb / 2 * Arena::kAlign, nb / 4 * Arena::kAlign, nb / 4 * Arena::kAlign};
        if (!a.fits(req) || a.used() != 0) return fail("fits", (long)a.size(), (long)a.used());
        const size_t too[2] = {a.size() / 2 + Arena::kAlign, a.size() / 2 + Arena::kAlign};
        if (a.size() >= 4 * Arena::kAlign && a.fits(too)) return fail("fits too much", (long)a.size(), 0);
    }
    std::printf("ok\n");
    return 0;
}
