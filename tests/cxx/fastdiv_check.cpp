// Host check of pk_fastdiv.h: the multiplier quotient equals n / d for every divisor used as a
// cell-grid extent (1 .. 4096 and products of two of them up to 2^24) on edge and random n < 2^31.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../../pica_amd/csrc/pk_fastdiv.h"

int main()
{
    std::vector<uint32_t> divs;
    for (uint32_t d = 1; d <= 4096; d++) divs.push_back(d);
    for (uint32_t a : {3u, 7u, 96u, 100u, 255u, 256u, 257u, 1000u, 1024u, 4095u, 4096u})
        for (uint32_t b : {1u, 5u, 96u, 128u, 999u, 4096u}) divs.push_back(a * b);
    std::mt19937_64 rng(12345);
    long bad = 0, checked = 0;
    for (uint32_t d : divs) {
        const FastDiv f = make_fastdiv(d);
        std::vector<uint32_t> ns = {0u, 1u, d - 1, d, d + 1, 2 * d - 1, 0x7fffffffu, 0x7ffffffeu,
                                    (0x7fffffffu / d) * d, (0x7fffffffu / d) * d - 1};
        for (int k = 0; k < 2000; k++) ns.push_back((uint32_t)(rng() & 0x7fffffffu));
        for (int k = 0; k < 200; k++) {   // multiples of d and their neighbours
            const uint64_t q = rng() % (0x7fffffffu / d + 1);
            for (int64_t e = -1; e <= 1; e++) {
                const int64_t n = (int64_t)(q * d) + e;
                if (n >= 0 && n <= 0x7fffffff) ns.push_back((uint32_t)n);
            }
        }
        for (uint32_t n : ns) {
            checked++;
            if (n > 0x7fffffffu) continue;
            if (fastdiv(n, f) != n / d) {
                if (bad < 10) std::printf("mismatch n=%u d=%u: %u vs %u\n", n, d, fastdiv(n, f), n / d);
                bad++;
            }
        }
    }
    std::printf("checked %ld quotients, %ld wrong\n", checked, bad);
    return bad ? 1 : 0;
}
