// Host checks of device helpers in bfm_device.hpp:
//  * the 9-site front arithmetic (front_sites) against a direct enumeration
//    with the lattice's site function: for every box, corner and direction
//    tested, the at most 4 (word, mask) pairs must hold exactly the bits of
//    the 9 front sites, and new_site / old_site must be the sites of the moved
//    and the old corner (exhaustive for small boxes, random corners for large);
//  * colour_at (one entry of the per-MCS colour permutation, computed without
//    the array) against the full shuffle colour_order.
// Prints "ok <cases>" and exits 0 on success.
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <set>

#include "host_system.hpp"

using namespace pscbfm;

namespace {

struct HostDepositTable {
    const Tables *t;
    uint32_t operator()(int a, uint32_t v) const { return t->dep[a][v & 31u] | t->dep[a][32 + (v >> 5)]; }
};

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint32_t next_u32()
{
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return uint32_t(g_state >> 32);
}

template <class Dep>
bool check_one(const uint64_t box[3], const Geometry &g, const Tables &t, const Dep &dep, int x, int y, int z, int dir)
{
    const int4 r = make_int4(x, y, z, 0);
    const Front f = front_sites(g, dep, r, dir);
    const int a = dir >> 1, s = (dir & 1) ? 1 : -1;
    int d[3] = {0, 0, 0};
    d[a] = s;
    std::set<uint32_t> want;
    const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
    for (int j = -1; j <= 1; ++j)
        for (int k = -1; k <= 1; ++k) {
            int p[3] = {x, y, z};
            p[a] += 2 * s;
            p[b] += j;
            p[c] += k;
            want.insert(host_site(g, t, p[0], p[1], p[2]));
        }
    std::set<uint32_t> got;
    int bits = 0;
    for (int q = 0; q < 4; ++q)
        for (int bit = 0; bit < 32; ++bit)
            if ((f.m[q] >> bit) & 1u) {
                got.insert((f.w[q] << 5) | uint32_t(bit));
                ++bits;
            }
    const uint32_t new_site = host_site(g, t, x + d[0], y + d[1], z + d[2]);
    const uint32_t old_site = host_site(g, t, x, y, z);
    if (got != want || bits != 9 || f.new_site != new_site || f.old_site != old_site) {
        std::printf("FAIL box (%llu,%llu,%llu) r (%d,%d,%d) dir %d: %zu/%d bits, new %u/%u old %u/%u\n",
                    (unsigned long long)box[0], (unsigned long long)box[1], (unsigned long long)box[2], x, y, z,
                    dir, got.size(), bits, f.new_site, new_site, f.old_site, old_site);
        return false;
    }
    return true;
}

}  // namespace

// "colour" mode: the same graph format as oracle/ref_coloring_driver.cpp on
// stdin ("uniform n", then n lines "k j_1 ... j_k"), the product's colouring
// printed as "n_colours c_0 ... c_{n-1}" per graph.
static int colour_graphs()
{
    int uniform = 0;
    unsigned long n = 0;
    while (std::scanf("%d %lu", &uniform, &n) == 2) {
        std::vector<uint8_t> nnbr(n, 0);
        std::vector<uint32_t> nbr(n * kMaxConnectivity, 0);
        for (unsigned long i = 0; i < n; ++i) {
            unsigned k = 0;
            if (std::scanf("%u", &k) != 1 || k > kMaxConnectivity)
                return 1;
            nnbr[i] = uint8_t(k);
            for (unsigned q = 0; q < k; ++q)
                if (std::scanf("%u", &nbr[i * kMaxConnectivity + q]) != 1)
                    return 1;
        }
        std::vector<uint32_t> colour;
        const uint32_t k = graph_colouring(uint32_t(n), nnbr, nbr, uniform != 0, colour);
        std::printf("%u", k);
        for (uint32_t c : colour)
            std::printf(" %u", c);
        std::printf("\n");
    }
    return 0;
}

int main(int argc, char **argv)
{
    if (argc > 1 && std::strcmp(argv[1], "colour") == 0)
        return colour_graphs();
    const uint64_t boxes[][3] = {{8, 8, 8},    {16, 16, 16}, {8, 16, 32},    {32, 8, 16},  {16, 64, 8},
                                 {64, 64, 64}, {512, 512, 512}, {1024, 8, 8}, {8, 1024, 64}, {128, 256, 1024}};
    const bool periodic[3] = {true, true, true};
    long cases = 0;
    for (const auto &box : boxes) {
        Geometry g{};
        Tables t{};
        uint64_t words = 0;
        make_geometry(box, periodic, g, t, words);
        const HostDepositTable table{&t};
        const DepositCubic cubic{};
        const uint64_t sites = box[0] * box[1] * box[2];
        const bool exhaustive = sites <= 16384;
        const long samples = exhaustive ? long(sites) : 20000;
        for (long n = 0; n < samples; ++n) {
            int x, y, z;
            if (exhaustive) {
                x = int(n % box[0]);
                y = int((n / box[0]) % box[1]);
                z = int(n / (box[0] * box[1]));
            } else {
                x = int(next_u32() % box[0]);
                y = int(next_u32() % box[1]);
                z = int(next_u32() % box[2]);
            }
            for (int dir = 0; dir < 6; ++dir) {
                if (!check_one(box, g, t, table, x, y, z, dir))
                    return 1;
                if (g.cubic && !check_one(box, g, t, cubic, x, y, z, dir))
                    return 1;
                ++cases;
            }
        }
    }
    for (uint32_t n = 1; n <= 40; ++n)
        for (uint64_t age = 0; age < 200; ++age) {
            const uint64_t seed = 0x5EED000000000000ull + n * 7919u + age;
            uint32_t order[40];
            colour_order(seed, age, n, order);
            for (uint32_t k = 0; k < n; ++k) {
                if (colour_at(seed, age, n, k) != order[k]) {
                    std::printf("FAIL colour_at n %u age %llu k %u: %u vs %u\n", n, (unsigned long long)age, k,
                                colour_at(seed, age, n, k), order[k]);
                    return 1;
                }
                ++cases;
            }
        }
    std::printf("ok %ld\n", cases);
    return 0;
}
