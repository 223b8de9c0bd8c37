// Host checks of device helpers in bfm_device.hpp:
//  * the 9-site front (front_of) against a direct
//    enumeration with the lattice's site function: for every box, corner and
//    direction tested, the 4 (word, mask) pairs must hold exactly the bits of
//    the 9 front sites, and new_site must be the site of the moved corner
//    (exhaustive for small boxes, random corners for large);
//  * the wall and periodic-crossing tests on the moved corner's site;
//  * bond_key (per-axis differences modulo 8 on the low 9 site bits) against
//    the bond vector of the two corners, and the forbidden mask plan() builds in
//    that order against the reference-order table (copyBondSet's index) for
//    every vector in [-4,3]^3;
//  * site_coord (coordinates gathered from a site) against the coordinates;
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

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint32_t next_u32()
{
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return uint32_t(g_state >> 32);
}

template <bool kCubic>
bool check_one(const uint64_t box[3], const Geometry &g, const Tables &t, int x, int y, int z, int dir)
{
    const uint32_t site = host_site(g, t, x, y, z);
    const Front f = front_of<kCubic>(g, site, dir);
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
    const uint32_t *m = f.m;
    for (int q = 0; q < 4; ++q)
        for (int bit = 0; bit < 32; ++bit)
            if ((m[q] >> bit) & 1u) {
                got.insert((f.w[q] << 5) | uint32_t(bit));
                ++bits;
            }
    const int r1[3] = {x + d[0], y + d[1], z + d[2]};
    const uint32_t new_site = host_site(g, t, r1[0], r1[1], r1[2]);
    // the moved corner's coordinate on the move axis is L - 1 exactly when the
    // move leaves [0, L - 2] (the wall test) and 0 / L - 1 when an upward /
    // downward move crosses the periodic boundary
    const int L = int(box[a]), ra = a == 0 ? x : (a == 1 ? y : z);
    const uint32_t a1 = f.new_site & f.ma;
    // (a corner on a walled axis lies in [0, L - 2])
    const bool wall_hit = ra <= L - 2 && a1 == f.ma, want_wall = ra <= L - 2 && (ra + s < 0 || ra + s > L - 2);
    const bool crossed = (dir & 1) ? a1 == 0u : a1 == f.ma, want_cross = ra + s < 0 || ra + s >= L;
    if (got != want || bits != 9 || f.new_site != new_site || f.ma != g.dmask[a] || wall_hit != want_wall ||
        crossed != want_cross) {
        std::printf("FAIL box (%llu,%llu,%llu) r (%d,%d,%d) dir %d: %zu/%d bits, new %u/%u wall %d/%d cross %d/%d\n",
                    (unsigned long long)box[0], (unsigned long long)box[1], (unsigned long long)box[2], x, y, z,
                    dir, got.size(), bits, f.new_site, new_site, int(wall_hit), int(want_wall), int(crossed),
                    int(want_cross));
        return false;
    }
    // bonds from the moved corner to corners at every vector within reach
    // (the key depends only on the vector modulo 8)
    const uint32_t r1x = f.new_site & kLow9[0], r1y = f.new_site & kLow9[1], r1z = f.new_site & kLow9[2];
    for (int k = 0; k < 4; ++k) {
        const int v[3] = {int(next_u32() % 9) - 4, int(next_u32() % 9) - 4, int(next_u32() % 9) - 4};
        const uint32_t q = host_site(g, t, r1[0] + v[0], r1[1] + v[1], r1[2] + v[2]);
        if (bond_key(q, r1x, r1y, r1z) != bond_key_of(v[0], v[1], v[2])) {
            std::printf("FAIL bond_key box (%llu,%llu,%llu) r1 (%d,%d,%d) v (%d,%d,%d)\n",
                        (unsigned long long)box[0], (unsigned long long)box[1], (unsigned long long)box[2], r1[0],
                        r1[1], r1[2], v[0], v[1], v[2]);
            return false;
        }
    }
    const int r[3] = {x, y, z};
    for (int k = 0; k < 3; ++k)
        if (site_coord(site, g.dmask[k]) != uint32_t(r[k])) {
            std::printf("FAIL site_coord box (%llu,%llu,%llu) r (%d,%d,%d) axis %d\n", (unsigned long long)box[0],
                        (unsigned long long)box[1], (unsigned long long)box[2], x, y, z, k);
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
                if (!check_one<false>(box, g, t, x, y, z, dir))
                    return 1;
                if (g.cubic && !check_one<true>(box, g, t, x, y, z, dir))
                    return 1;
                ++cases;
            }
        }
    }
    {
        // plan() on one free monomer: the device bond mask in bond_key order
        HostSystem h;
        h.box[0] = h.box[1] = h.box[2] = 16;
        h.n = 1;
        h.xyz = {3, 4, 5};
        h.attribute = {0};
        h.nnbr = {0};
        h.nbr.assign(kMaxConnectivity, 0);
        standard_bond_set(h.forbidden);
        plan(h);
        int allowed = 0;
        for (int dx = -4; dx <= 3; ++dx)
            for (int dy = -4; dy <= 3; ++dy)
                for (int dz = -4; dz <= 3; ++dz) {
                    const uint32_t k = bond_key_of(dx, dy, dz);
                    const bool dev = (h.tables.bond[k >> 5] >> (k & 31u)) & 1u;
                    const bool ref = h.forbidden[bond_table_index(dx, dy, dz)] != 0;
                    if (dev != ref) {
                        std::printf("FAIL bond mask (%d,%d,%d): device %d reference %d\n", dx, dy, dz, int(dev),
                                    int(ref));
                        return 1;
                    }
                    allowed += !dev;
                    ++cases;
                }
        if (allowed != 108) {
            std::printf("FAIL bond mask: %d allowed vectors\n", allowed);
            return 1;
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
