// host_system.cpp -- validation, colouring, Morton tables and layout.
#include "host_system.hpp"

#include <algorithm>
#include <cstring>
#include <sstream>

namespace pscbfm {

HostSystem::HostSystem() { reset(); }

void HostSystem::reset()
{
    box[0] = box[1] = box[2] = 0;
    periodic[0] = periodic[1] = periodic[2] = true;
    n = 0;
    steps_between_sortings = 5000;
    xyz.clear();
    attribute.clear();
    nnbr.clear();
    nbr.clear();
    // Until copyBondSet says otherwise every vector is forbidden; the
    // LeMonADE adapter sends all 512 entries (GPUScBFM_AB_Type.h:168-175).
    std::memset(forbidden, 1, sizeof forbidden);
    geom = Geometry{};
    tables = Tables{};
    lattice_words = 0;
    colour.clear();
    n_colours = 0;
    seg_begin.clear();
    seg_count.clear();
    n_pad = 0;
    max_segment = 0;
    to_sorted.clear();
    ppos_sorted.clear();
    ids_sorted.clear();
    image_sorted.clear();
    nbr_sorted.clear();
    lattice.clear();
    mean_bonds = 0.0;
}

int count_allowed_bonds(const uint8_t forbidden[512])
{
    int k = 0;
    for (int i = 0; i < 512; ++i)
        k += forbidden[i] == 0;
    return k;
}

void standard_bond_set(uint8_t forbidden[512])
{
    std::memset(forbidden, 1, 512);
    for (int dx = -3; dx <= 3; ++dx)
        for (int dy = -3; dy <= 3; ++dy)
            for (int dz = -3; dz <= 3; ++dz) {
                int a[3] = {std::abs(dx), std::abs(dy), std::abs(dz)};
                std::sort(a, a + 3);
                const int key = a[0] * 100 + a[1] * 10 + a[2];
                // sorted magnitudes of the six generating vectors
                if (key == 2 || key == 12 || key == 112 || key == 122 || key == 3 || key == 13)
                    forbidden[bond_table_index(dx, dy, dz)] = 0;
            }
}

static int log2_exact(uint64_t v)
{
    int k = 0;
    while ((uint64_t(1) << k) < v)
        ++k;
    return (uint64_t(1) << k) == v ? k : -1;
}

void make_geometry(const uint64_t box[3], const bool periodic[3], Geometry &g, Tables &t,
                   uint64_t &lattice_words)
{
    int k[3];
    for (int a = 0; a < 3; ++a) {
        k[a] = log2_exact(box[a]);
        if (k[a] < 3 || box[a] > kMaxBox) {
            std::ostringstream msg;
            msg << "box edge " << box[a] << " must be a power of two in [8," << kMaxBox << "]";
            throw std::runtime_error(msg.str());
        }
    }
    int bitpos[3][10];
    int next = 0;
    for (int lvl = 0; lvl < 10; ++lvl)
        for (int a = 0; a < 3; ++a)
            if (lvl < k[a])
                bitpos[a][lvl] = next++;
    auto dep = [&](int a, uint32_t v) {
        uint32_t d = 0;
        for (int lvl = 0; lvl < k[a]; ++lvl)
            if ((v >> lvl) & 1u)
                d |= 1u << bitpos[a][lvl];
        return d;
    };
    g.periodic = 0;
    g.cubic = (box[0] == box[1] && box[1] == box[2]) ? 1u : 0u;
    for (int a = 0; a < 3; ++a) {
        g.mask[a] = uint32_t(box[a] - 1);
        g.dmask[a] = dep(a, g.mask[a]);
        g.wall_hi[a] = int32_t(box[a]) - 2;
        if (periodic[a])
            g.periodic |= 1u << a;
        for (uint32_t j = 0; j < 32; ++j) {
            t.dep[a][j] = dep(a, j & g.mask[a]);
            t.dep[a][32 + j] = dep(a, (j << 5) & g.mask[a]);
        }
    }
    const uint64_t sites = uint64_t(1) << next;
    lattice_words = std::max<uint64_t>(1, sites / 32);
}

uint32_t host_site(const Geometry &g, const Tables &t, int32_t x, int32_t y, int32_t z)
{
    const int32_t v[3] = {x, y, z};
    uint32_t s = 0;
    for (int a = 0; a < 3; ++a) {
        const uint32_t u = uint32_t(v[a]) & g.mask[a];
        s |= t.dep[a][u & 31u] | t.dep[a][32 + (u >> 5)];
    }
    return s;
}

void colouring_adjacency(uint32_t n, const std::vector<uint8_t> &nnbr, const std::vector<uint32_t> &nbr,
                         std::vector<uint32_t> &off, std::vector<uint32_t> &adj)
{
    // links i -> j whose reverse j -> i is not listed: j gets i appended
    auto lists = [&](uint32_t a, uint32_t b) {
        for (uint32_t k = 0; k < nnbr[a]; ++k)
            if (nbr[size_t(a) * kMaxConnectivity + k] == b)
                return true;
        return false;
    };
    std::vector<uint32_t> extra_count(n, 0);
    std::vector<std::pair<uint32_t, uint32_t>> extra;   // (monomer, appended partner), by partner id
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = 0; k < nnbr[i]; ++k) {
            const uint32_t j = nbr[size_t(i) * kMaxConnectivity + k];
            if (!lists(j, i)) {
                extra.emplace_back(j, i);
                ++extra_count[j];
            }
        }
    off.assign(size_t(n) + 1, 0);
    for (uint32_t i = 0; i < n; ++i)
        off[i + 1] = off[i] + nnbr[i] + extra_count[i];
    adj.assign(off[n], 0);
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = 0; k < nnbr[i]; ++k)
            adj[fill[i]++] = nbr[size_t(i) * kMaxConnectivity + k];
    for (const auto &e : extra)
        adj[fill[e.first]++] = e.second;
}

uint32_t graph_colouring(uint32_t n, const std::vector<uint8_t> &nnbr, const std::vector<uint32_t> &nbr,
                         bool uniform, std::vector<uint32_t> &colour)
{
    std::vector<uint32_t> off, adj;
    colouring_adjacency(n, nnbr, nbr, off, adj);
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    constexpr uint32_t kColourLimit = 255;   // graphColoring.tpp:114: at most 255 colours
    colour.assign(n, kNone);
    std::vector<uint64_t> used_by;           // monomers per colour

    // Walk of graphColoring.tpp:131-294.  Each connected polymer is coloured
    // from its lowest id by following links: a monomer takes the lowest
    // colour none of its partners has; the walk goes on to its first
    // uncoloured partner and remembers the monomer while it has two or more;
    // at a dead end it resumes from the most recent remembered monomer that
    // still has an uncoloured partner.
    auto open_partners = [&](uint32_t v, uint32_t &first) {
        uint32_t open = 0;
        first = kNone;
        for (uint32_t k = off[v]; k < off[v + 1]; ++k)
            if (colour[adj[k]] == kNone) {
                if (first == kNone)
                    first = adj[k];
                ++open;
            }
        return open;
    };
    auto lowest_free = [&](uint32_t v) {
        uint64_t taken[4] = {0, 0, 0, 0};
        for (uint32_t k = off[v]; k < off[v + 1]; ++k) {
            const uint32_t c = colour[adj[k]];
            if (c != kNone)
                taken[c >> 6] |= uint64_t(1) << (c & 63u);
        }
        uint32_t c = 0;
        while (c < kColourLimit && ((taken[c >> 6] >> (c & 63u)) & 1u))
            ++c;
        if (c == kColourLimit)
            throw std::runtime_error("graph colouring needs more than 255 colours");
        return c;
    };
    std::vector<uint32_t> remembered;
    for (uint32_t root = 0; root < n; ++root) {
        if (colour[root] != kNone)
            continue;
        remembered.clear();
        uint32_t v = root;
        while (v != kNone) {
            uint32_t next = kNone;
            const uint32_t open = open_partners(v, next);
            const uint32_t c = lowest_free(v);
            colour[v] = c;
            if (c >= used_by.size())
                used_by.resize(c + 1, 0);
            ++used_by[c];
            if (open > 1)
                remembered.push_back(v);
            v = next;
            while (v == kNone && !remembered.empty()) {
                const uint32_t u = remembered.back();
                if (open_partners(u, v) < 2)
                    remembered.pop_back();
            }
        }
    }
    const uint32_t k_colours = uint32_t(used_by.size());

    // Uniform rebalance of graphColoring.tpp:323-418 (switched on by
    // initializeSpeciesSorting, UpdaterGPUScBFM_AB_Type.cu:1565): in id order,
    // a monomer of a colour used more than ceil(n / k) times tries the colours
    // used less than that, rarest first, and takes each one none of its
    // partners has; the ranking is re-sorted (stably) after every change, and
    // the monomer goes on through the ranking after a change, as the
    // reference's loop does.
    if (uniform && k_colours > 1) {
        const uint64_t target = (uint64_t(n) + k_colours - 1) / k_colours;
        std::vector<std::pair<uint32_t, uint64_t>> rank(k_colours);   // (colour, monomers)
        for (uint32_t c = 0; c < k_colours; ++c)
            rank[c] = {c, used_by[c]};
        auto by_count = [](const std::pair<uint32_t, uint64_t> &a, const std::pair<uint32_t, uint64_t> &b) {
            return a.second < b.second;
        };
        std::stable_sort(rank.begin(), rank.end(), by_count);
        for (uint32_t v = 0; v < n; ++v) {
            if (used_by[colour[v]] <= target)
                continue;
            for (size_t r = 0; r < rank.size() && rank[r].second < target; ++r) {
                const uint32_t want = rank[r].first;
                bool clash = false;
                for (uint32_t k = off[v]; k < off[v + 1] && !clash; ++k)
                    clash = colour[adj[k]] == want;
                if (clash)
                    continue;
                const uint32_t had = colour[v];
                --used_by[had];
                ++used_by[want];
                for (auto &e : rank) {
                    if (e.first == had)
                        --e.second;
                    if (e.first == want)
                        ++e.second;
                }
                std::stable_sort(rank.begin(), rank.end(), by_count);
                colour[v] = want;
            }
        }
    }
    return k_colours;
}

void plan(HostSystem &s)
{
    if (s.n == 0)
        throw std::runtime_error("setNrOfAllMonomers must be called with n > 0 before initialize");
    if (s.box[0] == 0 || s.box[1] == 0 || s.box[2] == 0)
        throw std::runtime_error("setLatticeSize must be called before initialize");
    const int allowed = count_allowed_bonds(s.forbidden);
    if (allowed != 108) {
        std::ostringstream msg;
        msg << "Wrong bond-set! Expected 108 allowed bonds, but got " << allowed;
        throw std::runtime_error(msg.str());
    }
    make_geometry(s.box, s.periodic, s.geom, s.tables, s.lattice_words);
    if (s.n > kIdMask)
        throw std::runtime_error("too many monomers for the 28-bit monomer id");
    // the forbidden mask in the kernels' order (bond_key: the three 3-bit
    // differences interleaved as in a site index)
    for (int i = 0; i < 512; ++i)
        if (s.forbidden[i]) {
            const uint32_t k = bond_key_of(i & 7, (i >> 3) & 7, (i >> 6) & 7);
            s.tables.bond[k >> 5] |= 1u << (k & 31);
        }

    // walls
    for (uint32_t i = 0; i < s.n; ++i)
        for (int a = 0; a < 3; ++a)
            if (!s.periodic[a]) {
                const int32_t v = s.xyz[3 * size_t(i) + a];
                if (v < 0 || v > int32_t(s.box[a]) - 2) {
                    std::ostringstream msg;
                    msg << "monomer " << i << " lies outside the non-periodic walls on axis " << a;
                    throw std::runtime_error(msg.str());
                }
            }

    // overlap check on the full cubes (checkSystem, UpdaterGPUScBFM_AB_Type.cu:
    // 2808-2851), then the device lattice: one bit per monomer at its lower
    // corner (the reference's mLatticeOut, initializeLattices :2142-2176)
    {
        std::vector<uint32_t> cubes(s.lattice_words, 0u);
        for (uint32_t i = 0; i < s.n; ++i) {
            const int32_t *r = &s.xyz[3 * size_t(i)];
            for (int dz = 0; dz < 2; ++dz)
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        const uint32_t site = host_site(s.geom, s.tables, r[0] + dx, r[1] + dy, r[2] + dz);
                        uint32_t &w = cubes[site >> 5];
                        const uint32_t bit = 1u << (site & 31u);
                        if (w & bit) {
                            std::ostringstream msg;
                            msg << "Occupation count in lattice is wrong: monomer " << i << " at (" << r[0] << ","
                                << r[1] << "," << r[2] << ") overlaps another monomer";
                            throw std::runtime_error(msg.str());
                        }
                        w |= bit;
                    }
        }
    }
    s.lattice.assign(s.lattice_words, 0u);
    for (uint32_t i = 0; i < s.n; ++i) {
        const int32_t *r = &s.xyz[3 * size_t(i)];
        const uint32_t site = host_site(s.geom, s.tables, r[0], r[1], r[2]);
        s.lattice[site >> 5] |= 1u << (site & 31u);
    }

    // bonds (UpdaterGPUScBFM_AB_Type.cu:2853-2899)
    uint64_t n_links = 0;
    for (uint32_t i = 0; i < s.n; ++i)
        for (uint32_t j = 0; j < s.nnbr[i]; ++j) {
            const uint32_t p = s.nbr[size_t(i) * kMaxConnectivity + j];
            if (p >= s.n) {
                std::ostringstream msg;
                msg << "monomer " << i << " is linked to monomer " << p << " which does not exist";
                throw std::invalid_argument(msg.str());
            }
            int32_t d[3];
            for (int a = 0; a < 3; ++a) {
                d[a] = s.xyz[3 * size_t(p) + a] - s.xyz[3 * size_t(i) + a];
                if (s.periodic[a]) {
                    // minimum image: a periodic network cannot be unwrapped, so a
                    // bond is its difference modulo the box (the kernels' (d & 7)
                    // index is the same modulo any box that is a multiple of 8)
                    const int64_t L = int64_t(s.box[a]);
                    int64_t m = ((int64_t(d[a]) % L) + L) % L;
                    if (m > L / 2)
                        m -= L;
                    d[a] = int32_t(m);
                }
            }
            const bool too_long = std::abs(d[0]) > 3 || std::abs(d[1]) > 3 || std::abs(d[2]) > 3;
            if (too_long || s.forbidden[bond_table_index(d[0], d[1], d[2])]) {
                std::ostringstream msg;
                msg << "Invalid bond (" << d[0] << "," << d[1] << "," << d[2] << ") between monomer " << i
                    << " and monomer " << p;
                throw std::runtime_error(msg.str());
            }
            ++n_links;
        }
    s.mean_bonds = double(n_links) / double(s.n);

    // colouring and colour-sorted layout (initializeSpeciesSorting, :1561-1686)
    s.n_colours = graph_colouring(s.n, s.nnbr, s.nbr, true, s.colour);
    s.seg_count.assign(s.n_colours, 0);
    for (uint32_t i = 0; i < s.n; ++i)
        ++s.seg_count[s.colour[i]];
    s.seg_begin.assign(s.n_colours, 0);
    uint64_t offset = 0;
    s.max_segment = 0;
    for (uint32_t c = 0; c < s.n_colours; ++c) {
        s.seg_begin[c] = uint32_t(offset);
        offset += (uint64_t(s.seg_count[c]) + kSegmentAlign - 1) / kSegmentAlign * kSegmentAlign;
        s.max_segment = std::max(s.max_segment, s.seg_count[c]);
    }
    if (offset > kSortedIndexMask)
        throw std::runtime_error("padded system exceeds the 29-bit sorted index");
    s.n_pad = uint32_t(std::max<uint64_t>(offset, kSegmentAlign));
    s.to_sorted.assign(s.n, 0);
    if (s.spatial_order) {
        // within a colour, monomers follow the lattice's Morton order of their
        // lower corners (ties by id): a wavefront then covers a compact patch
        // of space, so its lattice words, claims and partners share cache
        // lines (the reference's doSpatialSorting, :1836-1935, sorts each
        // species by the linearized position the same way)
        std::vector<uint64_t> key(s.n);
        for (uint32_t i = 0; i < s.n; ++i) {
            const int32_t *r = &s.xyz[3 * size_t(i)];
            key[i] = (uint64_t(host_site(s.geom, s.tables, r[0], r[1], r[2])) << 32) | i;
        }
        std::vector<std::vector<uint64_t>> per(s.n_colours);
        for (uint32_t c = 0; c < s.n_colours; ++c)
            per[c].reserve(s.seg_count[c]);
        for (uint32_t i = 0; i < s.n; ++i)
            per[s.colour[i]].push_back(key[i]);
        for (uint32_t c = 0; c < s.n_colours; ++c) {
            std::sort(per[c].begin(), per[c].end());
            for (uint32_t k = 0; k < per[c].size(); ++k)
                s.to_sorted[uint32_t(per[c][k])] = s.seg_begin[c] + k;
        }
    } else {
        std::vector<uint32_t> next(s.seg_begin);
        for (uint32_t i = 0; i < s.n; ++i)
            s.to_sorted[i] = next[s.colour[i]]++;
    }
    s.ppos_sorted.assign(s.n_pad, 0u);
    s.ids_sorted.assign(s.n_pad, 0xFFFFFFFFu);
    s.image_sorted.assign(3 * size_t(s.n_pad), 0);
    // slot-major, only the columns in use (the device table has all
    // kMaxConnectivity; the upload copies these and leaves the rest zero);
    // past its own partners a slot holds its own index, which the check
    // kernel reads unconditionally in place of a missing partner
    uint32_t cols = 0;
    for (const uint8_t c : s.nnbr)
        cols = std::max<uint32_t>(cols, c);
    s.nbr_sorted.assign(size_t(cols) * s.n_pad, 0);
    for (uint32_t i = 0; i < s.n; ++i) {
        const uint32_t t = s.to_sorted[i];
        uint32_t w[3];
        for (int a = 0; a < 3; ++a) {
            const int64_t v = s.xyz[3 * size_t(i) + a];
            const int64_t L = int64_t(s.box[a]);
            const int64_t img = (v >= 0 ? v / L : -((-v + L - 1) / L));   // floor(v / L)
            w[a] = uint32_t(v - img * L);
            s.image_sorted[size_t(a) * s.n_pad + t] = int32_t(img);
        }
        s.ppos_sorted[t] = host_site(s.geom, s.tables, int32_t(w[0]), int32_t(w[1]), int32_t(w[2]));
        s.ids_sorted[t] = i | (uint32_t(s.nnbr[i]) << kIdBits);
        for (uint32_t j = 0; j < s.nnbr[i]; ++j)
            s.nbr_sorted[size_t(j) * s.n_pad + t] = s.to_sorted[s.nbr[size_t(i) * kMaxConnectivity + j]];
        for (uint32_t j = s.nnbr[i]; j < cols; ++j)
            s.nbr_sorted[size_t(j) * s.n_pad + t] = t;
    }
}

void release_upload_staging(HostSystem &s)
{
    std::vector<uint32_t>().swap(s.to_sorted);
    std::vector<uint32_t>().swap(s.ppos_sorted);
    std::vector<uint32_t>().swap(s.ids_sorted);
    std::vector<int32_t>().swap(s.image_sorted);
    std::vector<uint32_t>().swap(s.nbr_sorted);
    std::vector<uint32_t>().swap(s.lattice);
}

void unpack_positions(HostSystem &s)
{
    for (uint32_t i = 0; i < s.n; ++i) {
        const uint32_t t = s.to_sorted[i];
        const uint32_t p = s.ppos_sorted[t];
        for (int a = 0; a < 3; ++a) {
            const int64_t w = site_coord(p, s.geom.dmask[a]);
            const int64_t img = s.image_sorted[size_t(a) * s.n_pad + t];
            s.xyz[3 * size_t(i) + a] = int32_t(w + img * int64_t(s.box[a]));
        }
    }
}

}  // namespace pscbfm
