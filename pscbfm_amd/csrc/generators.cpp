// generators.cpp -- seeded synthetic initial states (see pscbfm_amd_systems.h).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pscbfm_amd.h"
#include "../../include/pscbfm_amd_systems.h"

namespace pscbfm {
void set_last_error(const std::string &msg);
}

namespace {

int gen_fail(int code, const std::string &msg)
{
    pscbfm::set_last_error(msg);
    return code;
}

// splitmix64: a small, well-mixed generator for host-side construction.
struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    uint32_t below(uint32_t n) { return uint32_t(next() % n); }
};

int ilog2(uint64_t v)
{
    int k = 0;
    while ((uint64_t(1) << k) < v)
        ++k;
    return k;
}

bool pow2(uint64_t v) { return v && !(v & (v - 1)); }

const int kSteps[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};

// Self-avoiding walk of `len` cells inside a block of dims bd (cells),
// depth-first with random neighbour order and backtracking.  Returns false
// when the step budget is exhausted.
bool block_walk(const int bd[3], uint32_t len, SplitMix &rng, std::vector<int> &cells, uint64_t budget)
{
    const int nb = bd[0] * bd[1] * bd[2];
    std::vector<uint8_t> used(nb, 0);
    struct Frame {
        int cell;
        uint8_t order[6];
        uint8_t tried;
    };
    std::vector<Frame> stack;
    auto push = [&](int cell) {
        Frame f;
        f.cell = cell;
        for (int k = 0; k < 6; ++k)
            f.order[k] = uint8_t(k);
        for (int k = 5; k > 0; --k)
            std::swap(f.order[k], f.order[rng.below(k + 1)]);
        f.tried = 0;
        stack.push_back(f);
        used[cell] = 1;
    };
    push(int(rng.below(nb)));
    uint64_t steps = 0;
    while (stack.size() < len) {
        if (++steps > budget)
            return false;
        Frame &f = stack.back();
        if (f.tried == 6) {
            used[f.cell] = 0;
            stack.pop_back();
            if (stack.empty())
                push(int(rng.below(nb)));
            continue;
        }
        const int d = f.order[f.tried++];
        int c[3] = {f.cell % bd[0], (f.cell / bd[0]) % bd[1], f.cell / (bd[0] * bd[1])};
        c[0] += kSteps[d][0];
        c[1] += kSteps[d][1];
        c[2] += kSteps[d][2];
        if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= bd[0] || c[1] >= bd[1] || c[2] >= bd[2])
            continue;
        const int nc = c[0] + bd[0] * (c[1] + bd[1] * c[2]);
        if (used[nc])
            continue;
        push(nc);
    }
    cells.resize(len);
    for (uint32_t k = 0; k < len; ++k)
        cells[k] = stack[k].cell;
    return true;
}

}  // namespace

extern "C" int bfm_gen_linear_melt(uint32_t box, uint32_t n_chains, uint32_t chain_len, uint64_t seed, int32_t *xyz,
                                   uint32_t *bonds)
{
    if (!pow2(box) || box < 8 || box > 1024 || chain_len == 0 || n_chains == 0 || !xyz ||
        (chain_len > 1 && !bonds))
        return gen_fail(BFM_E_INVALID_ARGUMENT, "linear melt: box must be a power of two in [8,1024], "
                                                "chains and length positive, arrays non-null");
    const int lg = ilog2(box / 2);                       // coarse cells per axis = 2^lg
    const int lb = ilog2(2ull * chain_len);              // block holds 2^lb coarse cells
    if (lb > 3 * lg)
        return gen_fail(BFM_E_INVALID_ARGUMENT, "linear melt: a chain does not fit into the box");
    int kb[3] = {0, 0, 0};                               // block edge exponents, round-robin
    for (int k = 0; k < lb; ++k)
        ++kb[k % 3];
    for (int a = 0; a < 3; ++a)
        if (kb[a] > lg)
            return gen_fail(BFM_E_INVALID_ARGUMENT, "linear melt: a chain does not fit into the box");
    const int bd[3] = {1 << kb[0], 1 << kb[1], 1 << kb[2]};
    const uint32_t nbx = (box / 2) >> kb[0], nby = (box / 2) >> kb[1], nbz = (box / 2) >> kb[2];
    const uint64_t n_blocks = uint64_t(nbx) * nby * nbz;
    if (n_chains > n_blocks)
        return gen_fail(BFM_E_INVALID_ARGUMENT, "linear melt: more chains than blocks of 2*chain_len even-sublattice "
                                                "cells (volume fraction above 0.5)");

    // choose the blocks, then visit them in increasing order for locality
    SplitMix pick(seed ^ 0xB10C5EEDull);
    std::vector<uint32_t> blocks(n_blocks);
    for (uint32_t b = 0; b < n_blocks; ++b)
        blocks[b] = b;
    if (n_chains < n_blocks) {
        for (uint32_t k = 0; k < n_chains; ++k)
            std::swap(blocks[k], blocks[k + pick.below(uint32_t(n_blocks - k))]);
        blocks.resize(n_chains);
        std::sort(blocks.begin(), blocks.end());
    }

    std::vector<int> cells;
    for (uint32_t ch = 0; ch < n_chains; ++ch) {
        const uint32_t b = blocks[ch];
        const uint32_t bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
        SplitMix rng(seed * 0x9E3779B97F4A7C15ull + ch + 1);
        bool ok = false;
        for (int attempt = 0; attempt < 64 && !ok; ++attempt)
            ok = block_walk(bd, chain_len, rng, cells, 200ull * chain_len + 10000);
        if (!ok)
            return gen_fail(BFM_E_RUNTIME, "linear melt: self-avoiding walk growth failed");
        for (uint32_t k = 0; k < chain_len; ++k) {
            const int c = cells[k];
            const uint64_t id = uint64_t(ch) * chain_len + k;
            xyz[3 * id + 0] = int32_t(2 * (bx * bd[0] + c % bd[0]));
            xyz[3 * id + 1] = int32_t(2 * (by * bd[1] + (c / bd[0]) % bd[1]));
            xyz[3 * id + 2] = int32_t(2 * (bz * bd[2] + c / (bd[0] * bd[1])));
            if (k + 1 < chain_len) {
                const uint64_t e = uint64_t(ch) * (chain_len - 1) + k;
                bonds[2 * e] = uint32_t(id);
                bonds[2 * e + 1] = uint32_t(id + 1);
            }
        }
    }
    pscbfm::set_last_error("");
    return BFM_OK;
}

// ---- end-linked diamond network --------------------------------------------

namespace {

struct Link {
    uint32_t u, v;      // crosslinker ids
    float mid[3];       // midpoint (coarse cells, may lie outside [0,G))
};

struct Router {
    int G;
    uint32_t steps;                  // bonds per strand = strand_len + 1
    std::vector<int32_t> occ;        // -1 free, >= 0 owner (crosslinker id or 1<<30 | link id)
    std::vector<std::array<int, 3>> nodes;
    std::vector<Link> links;
    std::vector<std::vector<int>> paths;   // per link: cells (packed) strictly between u and v
    SplitMix rng;

    Router(int g, uint32_t st, uint64_t seed) : G(g), steps(st), occ(size_t(g) * g * g, -1), rng(seed) {}

    int pack(int x, int y, int z) const { return (z * G + y) * G + x; }
    int wrap(int v) const { return ((v % G) + G) % G; }
    int wd(int a, int b) const   // minimum-image b - a
    {
        int d = wrap(b - a);
        return d <= G / 2 ? d : d - G;
    }
    float fd(float a, float b) const
    {
        float d = std::fmod(b - a, float(G));
        if (d < 0)
            d += G;
        return d <= G / 2 ? d : G - d;
    }

    bool route(uint32_t li, uint64_t budget)
    {
        const Link &L = links[li];
        const auto &pu = nodes[L.u], &pv = nodes[L.v];
        struct Frame {
            int c[3];
            uint8_t cand[6];
            uint8_t ncand, next;
        };
        std::vector<Frame> st;
        st.reserve(steps + 1);
        std::vector<int> used;
        auto dist_to_v = [&](const int c[3]) {
            return std::abs(wd(c[0], pv[0])) + std::abs(wd(c[1], pv[1])) + std::abs(wd(c[2], pv[2]));
        };
        auto expand = [&](Frame &f, uint32_t rem) {
            // rem = steps still to take from f.c
            f.ncand = 0;
            f.next = 0;
            float key[6];
            for (int d = 0; d < 6; ++d) {
                int n[3] = {wrap(f.c[0] + kSteps[d][0]), wrap(f.c[1] + kSteps[d][1]), wrap(f.c[2] + kSteps[d][2])};
                if (rem == 1) {
                    if (n[0] == pv[0] && n[1] == pv[1] && n[2] == pv[2]) {
                        f.cand[f.ncand] = uint8_t(d);
                        key[f.ncand++] = 0.f;
                    }
                    continue;
                }
                if (occ[pack(n[0], n[1], n[2])] != -1)
                    continue;
                if (uint32_t(dist_to_v(n)) > rem - 1)
                    continue;
                const float off = fd(L.mid[0], float(n[0])) + fd(L.mid[1], float(n[1])) + fd(L.mid[2], float(n[2]));
                f.cand[f.ncand] = uint8_t(d);
                key[f.ncand++] = off + float(rng.next() >> 40) * (2.0f / float(1 << 24));
            }
            for (int a = 1; a < f.ncand; ++a)   // insertion sort by key
                for (int b = a; b > 0 && key[b] < key[b - 1]; --b) {
                    std::swap(key[b], key[b - 1]);
                    std::swap(f.cand[b], f.cand[b - 1]);
                }
        };
        Frame f0;
        f0.c[0] = pu[0];
        f0.c[1] = pu[1];
        f0.c[2] = pu[2];
        expand(f0, steps);
        st.push_back(f0);
        uint64_t work = 0;
        while (!st.empty()) {
            if (++work > budget)
                break;
            Frame &f = st.back();
            const uint32_t depth = uint32_t(st.size()) - 1;  // steps taken so far
            if (f.next == f.ncand) {
                if (depth > 0)
                    occ[pack(f.c[0], f.c[1], f.c[2])] = -1;
                st.pop_back();
                continue;
            }
            const int d = f.cand[f.next++];
            Frame n;
            n.c[0] = wrap(f.c[0] + kSteps[d][0]);
            n.c[1] = wrap(f.c[1] + kSteps[d][1]);
            n.c[2] = wrap(f.c[2] + kSteps[d][2]);
            if (depth + 1 == steps) {
                // reached v: record the path (frames 1..steps-1 are the strand)
                std::vector<int> cells;
                cells.reserve(steps - 1);
                for (size_t k = 1; k < st.size(); ++k)
                    cells.push_back(pack(st[k].c[0], st[k].c[1], st[k].c[2]));
                for (int c : cells)
                    occ[c] = (1 << 30) | int(li);
                paths[li] = std::move(cells);
                return true;
            }
            occ[pack(n.c[0], n.c[1], n.c[2])] = -2;   // tentatively taken by this search
            expand(n, steps - depth - 1);
            st.push_back(n);
        }
        for (size_t k = 1; k < st.size(); ++k)
            occ[pack(st[k].c[0], st[k].c[1], st[k].c[2])] = -1;
        return false;
    }

    void unroute(uint32_t li)
    {
        for (int c : paths[li])
            occ[c] = -1;
        paths[li].clear();
    }
};

}  // namespace

extern "C" int bfm_gen_diamond_network_size(uint32_t box, uint32_t cx, uint32_t cy, uint32_t cz, uint32_t strand_len,
                                            uint32_t *n_monomers, uint32_t *n_bonds)
{
    if (!n_monomers || !n_bonds || !pow2(box) || box < 8 || box > 1024 || !cx || !cy || !cz || !strand_len)
        return gen_fail(BFM_E_INVALID_ARGUMENT, "diamond network: bad arguments");
    const uint32_t G = box / 2;
    if (G % cx || G % cy || G % cz || (G / cx) % 4 || (G / cy) % 4 || (G / cz) % 4)
        return gen_fail(BFM_E_INVALID_ARGUMENT,
                        "diamond network: box/2 must be a multiple of 4*cells on every axis");
    const uint32_t q[3] = {G / cx / 4, G / cy / 4, G / cz / 4};
    const uint32_t manhattan = q[0] + q[1] + q[2];
    if (manhattan > strand_len + 1 || (manhattan ^ (strand_len + 1)) & 1u)
        return gen_fail(BFM_E_INVALID_ARGUMENT,
                        "diamond network: strand_len+1 must be >= and of the same parity as the crosslinker "
                        "spacing in even-sublattice steps");
    const uint64_t nodes = 8ull * cx * cy * cz;
    const uint64_t strands = 2 * nodes;
    const uint64_t n = nodes + strands * strand_len;
    const uint64_t b = strands * (strand_len + 1);
    if (n > uint64_t(G) * G * G || b > 0xFFFFFFFFull)
        return gen_fail(BFM_E_INVALID_ARGUMENT, "diamond network: does not fit into the box");
    *n_monomers = uint32_t(n);
    *n_bonds = uint32_t(b);
    pscbfm::set_last_error("");
    return BFM_OK;
}

extern "C" int bfm_gen_diamond_network(uint32_t box, uint32_t cx, uint32_t cy, uint32_t cz, uint32_t strand_len,
                                       uint64_t seed, int32_t *xyz, uint32_t *bonds)
{
    uint32_t n_mono = 0, n_bonds = 0;
    int rc = bfm_gen_diamond_network_size(box, cx, cy, cz, strand_len, &n_mono, &n_bonds);
    if (rc != BFM_OK)
        return rc;
    if (!xyz || !bonds)
        return gen_fail(BFM_E_INVALID_ARGUMENT, "diamond network: null output array");
    const int G = int(box / 2);
    const int q[3] = {G / int(cx) / 4, G / int(cy) / 4, G / int(cz) / 4};
    Router R(G, strand_len + 1, seed * 0x9E3779B97F4A7C15ull + 0xD1A40ull);

    // crosslinkers: conventional diamond cell, sublattice A then B
    static const int kSub[8][3] = {{0, 0, 0}, {0, 2, 2}, {2, 0, 2}, {2, 2, 0},
                                   {1, 1, 1}, {1, 3, 3}, {3, 1, 3}, {3, 3, 1}};
    for (uint32_t i = 0; i < cx; ++i)
        for (uint32_t j = 0; j < cy; ++j)
            for (uint32_t k = 0; k < cz; ++k)
                for (int s = 0; s < 8; ++s) {
                    const std::array<int, 3> p = {R.wrap(int(i) * 4 * q[0] + kSub[s][0] * q[0]),
                                                  R.wrap(int(j) * 4 * q[1] + kSub[s][1] * q[1]),
                                                  R.wrap(int(k) * 4 * q[2] + kSub[s][2] * q[2])};
                    int &o = R.occ[R.pack(p[0], p[1], p[2])];
                    if (o != -1)
                        return gen_fail(BFM_E_RUNTIME, "diamond network: crosslinkers coincide");
                    o = int(R.nodes.size());
                    R.nodes.push_back(p);
                }
    // links from every A crosslinker along the four tetrahedral directions
    static const int kTet[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
    for (uint32_t nid = 0; nid < R.nodes.size(); ++nid) {
        if (nid % 8 >= 4)
            continue;
        const auto &p = R.nodes[nid];
        for (const auto &t : kTet) {
            const int tgt[3] = {R.wrap(p[0] + t[0] * q[0]), R.wrap(p[1] + t[1] * q[1]), R.wrap(p[2] + t[2] * q[2])};
            const int o = R.occ[R.pack(tgt[0], tgt[1], tgt[2])];
            if (o < 0 || uint32_t(o) >= R.nodes.size())
                return gen_fail(BFM_E_RUNTIME, "diamond network: missing partner crosslinker");
            Link L;
            L.u = nid;
            L.v = uint32_t(o);
            for (int a = 0; a < 3; ++a)
                L.mid[a] = float(p[a]) + 0.5f * float(t[a] * q[a]);
            R.links.push_back(L);
        }
    }
    const uint32_t NL = uint32_t(R.links.size());
    R.paths.assign(NL, {});

    // route in random order; rip up and reroute around failures
    std::vector<uint32_t> order(NL);
    for (uint32_t k = 0; k < NL; ++k)
        order[k] = k;
    for (uint32_t k = NL; k-- > 1;)
        std::swap(order[k], order[R.rng.below(k + 1)]);
    std::vector<uint32_t> failed;
    const uint64_t budget = 40ull * (strand_len + 1) * 6 + 20000;
    for (uint32_t li : order)
        if (!R.route(li, budget))
            failed.push_back(li);
    // midpoint buckets for finding neighbours of a failed link
    const int bsz = 8, nb = (G + bsz - 1) / bsz;
    std::vector<std::vector<uint32_t>> bucket(size_t(nb) * nb * nb);
    auto bidx = [&](const float m[3]) {
        int c[3];
        for (int a = 0; a < 3; ++a)
            c[a] = R.wrap(int(std::floor(m[a]))) / bsz;
        return (size_t(c[2]) * nb + c[1]) * nb + c[0];
    };
    for (uint32_t li = 0; li < NL; ++li)
        bucket[bidx(R.links[li].mid)].push_back(li);
    const uint64_t max_rounds = 50ull * failed.size() + 1000;
    uint64_t round = 0;
    while (!failed.empty()) {
        if (++round > max_rounds)
            return gen_fail(BFM_E_RUNTIME, "diamond network: strand routing did not converge");
        const uint32_t li = failed.back();
        failed.pop_back();
        const Link &L = R.links[li];
        std::vector<uint32_t> near;
        int c[3];
        for (int a = 0; a < 3; ++a)
            c[a] = R.wrap(int(std::floor(L.mid[a]))) / bsz;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const size_t b = (size_t((c[2] + dz + nb) % nb) * nb + size_t((c[1] + dy + nb) % nb)) * nb +
                                     size_t((c[0] + dx + nb) % nb);
                    for (uint32_t lj : bucket[b]) {
                        if (lj == li || R.paths[lj].empty())
                            continue;
                        const float d = R.fd(L.mid[0], R.links[lj].mid[0]) + R.fd(L.mid[1], R.links[lj].mid[1]) +
                                        R.fd(L.mid[2], R.links[lj].mid[2]);
                        if (d <= 6.5f)
                            near.push_back(lj);
                    }
                }
        for (size_t k = near.size(); k-- > 1;)
            std::swap(near[k], near[R.rng.below(uint32_t(k + 1))]);
        const size_t rip = std::min<size_t>(near.size(), 1 + round / 10 % 8);
        for (size_t k = 0; k < rip; ++k)
            R.unroute(near[k]);
        if (!R.route(li, 4 * budget))
            failed.insert(failed.begin(), li);
        for (size_t k = 0; k < rip; ++k)
            if (!R.route(near[k], 4 * budget))
                failed.push_back(near[k]);
    }

    // ids: crosslinkers, then each strand from its A end; coordinates are
    // fine-lattice (2 x cell), strands unwrapped from their first end
    const uint32_t K = uint32_t(R.nodes.size());
    for (uint32_t nid = 0; nid < K; ++nid)
        for (int a = 0; a < 3; ++a)
            xyz[3 * size_t(nid) + a] = 2 * R.nodes[nid][a];
    uint64_t e = 0;
    for (uint32_t li = 0; li < NL; ++li) {
        const Link &L = R.links[li];
        int prev[3] = {R.nodes[L.u][0], R.nodes[L.u][1], R.nodes[L.u][2]};
        uint32_t prev_id = L.u;
        for (uint32_t k = 0; k < strand_len; ++k) {
            const int cell = R.paths[li][k];
            const int w[3] = {cell % G, (cell / G) % G, cell / (G * G)};
            int cur[3];
            for (int a = 0; a < 3; ++a)
                cur[a] = prev[a] + R.wd(prev[a], w[a]);
            const uint32_t id = K + li * strand_len + k;
            for (int a = 0; a < 3; ++a)
                xyz[3 * size_t(id) + a] = 2 * cur[a];
            bonds[2 * e] = prev_id;
            bonds[2 * e + 1] = id;
            ++e;
            prev_id = id;
            for (int a = 0; a < 3; ++a)
                prev[a] = cur[a];
        }
        bonds[2 * e] = prev_id;
        bonds[2 * e + 1] = L.v;
        ++e;
    }
    pscbfm::set_last_error("");
    return BFM_OK;
}
