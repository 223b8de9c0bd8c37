// generators.cpp -- seeded synthetic initial states (see pscbfm_amd_systems.h).
#include <algorithm>
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
