// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (built by
// __graft_entry__.build() with g++ -fsanitize=address,undefined): the system
// generators, the planning and checks of initialize(), the LeMonADE .bfm
// reader and writer, and the oracle's sweeps, on small melts and a network.
// Prints "ok <checks>" when every check passes; a sanitizer report ends the
// run with a non-zero status.
//
// usage: host_sanitizer_check <scratch directory>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "bfm_io.hpp"
#include "bfm_oracle.h"
#include "host_system.hpp"
#include "pscbfm_amd.h"
#include "pscbfm_amd_systems.h"

// The library keeps the message of the last failed call per thread (capi.cpp,
// not built here); the generators report through this.
namespace pscbfm {
void set_last_error(const std::string &msg)
{
    if (!msg.empty())
        std::fprintf(stderr, "last error: %s\n", msg.c_str());
}
}  // namespace pscbfm

namespace {

int g_checks = 0;

void expect(bool ok, const std::string &what)
{
    ++g_checks;
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what.c_str());
        std::exit(1);
    }
}

struct Sys {
    uint32_t box = 0;
    std::vector<int32_t> xyz;     // 3 * n
    std::vector<uint32_t> bonds;  // 2 * bonds
    uint32_t n() const { return uint32_t(xyz.size() / 3); }
};

Sys melt(uint32_t box, uint32_t chains, uint32_t len, uint64_t seed)
{
    Sys s;
    s.box = box;
    s.xyz.resize(size_t(3) * chains * len);
    s.bonds.resize(size_t(2) * chains * (len - 1));
    expect(bfm_gen_linear_melt(box, chains, len, seed, s.xyz.data(), s.bonds.data()) == BFM_OK, "melt generator");
    return s;
}

Sys network(uint32_t box, uint32_t cx, uint32_t cy, uint32_t cz, uint32_t strand, uint64_t seed)
{
    uint32_t nm = 0, nb = 0;
    expect(bfm_gen_diamond_network_size(box, cx, cy, cz, strand, &nm, &nb) == BFM_OK, "network size");
    Sys s;
    s.box = box;
    s.xyz.resize(size_t(3) * nm);
    s.bonds.resize(size_t(2) * nb);
    expect(bfm_gen_diamond_network(box, cx, cy, cz, strand, seed, s.xyz.data(), s.bonds.data()) == BFM_OK,
           "network generator");
    return s;
}

pscbfm::HostSystem host_of(const Sys &s)
{
    pscbfm::HostSystem h;
    for (int a = 0; a < 3; ++a) {
        h.box[a] = s.box;
        h.periodic[a] = true;
    }
    h.n = s.n();
    h.xyz = s.xyz;
    h.attribute.assign(h.n, 1);
    h.nnbr.assign(h.n, 0);
    h.nbr.assign(size_t(h.n) * pscbfm::kMaxConnectivity, 0);
    for (size_t k = 0; k + 1 < s.bonds.size(); k += 2) {
        const uint32_t a = s.bonds[k], b = s.bonds[k + 1];
        h.nbr[size_t(a) * pscbfm::kMaxConnectivity + h.nnbr[a]++] = b;
        h.nbr[size_t(b) * pscbfm::kMaxConnectivity + h.nnbr[b]++] = a;
    }
    pscbfm::standard_bond_set(h.forbidden);
    return h;
}

void check_plan(const Sys &s, const char *name)
{
    pscbfm::HostSystem h = host_of(s);
    pscbfm::plan(h);
    const std::string tag = name;
    expect(h.n_colours >= 2, tag + ": colours");
    uint64_t total = 0;
    for (uint32_t c = 0; c < h.n_colours; ++c)
        total += h.seg_count[c];
    expect(total == h.n, tag + ": colour segments cover every monomer");
    std::vector<uint8_t> seen(h.n_pad, 0);
    for (uint32_t i = 0; i < h.n; ++i) {
        const uint32_t t = h.to_sorted[i];
        expect(t < h.n_pad && !seen[t], tag + ": sorted slots are distinct");
        seen[t] = 1;
        expect((h.ids_sorted[t] & pscbfm::kIdMask) == i, tag + ": sorted slot holds its id");
    }
    const std::vector<int32_t> before = h.xyz;
    pscbfm::unpack_positions(h);
    expect(h.xyz == before, tag + ": packed positions unpack to the input");
}

template <class F>
void expect_throw(F f, const std::string &what)
{
    bool threw = false;
    try {
        f();
    } catch (const std::exception &) {
        threw = true;
    }
    expect(threw, what);
}

void check_plan_errors(const Sys &s)
{
    {
        pscbfm::HostSystem h = host_of(s);
        for (int a = 0; a < 3; ++a)
            h.xyz[3 + a] = h.xyz[a];   // monomer 1 on top of monomer 0
        expect_throw([&] { pscbfm::plan(h); }, "overlap rejected");
    }
    {
        pscbfm::HostSystem h = host_of(s);
        h.box[0] = 48;
        expect_throw([&] { pscbfm::plan(h); }, "box that is not a power of two rejected");
    }
    {
        pscbfm::HostSystem h = host_of(s);
        h.forbidden[pscbfm::bond_table_index(2, 0, 0)] = 1;
        expect_throw([&] { pscbfm::plan(h); }, "bond set of 107 vectors rejected");
    }
}

void check_bfm_io(const Sys &s, const std::string &dir)
{
    pscbfm::BfmState st;
    for (int a = 0; a < 3; ++a) {
        st.box[a] = s.box;
        st.periodic[a] = true;
    }
    st.n = s.n();
    st.xyz = s.xyz;
    for (size_t k = 0; k + 1 < s.bonds.size(); k += 2)
        st.bonds.emplace_back(s.bonds[k], s.bonds[k + 1]);
    uint8_t forbidden[512];
    pscbfm::standard_bond_set(forbidden);
    st.bondset_ids = pscbfm::bondset_ids_for(forbidden);
    st.attributes.assign(st.n, 2);
    st.age = 5;
    const std::string path = dir + "/sanitizer_check.bfm";
    pscbfm::write_bfm_file(path, st);
    pscbfm::BfmState back = pscbfm::read_bfm_file(path);
    expect(back.n == st.n && back.xyz == st.xyz && back.age == st.age, "bfm round trip: positions and age");
    auto sorted = [](std::vector<std::pair<uint32_t, uint32_t>> v) {
        for (auto &p : v)
            if (p.first > p.second)
                std::swap(p.first, p.second);
        std::sort(v.begin(), v.end());
        return v;
    };
    expect(sorted(back.bonds) == sorted(st.bonds), "bfm round trip: bonds");
    st.age = 9;
    pscbfm::append_bfm_conformation(path, st);
    back = pscbfm::read_bfm_file(path);
    expect(back.conformations == 2 && back.age == 9 && back.xyz == st.xyz, "bfm append: last conformation");
    std::remove(path.c_str());
}

void check_oracle(const Sys &s, uint64_t seed)
{
    const uint32_t n = s.n();
    std::vector<uint32_t> off(n + 1, 0), idx;
    std::vector<std::vector<uint32_t>> adj(n);
    for (size_t k = 0; k + 1 < s.bonds.size(); k += 2) {
        adj[s.bonds[k]].push_back(s.bonds[k + 1]);
        adj[s.bonds[k + 1]].push_back(s.bonds[k]);
    }
    for (uint32_t i = 0; i < n; ++i) {
        off[i + 1] = off[i] + uint32_t(adj[i].size());
        idx.insert(idx.end(), adj[i].begin(), adj[i].end());
    }
    std::vector<uint32_t> colour(n);
    uint32_t n_colours = 0;
    expect(orc_graph_coloring(n, off.data(), idx.data(), 1, colour.data(), &n_colours) == 0, "oracle colouring");
    uint8_t forbidden[512];
    orc_standard_bondset(forbidden);
    orc_system sys{};
    for (int a = 0; a < 3; ++a) {
        sys.box[a] = int32_t(s.box);
        sys.periodic[a] = 1;
    }
    sys.n = n;
    sys.nbr_off = off.data();
    sys.nbr_idx = idx.data();
    sys.color = colour.data();
    sys.n_colors = n_colours;
    sys.forbidden = forbidden;
    sys.seed = seed;
    sys.rng = ORC_RNG_INT_HASH;
    char msg[256];
    std::vector<int32_t> serial = s.xyz, parallel = s.xyz, rs = s.xyz;
    expect(orc_sweep(&sys, serial.data(), 0, 3, 1, nullptr, nullptr) == 0, "oracle sweep");
    expect(orc_check_system(&sys, serial.data(), msg, sizeof msg) == 0, "oracle sweep keeps a valid state");
    expect(orc_sweep_parallel(&sys, parallel.data(), 0, 3, 2, nullptr, nullptr) == 0, "oracle OpenMP sweep");
    expect(parallel == serial, "oracle OpenMP sweep equals the serial one");
    expect(orc_random_sequential(&sys, rs.data(), 0, 2, nullptr) == 0, "oracle random-sequential updater");
    expect(orc_check_system(&sys, rs.data(), msg, sizeof msg) == 0, "random-sequential keeps a valid state");
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const Sys m = melt(32, 64, 16, 11);
    const Sys net = network(64, 2, 2, 8, 32, 12);
    check_plan(m, "melt");
    check_plan(net, "network");
    check_plan_errors(m);
    check_bfm_io(m, dir);
    check_bfm_io(net, dir);
    check_oracle(m, 21);
    check_oracle(net, 22);
    std::printf("ok %d\n", g_checks);
    return 0;
}
