// bfm_io.cpp -- LeMonADE BFM file reader and writer.
#include "bfm_io.hpp"

#include <array>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>

#include "host_system.hpp"

namespace pscbfm {
namespace {

[[noreturn]] void bad(const std::string &path, size_t line_no, const std::string &what)
{
    std::ostringstream msg;
    msg << path << ":" << line_no << ": " << what;
    throw std::runtime_error(msg.str());
}

bool starts_with(const std::string &s, const char *prefix) { return s.compare(0, std::strlen(prefix), prefix) == 0; }

// Integer after "!key=" (the whole line is "!key=value").
int64_t value_of(const std::string &path, size_t line_no, const std::string &line)
{
    const size_t eq = line.find('=');
    if (eq == std::string::npos)
        bad(path, line_no, "expected '=' in '" + line + "'");
    errno = 0;
    char *end = nullptr;
    const long long v = std::strtoll(line.c_str() + eq + 1, &end, 10);
    if (errno != 0 || end == line.c_str() + eq + 1)
        bad(path, line_no, "expected an integer in '" + line + "'");
    return v;
}

// Parses a signed integer at p; advances p.
bool parse_int(const char *&p, const char *end, long long &v)
{
    while (p < end && *p == ' ')
        ++p;
    if (p >= end)
        return false;
    errno = 0;
    char *q = nullptr;
    v = std::strtoll(p, &q, 10);
    if (q == p || errno != 0)
        return false;
    p = q;
    return true;
}

}  // namespace

std::vector<std::pair<int, int>> bondset_ids_for(const uint8_t forbidden[512])
{
    std::vector<std::pair<int, int>> out;
    int id = 17;
    for (int idx = 0; idx < 512; ++idx) {
        if (forbidden[idx])
            continue;
        if (id == ' ')   // a trailing space can be lost by line-oriented tools
            ++id;
        if (id > 255)
            throw std::runtime_error("bond set too large for one-byte identifiers");
        out.emplace_back(idx, id++);
    }
    return out;
}

BfmState read_bfm_file(const std::string &path)
{
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f)
        throw std::runtime_error("cannot open '" + path + "'");
    BfmState s;
    std::map<int, std::array<int, 3>> vectors;   // id -> vector
    std::vector<std::pair<uint32_t, uint32_t>> extra;
    std::vector<std::pair<int, int>> ids;
    std::vector<int32_t> attributes;
    std::vector<std::pair<std::pair<uint64_t, uint64_t>, int32_t>> attr_ranges;
    bool have_n = false;
    std::string line;
    size_t line_no = 0;

    // conformation being read
    std::vector<int32_t> xyz;
    std::vector<std::pair<uint32_t, uint32_t>> chain_bonds;
    bool in_conformation = false;
    enum Block { kNone, kBondset, kBonds, kAttributes } block = kNone;
    int64_t conf_age = 0;
    std::vector<int32_t> last_xyz;
    std::vector<std::pair<uint32_t, uint32_t>> last_chain;

    auto finish_conformation = [&]() {
        if (!in_conformation)
            return;
        if (xyz.size() != 3 * size_t(s.n))
            bad(path, line_no, "a conformation holds " + std::to_string(xyz.size() / 3) + " monomers, not " +
                                   std::to_string(s.n));
        last_xyz.swap(xyz);
        last_chain.swap(chain_bonds);
        s.age = conf_age;
        ++s.conformations;
        xyz.clear();
        chain_bonds.clear();
        in_conformation = false;
    };

    while (std::getline(f, line)) {
        ++line_no;
        if (!line.empty() && line.back() == '\r')
            line.pop_back();
        if (line.empty()) {
            finish_conformation();
            block = kNone;
            continue;
        }
        if (line[0] == '!') {
            finish_conformation();
            block = kNone;
            if (starts_with(line, "!number_of_monomers")) {
                const int64_t n = value_of(path, line_no, line);
                if (n < 0 || n > int64_t(kIdMask))
                    bad(path, line_no, "number of monomers out of range");
                s.n = uint32_t(n);
                have_n = true;
            } else if (starts_with(line, "!box_x") || starts_with(line, "!box_y") || starts_with(line, "!box_z")) {
                const int64_t v = value_of(path, line_no, line);
                if (v <= 0)
                    bad(path, line_no, "box edge must be positive");
                s.box[line[5] - 'x'] = uint64_t(v);
            } else if (starts_with(line, "!periodic_x") || starts_with(line, "!periodic_y") ||
                       starts_with(line, "!periodic_z")) {
                s.periodic[line[10] - 'x'] = value_of(path, line_no, line) != 0;
            } else if (starts_with(line, "!set_of_bondvectors")) {
                block = kBondset;
            } else if (starts_with(line, "!bonds") || starts_with(line, "!add_bonds")) {
                block = kBonds;
            } else if (starts_with(line, "!attributes")) {
                block = kAttributes;
            } else if (starts_with(line, "!mcs")) {
                if (!have_n)
                    bad(path, line_no, "!mcs before !number_of_monomers");
                const size_t eq = line.find('=');
                conf_age = eq == std::string::npos ? 0 : value_of(path, line_no, line);
                in_conformation = true;
                xyz.reserve(3 * size_t(s.n));
            }
            // other keys (version, comments, viewer settings) carry nothing for the sweep
            continue;
        }
        const char *p = line.c_str(), *end = p + line.size();
        if (in_conformation) {
            long long x, y, z;
            if (!parse_int(p, end, x) || !parse_int(p, end, y) || !parse_int(p, end, z))
                bad(path, line_no, "expected 'x y z' at the start of a conformation line");
            if (xyz.size() / 3 >= s.n)
                bad(path, line_no, "more monomers than !number_of_monomers");
            xyz.push_back(int32_t(x));
            xyz.push_back(int32_t(y));
            xyz.push_back(int32_t(z));
            if (p < end && *p == ' ')
                ++p;   // one space separates the coordinates from the bond bytes
            for (; p < end; ++p) {
                const int id = static_cast<unsigned char>(*p);
                auto it = vectors.find(id);
                if (it == vectors.end())
                    bad(path, line_no, "bond byte " + std::to_string(id) + " is not in !set_of_bondvectors");
                const size_t k = xyz.size() / 3;
                if (k >= s.n)
                    bad(path, line_no, "more monomers than !number_of_monomers");
                for (int a = 0; a < 3; ++a)
                    xyz.push_back(xyz[3 * (k - 1) + a] + it->second[a]);
                chain_bonds.emplace_back(uint32_t(k - 1), uint32_t(k));
            }
            continue;
        }
        switch (block) {
        case kBondset: {
            long long dx, dy, dz, id;
            if (!parse_int(p, end, dx) || !parse_int(p, end, dy) || !parse_int(p, end, dz) || p >= end || *p != ':')
                bad(path, line_no, "expected 'dx dy dz:id' in !set_of_bondvectors");
            ++p;
            if (!parse_int(p, end, id) || id < 0 || id > 255)
                bad(path, line_no, "bond vector id must be a byte value");
            if (dx < -4 || dx > 3 || dy < -4 || dy > 3 || dz < -4 || dz > 3)
                bad(path, line_no, "bond vector components must lie in [-4, 3]");
            vectors[int(id)] = {int(dx), int(dy), int(dz)};
            ids.emplace_back(int(bond_table_index(int(dx), int(dy), int(dz))), int(id));
            break;
        }
        case kBonds: {
            long long a, b;
            if (!parse_int(p, end, a) || !parse_int(p, end, b))
                bad(path, line_no, "expected 'a b' in !bonds");
            if (a < 1 || b < 1 || (have_n && (uint64_t(a) > s.n || uint64_t(b) > s.n)))
                bad(path, line_no, "bond partner out of range");
            extra.emplace_back(uint32_t(a - 1), uint32_t(b - 1));
            break;
        }
        case kAttributes: {
            long long i, j, v;
            if (!parse_int(p, end, i) || p >= end || *p != '-')
                bad(path, line_no, "expected 'i-j:value' in !attributes");
            ++p;
            if (!parse_int(p, end, j) || p >= end || *p != ':')
                bad(path, line_no, "expected 'i-j:value' in !attributes");
            ++p;
            if (!parse_int(p, end, v) || i < 1 || j < i)
                bad(path, line_no, "bad attribute range");
            attr_ranges.push_back({{uint64_t(i), uint64_t(j)}, int32_t(v)});
            break;
        }
        default:
            bad(path, line_no, "data line outside any block: '" + line.substr(0, 40) + "'");
        }
    }
    finish_conformation();
    if (!have_n)
        throw std::runtime_error(path + ": no !number_of_monomers");
    if (s.conformations == 0)
        throw std::runtime_error(path + ": no conformation (starting with !mcs) found");
    s.xyz.swap(last_xyz);
    // bonds: the explicit ones plus those of the last conformation's chains,
    // each undirected pair once
    std::set<std::pair<uint32_t, uint32_t>> seen;
    for (auto b : extra)
        if (b.first != b.second && seen.insert({std::min(b.first, b.second), std::max(b.first, b.second)}).second)
            s.bonds.push_back({std::min(b.first, b.second), std::max(b.first, b.second)});
    for (auto b : last_chain)
        if (seen.insert(b).second)
            s.bonds.push_back(b);
    s.bondset_ids = ids;
    s.attributes.assign(s.n, 0);
    for (auto &r : attr_ranges)
        for (uint64_t i = r.first.first; i <= r.first.second && i <= s.n; ++i)
            s.attributes[i - 1] = r.second;
    return s;
}

namespace {

void write_conformation(std::ostream &o, const BfmState &s)
{
    std::map<int, int> id_of;   // bond-table index -> id
    for (auto &p : s.bondset_ids)
        id_of[p.first] = p.second;
    std::set<std::pair<uint32_t, uint32_t>> bonded;
    for (auto b : s.bonds)
        bonded.insert({std::min(b.first, b.second), std::max(b.first, b.second)});
    o << "!mcs=" << s.age << "\n";
    for (uint32_t i = 0; i < s.n;) {
        o << s.xyz[3 * size_t(i)] << " " << s.xyz[3 * size_t(i) + 1] << " " << s.xyz[3 * size_t(i) + 2];
        std::string tail;
        uint32_t k = i;
        while (k + 1 < s.n && bonded.count({k, k + 1})) {
            int d[3];
            bool fits = true;
            for (int a = 0; a < 3; ++a) {
                d[a] = s.xyz[3 * size_t(k + 1) + a] - s.xyz[3 * size_t(k) + a];
                fits = fits && d[a] >= -4 && d[a] <= 3;
            }
            auto it = fits ? id_of.find(int(bond_table_index(d[0], d[1], d[2]))) : id_of.end();
            if (it == id_of.end())
                break;
            tail.push_back(char(it->second));
            ++k;
        }
        if (!tail.empty())
            o << " " << tail;
        o << "\n";
        i = k + 1;
    }
    o << "\n";
}

}  // namespace

void write_bfm_file(const std::string &path, const BfmState &s)
{
    std::ofstream o(path, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!o)
        throw std::runtime_error("cannot write '" + path + "'");
    o << "!version=2.0\n";
    o << "!number_of_monomers=" << s.n << "\n";
    o << "!box_x=" << s.box[0] << "\n!box_y=" << s.box[1] << "\n!box_z=" << s.box[2] << "\n";
    o << "!periodic_x=" << (s.periodic[0] ? 1 : 0) << "\n!periodic_y=" << (s.periodic[1] ? 1 : 0)
      << "\n!periodic_z=" << (s.periodic[2] ? 1 : 0) << "\n\n";
    o << "!set_of_bondvectors\n";
    for (auto &p : s.bondset_ids) {
        const int idx = p.first;
        const int dx = ((idx & 7) ^ 4) - 4, dy = (((idx >> 3) & 7) ^ 4) - 4, dz = (((idx >> 6) & 7) ^ 4) - 4;
        o << dx << " " << dy << " " << dz << ":" << p.second << "\n";
    }
    o << "\n";
    if (!s.attributes.empty()) {
        o << "!attributes\n";
        for (uint32_t i = 0; i < s.n;) {
            uint32_t j = i;
            while (j + 1 < s.n && s.attributes[j + 1] == s.attributes[i])
                ++j;
            o << i + 1 << "-" << j + 1 << ":" << s.attributes[i] << "\n";
            i = j + 1;
        }
        o << "\n";
    }
    // bonds that the chain encoding of the conformation does not carry: not
    // between consecutive ids, or not a vector of the set (periodic images)
    std::map<int, int> id_of;
    for (auto &p : s.bondset_ids)
        id_of[p.first] = p.second;
    std::vector<std::pair<uint32_t, uint32_t>> listed;
    for (auto b : s.bonds) {
        const uint32_t a = std::min(b.first, b.second), c = std::max(b.first, b.second);
        bool chain = false;
        if (c == a + 1) {
            int d[3];
            bool fits = true;
            for (int k = 0; k < 3; ++k) {
                d[k] = s.xyz[3 * size_t(c) + k] - s.xyz[3 * size_t(a) + k];
                fits = fits && d[k] >= -4 && d[k] <= 3;
            }
            chain = fits && id_of.count(int(bond_table_index(d[0], d[1], d[2])));
        }
        if (!chain)
            listed.push_back({a, c});
    }
    if (!listed.empty()) {
        o << "!bonds\n";
        for (auto b : listed)
            o << b.first + 1 << " " << b.second + 1 << "\n";
        o << "\n";
    }
    write_conformation(o, s);
    if (!o)
        throw std::runtime_error("error while writing '" + path + "'");
}

void append_bfm_conformation(const std::string &path, const BfmState &s)
{
    std::ofstream o(path, std::ios::out | std::ios::binary | std::ios::app);
    if (!o)
        throw std::runtime_error("cannot append to '" + path + "'");
    write_conformation(o, s);
    if (!o)
        throw std::runtime_error("error while writing '" + path + "'");
}

}  // namespace pscbfm
