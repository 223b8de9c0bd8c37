// analyzers_host.cpp -- host half of the analyzers: chain discovery, bond list,
// AnalyzerMsd bookkeeping and its output file.  No HIP calls here.
#include "analyzers.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <stdexcept>

namespace pscbfm {

std::vector<std::vector<uint32_t>> link_lists(const HostSystem &s)
{
    std::vector<std::vector<uint32_t>> links(s.n);
    for (uint32_t i = 0; i < s.n; ++i)
        for (uint32_t j = 0; j < s.nnbr[i]; ++j) {
            const uint32_t p = s.nbr[size_t(i) * kMaxConnectivity + j];
            if (std::find(links[i].begin(), links[i].end(), p) == links[i].end())
                links[i].push_back(p);
        }
    for (uint32_t i = 0; i < s.n; ++i)
        for (uint32_t j = 0; j < s.nnbr[i]; ++j) {
            const uint32_t p = s.nbr[size_t(i) * kMaxConnectivity + j];
            if (std::find(links[p].begin(), links[p].end(), i) == links[p].end())
                links[p].push_back(i);
        }
    return links;
}

std::vector<std::pair<uint32_t, uint32_t>> bond_pairs(const HostSystem &s)
{
    const auto links = link_lists(s);
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (uint32_t i = 0; i < s.n; ++i)
        for (uint32_t p : links[i])
            if (p > i)
                pairs.emplace_back(i, p);
    std::sort(pairs.begin(), pairs.end());
    return pairs;
}

ChainTable find_linear_chains(const HostSystem &s)
{
    const auto links = link_lists(s);
    ChainTable t;
    t.linear = true;
    std::vector<uint8_t> visited(s.n, 0);
    std::vector<uint32_t> stack, ends;
    // the walk of AnalyzerMsd.tpp:321-342 and :247-265: go to link 0 unless it
    // is where we came from, else to link 1 if there is one
    auto step = [&](uint32_t cur, uint32_t prev, uint32_t &next) {
        const auto &l = links[cur];
        if (!l.empty() && l[0] != prev) {
            next = l[0];
            return true;
        }
        if (l.size() >= 2) {
            next = l[1];
            return true;
        }
        return false;
    };
    t.offsets.push_back(0);
    for (uint32_t start = 0; start < s.n; ++start) {
        if (visited[start])
            continue;
        ends.clear();
        uint32_t size = 1;
        stack.assign(1, start);
        visited[start] = 1;
        while (!stack.empty()) {
            const uint32_t m = stack.back();
            stack.pop_back();
            const size_t nl = links[m].size();
            if (nl > 2) {
                t.linear = false;
                t.msd_error = "Some monomers have more than two bonds, indicating that this polymer system does not "
                              "only consist of linear chains. This Analyzer module only works with linear chains!";
                return t;
            }
            if (nl <= 1)
                ends.push_back(m);
            for (uint32_t p : links[m])
                if (!visited[p]) {
                    ++size;
                    visited[p] = 1;
                    stack.push_back(p);
                }
        }
        if (size == 1 && ends.size() == 1)
            ends.push_back(ends[0]);
        if (ends.size() != 2) {
            t.linear = false;
            t.msd_error = "Found not exactly 2 end monomers! The polymer system might be weird, e.g. end monomers "
                          "with have loops or ring polymers.";
            return t;
        }
        // middle monomer: floor(size / 2) steps from the first end
        uint32_t mid = ends[0], prev = mid;
        for (uint32_t k = 0; k < size / 2 && t.msd_error.empty(); ++k) {
            uint32_t next = mid;
            step(mid, prev, next);
            prev = mid;
            mid = next;
            if (links[mid].size() < 2)
                t.msd_error = "Unexpected end monomer encountered when iterating over the polymer to find the "
                              "middle monomer.";
        }
        // members from the first end
        uint32_t cur = ends[0];
        prev = cur;
        for (uint32_t k = 0; k < size; ++k) {
            t.members.push_back(cur);
            uint32_t next = cur;
            if (!step(cur, prev, next))
                break;
            prev = cur;
            cur = next;
        }
        t.offsets.push_back(uint32_t(t.members.size()));
        t.min_len = t.n_chains == 0 ? size : std::min(t.min_len, size);
        t.max_len = std::max(t.max_len, size);
        ++t.n_chains;
    }
    return t;
}

double OnlineStatistics::sg() const
{
    // getCentralMoment(2) of AnalyzerMsd.tpp:72-85 with mMoments.size() == 3
    auto choose = [](unsigned long total, unsigned long k) {
        if (k > total / 2)
            k = total - k;
        double r = 1;
        for (unsigned long i = 0; i < k; ++i)
            r *= double(total - i) / double(i + 1);
        return r;
    };
    const double moment[3] = {m0 / m0, m1 / m0, m2 / m0};
    double cmom = 0;
    for (unsigned i = 0; i <= 2; ++i)
        cmom += choose(3, i) * moment[2 - i] * (i % 2 == 0 ? +1 : -1) * std::pow(moment[1], double(i));
    return std::sqrt(cmom * (m0 / (m0 - 1)));
}

std::vector<MsdPoint> DeviceAnalyzer::msdPoints() const
{
    std::vector<MsdPoint> out(mMsds.size());
    for (size_t e = 0; e < mMsds.size(); ++e) {
        out[e].dt = e < mEvalTimes.size() ? mEvalTimes[e] : 0;
        out[e].n_values = mMsds[e][0].m0;
        for (int g = 0; g < 5; ++g) {
            out[e].mean[g] = mMsds[e][g].mean();
            out[e].sg[g] = mMsds[e][g].sg();
        }
    }
    return out;
}

void DeviceAnalyzer::msdWrite(const HostSystem &s, const std::string &path) const
{
    if (!mMsd)
        throw std::runtime_error("no MSD analyzer attached");
    std::FILE *f = std::fopen(path.c_str(), "w");
    if (!f)
        throw std::runtime_error("cannot open '" + path + "' for writing");
    const double density = 8.0 * double(s.n) / (double(s.box[0]) * double(s.box[1]) * double(s.box[2]));
    std::fprintf(f, "# nMinMonomersPerChain = %u\n", mChains.min_len);
    std::fprintf(f, "# nMaxMonomersPerChain = %u\n", mChains.max_len);
    std::fprintf(f, "# nChains = %u\n", mChains.n_chains);
    std::fprintf(f, "# numberDensity = %g\n", density);
    std::fprintf(f, "# nMcsSkipped = %llu\n", (unsigned long long)mMcsToSkip);
    std::fprintf(f, "# deltaTRealization = %llu\n", (unsigned long long)mDtReal);
    std::fprintf(f, "# iMcsLastRealization = %llu\n", (unsigned long long)mLastRealization);
    std::fprintf(f, "# deltaTEvaluate = %llu\n", (unsigned long long)mDtEval);
    std::fprintf(f, "# iMcsLastEvaluation = %llu\n", (unsigned long long)mLastEvaluation);
    std::fprintf(f, "# nSubsequences = %zu\n", mFirstMcs.size());
    std::fprintf(f, "# g1 to g5 are in this order the MSDs of: first monomer, middle monomer, end monomer, chain "
                    "center of mass (COM), middle monomer - chain COM\n");
    std::fprintf(f, "# dt nSamples g1 sg1 g2 sg2 g3 sg3 g4 sg4 g5 sg5\n");
    for (const MsdPoint &p : msdPoints()) {
        std::fprintf(f, "%llu %g", (unsigned long long)p.dt, p.n_values);
        for (int g = 0; g < 5; ++g)
            std::fprintf(f, " %g %g", p.mean[g], p.sg[g]);
        std::fprintf(f, "\n");
    }
    const bool ok = std::ferror(f) == 0;
    if (std::fclose(f) != 0 || !ok)
        throw std::runtime_error("writing '" + path + "' failed");
}

}  // namespace pscbfm
