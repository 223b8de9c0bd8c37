// analyzers.hpp -- observables computed on the device from the resident state.
//
// The reference analyses its melts on the host, after copying every position
// back (AnalyzerMsd, src/AnalyzerMsd.tpp:99-421; LeMonADE's analyzers behind
// the TaskManager of mainCUDASimulatorGPUScBFM_AB_Type.cpp:166-177).  Here the
// same quantities are reduced on the device and only the reductions cross
// PCIe:
//  * the bond-vector histogram (every undirected bond once, oriented from the
//    lower to the higher id, by bond-table index (dx&7)|(dy&7)<<3|(dz&7)<<6);
//  * for systems of linear chains: the mean squared end-to-end distance and
//    radius of gyration, the box centre of mass (mean of the chains' centres
//    of mass, AnalyzerMsd.tpp:346-349);
//  * AnalyzerMsd itself: g1..g5 = MSDs of the first, middle and last monomer,
//    of the chain centre of mass and of the middle monomer relative to it,
//    over time-shifted subsequences, with the box COM drift removed from
//    g1..g4 (AnalyzerMsd.tpp:303-421), and its output file (:423-468).
//
// Integer parts (histogram, per-chain sums, the end-to-end sum) are exact.
// Floating-point means come from fixed-order device reductions: the same
// inputs give the same bits on every run, and they differ from a sequential
// host sum only by rounding.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "bfm_kernels.hpp"
#include "host_system.hpp"

namespace pscbfm {

// Linear chains as AnalyzerMsd::getLinearPolymerSystemInfo finds them
// (AnalyzerMsd.tpp:186-288): components in order of their lowest id, the two
// ends in the order a depth-first walk with a stack pops them, members listed
// by walking from the first end.  `linear` is false when some monomer has more
// than 2 links or a component has not exactly 2 ends; `msd_error` holds the
// message the reference's analyzer throws for such a system, and also for a
// two-monomer chain, whose middle-monomer walk ends on an end monomer (:263-264).
struct ChainTable {
    bool linear = false;
    std::string msd_error;
    uint32_t n_chains = 0;
    uint32_t min_len = 0, max_len = 0;
    std::vector<uint32_t> members;   // chain by chain, walked from the first end
    std::vector<uint32_t> offsets;   // n_chains + 1
};

// Per-monomer links in the order they were given, each partner once, followed
// by the reverse direction of links given only one way (the symmetric link
// lists of the reference's molecules container).
std::vector<std::vector<uint32_t>> link_lists(const HostSystem &s);
ChainTable find_linear_chains(const HostSystem &s);
// Undirected bonds (a < b), in increasing (a, b).
std::vector<std::pair<uint32_t, uint32_t>> bond_pairs(const HostSystem &s);

// One linear chain at one time, relative to its first monomer (exact).
struct ChainSums {
    int64_t s[3];        // sum over members of r - r_first
    int64_t q;           // sum over members of |r - r_first|^2
    int32_t first[3], middle[3], last[3];
    uint32_t n;
};

struct Observables {
    uint64_t bonds = 0;
    uint64_t hist[512] = {};
    uint32_t n_chains = 0;
    int64_t re2_sum = 0;    // sum over chains of |r_last - r_first|^2
    double re2 = 0, rg2 = 0;
    double box_com[3] = {0, 0, 0};
};

// The reference's OnlineStatistics with two moments (AnalyzerMsd.tpp:18-95):
// n, sum x, sum x^2.  `sg` is what its cleanup() prints next to each mean:
// sqrt(getCentralMoment(2)), whose binomial factors are choose(3, i) (the
// number of stored moments, :79), so it is the Bessel-scaled root of the raw
// second moment, evaluated term by term as there.
struct OnlineStatistics {
    double m0 = 0, m1 = 0, m2 = 0;
    void add(double count, double sum, double sum_sq)
    {
        m0 += count;
        m1 += sum;
        m2 += sum_sq;
    }
    double mean() const { return m1 / m0; }
    double sg() const;
};

struct MsdPoint {
    uint64_t dt;
    double n_values;
    double mean[5], sg[5];
};

// Device half: positions by id, per-chain sums, reductions, MSD snapshots.
class DeviceAnalyzer {
public:
    DeviceAnalyzer() = default;
    ~DeviceAnalyzer();
    DeviceAnalyzer(const DeviceAnalyzer &) = delete;
    DeviceAnalyzer &operator=(const DeviceAnalyzer &) = delete;

    // Host tables and device buffers for system `s` (once per topology).
    void prepare(const HostSystem &s);
    bool prepared() const { return mPrepared; }
    const ChainTable &chains() const { return mChains; }
    // Observables of the state in `buf` (stream-ordered, then synchronises).
    void observe(const HostSystem &s, const SweepBuffers &buf, hipStream_t stream, Observables &out);

    // AnalyzerMsd(rnMcsToSkip, rDeltaTRealization, rDeltaTEvaluate,
    // rnMaxSubsequences) followed by its initialize() (:141-177): throws
    // std::invalid_argument with the reference's message if the system is not
    // made of linear chains.
    void msdCreate(const HostSystem &s, const SweepBuffers &buf, hipStream_t stream, uint64_t age,
                   uint64_t mcs_to_skip, uint64_t delta_t_realization, uint64_t delta_t_evaluate,
                   uint64_t max_subsequences);
    bool msdAttached() const { return mMsd; }
    // AnalyzerMsd::execute at `age` (:303-421).
    void msdExecute(const HostSystem &s, const SweepBuffers &buf, hipStream_t stream, uint64_t age);
    std::vector<MsdPoint> msdPoints() const;
    // AnalyzerMsd::cleanup's file (:423-468).
    void msdWrite(const HostSystem &s, const std::string &path) const;
    uint32_t msdSubsequences() const { return uint32_t(mFirstMcs.size()); }

    // Per-chain sums of the state in `buf` (an extension used to check the
    // exact integer part).
    void chainSums(const HostSystem &s, const SweepBuffers &buf, hipStream_t stream, std::vector<ChainSums> &out);

    void release();

private:
    void ensureSnapshots(uint32_t count, hipStream_t stream);
    void computeChains(const HostSystem &s, const SweepBuffers &buf, hipStream_t stream);

    bool mPrepared = false;
    ChainTable mChains;
    uint64_t mBonds = 0;
    uint32_t mN = 0;
    void *mDev = nullptr;          // positions, members, offsets, bond pairs, chain sums, results
    int4 *mPos = nullptr;
    uint32_t *mMembers = nullptr, *mOffsets = nullptr;
    uint2 *mPairs = nullptr;
    ChainSums *mNow = nullptr;
    void *mResult = nullptr;       // DeviceResult
    void *mHostResult = nullptr;   // pinned copy

    // AnalyzerMsd state
    bool mMsd = false;
    uint64_t mMcsToSkip = 0, mDtReal = 1, mDtEval = 1;
    uint64_t mLastRealization = 0, mLastEvaluation = 0;
    bool mAnyRealization = false;
    std::vector<uint64_t> mFirstMcs;               // per subsequence
    std::vector<std::vector<OnlineStatistics>> mMsds;   // [eval point][g]
    std::vector<uint64_t> mEvalTimes;
    ChainSums *mSnap = nullptr;    // per subsequence: chain sums at its start
    double *mSnapCom = nullptr;    // per subsequence: box COM at its start
    double *mSums = nullptr;       // per subsequence: 5 x (sum, sum of squares)
    uint32_t mSnapCap = 0;
};

}  // namespace pscbfm
