// analyzers.hip -- device half of the analyzers (gfx950).
//
// Pipeline of one evaluation, all on the updater's stream:
//   k_positions    unwrapped position of every monomer, indexed by original id
//                  (slot order changes with re-sorts; the id does not)
//   k_bond_hist    histogram of bond-table indices over the undirected bonds
//   k_chain_sums   one wavefront per linear chain: exact integer sums
//   k_chain_reduce one workgroup: Re^2, Rg^2 and centre-of-mass sums
//   k_msd          one workgroup per MSD subsequence: sums of g1..g5 and of
//                  their squares over the chains
// Floating-point sums run in a fixed order (per-thread strided loops, then a
// fixed LDS tree), so a result never depends on scheduling.
#include "analyzers.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "updater.hpp"

namespace pscbfm {
namespace {

constexpr int kAnaBlock = 256;
constexpr int kReduceBlock = 1024;
constexpr int kHistBins = 512;

struct DeviceResult {
    unsigned long long hist[kHistBins];
    long long re2_sum;
    double rg2_sum;
    double com_sum[3];
    double box_com[3];
};

__global__ __launch_bounds__(kHistBins) void k_clear_result(DeviceResult *r)
{
    r->hist[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(kAnaBlock) void k_positions(const uint32_t *__restrict__ ppos,
                                                         const uint32_t *__restrict__ ids,
                                                         const int32_t *__restrict__ image, uint32_t n_pad,
                                                         Geometry g, int64_t bx, int64_t by, int64_t bz,
                                                         int4 *__restrict__ pos)
{
    const uint32_t t = blockIdx.x * kAnaBlock + threadIdx.x;
    if (t >= n_pad)
        return;
    const uint32_t id = ids[t];
    if (id == 0xFFFFFFFFu)
        return;
    const uint32_t site = ppos[t];
    const int64_t x = int64_t(site_coord(site, g.dmask[0])) + int64_t(image[t]) * bx;
    const int64_t y = int64_t(site_coord(site, g.dmask[1])) + int64_t(image[size_t(n_pad) + t]) * by;
    const int64_t z = int64_t(site_coord(site, g.dmask[2])) + int64_t(image[2 * size_t(n_pad) + t]) * bz;
    pos[id & kIdMask] = make_int4(int32_t(x), int32_t(y), int32_t(z), 0);
}

__global__ __launch_bounds__(kAnaBlock) void k_bond_hist(const uint2 *__restrict__ pairs, uint64_t n_bonds,
                                                         const int4 *__restrict__ pos,
                                                         unsigned long long *__restrict__ hist)
{
    __shared__ uint32_t h[kHistBins];
    for (int k = threadIdx.x; k < kHistBins; k += kAnaBlock)
        h[k] = 0u;
    __syncthreads();
    for (uint64_t b = uint64_t(blockIdx.x) * kAnaBlock + threadIdx.x; b < n_bonds;
         b += uint64_t(gridDim.x) * kAnaBlock) {
        const uint2 p = pairs[b];
        const int4 a = pos[p.x], c = pos[p.y];
        const uint32_t idx = uint32_t((c.x - a.x) & 7) | (uint32_t((c.y - a.y) & 7) << 3) |
                             (uint32_t((c.z - a.z) & 7) << 6);
        atomicAdd(&h[idx], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kHistBins; k += kAnaBlock)
        if (h[k])
            atomicAdd(&hist[k], (unsigned long long)h[k]);
}

__device__ __forceinline__ long long wave_sum(long long v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
        v += __shfl_xor(v, m, 64);
    return v;
}

// One wavefront per chain; members are listed from the first end.
__global__ __launch_bounds__(kAnaBlock) void k_chain_sums(const uint32_t *__restrict__ members,
                                                          const uint32_t *__restrict__ offsets, uint32_t n_chains,
                                                          const int4 *__restrict__ pos, ChainSums *__restrict__ out)
{
    const uint32_t c = blockIdx.x * (kAnaBlock / 64) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (c >= n_chains)
        return;
    const uint32_t b = offsets[c], n = offsets[c + 1] - b;
    const int4 f = pos[members[b]];
    long long sx = 0, sy = 0, sz = 0, q = 0;
    for (uint32_t k = lane; k < n; k += 64) {
        const int4 r = pos[members[b + k]];
        const long long dx = (long long)r.x - f.x, dy = (long long)r.y - f.y, dz = (long long)r.z - f.z;
        sx += dx;
        sy += dy;
        sz += dz;
        q += dx * dx + dy * dy + dz * dz;
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    sz = wave_sum(sz);
    q = wave_sum(q);
    if (lane == 0) {
        const int4 m = pos[members[b + n / 2]], l = pos[members[b + n - 1]];
        ChainSums o;
        o.s[0] = sx;
        o.s[1] = sy;
        o.s[2] = sz;
        o.q = q;
        o.first[0] = f.x, o.first[1] = f.y, o.first[2] = f.z;
        o.middle[0] = m.x, o.middle[1] = m.y, o.middle[2] = m.z;
        o.last[0] = l.x, o.last[1] = l.y, o.last[2] = l.z;
        o.n = n;
        out[c] = o;
    }
}

// Centre of mass of a chain along one axis: the exact member sum divided by
// the member count (AnalyzerMsd.tpp:327,343).
__device__ __forceinline__ double chain_com(const ChainSums &c, int a)
{
    return double((long long)c.first[a] * (long long)c.n + c.s[a]) / double(c.n);
}

// Sums v over the N threads of the workgroup into lds[k][0], pairing
// thread t with t + half for half = N/2, N/4, ..., 1.
template <int K, int N>
__device__ __forceinline__ void block_tree(double (*lds)[N], double (&v)[K])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < K; ++k)
        lds[k][t] = v[k];
    __syncthreads();
    for (int half = N / 2; half >= 1; half >>= 1) {
        if (t < half)
#pragma unroll
            for (int k = 0; k < K; ++k)
                lds[k][t] += lds[k][t + half];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kReduceBlock) void k_chain_reduce(const ChainSums *__restrict__ now, uint32_t n_chains,
                                                               DeviceResult *__restrict__ r)
{
#pragma clang fp contract(off)
    __shared__ double lds[4][kReduceBlock];
    __shared__ long long re2[kReduceBlock];
    double v[4] = {0, 0, 0, 0};   // rg2, com x, y, z
    long long e = 0;
    for (uint32_t c = threadIdx.x; c < n_chains; c += kReduceBlock) {
        const ChainSums x = now[c];
        const double n = double(x.n);
        const double mx = double(x.s[0]) / n, my = double(x.s[1]) / n, mz = double(x.s[2]) / n;
        v[0] += double(x.q) / n - ((mx * mx + my * my) + mz * mz);
        v[1] += chain_com(x, 0);
        v[2] += chain_com(x, 1);
        v[3] += chain_com(x, 2);
        const long long dx = (long long)x.last[0] - x.first[0], dy = (long long)x.last[1] - x.first[1],
                        dz = (long long)x.last[2] - x.first[2];
        e += dx * dx + dy * dy + dz * dz;
    }
    re2[threadIdx.x] = e;
    block_tree<4, kReduceBlock>(lds, v);
    for (int half = kReduceBlock / 2; half >= 1; half >>= 1) {
        if (int(threadIdx.x) < half)
            re2[threadIdx.x] += re2[threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        r->re2_sum = re2[0];
        r->rg2_sum = lds[0][0];
        for (int a = 0; a < 3; ++a) {
            r->com_sum[a] = lds[1 + a][0];
            r->box_com[a] = n_chains ? lds[1 + a][0] / double(n_chains) : 0.0;
        }
    }
}

constexpr int kMsdBlock = 256;

// Workgroup s: subsequence s against the current chain sums
// (AnalyzerMsd.tpp:389-420).  Output: for g1..g5 the sum over chains of the
// squared displacement and of its square.
__global__ __launch_bounds__(kMsdBlock) void k_msd(const ChainSums *__restrict__ now,
                                                   const ChainSums *__restrict__ snap,
                                                   const double *__restrict__ snap_com,
                                                   const DeviceResult *__restrict__ r, uint32_t n_chains,
                                                   double *__restrict__ sums)
{
#pragma clang fp contract(off)
    __shared__ double lds[10][kMsdBlock];
    const uint32_t s = blockIdx.x;
    const ChainSums *first = snap + size_t(s) * n_chains;
    double box[3];
    for (int a = 0; a < 3; ++a)
        box[a] = n_chains > 1 ? r->box_com[a] - snap_com[3 * s + a] : 0.0;
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    auto add = [&](int g, const double (&d)[3]) {
        const double x = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        v[2 * g] += x;
        v[2 * g + 1] += x * x;
    };
    for (uint32_t c = threadIdx.x; c < n_chains; c += kMsdBlock) {
        const ChainSums a = now[c], f = first[c];
        double d[3];
        for (int k = 0; k < 3; ++k)
            d[k] = (double(a.first[k]) - double(f.first[k])) - box[k];
        add(0, d);
        for (int k = 0; k < 3; ++k)
            d[k] = (double(a.middle[k]) - double(f.middle[k])) - box[k];
        add(1, d);
        for (int k = 0; k < 3; ++k)
            d[k] = (double(a.last[k]) - double(f.last[k])) - box[k];
        add(2, d);
        double ca[3], cf[3];
        for (int k = 0; k < 3; ++k) {
            ca[k] = chain_com(a, k);
            cf[k] = chain_com(f, k);
            d[k] = (ca[k] - cf[k]) - box[k];
        }
        add(3, d);
        for (int k = 0; k < 3; ++k)
            d[k] = (double(a.middle[k]) - ca[k]) - (double(f.middle[k]) - cf[k]);
        add(4, d);
    }
    block_tree<10, kMsdBlock>(lds, v);
    if (threadIdx.x < 10)
        sums[size_t(s) * 10 + threadIdx.x] = lds[threadIdx.x][0];
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

DeviceAnalyzer::~DeviceAnalyzer() { release(); }

void DeviceAnalyzer::release()
{
    if (mDev)
        hipFree(mDev);
    if (mHostResult)
        hipHostFree(mHostResult);
    if (mSnap)
        hipFree(mSnap);
    if (mSnapCom)
        hipFree(mSnapCom);
    if (mSums)
        hipFree(mSums);
    mDev = mHostResult = nullptr;
    mSnap = nullptr;
    mSnapCom = mSums = nullptr;
    mSnapCap = 0;
    mPrepared = false;
    mMsd = false;
}

void DeviceAnalyzer::prepare(const HostSystem &s)
{
    if (mPrepared)
        return;
    mChains = find_linear_chains(s);
    const auto pairs = bond_pairs(s);
    mBonds = pairs.size();
    mN = s.n;
    const uint32_t C = mChains.linear ? mChains.n_chains : 0;
    size_t off = 0;
    const size_t o_pos = off;
    off += align256(sizeof(int4) * size_t(s.n));
    const size_t o_members = off;
    off += align256(sizeof(uint32_t) * std::max<size_t>(1, mChains.members.size()));
    const size_t o_offsets = off;
    off += align256(sizeof(uint32_t) * (size_t(C) + 1));
    const size_t o_pairs = off;
    off += align256(sizeof(uint2) * std::max<size_t>(1, pairs.size()));
    const size_t o_now = off;
    off += align256(sizeof(ChainSums) * std::max<size_t>(1, C));
    const size_t o_result = off;
    off += align256(sizeof(DeviceResult));
    hip_check(hipMalloc(&mDev, off), "hipMalloc (analyzer)");
    char *base = static_cast<char *>(mDev);
    mPos = reinterpret_cast<int4 *>(base + o_pos);
    mMembers = reinterpret_cast<uint32_t *>(base + o_members);
    mOffsets = reinterpret_cast<uint32_t *>(base + o_offsets);
    mPairs = reinterpret_cast<uint2 *>(base + o_pairs);
    mNow = reinterpret_cast<ChainSums *>(base + o_now);
    mResult = base + o_result;
    hip_check(hipHostMalloc(&mHostResult, sizeof(DeviceResult), hipHostMallocDefault), "hipHostMalloc (analyzer)");
    if (C) {
        hip_check(hipMemcpy(mMembers, mChains.members.data(), sizeof(uint32_t) * mChains.members.size(),
                            hipMemcpyHostToDevice),
                  "upload chain members");
        hip_check(hipMemcpy(mOffsets, mChains.offsets.data(), sizeof(uint32_t) * (size_t(C) + 1),
                            hipMemcpyHostToDevice),
                  "upload chain offsets");
    }
    if (!pairs.empty()) {
        std::vector<uint2> p(pairs.size());
        for (size_t k = 0; k < pairs.size(); ++k)
            p[k] = make_uint2(pairs[k].first, pairs[k].second);
        hip_check(hipMemcpy(mPairs, p.data(), sizeof(uint2) * p.size(), hipMemcpyHostToDevice), "upload bonds");
    }
    mPrepared = true;
}

void DeviceAnalyzer::computeChains(const HostSystem &s, const SweepBuffers &buf, hipStream_t stream)
{
    const dim3 block(kAnaBlock);
    hipLaunchKernelGGL(k_positions, dim3((s.n_pad + kAnaBlock - 1) / kAnaBlock), block, 0, stream, buf.ppos,
                       buf.ids, buf.image, s.n_pad, buf.geom, int64_t(s.box[0]), int64_t(s.box[1]),
                       int64_t(s.box[2]), mPos);
    hip_check(hipGetLastError(), "analyzer positions");
    const uint32_t C = mChains.linear ? mChains.n_chains : 0;
    if (C) {
        const uint32_t waves = kAnaBlock / 64;
        hipLaunchKernelGGL(k_chain_sums, dim3((C + waves - 1) / waves), block, 0, stream, mMembers, mOffsets, C,
                           mPos, mNow);
        hip_check(hipGetLastError(), "analyzer chain sums");
        hipLaunchKernelGGL(k_chain_reduce, dim3(1), dim3(kReduceBlock), 0, stream, mNow, C,
                           static_cast<DeviceResult *>(mResult));
        hip_check(hipGetLastError(), "analyzer chain reduction");
    }
}

void DeviceAnalyzer::observe(const HostSystem &s, const SweepBuffers &buf, hipStream_t stream, Observables &out)
{
    prepare(s);
    DeviceResult *r = static_cast<DeviceResult *>(mResult);
    hipLaunchKernelGGL(k_clear_result, dim3(1), dim3(kHistBins), 0, stream, r);
    hip_check(hipGetLastError(), "analyzer clear");
    computeChains(s, buf, stream);
    if (mBonds) {
        const uint64_t blocks = std::min<uint64_t>(1024, (mBonds + kAnaBlock - 1) / kAnaBlock);
        hipLaunchKernelGGL(k_bond_hist, dim3(uint32_t(blocks)), dim3(kAnaBlock), 0, stream, mPairs, mBonds, mPos,
                           r->hist);
        hip_check(hipGetLastError(), "analyzer bond histogram");
    }
    hip_check(hipMemcpyAsync(mHostResult, r, sizeof(DeviceResult), hipMemcpyDeviceToHost, stream),
              "download observables");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize (observables)");
    const DeviceResult &h = *static_cast<const DeviceResult *>(mHostResult);
    out = Observables{};
    out.bonds = mBonds;
    std::memcpy(out.hist, h.hist, sizeof(out.hist));
    const uint32_t C = mChains.linear ? mChains.n_chains : 0;
    out.n_chains = C;
    if (C) {
        out.re2_sum = h.re2_sum;
        out.re2 = double(h.re2_sum) / double(C);
        out.rg2 = h.rg2_sum / double(C);
        for (int a = 0; a < 3; ++a)
            out.box_com[a] = h.box_com[a];
    } else {
        out.re2 = out.rg2 = std::nan("");
        for (int a = 0; a < 3; ++a)
            out.box_com[a] = std::nan("");
    }
}

void DeviceAnalyzer::chainSums(const HostSystem &s, const SweepBuffers &buf, hipStream_t stream,
                               std::vector<ChainSums> &out)
{
    prepare(s);
    if (!mChains.linear)
        throw std::invalid_argument(mChains.msd_error);
    computeChains(s, buf, stream);
    out.resize(mChains.n_chains);
    hip_check(hipMemcpyAsync(out.data(), mNow, sizeof(ChainSums) * out.size(), hipMemcpyDeviceToHost, stream),
              "download chain sums");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize (chain sums)");
}

void DeviceAnalyzer::ensureSnapshots(uint32_t count, hipStream_t stream)
{
    if (count <= mSnapCap)
        return;
    const uint32_t C = mChains.n_chains;
    const uint32_t cap = std::max(count, std::max(mSnapCap * 2, 4u));
    ChainSums *snap = nullptr;
    double *com = nullptr, *sums = nullptr;
    hip_check(hipMalloc(&snap, sizeof(ChainSums) * size_t(C) * cap), "hipMalloc (MSD snapshots)");
    hip_check(hipMalloc(&com, sizeof(double) * 3 * size_t(cap)), "hipMalloc (MSD snapshots)");
    hip_check(hipMalloc(&sums, sizeof(double) * 10 * size_t(cap)), "hipMalloc (MSD sums)");
    if (mSnapCap) {
        hip_check(hipMemcpyAsync(snap, mSnap, sizeof(ChainSums) * size_t(C) * mSnapCap, hipMemcpyDeviceToDevice,
                                 stream),
                  "copy MSD snapshots");
        hip_check(hipMemcpyAsync(com, mSnapCom, sizeof(double) * 3 * size_t(mSnapCap), hipMemcpyDeviceToDevice,
                                 stream),
                  "copy MSD snapshots");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize (MSD snapshots)");
        hipFree(mSnap);
        hipFree(mSnapCom);
        hipFree(mSums);
    }
    mSnap = snap;
    mSnapCom = com;
    mSums = sums;
    mSnapCap = cap;
}

void DeviceAnalyzer::msdCreate(const HostSystem &s, const SweepBuffers &buf, hipStream_t stream, uint64_t age,
                               uint64_t mcs_to_skip, uint64_t delta_t_realization, uint64_t delta_t_evaluate,
                               uint64_t max_subsequences)
{
    if (delta_t_evaluate == 0)
        throw std::invalid_argument("the MSD evaluation interval must be positive");
    prepare(s);
    if (!mChains.msd_error.empty())
        throw std::invalid_argument(mChains.msd_error);
    mMsd = true;
    mMcsToSkip = mcs_to_skip;
    mDtReal = delta_t_realization;
    mDtEval = delta_t_evaluate;
    mLastEvaluation = uint64_t(0) - delta_t_evaluate;   // AnalyzerMsd.tpp:155: evaluates at the first call
    mLastRealization = 0;
    mAnyRealization = false;
    mFirstMcs.clear();
    mMsds.clear();
    mEvalTimes.clear();
    // the reference reserves rnMaxSubsequences (:158-160); here up to 16, more on demand
    ensureSnapshots(uint32_t(std::min<uint64_t>(std::max<uint64_t>(max_subsequences, 1), 16)), stream);
    msdExecute(s, buf, stream, age);
}

void DeviceAnalyzer::msdExecute(const HostSystem &s, const SweepBuffers &buf, hipStream_t stream, uint64_t age)
{
    if (!mMsd)
        throw std::runtime_error("no MSD analyzer attached");
    if (age - mLastEvaluation < mDtEval)
        return;
    mLastEvaluation = age;
    computeChains(s, buf, stream);
    const uint32_t C = mChains.n_chains;
    const size_t n_sub = mFirstMcs.size();
    // The reference leaves the time of the last realization uninitialised
    // (AnalyzerMsd.tpp:107,150-160); here the first evaluation starts one.
    if (!mAnyRealization || age - mLastRealization >= mDtReal) {
        mAnyRealization = true;
        mLastRealization = age;
        ensureSnapshots(uint32_t(n_sub + 1), stream);
        hip_check(hipMemcpyAsync(mSnap + n_sub * C, mNow, sizeof(ChainSums) * C, hipMemcpyDeviceToDevice, stream),
                  "MSD snapshot");
        hip_check(hipMemcpyAsync(mSnapCom + 3 * n_sub, static_cast<DeviceResult *>(mResult)->box_com,
                                 sizeof(double) * 3, hipMemcpyDeviceToDevice, stream),
                  "MSD snapshot");
        mFirstMcs.push_back(age);
    }
    for (size_t k = 0; k < n_sub; ++k) {
        const uint64_t e = (age - mFirstMcs[k]) / mDtEval;
        if (e >= mMsds.size()) {
            while (e >= mMsds.size())
                mMsds.emplace_back(5);
            while (e >= mEvalTimes.size())
                mEvalTimes.push_back(0);
            mEvalTimes[e] = age - mFirstMcs[k];
        }
    }
    if (n_sub == 0) {
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize (MSD)");
        return;
    }
    hipLaunchKernelGGL(k_msd, dim3(uint32_t(n_sub)), dim3(kMsdBlock), 0, stream, mNow, mSnap, mSnapCom,
                       static_cast<const DeviceResult *>(mResult), C, mSums);
    hip_check(hipGetLastError(), "MSD kernel");
    std::vector<double> sums(10 * n_sub);
    hip_check(hipMemcpyAsync(sums.data(), mSums, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, stream),
              "download MSD sums");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize (MSD)");
    for (size_t k = 0; k < n_sub; ++k) {
        const uint64_t e = (age - mFirstMcs[k]) / mDtEval;
        for (int g = 0; g < 5; ++g)
            mMsds[e][g].add(double(C), sums[10 * k + 2 * g], sums[10 * k + 2 * g + 1]);
    }
}

}  // namespace pscbfm
