// batch.cpp -- many systems swept together on one device.
#include "batch.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "device_slab.hpp"
#include "updater.hpp"

namespace pscbfm {

PinnedPool::~PinnedPool()
{
    for (Buf &b : mBufs) {
        if (b.done) {
            hipEventSynchronize(b.done);
            hipEventDestroy(b.done);
        }
        if (b.p)
            hipHostFree(b.p);
    }
}

size_t PinnedPool::take(size_t bytes)
{
    // the smallest free buffer that fits; a buffer's event is queried only
    // while its copy is pending (once seen complete it stays free)
    size_t best = mBufs.size();
    for (size_t k = 0; k < mBufs.size(); ++k) {
        Buf &b = mBufs[k];
        if (b.held || b.bytes < bytes || (best < mBufs.size() && mBufs[best].bytes <= b.bytes))
            continue;
        if (b.pending && hipEventQuery(b.done) == hipSuccess)
            b.pending = false;
        if (!b.pending)
            best = k;
    }
    if (best == mBufs.size()) {
        Buf b;
        b.bytes = std::max<size_t>(size_t(1) << 16, bytes);
        while (b.bytes & (b.bytes - 1))
            b.bytes += b.bytes & ~(b.bytes - 1);   // next power of two
        hip_check(hipHostMalloc(&b.p, b.bytes, hipHostMallocDefault), "hipHostMalloc (staging)");
        hip_check(hipEventCreateWithFlags(&b.done, hipEventDisableTiming), "hipEventCreate (staging)");
        mBufs.push_back(b);
    }
    mBufs[best].held = true;
    return best;
}

void PinnedPool::record(size_t k, hipStream_t stream)
{
    hip_check(hipEventRecord(mBufs[k].done, stream), "hipEventRecord (staging)");
    mBufs[k].pending = true;
}

bool PinnedPool::ready(size_t k) const
{
    const hipError_t e = hipEventQuery(mBufs[k].done);
    if (e == hipErrorNotReady)
        return false;
    hip_check(e, "hipEventQuery (staging)");
    return true;
}

void PinnedPool::wait(size_t k) const
{
    hip_check(hipEventSynchronize(mBufs[k].done), "hipEventSynchronize (staging)");
}

BatchEngine::BatchEngine(int gpu) : mGpu(gpu)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice || (e == hipSuccess && n == 0))
        throw NoDeviceError("no HIP device is visible");
    hip_check(e, "hipGetDeviceCount");
    if (gpu < 0 || gpu >= n)
        throw std::invalid_argument("batch: GPU id out of range");
    hip_check(hipSetDevice(gpu), "hipSetDevice");
    hip_check(hipStreamCreate(&mStream), "hipStreamCreate");
    // members come and go: allocate their slabs in stream order from the
    // device's pool and keep freed memory in the pool for the next ones
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, gpu) == hipSuccess && pool) {
        uint64_t keep = ~0ull;
        hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        mPooled = true;
    }
    hipGetLastError();
}

BatchEngine::~BatchEngine()
{
    hipSetDevice(mGpu);
    if (mStream)
        hipStreamSynchronize(mStream);
    for (auto &g : mGraphs)
        hipGraphExecDestroy(g.exec);
    for (int b = 0; b < 2; ++b) {
        if (mDone[b])
            hipEventDestroy(mDone[b]);
        if (mHostChunk[b])
            hipHostFree(mHostChunk[b]);

    }
    if (mDevChunk)
        hipFree(mDevChunk);
    if (mDevSched)
        hipFree(mDevSched);
    for (auto &m : mMembers)
        if (m && m->slab)
            hipFree(m->slab);
    for (auto &f : mFreeSlabs)
        hipFree(f.second);
    if (mDevSys)
        hipFree(mDevSys);
    for (uint32_t c = 0; c < kClasses; ++c) {
        if (mDevCheckBlk[c])
            hipFree(mDevCheckBlk[c]);
        if (mDevPerformBlk[c])
            hipFree(mDevPerformBlk[c]);
    }
    if (mStream)
        hipStreamDestroy(mStream);
}

void *BatchEngine::allocate(size_t bytes, const char *what)
{
    void *p = nullptr;
    hipError_t e = mPooled ? hipMallocAsync(&p, bytes, mStream) : hipMalloc(&p, bytes);
    if (e != hipSuccess)
        hip_check(e, what);
    return p;
}

void BatchEngine::release(void *p)
{
    if (!p)
        return;
    if (mPooled)
        hip_check(hipFreeAsync(p, mStream), "hipFreeAsync");
    else
        hip_check(hipFree(p), "hipFree");
}

size_t BatchEngine::bytesFor(const HostSystem &sys) { return slab_layout(sys).total; }

// Size class of a slab: rounded up to 1/8 of the next power of two (at most
// 12.5 % over), so that members of similar size share freed slabs.
static size_t slab_class(size_t bytes)
{
    size_t p = 4096;
    while (p < bytes)
        p <<= 1;
    const size_t step = std::max<size_t>(p / 8, 4096);
    return (bytes + step - 1) / step * step;
}

void *BatchEngine::takeSlab(size_t bytes, size_t &klass)
{
    klass = slab_class(bytes);
    for (size_t k = mFreeSlabs.size(); k-- > 0;)
        if (mFreeSlabs[k].first == klass) {
            void *p = mFreeSlabs[k].second;
            mFreeSlabs[k] = mFreeSlabs.back();
            mFreeSlabs.pop_back();
            return p;
        }
    return allocate(klass, "allocation of a batch member");
}

void BatchEngine::giveSlab(void *slab, size_t klass)
{
    // freed slabs kept beyond 1024 go back to the device pool
    if (mFreeSlabs.size() >= 1024) {
        release(mFreeSlabs.front().second);
        mFreeSlabs.erase(mFreeSlabs.begin());
    }
    mFreeSlabs.emplace_back(klass, slab);
}

BatchEngine::Member &BatchEngine::member(uint32_t id)
{
    if (id >= mMembers.size() || !mMembers[id])
        throw std::invalid_argument("batch: no such member (never added, or released)");
    return *mMembers[id];
}

const BatchEngine::Member &BatchEngine::member(uint32_t id) const
{
    if (id >= mMembers.size() || !mMembers[id])
        throw std::invalid_argument("batch: no such member (never added, or released)");
    return *mMembers[id];
}

template <class T>
static void drop(std::vector<T> &v)
{
    std::vector<T>().swap(v);
}

template <class T>
static size_t held(const std::vector<T> &v)
{
    return sizeof(T) * v.capacity();
}

size_t BatchEngine::hostBytes() const
{
    size_t b = 0;
    for (const auto &m : mMembers)
        if (m) {
            const HostSystem &s = m->sys;
            b += held(s.xyz) + held(s.attribute) + held(s.nnbr) + held(s.nbr) + held(s.colour) + held(s.seg_begin) +
                 held(s.seg_count) + held(s.to_sorted) + held(s.ppos_sorted) + held(s.ids_sorted) +
                 held(s.image_sorted) + held(s.nbr_sorted) + held(s.lattice);
        }
    return b;
}

uint32_t BatchEngine::liveMembers() const
{
    uint32_t k = 0;
    for (const auto &m : mMembers)
        k += m ? 1u : 0u;
    return k;
}

uint32_t BatchEngine::add(HostSystem &&sys, uint64_t seed, int64_t age, bool keep_topology, bool may_sort)
{
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    auto m = std::make_unique<Member>();
    m->sys = std::move(sys);
    m->seed = seed;
    m->age = age;
    const SlabLayout L = slab_layout(m->sys);
    m->slab = takeSlab(L.total, m->slab_class);
    m->bytes = L.total;
    m->hot_bytes = L.bytes[kPartPPos] + L.bytes[kPartIds] + (L.upload - L.offset[kPartNbr]) +
                   L.bytes[kPartLattice] + L.bytes[kPartClaim0] + L.bytes[kPartClaim1];
    try {
        // staged through pinned memory, so that admitting a member never
        // waits for the chunks already queued
        const size_t k = mPinned.take(upload_bytes(m->sys));
        m->dev.buf = upload_slab(m->sys, m->slab, L, seed, mStream, mPinned.data(k));
        mPinned.record(k, mStream);
        mPinned.give(k);
    } catch (...) {
        giveSlab(m->slab, m->slab_class);
        throw;
    }
    // The upload went through pinned staging, so the host copies are done
    // with: a member keeps only its sizes, geometry and colour segments, and
    // its bond lists for the analyzers (downloads are unsorted and unwrapped
    // on the device; positions come back into arrays sized when they arrive).
    if (!keep_topology) {
        drop(m->sys.nnbr);
        drop(m->sys.nbr);
    }
    drop(m->sys.to_sorted);
    drop(m->sys.ppos_sorted);
    drop(m->sys.ids_sorted);
    drop(m->sys.image_sorted);
    drop(m->sys.nbr_sorted);
    drop(m->sys.lattice);
    drop(m->sys.attribute);
    drop(m->sys.xyz);
    drop(m->sys.colour);
    if (m->sys.spatial_order && m->sys.steps_between_sortings > 0 && may_sort) {
        m->sort_period = m->sys.steps_between_sortings;
        m->last_sort_age = age;   // laid out in spatial order by plan()
        mResorter.reserve(m->sys, L);
    }
    m->dev.seg_begin = static_cast<const uint32_t *>(slab_part(m->slab, L, kPartSegBegin));
    m->dev.seg_count = static_cast<const uint32_t *>(slab_part(m->slab, L, kPartSegCount));
    m->dev.n_colours = m->sys.n_colours;
    m->check_blocks_max = check_blocks(std::max<uint32_t>(m->sys.max_segment, 1));
    mBytes += L.total;
    mMembers.push_back(std::move(m));
    return uint32_t(mMembers.size() - 1);
}

void BatchEngine::remove(uint32_t id)
{
    Member &m = member(id);
    if (!m.slab)
        return;
    // stream order: launches already queued finish before the memory is reused
    giveSlab(m.slab, m.slab_class);
    m.slab = nullptr;
    m.analyzer.reset();
    mBytes -= m.bytes;
    mCompIds.clear();   // forces a rebuild of the tables
}

void BatchEngine::forget(uint32_t id)
{
    if (id >= mMembers.size() || !mMembers[id])
        return;
    if (mMembers[id]->slab)
        remove(id);
    mMembers[id].reset();
}

void BatchEngine::observe(uint32_t id, Observables &out)
{
    Member &m = member(id);
    if (!m.slab)
        throw std::invalid_argument("batch: member was removed");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    if (!m.analyzer)
        m.analyzer = std::make_unique<DeviceAnalyzer>();
    m.analyzer->observe(m.sys, m.dev.buf, mStream, out);
}

void BatchEngine::setPerformGroup(uint32_t group)
{
    if (!valid_perform_group(group))
        throw std::invalid_argument("perform group must be 2, 4 or 8");
    if (group != mPerformGroup) {
        mPerformGroup = group;
        mCompIds.clear();   // the perform block tables depend on it
    }
}

void BatchEngine::autotune(uint32_t probe_mcs, uint32_t chunk)
{
    if (probe_mcs == 0 || chunk == 0)
        throw std::invalid_argument("autotune: probe_mcs and chunk must be positive");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    const uint32_t group0 = mPerformGroup;
    const bool xcd0 = mXcdRanges;
    uint32_t best_group = group0;
    bool best_xcd = xcd0;
    try {
        // one untimed probe first, so that no configuration pays for cold caches
        run(probe_mcs, chunk);
        float ms[2][3] = {};
        for (int x = 0; x < 2; ++x)
            for (int k = 0; k < 3; ++k) {
                setXcdRanges(x == 0);
                setPerformGroup(kPerformGroups[k]);
                hip_check(hipEventRecord(e0, mStream), "hipEventRecord");
                run(probe_mcs, chunk);
                hip_check(hipEventRecord(e1, mStream), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                hip_check(hipEventElapsedTime(&ms[x][k], e0, e1), "hipEventElapsedTime");
            }
        // the configuration in force unless another is faster by more than 2 %
        // (a choice made on run-to-run noise can be slower than the default)
        int bx = xcd0 ? 0 : 1, bk = 0;
        while (bk < 2 && kPerformGroups[bk] != group0)
            ++bk;
        const float incumbent = ms[bx][bk];
        for (int x = 0; x < 2; ++x)
            for (int k = 0; k < 3; ++k)
                if (ms[x][k] < 0.98f * incumbent && ms[x][k] < ms[bx][bk]) {
                    bx = x;
                    bk = k;
                }
        best_group = kPerformGroups[bk];
        best_xcd = bx == 0;
    } catch (...) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        setPerformGroup(group0);
        setXcdRanges(xcd0);
        throw;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    setPerformGroup(best_group);
    setXcdRanges(best_xcd);
}

void BatchEngine::synchronize()
{
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    hip_check(hipStreamSynchronize(mStream), "hipStreamSynchronize (batch)");
}

template <class T>
void BatchEngine::ensureDevice(T *&p, size_t &cap, size_t n, const char *what)
{
    if (n <= cap && p)
        return;
    release(p);
    p = nullptr;
    cap = std::max<size_t>(n, 1);
    p = static_cast<T *>(allocate(sizeof(T) * cap, what));
    ++mTableGen;   // graphs captured with the old buffer are stale
}

// Capacity a table that holds `need` entries keeps: its current one while the
// need lies in [cap/2, cap], else another (decided by the caller).
static uint32_t table_capacity(uint32_t need, uint32_t cap)
{
    if (need == 0 || cap == 0)
        return need == 0 ? 0u : ~0u;
    return need <= cap && 2 * uint64_t(need) >= cap ? cap : ~0u;
}

// The block table of one class at capacity: with XCD ranges, the real
// blocks are dealt to the 8 XCDs' logical ranges in proportion to the range
// lengths (each XCD sweeps a contiguous run of them, its idle entries last);
// without, they fill the prefix.  An idle entry exits at its first test.
static void lay_out_blocks(const std::vector<BlockEntry> &real, uint32_t cap, bool xcd_ranges,
                           std::vector<BlockEntry> &out)
{
    out.assign(cap, BlockEntry{0u, 0xFFFFFFFFu});
    const uint64_t n = real.size();
    if (!xcd_ranges) {
        std::copy(real.begin(), real.end(), out.begin());
        return;
    }
    const uint32_t per = cap / kXcds, rem = cap % kXcds;
    uint32_t start = 0, placed = 0;
    uint64_t through = 0;
    for (uint32_t x = 0; x < kXcds; ++x) {
        const uint32_t len = per + (x < rem ? 1u : 0u);
        through += len;
        const uint32_t upto = uint32_t(n * through / cap);   // real blocks placed through XCD x
        for (uint32_t k = placed; k < upto; ++k)
            out[start + (k - placed)] = real[k];
        placed = upto;
        start += len;
    }
}

void BatchEngine::rebuildTables(const std::vector<uint32_t> &ids)
{
    // Launches already queued still read the device tables; the copies below
    // are queued behind them on the same stream, from pinned staging, so the
    // host never waits here for the device.
    std::vector<BatchSystem> &sys = mHostSys;
    std::vector<BlockEntry> real_c[kClasses], real_p[kClasses];
    sys.clear();
    mMaxColours = 0;
    for (uint32_t j = 0; j < ids.size(); ++j) {
        Member &m = member(ids[j]);
        if (!m.slab)
            throw std::invalid_argument("batch: member was removed");
        if (m.spent)
            throw std::logic_error("batch: member ran after its final download");
        const uint32_t cls = kernel_class(m.sys.geom);
        for (uint32_t b = 0; b < m.check_blocks_max; ++b)
            real_c[cls].push_back(BlockEntry{j, b});
        const uint32_t pb = perform_blocks(m.check_blocks_max, mPerformGroup);
        for (uint32_t b = 0; b < pb; ++b)
            real_p[cls].push_back(BlockEntry{j, b});
        sys.push_back(m.dev);
        mMaxColours = std::max(mMaxColours, m.sys.n_colours);
    }
    // The three tables are resized together (so that a growing population
    // changes the launch shape once, not three times): when any need leaves
    // its band [cap/2, cap], every table takes the need plus 1/2 (the exact
    // need the first time).
    bool resize = table_capacity(uint32_t(sys.size()), mSysCapN) != mSysCapN;
    for (uint32_t c = 0; c < kClasses; ++c)
        resize = resize || table_capacity(uint32_t(real_c[c].size()), mCheckCapN[c]) != mCheckCapN[c] ||
                 table_capacity(uint32_t(real_p[c].size()), mPerformCapN[c]) != mPerformCapN[c];
    if (resize) {
        const bool first = mSysCapN == 0;
        auto fit = [first](size_t need) {
            const uint64_t c = first ? need : need + need / 2;
            return need ? uint32_t((c + 63) / 64 * 64) : 0u;
        };
        mSysCapN = fit(sys.size());
        for (uint32_t c = 0; c < kClasses; ++c) {
            mCheckCapN[c] = fit(real_c[c].size());
            mPerformCapN[c] = fit(real_p[c].size());
        }
    }
    // idle systems: no colours, so the schedule gives them no work
    BatchSystem idle = sys.front();
    idle.n_colours = 0;
    sys.resize(mSysCapN, idle);
    ensureDevice(mDevSys, mSysCap, sys.size(), "allocation of the batch systems");
    size_t bytes = sizeof(BatchSystem) * sys.size();
    for (uint32_t c = 0; c < kClasses; ++c) {
        if (!mCheckCapN[c])
            continue;
        lay_out_blocks(real_c[c], mCheckCapN[c], mXcdRanges, mHostCheckBlk[c]);
        lay_out_blocks(real_p[c], mPerformCapN[c], mXcdRanges, mHostPerformBlk[c]);
        ensureDevice(mDevCheckBlk[c], mCheckCap[c], mCheckCapN[c], "allocation of the batch blocks");
        ensureDevice(mDevPerformBlk[c], mPerformCap[c], mPerformCapN[c], "allocation of the batch blocks");
        bytes += sizeof(BlockEntry) * (size_t(mCheckCapN[c]) + mPerformCapN[c]);
    }
    const size_t k = mPinned.take(bytes);
    char *at = static_cast<char *>(mPinned.data(k));
    auto stage = [&](void *dst, const void *src, size_t n, const char *what) {
        std::memcpy(at, src, n);
        hip_check(hipMemcpyAsync(dst, at, n, hipMemcpyHostToDevice, mStream), what);
        at += n;
    };
    stage(mDevSys, sys.data(), sizeof(BatchSystem) * sys.size(), "upload batch systems");
    for (uint32_t c = 0; c < kClasses; ++c) {
        if (!mCheckCapN[c])
            continue;
        stage(mDevCheckBlk[c], mHostCheckBlk[c].data(), sizeof(BlockEntry) * mCheckCapN[c], "upload batch blocks");
        stage(mDevPerformBlk[c], mHostPerformBlk[c].data(), sizeof(BlockEntry) * mPerformCapN[c],
              "upload batch blocks");
    }
    mPinned.record(k, mStream);
    mPinned.give(k);
    mCompIds = ids;
}

void BatchEngine::runChunk(uint32_t mcs)
{
    const uint32_t n_real = uint32_t(mCompIds.size());
    const uint32_t n_sys = mSysCapN;   // the tables' capacity; systems past n_real idle
    const int b = mBuf;
    mBuf ^= 1;

    // the host buffer b was last read by the copy of chunk k-2
    if (mDone[b])
        hip_check(hipEventSynchronize(mDone[b]), "hipEventSynchronize (chunk table)");
    else
        hip_check(hipEventCreateWithFlags(&mDone[b], hipEventDisableTiming), "hipEventCreate");
    if (n_sys > mHostChunkCap[b]) {
        if (mHostChunk[b])
            hip_check(hipHostFree(mHostChunk[b]), "hipHostFree");
        mHostChunk[b] = nullptr;
        hip_check(hipHostMalloc(&mHostChunk[b], sizeof(ChunkEntry) * n_sys, hipHostMallocDefault),
                  "hipHostMalloc (chunk table)");
        mHostChunkCap[b] = n_sys;
    }
    ensureDevice(mDevChunk, mChunkCap, n_sys, "allocation of a chunk table");

    // where every member starts this chunk; the kernels derive colour, parity
    // and substep seed of each slot from it
    ChunkEntry *chunk = mHostChunk[b];
    for (uint32_t j = 0; j < n_real; ++j) {
        Member &M = *mMembers[mCompIds[j]];
        chunk[j].age = uint64_t(M.age);
        chunk[j].substeps = M.substeps;
        M.age += mcs;
        M.substeps += uint64_t(mcs) * M.sys.n_colours;
    }
    for (uint32_t j = n_real; j < n_sys; ++j)
        chunk[j] = ChunkEntry{0, 0};
    hip_check(hipMemcpyAsync(mDevChunk, chunk, sizeof(ChunkEntry) * n_sys, hipMemcpyHostToDevice, mStream),
              "upload chunk table");
    const uint32_t slots = mcs * mMaxColours;
    const size_t entries = size_t(slots) * n_sys;
    ensureDevice(mDevSched, mSchedCap, entries, "allocation of a schedule");

    auto launch_all = [&]() {
        hip_check(launch_schedule(mDevSys, mDevChunk, n_sys, mcs, mMaxColours, mDevSched, mStream),
                  "schedule kernel");
        for (uint32_t slot = 0; slot < slots; ++slot)
            for (uint32_t c = 0; c < kClasses; ++c) {
                if (!mCheckCapN[c])
                    continue;
                hip_check(launch_check_batch(c, mDevSys, mDevCheckBlk[c], mCheckCapN[c], mDevSched, n_sys, slot,
                                             mXcdRanges ? 1u : 0u, mStream),
                          "batched check kernel");
                hip_check(launch_perform_batch(c, mDevSys, mDevPerformBlk[c], mPerformCapN[c], mDevSched, n_sys,
                                               slot, mXcdRanges ? 1u : 0u, mStream, mPerformGroup),
                          "batched perform kernel");
            }
    };
    uint32_t classes = 0;
    for (uint32_t c = 0; c < kClasses; ++c)
        classes += mCheckCapN[c] ? 1u : 0u;
    mKernelLaunches += 2ull * slots * classes + 1;
    if (!mUseGraphs) {
        launch_all();
        hip_check(hipEventRecord(mDone[b], mStream), "hipEventRecord");
        return;
    }
    // A chunk's launches depend on the table capacities, the chunk length,
    // the colour slots and the launch configuration, not on which members
    // fill the tables: one graph per such shape.
    uint64_t key[3] = {(uint64_t(mcs) << 32) | (uint64_t(mMaxColours) << 16) | (uint64_t(mPerformGroup) << 8) |
                           (mXcdRanges ? 1u : 0u),
                       uint64_t(n_sys), 0};
    for (uint32_t c = 0; c < kClasses; ++c)
        key[2] = key[2] * 1000003u + (uint64_t(mCheckCapN[c]) << 24) + mPerformCapN[c];
    CachedGraph *g = nullptr;
    for (auto &x : mGraphs)
        if (x.gen == mTableGen && x.key[0] == key[0] && x.key[1] == key[1] && x.key[2] == key[2])
            g = &x;
    if (!g) {
        hipGraph_t graph = nullptr;
        hip_check(hipStreamBeginCapture(mStream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            launch_all();
        } catch (...) {
            hipStreamEndCapture(mStream, &graph);
            if (graph)
                hipGraphDestroy(graph);
            throw;
        }
        hip_check(hipStreamEndCapture(mStream, &graph), "hipStreamEndCapture");
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        hip_check(e, "hipGraphInstantiate");
        ++mGraphInstantiations;
        // Drop stale graphs, then the least recently used beyond the cache
        // size.  A graph's last launch may be chunk k-1, still queued or
        // running (chunk k-2 has finished, its event was waited on above):
        // its executable must outlive it, so wait for chunk k-1 before
        // destroying any.
        const bool evict = mGraphs.size() >= kGraphCache ||
                           std::any_of(mGraphs.begin(), mGraphs.end(),
                                       [&](const CachedGraph &x) { return x.gen != mTableGen; });
        if (evict && mDone[b ^ 1])
            hip_check(hipEventSynchronize(mDone[b ^ 1]), "hipEventSynchronize (graph in flight)");
        for (size_t k = mGraphs.size(); k-- > 0;)
            if (mGraphs[k].gen != mTableGen) {
                hip_check(hipGraphExecDestroy(mGraphs[k].exec), "hipGraphExecDestroy");
                mGraphs.erase(mGraphs.begin() + long(k));
            }
        if (mGraphs.size() >= kGraphCache) {
            size_t old = 0;
            for (size_t k = 1; k < mGraphs.size(); ++k)
                if (mGraphs[k].used < mGraphs[old].used)
                    old = k;
            hip_check(hipGraphExecDestroy(mGraphs[old].exec), "hipGraphExecDestroy");
            mGraphs.erase(mGraphs.begin() + long(old));
        }
        CachedGraph x;
        x.exec = exec;
        std::copy(key, key + 3, x.key);
        x.gen = mTableGen;
        mGraphs.push_back(x);
        g = &mGraphs.back();
    }
    g->used = ++mUse;
    hip_check(hipGraphLaunch(g->exec, mStream), "hipGraphLaunch");
    ++mGraphLaunches;
    hip_check(hipEventRecord(mDone[b], mStream), "hipEventRecord");
}

void BatchEngine::runMembers(const std::vector<uint32_t> &ids, uint32_t n_mcs, uint32_t chunk)
{
    if (ids.empty() || n_mcs == 0)
        return;
    if (chunk == 0)
        throw std::invalid_argument("batch: chunk must be positive");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    if (ids != mCompIds)
        rebuildTables(ids);
    for (uint32_t done = 0; done < n_mcs;) {
        // a chunk ends where a member's next re-sort falls (rarely: every
        // steps_between_sortings MCS), so each member sorts at the same ages
        // as it would alone
        const uint32_t m = sortDue(std::min(chunk, n_mcs - done));
        runChunk(m);
        done += m;
    }
}

uint32_t BatchEngine::sortDue(uint32_t limit)
{
    uint32_t until = limit;
    for (const uint32_t id : mCompIds) {
        Member &M = *mMembers[id];
        if (M.sort_period <= 0)
            continue;
        const int64_t P = M.sort_period;
        if (M.age % P == 0 && M.age != M.last_sort_age) {
            mResorter.run(M.sys, M.slab, slab_layout(M.sys), M.dev.buf, mStream);
            M.last_sort_age = M.age;
            ++M.sorts;
        }
        const int64_t next = (M.age / P + 1) * P;   // the next multiple after its age
        until = uint32_t(std::min<int64_t>(until, next - M.age));
    }
    return until;
}

std::vector<std::vector<uint32_t>> BatchEngine::workingSetGroups(const std::vector<uint32_t> &ids) const
{
    size_t hot = 0;
    uint64_t monomers = 0;
    for (uint32_t id : ids) {
        hot += member(id).hot_bytes;
        monomers += member(id).sys.n;
    }
    uint64_t groups = 1;
    if (mWorkingSetBytes > 0 && hot > mWorkingSetBytes) {
        groups = (hot + mWorkingSetBytes - 1) / mWorkingSetBytes;
        groups = std::min<uint64_t>(groups, std::max<uint64_t>(1, monomers / kMinGroupMonomers));
        groups = std::min<uint64_t>(groups, ids.size());
    }
    if (groups <= 1)
        return {ids};
    // largest members first, each into the group with the fewest hot bytes;
    // every group keeps the members' order
    std::vector<uint32_t> order(ids);
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return member(a).hot_bytes > member(b).hot_bytes; });
    std::vector<std::vector<uint32_t>> out(groups);
    std::vector<size_t> load(groups, 0);
    for (uint32_t id : order) {
        const size_t g = size_t(std::min_element(load.begin(), load.end()) - load.begin());
        out[g].push_back(id);
        load[g] += member(id).hot_bytes;
    }
    for (auto &g : out)
        std::sort(g.begin(), g.end());
    return out;
}

void BatchEngine::run(uint32_t n_mcs, uint32_t chunk)
{
    std::vector<uint32_t> ids;
    for (uint32_t i = 0; i < mMembers.size(); ++i)
        if (active(i))
            ids.push_back(i);
    for (const auto &g : workingSetGroups(ids))
        runMembers(g, n_mcs, chunk);
}

void BatchEngine::copyBack(uint32_t id)
{
    Download d = download(id);
    wait(d);
    HostSystem &s = member(id).sys;
    s.xyz.resize(3 * size_t(s.n));
    copyOut(d, s.xyz.data());
    give(d);
}

BatchEngine::Download BatchEngine::download(uint32_t id, bool final)
{
    Member &m = member(id);
    if (!m.slab)
        throw std::invalid_argument("batch: member was removed");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    const HostSystem &s = m.sys;
    const size_t bytes = sizeof(int32_t) * 3 * size_t(s.n);
    Download d;
    d.n = s.n;
    d.buf = mPinned.take(std::max<size_t>(bytes, 16));
    d.held = true;
    if (s.n) {
        // unsort and unwrap on the device (as the plain copy-back does), then
        // one copy of 12 bytes per monomer into pinned memory
        // a final download stages in the slab from the lattice on (the
        // partner table behind it alone holds 32 bytes per slot); the unwrap
        // reads only positions, ids and images, which lie before it
        const SlabLayout L = slab_layout(s);
        int32_t *xyz = final ? static_cast<int32_t *>(slab_part(m.slab, L, kPartLattice))
                             : static_cast<int32_t *>(allocate(bytes, "allocation of a download buffer"));
        const uint32_t box[3] = {uint32_t(s.box[0]), uint32_t(s.box[1]), uint32_t(s.box[2])};
        hip_check(launch_unwrap(m.dev.buf, box, xyz, mStream), "unwrap kernel");
        hip_check(hipMemcpyAsync(mPinned.data(d.buf), xyz, bytes, hipMemcpyDeviceToHost, mStream),
                  "download positions");
        if (final)
            m.spent = true;
        else
            release(xyz);
    }
    mPinned.record(d.buf, mStream);
    return d;
}

void BatchEngine::copyOut(const Download &d, int32_t *xyz) const
{
    if (!d.held)
        throw std::logic_error("batch: download already given back");
    std::memcpy(xyz, mPinned.data(d.buf), sizeof(int32_t) * 3 * size_t(d.n));
}

void BatchEngine::give(Download &d)
{
    if (d.held)
        mPinned.give(d.buf);
    d.held = false;
}

void BatchEngine::counters(uint32_t id, uint64_t &accepted, uint64_t &proposals)
{
    Member &m = member(id);
    accepted = proposals = 0;
    if (!m.slab)
        return;
    synchronize();
    std::vector<unsigned long long> slots(2 * size_t(kStatSlots));
    hip_check(hipMemcpy(slots.data(), m.dev.buf.accepted, sizeof(unsigned long long) * kStatSlots, hipMemcpyDeviceToHost),
              "download counters");
    hip_check(hipMemcpy(slots.data() + kStatSlots, m.dev.buf.proposals, sizeof(unsigned long long) * kStatSlots,
                        hipMemcpyDeviceToHost),
              "download counters");
    for (uint32_t k = 0; k < kStatSlots; ++k) {
        accepted += slots[k];
        proposals += slots[kStatSlots + k];
    }
}

}  // namespace pscbfm
