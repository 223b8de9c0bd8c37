// batch.hpp -- many independent systems swept together on one device.
//
// Every member keeps its own device slab (positions, partners, lattices,
// survivor and collision lists: the same layout as one plain updater).  A
// substep slot of a chunk is one k_check_batch and one k_perform_batch launch
// over all members at once.  The host writes only where each member stands at
// the start of the chunk (age, substeps); a small kernel at the head of the
// chunk derives every member's colour, parity and substep seed for every slot
// (one thread per member and MCS), which the sweep kernels read.  A chunk
// runs as one hipGraph launch once its composition has a graph: a composition
// that ran before replays or captures one, and a new composition is written
// into the existing graph of the same shape by an update (new arguments and
// grid sizes) instead of a new instantiation.
//
// A member's trajectory depends only on its own seed and age (the RNG is keyed
// on seed, MCS and monomer id), so it is the same alone, in any batch and in
// any chunking -- the plain updater's trajectory, bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "analyzers.hpp"
#include "bfm_kernels.hpp"
#include "host_system.hpp"
#include "resort.hpp"

namespace pscbfm {

// Pinned host buffers for asynchronous copies.  A buffer is held by its user
// until given back; a given-back buffer is reused once the work recorded on
// its event has run.
class PinnedPool {
public:
    PinnedPool() = default;
    ~PinnedPool();
    PinnedPool(const PinnedPool &) = delete;
    PinnedPool &operator=(const PinnedPool &) = delete;

    size_t take(size_t bytes);   // index of a free buffer of at least `bytes`
    void *data(size_t k) const { return mBufs[k].p; }
    void record(size_t k, hipStream_t stream);   // marks the work queued so far
    bool ready(size_t k) const;                  // that work has run
    void wait(size_t k) const;
    void give(size_t k) { mBufs[k].held = false; }
    size_t bytes(size_t k) const { return mBufs[k].bytes; }

private:
    struct Buf {
        void *p = nullptr;
        size_t bytes = 0;
        hipEvent_t done = nullptr;
        bool held = false;
        bool pending = false;   // recorded work not yet seen complete (queried only then)
    };
    std::vector<Buf> mBufs;
};

// Working-set budget of BatchEngine::run: half of the MI355X's 256 MB
// memory-side cache, measured best for the 512-system mixed batch (its 0.48 GB
// of hot arrays in 4 groups: 7.8e10 against 6.9e10 attempts/s in one, 6.9e10
// in 2 groups, 6.4e10 in 8, where the launches get small).
constexpr size_t kDefaultWorkingSetBytes = size_t(128) << 20;
// A group holds at least this many monomers, so that its launches fill the
// device (2 M monomers: 4,096 check waves per colour of 1 M).
constexpr uint64_t kMinGroupMonomers = 2u << 20;

class BatchEngine {
public:
    explicit BatchEngine(int gpu);
    ~BatchEngine();
    BatchEngine(const BatchEngine &) = delete;
    BatchEngine &operator=(const BatchEngine &) = delete;

    // Takes a planned system (HostSystem after plan()); allocates and uploads
    // its device slab at once.  Returns the member id.  Without
    // keep_topology the member drops its bond lists (no observe()).
    // A member with spatial order re-sorts on the device at the start of
    // every MCS whose age is a multiple of its sys.steps_between_sortings, as
    // a plain updater does (never at the age it was added with, its layout is
    // fresh); may_sort = false promises that it will not run over such an
    // MCS (the engine then reserves no re-sort workspace for it).
    uint32_t add(HostSystem &&sys, uint64_t seed, int64_t age, bool keep_topology = true, bool may_sort = true);
    // Device bytes a planned system would take as a member.
    static size_t bytesFor(const HostSystem &sys);

    // Runs n_mcs MCS on every active member, in chunks of `chunk` MCS.  Where
    // the members' hot arrays (sites, ids, partner columns, lattices and
    // claim lattices) exceed the working-set budget, the members are split
    // into groups of about that many hot bytes (each of at least
    // kMinGroupMonomers monomers, to fill the device) and every group runs its
    // n_mcs before the next starts: a group's arrays then stay in the
    // memory-side cache (MALL) across its MCS.  Members are independent, so
    // every trajectory is the same as without groups.
    void run(uint32_t n_mcs, uint32_t chunk);
    // Working-set budget of run() in bytes (0: never split).
    void setWorkingSetBytes(size_t bytes) { mWorkingSetBytes = bytes; }
    size_t workingSetBytes() const { return mWorkingSetBytes; }
    // The groups run() would use for these members.
    std::vector<std::vector<uint32_t>> workingSetGroups(const std::vector<uint32_t> &ids) const;
    // Runs n_mcs MCS on the members listed (others idle).
    void runMembers(const std::vector<uint32_t> &ids, uint32_t n_mcs, uint32_t chunk);

    // Copies a member's positions back into its HostSystem's xyz (by id).
    void copyBack(uint32_t id);
    // The same in two halves: download() queues, behind the work queued so
    // far, the unsort and unwrap of the member's positions on the device and
    // the copy of the result (12 bytes per monomer, by id) into pinned memory,
    // and returns at once; once finished(), copyOut() reads it.  The member
    // may be removed in between (its memory is freed in stream order, after
    // the copy).  give() returns the pinned buffer.
    struct Download {
        size_t buf = 0;
        uint32_t n = 0;
        bool held = false;
    };
    // final = true: the member's last download before it is removed; the
    // result is staged in the member's own slab (behind its positions, over
    // the lattice and partner tables it no longer needs), so nothing is
    // allocated.  The member must not run again.
    Download download(uint32_t id, bool final = false);
    bool finished(const Download &d) const { return mPinned.ready(d.buf); }
    void wait(const Download &d) const { mPinned.wait(d.buf); }
    void copyOut(const Download &d, int32_t *xyz) const;
    void give(Download &d);
    size_t pinnedBytes(const Download &d) const { return d.held ? mPinned.bytes(d.buf) : 0; }
    // Frees a member's device memory; its id is not reused.
    void remove(uint32_t id);
    // Drops a member's host arrays too (removing it first if it is active);
    // its id stays unused.  No-op for a member already forgotten.
    void forget(uint32_t id);
    // Host memory the members' systems still hold.
    size_t hostBytes() const;
    // Members still tracked (added and not forgotten).
    uint32_t liveMembers() const;

    // Observables of a member's resident state (the device analyzers of a
    // plain updater, on the batch's stream; waits for the member's work).
    void observe(uint32_t id, Observables &out);

    const HostSystem &system(uint32_t id) const { return member(id).sys; }
    int64_t age(uint32_t id) const { return member(id).age; }
    bool active(uint32_t id) const { return id < mMembers.size() && mMembers[id] && mMembers[id]->slab; }
    size_t size() const { return mMembers.size(); }
    size_t deviceBytes() const { return mBytes; }
    void counters(uint32_t id, uint64_t &accepted, uint64_t &proposals);
    uint64_t sorts(uint32_t id) const { return member(id).sorts; }
    hipStream_t stream() const { return mStream; }
    void setUseGraphs(bool on) { mUseGraphs = on; }
    void setXcdRanges(bool on)
    {
        if (on != mXcdRanges)
            mCompIds.clear();   // the block tables are laid out for the placement
        mXcdRanges = on;
    }
    bool xcdRanges() const { return mXcdRanges; }
    // check waves per perform wave (2, 4 or 8); takes effect at the next chunk
    void setPerformGroup(uint32_t group);
    uint32_t performGroup() const { return mPerformGroup; }
    // Times probe_mcs MCS (in chunks of `chunk`) on every active member under
    // every launch configuration (perform group x XCD placement) and keeps the
    // fastest.  The members run on meanwhile (one untimed probe first, so
    // 7 * probe_mcs MCS each); their trajectories are the same under every
    // configuration.
    void autotune(uint32_t probe_mcs, uint32_t chunk);
    uint64_t graphLaunches() const { return mGraphLaunches; }
    uint64_t graphUpdates() const { return mGraphUpdates; }
    uint64_t graphInstantiations() const { return mGraphInstantiations; }
    uint64_t kernelLaunches() const { return mKernelLaunches; }
    void synchronize();

private:
    struct Member {
        HostSystem sys;
        uint64_t seed = 0;
        int64_t age = 0;
        uint64_t substeps = 0;
        void *slab = nullptr;
        size_t bytes = 0;
        size_t slab_class = 0;   // bytes actually held (the size class)
        bool spent = false;      // its final download has overwritten its lattice
        int64_t sort_period = 0; // re-sort every this many MCS (0: never)
        int64_t last_sort_age = 0;
        uint64_t sorts = 0;
        BatchSystem dev{};
        uint32_t check_blocks_max = 0;
        size_t hot_bytes = 0;    // arrays every MCS streams: sites, ids, partner columns, lattice, claims
        std::unique_ptr<DeviceAnalyzer> analyzer;   // made by the first observe()
    };

    Member &member(uint32_t id);
    const Member &member(uint32_t id) const;
    void rebuildTables(const std::vector<uint32_t> &ids);
    void runChunk(uint32_t mcs);
    // re-sorts the members whose sort falls on their current age; returns
    // the MCS until the next one falls due among them (at most `limit`)
    uint32_t sortDue(uint32_t limit);
    Resorter mResorter;   // shared by the members (stream order serialises them)
    // device memory in stream order (the pool of the device) when available
    void *allocate(size_t bytes, const char *what);
    void release(void *p);
    // Member slabs: a removed member's slab goes to a free list by size class
    // and the next member of that class takes it.  Every use is on mStream, so
    // a new member's upload runs after the old member's last work.
    void *takeSlab(size_t bytes, size_t &klass);
    void giveSlab(void *slab, size_t klass);
    std::vector<std::pair<size_t, void *>> mFreeSlabs;   // (size class, slab)
    template <class T>
    void ensureDevice(T *&p, size_t &cap, size_t n, const char *what);

    int mGpu;
    size_t mWorkingSetBytes = kDefaultWorkingSetBytes;
    hipStream_t mStream = nullptr;
    PinnedPool mPinned;   // staging of member uploads and downloads
    std::vector<std::unique_ptr<Member>> mMembers;
    size_t mBytes = 0;

    // tables of the current composition
    std::vector<uint32_t> mCompIds;          // member ids in launch order
    bool mPooled = false;                    // stream-ordered allocation available
    // host copies of the tables, built here and staged through pinned memory
    // (a copy from pageable memory would wait for the stream to drain)
    std::vector<BatchSystem> mHostSys;
    std::vector<BlockEntry> mHostCheckBlk[kClasses], mHostPerformBlk[kClasses];
    BatchSystem *mDevSys = nullptr;
    size_t mSysCap = 0;          // allocated entries
    uint32_t mMaxColours = 0;
    // per kernel-variant class: block -> system tables
    BlockEntry *mDevCheckBlk[kClasses] = {}, *mDevPerformBlk[kClasses] = {};
    size_t mCheckCap[kClasses] = {}, mPerformCap[kClasses] = {};   // allocated entries

    // chunk tables of the last two chunks (double-buffered so that the host
    // can write chunk k+1's table while chunk k runs)
    int mBuf = 0;
    ChunkEntry *mHostChunk[2] = {nullptr, nullptr};   // pinned
    size_t mHostChunkCap[2] = {0, 0};
    // one device chunk table and schedule: the upload and the schedule kernel
    // of chunk k+1 are queued behind chunk k's kernels, which read them
    ChunkEntry *mDevChunk = nullptr;
    size_t mChunkCap = 0;
    SchedEntry *mDevSched = nullptr;    // written by the schedule kernel
    size_t mSchedCap = 0;
    hipEvent_t mDone[2] = {nullptr, nullptr};

    // Tables are sized to capacities, with idle entries (a block entry that
    // exits at once, a system with no colours) beyond the composition, so a
    // chunk's launch sequence depends only on the capacities: one captured
    // graph per shape is replayed whatever members come and go.  Capacities
    // grow by 1/2 over the need and shrink when the need halves.
    uint32_t mSysCapN = 0;
    uint32_t mCheckCapN[kClasses] = {}, mPerformCapN[kClasses] = {};
    uint64_t mTableGen = 0;   // bumped whenever a table the graphs read is reallocated
    bool mUseGraphs = true;
    struct CachedGraph {
        hipGraphExec_t exec = nullptr;
        uint64_t key[3] = {0, 0, 0};
        uint64_t gen = 0;
        uint64_t used = 0;
    };
    std::vector<CachedGraph> mGraphs;   // at most kGraphCache, least recently used evicted
    static constexpr size_t kGraphCache = 8;
    uint64_t mUse = 0;
    bool mXcdRanges = true;
    uint32_t mPerformGroup = kPerformGroup;
    uint64_t mGraphLaunches = 0, mKernelLaunches = 0;
    uint64_t mGraphUpdates = 0, mGraphInstantiations = 0;
};

}  // namespace pscbfm
