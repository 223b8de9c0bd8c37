// updater.cpp -- host orchestration of the colour sweep on one device.
#include "updater.hpp"

#include "device_slab.hpp"
#include "noalloc.hpp"

#include <algorithm>
#include <cstring>
#include <sstream>

namespace pscbfm {

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        std::ostringstream msg;
        msg << what << ": " << hipGetErrorName(e) << " (" << hipGetErrorString(e) << ")";
        if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
            throw NoDeviceError(msg.str());
        throw HipError(e, msg.str());
    }
}

UpdaterGPUScBFM_AB_Type::UpdaterGPUScBFM_AB_Type() = default;

UpdaterGPUScBFM_AB_Type::~UpdaterGPUScBFM_AB_Type()
{
    try {
        freeDevice();
    } catch (...) {
    }
}

void UpdaterGPUScBFM_AB_Type::setGpu(int gpu)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice || (e == hipSuccess && n == 0))
        throw NoDeviceError("no HIP device is visible");
    hip_check(e, "hipGetDeviceCount");
    if (gpu < 0 || gpu >= n) {
        std::ostringstream msg;
        msg << "GPU with ID " << gpu << " not present. Only " << n << " GPUs are available.";
        throw std::invalid_argument(msg.str());
    }
    hip_check(hipSetDevice(gpu), "hipSetDevice");
    mGpu = gpu;
}

void UpdaterGPUScBFM_AB_Type::copyBondSet(int dx, int dy, int dz, bool forbidden)
{
    if (dx < -4 || dx > 3 || dy < -4 || dy > 3 || dz < -4 || dz > 3)
        throw std::invalid_argument("copyBondSet: components must lie in [-4,3]");
    mSys.forbidden[bond_table_index(dx, dy, dz)] = forbidden ? 1 : 0;
    mValidated = false;
}

void UpdaterGPUScBFM_AB_Type::setLatticeSize(T_BoxSize x, T_BoxSize y, T_BoxSize z)
{
    const T_BoxSize v[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        if (v[a] == 0 || v[a] > T_BoxSize(INT32_MAX)) {
            std::ostringstream msg;
            msg << "The box size (" << x << "," << y << "," << z << ") is outside the int32 coordinate range";
            throw std::invalid_argument(msg.str());
        }
        if ((v[a] & (v[a] - 1)) != 0) {
            std::ostringstream msg;
            msg << "Check whether the box size is a power of 2! (" << x << "," << y << "," << z << ")";
            throw std::runtime_error(msg.str());
        }
    }
    if (mInitialized)
        throw StateError("setLatticeSize after initialize; call cleanup first");
    mSys.box[0] = x;
    mSys.box[1] = y;
    mSys.box[2] = z;
    mValidated = false;
}

void UpdaterGPUScBFM_AB_Type::setNrOfAllMonomers(T_Id n)
{
    if (mMonomersSet) {
        std::ostringstream msg;
        msg << "Number of Monomers already set to " << mSys.n << "!";
        throw std::runtime_error(msg.str());
    }
    mSys.n = n;
    mSys.xyz.assign(3 * size_t(n), 0);
    mSys.attribute.assign(n, 0);
    mSys.nnbr.assign(n, 0);
    mSys.nbr.assign(size_t(n) * kMaxConnectivity, 0);
    mMonomersSet = true;
    mValidated = false;
}

static void check_index(uint32_t i, uint32_t n, const char *where)
{
    if (i >= n) {
        std::ostringstream msg;
        msg << where << ": monomer index " << i << " out of range (" << n << " monomers)";
        throw std::invalid_argument(msg.str());
    }
}

void UpdaterGPUScBFM_AB_Type::setAttribute(T_Id i, int32_t attribute)
{
    check_index(i, mSys.n, "setAttribute");
    mSys.attribute[i] = attribute;
}

void UpdaterGPUScBFM_AB_Type::setMonomerCoordinates(T_Id i, T_Coordinate x, T_Coordinate y, T_Coordinate z)
{
    check_index(i, mSys.n, "setMonomerCoordinates");
    if (mInitialized)
        throw StateError("setMonomerCoordinates after initialize; call cleanup first");
    mSys.xyz[3 * size_t(i) + 0] = x;
    mSys.xyz[3 * size_t(i) + 1] = y;
    mSys.xyz[3 * size_t(i) + 2] = z;
    mValidated = false;
}

void UpdaterGPUScBFM_AB_Type::setConnectivity(T_Id i, T_Id j)
{
    check_index(i, mSys.n, "setConnectivity");
    check_index(j, mSys.n, "setConnectivity");
    if (mInitialized)
        throw StateError("setConnectivity after initialize; call cleanup first");
    if (mSys.nnbr[i] >= kMaxConnectivity) {
        std::ostringstream msg;
        msg << "The maximum amount of bonds per monomer (" << kMaxConnectivity << ") has been exceeded!";
        throw std::invalid_argument(msg.str());
    }
    mSys.nbr[size_t(i) * kMaxConnectivity + mSys.nnbr[i]++] = j;
    mValidated = false;
}

void UpdaterGPUScBFM_AB_Type::setPeriodicity(bool px, bool py, bool pz)
{
    if (mInitialized)
        throw StateError("setPeriodicity after initialize; call cleanup first");
    mSys.periodic[0] = px;
    mSys.periodic[1] = py;
    mSys.periodic[2] = pz;
    mValidated = false;
}

void UpdaterGPUScBFM_AB_Type::setRng(int rng)
{
    if (rng != int(kRngIntHash) && rng != int(kRngPhilox)) {
        std::ostringstream msg;
        msg << "random number generator " << rng << " is not available; only " << kRngIntHash
            << " (IntHash) and " << kRngPhilox << " (Philox) are";
        throw std::invalid_argument(msg.str());
    }
    mSys.rng = uint32_t(rng);
    mBuf.rng = uint32_t(rng);
}

void UpdaterGPUScBFM_AB_Type::setStepsBetweenSortings(int64_t steps)
{
    if (steps <= 0)
        throw std::invalid_argument("setStepsBetweenSortings: steps must be positive");
    mStepsBetweenSortings = steps;
}

void UpdaterGPUScBFM_AB_Type::setPerformGroup(uint32_t group)
{
    if (!valid_perform_group(group))
        throw std::invalid_argument("perform group must be 2, 4 or 8");
    mPerformGroup = group;
    mPerformGroupSet = true;
    std::fill(mSegGroup.begin(), mSegGroup.end(), group);
}

uint32_t UpdaterGPUScBFM_AB_Type::defaultItems(uint32_t colour) const
{
    return default_check_items(check_waves(mSys.seg_count[colour], kBlock), mResidentWaves);
}

uint32_t UpdaterGPUScBFM_AB_Type::defaultGroup(uint32_t colour) const
{
    return default_perform_group(check_waves(mSys.seg_count[colour], kBlock), mResidentWaves);
}

void UpdaterGPUScBFM_AB_Type::setCheckBlock(uint32_t block)
{
    if (!valid_check_block(block))
        throw std::invalid_argument("check workgroup size must be 64, 128, 256, 512 or 1024");
    mCheckBlock = block;
    std::fill(mSegBlock.begin(), mSegBlock.end(), block);
}

void UpdaterGPUScBFM_AB_Type::setCheckItems(uint32_t items)
{
    if (!valid_check_items(items))
        throw std::invalid_argument("check monomers per lane must be 1 or 2");
    mCheckItems = items;
    mCheckItemsSet = true;
    std::fill(mSegItems.begin(), mSegItems.end(), items);
}

void UpdaterGPUScBFM_AB_Type::setSegmentItems(uint32_t colour, uint32_t items)
{
    if (!mInitialized)
        throw StateError("segment configuration before initialize");
    if (colour >= mSegItems.size())
        throw std::invalid_argument("no such colour");
    if (!valid_check_items(items))
        throw std::invalid_argument("check monomers per lane must be 1 or 2");
    mSegItems[colour] = items;
}

uint32_t UpdaterGPUScBFM_AB_Type::segmentItems(uint32_t colour) const
{
    if (!mInitialized)
        throw StateError("segment configuration before initialize");
    if (colour >= mSegItems.size())
        throw std::invalid_argument("no such colour");
    return check_items_for(mSegBlock[colour], mSegItems[colour]);
}

void UpdaterGPUScBFM_AB_Type::setSegmentConfig(uint32_t colour, uint32_t check_block, uint32_t group)
{
    if (!mInitialized)
        throw StateError("segment configuration before initialize");
    if (colour >= mSegBlock.size())
        throw std::invalid_argument("no such colour");
    if (!valid_check_block(check_block))
        throw std::invalid_argument("check workgroup size must be 64, 128, 256, 512 or 1024");
    if (!valid_perform_group(group))
        throw std::invalid_argument("perform group must be 2, 4 or 8");
    mSegBlock[colour] = check_block;
    mSegGroup[colour] = group;
}

void UpdaterGPUScBFM_AB_Type::segmentConfig(uint32_t colour, uint32_t &check_block, uint32_t &group) const
{
    if (!mInitialized)
        throw StateError("segment configuration before initialize");
    if (colour >= mSegBlock.size())
        throw std::invalid_argument("no such colour");
    check_block = mSegBlock[colour];
    group = mSegGroup[colour];
}

namespace {
// A device array of 32-bit words for the autotuner's own use (freed with it).
struct DeviceWords {
    uint32_t *p = nullptr;
    explicit DeviceWords(size_t n)
    {
        hip_check(hipMalloc(&p, sizeof(uint32_t) * std::max<size_t>(n, 1)), "hipMalloc (saved layout)");
    }
    ~DeviceWords()
    {
        if (p)
            (void)hipFree(p);
    }
    DeviceWords(const DeviceWords &) = delete;
    DeviceWords &operator=(const DeviceWords &) = delete;
};
}  // namespace

void UpdaterGPUScBFM_AB_Type::resortLayout(bool morton)
{
    mResorter.run(mSys, mDevSlab, mLayout, mBuf, mStream, !morton);
    mMorton = morton;
    ++mSorts;
}

void UpdaterGPUScBFM_AB_Type::timedProbe(uint32_t n_mcs, std::vector<double> &check_ms,
                                         std::vector<double> &perform_ms)
{
    collectTimings();
    const uint32_t every = mTimingEvery;
    mTimingEvery = 1;
    std::fill(mColourCheckMs.begin(), mColourCheckMs.end(), 0.0);
    std::fill(mColourPerformMs.begin(), mColourPerformMs.end(), 0.0);
    try {
        runResident(n_mcs, true);
        collectTimings();
    } catch (...) {
        mTimingEvery = every;
        throw;
    }
    mTimingEvery = every;
    check_ms = mColourCheckMs;
    perform_ms = mColourPerformMs;
}

double UpdaterGPUScBFM_AB_Type::wallProbe(uint32_t n_mcs)
{
    collectTimings();
    hipEvent_t e0 = takeEvent(), e1 = takeEvent();
    float ms = 0.0f;
    hipError_t err = hipEventRecord(e0, mStream);
    try {
        if (err == hipSuccess) {
            runResident(n_mcs, false);
            err = hipEventRecord(e1, mStream);
        }
        if (err == hipSuccess)
            err = hipEventSynchronize(e1);
        if (err == hipSuccess)
            err = hipEventElapsedTime(&ms, e0, e1);
    } catch (...) {
        mEventPool.push_back(e0);
        mEventPool.push_back(e1);
        throw;
    }
    mEventPool.push_back(e0);
    mEventPool.push_back(e1);
    hip_check(err, "probe timing");
    return double(ms);
}

// A tuned choice replaces the default only when its probe is faster by more
// than 2 %: short kernels time within a few per cent of each other from run
// to run, and a choice made on noise can be slower than the default.
static constexpr double kTuneMargin = 0.98;
// The probes of the per-segment search: every check workgroup size with one
// and with two monomers per lane (64 with one only), each paired with a
// perform group so that every group is timed (each segment's default group is
// default_perform_group of its size).
constexpr int kProbes = 9;
constexpr uint32_t kProbeBlock[kProbes] = {64, 128, 256, 512, 1024, 128, 256, 512, 1024};
constexpr uint32_t kProbeItems[kProbes] = {1, 1, 1, 1, 1, 2, 2, 2, 2};
constexpr uint32_t kProbeGroup[kProbes] = {4, 2, 8, 8, 4, 4, 4, 2, 8};
// the probes of the default workgroup size with one and with two monomers
// per lane (a segment's default is default_check_items of its size)
constexpr int kDefaultProbe1 = 2, kDefaultProbe2 = 6;
static_assert(kProbeBlock[kDefaultProbe1] == kBlock && kProbeItems[kDefaultProbe1] == 1 &&
                  kProbeBlock[kDefaultProbe2] == kBlock && kProbeItems[kDefaultProbe2] == 2,
              "the default probes have the default workgroup size");

void UpdaterGPUScBFM_AB_Type::autotune(uint32_t probe_mcs)
{
    if (!mInitialized)
        throw StateError("autotune before initialize");
    if (probe_mcs == 0)
        throw std::invalid_argument("autotune: probe_mcs must be positive");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    const uint32_t nc = mSys.n_colours;
    const std::vector<uint32_t> block0 = mSegBlock, items0 = mSegItems, group0 = mSegGroup;
    const bool xcd0 = mXcdRanges, morton0 = mMorton;
    auto restore = [&] {
        mSegBlock = block0;
        mSegItems = items0;
        mSegGroup = group0;
        setXcdRanges(xcd0);
    };
    auto set_config = [&](const std::vector<uint32_t> &block, const std::vector<uint32_t> &items,
                          const std::vector<uint32_t> &group, bool xcd) {
        mSegBlock = block;
        mSegItems = items;
        mSegGroup = group;
        setXcdRanges(xcd);
    };
    try {
        std::vector<double> cms, pms;
        // one untimed probe first, so that no configuration pays for cold caches
        runResident(probe_mcs);
        // 1. the monomer order within colours (only where spatial sorting is on:
        // without it the layout is id order throughout), by wall time: two
        // alternating probes per order, the faster of each pair.  The layout
        // the system has is restored after each probe of the other order
        // (every monomer back into its saved slot), not derived again from the
        // current positions: on the relaxed 512^3 melt a fresh Morton order
        // measured 30.6 us per check launch against 28.2 us for the order
        // sorted at initialize, which the probes would otherwise replace
        if (mSys.spatial_order) {
            DeviceWords slots(mSys.n);
            Resorter::saveSlots(mSys, mBuf, slots.p, mStream);
            const bool morton_now = mMorton;
            double t_now = 0.0, t_other = 0.0;
            for (int rep = 0; rep < 2; ++rep) {
                const double a = wallProbe(probe_mcs);
                resortLayout(!morton_now);
                const double b = wallProbe(probe_mcs);
                mResorter.run(mSys, mDevSlab, mLayout, mBuf, mStream, false, slots.p);
                mMorton = morton_now;
                ++mSorts;
                t_now = rep == 0 ? a : std::min(t_now, a);
                t_other = rep == 0 ? b : std::min(t_other, b);
            }
            if (t_other < kTuneMargin * t_now)
                resortLayout(!mMorton);
            hip_check(hipStreamSynchronize(mStream), "hipStreamSynchronize (layout probes)");
        }
        // 2. per segment: check shape (workgroup size, monomers per lane) and
        // perform group, per placement, from each colour's own kernel
        // durations (the perform kernel's work does not depend on the check
        // shape: run r of a check launch covers monomers [64 r, 64 r + 64)
        // whatever the shape)
        double best_total[2] = {0.0, 0.0};
        std::vector<uint32_t> best_block[2], best_items[2], best_group[2];
        for (int x = 0; x < 2; ++x) {
            setXcdRanges(x == 0);
            std::vector<double> check_t[kProbes], perform_t[kProbes];
            for (int k = 0; k < kProbes; ++k) {
                std::fill(mSegBlock.begin(), mSegBlock.end(), kProbeBlock[k]);
                std::fill(mSegItems.begin(), mSegItems.end(), kProbeItems[k]);
                std::fill(mSegGroup.begin(), mSegGroup.end(), kProbeGroup[k]);
                timedProbe(probe_mcs, check_t[k], perform_t[k]);
            }
            best_block[x].assign(nc, kBlock);
            best_items[x].resize(nc);
            best_group[x].resize(nc);
            for (uint32_t c = 0; c < nc; ++c) {
                // the default unless another is faster by more than the margin;
                // a group timed in several probes counts with its best time
                int bc = defaultItems(c) == 2u ? kDefaultProbe2 : kDefaultProbe1;
                for (int k = 0; k < kProbes; ++k)
                    if (check_t[k][c] < kTuneMargin * check_t[bc][c])
                        bc = k;
                double group_t[3] = {-1.0, -1.0, -1.0};
                for (int k = 0; k < kProbes; ++k) {
                    int gi = 0;
                    while (gi < 2 && kPerformGroups[gi] != kProbeGroup[k])
                        ++gi;
                    if (group_t[gi] < 0.0 || perform_t[k][c] < group_t[gi])
                        group_t[gi] = perform_t[k][c];
                }
                int bp = 0;   // the segment's default group
                while (bp < 2 && kPerformGroups[bp] != defaultGroup(c))
                    ++bp;
                for (int gi = 0; gi < 3; ++gi)
                    if (group_t[gi] >= 0.0 && group_t[gi] < kTuneMargin * group_t[bp])
                        bp = gi;
                best_block[x][c] = kProbeBlock[bc];
                best_items[x][c] = kProbeItems[bc];
                best_group[x][c] = kPerformGroups[bp];
                best_total[x] += check_t[bc][c] + group_t[bp];
            }
        }
        const int x = best_total[1] < kTuneMargin * best_total[0] ? 1 : 0;
        // 3. the pick against the defaults by wall time, which also sees what
        // per-kernel durations do not (the boundaries between dependent
        // kernels): two alternating probes each, the faster of each pair
        const std::vector<uint32_t> def_block(nc, kBlock);
        std::vector<uint32_t> def_items(nc), def_group(nc);
        for (uint32_t c = 0; c < nc; ++c) {
            def_items[c] = defaultItems(c);
            def_group[c] = defaultGroup(c);
        }
        const bool tuned_is_default =
            best_block[x] == def_block && best_items[x] == def_items && best_group[x] == def_group && x == 0;
        bool keep_tuned = false;
        if (!tuned_is_default) {
            double t_def = 0.0, t_tuned = 0.0;
            for (int rep = 0; rep < 2; ++rep) {
                set_config(def_block, def_items, def_group, true);
                const double a = wallProbe(probe_mcs);
                set_config(best_block[x], best_items[x], best_group[x], x == 0);
                const double b = wallProbe(probe_mcs);
                t_def = rep == 0 ? a : std::min(t_def, a);
                t_tuned = rep == 0 ? b : std::min(t_tuned, b);
            }
            keep_tuned = t_tuned < kTuneMargin * t_def;
        }
        if (keep_tuned)
            set_config(best_block[x], best_items[x], best_group[x], x == 0);
        else
            set_config(def_block, def_items, def_group, true);
    } catch (...) {
        restore();
        if (mMorton != morton0) {
            try {
                resortLayout(morton0);
            } catch (...) {
            }
        }
        throw;
    }
    mPerformGroup = mSegGroup.empty() ? kPerformGroup : mSegGroup[0];
    mCheckBlock = mSegBlock.empty() ? kBlock : mSegBlock[0];
    mCheckItems = mSegItems.empty() ? kCheckItems : mSegItems[0];
}

void UpdaterGPUScBFM_AB_Type::setSpatialSorting(bool on)
{
    if (mInitialized)
        throw StateError("setSpatialSorting after initialize; call cleanup first");
    mSys.spatial_order = on;
    mValidated = false;
}

void UpdaterGPUScBFM_AB_Type::validate()
{
    if (mInitialized)
        throw std::runtime_error("Initialize was already called and may not be called again until cleanup was called!");
    plan(mSys);
    mValidated = true;
}

HostSystem UpdaterGPUScBFM_AB_Type::takeSystem()
{
    if (mInitialized)
        throw StateError("takeSystem of an initialized updater");
    if (!mValidated)
        validate();
    HostSystem out = std::move(mSys);
    cleanup();
    return out;
}

void UpdaterGPUScBFM_AB_Type::initialize()
{
    validate();
    if (mGpu < 0)
        setGpu(0);
    else
        hip_check(hipSetDevice(mGpu), "hipSetDevice");

    const HostSystem &s = mSys;
    const SlabLayout L = slab_layout(s);
    mLayout = L;
    hipError_t e = hipMalloc(&mDevSlab, L.total);
    if (e != hipSuccess) {
        mDevSlab = nullptr;
        hip_check(e, "hipMalloc (system slab)");
    }
    mDevBytes = L.total;
    try {
        hip_check(hipStreamCreate(&mStream), "hipStreamCreate");
        mBuf = upload_slab(s, mDevSlab, L, mSeed, mStream);
        mBuf.xcd_ranges = mXcdRanges ? 1u : 0u;
        // the periodic re-sort's workspace, so that runs never allocate
        if (s.spatial_order)
            mResorter.reserve(s, L);
        hip_check(hipStreamSynchronize(mStream), "hipStreamSynchronize (initialize)");
    } catch (...) {
        freeDevice();
        throw;
    }
    // The sorted arrays were staging for the upload: from here on the device
    // holds the state (the copy-back unsorts and unwraps there) and the sweep
    // driver and the re-sort read only the colour segments.
    release_upload_staging(mSys);
    mSubsteps = 0;
    mAttempts = 0;
    mCheckMs = mPerformMs = 0.0;
    mTimedLaunches = 0;
    mLastSortAge = mAge;   // plan() laid the colours out in spatial order
    mSorts = 0;
    mMorton = s.spatial_order;
    // resident wavefronts of the device at the sweep kernels' occupancy (8
    // per SIMD, 4 SIMDs per CU): the default perform group depends on it
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, mGpu) == hipSuccess && cus > 0)
        mResidentWaves = uint32_t(cus) * 4u * 8u;
    mSegBlock.assign(s.n_colours, mCheckBlock);
    mSegItems.resize(s.n_colours);
    mSegGroup.resize(s.n_colours);
    for (uint32_t c = 0; c < s.n_colours; ++c)
        mSegGroup[c] = mPerformGroupSet ? mPerformGroup : defaultGroup(c);
    for (uint32_t c = 0; c < s.n_colours; ++c)
        mSegItems[c] = mCheckItemsSet ? mCheckItems : defaultItems(c);
    mColourCheckMs.assign(s.n_colours, 0.0);
    mColourPerformMs.assign(s.n_colours, 0.0);
    mInitialized = true;
}

void UpdaterGPUScBFM_AB_Type::runResident(uint32_t n_mcs, bool timing)
{
    if (!mInitialized)
        throw StateError("runSimulationOnGPU before initialize");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    mBuf.seed = mSeed;
    const HostSystem &s = mSys;
    uint32_t order[256];   // at most 255 colours; no allocation in the loop
    for (uint32_t m = 0; m < n_mcs; ++m) {
        // periodic re-sort at the start of every mnStepsBetweenSortings-th MCS
        // (runSimulationOnGPU, UpdaterGPUScBFM_AB_Type.cu:3005-3012)
        if (s.spatial_order && mMorton && mStepsBetweenSortings > 0 && mAge % mStepsBetweenSortings == 0 &&
            mAge != mLastSortAge) {
            mResorter.run(s, mDevSlab, mLayout, mBuf, mStream);
            mLastSortAge = mAge;
            ++mSorts;
        }
        const uint64_t age = uint64_t(mAge);
        colour_order(mSeed, age, s.n_colours, order);
        for (uint32_t k = 0; k < s.n_colours; ++k) {
            const uint32_t c = order[k];
            const uint32_t parity = uint32_t(mSubsteps & 1u);
            // timed substeps: start and stop events of each kernel's own dispatch
            hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
            const bool timed = timing && mTimingEvery != 0 && mSubsteps % mTimingEvery == 0;
            if (timed)
                for (auto &x : ev)
                    x = takeEvent();
            hip_check(launch_check(mBuf, s.seg_begin[c], s.seg_count[c], age, substep_seed(mSeed, age, k), parity,
                                   mStream, ev[0], ev[1], mSegBlock[c], mSegItems[c]),
                      "check kernel");
            hip_check(launch_perform(mBuf, s.seg_count[c], parity, mStream, mSegGroup[c], ev[2], ev[3], mSegBlock[c]),
                      "perform kernel");
            if (timed) {
                mEvents.insert(mEvents.end(), ev, ev + 4);
                mEventColours.push_back(c);
            }
            ++mSubsteps;
        }
        mAttempts += s.n;
        ++mAge;
        if (mEvents.size() > 4 * 4096)
            collectTimings();
    }
}

hipEvent_t UpdaterGPUScBFM_AB_Type::takeEvent()
{
    if (mEventPool.empty()) {
        hipEvent_t e = nullptr;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }
    hipEvent_t e = mEventPool.back();
    mEventPool.pop_back();
    return e;
}

void UpdaterGPUScBFM_AB_Type::collectTimings() const
{
    if (mEvents.empty())
        return;
    hip_check(hipEventSynchronize(mEvents.back()), "hipEventSynchronize");
    for (size_t k = 0; k + 3 < mEvents.size(); k += 4) {
        float a = 0.f, b = 0.f;
        hip_check(hipEventElapsedTime(&a, mEvents[k], mEvents[k + 1]), "hipEventElapsedTime");
        hip_check(hipEventElapsedTime(&b, mEvents[k + 2], mEvents[k + 3]), "hipEventElapsedTime");
        mCheckMs += a;
        mPerformMs += b;
        ++mTimedLaunches;
        const uint32_t c = mEventColours[k / 4];
        if (c < mColourCheckMs.size()) {
            mColourCheckMs[c] += a;
            mColourPerformMs[c] += b;
        }
    }
    auto &pool = const_cast<std::vector<hipEvent_t> &>(mEventPool);
    pool.insert(pool.end(), mEvents.begin(), mEvents.end());
    mEvents.clear();
    mEventColours.clear();
}

void UpdaterGPUScBFM_AB_Type::copyBack()
{
    if (!mInitialized)
        throw StateError("copy-back before initialize");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    HostSystem &s = mSys;
    // unsort and unwrap on the device (wherever a re-sort has moved the
    // monomers: the id words travel with them), then one copy of the result
    if (s.n == 0)
        return;
    if (!mDevXyz)
        hip_check(hipMalloc(&mDevXyz, sizeof(int32_t) * 3 * size_t(s.n)), "allocation of the copy-back buffer");
    const uint32_t box[3] = {uint32_t(s.box[0]), uint32_t(s.box[1]), uint32_t(s.box[2])};
    hip_check(launch_unwrap(mBuf, box, mDevXyz, mStream), "unwrap kernel");
    hip_check(hipMemcpyAsync(s.xyz.data(), mDevXyz, sizeof(int32_t) * 3 * size_t(s.n), hipMemcpyDeviceToHost,
                             mStream),
              "download positions");
    hip_check(hipStreamSynchronize(mStream), "hipStreamSynchronize (copy-back)");
    collectTimings();
}

void UpdaterGPUScBFM_AB_Type::runSimulationOnGPU(uint32_t n_mcs)
{
    runResident(n_mcs);
    copyBack();
}

void UpdaterGPUScBFM_AB_Type::runSimulationOnGPUNoAlloc(uint32_t n_mcs)
{
    HostSystem &s = mSys;
    if (s.n > kNoAllocMaxMonomers) {
        std::ostringstream msg;
        msg << "runSimulationOnGPU without allocation takes at most " << kNoAllocMaxMonomers
            << " monomers (its copy-back scratch holds " << (kNoAllocScratchBytes >> 20) << " MiB); this system has "
            << s.n;
        throw std::invalid_argument(msg.str());
    }
    if (!mInitialized)
        throw StateError("runSimulationOnGPU before initialize");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    // The sweep allocates nothing after initialize (its re-sort workspace is
    // reserved there); the copy-back is staged in the library's device scratch
    // of this GPU, shared by every call on it: calls that overlap on one GPU
    // race.  The unwrap kernel writes every monomer's entry, so nothing in the
    // scratch needs clearing first.
    int32_t *staged = static_cast<int32_t *>(noalloc_scratch(mGpu));
    runResident(n_mcs, false);
    if (s.n == 0)
        return;
    const uint32_t box[3] = {uint32_t(s.box[0]), uint32_t(s.box[1]), uint32_t(s.box[2])};
    hip_check(launch_unwrap(mBuf, box, staged, mStream), "unwrap kernel");
    hip_check(hipMemcpyAsync(s.xyz.data(), staged, sizeof(int32_t) * 3 * size_t(s.n), hipMemcpyDeviceToHost, mStream),
              "download positions");
    hip_check(hipStreamSynchronize(mStream), "hipStreamSynchronize (copy-back)");
}

void UpdaterGPUScBFM_AB_Type::deviceCounters(uint64_t &accepted, uint64_t &proposals) const
{
    accepted = proposals = 0;
    if (!mInitialized)
        return;
    std::vector<unsigned long long> slots(2 * size_t(kStatSlots));
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    hip_check(hipMemcpyAsync(slots.data(), mBuf.accepted, sizeof(unsigned long long) * kStatSlots, hipMemcpyDeviceToHost,
                             mStream),
              "download counters");
    hip_check(hipMemcpyAsync(slots.data() + kStatSlots, mBuf.proposals, sizeof(unsigned long long) * kStatSlots,
                             hipMemcpyDeviceToHost, mStream),
              "download counters");
    hip_check(hipStreamSynchronize(mStream), "hipStreamSynchronize (counters)");
    collectTimings();
    for (uint32_t k = 0; k < kStatSlots; ++k) {
        accepted += slots[k];
        proposals += slots[kStatSlots + k];
    }
}

int32_t UpdaterGPUScBFM_AB_Type::at(T_Id i, int a) const
{
    check_index(i, mSys.n, "getMonomerPosition");
    return mSys.xyz[3 * size_t(i) + a];
}

void UpdaterGPUScBFM_AB_Type::freeDevice()
{
    if (mGpu >= 0 && (mStream || mDevSlab))
        hipSetDevice(mGpu);
    for (auto e : mEvents)
        hipEventDestroy(e);
    mEvents.clear();
    mEventColours.clear();
    for (auto e : mEventPool)
        hipEventDestroy(e);
    mEventPool.clear();
    if (mStream) {
        hipStreamSynchronize(mStream);
        hipStreamDestroy(mStream);
        mStream = nullptr;
    }
    mResorter.release();
    mAnalyzer.release();
    if (mDevSlab) {
        hipFree(mDevSlab);
        mDevSlab = nullptr;
    }
    if (mDevXyz) {
        hipFree(mDevXyz);
        mDevXyz = nullptr;
    }
    mBuf = SweepBuffers{};
    mInitialized = false;
}

void UpdaterGPUScBFM_AB_Type::observe(Observables &out)
{
    if (!mInitialized)
        throw StateError("observables before initialize");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    mAnalyzer.observe(mSys, mBuf, mStream, out);
}

void UpdaterGPUScBFM_AB_Type::chainSums(std::vector<ChainSums> &out)
{
    if (!mInitialized)
        throw StateError("chain sums before initialize");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    mAnalyzer.chainSums(mSys, mBuf, mStream, out);
}

void UpdaterGPUScBFM_AB_Type::msdAttach(uint64_t mcs_to_skip, uint64_t delta_t_realization,
                                        uint64_t delta_t_evaluate, uint64_t max_subsequences)
{
    if (!mInitialized)
        throw StateError("MSD analyzer before initialize");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    mAnalyzer.msdCreate(mSys, mBuf, mStream, uint64_t(mAge), mcs_to_skip, delta_t_realization, delta_t_evaluate,
                        max_subsequences);
}

void UpdaterGPUScBFM_AB_Type::msdExecute()
{
    if (!mInitialized)
        throw StateError("MSD analyzer before initialize");
    hip_check(hipSetDevice(mGpu), "hipSetDevice");
    mAnalyzer.msdExecute(mSys, mBuf, mStream, uint64_t(mAge));
}

void UpdaterGPUScBFM_AB_Type::cleanup()
{
    freeDevice();
    mSys.reset();
    mMonomersSet = false;
    mValidated = false;
    mSubsteps = 0;
    mAttempts = 0;
}

}  // namespace pscbfm
