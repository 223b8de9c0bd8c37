// updater.hpp -- MI355X counterpart of UpdaterGPUScBFM_AB_Type
// (src/pscbfm/UpdaterGPUScBFM_AB_Type.h:176-518).  Same setters, same call
// order, same exceptions; one HIP stream and one device-resident system per
// object.  Coordinates are int32 on the device (the reference picks uint8 or
// uint16 plus periodic-image counters, GPUScBFM_AB_Type.h:253-280, and caps
// the box at 256; int32 removes the overflow repair and the cap).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "analyzers.hpp"
#include "bfm_kernels.hpp"
#include "device_slab.hpp"
#include "host_system.hpp"
#include "resort.hpp"

namespace pscbfm {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string &what) : std::runtime_error(what), code(c) {}
};
struct NoDeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what);

class UpdaterGPUScBFM_AB_Type {
public:
    using T_BoxSize = uint64_t;
    using T_Coordinate = int32_t;
    using T_Id = uint32_t;

    UpdaterGPUScBFM_AB_Type();
    ~UpdaterGPUScBFM_AB_Type();
    UpdaterGPUScBFM_AB_Type(const UpdaterGPUScBFM_AB_Type &) = delete;
    UpdaterGPUScBFM_AB_Type &operator=(const UpdaterGPUScBFM_AB_Type &) = delete;

    void setGpu(int gpu);
    void copyBondSet(int dx, int dy, int dz, bool forbidden);
    void setLatticeSize(T_BoxSize x, T_BoxSize y, T_BoxSize z);
    void setNrOfAllMonomers(T_Id n);
    void setAttribute(T_Id i, int32_t attribute);
    void setMonomerCoordinates(T_Id i, T_Coordinate x, T_Coordinate y, T_Coordinate z);
    void setConnectivity(T_Id i, T_Id j);
    void setPeriodicity(bool px, bool py, bool pz);
    void setAge(int64_t age) { mAge = age; }
    int64_t getAge() const { return mAge; }
    void setRng(int rng);
    int getRng() const { return int(mSys.rng); }
    void setSeed(uint64_t seed) { mSeed = seed; }
    uint64_t seed() const { return mSeed; }
    void setStepsBetweenSortings(int64_t steps);
    int64_t getStepsBetweenSortings() const { return mStepsBetweenSortings; }
    void setSpatialSorting(bool on);

    // host half of initialize(): checks and colouring, no device work
    void validate();
    // The validated system (validated here if needed) moved out of an updater
    // that is not initialized; the updater is left as after cleanup().
    HostSystem takeSystem();
    void initialize();
    void runSimulationOnGPU(uint32_t n_mcs);
    int32_t getMonomerPositionInX(T_Id i) const { return at(i, 0); }
    int32_t getMonomerPositionInY(T_Id i) const { return at(i, 1); }
    int32_t getMonomerPositionInZ(T_Id i) const { return at(i, 2); }
    void cleanup();

    // extensions
    void runResident(uint32_t n_mcs) { runResident(n_mcs, true); }
    // runSimulationOnGPU with its copy-back staged in the library's static
    // device scratch: no device or host allocation, no kernel timing.
    void runSimulationOnGPUNoAlloc(uint32_t n_mcs);
    void copyBack();
    void setKernelTiming(uint32_t every) { mTimingEvery = every; }
    void setXcdRanges(bool on) { mBuf.xcd_ranges = on ? 1u : 0u; mXcdRanges = on; }
    bool xcdRanges() const { return mXcdRanges; }
    // check waves per perform wave (2, 4 or 8), for every colour segment;
    // without it each segment starts with default_perform_group of its size
    void setPerformGroup(uint32_t group);
    uint32_t defaultGroup(uint32_t colour) const;
    uint32_t defaultItems(uint32_t colour) const;
    uint32_t performGroup() const { return mPerformGroup; }
    // check workgroup size (64, 128, 256, 512 or 1024 threads), for every colour segment
    void setCheckBlock(uint32_t block);
    // Monomers per lane of the check kernel (1 or 2) for every colour segment
    // (any time), or for one (after initialize).
    void setCheckItems(uint32_t items);
    void setSegmentItems(uint32_t colour, uint32_t items);
    uint32_t segmentItems(uint32_t colour) const;
    // the launch configuration of colour segment c (after initialize)
    void setSegmentConfig(uint32_t colour, uint32_t check_block, uint32_t group);
    void segmentConfig(uint32_t colour, uint32_t &check_block, uint32_t &group) const;
    // monomers of a colour in Morton order of position (true) or in id order
    bool mortonLayout() const { return mMorton; }
    // Chooses the launch configuration by timing short probes (the run goes
    // on meanwhile; the trajectory is the same under every configuration):
    //  1. the monomer order within colours, Morton versus id (with spatial
    //     sorting on: the monomers are re-sorted on the device between two
    //     alternating wall-time probes of each order);
    //  2. per colour segment, the check workgroup size (64 ... 1024) and
    //     the perform group (2, 4, 8), from each segment's own kernel
    //     durations (the check's do not depend on the group, the perform's
    //     not on the workgroup size), under each XCD placement; the placement
    //     with the smaller total wins;
    //  3. that pick against the defaults by wall time (two probes each),
    //     kept only if faster there too.
    // Every step keeps the default unless another choice is faster by more
    // than 2 %.  One untimed probe first; the age advances by 15 * probe_mcs
    // when the pick is the default, else 19 * probe_mcs (four fewer without
    // spatial sorting).
    void autotune(uint32_t probe_mcs);
    const HostSystem &system() const { return mSys; }
    bool initialized() const { return mInitialized; }
    bool validated() const { return mValidated; }
    uint64_t attempts() const { return mAttempts; }
    uint64_t substeps() const { return mSubsteps; }
    size_t deviceBytes() const { return mDevBytes; }
    // syncs the stream; also folds pending kernel timings into checkMs/performMs
    void deviceCounters(uint64_t &accepted, uint64_t &proposals) const;
    double checkMs() const { return mCheckMs; }
    double performMs() const { return mPerformMs; }
    uint64_t timedLaunches() const { return mTimedLaunches; }
    uint64_t sorts() const { return mSorts; }

    // analyzers on the resident device state (no copy-back needed)
    void observe(Observables &out);
    void chainSums(std::vector<ChainSums> &out);
    void msdAttach(uint64_t mcs_to_skip, uint64_t delta_t_realization, uint64_t delta_t_evaluate,
                   uint64_t max_subsequences);
    void msdExecute();
    const DeviceAnalyzer &analyzer() const { return mAnalyzer; }

private:
    int32_t at(T_Id i, int a) const;
    // the sweep driver; `timing` = honour setKernelTiming
    void runResident(uint32_t n_mcs, bool timing);
    // every substep of n_mcs MCS timed, per colour
    void timedProbe(uint32_t n_mcs, std::vector<double> &check_ms, std::vector<double> &perform_ms);
    double wallProbe(uint32_t n_mcs);   // ms of n_mcs MCS on the stream, no per-kernel timing
    void resortLayout(bool morton);
    void freeDevice();
    void collectTimings() const;

    HostSystem mSys;
    int mGpu = -1;
    int64_t mAge = 0;
    uint64_t mSeed = 0x5EED000000000001ull;
    bool mValidated = false;
    bool mInitialized = false;
    bool mMonomersSet = false;

    hipStream_t mStream = nullptr;
    SweepBuffers mBuf{};
    void *mDevSlab = nullptr;   // one allocation holds every device array
    int32_t *mDevXyz = nullptr; // unwrapped coordinates for the copy-back (3 * n), made on first use
    size_t mDevBytes = 0;
    uint64_t mSubsteps = 0;     // substep counter, its parity selects buffers
    uint64_t mAttempts = 0;

    SlabLayout mLayout{};
    Resorter mResorter;          // on-device spatial re-sort
    int64_t mLastSortAge = 0;    // age of the last sort (initialize sorts)
    uint64_t mSorts = 0;
    DeviceAnalyzer mAnalyzer;    // observables and AnalyzerMsd on the device

    uint32_t mTimingEvery = 0;   // time every n-th substep (0: off)
    bool mXcdRanges = true;
    uint32_t mPerformGroup = kPerformGroup;
    bool mPerformGroupSet = false;    // setPerformGroup called: every segment uses it
    bool mCheckItemsSet = false;      // setCheckItems called: every segment uses it
    uint32_t mResidentWaves = 8192;   // 256 CUs x 4 SIMDs x 8 waves; queried at initialize
    uint32_t mCheckBlock = kBlock;
    uint32_t mCheckItems = kCheckItems;
    std::vector<uint32_t> mSegBlock, mSegGroup, mSegItems;   // per colour segment
    bool mMorton = true;          // current order within colours (with spatial sorting)
    int64_t mStepsBetweenSortings = 5000;   // mnStepsBetweenSortings, UpdaterGPUScBFM_AB_Type.cu:1379
    std::vector<hipEvent_t> mEventPool;          // reusable events
    mutable std::vector<hipEvent_t> mEvents;     // pending (check start, stop, perform start, stop) quadruples
    mutable std::vector<uint32_t> mEventColours; // colour of each pending quadruple
    mutable std::vector<double> mColourCheckMs, mColourPerformMs;   // summed per colour
    mutable double mCheckMs = 0.0, mPerformMs = 0.0;
    mutable uint64_t mTimedLaunches = 0;
    hipEvent_t takeEvent();
};

}  // namespace pscbfm
