// bfm_kernels.hpp -- launch interface of the colour-sweep kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "bfm_device.hpp"

namespace pscbfm {

constexpr uint32_t kPerformGroup = 4;         // check waves whose survivors one perform wave takes (4: ~42 survivors, one round of 64 lanes)
constexpr uint32_t kStatSlots = 4096;          // spread of the accepted / proposed counters (one per perform wave up to 4096 waves)

// Everything one substep of one system needs on the device.
struct SweepBuffers {
    uint32_t *ppos;           // lattice sites of the lower corners
    const uint32_t *ids;      // original id | partner count << kIdBits
    int32_t *image;           // [3][n_pad] periodic images
    const uint32_t *nbr;
    uint32_t n_pad;
    uint32_t *lattice;        // occupancy: 1 bit per monomer, at its lower corner
    uint32_t lattice_words;   // words of lattice and of each claim lattice (a power of two >= 16)
    // moved corners of this substep's survivors, double-buffered by substep
    // parity; selected with pick(), never indexed (a runtime index into an
    // argument struct forces it into scratch memory)
    uint32_t *claim0, *claim1;
    uint2 *surv;              // [check wave][kWave] compacted proposals
    uint32_t *scount;         // [check wave] proposals
    unsigned long long *accepted;   // [kStatSlots]
    unsigned long long *proposals;  // [kStatSlots]
    const Tables *tables;     // Morton tables (read only for non-cubic boxes)
    Geometry geom;
    uint64_t seed;
    uint32_t rng;             // kRngIntHash or kRngPhilox
    uint32_t xcd_ranges;      // 1: logical blocks laid out as one contiguous range per XCD
};

constexpr uint32_t kXcds = 8;   // accelerator complex dies of an MI355X

// Workgroups are dealt to the XCDs round-robin by block index, and each XCD
// has its own L2.  Logical block of physical block b in a grid of n blocks
// such that every XCD sweeps one contiguous range of logical blocks (so
// neighbouring monomers, which share lattice lines, stay on one L2).
__host__ __device__ __forceinline__ uint32_t xcd_logical(uint32_t b, uint32_t n)
{
    const uint32_t x = b % kXcds, i = b / kXcds, per = n / kXcds, rem = n % kXcds;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

template <class T>
__host__ __device__ inline T pick(uint32_t parity, T even, T odd)
{
    return parity ? odd : even;
}

// Workgroups of a check launch over `count` monomers with `block` threads each.
__host__ __device__ inline uint32_t check_blocks(uint32_t count, uint32_t block = kBlock)
{
    return count == 0 ? 1u : (count + block - 1) / block;
}
// Wavefronts of that launch.  Wave w always covers monomers [64 w, 64 w + 64)
// of the segment, whatever the workgroup size, so the survivors of a wave (and
// the perform kernel's work) do not depend on it.
__host__ __device__ inline uint32_t check_waves(uint32_t count, uint32_t block = kBlock)
{
    return check_blocks(count, block) * (block / kWave);
}
// Perform blocks (kBlock threads) covering `waves` check waves in groups.
__host__ __device__ inline uint32_t perform_blocks_for(uint32_t waves, uint32_t group)
{
    return ((waves + group - 1) / group + kWavesPerBlock - 1) / kWavesPerBlock;
}
// perform blocks that cover the survivors of a check launch of cb blocks of kBlock
__host__ __device__ inline uint32_t perform_blocks(uint32_t cb, uint32_t group = kPerformGroup)
{
    return (cb + group - 1) / group;
}
// Workgroup sizes of the plain check kernel with a compiled variant, in
// monomers (a launch parameter the autotuner chooses among per colour
// segment); kBlock is the default, and surv / scount are sized for the
// largest.
constexpr uint32_t kCheckBlockSizes[] = {64, 128, 256, 512, 1024};
constexpr uint32_t kMaxCheckBlock = 1024;
inline bool valid_check_block(uint32_t b) { return b == 64 || b == 128 || b == 256 || b == 512 || b == 1024; }
// Monomers per lane of the plain check kernel, 1 or 2 (a launch parameter the
// autotuner chooses among per colour segment): with 2, a workgroup of B
// monomers has B / 2 threads and every wavefront takes two runs of 64
// monomers, issuing the loads of both before using any.  A workgroup of 64
// monomers takes one per lane whatever the setting.
constexpr uint32_t kCheckItems = 2;
inline bool valid_check_items(uint32_t items) { return items == 1 || items == 2; }
inline uint32_t check_items_for(uint32_t block, uint32_t items) { return block >= 128 ? items : 1u; }
// The monomers per lane a colour segment of `check_waves` runs of 64 starts
// with: 2 where the launch still has at least 2 generations of resident
// wavefronts with them (fewer, longer waves hide more of each wave's
// dependent loads: 27.3 against 28.7 us per check launch on the relaxed 512^3
// melt, 65,536 runs, 8 generations of one run), else 1 (a launch of about one
// generation, such as a 532,480-monomer colour, would leave half the wave
// slots empty: 9.0 against 8.2 us there).
inline uint32_t default_check_items(uint32_t check_waves, uint32_t resident_waves)
{
    return check_waves >= 4ull * resident_waves ? 2u : 1u;
}
// Perform group sizes with a compiled kernel (a launch parameter the
// autotuners choose among); kPerformGroup is the default.
constexpr uint32_t kPerformGroups[] = {2, 4, 8};
inline bool valid_perform_group(uint32_t g) { return g == 2 || g == 4 || g == 8; }
// The group a colour segment of `check_waves` check waves starts with: 8 where
// the perform launch still fills the device with one generation of resident
// wavefronts (each wave then keeps two rounds of loads in flight), else
// kPerformGroup.  Measured on the 512^3 melt (65,536 check waves per colour):
// 13.6 against 14.3 us per perform launch; on a 532,480-monomer colour group 8
// leaves one wave per SIMD and is slower.
inline uint32_t default_perform_group(uint32_t check_waves, uint32_t resident_waves)
{
    return check_waves >= 8ull * resident_waves ? 8u : kPerformGroup;
}

// Phase 1 of a substep of parity `parity` for the colour segment
// [begin, begin+count): every monomer proposes a move and checks it against
// walls, bonds and the occupancy lattice; survivors mark their moved corner in
// claim[parity] and are compacted per wave into surv.  Also clears
// claim[parity ^ 1], the marks of the previous substep.
// sub_seed: substep_seed() of this substep (used by the integer-hash RNG).
// With a start and a stop event, the events take the kernel's own dispatch
// timestamps (its execution time, as a kernel trace reports it).
// block: the check workgroup size in monomers (one of kCheckBlockSizes);
// items: monomers per lane (check_items_for(block, items) are used).
hipError_t launch_check(const SweepBuffers &buf, uint32_t begin, uint32_t count, uint64_t age, uint32_t sub_seed,
                        uint32_t parity, hipStream_t stream, hipEvent_t start = nullptr, hipEvent_t stop = nullptr,
                        uint32_t block = kBlock, uint32_t items = kCheckItems);

// Phase 2: every survivor with no other moved corner in its front moves.
// count is the colour segment size of the check launch, check_block its
// workgroup size.
hipError_t launch_perform(const SweepBuffers &buf, uint32_t count, uint32_t parity, hipStream_t stream,
                          uint32_t group = kPerformGroup, hipEvent_t start = nullptr, hipEvent_t stop = nullptr,
                          uint32_t check_block = kBlock);

// ---- many systems per launch -------------------------------------------------

constexpr uint32_t kNoColour = 0xFFFFFFFFu;

// One system of a batch, as the batched kernels see it.
struct BatchSystem {
    SweepBuffers buf;
    const uint32_t *seg_begin;   // per colour, sorted index of the first monomer
    const uint32_t *seg_count;   // per colour
    uint32_t n_colours;
    uint32_t pad;
};

// Block `blk` of system `sys` in a batched launch (one entry per block, so
// that a block finds its system and its place in it with one load).
struct BlockEntry {
    uint32_t sys;
    uint32_t blk;
};

// Where a system stands at the start of a chunk; substep k of MCS m of the
// chunk sweeps colour colour_at(seed, age + m, n_colours, k) with parity
// (substeps + m * n_colours + k) & 1.
struct ChunkEntry {
    uint64_t age;
    uint64_t substeps;
};

// What one system does in substep slot m * max_colours + k of a chunk,
// written on the device by launch_schedule.
struct SchedEntry {
    uint64_t age;          // MCS that keys the generator
    uint32_t colour;       // kNoColour: the system idles in this slot
    uint32_t parity;       // substep parity
    uint32_t sub_seed;     // substep_seed() of this substep
    uint32_t begin;        // the colour's segment: first sorted index ...
    uint32_t count;        // ... and size
    uint32_t pad;
};

// Kernel variant class of a system: periodic on every axis, cubic box.
enum : uint32_t { kClassPeriodicCubic = 0, kClassPeriodic = 1, kClassCubic = 2, kClassOther = 3, kClasses = 4 };
inline uint32_t kernel_class(const Geometry &g)
{
    const bool p = g.periodic == 7u, c = g.cubic != 0u;
    return p && c ? kClassPeriodicCubic : (p ? kClassPeriodic : (c ? kClassCubic : kClassOther));
}

// The schedule of a chunk of `mcs` MCS: sched[(m * max_colours + k) * n_sys
// + j] for every system j, from chunk[j] and the systems' colour segments.
hipError_t launch_schedule(const BatchSystem *sys, const ChunkEntry *chunk, uint32_t n_sys, uint32_t mcs,
                           uint32_t max_colours, SchedEntry *sched, hipStream_t stream);

// Phase 1 / phase 2 of one substep slot for every system of one class of a
// batch.  n_blocks = sum over the class's systems of their reserved blocks;
// blk_sys maps a block to its system (index into sys and into each slot's
// n_sys schedule entries).
hipError_t launch_check_batch(uint32_t cls, const BatchSystem *sys, const BlockEntry *blk_tab, uint32_t n_blocks,
                              const SchedEntry *sched, uint32_t n_sys, uint32_t slot, uint32_t xcd_ranges,
                              hipStream_t stream);
hipError_t launch_perform_batch(uint32_t cls, const BatchSystem *sys, const BlockEntry *blk_tab, uint32_t n_blocks,
                                const SchedEntry *sched, uint32_t n_sys, uint32_t slot, uint32_t xcd_ranges,
                                hipStream_t stream,
                                uint32_t group = kPerformGroup);

// Unwrapped coordinates (3 int32 per monomer, original-id order) of the
// system in buf into xyz (3 * monomers entries on the device).
hipError_t launch_unwrap(const SweepBuffers &buf, const uint32_t box[3], int32_t *xyz, hipStream_t stream);

// Streaming copy of `bytes` (a multiple of 16) from src to dst.
hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t stream);

}  // namespace pscbfm
