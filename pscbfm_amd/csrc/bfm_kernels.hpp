// bfm_kernels.hpp -- launch interface of the colour-sweep kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "bfm_device.hpp"

namespace pscbfm {

constexpr uint32_t kPerformGroup = 4;   // check blocks whose survivors one perform block takes
constexpr uint32_t kCollSlots = 4 * kBlock;  // collision-list slots of one check block
constexpr uint32_t kStatSlots = 256;    // spread of the accepted / proposed counters

// Everything one substep of one system needs on the device.
struct SweepBuffers {
    int4 *pos;
    const uint32_t *nbr;
    const uint8_t *nnbr;
    uint32_t n_pad;
    uint32_t *lattice;
    uint32_t *claim;
    uint32_t *coll[2];
    uint32_t *clist[2];       // [check block][kCollSlots] collision words
    uint32_t *ccount[2];      // [check block] entries in clist
    uint32_t *surv;           // [check block][kBlock] compacted proposals
    uint32_t *scount;         // [check block] proposals
    unsigned long long *accepted;   // [kStatSlots]
    unsigned long long *proposals;  // [kStatSlots]
    const Tables *tables;
    Geometry geom;
    uint64_t seed;
};

inline uint32_t check_blocks(uint32_t count) { return count == 0 ? 1u : (count + kBlock - 1) / kBlock; }

// Phase 1 of a substep of parity `parity` for the colour segment
// [begin, begin+count): every monomer proposes a move and checks it against
// walls, bonds and the occupancy lattice; survivors claim their face sites and
// are compacted per block into buf.surv.  Also clears the collision bits that
// the previous check launch (prev_blocks blocks, other parity) listed.
hipError_t launch_check(const SweepBuffers &buf, uint32_t begin, uint32_t count, uint64_t age, uint32_t parity,
                        uint32_t prev_blocks, hipStream_t stream);

// Phase 2: every survivor whose face sites no other survivor claimed moves;
// releases the claims.  count is the colour segment size of the check launch.
hipError_t launch_perform(const SweepBuffers &buf, uint32_t count, uint32_t parity, hipStream_t stream);

}  // namespace pscbfm
