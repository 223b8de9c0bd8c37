// device_slab.hpp -- one system's device arrays in a single allocation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "bfm_kernels.hpp"
#include "host_system.hpp"

namespace pscbfm {

enum SlabPart {
    kPartPPos,
    kPartIds,
    kPartImage,
    kPartNbr,
    kPartLattice,
    kPartClaim0,
    kPartClaim1,
    kPartSurv,
    kPartScount,
    kPartAccepted,
    kPartProposals,
    kPartTables,
    kPartSegBegin,
    kPartSegCount,
    kPartCount
};

struct SlabLayout {
    size_t offset[kPartCount];
    size_t bytes[kPartCount];
    size_t total;
};

// Offsets (256-byte aligned) of every array of a planned system.
SlabLayout slab_layout(const HostSystem &s);
// Bytes of host staging memory upload_slab needs for s: the parts it copies,
// packed (the partner table only up to the largest partner count).
size_t upload_bytes(const HostSystem &s);

// Zeroes the slab, uploads positions, partners, lattice, tables and colour
// segments on `stream` (asynchronously: the HostSystem must stay alive until
// the stream has passed this point) and returns the kernels' view of it.
// Clears the slab and uploads the system's arrays on `stream`.  With a pinned
// `staging` buffer of L.total bytes the host arrays are first copied into it
// (at their slab offsets), so that the uploads never wait for the stream;
// the caller keeps the buffer until the uploads have run.
SweepBuffers upload_slab(const HostSystem &s, void *slab, const SlabLayout &L, uint64_t seed, hipStream_t stream,
                         void *staging = nullptr);

// Pointer to one part of the slab.
inline void *slab_part(void *slab, const SlabLayout &L, SlabPart p)
{
    return static_cast<char *>(slab) + L.offset[p];
}

}  // namespace pscbfm
