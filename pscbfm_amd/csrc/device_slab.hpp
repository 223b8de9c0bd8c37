// device_slab.hpp -- one system's device arrays in a single allocation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "bfm_kernels.hpp"
#include "host_system.hpp"

namespace pscbfm {

// Parts in slab order.  The parts an upload writes come first and are
// followed by the partner table, whose first columns (the largest partner
// count) end the uploaded prefix; everything after it starts zeroed.
enum SlabPart {
    kPartPPos,
    kPartImage,
    kPartIds,
    kPartTables,
    kPartSegBegin,
    kPartSegCount,
    kPartLattice,
    kPartNbr,
    kPartClaim0,
    kPartClaim1,
    kPartSurv,
    kPartScount,
    kPartAccepted,
    kPartProposals,
    kPartCount
};

struct SlabLayout {
    size_t offset[kPartCount];
    size_t bytes[kPartCount];
    size_t total;
    size_t upload;   // bytes of the uploaded prefix (through the used partner columns)
};

// Offsets (256-byte aligned) of every array of a planned system.
SlabLayout slab_layout(const HostSystem &s);
// Bytes of host staging memory upload_slab needs for s: the uploaded prefix
// of the slab, laid out as in the slab.
size_t upload_bytes(const HostSystem &s);

// Uploads positions, images, ids, tables, colour segments, lattice and
// partners on `stream` and zeroes the rest of the slab; returns the kernels'
// view of it.  Without staging the host arrays are copied part by part
// (asynchronously: the HostSystem must stay alive until the stream has passed
// this point).  With a pinned `staging` buffer of upload_bytes(s) the arrays
// are first written into it at their slab offsets, so that the upload is one
// copy that never waits for the stream, plus one clear; the caller keeps the
// buffer until the upload has run.
SweepBuffers upload_slab(const HostSystem &s, void *slab, const SlabLayout &L, uint64_t seed, hipStream_t stream,
                         void *staging = nullptr);

// Pointer to one part of the slab.
inline void *slab_part(void *slab, const SlabLayout &L, SlabPart p)
{
    return static_cast<char *>(slab) + L.offset[p];
}

}  // namespace pscbfm
