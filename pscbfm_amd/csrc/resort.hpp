// resort.hpp -- on-device spatial re-sort of a system's monomers.
//
// The reference re-sorts each species by the linearized position every
// mnStepsBetweenSortings MCS (doSpatialSorting, UpdaterGPUScBFM_AB_Type.cu:
// 1836-1935).  Here the key is the lattice site of a monomer's lower corner
// (the compact Morton order the lattice uses); corners never coincide, so the
// order is unique.  Each colour segment is radix-sorted on the device (rocPRIM), the
// per-monomer arrays are gathered into the new order and the partner indices
// are remapped through the inverse permutation.  The generator is keyed on the
// original id, so a re-sort never changes the trajectory.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "bfm_kernels.hpp"
#include "device_slab.hpp"
#include "host_system.hpp"

namespace pscbfm {

class Resorter {
public:
    Resorter() = default;
    ~Resorter();
    Resorter(const Resorter &) = delete;
    Resorter &operator=(const Resorter &) = delete;

    // Allocates (or grows) the device workspace a re-sort of s needs; the
    // sweep drivers call it before their run loops, which then allocate nothing.
    void reserve(const HostSystem &s, const SlabLayout &L);
    // Re-sorts the system held in `slab` (layout L, kernels' view buf) on
    // `stream`, in the workspace reserve() made (radix sort of rocPRIM): by
    // the lattice site of each monomer's corner, or with by_id by original id
    // (the layout without spatial order), or with slot_of_id into the slots
    // it gives by original id (a layout saved by saveSlots, restored).
    void run(const HostSystem &s, void *slab, const SlabLayout &L, const SweepBuffers &buf, hipStream_t stream,
             bool by_id = false, const uint32_t *slot_of_id = nullptr);
    // slot_of_id[id] = the slot of monomer id now (s.n entries), on `stream`.
    static void saveSlots(const HostSystem &s, const SweepBuffers &buf, uint32_t *slot_of_id, hipStream_t stream);
    // Frees the workspace (call with the device current).
    void release();
    size_t workspaceBytes() const { return mWorkBytes + mSortBytes; }

private:
    void *mWork = nullptr;
    size_t mWorkBytes = 0;
    void *mSortTemp = nullptr;
    size_t mSortBytes = 0;
};

}  // namespace pscbfm
