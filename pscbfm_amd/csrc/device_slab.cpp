// device_slab.cpp -- layout and upload of one system's device arrays.
#include "device_slab.hpp"

#include <algorithm>
#include <cstring>

#include "updater.hpp"

namespace pscbfm {

// Partner-table columns in use (the largest partner count); the columns
// above are never read (zero, or the slot's own index after a re-sort on the
// device).
static uint32_t used_partner_columns(const HostSystem &s)
{
    uint32_t cols = 0;
    for (const uint8_t c : s.nnbr)
        cols = std::max<uint32_t>(cols, c);
    return cols;
}

SlabLayout slab_layout(const HostSystem &s)
{
    // survivor slots for the check launch with the most (padded) wavefronts
    const size_t waves = check_waves(std::max<uint32_t>(s.max_segment, 1), kMaxCheckBlock);
    SlabLayout L{};
    L.bytes[kPartPPos] = sizeof(uint32_t) * s.n_pad;
    L.bytes[kPartImage] = sizeof(int32_t) * 3 * size_t(s.n_pad);
    L.bytes[kPartIds] = sizeof(uint32_t) * s.n_pad;
    L.bytes[kPartTables] = sizeof(Tables);
    L.bytes[kPartSegBegin] = sizeof(uint32_t) * std::max<uint32_t>(s.n_colours, 1);
    L.bytes[kPartSegCount] = sizeof(uint32_t) * std::max<uint32_t>(s.n_colours, 1);
    L.bytes[kPartLattice] = sizeof(uint32_t) * s.lattice_words;
    L.bytes[kPartNbr] = sizeof(uint32_t) * kMaxConnectivity * s.n_pad;
    L.bytes[kPartClaim0] = sizeof(uint32_t) * s.lattice_words;
    L.bytes[kPartClaim1] = sizeof(uint32_t) * s.lattice_words;
    L.bytes[kPartSurv] = sizeof(uint2) * kWave * waves;
    L.bytes[kPartScount] = sizeof(uint32_t) * waves;
    L.bytes[kPartAccepted] = sizeof(unsigned long long) * kStatSlots;
    L.bytes[kPartProposals] = sizeof(unsigned long long) * kStatSlots;
    size_t total = 0;
    for (int p = 0; p < kPartCount; ++p) {
        L.offset[p] = total;
        total += (L.bytes[p] + 255) / 256 * 256;
    }
    L.total = total;
    L.upload = L.offset[kPartNbr] + sizeof(uint32_t) * size_t(used_partner_columns(s)) * s.n_pad;
    return L;
}

size_t upload_bytes(const HostSystem &s) { return slab_layout(s).upload; }

SweepBuffers upload_slab(const HostSystem &s, void *slab, const SlabLayout &L, uint64_t seed, hipStream_t stream,
                         void *staging)
{
    const size_t cols_bytes = sizeof(uint32_t) * size_t(used_partner_columns(s)) * s.n_pad;
    struct Piece {
        SlabPart part;
        const void *src;
        size_t bytes;
    } pieces[] = {
        {kPartPPos, s.ppos_sorted.data(), L.bytes[kPartPPos]},
        {kPartImage, s.image_sorted.data(), L.bytes[kPartImage]},
        {kPartIds, s.ids_sorted.data(), L.bytes[kPartIds]},
        {kPartTables, &s.tables, sizeof(Tables)},
        {kPartSegBegin, s.seg_begin.data(), sizeof(uint32_t) * s.n_colours},
        {kPartSegCount, s.seg_count.data(), sizeof(uint32_t) * s.n_colours},
        {kPartLattice, s.lattice.data(), L.bytes[kPartLattice]},
        {kPartNbr, s.nbr_sorted.data(), cols_bytes},
    };
    char *dev = static_cast<char *>(slab);
    if (staging) {
        // the prefix as it lies in the slab (alignment gaps are never read)
        char *host = static_cast<char *>(staging);
        for (const Piece &q : pieces)
            if (q.bytes)
                std::memcpy(host + L.offset[q.part], q.src, q.bytes);
        hip_check(hipMemcpyAsync(dev, host, L.upload, hipMemcpyHostToDevice, stream), "upload system");
        hip_check(hipMemsetAsync(dev + L.upload, 0, L.total - L.upload, stream), "hipMemsetAsync (slab)");
    } else {
        hip_check(hipMemsetAsync(dev, 0, L.total, stream), "hipMemsetAsync (slab)");
        for (const Piece &q : pieces)
            if (q.bytes)
                hip_check(hipMemcpyAsync(dev + L.offset[q.part], q.src, q.bytes, hipMemcpyHostToDevice, stream),
                          "upload system");
    }

    SweepBuffers b{};
    b.ppos = static_cast<uint32_t *>(slab_part(slab, L, kPartPPos));
    b.ids = static_cast<const uint32_t *>(slab_part(slab, L, kPartIds));
    b.image = static_cast<int32_t *>(slab_part(slab, L, kPartImage));
    b.nbr = static_cast<const uint32_t *>(slab_part(slab, L, kPartNbr));
    b.n_pad = s.n_pad;
    b.lattice = static_cast<uint32_t *>(slab_part(slab, L, kPartLattice));
    b.lattice_words = uint32_t(s.lattice_words);
    b.claim0 = static_cast<uint32_t *>(slab_part(slab, L, kPartClaim0));
    b.claim1 = static_cast<uint32_t *>(slab_part(slab, L, kPartClaim1));
    b.surv = static_cast<uint2 *>(slab_part(slab, L, kPartSurv));
    b.scount = static_cast<uint32_t *>(slab_part(slab, L, kPartScount));
    b.accepted = static_cast<unsigned long long *>(slab_part(slab, L, kPartAccepted));
    b.proposals = static_cast<unsigned long long *>(slab_part(slab, L, kPartProposals));
    b.tables = static_cast<const Tables *>(slab_part(slab, L, kPartTables));
    b.geom = s.geom;
    b.seed = seed;
    b.rng = s.rng;
    return b;
}

}  // namespace pscbfm
