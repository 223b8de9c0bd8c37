// device_slab.cpp -- layout and upload of one system's device arrays.
#include "device_slab.hpp"

#include <algorithm>
#include <cstring>

#include "updater.hpp"

namespace pscbfm {

SlabLayout slab_layout(const HostSystem &s)
{
    const size_t blocks = check_blocks(std::max<uint32_t>(s.max_segment, 1));
    SlabLayout L{};
    L.bytes[kPartPPos] = sizeof(uint32_t) * s.n_pad;
    L.bytes[kPartIds] = sizeof(uint32_t) * s.n_pad;
    L.bytes[kPartImage] = sizeof(int32_t) * 3 * size_t(s.n_pad);
    L.bytes[kPartNbr] = sizeof(uint32_t) * kMaxConnectivity * s.n_pad;
    L.bytes[kPartLattice] = sizeof(uint32_t) * s.lattice_words;
    L.bytes[kPartClaim0] = sizeof(uint32_t) * s.lattice_words;
    L.bytes[kPartClaim1] = sizeof(uint32_t) * s.lattice_words;
    L.bytes[kPartSurv] = sizeof(uint2) * kBlock * blocks;
    L.bytes[kPartScount] = sizeof(uint32_t) * kWavesPerBlock * blocks;
    L.bytes[kPartAccepted] = sizeof(unsigned long long) * kStatSlots;
    L.bytes[kPartProposals] = sizeof(unsigned long long) * kStatSlots;
    L.bytes[kPartTables] = sizeof(Tables);
    L.bytes[kPartSegBegin] = sizeof(uint32_t) * std::max<uint32_t>(s.n_colours, 1);
    L.bytes[kPartSegCount] = sizeof(uint32_t) * std::max<uint32_t>(s.n_colours, 1);
    size_t total = 0;
    for (int p = 0; p < kPartCount; ++p) {
        L.offset[p] = total;
        total += (L.bytes[p] + 255) / 256 * 256;
    }
    L.total = total;
    return L;
}

// Partner-table columns in use (the largest partner count); the slab is
// zeroed first, so the columns above need no copy.
static uint32_t used_partner_columns(const HostSystem &s)
{
    uint32_t cols = 0;
    for (const uint8_t c : s.nnbr)
        cols = std::max<uint32_t>(cols, c);
    return cols;
}

static size_t staged(size_t bytes) { return (bytes + 15) / 16 * 16; }

size_t upload_bytes(const HostSystem &s)
{
    return staged(sizeof(uint32_t) * s.n_pad) * 2 + staged(sizeof(int32_t) * 3 * size_t(s.n_pad)) +
           staged(sizeof(uint32_t) * used_partner_columns(s) * size_t(s.n_pad)) +
           staged(sizeof(uint32_t) * s.lattice_words) + staged(sizeof(Tables)) +
           2 * staged(sizeof(uint32_t) * s.n_colours);
}

SweepBuffers upload_slab(const HostSystem &s, void *slab, const SlabLayout &L, uint64_t seed, hipStream_t stream,
                         void *staging)
{
    // with staging: the copied parts packed one after the other (upload_bytes)
    size_t at = 0;
    auto up = [&](SlabPart p, const void *src, size_t bytes, const char *what) {
        if (bytes == 0)
            return;
        if (staging) {
            void *dst = static_cast<char *>(staging) + at;
            std::memcpy(dst, src, bytes);
            src = dst;
            at += staged(bytes);
        }
        hip_check(hipMemcpyAsync(slab_part(slab, L, p), src, bytes, hipMemcpyHostToDevice, stream), what);
    };
    hip_check(hipMemsetAsync(slab, 0, L.total, stream), "hipMemsetAsync (slab)");
    up(kPartPPos, s.ppos_sorted.data(), L.bytes[kPartPPos], "upload positions");
    up(kPartIds, s.ids_sorted.data(), L.bytes[kPartIds], "upload ids");
    up(kPartImage, s.image_sorted.data(), L.bytes[kPartImage], "upload images");
    up(kPartNbr, s.nbr_sorted.data(), sizeof(uint32_t) * used_partner_columns(s) * size_t(s.n_pad), "upload nbr");
    up(kPartLattice, s.lattice.data(), L.bytes[kPartLattice], "upload lattice");
    up(kPartTables, &s.tables, sizeof(Tables), "upload tables");
    up(kPartSegBegin, s.seg_begin.data(), sizeof(uint32_t) * s.n_colours, "upload segments");
    up(kPartSegCount, s.seg_count.data(), sizeof(uint32_t) * s.n_colours, "upload segments");

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
