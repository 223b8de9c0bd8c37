// resort.hip -- on-device spatial re-sort within colour segments (gfx950).
#include "resort.hpp"

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <stdexcept>
#include <utility>

#include "updater.hpp"

namespace pscbfm {
namespace {

constexpr int kResortBlock = 256;

// Sort key of every slot (padding slots included; they lie outside the
// sorted ranges): the site of its lower corner, its id, or the slot saved for
// its id; and the identity permutation.
__global__ __launch_bounds__(kResortBlock) void k_keys(const uint32_t *__restrict__ ppos,
                                                       const uint32_t *__restrict__ ids, uint32_t n_pad,
                                                       uint32_t by_id, const uint32_t *__restrict__ slot_of_id,
                                                       uint32_t *__restrict__ keys, uint32_t *__restrict__ ident,
                                                       uint32_t *__restrict__ perm)
{
    const uint32_t t = blockIdx.x * kResortBlock + threadIdx.x;
    if (t >= n_pad)
        return;
    const uint32_t idn = ids[t];
    if (slot_of_id)
        keys[t] = idn == 0xFFFFFFFFu ? t : slot_of_id[idn & kIdMask];
    else if (by_id)
        keys[t] = idn & kIdMask;
    else
        keys[t] = ppos[t];   // the corner's site: Morton order
    ident[t] = t;
    perm[t] = t;
}

// slot_of_id[id] = t for every slot t that holds a monomer.
__global__ __launch_bounds__(kResortBlock) void k_save_slots(const uint32_t *__restrict__ ids, uint32_t n_pad,
                                                             uint32_t *__restrict__ slot_of_id)
{
    const uint32_t t = blockIdx.x * kResortBlock + threadIdx.x;
    if (t >= n_pad)
        return;
    const uint32_t idn = ids[t];
    if (idn != 0xFFFFFFFFu)
        slot_of_id[idn & kIdMask] = t;
}

// inv[perm[t]] = t: where every old slot went.
__global__ __launch_bounds__(kResortBlock) void k_inverse(const uint32_t *__restrict__ perm, uint32_t n_pad,
                                                          uint32_t *__restrict__ inv)
{
    const uint32_t t = blockIdx.x * kResortBlock + threadIdx.x;
    if (t < n_pad)
        inv[perm[t]] = t;
}

// New slot t takes old slot perm[t]; partner indices move with their owners,
// and past a slot's own partners the table holds the slot's own index.
__global__ __launch_bounds__(kResortBlock) void k_gather(const uint32_t *__restrict__ ppos,
                                                         const uint32_t *__restrict__ ids,
                                                         const int32_t *__restrict__ image,
                                                         const uint32_t *__restrict__ nbr,
                                                         const uint32_t *__restrict__ perm,
                                                         const uint32_t *__restrict__ inv, uint32_t n_pad,
                                                         uint32_t *__restrict__ ppos2, uint32_t *__restrict__ ids2,
                                                         int32_t *__restrict__ image2, uint32_t *__restrict__ nbr2)
{
    const uint32_t t = blockIdx.x * kResortBlock + threadIdx.x;
    if (t >= n_pad)
        return;
    const uint32_t o = perm[t];
    const uint32_t idn = ids[o];
    ppos2[t] = ppos[o];
    ids2[t] = idn;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        image2[size_t(a) * n_pad + t] = image[size_t(a) * n_pad + o];
    const uint32_t nn = idn == 0xFFFFFFFFu ? 0u : idn >> kIdBits;
#pragma unroll
    for (int j = 0; j < kMaxConnectivity; ++j)
        nbr2[size_t(j) * n_pad + t] = uint32_t(j) < nn ? inv[nbr[size_t(j) * n_pad + o]] : t;
}

uint32_t site_bits(const HostSystem &s)
{
    uint32_t b = 0;
    while ((uint64_t(1) << b) < uint64_t(s.lattice_words) * 32u)
        ++b;
    return b;
}

}  // namespace

Resorter::~Resorter() { release(); }

void Resorter::release()
{
    if (mWork)
        hipFree(mWork);
    if (mSortTemp)
        hipFree(mSortTemp);
    mWork = mSortTemp = nullptr;
    mWorkBytes = mSortBytes = 0;
}

namespace {

// The workspace: the four gathered arrays (each 256-byte aligned, sized as
// in the slab), then keys, sorted keys, identity, permutation and inverse.
struct ResortSpace {
    size_t ppos, ids, image, nbr, keys, words, total;
};

ResortSpace resort_space(const HostSystem &s, const SlabLayout &L)
{
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    ResortSpace w{};
    w.ppos = 0;
    w.ids = w.ppos + up(L.bytes[kPartPPos]);
    w.image = w.ids + up(L.bytes[kPartIds]);
    w.nbr = w.image + up(L.bytes[kPartImage]);
    w.keys = w.nbr + up(L.bytes[kPartNbr]);
    w.words = up(size_t(s.n_pad) * sizeof(uint32_t));
    w.total = w.keys + 5 * w.words;
    return w;
}

}  // namespace

void Resorter::reserve(const HostSystem &s, const SlabLayout &L)
{
    const size_t need = resort_space(s, L).total;
    if (mWorkBytes < need) {
        if (mWork)
            hip_check(hipFree(mWork), "hipFree");
        mWork = nullptr;
        mWorkBytes = 0;
        hip_check(hipMalloc(&mWork, need), "hipMalloc (re-sort workspace)");
        mWorkBytes = need;
    }
    size_t temp = 0;
    uint32_t *none = nullptr;
    hip_check(rocprim::radix_sort_pairs(nullptr, temp, none, none, none, none,
                                        size_t(std::max<uint32_t>(s.max_segment, 1)), 0u,
                                        std::max<uint32_t>(site_bits(s), kIdBits)),
              "radix sort (size query)");
    temp = std::max<size_t>(temp, 256);
    if (mSortBytes < temp) {
        if (mSortTemp)
            hip_check(hipFree(mSortTemp), "hipFree");
        mSortTemp = nullptr;
        mSortBytes = 0;
        hip_check(hipMalloc(&mSortTemp, temp), "hipMalloc (re-sort temporary)");
        mSortBytes = temp;
    }
}

void Resorter::run(const HostSystem &s, void *slab, const SlabLayout &L, const SweepBuffers &buf, hipStream_t stream,
                   bool by_id, const uint32_t *slot_of_id)
{
    const uint32_t n_pad = s.n_pad;
    const ResortSpace sp = resort_space(s, L);
    if (mWorkBytes < sp.total || !mSortTemp)
        throw std::logic_error("re-sort workspace not reserved for this system");
    char *w = static_cast<char *>(mWork);
    uint32_t *keys = reinterpret_cast<uint32_t *>(w + sp.keys);
    uint32_t *keys_sorted = reinterpret_cast<uint32_t *>(w + sp.keys + sp.words);
    uint32_t *ident = reinterpret_cast<uint32_t *>(w + sp.keys + 2 * sp.words);
    uint32_t *perm = reinterpret_cast<uint32_t *>(w + sp.keys + 3 * sp.words);
    uint32_t *inv = reinterpret_cast<uint32_t *>(w + sp.keys + 4 * sp.words);
    const uint32_t bits = slot_of_id ? kSortedIndexBits : (by_id ? kIdBits : site_bits(s));

    const dim3 grid((n_pad + kResortBlock - 1) / kResortBlock), block(kResortBlock);
    hipLaunchKernelGGL(k_keys, grid, block, 0, stream, buf.ppos, buf.ids, n_pad, by_id ? 1u : 0u, slot_of_id, keys,
                       ident, perm);
    hip_check(hipGetLastError(), "re-sort keys");
    for (uint32_t c = 0; c < s.n_colours; ++c) {
        const uint32_t b = s.seg_begin[c], n = s.seg_count[c];
        if (n < 2)
            continue;
        size_t tb = mSortBytes;
        hip_check(rocprim::radix_sort_pairs(mSortTemp, tb, keys + b, keys_sorted + b, ident + b, perm + b, size_t(n),
                                            0u, bits, stream),
                  "radix sort");
    }
    hipLaunchKernelGGL(k_inverse, grid, block, 0, stream, perm, n_pad, inv);
    hip_check(hipGetLastError(), "re-sort inverse");
    hipLaunchKernelGGL(k_gather, grid, block, 0, stream, buf.ppos, buf.ids, buf.image, buf.nbr, perm, inv, n_pad,
                       reinterpret_cast<uint32_t *>(w + sp.ppos), reinterpret_cast<uint32_t *>(w + sp.ids),
                       reinterpret_cast<int32_t *>(w + sp.image), reinterpret_cast<uint32_t *>(w + sp.nbr));
    hip_check(hipGetLastError(), "re-sort gather");
    const std::pair<SlabPart, size_t> back[] = {
        {kPartPPos, sp.ppos}, {kPartIds, sp.ids}, {kPartImage, sp.image}, {kPartNbr, sp.nbr}};
    for (const auto &pb : back)
        hip_check(hipMemcpyAsync(slab_part(slab, L, pb.first), w + pb.second, L.bytes[pb.first],
                                 hipMemcpyDeviceToDevice, stream),
                  "re-sort copy");
}

void Resorter::saveSlots(const HostSystem &s, const SweepBuffers &buf, uint32_t *slot_of_id, hipStream_t stream)
{
    const dim3 grid((s.n_pad + kResortBlock - 1) / kResortBlock), block(kResortBlock);
    hipLaunchKernelGGL(k_save_slots, grid, block, 0, stream, buf.ids, s.n_pad, slot_of_id);
    hip_check(hipGetLastError(), "save slots");
}

}  // namespace pscbfm
