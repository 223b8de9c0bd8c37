// bfm_kernels.hip -- the two kernels of one colour substep, for gfx950.
//
// The acceptance rule is that of the reference's check / perform pair
// (UpdaterGPUScBFM_AB_Type.cu:930-1028 and :1148-1190): a monomer of the
// current colour moves iff (1) its proposal passes walls, bonds and excluded
// volume against the state before the substep, and (2) no other proposal of
// the same substep enters a site its cube enters.  The outcome depends only on
// the set of proposals, never on thread order, so every run is bit-exact.
//
// Differences in mechanism (same result):
//  * excluded volume reads 4 bits of an 8-site occupancy lattice instead of 9
//    corner bytes;
//  * conflicts are found with face-site claims: a survivor ORs its 4 entry
//    sites into `claim`; bits that were already claimed are ORed into `coll`;
//    phase 2 rejects a survivor with any of its sites in `coll`;
//  * survivors are compacted per workgroup (wave64 ballot, mbcnt and an LDS
//    prefix over the 4 waves), so phase 2 only runs over moves that can still
//    happen (the reference's TODO, :1222-1228) and no global counter is
//    contended;
//  * no memset: claims are released by their owners in phase 2 and collision
//    words are cleared from per-workgroup lists one substep later;
//  * an accepted move flips its 8 lattice bits with one atomic XOR per word.
#include "bfm_kernels.hpp"

#include <algorithm>

namespace pscbfm {
namespace {

__device__ inline void load_tables(const Tables *t, uint32_t *lds)
{
    const uint32_t *src = reinterpret_cast<const uint32_t *>(t);
    for (int k = threadIdx.x; k < kTableWords; k += blockDim.x)
        lds[k] = src[k];
}

// Number of set bits of `mask` below the calling lane.
__device__ inline uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

__device__ inline int4 moved(const int4 &r, int dir)
{
    const int32_t s = (dir & 1) ? 1 : -1;
    int4 q = r;
    const int a = dir >> 1;
    q.x += a == 0 ? s : 0;
    q.y += a == 1 ? s : 0;
    q.z += a == 2 ? s : 0;
    return q;
}

struct CheckShared {
    uint32_t tables[kTableWords];
    uint32_t wave[kBlock / 64];
    uint32_t coll;
};

// Phase 1 for check block `blk` of `nblk` covering the colour segment
// [begin, begin+count).
template <bool kAllPeriodic>
__device__ inline void check_body(const SweepBuffers &buf, uint32_t begin, uint32_t count, uint64_t age,
                                  uint32_t parity, uint32_t prev_blocks, uint32_t blk, uint32_t nblk,
                                  CheckShared &sh)
{
    if (threadIdx.x == 0)
        sh.coll = 0u;
    load_tables(buf.tables, sh.tables);
    __syncthreads();
    const uint32_t *dep = sh.tables;
    const uint32_t *bond = sh.tables + 3 * 64;
    const Geometry &g = buf.geom;

    // Collision bits listed by the previous check launch (other parity):
    // coll[parity^1] is not read in this kernel.
    {
        uint32_t *coll_prev = buf.coll[parity ^ 1u];
        const uint32_t *list_prev = buf.clist[parity ^ 1u];
        const uint32_t *cnt_prev = buf.ccount[parity ^ 1u];
        for (uint32_t pb = blk; pb < prev_blocks; pb += nblk) {
            const uint32_t n = cnt_prev[pb];
            for (uint32_t k = threadIdx.x; k < n; k += kBlock)
                coll_prev[list_prev[size_t(pb) * kCollSlots + k]] = 0u;
        }
    }

    const uint32_t t = blk * kBlock + threadIdx.x;
    bool pass = false;
    uint2 entry = make_uint2(0u, 0u);
    if (t < count) {
        const uint32_t i = begin + t;
        const uint32_t p0 = buf.ppos[i];
        const int4 r0 = unpack_xyz(p0);
        const uint32_t rn = philox_word0(buf.seed, buf.ids[i], age, kTagMove);
        const int dir = int(rn % 6u);
        const int4 r1 = moved(r0, dir);

        bool ok = true;
        if (!kAllPeriodic) {
            const int a = dir >> 1;
            const int32_t v = comp(r1, a);
            const int32_t hi = a == 0 ? g.wall_hi[0] : (a == 1 ? g.wall_hi[1] : g.wall_hi[2]);
            if (!((g.periodic >> a) & 1u) && (v < 0 || v > hi))
                ok = false;
        }

        // bond partners belong to other colours and do not move this substep
        const uint32_t nn = buf.nnbr[i];
        uint32_t partner[kMaxConnectivity];
#pragma unroll
        for (int j = 0; j < kMaxConnectivity; ++j)
            if (uint32_t(j) < nn)
                partner[j] = buf.nbr[size_t(j) * buf.n_pad + i];
#pragma unroll
        for (int j = 0; j < kMaxConnectivity; ++j)
            if (uint32_t(j) < nn) {
                const int4 q = unpack_xyz(buf.ppos[partner[j]]);
                if (bond_forbidden(bond, q.x - r1.x, q.y - r1.y, q.z - r1.z))
                    ok = false;
            }

        uint32_t s[4], d_face;
        face_sites(g, dep, r0, dir, s, d_face);
        const uint32_t *lat = buf.lattice;
        if (bit_test(lat, s[0]) | bit_test(lat, s[1]) | bit_test(lat, s[2]) | bit_test(lat, s[3]))
            ok = false;

        if (ok) {
            uint32_t w[4], m[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                w[k] = s[k] >> 5;
                m[k] = 1u << (s[k] & 31u);
            }
            merge_words(w, m);
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (m[k]) {
                    const uint32_t hit = atomicOr(&buf.claim[w[k]], m[k]) & m[k];
                    if (hit) {
                        atomicOr(&buf.coll[parity][w[k]], hit);
                        const uint32_t slot = atomicAdd(&sh.coll, 1u);
                        buf.clist[parity][size_t(blk) * kCollSlots + slot] = w[k];
                    }
                }
            pass = true;
            entry = make_uint2(i | (uint32_t(dir) << kSortedIndexBits), p0);
        }
    }

    // per-workgroup compaction: ballot within the wave, LDS prefix over waves
    const uint64_t mball = __ballot(pass);
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0)
        sh.wave[wave] = uint32_t(__popcll(mball));
    __syncthreads();
    uint32_t base = 0;
#pragma unroll
    for (uint32_t k = 0; k < kBlock / 64; ++k)
        base += k < wave ? sh.wave[k] : 0u;
    if (pass)
        buf.surv[size_t(blk) * kBlock + base + lanes_below(mball)] = entry;
    if (threadIdx.x == 0) {
        uint32_t total = 0;
#pragma unroll
        for (uint32_t k = 0; k < kBlock / 64; ++k)
            total += sh.wave[k];
        buf.scount[blk] = total;
        buf.ccount[parity][blk] = sh.coll;
    }
}

template <bool kAllPeriodic>
__global__ __launch_bounds__(kBlock) void k_check(SweepBuffers buf, uint32_t begin, uint32_t count, uint64_t age,
                                                  uint32_t parity, uint32_t prev_blocks)
{
    __shared__ CheckShared sh;
    check_body<kAllPeriodic>(buf, begin, count, age, parity, prev_blocks, blockIdx.x, gridDim.x, sh);
}

struct PerformShared {
    uint32_t tables[kTableWords];
    uint32_t pre[kPerformGroup + 1];
    uint32_t acc;
};

// Phase 2 for perform block `blk`: the survivors of check blocks
// [blk*kPerformGroup, (blk+1)*kPerformGroup) of a check launch of
// check_blocks blocks.
__device__ inline void perform_body(const SweepBuffers &buf, uint32_t check_blocks, uint32_t parity, uint32_t blk,
                                    PerformShared &sh)
{
    uint32_t *s_pre = sh.pre;
    load_tables(buf.tables, sh.tables);
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        s_pre[0] = 0;
        for (uint32_t k = 0; k < kPerformGroup; ++k) {
            const uint32_t cb = blk * kPerformGroup + k;
            acc += cb < check_blocks ? buf.scount[cb] : 0u;
            s_pre[k + 1] = acc;
        }
        sh.acc = 0;
    }
    __syncthreads();
    const uint32_t *dep = sh.tables;
    const Geometry &g = buf.geom;
    const uint32_t *coll = buf.coll[parity];
    const uint32_t total = s_pre[kPerformGroup];

    for (uint32_t t = threadIdx.x; t < total; t += kBlock) {
        uint32_t k = 0;
#pragma unroll
        for (uint32_t q = 1; q < kPerformGroup; ++q)
            k += t >= s_pre[q] ? 1u : 0u;
        const uint2 e = buf.surv[size_t(blk * kPerformGroup + k) * kBlock + (t - s_pre[k])];
        const uint32_t i = e.x & kSortedIndexMask;
        const int dir = int(e.x >> kSortedIndexBits);
        const int4 r0 = unpack_xyz(e.y);
        uint32_t s[4], d_face;
        face_sites(g, dep, r0, dir, s, d_face);
        const bool clash = bit_test(coll, s[0]) | bit_test(coll, s[1]) | bit_test(coll, s[2]) | bit_test(coll, s[3]);

        uint32_t w[8], m[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            w[q] = s[q] >> 5;
            m[q] = 1u << (s[q] & 31u);
        }
        {
            // release the claims (claim[] is not read in this kernel)
            uint32_t cw[4] = {w[0], w[1], w[2], w[3]}, cm[4] = {m[0], m[1], m[2], m[3]};
            merge_words(cw, cm);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (cm[q])
                    buf.claim[cw[q]] = 0u;
        }
        if (!clash) {
            const int a = dir >> 1;
            const uint32_t dm = sel3(g.dmask, a);
            // the face the cube leaves lies two planes behind the one it enters
            const uint32_t d_back = (dir & 1) ? dilated_dec(dilated_dec(d_face, dm), dm)
                                              : dilated_inc(dilated_inc(d_face, dm), dm);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t b = (s[q] & ~dm) | d_back;
                w[4 + q] = b >> 5;
                m[4 + q] = 1u << (b & 31u);
            }
            // the 8 flipped sites are distinct and every other accepted move
            // flips different ones, so XOR sets the entered and clears the
            // vacated bits
            merge_words(w, m);
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (m[q])
                    atomicXor(&buf.lattice[w[q]], m[q]);
            // new wrapped coordinate on the move axis; crossing the box
            // boundary changes the periodic image (only on periodic axes:
            // walls keep the others inside)
            const uint32_t am = sel3(g.mask, a);
            const uint32_t v = uint32_t(comp(r0, a) + ((dir & 1) ? 1 : -1)) & am;
            const uint32_t shift = kCoordBits * uint32_t(a);
            buf.ppos[i] = (e.y & ~(kCoordMask << shift)) | (v << shift);
            if ((dir & 1) ? v == 0u : v == am)
                buf.image[size_t(a) * buf.n_pad + i] += (dir & 1) ? 1 : -1;
        }
        const uint64_t acc = __ballot(!clash);
        if ((threadIdx.x & 63u) == uint32_t(__ffsll((unsigned long long)__ballot(1)) - 1))
            atomicAdd(&sh.acc, uint32_t(__popcll(acc)));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t slot = blk & (kStatSlots - 1);
        if (sh.acc)
            atomicAdd(&buf.accepted[slot], (unsigned long long)sh.acc);
        if (total)
            atomicAdd(&buf.proposals[slot], (unsigned long long)total);
    }
}

__global__ __launch_bounds__(kBlock) void k_perform(SweepBuffers buf, uint32_t check_blocks, uint32_t parity)
{
    __shared__ PerformShared sh;
    perform_body(buf, check_blocks, parity, blockIdx.x, sh);
}

// ---- many systems per launch ------------------------------------------------
// Block b of a batched launch belongs to system blk_sys[b]; the system's own
// blocks start at its check_base / perform_base.  The schedule says which
// colour (if any) each system sweeps in this substep slot.

template <bool kAllPeriodic>
__device__ inline void check_batch_body(const BatchSystem *__restrict__ sys, const uint32_t *__restrict__ blk_sys,
                                        const SchedEntry *__restrict__ sched, uint32_t n_sys, uint32_t slot,
                                        CheckShared &sh)
{
    const uint32_t s = blk_sys[blockIdx.x];
    const SchedEntry e = sched[size_t(slot) * n_sys + s];
    if (e.colour == kNoColour)
        return;
    const BatchSystem &S = sys[s];
    const uint32_t count = S.seg_count[e.colour];
    const uint32_t nb = check_blocks(count);
    const uint32_t blk = blockIdx.x - S.check_base;
    if (blk >= nb)
        return;
    check_body<kAllPeriodic>(S.buf, S.seg_begin[e.colour], count, e.age, e.parity, e.prev_blocks, blk, nb, sh);
}

__global__ __launch_bounds__(kBlock) void k_check_batch(const BatchSystem *__restrict__ sys,
                                                        const uint32_t *__restrict__ blk_sys,
                                                        const SchedEntry *__restrict__ sched, uint32_t n_sys,
                                                        uint32_t slot)
{
    __shared__ CheckShared sh;
    if (sys[blk_sys[blockIdx.x]].buf.geom.periodic == 7u)
        check_batch_body<true>(sys, blk_sys, sched, n_sys, slot, sh);
    else
        check_batch_body<false>(sys, blk_sys, sched, n_sys, slot, sh);
}

__global__ __launch_bounds__(kBlock) void k_perform_batch(const BatchSystem *__restrict__ sys,
                                                          const uint32_t *__restrict__ blk_sys,
                                                          const SchedEntry *__restrict__ sched, uint32_t n_sys,
                                                          uint32_t slot)
{
    __shared__ PerformShared sh;
    const uint32_t s = blk_sys[blockIdx.x];
    const SchedEntry e = sched[size_t(slot) * n_sys + s];
    if (e.colour == kNoColour)
        return;
    const BatchSystem &S = sys[s];
    const uint32_t cb = check_blocks(S.seg_count[e.colour]);
    const uint32_t blk = blockIdx.x - S.perform_base;
    if (blk >= (cb + kPerformGroup - 1) / kPerformGroup)
        return;
    perform_body(S.buf, cb, e.parity, blk, sh);
}

}  // namespace

hipError_t launch_check(const SweepBuffers &buf, uint32_t begin, uint32_t count, uint64_t age, uint32_t parity,
                        uint32_t prev_blocks, hipStream_t stream)
{
    const dim3 grid(check_blocks(count)), block(kBlock);
    if (buf.geom.periodic == 7u)
        hipLaunchKernelGGL(k_check<true>, grid, block, 0, stream, buf, begin, count, age, parity, prev_blocks);
    else
        hipLaunchKernelGGL(k_check<false>, grid, block, 0, stream, buf, begin, count, age, parity, prev_blocks);
    return hipGetLastError();
}

hipError_t launch_perform(const SweepBuffers &buf, uint32_t count, uint32_t parity, hipStream_t stream)
{
    const uint32_t cb = check_blocks(count);
    const uint32_t pb = (cb + kPerformGroup - 1) / kPerformGroup;
    hipLaunchKernelGGL(k_perform, dim3(pb), dim3(kBlock), 0, stream, buf, cb, parity);
    return hipGetLastError();
}

hipError_t launch_check_batch(const BatchSystem *sys, const uint32_t *blk_sys, uint32_t n_blocks,
                              const SchedEntry *sched, uint32_t n_sys, uint32_t slot, hipStream_t stream)
{
    hipLaunchKernelGGL(k_check_batch, dim3(n_blocks), dim3(kBlock), 0, stream, sys, blk_sys, sched, n_sys, slot);
    return hipGetLastError();
}

hipError_t launch_perform_batch(const BatchSystem *sys, const uint32_t *blk_sys, uint32_t n_blocks,
                                const SchedEntry *sched, uint32_t n_sys, uint32_t slot, hipStream_t stream)
{
    hipLaunchKernelGGL(k_perform_batch, dim3(n_blocks), dim3(kBlock), 0, stream, sys, blk_sys, sched, n_sys, slot);
    return hipGetLastError();
}

}  // namespace pscbfm

namespace pscbfm {
namespace {

// Plain streaming copy, 16 bytes per lane per iteration: the reference point
// for HBM counters (a copy of B bytes reads B and writes B).
__global__ __launch_bounds__(kBlock) void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock)
        dst[i] = src[i];
}

}  // namespace

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t stream)
{
    const size_t n = bytes / sizeof(uint4);
    const uint32_t blocks = uint32_t(std::min<size_t>((n + kBlock - 1) / kBlock, 256 * 64));
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(kBlock), 0, stream, static_cast<const uint4 *>(src),
                       static_cast<uint4 *>(dst), n);
    return hipGetLastError();
}

}  // namespace pscbfm
