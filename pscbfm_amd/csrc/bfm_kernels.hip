// bfm_kernels.hip -- the two kernels of one colour substep, for gfx950.
//
// The acceptance rule is that of the reference's check / perform pair
// (UpdaterGPUScBFM_AB_Type.cu:930-1028 and :1148-1190): a monomer of the
// current colour moves iff (1) its proposal passes walls, bonds and excluded
// volume against the state before the substep, and (2) no other proposal of
// the same substep enters a site its cube enters.  The outcome depends only on
// the set of proposals, never on thread order, so every run is bit-exact.
//
// Both phases use the reference's own geometry: the lattice holds one bit per
// monomer at its lower corner, and a move is tested on the 9-site front of
// the moved cube (checkFront, :420-594), first in the occupancy lattice, then
// in a lattice of this substep's moved corners (the reference's mLatticeTmp).
// Differences in mechanism (same result):
//  * 1 bit per site instead of 1 byte, in compact Morton order; the 9 front
//    sites fall into at most 4 words;
//  * a position is the site of its corner, moved and compared by dilated
//    arithmetic on one axis's bits, so neither kernel spreads coordinates;
//  * a survivor marks its moved corner with one atomic OR (no return value);
//  * survivors are compacted per wavefront (wave64 ballot + mbcnt), so phase 2
//    only runs over moves that can still happen (the reference's TODO,
//    :1222-1228) and no global counter is contended; neither kernel uses LDS
//    or a workgroup barrier;
//  * no memset: the marks of a substep are cleared by the next check launch
//    (the claim lattice is double-buffered by substep parity);
//  * an accepted move writes its old and new lattice bytes with plain byte
//    stores: each byte is an aligned 2x2x2 block of sites, which holds at
//    most one corner, and no two accepted moves of a substep share a block.
#include "bfm_kernels.hpp"

#include <hip/hip_ext.h>

namespace pscbfm {
namespace {

// Number of set bits of `mask` below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Value of lane i - D of the same row of 16 lanes (DPP row_shr:D), 0 where
// that lane lies before the row's start.
template <int D>
__device__ __forceinline__ uint32_t row_shr(uint32_t v)
{
    static_assert(D >= 1 && D <= 15, "row_shr shifts by 1..15 lanes");
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x110 + D, 0xF, 0xF, false));
}

// Pointers into global memory.  The kernel bodies access their buffers through
// these: a batched kernel reads its member's pointers from memory, and without
// the address space the compiler would use flat instructions for them (slower,
// and waited for on both the vector and the scalar counter).
template <class T>
using gptr = __attribute__((address_space(1))) T *;

template <class T>
__device__ __forceinline__ gptr<T> as_global(T *p)
{
    return (gptr<T>)p;
}

// Generic view of a global element, for class types (HIP vector types) whose
// member operators expect a generic `this`; the compiler still sees that the
// address is global.
template <class T>
__device__ __forceinline__ T &element(gptr<T> p)
{
    return *(T *)p;
}

// An empty asm statement that takes wave-uniform values in SGPRs: the loads
// that produce them must have completed where it stands.  The batched kernels
// place it before their exit tests, which the compiler would otherwise let the
// loads of a member's buffers sink below (each sunk group is one more
// dependent level of scalar loads before a block's first data load).
__device__ __forceinline__ void in_sgprs(uint32_t v) { asm volatile("" ::"s"(v)); }
template <class T>
__device__ __forceinline__ void in_sgprs(T *p)
{
    asm volatile("" ::"s"(p));
}

// Everything check_body and perform_body read from a member's buffers.
__device__ __forceinline__ void buffers_in_sgprs(const SweepBuffers &b)
{
    in_sgprs(b.ppos);
    in_sgprs(b.ids);
    in_sgprs(b.image);
    in_sgprs(b.nbr);
    in_sgprs(b.n_pad);
    in_sgprs(b.lattice);
    in_sgprs(b.lattice_words);
    in_sgprs(b.claim0);
    in_sgprs(b.claim1);
    in_sgprs(b.surv);
    in_sgprs(b.scount);
    in_sgprs(b.accepted);
    in_sgprs(b.proposals);
    in_sgprs(b.tables);
    in_sgprs(b.rng);
    in_sgprs(uint32_t(b.seed));
    in_sgprs(uint32_t(b.seed >> 32));
    for (int a = 0; a < 3; ++a) {
        in_sgprs(b.geom.mask[a]);
        in_sgprs(b.geom.dmask[a]);
    }
}

// The 512-bit forbidden-bond mask held across a wavefront: lane k holds word
// k & 15, and a lookup is one cross-lane read.  Unlike a table in LDS it needs
// no workgroup barrier before the first data load, and its own load overlaps
// the body's.  Every lane of the wave must be active at a lookup (an inactive
// source lane supplies no data).
struct WaveBondMask {
    uint32_t word;
    __device__ __forceinline__ explicit WaveBondMask(const Tables *t)
        : word(as_global(t)->bond[threadIdx.x & 15u]) {}
    // key: bond_key() of the bond vector (a 9-bit index)
    __device__ __forceinline__ bool forbidden(uint32_t key) const
    {
        // lane key >> 5 (below 16) holds the word; ds_bpermute takes the
        // source lane as a byte address
        return (uint32_t(__builtin_amdgcn_ds_bpermute(int((key >> 5) << 2), int(word))) >> (key & 31u)) & 1u;
    }
};

// A raw buffer resource over `words` 32-bit words at p.  dword 3 of the
// descriptor selects the 32-bit data format and nothing else.  A load through
// it whose byte offset lies at or past the end reads zero and makes no memory
// access (the resource's range check), so a lane can drop out of a load
// without a branch.
constexpr int kRawBufferWord3 = 0x00020000;
constexpr uint32_t kPastEnd = 0xFFFFFFFCu;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t word_rsrc(const uint32_t *p, uint32_t words)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(p), short(0), int(words * 4u), kRawBufferWord3);
}
// Word `word` of the resource (an offset below 2^32 bytes: 32-bit address
// arithmetic per lane).
__device__ __forceinline__ uint32_t load_word(__amdgpu_buffer_rsrc_t r, uint32_t word)
{
    return uint32_t(__builtin_amdgcn_raw_buffer_load_b32(r, int(word * 4u), 0, 0));
}
// Word `word` of the resource where `use`, else zero without an access.
__device__ __forceinline__ uint32_t load_or_zero(__amdgpu_buffer_rsrc_t r, bool use, uint32_t word)
{
    return uint32_t(__builtin_amdgcn_raw_buffer_load_b32(r, int(use ? word * 4u : kPastEnd), 0, 0));
}

// ---- phase 1: the check ------------------------------------------------------
// Pieces shared by the two forms of the check body below.

// The arrays a check reads per lane, as buffer resources (32-bit offsets),
// and the claim lattices of this substep and of the previous one.
struct CheckArrays {
    __amdgpu_buffer_rsrc_t ppos, ids, lattice;
    gptr<uint32_t> claim_cur, claim_prev;
    __device__ __forceinline__ CheckArrays(const SweepBuffers &buf, uint32_t parity)
        : ppos(word_rsrc(buf.ppos, buf.n_pad)), ids(word_rsrc(buf.ids, buf.n_pad)),
          lattice(word_rsrc(buf.lattice, buf.lattice_words)), claim_cur(as_global(pick(parity, buf.claim0, buf.claim1))),
          claim_prev(as_global(pick(parity, buf.claim1, buf.claim0)))
    {
    }
};

// Column j of the partner table.  It has kMaxConnectivity columns; past a
// slot's own partners, up to the largest partner count of the system, it
// holds the slot's own index, so every read below stays in bounds (a lane
// with fewer partners reads its own slot and ignores it), and a lane uses
// only its first nn.  The site load's address is then the partner load alone,
// which the compiler cannot turn into a branch around the load (that would
// make it wait for every load in flight).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t nbr_column(const SweepBuffers &buf, int j)
{
    return word_rsrc(buf.nbr + size_t(j) * buf.n_pad, buf.n_pad);
}

// direction = rn % 6 (.cu:964-967); the generator is uniform per launch
__device__ __forceinline__ int move_direction(const SweepBuffers &buf, uint32_t id, uint64_t age, uint32_t sub_seed)
{
    uint32_t rn;
    if (buf.rng == kRngPhilox)
        rn = philox_word0(buf.seed, id, age, kTagMove);
    else
        rn = int_hash(int_hash(id) ^ sub_seed);
    return int(rn % 6u);
}

// Excluded volume: no corner in the 9-site front.  Only the words that hold
// front sites are read (mask non-zero); the other lanes give their load an
// offset past the lattice's end, which reads zero without an access, so the
// loads stay unconditional.
__device__ __forceinline__ void load_front(const CheckArrays &a, const Front &f, uint32_t (&l)[4])
{
#pragma unroll
    for (int q = 0; q < 4; ++q)
        l[q] = load_or_zero(a.lattice, f.m[q] != 0u, f.w[q]);
}
__device__ __forceinline__ bool front_free(const Front &f, const uint32_t (&l)[4])
{
    return ((l[0] & f.m[0]) | (l[1] & f.m[1]) | (l[2] & f.m[2]) | (l[3] & f.m[3])) == 0u;
}

// Walls and the bonds to the first kEarlyPartners partners (sites q) from
// the moved corner.  Walls: a move along a walled axis may not leave
// [0, L - 2]; the moved corner's coordinate is then L - 1 (from L - 2 upward,
// or from 0 downward modulo L), whose dilated form is the axis's mask.  The
// bond lookups run on every lane; a lane with fewer partners tests its own
// slot and ignores the result.
template <bool kAllPeriodic>
__device__ __forceinline__ bool walls_and_early_bonds_ok(const Geometry &g, const WaveBondMask &bonds, const Front &f,
                                                         int dir, uint32_t nn, const uint32_t (&q)[kEarlyPartners])
{
    bool ok = true;
    if (!kAllPeriodic && !((g.periodic >> (dir >> 1)) & 1u) && (f.new_site & f.ma) == f.ma)
        ok = false;
    const uint32_t r1x = f.new_site & kLow9[0], r1y = f.new_site & kLow9[1], r1z = f.new_site & kLow9[2];
#pragma unroll
    for (int j = 0; j < kEarlyPartners; ++j) {
        const bool bad = bonds.forbidden(bond_key(q[j], r1x, r1y, r1z));
        if (uint32_t(j) < nn && bad)
            ok = false;
    }
    return ok;
}

// The bonds to the partners beyond the early columns, in a wavefront with a
// lane that has some (a network's crosslinker); every lane of the wave calls
// it.  Their indices and sites are read here, after the counts, and live only
// in this call, so waves without such lanes carry none.
__device__ __forceinline__ bool late_bonds_ok(const SweepBuffers &buf, const CheckArrays &a,
                                              const WaveBondMask &bonds, const Front &f, uint32_t i, uint32_t nn)
{
    bool ok = true;
    uint32_t nn_wave = uint32_t(kEarlyPartners);
#pragma unroll
    for (int j = kEarlyPartners; j < kMaxConnectivity; ++j)
        nn_wave = __ballot(nn > uint32_t(j)) ? uint32_t(j + 1) : nn_wave;
    uint32_t partner[kMaxConnectivity - kEarlyPartners], q[kMaxConnectivity - kEarlyPartners];
#pragma unroll
    for (int j = kEarlyPartners; j < kMaxConnectivity; ++j)
        if (uint32_t(j) < nn_wave)
            partner[j - kEarlyPartners] = load_word(nbr_column(buf, j), i);
#pragma unroll
    for (int j = kEarlyPartners; j < kMaxConnectivity; ++j)
        if (uint32_t(j) < nn_wave)
            q[j - kEarlyPartners] = load_word(a.ppos, partner[j - kEarlyPartners]);
    const uint32_t r1x = f.new_site & kLow9[0], r1y = f.new_site & kLow9[1], r1z = f.new_site & kLow9[2];
#pragma unroll
    for (int j = kEarlyPartners; j < kMaxConnectivity; ++j)
        if (uint32_t(j) < nn_wave) {
            const bool bad = bonds.forbidden(bond_key(q[j - kEarlyPartners], r1x, r1y, r1z));
            if (uint32_t(j) < nn && bad)
                ok = false;
        }
    return ok;
}

// A survivor marks its moved corner in the reference's tentative lattice
// (dpLatticeTmp, :1012-1023), one bit per moved corner, and keeps its entry
// (sorted index | direction, old corner) for the wave's compaction.
__device__ __forceinline__ uint2 claim(const CheckArrays &a, const Front &f, uint32_t i, int dir, uint32_t s0)
{
    atomicOr((uint32_t *)&a.claim_cur[f.new_site >> 5], 1u << (f.new_site & 31u));
    return make_uint2(i | (uint32_t(dir) << kSortedIndexBits), s0);
}

// Wave compaction of one run's survivors into its 64 slots of surv.
__device__ __forceinline__ void compact(const SweepBuffers &buf, uint32_t run, bool pass, uint2 entry)
{
    const uint64_t mball = __ballot(pass);
    if (pass)
        element(as_global(buf.surv) + size_t(run) * kWave + lanes_below(mball)) = entry;
    if ((threadIdx.x & (kWave - 1)) == 0)
        as_global(buf.scount)[run] = uint32_t(__popcll(mball));
}

// The previous substep's claims (the reference zeroes its tentative lattice in
// a third kernel, :1240-1262): nobody reads claim[parity^1] in this kernel, so
// the grid's threads clear all of it, each its share of one part of the
// lattice (a ClearPart), with coalesced 16-byte stores; at the densities of a
// melt a survivor's mark shares every 64-byte line with others, so clearing
// marks one by one writes more.  A wave issues its share first, before its own
// loads: the stores then drain while it waits on them, instead of at the end
// of the last waves (27.2 -> 26.5 us per check launch on the relaxed 512^3
// melt).
struct ClearPart {
    uint32_t t, threads;   // this thread's number among the part's threads
    uint32_t lo, hi;       // the part: 16-byte words [lo, hi) of the claim lattice
};
__device__ __forceinline__ void clear_previous_claims(const CheckArrays &a, const ClearPart &p)
{
    const gptr<uint4> cp = reinterpret_cast<gptr<uint4>>(a.claim_prev);
    for (uint32_t k = p.lo + p.t; k < p.hi; k += p.threads)
        element(cp + k) = make_uint4(0u, 0u, 0u, 0u);
}
// The whole lattice over the launch's nblk blocks, block blk of them.
__device__ __forceinline__ ClearPart whole_clear(const SweepBuffers &buf, uint32_t blk, uint32_t nblk,
                                                 uint32_t threads_per_block)
{
    return ClearPart{blk * threads_per_block + threadIdx.x, nblk * threads_per_block, 0u, buf.lattice_words / 4u};
}
// With XCD ranges: the blocks on XCD x (physical blocks x, x + 8, ...) clear
// the x-th eighth of the lattice, the region its own range of Morton-ordered
// monomers lies in, so that every XCD stores an equal share.  Numbered by
// logical block over the whole lattice, the first blocks, which alone store
// where the lattice has fewer 16-byte words than the launch has threads, are
// the ranges of the first XCDs.  For the launches of two runs per lane (4
// and more generations of waves) the equal shares measured faster (26.5 ->
// 25.7 us on the relaxed 512^3 melt); for one run per lane (about one
// generation: the network's 532,480-monomer colours) slower (7.7 -> 8.3 us),
// and so did every other spread tried there, so those keep the whole-lattice
// numbering.
__device__ __forceinline__ ClearPart xcd_clear(const SweepBuffers &buf, uint32_t nblk, uint32_t threads_per_block)
{
    if (!buf.xcd_ranges || nblk < kXcds)
        return whole_clear(buf, blockIdx.x, nblk, threads_per_block);
    const uint32_t x = blockIdx.x % kXcds, i = blockIdx.x / kXcds;
    const uint32_t n_x = nblk / kXcds + (x < nblk % kXcds ? 1u : 0u);
    const uint32_t w = buf.lattice_words / 4u;
    return ClearPart{i * threads_per_block + threadIdx.x, n_x * threads_per_block, uint32_t(uint64_t(x) * w / kXcds),
                     uint32_t(uint64_t(x + 1) * w / kXcds)};
}

// Phase 1 for check block `blk` of `nblk` covering the colour segment
// [begin, begin+count), one monomer per lane; every wavefront compacts its
// own survivors into surv.  Two levels of dependent loads per wave: the
// slot's own words and its first kEarlyPartners partner indices, then the
// front words and the partner sites.  Every load of the two stages is
// unconditional, so the compiler knows how many are in flight and waits for
// exactly the ones each stage needs (a load under a branch makes it wait for
// all of them).  Every lane of the wave runs the body (lanes past the segment
// on its last monomer), so that the cross-lane bond lookups find every source
// lane active; only lanes inside the segment may pass.
template <bool kAllPeriodic, bool kCubic, uint32_t kThreads>
__device__ __forceinline__ void check_body_one(const SweepBuffers &buf, uint32_t begin, uint32_t count, uint64_t age,
                                               uint32_t sub_seed, uint32_t parity, uint32_t blk, uint32_t nblk,
                                               const ClearPart &clear)
{
    const CheckArrays a(buf, parity);
    const uint32_t gwave = wave_uniform(blk * (kThreads / kWave) + (threadIdx.x >> 6));
    const uint32_t t = blk * kThreads + threadIdx.x;
    const WaveBondMask bonds(buf.tables);
    clear_previous_claims(a, clear);
    bool pass = false;
    uint2 entry = make_uint2(0u, 0u);
    if (count != 0) {
        const bool inside = t < count;
        const uint32_t i = begin + (inside ? t : count - 1u);
        const uint32_t idn = load_word(a.ids, i);
        const uint32_t s0 = load_word(a.ppos, i);
        uint32_t partner[kEarlyPartners];
#pragma unroll
        for (int j = 0; j < kEarlyPartners; ++j)
            partner[j] = load_word(nbr_column(buf, j), i);
        const uint32_t nn = idn >> kIdBits;
        const bool more = __ballot(nn > uint32_t(kEarlyPartners)) != 0ull;
        const int dir = move_direction(buf, idn & kIdMask, age, sub_seed);
        const Front f = front_of<kCubic>(buf.geom, s0, dir);
        uint32_t l[4];
        load_front(a, f, l);
        // bond partners belong to other colours and do not move this substep;
        // every site load is issued before the first is used
        uint32_t q[kEarlyPartners];
#pragma unroll
        for (int j = 0; j < kEarlyPartners; ++j)
            q[j] = load_word(a.ppos, partner[j]);
        bool ok = walls_and_early_bonds_ok<kAllPeriodic>(buf.geom, bonds, f, dir, nn, q);
        if (more && !late_bonds_ok(buf, a, bonds, f, i, nn))
            ok = false;
        if (ok && front_free(f, l) && inside) {
            entry = claim(a, f, i, dir, s0);
            pass = true;
        }
    }
    compact(buf, gwave, pass, entry);
}

// The same with kItems >= 2 monomers per lane: a wavefront takes kItems runs
// of 64 consecutive monomers (one per lane in each), issues the loads of all
// of them before using any, and compacts each run's survivors into its own 64
// slots of surv.  Run r of the segment always covers monomers
// [64 r, 64 r + 64), so the survivors (and the perform kernel's work) do not
// depend on the workgroup size or on kItems.  (With one run per lane this
// form measured slower than check_body_one, whose order of loads and
// arithmetic the compiler keeps; with two, 6 % faster on the relaxed 512^3
// melt.)
template <bool kAllPeriodic, bool kCubic, uint32_t kThreads, uint32_t kItems>
__device__ __forceinline__ void check_body_runs(const SweepBuffers &buf, uint32_t begin, uint32_t count, uint64_t age,
                                                uint32_t sub_seed, uint32_t parity, uint32_t blk, uint32_t nblk,
                                                const ClearPart &clear)
{
    const CheckArrays a(buf, parity);
    const uint32_t gwave = wave_uniform(blk * (kThreads / kWave) + (threadIdx.x >> 6));
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const WaveBondMask bonds(buf.tables);
    clear_previous_claims(a, clear);
    bool pass[kItems];
    uint2 entry[kItems];
#pragma unroll
    for (uint32_t it = 0; it < kItems; ++it) {
        pass[it] = false;
        entry[it] = make_uint2(0u, 0u);
    }
    if (count != 0) {
        bool inside[kItems];
        uint32_t i[kItems], idn[kItems], s0[kItems], partner[kItems][kEarlyPartners];
#pragma unroll
        for (uint32_t it = 0; it < kItems; ++it) {
            const uint32_t ti = (gwave * kItems + it) * kWave + lane;
            inside[it] = ti < count;
            i[it] = begin + (inside[it] ? ti : count - 1u);
            idn[it] = load_word(a.ids, i[it]);
            s0[it] = load_word(a.ppos, i[it]);
#pragma unroll
            for (int j = 0; j < kEarlyPartners; ++j)
                partner[it][j] = load_word(nbr_column(buf, j), i[it]);
        }
        int dir[kItems];
        bool more[kItems];
        Front f[kItems];
        uint32_t l[kItems][4], q[kItems][kEarlyPartners];
#pragma unroll
        for (uint32_t it = 0; it < kItems; ++it) {
            more[it] = __ballot((idn[it] >> kIdBits) > uint32_t(kEarlyPartners)) != 0ull;
            dir[it] = move_direction(buf, idn[it] & kIdMask, age, sub_seed);
            f[it] = front_of<kCubic>(buf.geom, s0[it], dir[it]);
            load_front(a, f[it], l[it]);
#pragma unroll
            for (int j = 0; j < kEarlyPartners; ++j)
                q[it][j] = load_word(a.ppos, partner[it][j]);
        }
#pragma unroll
        for (uint32_t it = 0; it < kItems; ++it) {
            const uint32_t nn = idn[it] >> kIdBits;
            bool ok = walls_and_early_bonds_ok<kAllPeriodic>(buf.geom, bonds, f[it], dir[it], nn, q[it]);
            if (more[it] && !late_bonds_ok(buf, a, bonds, f[it], i[it], nn))
                ok = false;
            if (ok && front_free(f[it], l[it]) && inside[it]) {
                entry[it] = claim(a, f[it], i[it], dir[it], s0[it]);
                pass[it] = true;
            }
        }
    }
#pragma unroll
    for (uint32_t it = 0; it < kItems; ++it)
        compact(buf, gwave * kItems + it, pass[it], entry[it]);
}

// Phase 1 with kItems monomers per lane.
template <bool kAllPeriodic, bool kCubic, uint32_t kThreads = kBlock, uint32_t kItems = 1>
__device__ __forceinline__ void check_body(const SweepBuffers &buf, uint32_t begin, uint32_t count, uint64_t age,
                                           uint32_t sub_seed, uint32_t parity, uint32_t blk, uint32_t nblk,
                                           const ClearPart &clear)
{
    if constexpr (kItems == 1)
        check_body_one<kAllPeriodic, kCubic, kThreads>(buf, begin, count, age, sub_seed, parity, blk, nblk, clear);
    else
        check_body_runs<kAllPeriodic, kCubic, kThreads, kItems>(buf, begin, count, age, sub_seed, parity, blk, nblk,
                                                                clear);
}

template <bool kAllPeriodic, bool kCubic, uint32_t kThreads, uint32_t kItems>
__global__ __launch_bounds__(kThreads) void k_check(SweepBuffers buf, uint32_t begin, uint32_t count, uint64_t age,
                                                    uint32_t sub_seed, uint32_t parity)
{
    // the grid size from the arguments (launch_check sizes the grid so), not
    // from the launch's implicit arguments, which are read after a branch
    const uint32_t nblk = check_blocks(count, kThreads * kItems);
    const uint32_t blk = buf.xcd_ranges ? xcd_logical(blockIdx.x, nblk) : blockIdx.x;
    check_body<kAllPeriodic, kCubic, kThreads, kItems>(
        buf, begin, count, age, sub_seed, parity, blk, nblk,
        kItems == 1 ? whole_clear(buf, blk, nblk, kThreads) : xcd_clear(buf, nblk, kThreads));
}

// Phase 2: every wave of perform block `blk` takes the survivors of kGroup
// consecutive check waves of a check launch of check_waves wavefronts.
template <bool kCubic, uint32_t kGroup>
__device__ __forceinline__ void perform_body(const SweepBuffers &buf, uint32_t check_waves, uint32_t parity,
                                             uint32_t blk)
{
    const Geometry &g = buf.geom;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t pwave = wave_uniform(blk * kWavesPerBlock + (threadIdx.x >> 6));
    const uint32_t cw0 = pwave * kGroup;
    const __amdgpu_buffer_rsrc_t claim = word_rsrc(pick(parity, buf.claim0, buf.claim1), buf.lattice_words);
    const gptr<const uint2> surv = as_global<const uint2>(buf.surv);
    const gptr<const uint32_t> scount = as_global<const uint32_t>(buf.scount);
    const gptr<uint32_t> lattice = as_global(buf.lattice);
    const gptr<uint32_t> ppos = as_global(buf.ppos);

    // survivor counts of the group's check waves and their inclusive prefix
    // sums, held by lanes 0..kGroup-1 (within the first row of 16 lanes, so
    // the scan is row shifts between lanes, not a trip through LDS; a lane
    // whose source lies before the row start adds 0)
    static_assert(kGroup <= 16, "the survivor-count scan runs within one DPP row");
    uint32_t incl = 0;
    if (lane < kGroup && cw0 + lane < check_waves)
        incl = scount[cw0 + lane];
    incl += row_shr<1>(incl);
    if (kGroup > 2)
        incl += row_shr<2>(incl);
    if (kGroup > 4)
        incl += row_shr<4>(incl);
    if (kGroup > 8)
        incl += row_shr<8>(incl);
    uint32_t ends[kGroup];
#pragma unroll
    for (uint32_t k = 0; k < kGroup; ++k)
        ends[k] = uint32_t(__builtin_amdgcn_readlane(int(incl), int(k)));
    const uint32_t total = ends[kGroup - 1];

    // slot in surv of the group's survivor t: check wave k of the group holds
    // survivors [ends[k-1], ends[k])
    auto slot = [&](uint32_t t) {
        uint32_t k = 0, base = 0;
#pragma unroll
        for (uint32_t q = 0; q + 1 < kGroup; ++q)
            if (t >= ends[q]) {
                k = q + 1;
                base = ends[q];
            }
        return size_t(cw0 + k) * kWave + (t - base);
    };
    // another survivor's moved corner in the front: the moved cubes overlap
    // and both stay (the reference's second front check on the tentative
    // lattice, :1148-1190)
    auto clashes = [&](const Front &f, const uint32_t (&k4)[4]) {
        return ((k4[0] & f.m[0]) | (k4[1] & f.m[1]) | (k4[2] & f.m[2]) | (k4[3] & f.m[3])) != 0u;
    };
    auto apply = [&](const uint2 &e, const Front &f) {
        // Vacate the old corner, occupy the new one.  Byte j of the lattice
        // is the aligned 2x2x2 block of sites j*8 .. j*8+7 (the lowest Morton
        // bits are x0, y0, z0 in every geometry), and two corners in one
        // block would be overlapping cubes: a valid state has at most one
        // corner per byte.  An accepted move's old block holds only its own
        // corner, its new block holds none, and no other accepted move of the
        // substep touches either (the front tests reject every pair whose
        // cubes would overlap).  So each move owns its two bytes and writes
        // their new values with plain byte stores: no atomics, no reads.
        const uint32_t i = e.x & kSortedIndexMask;
        const int dir = int(e.x >> kSortedIndexBits);
        const uint32_t old_site = e.y;
        const gptr<uint8_t> blocks = (gptr<uint8_t>)lattice;
        const uint32_t nb = f.new_site >> 3;
        if ((old_site >> 3) != nb)
            blocks[old_site >> 3] = uint8_t(0);
        blocks[nb] = uint8_t(1u << (f.new_site & 7u));
        ppos[i] = f.new_site;
        // crossing the box boundary changes the periodic image (only on
        // periodic axes: walls keep the others inside): the moved corner's
        // coordinate on the move axis wrapped to 0 upward or to L - 1 downward
        const uint32_t a1 = f.new_site & f.ma;
        if ((dir & 1) ? a1 == 0u : a1 == f.ma)
            as_global(buf.image)[size_t(dir >> 1) * buf.n_pad + i] += (dir & 1) ? 1 : -1;
    };
    auto front_of_entry = [&](const uint2 &e) {
        return front_of<kCubic>(g, e.y, int(e.x >> kSortedIndexBits));
    };

    // The survivors in rounds of 64 lanes.  The first kRounds rounds (enough
    // for a group's usual count) issue all their loads before any store: a
    // store may alias a later round's loads, so in a plain loop every round
    // would wait for the one before.
    constexpr uint32_t kRounds = kGroup >= 8 ? kGroup / 4 : 1;
    uint32_t acc = 0;
    uint2 e[kRounds];
    bool live[kRounds];
#pragma unroll
    for (uint32_t r = 0; r < kRounds; ++r) {
        live[r] = lane + r * kWave < total;
        e[r] = live[r] ? element(surv + slot(lane + r * kWave)) : make_uint2(0u, 0u);
    }
    Front f[kRounds];
    uint32_t k4[kRounds][4];
#pragma unroll
    for (uint32_t r = 0; r < kRounds; ++r) {
        f[r] = front_of_entry(e[r]);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            k4[r][q] = load_or_zero(claim, live[r] && f[r].m[q] != 0u, f[r].w[q]);   // no access where unused
    }
#pragma unroll
    for (uint32_t r = 0; r < kRounds; ++r) {
        const bool go = live[r] && !clashes(f[r], k4[r]);
        if (go)
            apply(e[r], f[r]);
        acc += uint32_t(__popcll(__ballot(go)));
    }
    // the rest, one round at a time (more survivors than kRounds rounds hold)
    for (uint32_t t = lane + kRounds * kWave; t < total; t += kWave) {
        const uint2 e1 = element(surv + slot(t));
        const Front f1 = front_of_entry(e1);
        uint32_t c4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            c4[q] = load_or_zero(claim, f1.m[q] != 0u, f1.w[q]);
        const bool go = !clashes(f1, c4);
        if (go)
            apply(e1, f1);
        acc += uint32_t(__popcll(__ballot(go)));
    }
    if (lane == 0 && total) {
        atomicAdd((unsigned long long *)&as_global(buf.accepted)[pwave & (kStatSlots - 1)], (unsigned long long)acc);
        atomicAdd((unsigned long long *)&as_global(buf.proposals)[pwave & (kStatSlots - 1)],
                  (unsigned long long)total);
    }
}

template <bool kCubic, uint32_t kGroup>
__global__ __launch_bounds__(kBlock) void k_perform(SweepBuffers buf, uint32_t check_waves, uint32_t parity)
{
    // the grid size from the arguments, as in k_check
    const uint32_t blk = buf.xcd_ranges ? xcd_logical(blockIdx.x, perform_blocks_for(check_waves, kGroup)) : blockIdx.x;
    perform_body<kCubic, kGroup>(buf, check_waves, parity, blk);
}

// ---- many systems per launch ------------------------------------------------
// One thread per (system, MCS of the chunk) writes that MCS's slots.
__global__ __launch_bounds__(kBlock) void k_schedule(const BatchSystem *__restrict__ sys,
                                                     const ChunkEntry *__restrict__ chunk, uint32_t n_sys,
                                                     uint32_t mcs, uint32_t max_colours, SchedEntry *__restrict__ sched)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n_sys * mcs)
        return;
    const uint32_t j = t % n_sys, m = t / n_sys;
    const BatchSystem &S = sys[j];
    const ChunkEntry ce = chunk[j];
    const uint32_t nc = S.n_colours;
    const uint64_t age = ce.age + m;
    for (uint32_t k = 0; k < max_colours; ++k) {
        SchedEntry e;
        e.age = age;
        e.pad = 0;
        if (k < nc) {
            const uint32_t c = colour_at(S.buf.seed, age, nc, k);
            e.colour = c;
            e.parity = uint32_t(ce.substeps + uint64_t(m) * nc + k) & 1u;
            e.sub_seed = substep_seed(S.buf.seed, age, k);
            e.begin = as_global(S.seg_begin)[c];
            e.count = as_global(S.seg_count)[c];
        } else {
            e.colour = kNoColour;
            e.parity = 0;
            e.sub_seed = 0;
            e.begin = 0;
            e.count = 0;
        }
        sched[size_t(m * max_colours + k) * n_sys + j] = e;
    }
}

// The prologue of a batched block is a chain of dependent loads (block entry,
// then the system's slot entry and buffers) before the body's first data
// load; the slot entry and the buffers depend on the system only, so both are
// read before either exit test.  n_blocks is the grid size, passed with the
// arguments rather than read from the launch's implicit arguments later.
// A batched block takes one monomer per lane (two measured the same on the
// 512-system mix, 173.5 against 173.2 ms of check kernels per 1000 MCS).
template <bool kAllPeriodic, bool kCubic>
__global__ __launch_bounds__(kBlock) void k_check_batch(const BatchSystem *__restrict__ sys,
                                                        const BlockEntry *__restrict__ blk_tab,
                                                        const SchedEntry *__restrict__ sched, uint32_t n_sys,
                                                        uint32_t slot, uint32_t xcd_ranges, uint32_t n_blocks)
{
    const uint32_t lb = xcd_ranges ? xcd_logical(blockIdx.x, n_blocks) : blockIdx.x;
    const BlockEntry be = blk_tab[lb];
    // the system's buffers by value: the body's stores cannot alias them, so
    // they stay in registers as a plain launch's arguments do; loaded before
    // the slot entry, so that both are in flight together
    const SweepBuffers buf = sys[be.sys].buf;
    const SchedEntry e = sched[size_t(slot) * n_sys + be.sys];
    buffers_in_sgprs(buf);
    in_sgprs(e.begin);
    in_sgprs(e.sub_seed);
    in_sgprs(e.parity);
    in_sgprs(uint32_t(e.age));
    in_sgprs(uint32_t(e.age >> 32));
    const uint32_t nb = check_blocks(e.count);
    if (e.colour == kNoColour || be.blk >= nb)
        return;
    const uint32_t blk = be.blk;
    check_body<kAllPeriodic, kCubic>(buf, e.begin, e.count, e.age, e.sub_seed, e.parity, blk, nb,
                                     whole_clear(buf, blk, nb, kBlock));
}

template <bool kCubic, uint32_t kGroup>
__global__ __launch_bounds__(kBlock) void k_perform_batch(const BatchSystem *__restrict__ sys,
                                                          const BlockEntry *__restrict__ blk_tab,
                                                          const SchedEntry *__restrict__ sched, uint32_t n_sys,
                                                          uint32_t slot, uint32_t xcd_ranges, uint32_t n_blocks)
{
    const uint32_t lb = xcd_ranges ? xcd_logical(blockIdx.x, n_blocks) : blockIdx.x;
    const BlockEntry be = blk_tab[lb];
    const SweepBuffers buf = sys[be.sys].buf;
    const SchedEntry e = sched[size_t(slot) * n_sys + be.sys];
    buffers_in_sgprs(buf);
    in_sgprs(e.parity);
    in_sgprs(e.count);
    const uint32_t cb = check_blocks(e.count);
    if (e.colour == kNoColour || be.blk >= perform_blocks(cb, kGroup))
        return;
    const uint32_t blk = be.blk;
    perform_body<kCubic, kGroup>(buf, cb * kWavesPerBlock, e.parity, blk);
}

// Unwrapped coordinates of every monomer in original-id order: the
// reference's doCopyBack (UpdaterGPUScBFM_AB_Type.cu:3402-3484: unsort, and
// position + image * box from :1765-1769) done on the device, so that one
// 12-byte record per monomer crosses to the host.  One lane per sorted slot;
// padding slots (id word ~0) write nothing.  The coordinates are gathered
// from the corner's site; the sum is taken modulo 2^32, as the int32 result
// of the 64-bit sum.
__global__ __launch_bounds__(kBlock) void k_unwrap(const uint32_t *__restrict__ ppos, const uint32_t *__restrict__ ids,
                                                   const int32_t *__restrict__ image, uint32_t n_pad, Geometry g,
                                                   uint32_t bx, uint32_t by, uint32_t bz, int32_t *__restrict__ xyz)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n_pad)
        return;
    const uint32_t idn = ids[t];
    if (idn == 0xFFFFFFFFu)
        return;
    const size_t id = idn & kIdMask;
    const uint32_t p = ppos[t];
    const uint32_t box[3] = {bx, by, bz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t w = site_coord(p, g.dmask[a]);
        const uint32_t img = uint32_t(image[size_t(a) * n_pad + t]);
        xyz[3 * id + size_t(a)] = int32_t(w + img * box[a]);
    }
}

// Plain streaming copy, one 16-byte element per lane and one pass of the
// grid: the reference point for HBM counters (a copy of B bytes reads B and
// writes B) and for the achievable bandwidth.  On the MI355X this form
// streams 6.2 TB/s; a grid-stride loop over 16384 blocks reached 4.6 TB/s.
__global__ __launch_bounds__(kBlock) void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n)
        dst[i] = src[i];
}

}  // namespace

// Launches `kernel`; given a start and a stop event, the events take the
// timestamps of the kernel's own dispatch (hipExtLaunchKernelGGL), the
// interval a kernel trace reports, instead of the span between two markers
// queued around it (which adds the dispatch and the end-of-kernel flush).
template <class... P, class... A>
void dispatch(void (*kernel)(P...), dim3 grid, uint32_t threads, hipStream_t stream, hipEvent_t start,
              hipEvent_t stop, A... args)
{
    if (start && stop)
        hipExtLaunchKernelGGL(kernel, grid, dim3(threads), 0, stream, start, stop, 0, P(args)...);
    else
        hipLaunchKernelGGL(kernel, grid, dim3(threads), 0, stream, P(args)...);
}

// B monomers per workgroup: B / I threads with I runs of 64 per wavefront
template <uint32_t B, uint32_t I>
void launch_check_threads(const SweepBuffers &buf, uint32_t begin, uint32_t count, uint64_t age, uint32_t sub_seed,
                          uint32_t parity, hipStream_t stream, hipEvent_t start, hipEvent_t stop)
{
    constexpr uint32_t T = B / I;
    const dim3 grid(check_blocks(count, B));
    const bool periodic = buf.geom.periodic == 7u, cubic = buf.geom.cubic != 0u;
    auto kernel = periodic ? (cubic ? k_check<true, true, T, I> : k_check<true, false, T, I>)
                           : (cubic ? k_check<false, true, T, I> : k_check<false, false, T, I>);
    dispatch(kernel, grid, T, stream, start, stop, buf, begin, count, age, sub_seed, parity);
}

hipError_t launch_check(const SweepBuffers &buf, uint32_t begin, uint32_t count, uint64_t age, uint32_t sub_seed,
                        uint32_t parity, hipStream_t stream, hipEvent_t start, hipEvent_t stop, uint32_t block,
                        uint32_t items)
{
    if (!valid_check_items(items))
        return hipErrorInvalidValue;
    const bool two = check_items_for(block, items) == 2u;
    switch (block) {
    case 64: launch_check_threads<64, 1>(buf, begin, count, age, sub_seed, parity, stream, start, stop); break;
    case 128:
        two ? launch_check_threads<128, 2>(buf, begin, count, age, sub_seed, parity, stream, start, stop)
            : launch_check_threads<128, 1>(buf, begin, count, age, sub_seed, parity, stream, start, stop);
        break;
    case 256:
        two ? launch_check_threads<256, 2>(buf, begin, count, age, sub_seed, parity, stream, start, stop)
            : launch_check_threads<256, 1>(buf, begin, count, age, sub_seed, parity, stream, start, stop);
        break;
    case 512:
        two ? launch_check_threads<512, 2>(buf, begin, count, age, sub_seed, parity, stream, start, stop)
            : launch_check_threads<512, 1>(buf, begin, count, age, sub_seed, parity, stream, start, stop);
        break;
    case 1024:
        two ? launch_check_threads<1024, 2>(buf, begin, count, age, sub_seed, parity, stream, start, stop)
            : launch_check_threads<1024, 1>(buf, begin, count, age, sub_seed, parity, stream, start, stop);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <uint32_t G>
void launch_perform_group(const SweepBuffers &buf, uint32_t waves, uint32_t parity, hipStream_t stream,
                          hipEvent_t start, hipEvent_t stop)
{
    const dim3 grid(perform_blocks_for(waves, G));
    auto kernel = buf.geom.cubic != 0u ? k_perform<true, G> : k_perform<false, G>;
    dispatch(kernel, grid, kBlock, stream, start, stop, buf, waves, parity);
}

hipError_t launch_perform(const SweepBuffers &buf, uint32_t count, uint32_t parity, hipStream_t stream,
                          uint32_t group, hipEvent_t start, hipEvent_t stop, uint32_t check_block)
{
    if (!valid_check_block(check_block))
        return hipErrorInvalidValue;
    const uint32_t waves = check_waves(count, check_block);
    switch (group) {
    case 2: launch_perform_group<2>(buf, waves, parity, stream, start, stop); break;
    case 8: launch_perform_group<8>(buf, waves, parity, stream, start, stop); break;
    default: launch_perform_group<4>(buf, waves, parity, stream, start, stop); break;
    }
    return hipGetLastError();
}

hipError_t launch_schedule(const BatchSystem *sys, const ChunkEntry *chunk, uint32_t n_sys, uint32_t mcs,
                           uint32_t max_colours, SchedEntry *sched, hipStream_t stream)
{
    const uint32_t n = n_sys * mcs;
    hipLaunchKernelGGL(k_schedule, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, sys, chunk, n_sys, mcs,
                       max_colours, sched);
    return hipGetLastError();
}

hipError_t launch_check_batch(uint32_t cls, const BatchSystem *sys, const BlockEntry *blk_tab, uint32_t n_blocks,
                              const SchedEntry *sched, uint32_t n_sys, uint32_t slot, uint32_t xcd_ranges,
                              hipStream_t stream)
{
    const dim3 grid(n_blocks), block(kBlock);
    switch (cls) {
    case kClassPeriodicCubic:
        hipLaunchKernelGGL((k_check_batch<true, true>), grid, block, 0, stream, sys, blk_tab, sched, n_sys, slot,
                           xcd_ranges, n_blocks);
        break;
    case kClassPeriodic:
        hipLaunchKernelGGL((k_check_batch<true, false>), grid, block, 0, stream, sys, blk_tab, sched, n_sys, slot,
                           xcd_ranges, n_blocks);
        break;
    case kClassCubic:
        hipLaunchKernelGGL((k_check_batch<false, true>), grid, block, 0, stream, sys, blk_tab, sched, n_sys, slot,
                           xcd_ranges, n_blocks);
        break;
    default:
        hipLaunchKernelGGL((k_check_batch<false, false>), grid, block, 0, stream, sys, blk_tab, sched, n_sys, slot,
                           xcd_ranges, n_blocks);
        break;
    }
    return hipGetLastError();
}

template <uint32_t G>
void launch_perform_batch_group(bool cubic, dim3 grid, const BatchSystem *sys, const BlockEntry *blk_tab,
                                const SchedEntry *sched, uint32_t n_sys, uint32_t slot, uint32_t xcd_ranges,
                                hipStream_t stream)
{
    if (cubic)
        hipLaunchKernelGGL((k_perform_batch<true, G>), grid, dim3(kBlock), 0, stream, sys, blk_tab, sched, n_sys,
                           slot, xcd_ranges, grid.x);
    else
        hipLaunchKernelGGL((k_perform_batch<false, G>), grid, dim3(kBlock), 0, stream, sys, blk_tab, sched, n_sys,
                           slot, xcd_ranges, grid.x);
}

hipError_t launch_perform_batch(uint32_t cls, const BatchSystem *sys, const BlockEntry *blk_tab, uint32_t n_blocks,
                                const SchedEntry *sched, uint32_t n_sys, uint32_t slot, uint32_t xcd_ranges,
                                hipStream_t stream, uint32_t group)
{
    const bool cubic = cls == kClassPeriodicCubic || cls == kClassCubic;
    const dim3 grid(n_blocks);
    switch (group) {
    case 2: launch_perform_batch_group<2>(cubic, grid, sys, blk_tab, sched, n_sys, slot, xcd_ranges, stream); break;
    case 8: launch_perform_batch_group<8>(cubic, grid, sys, blk_tab, sched, n_sys, slot, xcd_ranges, stream); break;
    default: launch_perform_batch_group<4>(cubic, grid, sys, blk_tab, sched, n_sys, slot, xcd_ranges, stream); break;
    }
    return hipGetLastError();
}

hipError_t launch_unwrap(const SweepBuffers &buf, const uint32_t box[3], int32_t *xyz, hipStream_t stream)
{
    if (buf.n_pad == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_unwrap, dim3((buf.n_pad + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, buf.ppos, buf.ids,
                       buf.image, buf.n_pad, buf.geom, box[0], box[1], box[2], xyz);
    return hipGetLastError();
}

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t stream)
{
    const size_t n = bytes / sizeof(uint4);
    const size_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks == 0 || blocks > 0x7fffffffu)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy, dim3(uint32_t(blocks)), dim3(kBlock), 0, stream, static_cast<const uint4 *>(src),
                       static_cast<uint4 *>(dst), n);
    return hipGetLastError();
}

}  // namespace pscbfm
