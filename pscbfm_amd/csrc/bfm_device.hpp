// bfm_device.hpp -- data layout and device helpers of the MI355X BFM sweep.
//
// Layout of one system in HBM (all arrays are indexed by the *sorted* monomer
// index: monomers grouped by colour, each colour segment padded to a multiple
// of kSegmentAlign so that every colour starts on an aligned boundary):
//   uint32_t ppos[nPad]        wrapped x | y << 10 | z << 20 (box edges <= 1024)
//   uint32_t ids[nPad]         original monomer id (keys the RNG) | number of
//                              bond partners << 28
//   int32_t  image[3][nPad]    periodic image of each axis: unwrapped = wrapped
//                              + L * image; changed only when a move crosses
//                              the box boundary
//   uint32_t nbr[8][nPad]      sorted index of the j-th bond partner (slot-major,
//                              so lane i reads nbr[j*nPad+i] coalesced)
//   uint32_t lattice[W]        occupancy bit lattice: the 8 sites of every
//                              monomer cube, compact Morton order
//   uint32_t claim[W]          sites entered by this substep's proposals
//   uint32_t coll[2][W]        sites entered by more than one proposal
//   uint2    surv[blocks*256]  compacted proposals: (sorted index | dir << 29,
//                              packed position), per check workgroup
//   uint32_t clist[2][...]     per-workgroup lists of coll[] words to clear
//
// Positions are wrapped on the device: a bond vector is only needed modulo 8
// (the bond-table index (d & 7), UpdaterGPUScBFM_AB_Type.cu:871-889) and every
// box edge is a multiple of 8, so wrapped coordinates give the same index.
//
// The reference keeps 1 lattice byte per monomer (its lower corner) and tests
// a 9-site front (UpdaterGPUScBFM_AB_Type.cu:420-594); this engine stores the
// full cube, 1 bit per site, so the same excluded-volume rule needs only the
// 4 sites of the face the cube enters.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pscbfm {

constexpr int kMaxConnectivity = 8;      // MAX_CONNECTIVITY, UpdaterGPUScBFM_AB_Type.h:85
// Partner columns the check kernel reads together with a slot's own words,
// before the partner counts arrive (a linear chain's monomer has at most 2).
constexpr int kEarlyPartners = 2;
constexpr uint32_t kSegmentAlign = 64;   // monomers; 64 x 16 B = 1 KiB
constexpr int kBlock = 256;              // threads per workgroup (4 wave64)
constexpr int kWave = 64;                // lanes per wavefront
constexpr uint32_t kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kSortedIndexBits = 29;
constexpr uint32_t kSortedIndexMask = (1u << kSortedIndexBits) - 1u;
// The id word of a sorted slot holds the original monomer id in its low
// kIdBits bits and the number of bond partners (0..kMaxConnectivity) above.
constexpr uint32_t kIdBits = 28;
constexpr uint32_t kIdMask = (1u << kIdBits) - 1u;
constexpr uint32_t kMaxBox = 1024;       // 10 bits per axis, 30-bit site index
constexpr uint32_t kCoordBits = 10;      // bits per axis of a packed position
constexpr uint32_t kCoordMask = (1u << kCoordBits) - 1u;

__host__ __device__ __forceinline__ uint32_t pack_xyz(uint32_t x, uint32_t y, uint32_t z)
{
    return x | (y << kCoordBits) | (z << (2 * kCoordBits));
}
__host__ __device__ __forceinline__ int4 unpack_xyz(uint32_t p)
{
    return make_int4(int32_t(p & kCoordMask), int32_t((p >> kCoordBits) & kCoordMask),
                     int32_t((p >> (2 * kCoordBits)) & kCoordMask), 0);
}

// Geometry of one periodic (or walled) power-of-two box.
struct Geometry {
    uint32_t mask[3];    // L_a - 1
    uint32_t dmask[3];   // index bits that belong to axis a (deposited L_a - 1)
    int32_t wall_hi[3];  // L_a - 2: largest lower-corner coordinate between walls
    uint32_t periodic;   // bit a set: axis a is periodic
    uint32_t cubic;      // 1: all edges equal, the index is the classic Z-curve
};

// Small per-system tables; each workgroup copies them to LDS.
struct Tables {
    uint32_t dep[3][64];  // deposit of coordinate bits: [a][v & 31] | [a][32 + (v >> 5)]
    uint32_t bond[16];    // 512-bit forbidden-bond mask, bit (dx&7)|(dy&7)<<3|(dz&7)<<6
};
constexpr int kTableWords = 3 * 64 + 16;

// ---- Philox4x32-10 (Salmon et al., SC'11) -----------------------------------
__host__ __device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c[3] ^ k1;
        c[1] = uint32_t(p1);
        c[3] = uint32_t(p0);
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Word 0 of Philox(key = seed, ctr = (c0, age_lo, age_hi, tag)).
__host__ __device__ __forceinline__ uint32_t philox_word0(uint64_t seed, uint32_t c0, uint64_t age, uint32_t tag)
{
    uint32_t c[4] = {c0, uint32_t(age), uint32_t(age >> 32), tag};
    philox4x32_10(c, uint32_t(seed), uint32_t(seed >> 32));
    return c[0];
}

constexpr uint32_t kTagMove = 0;
constexpr uint32_t kTagColour = 1;
constexpr uint32_t kTagSubstep = 3;

// Random number generators, numbered as the reference's Rng enum
// (UpdaterGPUScBFM_AB_Type.h:413-420).
constexpr uint32_t kRngIntHash = 1;   // the reference's default (.cu:1407)
constexpr uint32_t kRngPhilox = 3;

// Entry k of colour_order(seed, age, n) without the array: the value that ends
// at position k is found by following k back through the transpositions of
// the shuffle (the last one applied, i = 1, first).
__host__ __device__ __forceinline__ uint32_t colour_at(uint64_t seed, uint64_t age, uint32_t n, uint32_t k)
{
    uint32_t p = k;
    for (uint32_t i = 1; i < n; ++i) {
        const uint32_t j = philox_word0(seed, i, age, kTagColour) % (i + 1u);
        p = p == i ? j : (p == j ? i : p);
    }
    return p;
}

// 32-bit integer mix of the reference's Rngs::Hash (rngs/Hash.h:39-44):
// Robert Jenkins' six-step integer hash with its published constants.
__host__ __device__ __forceinline__ uint32_t int_hash(uint32_t a)
{
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}

// Seed of substep `substep` of MCS `age` for the integer-hash generator (the
// reference draws it from LeMonADE's r250 on the host, .cu:3066).
__host__ __device__ __forceinline__ uint32_t substep_seed(uint64_t seed, uint64_t age, uint32_t substep)
{
    return philox_word0(seed, substep, age, kTagSubstep);
}

// Colour order of one MCS: Fisher-Yates, slot i swaps with word % (i+1).
inline void colour_order(uint64_t seed, uint64_t age, uint32_t n, uint32_t *order)
{
    for (uint32_t i = 0; i < n; ++i)
        order[i] = i;
    for (uint32_t i = n; i-- > 1;) {
        const uint32_t j = philox_word0(seed, i, age, kTagColour) % (i + 1u);
        const uint32_t t = order[i];
        order[i] = order[j];
        order[j] = t;
    }
}

// ---- dilated (Morton) integer arithmetic ------------------------------------
// Spreads the low 10 bits of v to bits 0, 3, 6, ... (part1by2).
__host__ __device__ __forceinline__ uint32_t spread3(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Index bits of coordinate v (already masked) on axis a for a cubic box: the
// compact Morton order is then the classic Z-curve, x in bit 0, computed in
// registers.  Non-cubic boxes use per-axis tables in LDS (bfm_kernels.hip).
struct DepositCubic {
    __host__ __device__ __forceinline__ uint32_t operator()(int a, uint32_t v) const { return spread3(v) << a; }
};
__host__ __device__ __forceinline__ uint32_t dilated_inc(uint32_t d, uint32_t m) { return ((d | ~m) + 1u) & m; }
__host__ __device__ __forceinline__ uint32_t dilated_dec(uint32_t d, uint32_t m) { return (d - 1u) & m; }

__host__ __device__ __forceinline__ int32_t comp(const int4 &r, int a) { return a == 0 ? r.x : (a == 1 ? r.y : r.z); }
// v[a] for a per-lane axis without indexing memory (kernel arguments live in
// SGPRs; a dynamic index would turn them into scratch or vector loads).
__host__ __device__ __forceinline__ uint32_t sel3(const uint32_t v[3], int a) { return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]); }

// The lattice holds one bit per monomer, at its lower corner.  A cube with
// corner r moving in direction dir (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z, as the
// reference's DXTable_d/DYTable_d/DZTable_d, UpdaterGPUScBFM_AB_Type.cu:
// 1537-1541) can only overlap a cube whose corner lies in the 3x3 "front":
// the plane two sites ahead of r on the move axis, one site either side of r
// on the other two (checkFront, UpdaterGPUScBFM_AB_Type.cu:420-594).  The
// same 9 sites are where another moved corner makes two moved cubes overlap.
//
// Index bits 0..4 of a site are x0 y0 z0 x1 y1 for every box (edges >= 8), so
// a 32-bit word spans 4 x 4 x 2 sites and 3 consecutive coordinates of one
// axis fall into at most 2 words: the 9 sites lie in at most 4 words.  w[q]
// are those words, m[q] the bits of the front in each (m[q] == 0: not needed).
struct Front {
    uint32_t w[4], m[4];
    uint32_t new_site;   // corner after the move
    uint32_t old_site;   // corner before the move
};

// Product of two factors below 2^24 (one v_mul_u32_u24 on the device).
__host__ __device__ __forceinline__ uint32_t mul24(uint32_t a, uint32_t b)
{
    return (a & 0xFFFFFFu) * (b & 0xFFFFFFu);
}

template <class Dep>
__host__ __device__ __forceinline__ Front front_sites(const Geometry &g, const Dep &dep, const int4 &r, int dir)
{
    const int a = dir >> 1;
    const int b = a == 0 ? 1 : 0;
    const int c = a == 2 ? 1 : 2;
    const bool up = (dir & 1) != 0;
    const uint32_t dma = sel3(g.dmask, a), dmb = sel3(g.dmask, b), dmc = sel3(g.dmask, c);
    const uint32_t da = dep(a, uint32_t(up ? comp(r, a) + 2 : comp(r, a) - 2) & sel3(g.mask, a));
    const uint32_t b0 = dep(b, uint32_t(comp(r, b) - 1) & sel3(g.mask, b));
    const uint32_t c0 = dep(c, uint32_t(comp(r, c) - 1) & sel3(g.mask, c));
    const uint32_t b1 = dilated_inc(b0, dmb), b2 = dilated_inc(b1, dmb);
    const uint32_t c1 = dilated_inc(c0, dmc), c2 = dilated_inc(c1, dmc);
    // the moved corner is one step behind the front plane, the old one two
    const uint32_t a1 = up ? dilated_dec(da, dma) : dilated_inc(da, dma);
    const uint32_t a0 = up ? dilated_dec(a1, dma) : dilated_inc(a1, dma);
    Front f;
    f.new_site = a1 | b1 | c1;
    f.old_site = a0 | b1 | c1;
    // Site bit = 2^(la + lb + lc) with la, lb, lc the in-word parts of the
    // three axes (disjoint bit positions, so | is +).  Summing 2^lb over the
    // b-rows that share a word (sb_*) and 2^lc likewise (sc_*), the mask of a
    // word is sb * sc << la: the products of distinct powers never carry.
    // Every factor is < 2^19 (y1 is index bit 4), so 24-bit multiplies do.
    const uint32_t hb0 = b0 >> 5, hc0 = c0 >> 5;
    const bool b1_hi = (b1 >> 5) != hb0, b2_hi = (b2 >> 5) != hb0;
    const bool c1_hi = (c1 >> 5) != hc0, c2_hi = (c2 >> 5) != hc0;
    const uint32_t pb0 = 1u << (b0 & 31u), pb1 = 1u << (b1 & 31u), pb2 = 1u << (b2 & 31u);
    const uint32_t pc0 = 1u << (c0 & 31u), pc1 = 1u << (c1 & 31u), pc2 = 1u << (c2 & 31u);
    const uint32_t sb_lo = pb0 | (b1_hi ? 0u : pb1) | (b2_hi ? 0u : pb2);
    const uint32_t sb_hi = (b1_hi ? pb1 : 0u) | (b2_hi ? pb2 : 0u);
    const uint32_t sc_lo = pc0 | (c1_hi ? 0u : pc1) | (c2_hi ? 0u : pc2);
    const uint32_t sc_hi = (c1_hi ? pc1 : 0u) | (c2_hi ? pc2 : 0u);
    const uint32_t la = da & 31u, ha = da >> 5;
    const uint32_t hb2 = b2 >> 5, hc2 = c2 >> 5;
    f.w[0] = ha | hb0 | hc0;
    f.w[1] = ha | hb2 | hc0;
    f.w[2] = ha | hb0 | hc2;
    f.w[3] = ha | hb2 | hc2;
    f.m[0] = mul24(sb_lo, sc_lo) << la;
    f.m[1] = mul24(sb_hi, sc_lo) << la;
    f.m[2] = mul24(sb_lo, sc_hi) << la;
    f.m[3] = mul24(sb_hi, sc_hi) << la;
    return f;
}

// Site of corner (x, y, z) (wrapped coordinates).
template <class Dep>
__host__ __device__ __forceinline__ uint32_t corner_site(const Dep &dep, const int4 &r)
{
    return dep(0, uint32_t(r.x)) | dep(1, uint32_t(r.y)) | dep(2, uint32_t(r.z));
}

// Bond-table index of a bond vector: (dx & 7) | (dy & 7) << 3 | (dz & 7) << 6.
__host__ __device__ __forceinline__ uint32_t bond_index(int dx, int dy, int dz)
{
    return uint32_t(dx & 7) | (uint32_t(dy & 7) << 3) | (uint32_t(dz & 7) << 6);
}

}  // namespace pscbfm
