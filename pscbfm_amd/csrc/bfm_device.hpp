// bfm_device.hpp -- data layout and device helpers of the MI355X BFM sweep.
//
// Layout of one system in HBM (all arrays are indexed by the *sorted* monomer
// index: monomers grouped by colour, each colour segment padded to a multiple
// of kSegmentAlign so that every colour starts on an aligned boundary):
//   uint32_t ppos[nPad]        lattice site of the monomer's lower corner
//                              (compact Morton index, the bit of the corner
//                              in the occupancy lattice)
//   uint32_t ids[nPad]         original monomer id (keys the RNG) | number of
//                              bond partners << 28
//   int32_t  image[3][nPad]    periodic image of each axis: unwrapped = wrapped
//                              + L * image; changed only when a move crosses
//                              the box boundary
//   uint32_t nbr[8][nPad]      sorted index of the j-th bond partner (slot-major,
//                              so lane i reads nbr[j*nPad+i] coalesced)
//   uint32_t lattice[W]        occupancy bit lattice: one bit per monomer, at
//                              its lower corner, compact Morton order
//   uint32_t claim[2][W]       corners entered by a substep's proposals,
//                              double-buffered by substep parity
//   uint2    surv[waves*64]    compacted proposals: (sorted index | dir << 29,
//                              corner site), per check wavefront
//   uint32_t scount[waves]     proposals per check wavefront
//
// A position is kept as its site index and moved with dilated (Morton)
// arithmetic on one axis's bits: a bond vector is only needed modulo 8 (the
// bond-table index (d & 7), UpdaterGPUScBFM_AB_Type.cu:871-889), and the low
// 3 bits of every axis are the low 9 bits of the site in every geometry.
//
// The lattice keeps the reference's 1 site per monomer (its lower corner) and
// tests the reference's 9-site front (UpdaterGPUScBFM_AB_Type.cu:420-594), in
// at most 4 words of 32 sites.
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

// Geometry of one periodic (or walled) power-of-two box.
struct Geometry {
    uint32_t mask[3];    // L_a - 1
    uint32_t dmask[3];   // index bits that belong to axis a (deposited L_a - 1)
    int32_t wall_hi[3];  // L_a - 2: largest lower-corner coordinate between walls
    uint32_t periodic;   // bit a set: axis a is periodic
    uint32_t cubic;      // 1: all edges equal, the index is the classic Z-curve
};

// Small per-system tables.
struct Tables {
    uint32_t dep[3][64];  // deposit of coordinate bits: [a][v & 31] | [a][32 + (v >> 5)] (host side)
    // 512-bit forbidden-bond mask in the site's bit order: bit bond_key() of
    // the vector (dx & 7, dy & 7, dz & 7), i.e. those three 3-bit numbers
    // interleaved x, y, z from bit 0 (see bond_key)
    uint32_t bond[16];
};

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

// Arithmetic on one axis of a site index.  x holds only bits of the axis's
// mask m (dmask), k is a dilated constant of the axis; the result wraps modulo
// the box edge.  Adding sets every other bit so that carries run through them.
__host__ __device__ __forceinline__ uint32_t dil_add(uint32_t x, uint32_t k, uint32_t m) { return ((x | ~m) + k) & m; }
__host__ __device__ __forceinline__ uint32_t dil_sub(uint32_t x, uint32_t k, uint32_t m) { return (x - k) & m; }

// Coordinate of axis `dmask` of a site: the axis's bits gathered downward.
// (For the copy-back and the analyzers; the sweep never needs coordinates.)
__host__ __device__ inline uint32_t site_coord(uint32_t site, uint32_t dmask)
{
    uint32_t v = 0, bit = 1;
    for (uint32_t m = dmask; m != 0u; m &= m - 1u, bit <<= 1)
        if (site & m & (0u - m))
            v |= bit;
    return v;
}

// v[a] for a per-lane axis without indexing memory (kernel arguments live in
// SGPRs; a dynamic index would turn them into scratch or vector loads).
__host__ __device__ __forceinline__ uint32_t sel3(const uint32_t v[3], int a) { return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]); }

// ---- the 9-site front ---------------------------------------------------------
// The lattice holds one bit per monomer, at its lower corner.  A cube with
// corner r moving in direction dir (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z, as the
// reference's DXTable_d/DYTable_d/DZTable_d, UpdaterGPUScBFM_AB_Type.cu:
// 1537-1541) can only overlap a cube whose corner lies in the 3x3 "front":
// the plane two sites ahead of r on the move axis a, one site either side of
// r on the other two axes b and c (checkFront, UpdaterGPUScBFM_AB_Type.cu:
// 420-594).  The same 9 sites are where another moved corner makes two moved
// cubes overlap.
//
// Index bits 0..4 of a site are x0 y0 z0 x1 y1 for every box (edges >= 8,
// the interleave takes one bit of each axis per round while the axis has
// bits), so a 32-bit word spans 4 x 4 x 2 sites.  The three b values of the
// front fall into at most 2 words, as do the three c values: the 9 sites lie
// in the words of (plane, b-1 or b+1, c-1 or c+1), w[0..3] below, and m[q]
// are the front's bits in w[q] (0 where w[q] repeats an earlier word).
struct Front {
    uint32_t w[4], m[4];
    uint32_t new_site;   // corner after the move
    uint32_t ma;         // dmask of the move axis
};

// Product of two factors below 2^24 (one v_mul_u32_u24 on the device).
__host__ __device__ __forceinline__ uint32_t mul24(uint32_t a, uint32_t b)
{
    return (a & 0xFFFFFFu) * (b & 0xFFFFFFu);
}

// Dilated masks of the move axis a and the two front axes b, c of direction dir.
template <bool kCubic>
__host__ __device__ __forceinline__ void front_axes(const Geometry &g, int dir, uint32_t &ma, uint32_t &mb,
                                                    uint32_t &mc)
{
    const int a = dir >> 1;
    if (kCubic) {
        // every axis's bits are the x bits shifted by the axis number
        ma = g.dmask[0] << a;
        mb = g.dmask[0] << (a == 0 ? 1 : 0);
        mc = g.dmask[0] << (a == 2 ? 1 : 2);
    } else {
        ma = sel3(g.dmask, a);
        mb = a == 0 ? g.dmask[1] : g.dmask[0];
        mc = a == 2 ? g.dmask[1] : g.dmask[2];
    }
}

template <bool kCubic>
__host__ __device__ __forceinline__ Front front_of(const Geometry &g, uint32_t site, int dir)
{
    uint32_t ma, mb, mc;
    front_axes<kCubic>(g, dir, ma, mb, mc);
    const int a = dir >> 1;
    const bool up = (dir & 1) != 0;
    // dilated 1 and 2 of an axis are its first two bits: bits a and a + 3
    // for the move axis, the lowest bit of the mask for b and c
    const uint32_t one_a = 1u << a, two_a = 1u << (a + 3);
    const uint32_t one_b = mb & (0u - mb), one_c = mc & (0u - mc);
    const uint32_t sa = site & ma, sb = site & mb, sc = site & mc;
    // a - 2 is a + (L - 2), whose dilated form is the mask without its lowest
    // bit; a - 1 is a + (L - 1), the mask itself
    const uint32_t plane = dil_add(sa, up ? two_a : (ma ^ one_a), ma);
    const uint32_t a1 = dil_add(sa, up ? one_a : ma, ma);
    const uint32_t b0 = dil_sub(sb, one_b, mb), b2 = dil_add(sb, one_b, mb);
    const uint32_t c0 = dil_sub(sc, one_c, mc), c2 = dil_add(sc, one_c, mc);
    Front f;
    f.w[0] = (plane | b0 | c0) >> 5;
    f.w[1] = (plane | b2 | c0) >> 5;
    f.w[2] = (plane | b0 | c2) >> 5;
    f.w[3] = (plane | b2 | c2) >> 5;
    // The bit of a site in its word is the sum of its three axes' in-word
    // index bits (disjoint positions), so it is 2^la * 2^lb * 2^lc.  Summing
    // 2^lb over the b values that share a word (sb_*) and 2^lc likewise
    // (sc_*), the mask of a word is sb * sc << la: the products of distinct
    // powers never carry.  Every factor is < 2^19 (y1 is index bit 4), so
    // 24-bit multiplies do.  A value has left the word of b - 1 when its
    // index bits above the word differ.
    const bool b1_hi = ((sb ^ b0) >> 5) != 0u, b2_hi = ((b2 ^ b0) >> 5) != 0u;
    const bool c1_hi = ((sc ^ c0) >> 5) != 0u, c2_hi = ((c2 ^ c0) >> 5) != 0u;
    const uint32_t pb0 = 1u << (b0 & 31u), pb1 = 1u << (sb & 31u), pb2 = 1u << (b2 & 31u);
    const uint32_t pc0 = 1u << (c0 & 31u), pc1 = 1u << (sc & 31u), pc2 = 1u << (c2 & 31u);
    const uint32_t sb_lo = pb0 | (b1_hi ? 0u : pb1) | (b2_hi ? 0u : pb2);
    const uint32_t sb_hi = (b1_hi ? pb1 : 0u) | (b2_hi ? pb2 : 0u);
    const uint32_t sc_lo = pc0 | (c1_hi ? 0u : pc1) | (c2_hi ? 0u : pc2);
    const uint32_t sc_hi = (c1_hi ? pc1 : 0u) | (c2_hi ? pc2 : 0u);
    const uint32_t la = plane & 31u;
    f.m[0] = mul24(sb_lo, sc_lo) << la;
    f.m[1] = mul24(sb_hi, sc_lo) << la;
    f.m[2] = mul24(sb_lo, sc_hi) << la;
    f.m[3] = mul24(sb_hi, sc_hi) << la;
    f.new_site = a1 | sb | sc;
    f.ma = ma;
    return f;
}

// Key of a bond vector from the partner's site and the moved corner's site:
// the three per-axis differences modulo 8, taken on the low 9 site bits
// (x0 y0 z0 x1 y1 z1 x2 y2 z2 in every geometry), interleaved as there.
// r1x/r1y/r1z are the moved corner's site masked with kLow9[axis].
constexpr uint32_t kLow9[3] = {0x049u, 0x092u, 0x124u};
__host__ __device__ __forceinline__ uint32_t bond_key(uint32_t partner_site, uint32_t r1x, uint32_t r1y, uint32_t r1z)
{
    return dil_sub(partner_site & kLow9[0], r1x, kLow9[0]) | dil_sub(partner_site & kLow9[1], r1y, kLow9[1]) |
           dil_sub(partner_site & kLow9[2], r1z, kLow9[2]);
}
// The same key from a bond vector.
__host__ __device__ __forceinline__ uint32_t bond_key_of(int dx, int dy, int dz)
{
    return spread3(uint32_t(dx) & 7u) | (spread3(uint32_t(dy) & 7u) << 1) | (spread3(uint32_t(dz) & 7u) << 2);
}

// Bond-table index of a bond vector: (dx & 7) | (dy & 7) << 3 | (dz & 7) << 6.
__host__ __device__ __forceinline__ uint32_t bond_index(int dx, int dy, int dz)
{
    return uint32_t(dx & 7) | (uint32_t(dy & 7) << 3) | (uint32_t(dz & 7) << 6);
}

}  // namespace pscbfm
