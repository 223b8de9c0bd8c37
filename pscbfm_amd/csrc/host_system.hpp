// host_system.hpp -- host half of UpdaterGPUScBFM_AB_Type::initialize:
// validation, graph colouring, compact Morton tables and the colour-sorted
// device layout.  Pure C++ (no HIP calls), so every rule here is testable on a
// machine without a GPU.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "bfm_device.hpp"

namespace pscbfm {

struct HostSystem {
    // ---- inputs (setters) ---------------------------------------------------
    uint64_t box[3] = {0, 0, 0};
    bool periodic[3] = {true, true, true};
    uint32_t n = 0;
    std::vector<int32_t> xyz;        // 3*n unwrapped coordinates
    std::vector<int32_t> attribute;  // n (stored only)
    std::vector<uint8_t> nnbr;       // n
    std::vector<uint32_t> nbr;       // n * kMaxConnectivity, directed links as given
    uint8_t forbidden[512];          // 1 = forbidden (copyBondSet)
    bool spatial_order = true;       // lay colours out in Morton order of position
    uint32_t rng = kRngIntHash;      // setRng
    int64_t steps_between_sortings = 5000;   // re-sort period a batch member keeps (setStepsBetweenSortings)

    // ---- derived by plan() --------------------------------------------------
    Geometry geom{};
    Tables tables{};
    uint64_t lattice_words = 0;
    std::vector<uint32_t> colour;    // n
    uint32_t n_colours = 0;
    std::vector<uint32_t> seg_begin; // per colour, sorted index of first monomer
    std::vector<uint32_t> seg_count; // per colour
    uint32_t n_pad = 0;              // padded sorted length
    uint32_t max_segment = 0;
    std::vector<uint32_t> to_sorted; // original id -> sorted index
    std::vector<uint32_t> ppos_sorted;  // n_pad packed wrapped positions
    std::vector<uint32_t> ids_sorted;   // n_pad: original id | partner count << kIdBits (padding: ~0)
    std::vector<int32_t> image_sorted;  // 3 * n_pad periodic images, axis-major
    std::vector<uint32_t> nbr_sorted;  // (largest partner count) * n_pad, slot-major
    std::vector<uint32_t> lattice;     // lattice_words occupancy bits
    double mean_bonds = 0.0;           // mean number of listed partners per monomer

    HostSystem();
    void reset();
};

// Number of allowed entries of a 512-entry bond table.
int count_allowed_bonds(const uint8_t forbidden[512]);
// Standard 108-vector bond set (permutations and sign flips of (2,0,0),
// (2,1,0), (2,1,1), (2,2,1), (3,0,0), (3,1,0)).
void standard_bond_set(uint8_t forbidden[512]);
inline int bond_table_index(int dx, int dy, int dz) { return (dx & 7) | ((dy & 7) << 3) | ((dz & 7) << 6); }

// Compact Morton geometry of a power-of-two box (edges 8..1024): bits of x,
// y, z interleave from bit 0 upward in the order x, y, z, skipping an axis
// whose bits are used up.  A cubic box gives the classic Z-curve with x in
// bit 0 (UpdaterGPUScBFM_AB_Type.cu:289-316).
void make_geometry(const uint64_t box[3], const bool periodic[3], Geometry &g, Tables &t,
                   uint64_t &lattice_words);
uint32_t host_site(const Geometry &g, const Tables &t, int32_t x, int32_t y, int32_t z);

// Colouring graph of the bond lists: every monomer's listed partners in the
// order they were listed (setConnectivity), then the monomers that list it
// without being listed by it (in increasing id), so that a one-way link still
// keeps its two ends apart.  For symmetric lists, as the LeMonADE adapter
// sets them (GPUScBFM_AB_Type.h:159-161), these are the lists themselves.
void colouring_adjacency(uint32_t n, const std::vector<uint8_t> &nnbr, const std::vector<uint32_t> &nbr,
                         std::vector<uint32_t> &off, std::vector<uint32_t> &adj);

// The reference's graph colouring (graphColoring.tpp:103-439, called with
// bUniformColors = true by initializeSpeciesSorting,
// UpdaterGPUScBFM_AB_Type.cu:1561-1566): a walk along each connected polymer
// giving every monomer the lowest colour its partners do not have, then (with
// `uniform`) a rebalance toward ceil(n / colours) monomers per colour.
// A bipartite bond graph gets 2 colours.  Returns the colour count; throws
// std::runtime_error above 255 colours.
uint32_t graph_colouring(uint32_t n, const std::vector<uint8_t> &nnbr, const std::vector<uint32_t> &nbr,
                         bool uniform, std::vector<uint32_t> &colour);

// Unwrapped coordinates of every monomer (by original id) from the packed
// wrapped positions and periodic images in sorted order.
void unpack_positions(HostSystem &s);

// Frees the sorted per-monomer arrays and the lattice that plan() built for
// the upload (keeps the colours, the segments and the geometry).
void release_upload_staging(HostSystem &s);

// Runs every check of initialize() and fills the derived fields.
// Throws std::invalid_argument / std::runtime_error like the reference.
void plan(HostSystem &s);

}  // namespace pscbfm
