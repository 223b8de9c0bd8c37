// bfm_io.hpp -- LeMonADE's BFM file format (reader and writer).
//
// The format as the reference reads it (src/pscbfm/mainColorBFM.cpp:135-276;
// LeMonADE's FileImport/MoleculesRead behind UpdaterReadBfmFile, used by
// mainCUDASimulatorGPUScBFM_AB_Type.cpp:166): text lines
//   !key=value                  header (number_of_monomers, box_x/y/z,
//                               periodic_x/y/z, version)
//   !set_of_bondvectors         block of "dx dy dz:id" (id: the byte that
//                               stands for the vector in !mcs lines)
//   !bonds                      block of "a b" (1-based), bonds not implied
//                               by the chains of a conformation
//   !attributes                 block of "i-j:value" (1-based ranges)
//   !mcs=<age>                  a conformation: every line is "x y z" (the
//                               next monomer's absolute position), then one
//                               byte per bond-vector id, each adding the next
//                               monomer at the previous one plus that vector,
//                               bonded to it
// Blocks end at an empty line or at the next '!' line.  A file may hold many
// conformations; the last one is the state to continue from
// (READ_LAST_CONFIG_SAVE).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace pscbfm {

struct BfmState {
    uint64_t box[3] = {0, 0, 0};
    bool periodic[3] = {true, true, true};
    uint32_t n = 0;
    std::vector<int32_t> xyz;                               // 3 * n, unwrapped
    std::vector<std::pair<uint32_t, uint32_t>> bonds;       // undirected, 0-based
    std::vector<std::pair<int, int>> bondset_ids;           // (bond-table index, id) of the allowed vectors
    std::vector<int32_t> attributes;                        // n
    int64_t age = 0;
    uint32_t conformations = 0;                             // !mcs blocks in the file
};

// Reads the header, the bond-vector set, bonds, attributes and the last
// conformation.  Throws std::runtime_error with the line number on bad input.
BfmState read_bfm_file(const std::string &path);

// Writes a new file: header, bond-vector set, attributes, the bonds that the
// chain encoding does not carry, then the first conformation.
void write_bfm_file(const std::string &path, const BfmState &s);

// Appends a conformation "!mcs=<age>" for the state's bonds and bond set.
void append_bfm_conformation(const std::string &path, const BfmState &s);

// Bond vectors of the file's set: the standard set gets ids 17, 18, ... in
// the order of the bond-table index (any reader takes the ids from the file).
std::vector<std::pair<int, int>> bondset_ids_for(const uint8_t forbidden[512]);

}  // namespace pscbfm
