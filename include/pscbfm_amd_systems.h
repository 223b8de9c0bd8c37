/*
 * pscbfm_amd_systems.h -- seeded synthetic BFM systems.
 *
 * No counterpart in the reference: it reads its systems from LeMonADE .bfm
 * files (src/pscbfm/mainCUDASimulatorGPUScBFM_AB_Type.cpp:157-171), which
 * this repository does not have.  These generators produce valid, seeded
 * initial states of the same kinds (linear melts, end-linked networks).
 */
#ifndef PSCBFM_AMD_SYSTEMS_H
#define PSCBFM_AMD_SYSTEMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Linear melt of n_chains chains of chain_len monomers in a periodic cubic box
 * of edge box (power of two, 8..1024).  The box is cut into blocks of
 * next_pow2(2*chain_len) cells of the even sublattice (cells of edge 2); each
 * chain is a seeded self-avoiding walk confined to its own randomly chosen
 * block, with steps (+-2,0,0) and permutations, i.e. 6 of the 108 bond
 * vectors, so no two cubes can overlap.  Requires n_chains <= number of
 * blocks (volume fraction 8*n/box^3 <= 0.5 for power-of-two chain_len).
 * Monomer ids run chain by chain, chains in increasing block order.
 * xyz: 3*n_chains*chain_len int32.  bonds: 2*n_chains*(chain_len-1) uint32,
 * undirected pairs (i, i+1). */
int bfm_gen_linear_melt(uint32_t box, uint32_t n_chains, uint32_t chain_len, uint64_t seed,
                        int32_t *xyz, uint32_t *bonds);

/* Size of the network of bfm_gen_diamond_network for these arguments. */
int bfm_gen_diamond_network_size(uint32_t box, uint32_t cells_x, uint32_t cells_y, uint32_t cells_z,
                                 uint32_t strand_len, uint32_t *n_monomers, uint32_t *n_bonds);

/* End-linked network: 4-functional crosslinkers on a diamond topology of
 * cells_x*cells_y*cells_z conventional cells (8 crosslinkers each) in a
 * periodic cubic box, joined by strands of strand_len monomers.  Each cell
 * spans (box/2)/cells_a even-sublattice sites on axis a, which must be a
 * multiple of 4; neighbouring crosslinkers are q = ((box/2)/(4*cells_a))_a
 * sites apart, and strand_len+1 must be >= q_x+q_y+q_z and of the same parity.
 * Every bond is a (+-2,0,0)-class vector on the even sublattice; strands are
 * routed as seeded exact-length self-avoiding paths (depth-first, with
 * rip-up and reroute).  Ids: crosslinkers first (8 per cell, sublattice A
 * then B), then the strands, each from its A end.  Crosslinker coordinates lie
 * in [0, box); strand coordinates are unwrapped from their A end, so the bond
 * that closes a strand may cross the periodic boundary (bond vectors are
 * defined modulo the box, as for any percolating periodic network).
 * xyz: 3*n int32, bonds: 2*n_bonds uint32 undirected pairs. */
int bfm_gen_diamond_network(uint32_t box, uint32_t cells_x, uint32_t cells_y, uint32_t cells_z,
                            uint32_t strand_len, uint64_t seed, int32_t *xyz, uint32_t *bonds);

#ifdef __cplusplus
}
#endif
#endif
