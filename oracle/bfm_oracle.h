/*
 * bfm_oracle.h -- CPU restatement of the parallel bond-fluctuation-model
 * colour sweep.  TEST INFRASTRUCTURE ONLY: this library is the checker that
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py compare
 * the HIP product against.  Nothing in pscbfm_amd/ links, loads or calls it.
 *
 * Everything here is written for clarity, not speed: a byte lattice with
 * linear (x fastest) indexing, plain loops, one monomer at a time.  It shares
 * no code with the product (which uses a bit-packed Morton lattice and
 * atomics), so agreement between the two is evidence, not tautology.
 *
 * Reference anchors (paths relative to the pscbfm repository):
 *   check phase       src/pscbfm/UpdaterGPUScBFM_AB_Type.cu:930-1028
 *   front test (v6)   src/pscbfm/UpdaterGPUScBFM_AB_Type.cu:420-594
 *   perform + apply   src/pscbfm/UpdaterGPUScBFM_AB_Type.cu:1148-1190
 *   bond index        src/pscbfm/UpdaterGPUScBFM_AB_Type.cu:871-889
 *   direction table   src/pscbfm/UpdaterGPUScBFM_AB_Type.cu:1537-1541
 *   invariants        src/pscbfm/UpdaterGPUScBFM_AB_Type.cu:2789-2900
 *   colouring         src/pscbfm/graphColoring.tpp:131-294, :323-418
 *   integer-hash RNG  src/pscbfm/rngs/Hash.h:39-44,64-68 (default, .cu:1407)
 *
 * Parity status of each function is stated next to its declaration.
 */
#ifndef BFM_ORACLE_H
#define BFM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_system {
    int32_t  box[3];           /* power-of-two edge lengths, each >= 8          */
    int32_t  periodic[3];      /* 1 = periodic axis, 0 = hard walls             */
    uint32_t n;                /* number of monomers                            */
    const uint32_t *nbr_off;   /* CSR offsets, n+1 entries                      */
    const uint32_t *nbr_idx;   /* CSR neighbour ids (both directions present)   */
    const uint32_t *color;     /* colour of every monomer, < n_colors           */
    uint32_t n_colors;
    const uint8_t *forbidden;  /* 512 entries, index (dx&7)|(dy&7)<<3|(dz&7)<<6 */
    uint64_t seed;             /* Philox key                                    */
    uint32_t rng;              /* ORC_RNG_INT_HASH or ORC_RNG_PHILOX            */
} orc_system;

/* Values of the reference's Rng enum (UpdaterGPUScBFM_AB_Type.h:413-420). */
enum { ORC_RNG_INT_HASH = 1, ORC_RNG_PHILOX = 3 };

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
 * Pinned by the published Random123 known-answer vectors (tests/test_oracle.py)
 * and an independent Python-integer implementation. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Random word that decides the direction of monomer `id` in MCS `age`:
 * word 0 of Philox(key = seed, ctr = (id, age_lo, age_hi, 0)).
 * Direction = word % 6 as in UpdaterGPUScBFM_AB_Type.cu:966. */
uint32_t orc_move_word(uint64_t seed, uint64_t age, uint32_t id);

/* The 32-bit integer mix of the reference's default RNG (Rngs::Hash,
 * rngs/Hash.h:39-44; Robert Jenkins' six-step integer hash).  Pinned by
 * vectors of the reference header itself, built from its source into
 * oracle/_ref (tests/golden/int_hash_vectors.json), and by being a bijection. */
uint32_t orc_int_hash(uint32_t a);

/* Seed of substep `substep` (0-based slot in the MCS) of MCS `age`: word 0 of
 * Philox(seed, (substep, age_lo, age_hi, 3)).  The reference draws this
 * 32-bit value from LeMonADE's r250 on the host (.cu:3066). */
uint32_t orc_substep_seed(uint64_t seed, uint64_t age, uint32_t substep);

/* Direction word of monomer `id` in substep `substep` of MCS `age`:
 * ORC_RNG_INT_HASH: H(H(id) ^ orc_substep_seed(...)), the reference's
 *   Hash::rng32 with subsequence = monomer (Hash.h:64-68, .cu:961-964; keyed
 *   here on the original id instead of the sorted index);
 * ORC_RNG_PHILOX: orc_move_word(seed, age, id). */
uint32_t orc_direction_word(const orc_system *sys, uint64_t age, uint32_t substep, uint32_t id);

/* Colour order of MCS `age`: Fisher-Yates over 0..n-1, draw for slot i is
 * word 0 of Philox(seed, (i, age_lo, age_hi, 1)) % (i+1).  Every colour is
 * swept exactly once per MCS (the reference draws colours with replacement,
 * UpdaterGPUScBFM_AB_Type.cu:3065; the per-MCS permutation is this engine's
 * documented rule).  Pinned by tests: it is a permutation, and chi^2 over
 * many ages is uniform. */
void orc_color_order(uint64_t seed, uint64_t age, uint32_t n_colors, uint32_t *order);

/* The standard 108-vector BFM bond set in closed form: every permutation and
 * sign flip of (2,0,0),(2,1,0),(2,1,1),(2,2,1),(3,0,0),(3,1,0).
 * Pinned by the count check of UpdaterGPUScBFM_AB_Type.cu:1526 and by
 * brute-force enumeration in tests. */
void orc_standard_bondset(uint8_t forbidden[512]);

/* Invariants of UpdaterGPUScBFM_AB_Type.cu:2789-2900 on an 8-site cube
 * occupancy lattice: no two cubes share a site, every bond |d_i|<=3 and
 * allowed (difference taken as the minimum image on periodic axes),
 * non-periodic axes keep the cube inside [0, L-1].
 * Returns 0 if valid; otherwise a positive code and a message. */
int orc_check_system(const orc_system *sys, const int32_t *pos, char *msg, int msglen);

/* The reference's graph colouring, graphColoring.tpp:103-439, as
 * initializeSpeciesSorting calls it (UpdaterGPUScBFM_AB_Type.cu:1561-1566,
 * bUniformColors = true): each connected polymer is walked from its lowest id
 * (go to the first uncoloured partner, keep a stack of monomers with two or
 * more, resume from the stack at a dead end), every monomer taking the
 * lowest colour none of its partners has; with `uniform` a second pass in id
 * order moves monomers of over-full colours (more than ceil(n/k)) to the
 * rarest colours their partners do not have.
 * (nbr_off, nbr_idx): each monomer's links in the order they were set
 * (directed, as setConnectivity receives them); a link whose reverse is not
 * listed also counts for its target, appended after that monomer's own list
 * in increasing id.  Returns 0, or -1 when 255 colours do not suffice.
 * Pinned by colourings that the reference's own graphColoring.tpp, compiled
 * into oracle/_ref/coloring_ref, produced for the graphs in
 * tests/golden/coloring_vectors.json; by 2 colours on bipartite graphs; by
 * the chromatic number of complete graphs and odd cycles. */
int orc_graph_coloring(uint32_t n, const uint32_t *nbr_off, const uint32_t *nbr_idx, int uniform,
                       uint32_t *color, uint32_t *n_colors);

/* Serial replay of the colour sweep: for every MCS, every colour in
 * orc_color_order, (1) every monomer of the colour proposes a move, checked
 * against walls, bonds and the current excluded volume; (2) a survivor is
 * accepted iff no other survivor's new cube overlaps its new cube; (3) all
 * accepted moves apply at once.
 *   form 0: face-claim counting on the 8-site lattice (4 sites per move).
 *   form 1: the reference's own formulation: one corner site per monomer, the
 *           9-site front of UpdaterGPUScBFM_AB_Type.cu:546-594 for phase 1 and
 *           the tentative corner lattice of :1148-1190 for phase 2.
 *   form 2: brute force: phase 2 by pairwise cube overlap of all survivors.
 * All three must give identical trajectories (tests/test_oracle.py).
 * pos is int32 xyz per monomer (unwrapped), updated in place.
 * accepted / proposals (may be NULL) count accepted moves and phase-1 survivors.
 * Returns 0, or a negative error code on bad input / allocation failure. */
int orc_sweep(const orc_system *sys, int32_t *pos, uint64_t age0, uint32_t n_mcs,
              int form, uint64_t *accepted, uint64_t *proposals);

/* The same sweep (form 1: corner lattice, 9-site fronts) with every phase
 * of a substep run as an OpenMP loop over the colour's monomers on `threads`
 * threads (0: the OpenMP default).  Each loop only reads what earlier loops
 * wrote; the outcome depends only on the set of proposals, so the result is
 * that of orc_sweep for any thread count (pinned by comparing the two,
 * tests/test_oracle.py).  The all-cores CPU timing of the sweep. */
int orc_sweep_parallel(const orc_system *sys, int32_t *pos, uint64_t age0, uint32_t n_mcs,
                       int threads, uint64_t *accepted, uint64_t *proposals);

/* LeMonADE-style random-sequential updater: per MCS, n attempts; attempt k of
 * MCS `age` draws Philox(seed, (k, age_lo, age_hi, 2)): monomer = w0 % n,
 * direction = w1 % 6; a legal move is applied at once.  This is the CPU
 * baseline and the statistical reference.  Pinned by closed-form
 * trajectories of free and walled monomers from the Philox words
 * (tests/test_oracle.py) and by its stationary laws, which detailed balance
 * fixes: uniform over the 108 bond vectors for a free dimer and over the legal
 * relative positions for two free monomers (tests/test_statistics.py).
 * Trajectory parity with LeMonADE's own UpdaterSimpleSimulator is unpinned
 * (LeMonADE is not available, and it draws from r250); its move rules are
 * those of the sweep above. */
int orc_random_sequential(const orc_system *sys, int32_t *pos, uint64_t age0,
                          uint32_t n_mcs, uint64_t *accepted);

#ifdef __cplusplus
}
#endif
#endif
