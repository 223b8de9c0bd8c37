/*
 * pscbfm_amd.h -- C ABI of the MI355X bond-fluctuation-model updater.
 *
 * Every entry point mirrors one public member of the reference updater
 * UpdaterGPUScBFM_AB_Type (src/pscbfm/UpdaterGPUScBFM_AB_Type.h:453-514),
 * which the LeMonADE adapter GPUScBFM_AB_Type drives
 * (src/pscbfm/GPUScBFM_AB_Type.h:120-251).  Plain pointers and sizes only.
 *
 * Errors: no C++ exception and no exit() crosses this boundary.  Every call
 * returns BFM_OK (0) or a negative status; bfm_last_error() then returns a
 * thread-local message.  BFM_E_INVALID_ARGUMENT stands for the reference's
 * std::invalid_argument, BFM_E_RUNTIME for its std::runtime_error, BFM_E_HIP
 * for a failed HIP call (the reference would exit(EXIT_FAILURE),
 * src/pscbfm/cudacommon.hpp:36-51).
 *
 * Threading: every handle is independent (no module-global tables, unlike the
 * reference's __constant__ box/bond symbols, UpdaterGPUScBFM_AB_Type.cu:84-133).
 * A handle may be used by one host thread at a time.
 */
#ifndef PSCBFM_AMD_H
#define PSCBFM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BFM_OK = 0,
    BFM_E_INVALID_ARGUMENT = -1,
    BFM_E_RUNTIME = -2,
    BFM_E_HIP = -3,
    BFM_E_NO_DEVICE = -4,
    BFM_E_OUT_OF_MEMORY = -5,
    BFM_E_STATE = -6
};

/* Random number generators, numbered as the reference's Rng enum
 * (UpdaterGPUScBFM_AB_Type.h:413-420):
 *   BFM_RNG_INT_HASH (default, as .cu:1407): rn = H(H(id) ^ s), H the
 *     integer hash of rngs/Hash.h:39-44, s a 32-bit seed per substep drawn
 *     from Philox4x32-10(seed, (substep, MCS lo, MCS hi, 3));
 *   BFM_RNG_PHILOX: rn = word 0 of Philox4x32-10(seed, (id, MCS lo, MCS hi, 0)).
 * Both are keyed on the original monomer id and the MCS, never on the
 * layout, so trajectories do not depend on sorting or batching. */
enum { BFM_RNG_INT_HASH = 1, BFM_RNG_PHILOX = 3 };

typedef struct bfm_updater bfm_updater;

/* Thread-local description of the last failed call ("" if none). */
const char *bfm_last_error(void);

/* Constructor / destructor (UpdaterGPUScBFM_AB_Type.cu:1372-1490). */
int bfm_create(bfm_updater **out);
int bfm_destroy(bfm_updater *h);

/* setGpu, UpdaterGPUScBFM_AB_Type.cu:1498-1514. */
int bfm_set_gpu(bfm_updater *h, int gpu);

/* copyBondSet, UpdaterGPUScBFM_AB_Type.cu:2540-2545; forbidden != 0 marks the
 * vector as not allowed.  dx,dy,dz in [-4,3]. */
int bfm_copy_bond_set(bfm_updater *h, int dx, int dy, int dz, int forbidden);

/* setLatticeSize, UpdaterGPUScBFM_AB_Type.cu:2639-2684: powers of two. */
int bfm_set_lattice_size(bfm_updater *h, uint64_t box_x, uint64_t box_y, uint64_t box_z);

/* setNrOfAllMonomers, UpdaterGPUScBFM_AB_Type.cu:2548-2563 (may be called once). */
int bfm_set_nr_of_all_monomers(bfm_updater *h, uint32_t n);

/* setAttribute, UpdaterGPUScBFM_AB_Type.cu:2579 (stored, not used by the sweep). */
int bfm_set_attribute(bfm_updater *h, uint32_t i, int32_t attribute);

/* setMonomerCoordinates, UpdaterGPUScBFM_AB_Type.cu:2582-2610. */
int bfm_set_monomer_coordinates(bfm_updater *h, uint32_t i, int32_t x, int32_t y, int32_t z);

/* setConnectivity, UpdaterGPUScBFM_AB_Type.cu:2617-2635: directed link i->j,
 * at most 8 per monomer.  The LeMonADE adapter adds both directions. */
int bfm_set_connectivity(bfm_updater *h, uint32_t i, uint32_t j);

/* setPeriodicity, UpdaterGPUScBFM_AB_Type.cu:2565-2576. */
int bfm_set_periodicity(bfm_updater *h, int periodic_x, int periodic_y, int periodic_z);

/* setAge / getAge, UpdaterGPUScBFM_AB_Type.h:504-505.  The age is the MCS
 * counter that keys the RNG, so resuming at an age continues the trajectory. */
int bfm_set_age(bfm_updater *h, int64_t age);
int bfm_get_age(const bfm_updater *h, int64_t *age);

/* setRng, UpdaterGPUScBFM_AB_Type.h:506: BFM_RNG_INT_HASH or BFM_RNG_PHILOX
 * (Saru, PCG and XORWOW are not provided). */
int bfm_set_rng(bfm_updater *h, int rng);
/* getRng, UpdaterGPUScBFM_AB_Type.h:509. */
int bfm_get_rng(const bfm_updater *h, int *rng);

/* Seed of the RNGs.  The reference seeds LeMonADE's r250 from a seed file
 * (src/pscbfm/mainCUDASimulatorGPUScBFM_AB_Type.cpp:123-134); here the 64-bit
 * seed is the Philox key (which also draws the integer hash's substep seeds). */
int bfm_set_seed(bfm_updater *h, uint64_t seed);

/* setStepsBetweenSortings, UpdaterGPUScBFM_AB_Type.h:510-513 (default 5000):
 * every `steps`-th MCS the monomers of each colour are re-sorted on the device
 * by the lattice site of their position (doSpatialSorting, .cu:1836-1935).
 * Re-sorting never changes a trajectory here because the generator is keyed
 * on the original monomer id. */
int bfm_set_steps_between_sortings(bfm_updater *h, int64_t steps);
/* getStepsBetweenSortings, UpdaterGPUScBFM_AB_Type.h:514. */
int bfm_get_steps_between_sortings(const bfm_updater *h, int64_t *steps);

/* Layout of each colour segment (no counterpart in the reference's API; its
 * doSpatialSorting, UpdaterGPUScBFM_AB_Type.cu:1836-1935, is always on).  On
 * (the default), monomers of a colour are stored in the lattice's Morton
 * order of their positions, so a wavefront sweeps a compact patch of space;
 * off, in order of id.  The trajectory is the same either way.  Before
 * initialize / bfm_validate only. */
int bfm_set_spatial_sorting(bfm_updater *h, int on);

/* LeMonADE BFM files (UpdaterReadBfmFile / AnalyzerWriteBfmFile around the
 * reference's simulator, mainCUDASimulatorGPUScBFM_AB_Type.cpp:166-171; the
 * format as mainColorBFM.cpp:135-276 parses it).  read: configures an updater
 * that is not initialized yet from the file's header, bond-vector set,
 * attributes, bonds and last conformation (its !mcs value becomes the age).
 * write: append = 0 writes a new file (header, bond set, attributes, the bonds
 * the chain encoding does not carry, the current conformation); append != 0
 * appends the current conformation as "!mcs=<age>". */
int bfm_read_bfm_file(bfm_updater *h, const char *path);
int bfm_write_bfm_file(bfm_updater *h, const char *path, int append);

/* The host half of initialize (no device work): the checks of checkSystem
 * (UpdaterGPUScBFM_AB_Type.cu:2789-2900) and the bond-set count (:1526), then
 * the colouring of initializeSpeciesSorting (:1561-1686).  Lets a caller
 * validate a system and read its colours on a machine without a GPU. */
int bfm_validate(bfm_updater *h);

/* initialize, UpdaterGPUScBFM_AB_Type.cu:2393-2537: validates the system
 * (bond set of 108 vectors, overlap-free cubes, bonds in the set), colours
 * the bond graph, lays the system out on the device. */
int bfm_initialize(bfm_updater *h);

/* runSimulationOnGPU, UpdaterGPUScBFM_AB_Type.cu:2902-3400: n_mcs Monte-Carlo
 * steps, then positions are copied back (doCopyBack, :3402-3484) and the age
 * advances by n_mcs. */
int bfm_run_simulation_on_gpu(bfm_updater *h, uint32_t n_mcs);

/* No counterpart in the reference: bfm_run_simulation_on_gpu for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same result as
 * bfm_run_simulation_on_gpu (n_mcs MCS, positions copied back into the
 * handle's host mirror, the age advanced), but it calls no hipMalloc, hipFree
 * or other allocating routine: the copy-back is staged in the library's
 * static device scratch of 32 MiB per GPU (held until the process ends), so
 * the system may have at most bfm_noalloc_max_monomers() = 2,796,202 monomers
 * (BFM_E_INVALID_ARGUMENT otherwise).  No kernel timing.  Every call on one
 * GPU shares that GPU's scratch: calls that overlap on one GPU (from several
 * threads) race; calls one after another do not.  After bfm_initialize, whose
 * allocations it relies on. */
int bfm_run_simulation_on_gpu_noalloc(bfm_updater *h, uint32_t n_mcs);
uint32_t bfm_noalloc_max_monomers(void);

/* getMonomerPositionIn{X,Y,Z}, UpdaterGPUScBFM_AB_Type.cu:2612-2614
 * (unwrapped coordinates). */
int bfm_get_monomer_position_in_x(const bfm_updater *h, uint32_t i, int32_t *x);
int bfm_get_monomer_position_in_y(const bfm_updater *h, uint32_t i, int32_t *y);
int bfm_get_monomer_position_in_z(const bfm_updater *h, uint32_t i, int32_t *z);

/* cleanup, UpdaterGPUScBFM_AB_Type.cu:3493-3502: frees everything set up by
 * setNrOfAllMonomers and initialize; the handle can be set up again. */
int bfm_cleanup(bfm_updater *h);

/* ---- bulk extensions (no one-call-per-monomer loop over the ABI) ---------- */

/* xyz: 3*n int32, monomer-major.  Same checks as the single-monomer calls. */
int bfm_set_monomer_coordinates_bulk(bfm_updater *h, const int32_t *xyz, uint32_t n);
/* pairs: 2*n_links uint32 (i, j), each a directed link i->j. */
int bfm_set_connectivity_bulk(bfm_updater *h, const uint32_t *pairs, uint64_t n_links);
int bfm_get_monomer_positions(const bfm_updater *h, int32_t *xyz, uint32_t n);

/* Colour of every monomer after bfm_initialize, and the colour count. */
int bfm_get_colors(const bfm_updater *h, uint32_t *color, uint32_t n, uint32_t *n_colors);

/* Counters accumulated since bfm_initialize. */
typedef struct bfm_stats {
    uint64_t attempts;       /* monomer move attempts (MCS x monomers)        */
    uint64_t proposals;      /* attempts that passed walls, bonds, volume     */
    uint64_t accepted;       /* moves applied                                  */
    uint64_t substeps;       /* colour substeps run                            */
    double   check_ms;       /* summed duration of check kernels (if timed)   */
    double   perform_ms;     /* summed duration of perform kernels (if timed) */
    uint64_t timed_launches; /* check launches covered by check_ms             */
    uint64_t sorts;          /* on-device spatial re-sorts so far               */
} bfm_stats;
int bfm_get_stats(const bfm_updater *h, bfm_stats *out);

/* every > 0: the check and perform launches of every `every`-th substep are
 * timed by HIP events on the handle's stream that take each kernel's own
 * dispatch timestamps (hipExtLaunchKernelGGL: the execution time a kernel
 * trace reports), summed into bfm_stats (timed_launches counts the timed
 * substeps); 0: off.  Timed launches cost some launch pipeline, so sampling
 * (e.g. 8) keeps the sweep rate undisturbed. */
int bfm_set_kernel_timing(bfm_updater *h, int every);

/* Block-to-XCD placement of the sweep kernels (no counterpart in the
 * reference).  On (the default), each of the 8 XCDs sweeps one contiguous
 * range of the colour segment, so spatially close monomers share one L2;
 * off, blocks are dealt to the XCDs round-robin as launched.  The trajectory
 * is the same either way. */
int bfm_set_xcd_ranges(bfm_updater *h, int on);

/* Launch configuration (the trajectory is the same under every one).
 * perform group: check waves whose survivors one perform wave takes, 2, 4
 * (default) or 8, for every colour segment.  bfm_autotune (below) reports the
 * group it chose for colour 0 and the placement.  On an error the
 * configuration in force before the call stays.  After initialize. */
int bfm_set_perform_group(bfm_updater *h, uint32_t group);
int bfm_autotune(bfm_updater *h, uint32_t probe_mcs, uint32_t *group, int *xcd_ranges);
/* Check workgroup size in monomers (64, 128, 256 = default, 512 or 1024) for
 * every colour segment; the reference tunes threads per block per species
 * (UpdaterGPUScBFM_AB_Type.cu:2940-2990).  With one monomer per lane (below)
 * that is the number of threads.  Any time (after initialize it applies to
 * every segment). */
int bfm_set_check_block(bfm_updater *h, uint32_t block);
/* Monomers per lane of the check kernel, 1 or 2, for every colour segment (no
 * counterpart in the reference, one thread per monomer): with 2 a workgroup of
 * B monomers has B / 2 threads and each wavefront takes two runs of 64
 * monomers, with the memory loads of both in flight together.  A workgroup of
 * 64 monomers takes one per lane.  Without this call a segment takes 2 where
 * it has at least 4 resident generations of 64-monomer runs (about 2 M
 * monomers on an MI355X), else 1.  Any time. */
int bfm_set_check_items(bfm_updater *h, uint32_t items);
/* Launch configuration of one colour segment (after initialize): its check
 * workgroup size, monomers per lane and perform group; get also reports the
 * block-to-XCD placement and whether the monomers of each colour lie in
 * Morton order of position (1) or in id order (0), and bfm_get_check_items
 * the monomers per lane in force.  bfm_autotune chooses all of them: the
 * order within colours by timing both (re-sorting on the device), then per
 * segment the fastest check shape (workgroup size with one and with two
 * monomers per lane) and perform group from the segment's own kernel
 * durations, under the faster placement, and keeps that pick only if it also
 * beats the defaults by wall time; each choice needs a 2 % lead over the
 * default.  The system advances by 23 * probe_mcs MCS when the pick is the
 * default, else 27 * probe_mcs (four fewer without spatial sorting). */
int bfm_set_segment_config(bfm_updater *h, uint32_t colour, uint32_t block, uint32_t group);
int bfm_set_segment_items(bfm_updater *h, uint32_t colour, uint32_t items);
int bfm_get_launch_config(const bfm_updater *h, uint32_t colour, uint32_t *block, uint32_t *group, int *xcd_ranges,
                          int *morton_order);
int bfm_get_check_items(const bfm_updater *h, uint32_t colour, uint32_t *items);

/* Runs n_mcs steps without the copy-back of runSimulationOnGPU (positions
 * stay on the device; the age still advances).  For timing loops. */
int bfm_run_resident(bfm_updater *h, uint32_t n_mcs);
/* Copies the device state back to the host (the doCopyBack half). */
int bfm_copy_back(bfm_updater *h);

/* ---- analyzers on the device -------------------------------------------------
 * The reference analyses its systems on the host after copying every position
 * back (AnalyzerMsd, src/AnalyzerMsd.tpp:99-421, behind LeMonADE's
 * TaskManager, mainCUDASimulatorGPUScBFM_AB_Type.cpp:166-177).  These reduce
 * the resident state on the device and return only the results, so they work
 * between bfm_run_resident calls without a copy-back.  Integer results are
 * exact; floating-point means come from fixed-order device sums. */
typedef struct bfm_observables {
    uint64_t bonds;           /* undirected bonds                                         */
    uint64_t bond_hist[512];  /* bonds per bond-table index (dx&7)|(dy&7)<<3|(dz&7)<<6,
                                 each oriented from the lower to the higher id           */
    uint32_t n_chains;        /* linear chains; 0 when the bonds are not linear chains    */
    uint32_t reserved;
    int64_t  re2_sum;         /* sum over chains of |r_last - r_first|^2 (exact)          */
    double   re2;             /* mean squared end-to-end distance (NaN without chains)    */
    double   rg2;             /* mean squared radius of gyration (NaN without chains)     */
    double   box_com[3];      /* mean of the chains' centres of mass (AnalyzerMsd.tpp:346-349) */
} bfm_observables;
int bfm_observe(bfm_updater *h, bfm_observables *out);

/* Linear chains as AnalyzerMsd::getLinearPolymerSystemInfo finds them
 * (AnalyzerMsd.tpp:186-288): components in order of their lowest id, ends in
 * the order of its stack walk, the middle monomer floor(n/2) links from the
 * first end.  Host only (no device needed).  Writes up to `capacity` entries
 * of first/middle/last ids (any may be NULL) and the chain count; fails with
 * the reference analyzer's message if the bonds are not linear chains. */
int bfm_chain_table(bfm_updater *h, uint32_t *n_chains, uint32_t *first, uint32_t *middle, uint32_t *last,
                    uint32_t capacity);

/* Per linear chain, relative to its first monomer (exact integers). */
typedef struct bfm_chain_sums {
    int64_t  s[3];            /* sum over members of r - r_first   */
    int64_t  q;               /* sum over members of |r - r_first|^2 */
    int32_t  first[3], middle[3], last[3];
    uint32_t n;               /* members                             */
} bfm_chain_sums;
int bfm_get_chain_sums(bfm_updater *h, bfm_chain_sums *out, uint32_t capacity);

/* AnalyzerMsd on the device: attach = its constructor (rnMcsToSkip,
 * rDeltaTRealization, rDeltaTEvaluate, rnMaxSubsequences; :141-161) plus
 * initialize() (:173-177), which evaluates at the current age; execute =
 * AnalyzerMsd::execute at the current age (:303-421), to be called after each
 * run as the TaskManager does; points = the rows cleanup() writes (:456-467):
 * dt, samples, and for g1..g5 (first, middle, last monomer, chain COM, middle
 * monomer - chain COM) the mean and the reference's "sg" column.  The
 * reference leaves the time of the last realization uninitialised (:107);
 * here the first evaluation starts the first subsequence. */
typedef struct bfm_msd_point {
    uint64_t dt;
    double   n_values;
    double   mean[5];
    double   sg[5];
} bfm_msd_point;
int bfm_msd_attach(bfm_updater *h, uint64_t mcs_to_skip, uint64_t delta_t_realization, uint64_t delta_t_evaluate,
                   uint64_t max_subsequences);
int bfm_msd_execute(bfm_updater *h);
int bfm_msd_get_points(bfm_updater *h, bfm_msd_point *out, uint32_t capacity, uint32_t *n_points);
/* AnalyzerMsd::cleanup's file (:423-468). */
int bfm_msd_write(bfm_updater *h, const char *path);

/* ---- batched form: many independent systems per launch ----------------------
 * No counterpart in the reference, whose box and bond tables are module-global
 * __constant__ symbols (UpdaterGPUScBFM_AB_Type.cu:84-133), so one process
 * can run only one system size at a time.  A batch sweeps every member in the
 * same launches (one check + one perform launch per colour slot for all
 * members), captured as a hipGraph per chunk of MCS.  Each member's trajectory
 * is bit-identical to running it alone with bfm_run_simulation_on_gpu. */
typedef struct bfm_batch bfm_batch;
int bfm_batch_create(bfm_batch **out, int gpu);
int bfm_batch_destroy(bfm_batch *b);
/* Adds the system a configured updater describes (setters done; it need not
 * be initialized; the batch keeps its own copy).  Uses its seed and age.  An
 * initialized updater is copied back first and hands over its current state,
 * so the member continues its trajectory. */
int bfm_batch_add(bfm_batch *b, bfm_updater *h, uint32_t *member);
/* Device bytes the updater's system takes as a batch member. */
int bfm_member_bytes(bfm_updater *h, uint64_t *bytes);
/* n_mcs MCS on every member, in chunks of `chunk` MCS (one graph per chunk). */
int bfm_batch_run(bfm_batch *b, uint32_t n_mcs, uint32_t chunk);
/* The same for the listed members only. */
int bfm_batch_run_members(bfm_batch *b, const uint32_t *members, uint32_t n_members, uint32_t n_mcs,
                          uint32_t chunk);
int bfm_batch_get_positions(bfm_batch *b, uint32_t member, int32_t *xyz, uint32_t n);
int bfm_batch_get_age(bfm_batch *b, uint32_t member, int64_t *age);
/* attempts = 0 here (the batch does not track them); accepted and proposals
 * are the member's device counters; sorts its on-device re-sorts.  A member
 * keeps its updater's setStepsBetweenSortings period and re-sorts at the same
 * ages as the updater would alone (a chunk is split where one falls due). */
int bfm_batch_get_stats(bfm_batch *b, uint32_t member, bfm_stats *out);
/* bfm_observe for a member (its resident state; waits for the member's queued
 * work). */
int bfm_batch_observe(bfm_batch *b, uint32_t member, bfm_observables *out);
/* Frees the member's device memory (in stream order, after its queued work)
 * and its host arrays; its id is not reused. */
int bfm_batch_remove(bfm_batch *b, uint32_t member);
int bfm_batch_set_graphs(bfm_batch *b, int on);
/* Working-set budget of bfm_batch_run in bytes (no counterpart in the
 * reference, which runs one system): where the members' hot arrays exceed it,
 * bfm_batch_run splits them into groups of about that size and runs each
 * group's n_mcs in turn, so that a group stays in the memory-side cache; the
 * trajectories are unchanged.  0 never splits; the default is 128 MiB. */
int bfm_batch_set_working_set_bytes(bfm_batch *b, uint64_t bytes);
/* Number of groups bfm_batch_run would use now. */
int bfm_batch_get_working_set_groups(bfm_batch *b, uint32_t *groups);
/* As bfm_set_xcd_ranges, over the batch's whole grid (each member's blocks
 * stay on one or two XCDs). */
int bfm_batch_set_xcd_ranges(bfm_batch *b, int on);
/* As bfm_set_perform_group / bfm_autotune for the whole batch (probes run in
 * chunks of `chunk` MCS on every active member). */
int bfm_batch_set_perform_group(bfm_batch *b, uint32_t group);
int bfm_batch_autotune(bfm_batch *b, uint32_t probe_mcs, uint32_t chunk, uint32_t *group, int *xcd_ranges);
int bfm_batch_synchronize(bfm_batch *b);

/* ---- served form: continuous batching of simulation jobs --------------------
 * No counterpart in the reference.  Jobs queue in arrival order; every
 * bfm_server_step admits queued jobs while their device bytes fit the budget,
 * runs one chunk of MCS on all admitted jobs together and retires finished
 * ones (positions copied back, device memory freed).  States: */
enum { BFM_JOB_QUEUED = 0, BFM_JOB_RUNNING = 1, BFM_JOB_DONE = 2, BFM_JOB_FAILED = 3, BFM_JOB_CANCELLED = 4 };
typedef struct bfm_server bfm_server;
/* hbm_budget = 0: 80% of the free device memory at creation. */
int bfm_server_create(bfm_server **out, int gpu, uint64_t hbm_budget, uint32_t chunk_mcs);
int bfm_server_destroy(bfm_server *s);
/* Queues the configured updater's system (validated here) for n_mcs MCS;
 * an initialized updater hands over its positions as bfm_batch_add does.
 * A job larger than the whole budget is accepted as BFM_JOB_FAILED. */
int bfm_server_submit(bfm_server *s, bfm_updater *h, uint32_t n_mcs, uint64_t *job);
/* As bfm_server_submit, but a configured updater that is not initialized
 * hands its validated system over instead of a copy: nothing is copied, and
 * the updater is left as after bfm_cleanup (its seed and age kept).  An
 * initialized updater is submitted as by bfm_server_submit.  For serving
 * loops whose clients build one updater per job. */
int bfm_server_submit_take(bfm_server *s, bfm_updater *h, uint32_t n_mcs, uint64_t *job);
/* One scheduling round; *pending = jobs still queued or running. */
int bfm_server_step(bfm_server *s, uint32_t *pending);
int bfm_server_job_state(bfm_server *s, uint64_t job, int *state, uint32_t *mcs_done);
/* Seconds since server creation at submission, admission and completion (-1 if not yet). */
int bfm_server_job_times(bfm_server *s, uint64_t job, double *submitted, double *admitted, double *finished);
/* Final positions (3 * n int32, by id; n = the job's monomer count).  They are
 * unsorted and unwrapped on the device when the job retires and held in
 * pinned host memory until bfm_server_release. */
int bfm_server_job_result(bfm_server *s, uint64_t job, int32_t *xyz, uint32_t n);
/* Thread-local message of a failed job. */
int bfm_server_job_error(bfm_server *s, uint64_t job, char *buf, uint32_t len);
int bfm_server_cancel(bfm_server *s, uint64_t job);
int bfm_server_release(bfm_server *s, uint64_t job);
/* Queued / running jobs, device bytes in use and budget, attempts done. */
int bfm_server_info(bfm_server *s, uint32_t *queued, uint32_t *running, uint64_t *bytes_in_use,
                    uint64_t *budget, uint64_t *attempts);
/* Host memory the server holds for its jobs (queued systems, results not yet
 * released, the systems of admitted jobs), the jobs it still holds (every job
 * until bfm_server_release) and the batch members it still tracks.  A job's
 * host arrays go with bfm_server_job_result's first call (the positions stay
 * until release) or with release; a served process that releases its jobs
 * returns to its baseline. */
int bfm_server_host_bytes(bfm_server *s, uint64_t *host_bytes, uint32_t *jobs, uint32_t *members);
/* How the chunks were launched: graph launches, graphs written by an update of
 * an existing executable graph, graphs instantiated anew (any may be NULL). */
int bfm_server_graph_stats(bfm_server *s, uint64_t *launches, uint64_t *updates, uint64_t *instantiations);
int bfm_batch_graph_stats(bfm_batch *b, uint64_t *launches, uint64_t *updates, uint64_t *instantiations);

/* Measurement helper, no counterpart in the reference: times `repeats`
 * launches of a plain copy kernel (one 16-byte element per lane, one pass of
 * the grid, at most 2^31 - 1 blocks of 256) over `bytes` on device `gpu`
 * (after one warm-up launch).  A copy of B bytes reads B and writes B
 * bytes of HBM, which calibrates the FETCH_SIZE / WRITE_SIZE counters and
 * gives the achievable streaming bandwidth 2*B / avg_ms. */
int bfm_calibration_copy(int gpu, uint64_t bytes, int repeats, double *avg_ms);

/* Number of HIP devices visible (0 on a machine without a GPU). */
int bfm_device_count(int *count);

#ifdef __cplusplus
}
#endif
#endif
