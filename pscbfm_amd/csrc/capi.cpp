// capi.cpp -- extern "C" boundary: status codes and a thread-local message,
// never an exception or exit() across the ABI.
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/pscbfm_amd.h"
#include "bfm_io.hpp"
#include "server.hpp"
#include "noalloc.hpp"
#include "updater.hpp"

struct bfm_updater {
    pscbfm::UpdaterGPUScBFM_AB_Type u;
};

struct bfm_batch {
    std::unique_ptr<pscbfm::BatchEngine> e;
};

struct bfm_server {
    std::unique_ptr<pscbfm::Server> s;
};

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace pscbfm {
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace pscbfm

namespace {

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

template <class F>
int guarded(F &&f)
{
    try {
        f();
        g_last_error.clear();
        return BFM_OK;
    } catch (const pscbfm::NoDeviceError &e) {
        return fail(BFM_E_NO_DEVICE, e.what());
    } catch (const pscbfm::HipError &e) {
        return fail(e.code == hipErrorOutOfMemory ? BFM_E_OUT_OF_MEMORY : BFM_E_HIP, e.what());
    } catch (const pscbfm::StateError &e) {
        return fail(BFM_E_STATE, e.what());
    } catch (const std::invalid_argument &e) {
        return fail(BFM_E_INVALID_ARGUMENT, e.what());
    } catch (const std::bad_alloc &) {
        return fail(BFM_E_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(BFM_E_RUNTIME, e.what());
    } catch (...) {
        return fail(BFM_E_RUNTIME, "unknown error");
    }
}

#define NEED_HANDLE(h)                                                   \
    do {                                                                 \
        if (!(h))                                                        \
            return fail(BFM_E_INVALID_ARGUMENT, "null updater handle"); \
    } while (0)

}  // namespace

extern "C" {

const char *bfm_last_error(void) { return g_last_error.c_str(); }

int bfm_create(bfm_updater **out)
{
    if (!out)
        return fail(BFM_E_INVALID_ARGUMENT, "null output pointer");
    return guarded([&] { *out = new bfm_updater(); });
}

int bfm_destroy(bfm_updater *h)
{
    delete h;
    return BFM_OK;
}

int bfm_set_gpu(bfm_updater *h, int gpu)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setGpu(gpu); });
}

int bfm_copy_bond_set(bfm_updater *h, int dx, int dy, int dz, int forbidden)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.copyBondSet(dx, dy, dz, forbidden != 0); });
}

int bfm_set_lattice_size(bfm_updater *h, uint64_t x, uint64_t y, uint64_t z)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setLatticeSize(x, y, z); });
}

int bfm_set_nr_of_all_monomers(bfm_updater *h, uint32_t n)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setNrOfAllMonomers(n); });
}

int bfm_set_attribute(bfm_updater *h, uint32_t i, int32_t a)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setAttribute(i, a); });
}

int bfm_set_monomer_coordinates(bfm_updater *h, uint32_t i, int32_t x, int32_t y, int32_t z)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setMonomerCoordinates(i, x, y, z); });
}

int bfm_set_connectivity(bfm_updater *h, uint32_t i, uint32_t j)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setConnectivity(i, j); });
}

int bfm_set_periodicity(bfm_updater *h, int px, int py, int pz)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setPeriodicity(px != 0, py != 0, pz != 0); });
}

int bfm_set_age(bfm_updater *h, int64_t age)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setAge(age); });
}

int bfm_get_age(const bfm_updater *h, int64_t *age)
{
    NEED_HANDLE(h);
    if (!age)
        return fail(BFM_E_INVALID_ARGUMENT, "null output pointer");
    return guarded([&] { *age = h->u.getAge(); });
}

int bfm_set_rng(bfm_updater *h, int rng)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setRng(rng); });
}

int bfm_get_rng(const bfm_updater *h, int *rng)
{
    NEED_HANDLE(h);
    if (!rng)
        return fail(BFM_E_INVALID_ARGUMENT, "null output pointer");
    return guarded([&] { *rng = h->u.getRng(); });
}

int bfm_set_seed(bfm_updater *h, uint64_t seed)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setSeed(seed); });
}

int bfm_set_steps_between_sortings(bfm_updater *h, int64_t steps)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setStepsBetweenSortings(steps); });
}

int bfm_get_steps_between_sortings(const bfm_updater *h, int64_t *steps)
{
    NEED_HANDLE(h);
    if (!steps)
        return fail(BFM_E_INVALID_ARGUMENT, "null output pointer");
    return guarded([&] { *steps = h->u.getStepsBetweenSortings(); });
}

int bfm_set_spatial_sorting(bfm_updater *h, int on)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setSpatialSorting(on != 0); });
}

int bfm_validate(bfm_updater *h)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.validate(); });
}

int bfm_initialize(bfm_updater *h)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.initialize(); });
}

int bfm_run_simulation_on_gpu(bfm_updater *h, uint32_t n_mcs)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.runSimulationOnGPU(n_mcs); });
}

int bfm_run_simulation_on_gpu_noalloc(bfm_updater *h, uint32_t n_mcs)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.runSimulationOnGPUNoAlloc(n_mcs); });
}

uint32_t bfm_noalloc_max_monomers(void) { return pscbfm::kNoAllocMaxMonomers; }

int bfm_run_resident(bfm_updater *h, uint32_t n_mcs)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.runResident(n_mcs); });
}

int bfm_copy_back(bfm_updater *h)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.copyBack(); });
}

int bfm_get_monomer_position_in_x(const bfm_updater *h, uint32_t i, int32_t *v)
{
    NEED_HANDLE(h);
    if (!v)
        return fail(BFM_E_INVALID_ARGUMENT, "null output pointer");
    return guarded([&] { *v = h->u.getMonomerPositionInX(i); });
}

int bfm_get_monomer_position_in_y(const bfm_updater *h, uint32_t i, int32_t *v)
{
    NEED_HANDLE(h);
    if (!v)
        return fail(BFM_E_INVALID_ARGUMENT, "null output pointer");
    return guarded([&] { *v = h->u.getMonomerPositionInY(i); });
}

int bfm_get_monomer_position_in_z(const bfm_updater *h, uint32_t i, int32_t *v)
{
    NEED_HANDLE(h);
    if (!v)
        return fail(BFM_E_INVALID_ARGUMENT, "null output pointer");
    return guarded([&] { *v = h->u.getMonomerPositionInZ(i); });
}

int bfm_cleanup(bfm_updater *h)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.cleanup(); });
}

int bfm_read_bfm_file(bfm_updater *h, const char *path)
{
    NEED_HANDLE(h);
    if (!path)
        return fail(BFM_E_INVALID_ARGUMENT, "null path");
    return guarded([&] {
        const pscbfm::BfmState f = pscbfm::read_bfm_file(path);
        auto &u = h->u;
        u.setLatticeSize(f.box[0], f.box[1], f.box[2]);
        u.setPeriodicity(f.periodic[0], f.periodic[1], f.periodic[2]);
        u.setNrOfAllMonomers(f.n);
        for (uint32_t i = 0; i < f.n; ++i) {
            u.setMonomerCoordinates(i, f.xyz[3 * size_t(i)], f.xyz[3 * size_t(i) + 1], f.xyz[3 * size_t(i) + 2]);
            u.setAttribute(i, f.attributes[i]);
        }
        // both directions, as the LeMonADE adapter forwards every link of every
        // monomer (GPUScBFM_AB_Type.h:159-161)
        for (auto b : f.bonds) {
            u.setConnectivity(b.first, b.second);
            u.setConnectivity(b.second, b.first);
        }
        bool allowed[512] = {};
        for (auto &v : f.bondset_ids)
            allowed[v.first] = true;
        for (int dx = -4; dx < 4; ++dx)
            for (int dy = -4; dy < 4; ++dy)
                for (int dz = -4; dz < 4; ++dz)
                    u.copyBondSet(dx, dy, dz, !allowed[pscbfm::bond_table_index(dx, dy, dz)]);
        u.setAge(f.age);
    });
}

int bfm_write_bfm_file(bfm_updater *h, const char *path, int append)
{
    NEED_HANDLE(h);
    if (!path)
        return fail(BFM_E_INVALID_ARGUMENT, "null path");
    return guarded([&] {
        const pscbfm::HostSystem &sys = h->u.system();
        if (sys.n == 0 || sys.xyz.size() != 3 * size_t(sys.n))
            throw std::runtime_error("write_bfm_file: the updater holds no monomers");
        pscbfm::BfmState f;
        for (int a = 0; a < 3; ++a) {
            f.box[a] = sys.box[a];
            f.periodic[a] = sys.periodic[a];
        }
        f.n = sys.n;
        f.xyz = sys.xyz;
        for (uint32_t i = 0; i < sys.n; ++i)
            for (uint32_t j = 0; j < sys.nnbr[i]; ++j) {
                const uint32_t k = sys.nbr[size_t(i) * pscbfm::kMaxConnectivity + j];
                if (i < k || std::find(sys.nbr.begin() + size_t(k) * pscbfm::kMaxConnectivity,
                                       sys.nbr.begin() + size_t(k) * pscbfm::kMaxConnectivity + sys.nnbr[k],
                                       i) == sys.nbr.begin() + size_t(k) * pscbfm::kMaxConnectivity + sys.nnbr[k])
                    f.bonds.push_back({std::min(i, k), std::max(i, k)});
            }
        std::sort(f.bonds.begin(), f.bonds.end());
        f.bonds.erase(std::unique(f.bonds.begin(), f.bonds.end()), f.bonds.end());
        f.bondset_ids = pscbfm::bondset_ids_for(sys.forbidden);
        f.attributes = sys.attribute;
        f.age = h->u.getAge();
        if (append)
            pscbfm::append_bfm_conformation(path, f);
        else
            pscbfm::write_bfm_file(path, f);
    });
}

int bfm_set_monomer_coordinates_bulk(bfm_updater *h, const int32_t *xyz, uint32_t n)
{
    NEED_HANDLE(h);
    if (!xyz && n)
        return fail(BFM_E_INVALID_ARGUMENT, "null coordinate array");
    return guarded([&] {
        for (uint32_t i = 0; i < n; ++i)
            h->u.setMonomerCoordinates(i, xyz[3 * size_t(i)], xyz[3 * size_t(i) + 1], xyz[3 * size_t(i) + 2]);
    });
}

int bfm_set_connectivity_bulk(bfm_updater *h, const uint32_t *pairs, uint64_t n_links)
{
    NEED_HANDLE(h);
    if (!pairs && n_links)
        return fail(BFM_E_INVALID_ARGUMENT, "null link array");
    return guarded([&] {
        for (uint64_t k = 0; k < n_links; ++k)
            h->u.setConnectivity(pairs[2 * k], pairs[2 * k + 1]);
    });
}

int bfm_get_monomer_positions(const bfm_updater *h, int32_t *xyz, uint32_t n)
{
    NEED_HANDLE(h);
    if (!xyz && n)
        return fail(BFM_E_INVALID_ARGUMENT, "null coordinate array");
    return guarded([&] {
        const auto &s = h->u.system();
        if (n != s.n)
            throw std::invalid_argument("bfm_get_monomer_positions: n must equal the number of monomers");
        std::memcpy(xyz, s.xyz.data(), sizeof(int32_t) * 3 * size_t(n));
    });
}

int bfm_get_colors(const bfm_updater *h, uint32_t *color, uint32_t n, uint32_t *n_colors)
{
    NEED_HANDLE(h);
    return guarded([&] {
        const auto &s = h->u.system();
        if (!h->u.validated() && !h->u.initialized())
            throw pscbfm::StateError("colours exist only after bfm_validate or bfm_initialize");
        if (color) {
            if (n != s.n)
                throw std::invalid_argument("bfm_get_colors: n must equal the number of monomers");
            std::memcpy(color, s.colour.data(), sizeof(uint32_t) * n);
        }
        if (n_colors)
            *n_colors = s.n_colours;
    });
}

int bfm_get_stats(const bfm_updater *h, bfm_stats *out)
{
    NEED_HANDLE(h);
    if (!out)
        return fail(BFM_E_INVALID_ARGUMENT, "null output pointer");
    return guarded([&] {
        std::memset(out, 0, sizeof *out);
        uint64_t acc = 0, prop = 0;
        h->u.deviceCounters(acc, prop);
        out->attempts = h->u.attempts();
        out->proposals = prop;
        out->accepted = acc;
        out->substeps = h->u.substeps();
        out->check_ms = h->u.checkMs();
        out->perform_ms = h->u.performMs();
        out->timed_launches = h->u.timedLaunches();
        out->sorts = h->u.sorts();
    });
}

int bfm_set_kernel_timing(bfm_updater *h, int every)
{
    NEED_HANDLE(h);
    if (every < 0)
        return fail(BFM_E_INVALID_ARGUMENT, "kernel timing interval must be >= 0");
    return guarded([&] { h->u.setKernelTiming(uint32_t(every)); });
}

int bfm_set_xcd_ranges(bfm_updater *h, int on)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setXcdRanges(on != 0); });
}

namespace {
void export_observables(const pscbfm::Observables &o, bfm_observables *out)
{
    std::memset(out, 0, sizeof(*out));
    out->bonds = o.bonds;
    std::memcpy(out->bond_hist, o.hist, sizeof(out->bond_hist));
    out->n_chains = o.n_chains;
    out->re2_sum = o.re2_sum;
    out->re2 = o.re2;
    out->rg2 = o.rg2;
    for (int a = 0; a < 3; ++a)
        out->box_com[a] = o.box_com[a];
}
}  // namespace

int bfm_observe(bfm_updater *h, bfm_observables *out)
{
    NEED_HANDLE(h);
    if (!out)
        return fail(BFM_E_INVALID_ARGUMENT, "null output");
    return guarded([&] {
        pscbfm::Observables o;
        h->u.observe(o);
        export_observables(o, out);
    });
}

int bfm_chain_table(bfm_updater *h, uint32_t *n_chains, uint32_t *first, uint32_t *middle, uint32_t *last,
                    uint32_t capacity)
{
    NEED_HANDLE(h);
    if (!n_chains)
        return fail(BFM_E_INVALID_ARGUMENT, "null output");
    return guarded([&] {
        const pscbfm::ChainTable t = pscbfm::find_linear_chains(h->u.system());
        if (!t.msd_error.empty())
            throw std::invalid_argument(t.msd_error);
        *n_chains = t.n_chains;
        for (uint32_t c = 0; c < std::min(capacity, t.n_chains); ++c) {
            const uint32_t b = t.offsets[c], n = t.offsets[c + 1] - b;
            if (first)
                first[c] = t.members[b];
            if (middle)
                middle[c] = t.members[b + n / 2];
            if (last)
                last[c] = t.members[b + n - 1];
        }
    });
}

static_assert(sizeof(bfm_chain_sums) == sizeof(pscbfm::ChainSums), "chain sums layout");

int bfm_get_chain_sums(bfm_updater *h, bfm_chain_sums *out, uint32_t capacity)
{
    NEED_HANDLE(h);
    if (!out)
        return fail(BFM_E_INVALID_ARGUMENT, "null output");
    return guarded([&] {
        std::vector<pscbfm::ChainSums> v;
        h->u.chainSums(v);
        if (capacity < v.size())
            throw std::invalid_argument("output holds " + std::to_string(capacity) + " chains, the system has " +
                                        std::to_string(v.size()));
        std::memcpy(out, v.data(), sizeof(bfm_chain_sums) * v.size());
    });
}

int bfm_msd_attach(bfm_updater *h, uint64_t mcs_to_skip, uint64_t delta_t_realization, uint64_t delta_t_evaluate,
                   uint64_t max_subsequences)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.msdAttach(mcs_to_skip, delta_t_realization, delta_t_evaluate, max_subsequences); });
}

int bfm_msd_execute(bfm_updater *h)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.msdExecute(); });
}

int bfm_msd_get_points(bfm_updater *h, bfm_msd_point *out, uint32_t capacity, uint32_t *n_points)
{
    NEED_HANDLE(h);
    if (!n_points)
        return fail(BFM_E_INVALID_ARGUMENT, "null output");
    return guarded([&] {
        if (!h->u.analyzer().msdAttached())
            throw pscbfm::StateError("no MSD analyzer attached");
        const auto pts = h->u.analyzer().msdPoints();
        *n_points = uint32_t(pts.size());
        for (uint32_t k = 0; out && k < std::min<size_t>(capacity, pts.size()); ++k) {
            out[k].dt = pts[k].dt;
            out[k].n_values = pts[k].n_values;
            for (int g = 0; g < 5; ++g) {
                out[k].mean[g] = pts[k].mean[g];
                out[k].sg[g] = pts[k].sg[g];
            }
        }
    });
}

int bfm_msd_write(bfm_updater *h, const char *path)
{
    NEED_HANDLE(h);
    if (!path)
        return fail(BFM_E_INVALID_ARGUMENT, "null path");
    return guarded([&] {
        if (!h->u.analyzer().msdAttached())
            throw pscbfm::StateError("no MSD analyzer attached");
        h->u.analyzer().msdWrite(h->u.system(), path);
    });
}

int bfm_set_perform_group(bfm_updater *h, uint32_t group)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setPerformGroup(group); });
}

int bfm_autotune(bfm_updater *h, uint32_t probe_mcs, uint32_t *group, int *xcd_ranges)
{
    NEED_HANDLE(h);
    return guarded([&] {
        h->u.autotune(probe_mcs);
        if (group)
            *group = h->u.performGroup();
        if (xcd_ranges)
            *xcd_ranges = h->u.xcdRanges() ? 1 : 0;
    });
}

int bfm_set_check_block(bfm_updater *h, uint32_t block)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setCheckBlock(block); });
}

int bfm_set_check_items(bfm_updater *h, uint32_t items)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setCheckItems(items); });
}

int bfm_set_segment_config(bfm_updater *h, uint32_t colour, uint32_t block, uint32_t group)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setSegmentConfig(colour, block, group); });
}

int bfm_set_segment_items(bfm_updater *h, uint32_t colour, uint32_t items)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setSegmentItems(colour, items); });
}

int bfm_get_check_items(const bfm_updater *h, uint32_t colour, uint32_t *items)
{
    NEED_HANDLE(h);
    if (!items)
        return fail(BFM_E_INVALID_ARGUMENT, "null items");
    return guarded([&] { *items = h->u.segmentItems(colour); });
}

int bfm_get_launch_config(const bfm_updater *h, uint32_t colour, uint32_t *block, uint32_t *group, int *xcd_ranges,
                          int *morton_order)
{
    NEED_HANDLE(h);
    return guarded([&] {
        uint32_t t = 0, g = 0;
        h->u.segmentConfig(colour, t, g);
        if (block)
            *block = t;
        if (group)
            *group = g;
        if (xcd_ranges)
            *xcd_ranges = h->u.xcdRanges() ? 1 : 0;
        if (morton_order)
            *morton_order = h->u.mortonLayout() ? 1 : 0;
    });
}

int bfm_batch_create(bfm_batch **out, int gpu)
{
    if (!out)
        return fail(BFM_E_INVALID_ARGUMENT, "null output pointer");
    return guarded([&] {
        auto b = std::make_unique<bfm_batch>();
        b->e = std::make_unique<pscbfm::BatchEngine>(gpu);
        *out = b.release();
    });
}

int bfm_batch_destroy(bfm_batch *b)
{
    delete b;
    return BFM_OK;
}

#define NEED(p, what)                                              \
    do {                                                           \
        if (!(p))                                                  \
            return fail(BFM_E_INVALID_ARGUMENT, "null " what);    \
    } while (0)

int bfm_member_bytes(bfm_updater *h, uint64_t *bytes)
{
    NEED(h, "updater handle");
    NEED(bytes, "output pointer");
    return guarded([&] {
        if (!h->u.validated())
            h->u.validate();
        *bytes = pscbfm::BatchEngine::bytesFor(h->u.system());
    });
}

// The system an updater hands to a batch or a server: planned from the
// updater's inputs, or, for an initialized updater, copied back and re-planned
// from its current positions (its own planned arrays were staging for its
// upload and are gone; the generator is keyed on id and age, so the member
// continues the updater's trajectory whatever its layout).
static pscbfm::HostSystem system_for_batch(pscbfm::UpdaterGPUScBFM_AB_Type &u)
{
    if (u.initialized()) {
        u.copyBack();
        pscbfm::HostSystem s(u.system());
        pscbfm::plan(s);
        s.steps_between_sortings = u.getStepsBetweenSortings();
        return s;
    }
    if (!u.validated())
        u.validate();
    pscbfm::HostSystem s(u.system());
    s.steps_between_sortings = u.getStepsBetweenSortings();
    return s;
}

int bfm_batch_add(bfm_batch *b, bfm_updater *h, uint32_t *member)
{
    NEED(b, "batch handle");
    NEED(h, "updater handle");
    NEED(member, "output pointer");
    return guarded([&] { *member = b->e->add(system_for_batch(h->u), h->u.seed(), h->u.getAge()); });
}

int bfm_batch_run(bfm_batch *b, uint32_t n_mcs, uint32_t chunk)
{
    NEED(b, "batch handle");
    return guarded([&] { b->e->run(n_mcs, chunk); });
}

int bfm_batch_run_members(bfm_batch *b, const uint32_t *members, uint32_t n_members, uint32_t n_mcs,
                          uint32_t chunk)
{
    NEED(b, "batch handle");
    if (!members && n_members)
        return fail(BFM_E_INVALID_ARGUMENT, "null member list");
    return guarded([&] {
        std::vector<uint32_t> ids(members, members + n_members);
        for (uint32_t id : ids)
            if (id >= b->e->size() || !b->e->active(id))
                throw std::invalid_argument("batch: no such active member");
        b->e->runMembers(ids, n_mcs, chunk);
    });
}

int bfm_batch_get_positions(bfm_batch *b, uint32_t member, int32_t *xyz, uint32_t n)
{
    NEED(b, "batch handle");
    NEED(xyz, "coordinate array");
    return guarded([&] {
        if (member >= b->e->size())
            throw std::invalid_argument("batch: no such member");
        b->e->copyBack(member);
        const auto &s = b->e->system(member);
        if (n != s.n)
            throw std::invalid_argument("bfm_batch_get_positions: n must equal the member's monomer count");
        std::memcpy(xyz, s.xyz.data(), sizeof(int32_t) * 3 * size_t(n));
    });
}

int bfm_batch_get_age(bfm_batch *b, uint32_t member, int64_t *age)
{
    NEED(b, "batch handle");
    NEED(age, "output pointer");
    return guarded([&] {
        if (member >= b->e->size())
            throw std::invalid_argument("batch: no such member");
        *age = b->e->age(member);
    });
}

int bfm_batch_observe(bfm_batch *b, uint32_t member, bfm_observables *out)
{
    NEED(b, "batch handle");
    NEED(out, "output pointer");
    return guarded([&] {
        if (member >= b->e->size())
            throw std::invalid_argument("batch: no such member");
        pscbfm::Observables o;
        b->e->observe(member, o);
        export_observables(o, out);
    });
}

int bfm_batch_get_stats(bfm_batch *b, uint32_t member, bfm_stats *out)
{
    NEED(b, "batch handle");
    NEED(out, "output pointer");
    return guarded([&] {
        if (member >= b->e->size())
            throw std::invalid_argument("batch: no such member");
        std::memset(out, 0, sizeof *out);
        uint64_t acc = 0, prop = 0;
        b->e->counters(member, acc, prop);
        out->accepted = acc;
        out->proposals = prop;
        out->sorts = b->e->sorts(member);
    });
}

int bfm_batch_remove(bfm_batch *b, uint32_t member)
{
    NEED(b, "batch handle");
    return guarded([&] {
        if (member >= b->e->size())
            throw std::invalid_argument("batch: no such member");
        b->e->forget(member);
    });
}

int bfm_batch_set_graphs(bfm_batch *b, int on)
{
    NEED(b, "batch handle");
    return guarded([&] { b->e->setUseGraphs(on != 0); });
}

int bfm_batch_set_working_set_bytes(bfm_batch *b, uint64_t bytes)
{
    NEED(b, "batch handle");
    return guarded([&] { b->e->setWorkingSetBytes(size_t(bytes)); });
}

int bfm_batch_get_working_set_groups(bfm_batch *b, uint32_t *groups)
{
    NEED(b, "batch handle");
    NEED(groups, "groups");
    return guarded([&] {
        std::vector<uint32_t> ids;
        for (uint32_t i = 0; i < b->e->size(); ++i)
            if (b->e->active(i))
                ids.push_back(i);
        *groups = uint32_t(b->e->workingSetGroups(ids).size());
    });
}

int bfm_batch_set_xcd_ranges(bfm_batch *b, int on)
{
    NEED(b, "batch handle");
    return guarded([&] { b->e->setXcdRanges(on != 0); });
}

int bfm_batch_set_perform_group(bfm_batch *b, uint32_t group)
{
    NEED(b, "batch handle");
    return guarded([&] { b->e->setPerformGroup(group); });
}

int bfm_batch_autotune(bfm_batch *b, uint32_t probe_mcs, uint32_t chunk, uint32_t *group, int *xcd_ranges)
{
    NEED(b, "batch handle");
    return guarded([&] {
        b->e->autotune(probe_mcs, chunk);
        if (group)
            *group = b->e->performGroup();
        if (xcd_ranges)
            *xcd_ranges = b->e->xcdRanges() ? 1 : 0;
    });
}

int bfm_batch_synchronize(bfm_batch *b)
{
    NEED(b, "batch handle");
    return guarded([&] { b->e->synchronize(); });
}

int bfm_server_create(bfm_server **out, int gpu, uint64_t hbm_budget, uint32_t chunk_mcs)
{
    NEED(out, "output pointer");
    return guarded([&] {
        auto s = std::make_unique<bfm_server>();
        s->s = std::make_unique<pscbfm::Server>(gpu, hbm_budget, chunk_mcs);
        *out = s.release();
    });
}

int bfm_server_destroy(bfm_server *s)
{
    delete s;
    return BFM_OK;
}

int bfm_server_submit(bfm_server *s, bfm_updater *h, uint32_t n_mcs, uint64_t *job)
{
    NEED(s, "server handle");
    NEED(h, "updater handle");
    NEED(job, "output pointer");
    return guarded([&] { *job = s->s->submit(system_for_batch(h->u), h->u.seed(), h->u.getAge(), n_mcs); });
}

int bfm_server_submit_take(bfm_server *s, bfm_updater *h, uint32_t n_mcs, uint64_t *job)
{
    NEED(s, "server handle");
    NEED(h, "updater handle");
    NEED(job, "output pointer");
    return guarded([&] {
        const uint64_t seed = h->u.seed();
        const int64_t age = h->u.getAge();
        if (h->u.initialized()) {
            *job = s->s->submit(system_for_batch(h->u), seed, age, n_mcs);
            return;
        }
        // validated here if needed, then moved: nothing is copied
        const int64_t period = h->u.getStepsBetweenSortings();
        pscbfm::HostSystem sys = h->u.takeSystem();
        sys.steps_between_sortings = period;
        h->u.setSeed(seed);
        h->u.setAge(age);
        *job = s->s->submit(std::move(sys), seed, age, n_mcs);
    });
}

int bfm_server_step(bfm_server *s, uint32_t *pending)
{
    NEED(s, "server handle");
    return guarded([&] {
        const uint32_t p = s->s->step();
        if (pending)
            *pending = p;
    });
}

int bfm_server_job_state(bfm_server *s, uint64_t job, int *state, uint32_t *mcs_done)
{
    NEED(s, "server handle");
    return guarded([&] {
        if (state)
            *state = int(s->s->state(job));
        if (mcs_done)
            *mcs_done = s->s->mcsDone(job);
    });
}

int bfm_server_job_times(bfm_server *s, uint64_t job, double *submitted, double *admitted, double *finished)
{
    NEED(s, "server handle");
    NEED(submitted, "output pointer");
    NEED(admitted, "output pointer");
    NEED(finished, "output pointer");
    return guarded([&] { s->s->times(job, *submitted, *admitted, *finished); });
}

int bfm_server_job_result(bfm_server *s, uint64_t job, int32_t *xyz, uint32_t n)
{
    NEED(s, "server handle");
    NEED(xyz, "coordinate array");
    return guarded([&] { s->s->result(job, xyz, n); });
}

int bfm_server_job_error(bfm_server *s, uint64_t job, char *buf, uint32_t len)
{
    NEED(s, "server handle");
    NEED(buf, "buffer");
    return guarded([&] {
        const std::string &e = s->s->error(job);
        if (len) {
            const size_t k = std::min<size_t>(len - 1, e.size());
            std::memcpy(buf, e.data(), k);
            buf[k] = 0;
        }
    });
}

int bfm_server_cancel(bfm_server *s, uint64_t job)
{
    NEED(s, "server handle");
    return guarded([&] { s->s->cancel(job); });
}

int bfm_server_release(bfm_server *s, uint64_t job)
{
    NEED(s, "server handle");
    return guarded([&] { s->s->release(job); });
}

int bfm_server_host_bytes(bfm_server *s, uint64_t *host_bytes, uint32_t *jobs, uint32_t *members)
{
    NEED(s, "server handle");
    return guarded([&] {
        if (host_bytes)
            *host_bytes = s->s->hostBytes();
        if (jobs)
            *jobs = s->s->jobsHeld();
        if (members)
            *members = s->s->engine().liveMembers();
    });
}

int bfm_server_info(bfm_server *s, uint32_t *queued, uint32_t *running, uint64_t *bytes_in_use, uint64_t *budget,
                    uint64_t *attempts)
{
    NEED(s, "server handle");
    return guarded([&] {
        if (queued)
            *queued = s->s->queued();
        if (running)
            *running = s->s->running();
        if (bytes_in_use)
            *bytes_in_use = s->s->bytesInUse();
        if (budget)
            *budget = s->s->budget();
        if (attempts)
            *attempts = s->s->attemptsDone();
    });
}

namespace {
void graph_stats(pscbfm::BatchEngine &e, uint64_t *launches, uint64_t *updates, uint64_t *instantiations)
{
    if (launches)
        *launches = e.graphLaunches();
    if (updates)
        *updates = e.graphUpdates();
    if (instantiations)
        *instantiations = e.graphInstantiations();
}
}  // namespace

int bfm_server_graph_stats(bfm_server *s, uint64_t *launches, uint64_t *updates, uint64_t *instantiations)
{
    NEED(s, "server handle");
    return guarded([&] { graph_stats(s->s->engine(), launches, updates, instantiations); });
}

int bfm_batch_graph_stats(bfm_batch *b, uint64_t *launches, uint64_t *updates, uint64_t *instantiations)
{
    NEED(b, "batch handle");
    return guarded([&] { graph_stats(*b->e, launches, updates, instantiations); });
}

int bfm_calibration_copy(int gpu, uint64_t bytes, int repeats, double *avg_ms)
{
    if (!avg_ms || repeats <= 0 || bytes == 0 || bytes % 16)
        return fail(BFM_E_INVALID_ARGUMENT, "calibration copy: bytes must be a positive multiple of 16");
    return guarded([&] {
        pscbfm::hip_check(hipSetDevice(gpu), "hipSetDevice");
        void *src = nullptr, *dst = nullptr;
        hipStream_t st = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto release = [&] {
            if (e0) hipEventDestroy(e0);
            if (e1) hipEventDestroy(e1);
            if (st) hipStreamDestroy(st);
            if (src) hipFree(src);
            if (dst) hipFree(dst);
        };
        try {
            pscbfm::hip_check(hipMalloc(&src, bytes), "hipMalloc");
            pscbfm::hip_check(hipMalloc(&dst, bytes), "hipMalloc");
            pscbfm::hip_check(hipStreamCreate(&st), "hipStreamCreate");
            pscbfm::hip_check(hipMemsetAsync(src, 1, bytes, st), "hipMemsetAsync");
            pscbfm::hip_check(hipEventCreate(&e0), "hipEventCreate");
            pscbfm::hip_check(hipEventCreate(&e1), "hipEventCreate");
            pscbfm::hip_check(pscbfm::launch_copy(src, dst, bytes, st), "copy kernel");  // warm-up
            pscbfm::hip_check(hipEventRecord(e0, st), "hipEventRecord");
            for (int r = 0; r < repeats; ++r)
                pscbfm::hip_check(pscbfm::launch_copy(src, dst, bytes, st), "copy kernel");
            pscbfm::hip_check(hipEventRecord(e1, st), "hipEventRecord");
            pscbfm::hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.f;
            pscbfm::hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            *avg_ms = double(ms) / repeats;
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}

int bfm_device_count(int *count)
{
    if (!count)
        return fail(BFM_E_INVALID_ARGUMENT, "null output pointer");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    g_last_error.clear();
    return BFM_OK;
}

}  // extern "C"
