// capi.cpp -- extern "C" boundary: status codes and a thread-local message,
// never an exception or exit() across the ABI.
#include <cstring>
#include <new>
#include <string>

#include "../../include/pscbfm_amd.h"
#include "updater.hpp"

struct bfm_updater {
    pscbfm::UpdaterGPUScBFM_AB_Type u;
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
    });
}

int bfm_set_kernel_timing(bfm_updater *h, int on)
{
    NEED_HANDLE(h);
    return guarded([&] { h->u.setKernelTiming(on != 0); });
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
