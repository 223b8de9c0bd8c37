/*
 * pscbfm_amd_lemonade.hpp -- header-only LeMonADE updater adapter over the C
 * ABI; the drop-in for GPUScBFM_AB_Type<T_IngredientsType>
 * (src/pscbfm/GPUScBFM_AB_Type.h:27-286).
 *
 * Usage inside LeMonADE:
 *     #include <LeMonADE/updater/AbstractUpdater.h>
 *     #include <LeMonADE/utility/Vector3D.h>
 *     #include "pscbfm_amd_lemonade.hpp"
 *     using Updater = pscbfm_amd::GPUScBFM_AB_Type<Ing, VectorInt3, AbstractUpdater>;
 *     taskManager.addUpdater(new Updater(ingredients, saveInterval, gpu, seed));
 *
 * The third template argument is the base class (LeMonADE's AbstractUpdater
 * with virtual initialize/execute/cleanup); Vec3 is the vector type the bond
 * set's isValid() takes.  Errors become std::invalid_argument /
 * std::runtime_error like the reference updater's.
 */
#ifndef PSCBFM_AMD_LEMONADE_HPP
#define PSCBFM_AMD_LEMONADE_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "pscbfm_amd.h"

namespace pscbfm_amd {

struct NoBase {
    virtual ~NoBase() = default;
};

inline void check(int rc, const char *what)
{
    if (rc == BFM_OK)
        return;
    const std::string msg = std::string(what) + ": " + bfm_last_error();
    if (rc == BFM_E_INVALID_ARGUMENT)
        throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

template <class T_IngredientsType, class Vec3, class Base = NoBase>
class GPUScBFM_AB_Type : public Base {
public:
    using IngredientsType = T_IngredientsType;

    GPUScBFM_AB_Type(T_IngredientsType &ingredients, uint32_t n_steps, int gpu = 0,
                     uint64_t seed = 0x5EED000000000001ull)
        : mIngredients(ingredients), mnSteps(n_steps), miGpu(gpu), mSeed(seed)
    {
    }
    ~GPUScBFM_AB_Type() override
    {
        if (mH)
            bfm_destroy(mH);
    }
    GPUScBFM_AB_Type(const GPUScBFM_AB_Type &) = delete;
    GPUScBFM_AB_Type &operator=(const GPUScBFM_AB_Type &) = delete;

    void setGpu(int gpu) { miGpu = gpu; }
    void setSeed(uint64_t seed) { mSeed = seed; }
    // GPUScBFM_AB_Type.h:110-111: the generator to use (BFM_RNG_INT_HASH or
    // BFM_RNG_PHILOX); -1, the default, keeps the updater's own default.
    void setRng(int rng) { miRng = rng; }

    // GPUScBFM_AB_Type.h:120-184: forward box, periodicity, age, positions,
    // attributes, both link directions and the bond set, then initialize.
    void initialize()
    {
        if (!mH)
            check(bfm_create(&mH), "bfm_create");
        auto &ing = mIngredients;
        const auto &mol = ing.getMolecules();
        check(bfm_set_gpu(mH, miGpu), "setGpu");
        if (miRng != -1)   // forwarded only when set, as GPUScBFM_AB_Type.h:132-133
            check(bfm_set_rng(mH, miRng), "setRng");
        check(bfm_set_seed(mH, mSeed), "setSeed");
        check(bfm_set_age(mH, int64_t(ing.modifyMolecules().getAge())), "setAge");
        check(bfm_set_periodicity(mH, ing.isPeriodicX(), ing.isPeriodicY(), ing.isPeriodicZ()), "setPeriodicity");
        check(bfm_set_lattice_size(mH, uint64_t(ing.getBoxX()), uint64_t(ing.getBoxY()), uint64_t(ing.getBoxZ())),
              "setLatticeSize");
        const uint32_t n = uint32_t(mol.size());
        check(bfm_set_nr_of_all_monomers(mH, n), "setNrOfAllMonomers");
        for (uint32_t i = 0; i < n; ++i)
            check(bfm_set_monomer_coordinates(mH, i, mol[i].getX(), mol[i].getY(), mol[i].getZ()),
                  "setMonomerCoordinates");
        for (uint32_t i = 0; i < n; ++i)
            check(bfm_set_attribute(mH, i, int32_t(mol[i].getAttributeTag())), "setAttribute");
        for (uint32_t i = 0; i < n; ++i)
            for (size_t b = 0; b < mol.getNumLinks(i); ++b)
                check(bfm_set_connectivity(mH, i, uint32_t(mol.getNeighborIdx(i, b))), "setConnectivity");
        for (int dx = -4; dx < 4; ++dx)
            for (int dy = -4; dy < 4; ++dy)
                for (int dz = -4; dz < 4; ++dz)
                    check(bfm_copy_bond_set(mH, dx, dy, dz, !ing.getBondset().isValid(Vec3(dx, dy, dz))),
                          "copyBondSet");
        check(bfm_initialize(mH), "initialize");
    }

    // GPUScBFM_AB_Type.h:191-242: run mnSteps MCS, copy positions back, age += mnSteps.
    bool execute()
    {
        auto &ing = mIngredients;
        auto &mol = ing.modifyMolecules();
        check(bfm_set_age(mH, int64_t(mol.getAge())), "setAge");
        check(bfm_run_simulation_on_gpu(mH, mnSteps), "runSimulationOnGPU");
        // one bulk read of the copied-back positions instead of the
        // reference's three getters per monomer (GPUScBFM_AB_Type.h:207-214)
        const uint32_t n = uint32_t(mol.size());
        mPositions.resize(3 * size_t(n));
        check(bfm_get_monomer_positions(mH, mPositions.data(), n), "getMonomerPositions");
        for (uint32_t i = 0; i < n; ++i)
            mol[i].setAllCoordinates(mPositions[3 * size_t(i)], mPositions[3 * size_t(i) + 1],
                                     mPositions[3 * size_t(i) + 2]);
        mol.setAge(mol.getAge() + mnSteps);
        return true;
    }

    void cleanup()
    {
        if (mH)
            check(bfm_cleanup(mH), "cleanup");
    }

private:
    T_IngredientsType &mIngredients;
    uint32_t mnSteps;
    int miGpu;
    uint64_t mSeed;
    int miRng = -1;
    std::vector<int32_t> mPositions;
    bfm_updater *mH = nullptr;
};

}  // namespace pscbfm_amd

#endif
