// Drives include/pscbfm_amd_lemonade.hpp with a minimal stand-in for the
// LeMonADE Ingredients interface the reference adapter uses
// (src/pscbfm/GPUScBFM_AB_Type.h:120-242), the way a LeMonADE TaskManager
// would: initialize, execute x calls, cleanup.
//   lemonade_adapter_demo <in.txt> <out.txt> [rng]
// (rng: the generator forwarded through setRng, GPUScBFM_AB_Type.h:110-111)
// in.txt:  box n_monomers n_bonds seed steps_per_call calls age
//          then n_monomers lines "x y z", then n_bonds lines "i j"
// out.txt: final age, then n_monomers lines "x y z"
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "pscbfm_amd_lemonade.hpp"

struct Vec3 {
    int x, y, z;
    Vec3(int a, int b, int c) : x(a), y(b), z(c) {}
};

struct Monomer {
    int32_t x = 0, y = 0, z = 0, attribute = 1;
    int32_t getX() const { return x; }
    int32_t getY() const { return y; }
    int32_t getZ() const { return z; }
    int32_t getAttributeTag() const { return attribute; }
    void setAllCoordinates(int32_t a, int32_t b, int32_t c) { x = a; y = b; z = c; }
};

struct Molecules {
    std::vector<Monomer> m;
    std::vector<std::vector<uint32_t>> links;
    uint64_t age = 0;
    size_t size() const { return m.size(); }
    Monomer &operator[](size_t i) { return m[i]; }
    const Monomer &operator[](size_t i) const { return m[i]; }
    size_t getNumLinks(size_t i) const { return links[i].size(); }
    uint32_t getNeighborIdx(size_t i, size_t b) const { return links[i][b]; }
    uint64_t getAge() const { return age; }
    void setAge(uint64_t a) { age = a; }
};

struct BondSet {
    // the standard BFM set: |v|^2 in {4, 5, 6, 9, 10}
    bool isValid(const Vec3 &v) const
    {
        const int s = v.x * v.x + v.y * v.y + v.z * v.z;
        return s == 4 || s == 5 || s == 6 || s == 9 || s == 10;
    }
};

struct Ingredients {
    uint64_t box = 0;
    Molecules mol;
    BondSet bonds;
    uint64_t getBoxX() const { return box; }
    uint64_t getBoxY() const { return box; }
    uint64_t getBoxZ() const { return box; }
    bool isPeriodicX() const { return true; }
    bool isPeriodicY() const { return true; }
    bool isPeriodicZ() const { return true; }
    const Molecules &getMolecules() const { return mol; }
    Molecules &modifyMolecules() { return mol; }
    const BondSet &getBondset() const { return bonds; }
};

int main(int argc, char **argv)
{
    if (argc != 3 && argc != 4) {
        std::cerr << "usage: lemonade_adapter_demo in.txt out.txt [rng]\n";
        return 2;
    }
    std::ifstream in(argv[1]);
    uint64_t box, n, nb, seed, steps, calls, age;
    in >> box >> n >> nb >> seed >> steps >> calls >> age;
    Ingredients ing;
    ing.box = box;
    ing.mol.m.resize(n);
    ing.mol.links.resize(n);
    ing.mol.age = age;
    for (uint64_t i = 0; i < n; ++i)
        in >> ing.mol.m[i].x >> ing.mol.m[i].y >> ing.mol.m[i].z;
    for (uint64_t k = 0; k < nb; ++k) {
        uint32_t i, j;
        in >> i >> j;
        ing.mol.links[i].push_back(j);
        ing.mol.links[j].push_back(i);
    }
    if (!in) {
        std::cerr << "bad input\n";
        return 2;
    }
    pscbfm_amd::GPUScBFM_AB_Type<Ingredients, Vec3> updater(ing, uint32_t(steps), 0, seed);
    if (argc == 4)
        updater.setRng(std::atoi(argv[3]));
    try {
        updater.initialize();
        for (uint64_t c = 0; c < calls; ++c)
            updater.execute();
        updater.cleanup();
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 3;
    }
    std::ofstream out(argv[2]);
    out << ing.mol.age << "\n";
    for (uint64_t i = 0; i < n; ++i)
        out << ing.mol.m[i].x << " " << ing.mol.m[i].y << " " << ing.mol.m[i].z << "\n";
    return 0;
}
