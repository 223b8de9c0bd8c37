// sim_main.cpp -- command-line simulator with the reference's flags
// (src/pscbfm/mainCUDASimulatorGPUScBFM_AB_Type.cpp:20-195), on the C ABI.
//
// Reads the last conformation of a LeMonADE BFM file, runs the colour sweep on
// one GPU in cycles of --save-interval MCS and appends the state after every
// cycle to the output file (the reference's TaskManager of
// UpdaterReadBfmFile -> GPUScBFM_AB_Type -> AnalyzerWriteBfmFile, :166-177).
#include <getopt.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "../../include/pscbfm_amd.h"

namespace {

void print_help()
{
    std::printf(
        "usage: pscbfm_amd_sim [options]\n"
        "\n"
        "Colour-sweep bond-fluctuation-model simulator on one MI355X.\n"
        "\n"
        "    -e, --seeds <file path>\n"
        "        seed file: its first two unsigned 32-bit values form the 64-bit seed\n"
        "    -i, --initial-state <file path>\n"
        "        (required) BFM file to continue from (its last conformation)\n"
        "    -m, --max-mcs <integer>\n"
        "        (required) total Monte-Carlo steps to simulate\n"
        "    -r, --rng <integer>\n"
        "        random number generator: %d: IntHash (default), %d: Philox\n"
        "    -s, --save-interval <integer>\n"
        "        append the state to the output file every <integer> MCS (default: max-mcs)\n"
        "    -o, --output <file path>\n"
        "        output BFM file (default outfile.bfm); conformations are appended if it exists\n"
        "    -g, --gpu <integer>\n"
        "        GPU to use, 0 .. number of GPUs - 1\n"
        "    -v, --version\n",
        BFM_RNG_INT_HASH, BFM_RNG_PHILOX);
}

bool exists_nonempty(const std::string &path)
{
    struct stat st;
    return stat(path.c_str(), &st) == 0 && st.st_size > 0;
}

int check(int rc, const char *what)
{
    if (rc != BFM_OK)
        std::fprintf(stderr, "[pscbfm_amd_sim] %s failed: %s\n", what, bfm_last_error());
    return rc;
}

}  // namespace

int main(int argc, char **argv)
{
    using clock = std::chrono::steady_clock;
    const auto t_program0 = clock::now();
    std::string infile, outfile = "outfile.bfm", seed_file;
    long long max_mcs = -1, save_interval = 0;
    int gpu = 0, rng = -1;
    if (argc <= 1) {
        print_help();
        return 0;
    }
    static struct option long_options[] = {
        {"seeds", required_argument, 0, 'e'},  {"gpu", required_argument, 0, 'g'},
        {"help", no_argument, 0, 'h'},         {"initial-state", required_argument, 0, 'i'},
        {"max-mcs", required_argument, 0, 'm'}, {"output", required_argument, 0, 'o'},
        {"rng", required_argument, 0, 'r'},    {"save-interval", required_argument, 0, 's'},
        {"version", no_argument, 0, 'v'},      {0, 0, 0, 0}};
    while (true) {
        int option_index = 0;
        const int c = getopt_long(argc, argv, "e:g:hi:m:o:r:s:v", long_options, &option_index);
        if (c == -1)
            break;
        switch (c) {
        case 'e': seed_file = optarg; break;
        case 'g': gpu = std::atoi(optarg); break;
        case 'h': print_help(); return 0;
        case 'i': infile = optarg; break;
        case 'm': max_mcs = std::atoll(optarg); break;
        case 'o': outfile = optarg; break;
        case 'r': rng = std::atoi(optarg); break;
        case 's': save_interval = std::atoll(optarg); break;
        case 'v': std::printf("pscbfm_amd_sim (colour sweep for gfx950)\n"); return 0;
        default: print_help(); return 1;
        }
    }
    if (infile.empty() || max_mcs < 0) {
        std::fprintf(stderr, "[pscbfm_amd_sim] --initial-state and --max-mcs are required\n");
        return 1;
    }
    if (save_interval <= 0)
        save_interval = max_mcs > 0 ? max_mcs : 1;

    uint64_t seed = 0x5EED000000000001ull;
    if (!seed_file.empty()) {
        std::ifstream f(seed_file);
        unsigned long long hi = 0, lo = 0;
        if (!(f >> hi >> lo)) {
            std::fprintf(stderr, "[pscbfm_amd_sim] seed file '%s' needs two unsigned 32-bit values\n",
                         seed_file.c_str());
            return 1;
        }
        seed = ((hi & 0xFFFFFFFFull) << 32) | (lo & 0xFFFFFFFFull);
    }

    bfm_updater *h = nullptr;
    if (check(bfm_create(&h), "bfm_create"))
        return 1;
    int rc = 0;
    do {
        if ((rc = check(bfm_set_gpu(h, gpu), "setGpu")))
            break;
        if (rng != -1 && (rc = check(bfm_set_rng(h, rng), "setRng")))
            break;
        if ((rc = check(bfm_set_seed(h, seed), "seed")))
            break;
        if ((rc = check(bfm_read_bfm_file(h, infile.c_str()), "reading the initial state")))
            break;
        if ((rc = check(bfm_initialize(h), "initialize")))
            break;
        const bool append = exists_nonempty(outfile);
        if ((rc = check(bfm_write_bfm_file(h, outfile.c_str(), append ? 1 : 0), "writing the output")))
            break;
        const long long cycles = max_mcs / save_interval;
        const auto t0 = clock::now();
        uint32_t n = 0;
        for (long long k = 0; k < cycles && rc == 0; ++k) {
            if ((rc = check(bfm_run_simulation_on_gpu(h, uint32_t(save_interval)), "runSimulationOnGPU")))
                break;
            rc = check(bfm_write_bfm_file(h, outfile.c_str(), 1), "appending to the output");
        }
        if (rc)
            break;
        const double dt = std::chrono::duration<double>(clock::now() - t0).count();
        bfm_stats st{};
        bfm_get_stats(h, &st);
        n = uint32_t(st.attempts / std::max<long long>(1, cycles * save_interval));
        std::fprintf(stderr, "tTaskLoop = %gs\n", dt);
        std::fprintf(stderr, "monomers %u, %lld MCS: %.4g attempted moves/s, acceptance %.4f\n", n,
                     cycles * save_interval, dt > 0 ? double(st.attempts) / dt : 0.0,
                     st.attempts ? double(st.accepted) / double(st.attempts) : 0.0);
    } while (false);
    bfm_cleanup(h);
    bfm_destroy(h);
    std::fprintf(stderr, "tProgram = %gs\n",
                 std::chrono::duration<double>(clock::now() - t_program0).count());
    return rc ? 1 : 0;
}
