// Driver for the reference's own graph colouring (src/pscbfm/graphColoring.tpp,
// included from where it lies under the reference tree by oracle/ref.mk).
// Reads graphs from stdin, each as "uniform n" followed by n lines
// "k j_1 ... j_k" (monomer i's links in the order setConnectivity received
// them), and prints each graph's colours on one line, "n_colours c_0 ... c_{n-1}".
// The call is the one initializeSpeciesSorting makes
// (UpdaterGPUScBFM_AB_Type.cu:1561-1566) on per-monomer link lists.
// Test infrastructure only: its output pins the oracle's orc_graph_coloring.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "graphColoring.tpp"

int main()
{
    int uniform = 0;
    unsigned long n = 0;
    while (std::scanf("%d %lu", &uniform, &n) == 2) {
        std::vector<std::vector<uint32_t>> links(n);
        for (unsigned long i = 0; i < n; ++i) {
            unsigned k = 0;
            if (std::scanf("%u", &k) != 1)
                return 1;
            links[i].resize(k);
            for (unsigned q = 0; q < k; ++q)
                if (std::scanf("%u", &links[i][q]) != 1)
                    return 1;
        }
        typedef std::vector<std::vector<uint32_t>> Links;
        const std::vector<uint32_t> colours = graphColoring<Links, uint32_t, uint32_t>(
            links, links.size(), uniform != 0,
            [](Links const &x, uint32_t const &i) { return x[i].size(); },
            [](Links const &x, uint32_t const &i, size_t const &j) { return x[i][j]; });
        uint32_t k = 0;
        for (uint32_t c : colours)
            k = c + 1 > k ? c + 1 : k;
        std::printf("%u", k);
        for (uint32_t c : colours)
            std::printf(" %u", c);
        std::printf("\n");
    }
    return 0;
}
