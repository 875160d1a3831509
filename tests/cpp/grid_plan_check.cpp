// Checks plan_grid_runs (mesh_to_sdf_amd/csrc/sdfx_grid_plan.hpp) on random
// batches: every request with cells is in exactly one run, in order; a run
// holds at most `cap` cells unless it is a single request; and runs are
// greedy (the next run's first request would not have fit in the previous
// run). Prints "ok <runs>" or the first violation; exit status 0 on success.
#include <cstdio>
#include <random>

#include "../../mesh_to_sdf_amd/csrc/sdfx_grid_plan.hpp"

int main() {
    std::mt19937_64 g(12345);
    size_t total_runs = 0;
    for (int trial = 0; trial < 2000; ++trial) {
        const uint64_t cap = 1 + g() % 5000;
        std::vector<uint64_t> cells(g() % 40);
        for (uint64_t& c : cells) c = (g() % 4 == 0) ? 0 : 1 + g() % (2 * cap);
        const auto runs = sdfx::plan_grid_runs(cells, cap);
        total_runs += runs.size();
        size_t next = 0;  // every nonzero request, in order
        for (size_t r = 0; r < runs.size(); ++r) {
            if (runs[r].empty()) return std::printf("empty run\n"), 1;
            uint64_t sum = 0;
            for (size_t i : runs[r]) {
                while (next < cells.size() && cells[next] == 0) ++next;
                if (i != next) return std::printf("trial %d: request %zu out of order\n", trial, i), 1;
                ++next;
                sum += cells[i];
            }
            if (sum > cap && runs[r].size() > 1) return std::printf("trial %d: run over the cap\n", trial), 1;
            if (r > 0) {
                uint64_t prev = 0;
                for (size_t i : runs[r - 1]) prev += cells[i];
                if (prev + cells[runs[r][0]] <= cap) return std::printf("trial %d: runs not greedy\n", trial), 1;
            }
        }
        while (next < cells.size() && cells[next] == 0) ++next;
        if (next != cells.size()) return std::printf("trial %d: a request is missing\n", trial), 1;
    }
    std::printf("ok %zu\n", total_runs);
    return 0;
}
