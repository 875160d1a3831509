// sdfx_grid_plan.hpp — how a batch of grid requests is cut into runs (host
// code only, no HIP; tests compile it on its own).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace sdfx {

// cells[i]: the tile-padded cells request i puts in a run, 0 for a request
// that needs none. Requests are kept in order; a run takes requests while
// its cells stay within `cap`, and a request larger than `cap` runs alone.
// Returns the runs as lists of request indices.
inline std::vector<std::vector<size_t>> plan_grid_runs(const std::vector<uint64_t>& cells, uint64_t cap) {
    std::vector<std::vector<size_t>> runs;
    uint64_t run_cells = 0;
    for (size_t i = 0; i < cells.size(); ++i) {
        if (!cells[i]) continue;
        if (runs.empty() || run_cells + cells[i] > cap) {
            runs.emplace_back();
            run_cells = 0;
        }
        runs.back().push_back(i);
        run_cells += cells[i];
    }
    return runs;
}

}  // namespace sdfx
