// Schedule planner interface (see plan.cpp).
#pragma once
#include <string>
#include <vector>

namespace lux {

struct Plan {
    int W = 0, H = 0, T = 0, xsw = 0;
    int status = 0;          // 0 ok, -1 bad argument, -2 unsupported
    std::string error;
    int seed_row = -1;       // last pass +1 row written to line-buffer half 0 (-1: none)
    // border replay lists, 4 ints per visit: y, half, path-0 label, path-1 label
    // (label -1: initial fill; y' >= 0: neighbour column at row y', same pass; -2-c: pass +1 seed of cell c)
    std::vector<int> plus;   // pass +1, column W-1
    std::vector<int> minus;  // pass -1, column 0
    // per border-column row, the indices of its visits in plus / minus (two slots, -1 = none)
    std::vector<int> cellvis_plus, cellvis_minus;
};

Plan make_plan(int W, int H, int T, int xsw);
int seed_row_of(int H, int T);

}  // namespace lux
