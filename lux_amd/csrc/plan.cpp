// Schedule planner: what the reference's sheared-tile wavefront (DenseStereoSgm.cs:206-222) does at the two
// border columns, in the form the device kernels replay.
//
// The reference runs every Pixel call (:234-317) against two persistent line-buffer halves (:28-29,
// :158-166). Away from the border columns each call reads the states of its geometric neighbours, so the
// eight path recurrences can run as independent sweeps on the device. Three things differ, all decided
// by the schedule alone (W, H, T, LUX_SGM_XSW):
//   * pass +1 clamps x-start to W-1 (:217), so some cells of column W-1 are visited twice; a second
//     visit reads whatever the two halves hold at that moment (rows below, other visits);
//   * pass -1 mirrors this at column 0;
//   * pass -1's first row reads half 0 left over from pass +1 (the buffers are never reset, :60).
// The planner walks the schedule restricted to the columns {0, 1, W-2, W-1}, tracking for each line-buffer
// slot which visit wrote it last ("labels"), and emits for each border column the ordered visit list with
// the label every path read. It also checks the properties the device decomposition relies on and reports
// LUX_SGM_EUNSUPPORTED when one does not hold (odd T breaks the row/half alternation).
#include "plan.h"

#include <algorithm>
#include <cstdio>

namespace lux {

namespace {

struct Label {
    int pass = 0;  // 0: the initial fill (2000); +1 / -1: the pass of the visit that wrote the slot
    int x = 0, y = 0, v = 0;
};

struct Tracker {
    int W;
    std::vector<int> cells;                   // tracked cells, ascending
    std::vector<Label> lab[2];                // per half, per tracked cell
    std::vector<int> visits[2];               // per pass (0: +1, 1: -1), per tracked cell, per row
    int H;
    int idx(int c) const {
        for (size_t i = 0; i < cells.size(); i++)
            if (cells[i] == c) return (int)i;
        return -1;
    }
};

}  // namespace

int seed_row_of(int H, int T)
{
    for (int y = H - 1; y >= 0; y--) {
        int r = y % T;
        if ((r & 1) == 1) return y;  // rows with odd tile-row index r write half 0 (:219)
    }
    return -1;
}

Plan make_plan(int W, int H, int T, int xsw)
{
    Plan P;
    P.W = W;
    P.H = H;
    P.T = T;
    P.xsw = xsw;
    if (W <= 0 || H <= 0 || T <= 0) {
        P.status = -1;
        P.error = "bad layer size or tile";
        return P;
    }
    if (T & 1) {
        P.status = -2;
        P.error = "odd tile sizes are not supported on the device (the line-buffer halves stop alternating by row)";
        return P;
    }
    P.seed_row = seed_row_of(H, T);
    Tracker tr;
    tr.W = W;
    tr.H = H;
    for (int c : {0, 1, W - 2, W - 1})
        if (c >= 0 && c < W && tr.idx(c) < 0) tr.cells.push_back(c);
    std::sort(tr.cells.begin(), tr.cells.end());
    for (int h = 0; h < 2; h++) tr.lab[h].assign(tr.cells.size(), Label());
    for (int p = 0; p < 2; p++) tr.visits[p].assign(tr.cells.size() * (size_t)H, 0);

    auto fail = [&](const char* what) {
        if (P.status == 0) {
            P.status = -2;
            char buf[256];
            std::snprintf(buf, sizeof buf, "schedule W=%d H=%d T=%d: %s", W, H, T, what);
            P.error = buf;
        }
    };

    auto visit = [&](int dir, int x, int y, int cur) {
        int ci = tr.idx(x);
        int pi = dir > 0 ? 0 : 1;
        int& nv = tr.visits[pi][(size_t)ci * H + y];
        nv++;
        int other = 1 - cur;
        bool border = dir > 0 ? x == W - 1 : x == 0;
        if (border) {
            int nbc = x - dir;  // column of the path 0 / path 1 neighbours
            int nbi = (nbc >= 0 && nbc < W) ? tr.idx(nbc) : -1;
            int code[2];
            for (int p = 0; p < 2; p++) {
                int half = p == 0 ? cur : other;
                if (nbi < 0) { code[p] = -1; continue; }
                const Label& L = tr.lab[half][nbi];
                if (L.pass == 0) code[p] = -1;
                else if (L.pass == dir) {
                    if (L.x != nbc || L.v != 1) fail("a border visit reads a neighbour column visited twice");
                    code[p] = L.y;
                } else {  // pass -1 reading pass +1's leftovers: only half 0 of path 1 is kept as a seed
                    if (p != 1 || half != 0) fail("a border visit reads pass +1 state the seed table does not hold");
                    code[p] = -2 - nbc;
                }
            }
            // path 2 reads the border slot itself (kept by the replay); pass +1 leftovers only from half 0
            const Label& L2 = tr.lab[other][ci];
            if (L2.pass != 0 && L2.pass != dir && other != 0) fail("a border visit reads pass +1's half 1");
            // path 3 reads the margin slot (x + dir is outside the image)
            if (x + dir >= 0 && x + dir < W) fail("path 3 of a border visit is not the margin");
            auto& list = dir > 0 ? P.plus : P.minus;
            list.push_back(y);
            list.push_back(cur);
            list.push_back(code[0]);
            list.push_back(code[1]);
        }
        tr.lab[cur][ci] = Label{dir, x, y, nv};
    };

    for (int pass = 0; pass < 2 && P.status == 0; pass++) {
        int dir = pass == 0 ? 1 : -1;
        int gwN = (W + T + T - 1) / T, ghN = (H + T - 1) / T;
        for (int j = 0; j < ghN; j++)
            for (int i = 0; i < gwN; i++) {
                int x0 = dir > 0 ? i * T : W - 1 - i * T, y0 = dir > 0 ? j * T : H - 1 - j * T;
                int yEnd = std::clamp(y0 + dir * T, -1, H);
                for (int r = 0;; r++) {
                    int y = y0 + dir * r;
                    if (y == yEnd) break;
                    int xs = std::clamp(x0 - dir * r, 0, (xsw && dir > 0) ? W : W - 1);
                    int xe = std::clamp(x0 + dir * (T - r), -1, W);
                    if (xs == xe) continue;
                    int cur = (r & 1) == 0 ? 1 : 0;
                    if (dir > 0) {
                        for (int c : tr.cells)
                            if (c >= xs && c < xe) visit(dir, c, y, cur);
                    } else {
                        for (auto it = tr.cells.rbegin(); it != tr.cells.rend(); ++it)
                            if (*it <= xs && *it > xe) visit(dir, *it, y, cur);
                    }
                }
            }
        if (dir > 0) {
            // the seed table holds, for interior cells, the state of row seed_row: check the tracked ones
            for (size_t k = 0; k < tr.cells.size(); k++) {
                int c = tr.cells[k];
                if (c == W - 1) continue;
                const Label& L = tr.lab[0][k];
                bool ok = P.seed_row < 0 ? L.pass == 0 : (L.pass == 1 && L.x == c && L.y == P.seed_row && L.v == 1);
                if (!ok) fail("pass +1 leaves half 0 of an interior cell at an unexpected row");
            }
        }
    }
    // per-row visit index lists of the border columns (at most two visits per cell and pass)
    for (int pi = 0; pi < 2 && P.status == 0; pi++) {
        const std::vector<int>& list = pi == 0 ? P.plus : P.minus;
        std::vector<int>& cv = pi == 0 ? P.cellvis_plus : P.cellvis_minus;
        cv.assign((size_t)2 * H, -1);
        for (size_t i = 0; i < list.size() / 4; i++) {
            int y = list[4 * i];
            if (cv[2 * y] < 0) cv[2 * y] = (int)i;
            else if (cv[2 * y + 1] < 0) cv[2 * y + 1] = (int)i;
            else fail("a border cell is visited more than twice in one pass");
        }
    }
    // every cell must have been visited in both passes; interior tracked cells exactly once
    for (int pi = 0; pi < 2; pi++)
        for (size_t k = 0; k < tr.cells.size(); k++) {
            int c = tr.cells[k];
            bool border = pi == 0 ? c == W - 1 : c == 0;
            for (int y = 0; y < H; y++) {
                int n = tr.visits[pi][k * H + y];
                if (n < 1 || (!border && n != 1)) fail("a cell is not visited the way the device path assumes");
            }
        }
    return P;
}

}  // namespace lux
