// Host-side work lists of the two-step engine (TbTiling in rtm_core.hpp).
#pragma once
#include <algorithm>
#include <map>
#include <vector>

#include "rtm_core.hpp"

namespace psdm {

// Injection events of one shot's points (px, py: 1-based window columns and rows of n
// points) for the two-step work items: points outside the interior get no source (the
// reference injects only interior cells) and duplicate cells keep the last point (last
// writer wins).  For each item -- edge tiles 0 .. ne-1, then core items ne .. ne+ncore-1 --
// every such cell it computes: an edge tile's own rectangle +-3, a core item's columns
// c0-3 .. cend+2 and the rows of its lanes.  Out: ptr (ne + ncore + 1 offsets), cell (iw << 16
// | jw, ordered by column then row within an item), inj (the cell's value column) and cols
// (the point index of each value column; a cell in several items has one column).
inline void tb_events(const TbTiling &TB, int ny, int nx, int dn, const int *px, const int *py, int n,
                      std::vector<int> &ptr, std::vector<int> &cell, std::vector<int> &inj, std::vector<int> &cols) {
    std::map<int, int> last;  // (iw << 16 | jw) -> point; ordered by column, then row
    for (int q = 0; q < n; q++) {
        const int jw = py[q] - 1, iw = px[q] - 1;
        if (jw < 0 || jw >= ny || iw < 0 || iw >= nx) continue;
        last[(iw << 16) | jw] = q;
    }
    std::vector<int> keys, col_of;
    for (auto &kv : last) {
        keys.push_back(kv.first);
        col_of.push_back((int)cols.size());
        cols.push_back(kv.second);
    }
    // the cells with iw in [a, b) and jw in [r0, r1), in key order
    auto add = [&](int a, int b, int r0, int r1) {
        auto it = std::lower_bound(keys.begin(), keys.end(), a < 0 ? 0 : a << 16);
        for (; it != keys.end() && (*it >> 16) < b; ++it) {
            const int jw = *it & 0xffff;
            if (jw >= r0 && jw < r1) {
                cell.push_back(*it);
                inj.push_back(col_of[it - keys.begin()]);
            }
        }
    };
    ptr.assign(TB.ne + TB.ncore + 1, 0);
    for (int t = 0; t < TB.ne; t++) {
        int j0, j1, i0, i1;
        tb_edge_tile(TB, t, ny, nx, dn, &j0, &j1, &i0, &i1);
        add(i0 - 3 - dn, i1 + 3 - dn, j0 - 3 - dn, j1 + 3 - dn);
        ptr[t + 1] = (int)cell.size();
    }
    for (int u = 0; u < TB.ncore; u++) {
        const int CS = u / TB.cr, R = u - CS * TB.cr, o = tb_origin(R, ny);
        const int c0 = 3 + CS * TB.sw, cend = std::min(c0 + TB.sw, nx - 3);
        add(c0 - 3, cend + 3, o - 8, o + 248);
        ptr[TB.ne + u + 1] = (int)cell.size();
    }
}

}  // namespace psdm
