// qt_graph.hip -- host-side slicing of a device graph into layers that
// connect only to their neighbours (the block-tridiagonal structure the RGF
// sweeps need), the role of graph_partition::slice (graph_partition.f90:426-541).
//
// Points are 0-based.  The graph is held as CSR with its edges deduplicated.
//   one contact:  slice k = points at breadth-first distance k from the
//                 contact (slice_1contact, :426-450);
//   two contacts: recursive bisection (slice_2contacts, :452-541): a point
//                 joins the left part when it is strictly closer to the left
//                 contact than to the right one (ties go right, :318-328);
//                 each part's new contact is its points with a neighbour in
//                 the other part (:347-367); a part is bisected again while it
//                 has more than nmax points and more than its outer contact;
//                 the slices of the left part come first.  Contacts sharing a
//                 point give one slice.
// Each part is a connected subgraph holding a path from every one of its
// points to its outer contact (distances to that contact fall by one per
// step along a shortest path while the difference to the other contact's
// distance keeps its sign), so the recursion ends: both parts are non-empty.
// Sub-problems are not copied: points carry the tag of the part they belong
// to, and searches follow only neighbours with the current tag.
#include <algorithm>
#include <cstdint>
#include <limits>
#include <vector>

#include "qt_rgf.h"

namespace {

constexpr int64_t kFar = std::numeric_limits<int64_t>::max();

struct Graph {
    int32_t n = 0;
    std::vector<int64_t> off;  // n + 1
    std::vector<int32_t> adj;
};

Graph build_graph(int32_t n, int64_t ne, const int32_t* e) {
    std::vector<std::pair<int32_t, int32_t>> pairs;
    pairs.reserve((size_t)ne * 2);
    for (int64_t k = 0; k < ne; ++k) {
        const int32_t i = e[2 * k], j = e[2 * k + 1];
        if (i == j) continue;
        pairs.emplace_back(i, j);
        pairs.emplace_back(j, i);
    }
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    Graph g;
    g.n = n;
    g.off.assign((size_t)n + 1, 0);
    for (const auto& p : pairs) ++g.off[(size_t)p.first + 1];
    for (int32_t i = 0; i < n; ++i) g.off[(size_t)i + 1] += g.off[(size_t)i];
    g.adj.resize(pairs.size());
    for (size_t k = 0; k < pairs.size(); ++k) g.adj[k] = pairs[k].second;  // sorted by source
    return g;
}

struct Slicer {
    const Graph& g;
    int32_t nmax;
    std::vector<int32_t> tag;      // part each point belongs to
    std::vector<int64_t> dl, dr;   // distances inside the current part
    std::vector<int32_t> queue;
    std::vector<std::vector<int32_t>> slices;
    int32_t next_tag = 0;

    Slicer(const Graph& gr, int32_t nm) : g(gr), nmax(nm), tag(gr.n, 0), dl(gr.n), dr(gr.n) { queue.resize(gr.n); }

    // breadth-first distances from src over the points tagged t
    void bfs(const std::vector<int32_t>& pts, const std::vector<int32_t>& src, int32_t t, std::vector<int64_t>& d) {
        for (int32_t p : pts) d[p] = kFar;
        size_t head = 0, tail = 0;
        for (int32_t s : src) {
            if (d[s] != kFar) continue;
            d[s] = 0;
            queue[tail++] = s;
        }
        while (head < tail) {
            const int32_t p = queue[head++];
            for (int64_t k = g.off[p]; k < g.off[(size_t)p + 1]; ++k) {
                const int32_t q = g.adj[k];
                if (tag[q] != t || d[q] != kFar) continue;
                d[q] = d[p] + 1;
                queue[tail++] = q;
            }
        }
    }

    // pts: the part's points in ascending order, all tagged t
    void bisect(const std::vector<int32_t>& pts, const std::vector<int32_t>& e1, const std::vector<int32_t>& e2,
                int32_t t) {
        {
            bool shared = false;
            for (int32_t p : e1) tag[p] = -1 - t;  // mark the left contact
            for (int32_t p : e2) shared = shared || tag[p] == -1 - t;
            for (int32_t p : e1) tag[p] = t;
            if (shared) {
                slices.push_back(pts);
                return;
            }
        }
        bfs(pts, e1, t, dl);
        bfs(pts, e2, t, dr);
        std::vector<int32_t> part[2], edge[2];
        for (int32_t p : pts) part[dl[p] < dr[p] ? 0 : 1].push_back(p);
        for (int32_t p : pts) {
            const int side = dl[p] < dr[p] ? 0 : 1;
            for (int64_t k = g.off[p]; k < g.off[(size_t)p + 1]; ++k) {
                const int32_t q = g.adj[k];
                if (tag[q] == t && (dl[q] < dr[q] ? 0 : 1) != side) {
                    edge[side].push_back(p);
                    break;
                }
            }
        }
        const int32_t t1 = ++next_tag, t2 = ++next_tag;
        for (int32_t p : part[0]) tag[p] = t1;
        for (int32_t p : part[1]) tag[p] = t2;
        if ((int64_t)part[0].size() > nmax && part[0].size() > e1.size())
            bisect(part[0], e1, edge[0], t1);
        else
            slices.push_back(part[0]);
        if ((int64_t)part[1].size() > nmax && part[1].size() > e2.size())
            bisect(part[1], edge[1], e2, t2);
        else
            slices.push_back(part[1]);
    }
};

bool ids_ok(int32_t n, int32_t cnt, const int32_t* ids) {
    for (int32_t i = 0; i < cnt; ++i)
        if (ids[i] < 0 || ids[i] >= n) return false;
    return true;
}

}  // namespace

extern "C" int qt_graph_slice(int32_t npts, int64_t nedge, const int32_t* edges, int32_t nleft, const int32_t* left,
                              int32_t nright, const int32_t* right, int32_t nmax, int32_t* slice_of,
                              int32_t* nslices) {
    if (npts <= 0 || nedge < 0 || (nedge > 0 && !edges) || nleft <= 0 || !left || nright < 0 ||
        (nright > 0 && !right) || nmax < 1 || !slice_of || !nslices)
        return QT_ERR_INVALID;
    if (!ids_ok(npts, nleft, left) || !ids_ok(npts, nright, right)) return QT_ERR_INVALID;
    for (int64_t k = 0; k < 2 * nedge; ++k)
        if (edges[k] < 0 || edges[k] >= npts) return QT_ERR_INVALID;
    try {
        const Graph g = build_graph(npts, nedge, edges);
        Slicer s(g, nmax);
        std::vector<int32_t> all(npts), e1(left, left + nleft), e2(right, right + nright);
        for (int32_t i = 0; i < npts; ++i) all[i] = i;
        // the graph must be connected: every point reachable from the left
        // contact (a bisection between unconnected contacts would not end)
        s.bfs(all, e1, 0, s.dl);
        for (int32_t i = 0; i < npts; ++i)
            if (s.dl[i] == kFar) return QT_ERR_INVALID;
        if (nright == 0) {
            int64_t depth = 0;
            for (int32_t i = 0; i < npts; ++i) depth = std::max(depth, s.dl[i]);
            for (int32_t i = 0; i < npts; ++i) slice_of[i] = (int32_t)s.dl[i];
            *nslices = (int32_t)(depth + 1);
            return QT_OK;
        }
        s.bisect(all, e1, e2, 0);
        for (size_t k = 0; k < s.slices.size(); ++k)
            for (int32_t p : s.slices[k]) slice_of[p] = (int32_t)k;
        *nslices = (int32_t)s.slices.size();
        return QT_OK;
    } catch (const std::bad_alloc&) {
        return QT_ERR_NOMEM;
    }
}
