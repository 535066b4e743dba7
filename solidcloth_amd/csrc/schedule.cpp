// Constraint schedules (see schedule.h).
#include "schedule.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <string>

namespace sc {

namespace {

// Greedy first-fit colouring of `ids` (in the given order); colours limited
// to 64 * kWords. `used` is indexed by vertex and must be zero on entry for
// every vertex the constraints touch; it is cleared again on exit.
struct Greedy {
    static const int kWords = 4;
    std::vector<uint64_t> used;
    explicit Greedy(int nv) : used((size_t)nv * kWords, 0) {}
    int colour(const std::vector<int>& ids, int arity, const uint32_t* verts, std::vector<int>& col)
    {
        int nc = 0;
        for (size_t k = 0; k < ids.size(); k++) {
            const uint32_t* v = verts + (size_t)ids[k] * arity;
            uint64_t forb[kWords] = {0, 0, 0, 0};
            for (int a = 0; a < arity; a++)
                for (int w = 0; w < kWords; w++) forb[w] |= used[(size_t)v[a] * kWords + w];
            int c = -1;
            for (int w = 0; w < kWords && c < 0; w++)
                if (~forb[w]) c = w * 64 + __builtin_ctzll(~forb[w]);
            if (c < 0) throw std::runtime_error("constraint colouring needs more than 256 colours");
            col[k] = c;
            for (int a = 0; a < arity; a++) used[(size_t)v[a] * kWords + c / 64] |= 1ull << (c % 64);
            nc = std::max(nc, c + 1);
        }
        for (size_t k = 0; k < ids.size(); k++) {
            const uint32_t* v = verts + (size_t)ids[k] * arity;
            for (int a = 0; a < arity; a++)
                for (int w = 0; w < kWords; w++) used[(size_t)v[a] * kWords + w] = 0;
        }
        return nc;
    }
};

}  // namespace

Schedule schedule_colours(int n, int arity, const uint32_t* verts, int nv)
{
    Schedule s;
    s.arity = arity;
    s.slabbed = false;
    Greedy g(nv);
    std::vector<int> ids(n), col(n);
    std::iota(ids.begin(), ids.end(), 0);
    int nc = n ? g.colour(ids, arity, verts, col) : 0;
    std::vector<int> cnt(nc + 1, 0);
    for (int c : col) cnt[c + 1]++;
    for (int k = 0; k < nc; k++) cnt[k + 1] += cnt[k];
    s.phaseOffsets = cnt;
    if (s.phaseOffsets.empty()) s.phaseOffsets.push_back(0);
    s.order.resize(n);
    s.phaseOf.resize(n);
    s.slabOf.assign(n, -1);
    s.colourOf.resize(n);
    std::vector<int> fill(cnt.begin(), cnt.end() - (nc ? 1 : 0));
    for (int k = 0; k < n; k++) {
        int p = fill[col[k]]++;
        s.order[p] = k;
        s.phaseOf[p] = col[k];
        s.colourOf[p] = 0;
    }
    return s;
}

namespace {

int longest_axis(const f3* rest, int nv)
{
    f3 lo = rest[0], hi = rest[0];
    for (int i = 1; i < nv; i++) {
        lo = f3{std::min(lo.x, rest[i].x), std::min(lo.y, rest[i].y), std::min(lo.z, rest[i].z)};
        hi = f3{std::max(hi.x, rest[i].x), std::max(hi.y, rest[i].y), std::max(hi.z, rest[i].z)};
    }
    const float ext[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
    return (ext[1] > ext[0] && ext[1] >= ext[2]) ? 1 : (ext[2] > ext[0] && ext[2] > ext[1]) ? 2 : 0;
}

float axis_coord(const f3& p, int axis) { return axis == 0 ? p.x : axis == 1 ? p.y : p.z; }

// Strips: constraints sorted by centroid along the axis and cut every
// `slabSize` constraints.
std::vector<std::vector<int>> strip_slabs(int n, int arity, const uint32_t* verts, const f3* rest, int axis,
                                          int slabSize)
{
    // sort by centroid along the axis (sum of coordinates; ties by index)
    std::vector<float> key(n);
    for (int c = 0; c < n; c++) {
        float k = 0.f;
        for (int a = 0; a < arity; a++) k += axis_coord(rest[verts[(size_t)c * arity + a]], axis);
        key[c] = k;
    }
    std::vector<int> byKey(n);
    std::iota(byKey.begin(), byKey.end(), 0);
    std::stable_sort(byKey.begin(), byKey.end(), [&](int a, int b) { return key[a] < key[b]; });
    const int numSlabs = (n + slabSize - 1) / slabSize;
    std::vector<std::vector<int>> slabs(numSlabs);
    for (int k = 0; k < n; k++) slabs[k / slabSize].push_back(byKey[k]);
    return slabs;
}

// Interiors and seams: the vertices are sorted along the axis (ties by index)
// and cut into strips of about slabSize * nv / n vertices, and at least the
// widest constraint's span in that order, so that no constraint spans three
// strips. The constraints whose vertices all lie in strip i form interior
// slab 2i; those spanning
// strips i and i+1 form seam slab 2i+1. Interiors of different strips share
// no vertex, so they all run in one phase, and the seams (narrow bands
// around the cuts) in the next: two phases, of which only the interior one
// has the full local colour count.
std::vector<std::vector<int>> seamed_slabs(int n, int arity, const uint32_t* verts, int nv, const f3* rest,
                                           int axis, int slabSize)
{
    std::vector<int> byCoord(nv);
    std::iota(byCoord.begin(), byCoord.end(), 0);
    std::stable_sort(byCoord.begin(), byCoord.end(),
                     [&](int a, int b) { return axis_coord(rest[a], axis) < axis_coord(rest[b], axis); });
    std::vector<int> rank(nv);
    for (int k = 0; k < nv; k++) rank[byCoord[k]] = k;
    // a strip at least as wide (in sorted vertices) as any constraint, so that
    // no constraint spans three strips
    long widest = 1;
    for (int c = 0; c < n; c++) {
        int lo = nv, hi = -1;
        for (int a = 0; a < arity; a++) {
            const int r = rank[verts[(size_t)c * arity + a]];
            lo = std::min(lo, r);
            hi = std::max(hi, r);
        }
        widest = std::max(widest, (long)(hi - lo));
    }
    const long stripVerts = std::max(widest, (long)slabSize * nv / std::max(1, n));
    // strips of stripVerts vertices in axis order; with `snap`, each cut is
    // moved back to the first vertex of its coordinate, so that vertices of
    // one coordinate (a grid's row) share a strip
    std::vector<int> strip(nv);
    int numStrips = 0;
    auto cut_strips = [&](bool snap) {
        numStrips = 0;
        for (long begin = 0; begin < nv;) {
            long end = std::min<long>(begin + stripVerts, nv);
            if (snap && end < nv) {
                long cut = end;
                while (cut > begin && axis_coord(rest[byCoord[cut - 1]], axis) == axis_coord(rest[byCoord[cut]], axis))
                    cut--;
                if (cut > begin) end = cut;
            }
            for (long k = begin; k < end; k++) strip[byCoord[k]] = numStrips;
            numStrips++;
            begin = end;
        }
    };
    auto spans_three = [&]() {
        for (int c = 0; c < n; c++) {
            int lo = 1 << 30, hi = -1;
            for (int a = 0; a < arity; a++) {
                const int st = strip[verts[(size_t)c * arity + a]];
                lo = std::min(lo, st);
                hi = std::max(hi, st);
            }
            if (hi - lo > 1) return true;
        }
        return false;
    };
    // (snapped strips can be narrower than the widest constraint; then the
    // plain cuts are kept)
    cut_strips(true);
    if (spans_three()) cut_strips(false);
    std::vector<std::vector<int>> slabs(2 * (size_t)numStrips);
    for (int c = 0; c < n; c++) {
        int lo = 1 << 30, hi = -1;
        for (int a = 0; a < arity; a++) {
            const int st = strip[verts[(size_t)c * arity + a]];
            lo = std::min(lo, st);
            hi = std::max(hi, st);
        }
        if (hi - lo > 1) throw std::logic_error("a constraint spans more than two strips");
        slabs[2 * (size_t)lo + (hi > lo)].push_back(c);
    }
    // interiors first, so that the phase colouring gives them phase 0
    std::vector<std::vector<int>> out;
    for (int pass = 0; pass < 2; pass++)
        for (size_t k = pass; k < slabs.size(); k += 2)
            if (!slabs[k].empty()) out.push_back(std::move(slabs[k]));
    return out;
}

}  // namespace

Schedule schedule_slabs(int n, int arity, const uint32_t* verts, int nv, const f3* rest, int slabSize, int maxVerts,
                        bool seamed)
{
    Schedule s;
    s.arity = arity;
    s.slabbed = true;
    s.slabSize = slabSize;
    s.seamed = seamed;
    if (n == 0) {
        s.phaseOffsets = {0};
        s.phaseSlabBegin = {0};
        s.slabVertBegin = {0};
        s.slabColBegin = {0};
        s.colOffsets = {0};
        return s;
    }
    const int axis = longest_axis(rest, nv);
    std::vector<std::vector<int>> slabs = seamed ? seamed_slabs(n, arity, verts, nv, rest, axis, slabSize)
                                                 : strip_slabs(n, arity, verts, rest, axis, slabSize);
    const int numSlabs = (int)slabs.size();
    for (auto& sl : slabs) std::sort(sl.begin(), sl.end());
    // slab conflict graph through shared vertices, greedy-coloured in slab order
    std::vector<std::vector<int>> vslabs(nv);
    for (int sI = 0; sI < numSlabs; sI++)
        for (int c : slabs[sI])
            for (int a = 0; a < arity; a++) {
                auto& l = vslabs[verts[(size_t)c * arity + a]];
                if (l.empty() || l.back() != sI) l.push_back(sI);
            }
    std::vector<std::vector<int>> nbr(numSlabs);
    for (auto& l : vslabs)
        for (size_t i = 0; i < l.size(); i++)
            for (size_t j = 0; j < l.size(); j++)
                if (i != j) nbr[l[i]].push_back(l[j]);
    std::vector<int> phase(numSlabs, -1);
    int numPhases = 0;
    for (int sI = 0; sI < numSlabs; sI++) {
        std::vector<char> taken(nbr[sI].size() + 1, 0);
        for (int o : nbr[sI])
            if (phase[o] >= 0 && phase[o] < (int)taken.size()) taken[phase[o]] = 1;
        int p = 0;
        while (taken[p]) p++;
        phase[sI] = p;
        numPhases = std::max(numPhases, p + 1);
    }
    // renumber slabs phase-major (stable in slab order)
    std::vector<int> slabOrder(numSlabs);
    std::iota(slabOrder.begin(), slabOrder.end(), 0);
    std::stable_sort(slabOrder.begin(), slabOrder.end(), [&](int a, int b) { return phase[a] < phase[b]; });
    s.phaseSlabBegin.assign(numPhases + 1, 0);
    for (int sI = 0; sI < numSlabs; sI++) s.phaseSlabBegin[phase[sI] + 1]++;
    for (int p = 0; p < numPhases; p++) s.phaseSlabBegin[p + 1] += s.phaseSlabBegin[p];

    Greedy g(nv);
    std::vector<int> localOf(nv, -1);
    s.order.reserve(n);
    s.localIdx.reserve((size_t)n * arity);
    s.slabVertBegin.push_back(0);
    s.slabColBegin.push_back(0);
    s.phaseOffsets.assign(1, 0);
    int curPhase = 0;
    for (int r = 0; r < numSlabs; r++) {
        const int sI = slabOrder[r];
        while (phase[sI] > curPhase) {
            s.phaseOffsets.push_back((int)s.order.size());
            curPhase++;
        }
        const std::vector<int>& ids = slabs[sI];
        // vertices of the slab, ascending
        std::vector<uint32_t> vs;
        for (int c : ids)
            for (int a = 0; a < arity; a++) vs.push_back(verts[(size_t)c * arity + a]);
        std::sort(vs.begin(), vs.end());
        vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
        if ((int)vs.size() > maxVerts)
            throw std::runtime_error("slab has " + std::to_string(vs.size()) + " vertices, more than " +
                                     std::to_string(maxVerts));
        for (size_t i = 0; i < vs.size(); i++) localOf[vs[i]] = (int)i;
        s.maxSlabVerts = std::max(s.maxSlabVerts, (int)vs.size());
        // local colours
        std::vector<int> col(ids.size());
        int nc = g.colour(ids, arity, verts, col);
        if (nc > 63) throw std::runtime_error("a slab needs more than 63 colours");
        s.maxSlabColours = std::max(s.maxSlabColours, nc);
        std::vector<int> cnt(nc + 1, 0);
        for (int c : col) cnt[c + 1]++;
        for (int k = 0; k < nc; k++) cnt[k + 1] += cnt[k];
        const int base = (int)s.order.size();
        std::vector<int> pos(ids.size());
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (size_t k = 0; k < ids.size(); k++) pos[fill[col[k]]++] = (int)k;
        for (int k = 0; k < nc; k++) s.colOffsets.push_back(base + cnt[k]);
        for (size_t q = 0; q < ids.size(); q++) {
            const int k = pos[q];
            const int c = ids[k];
            s.order.push_back(c);
            s.phaseOf.push_back(phase[sI]);
            s.slabOf.push_back(r);
            s.colourOf.push_back(col[k]);
            for (int a = 0; a < arity; a++) s.localIdx.push_back((uint16_t)localOf[verts[(size_t)c * arity + a]]);
        }
        for (uint32_t v : vs) { s.slabVerts.push_back(v); localOf[v] = -1; }
        s.slabVertBegin.push_back((int)s.slabVerts.size());
        s.slabColBegin.push_back((int)s.colOffsets.size());
    }
    s.colOffsets.push_back((int)s.order.size());
    s.phaseOffsets.push_back((int)s.order.size());
    return s;
}

Schedule schedule_tiles(int n, int arity, const uint32_t* verts, int nv, const f3* rest, int tileVerts,
                        int maxVerts)
{
    Schedule s;
    s.arity = arity;
    s.slabbed = true;
    s.tiled = true;
    // one global colouring; the sequential order is colour-major
    Greedy gr(nv);
    std::vector<int> ids(n), col(n);
    std::iota(ids.begin(), ids.end(), 0);
    const int nc = n ? gr.colour(ids, arity, verts, col) : 0;
    std::vector<int> cstart(nc + 1, 0);
    for (int c : col) cstart[c + 1]++;
    for (int k = 0; k < nc; k++) cstart[k + 1] += cstart[k];
    s.order.resize(n);
    s.phaseOf.assign(n, 0);
    s.slabOf.assign(n, -1);
    s.colourOf.resize(n);
    {
        std::vector<int> fill(cstart.begin(), cstart.end() - (nc ? 1 : 0));
        for (int k = 0; k < n; k++) {
            const int p = fill[col[k]]++;
            s.order[p] = k;
            s.colourOf[p] = col[k];
        }
    }
    s.phaseOffsets = {0, n};
    s.halo = nc;
    // vertex -> constraints and vertex -> neighbours (sharing a constraint)
    std::vector<int> vcOff(nv + 1, 0), vc((size_t)n * arity);
    for (int k = 0; k < n * arity; k++) vcOff[verts[k] + 1]++;
    for (int v = 0; v < nv; v++) vcOff[v + 1] += vcOff[v];
    {
        std::vector<int> fill(vcOff.begin(), vcOff.end() - 1);
        for (int c = 0; c < n; c++)
            for (int a = 0; a < arity; a++) vc[fill[verts[(size_t)c * arity + a]]++] = c;
    }
    // tiles: a grid over the rest shape's two widest axes
    int ax0 = 0, ax1 = 2;
    f3 lo{0, 0, 0}, hi{0, 0, 0};
    if (nv > 0) {
        lo = hi = rest[0];
        for (int i = 1; i < nv; i++) {
            lo = f3{std::min(lo.x, rest[i].x), std::min(lo.y, rest[i].y), std::min(lo.z, rest[i].z)};
            hi = f3{std::max(hi.x, rest[i].x), std::max(hi.y, rest[i].y), std::max(hi.z, rest[i].z)};
        }
        const float ext[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
        const int narrow = (ext[0] <= ext[1] && ext[0] <= ext[2]) ? 0 : (ext[1] <= ext[2] ? 1 : 2);
        ax0 = narrow == 0 ? 1 : 0;
        ax1 = narrow == 2 ? 1 : 2;
    }
    const float e0 = axis_coord(hi, ax0) - axis_coord(lo, ax0), e1 = axis_coord(hi, ax1) - axis_coord(lo, ax1);
    const double bins = std::max(1.0, (double)nv / std::max(1, tileVerts));
    int nb0 = 1, nb1 = 1;
    if (e0 > 0.f && e1 > 0.f) {
        nb0 = std::max(1, (int)std::lround(std::sqrt(bins * e0 / e1)));
        nb1 = std::max(1, (int)std::lround(bins / nb0));
    } else if (e0 > 0.f) {
        nb0 = std::max(1, (int)std::lround(bins));
    } else if (e1 > 0.f) {
        nb1 = std::max(1, (int)std::lround(bins));
    }
    auto bin = [&](float x, float l, float e, int nb) {
        if (!(e > 0.f)) return 0;
        return std::min(nb - 1, std::max(0, (int)((x - l) / e * (float)nb)));
    };
    std::vector<std::vector<int>> owned((size_t)nb0 * nb1);
    for (int v = 0; v < nv; v++) {
        const int b0 = bin(axis_coord(rest[v], ax0), axis_coord(lo, ax0), e0, nb0);
        const int b1 = bin(axis_coord(rest[v], ax1), axis_coord(lo, ax1), e1, nb1);
        owned[(size_t)b1 * nb0 + b0].push_back(v);
    }
    // per tile: the halo rings, the constraints inside, local indices
    std::vector<int> stamp(nv, -1), local(nv, 0), cstamp(n, -1);
    s.phaseSlabBegin = {0};
    s.slabVertBegin = {0};
    s.slabColBegin = {0};
    s.colOffsets.clear();
    int tile = 0;
    for (const std::vector<int>& own : owned) {
        if (own.empty()) continue;
        std::vector<int> ext(own.begin(), own.end());
        for (int v : own) stamp[v] = tile;
        size_t frontBegin = 0;
        for (int ring = 0; ring < s.halo; ring++) {
            const size_t frontEnd = ext.size();
            for (size_t i = frontBegin; i < frontEnd; i++) {
                const int v = ext[i];
                for (int k = vcOff[v]; k < vcOff[v + 1]; k++) {
                    const uint32_t* cv = verts + (size_t)vc[k] * arity;
                    for (int a = 0; a < arity; a++) {
                        const int u = (int)cv[a];
                        if (stamp[u] != tile) {
                            stamp[u] = tile;
                            ext.push_back(u);
                        }
                    }
                }
            }
            frontBegin = frontEnd;
        }
        if ((int)ext.size() > maxVerts)
            throw std::runtime_error("an extended tile has " + std::to_string(ext.size()) + " vertices (at most " +
                                     std::to_string(maxVerts) + ")");
        std::sort(ext.begin(), ext.end());
        std::vector<char> isOwn(ext.size(), 0);
        for (size_t i = 0; i < ext.size(); i++) local[ext[i]] = (int)i;
        for (int v : own) isOwn[local[v]] = 1;
        // constraints with every vertex in the extended region, by colour
        std::vector<std::vector<int>> byCol(nc);
        for (int v : ext)
            for (int k = vcOff[v]; k < vcOff[v + 1]; k++) {
                const int c = vc[k];
                if (cstamp[c] == tile) continue;
                cstamp[c] = tile;
                bool inside = true;
                for (int a = 0; a < arity && inside; a++) inside = stamp[verts[(size_t)c * arity + a]] == tile;
                if (inside) byCol[col[c]].push_back(c);
            }
        for (int c = 0; c < nc; c++) {
            std::sort(byCol[c].begin(), byCol[c].end());
            s.colOffsets.push_back((int)s.instCons.size());
            for (int k : byCol[c]) {
                s.instCons.push_back(k);
                for (int a = 0; a < arity; a++) s.localIdx.push_back((uint16_t)local[verts[(size_t)k * arity + a]]);
            }
        }
        for (size_t i = 0; i < ext.size(); i++) {
            s.slabVerts.push_back((uint32_t)ext[i]);
            s.slabOwned.push_back(isOwn[i]);
        }
        s.slabVertBegin.push_back((int)s.slabVerts.size());
        s.slabColBegin.push_back((int)s.colOffsets.size());
        s.maxSlabVerts = std::max(s.maxSlabVerts, (int)ext.size());
        tile++;
    }
    s.colOffsets.push_back((int)s.instCons.size());
    s.phaseSlabBegin.push_back(tile);
    s.maxSlabColours = nc;
    if (nc > kMaxTileColours)
        throw std::runtime_error("tile schedule: " + std::to_string(nc) + " colours (at most " +
                                 std::to_string(kMaxTileColours) + ")");
    return s;
}

int schedule_cost(const Schedule& s, int threads)
{
    // In tenths of a sweep step. A step (one local colour of the phase's
    // slowest slab) costs about one projection's latency whatever its width,
    // plus one more per extra constraint a thread takes when the colour is
    // wider than the workgroup; a bend colour that needs more than half the
    // workgroup cannot split each bend over a pair of lanes and costs about
    // 1.6 steps (measured on the 256^2 drape: 15 such colours in one launch
    // took longer than 5 + 9 + 9 paired ones in three). Staging and writing
    // back a phase costs about three steps.
    int cost = 0;
    for (int p = 0; p < s.numPhases(); p++) {
        int worst = 0;
        for (int sl = s.phaseSlabBegin[p]; sl < s.phaseSlabBegin[p + 1]; sl++) {
            int steps = 0;
            for (int c = s.slabColBegin[sl]; c < s.slabColBegin[sl + 1]; c++) {
                const int width = s.colOffsets[c + 1] - s.colOffsets[c];
                const int step = s.arity == 4 && width > threads / 2 ? 16 : 10;
                steps += step * ((width + threads - 1) / threads);
            }
            worst = std::max(worst, steps);
        }
        cost += worst + 30;
    }
    return cost;
}

Schedule schedule_slabs_auto(int n, int arity, const uint32_t* verts, int nv, const f3* rest, int maxVerts,
                             int threads)
{
    // Strips of 1024 or 512 constraints, or interiors and seams at several
    // interior sizes: the candidate with the least sweep cost is kept (the
    // first on ties). Thin strips have few local colours but need two or
    // three phases; interiors and seams need two phases, only one of them
    // with the full colour count. Smaller strips are the fallback when a
    // slab's vertices would not fit the LDS budget.
    struct Cand {
        int size;
        bool seamed;
    };
    const Cand cands[] = {{1024, true}, {2048, true}, {3072, true}, {1024, false}, {512, false}, {256, false}};
    Schedule best;
    int bestCost = -1;
    for (const Cand& c : cands) {
        try {
            Schedule s = schedule_slabs(n, arity, verts, nv, rest, c.size, maxVerts, c.seamed);
            const int cost = schedule_cost(s, threads);
            if (bestCost < 0 || cost < bestCost) {
                bestCost = cost;
                best = std::move(s);
            }
        } catch (const std::runtime_error&) {
        }
    }
    if (bestCost < 0) throw std::runtime_error("no slab size fits the LDS budget");
    return best;
}

}  // namespace sc
