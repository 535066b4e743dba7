// Gauss-Seidel schedules for a set of constraints (FEM triangles or bends).
//
// Two schedules are built, both deterministic and both equal in effect to
// processing the constraints one after another in the exported order:
//
//  * colour schedule: greedy first-fit colouring in index order; one launch
//    per colour, one constraint per thread, positions in HBM.
//
//  * slab schedule: the constraints are grouped into slabs, either strips
//    (sorted along the longest axis of the rest shape and cut every
//    `slabSize` constraints) or strip interiors and the seams between strips;
//    slabs sharing a vertex are coloured apart (a chain of strips needs two
//    phases, interiors one and seams one); inside a slab the constraints are
//    greedily coloured again. One launch per phase,
//    one workgroup per slab: the workgroup stages the slab's vertices in LDS,
//    sweeps its local colours with workgroup barriers and writes the vertices
//    back. Slabs of one phase touch disjoint vertices.
//
//  * tile schedule (FEM triangles): one global greedy colouring; the
//    vertices are cut into tiles, each extended by a halo of as many rings of
//    neighbours as there are colours, and each tile sweeps every colour over
//    the constraints inside its extended region. A stale value at the
//    region's border moves inwards by at most one ring per colour step, so
//    the tile's own vertices end exactly as the sequential colour-major
//    sweep leaves them; halos are computed redundantly and only owned
//    vertices are written back. One phase, one launch, and the fewest
//    sequential colour steps (the global colour count).
//
// The exported order is phase-major, then slab, then local colour, then
// ascending constraint index; for tiles, colour-major then ascending index.
#pragma once

#include <cstdint>
#include <vector>

#include "scmath.h"

namespace sc {

// Largest number of colours of a tile schedule (the slab kernels' limit).
constexpr int kMaxTileColours = 63;

struct Schedule {
    int arity = 0;
    bool slabbed = false;
    std::vector<int> order;        // constraint indices in Gauss-Seidel order
    std::vector<int> phaseOffsets;  // positions in `order` where each phase starts (+ end)
    // per order position: phase, slab (global slab id, -1 in colour mode), local colour
    std::vector<int> phaseOf, slabOf, colourOf;
    // slab mode
    std::vector<int> phaseSlabBegin;   // slabs of phase p: [phaseSlabBegin[p], phaseSlabBegin[p+1])
    std::vector<int> slabVertBegin;    // into slabVerts (+ end)
    std::vector<int> slabColBegin;     // into colOffsets (+ end): colours of slab s
    std::vector<int> colOffsets;       // positions in `order`; colour k of slab s spans [colOffsets[i], colOffsets[i+1])
    std::vector<uint32_t> slabVerts;   // global vertex ids per slab, ascending
    std::vector<uint16_t> localIdx;    // per order position, `arity` local vertex indices
    int maxSlabVerts = 0;
    int maxSlabColours = 0;
    int slabSize = 0;
    bool seamed = false;
    // tile mode: slabs are tiles whose constraint lists are instances that
    // may repeat a constraint in several tiles (instCons: the constraint of
    // each instance; colOffsets and localIdx index instances, not `order`),
    // and slabOwned marks each tile's own vertices in slabVerts
    bool tiled = false;
    int halo = 0;
    std::vector<int> instCons;
    std::vector<uint8_t> slabOwned;
    int numPhases() const { return (int)phaseOffsets.size() - 1; }
};

Schedule schedule_colours(int numConstraints, int arity, const uint32_t* verts, int numVertices);

// Returns a slab schedule with at most `maxVerts` vertices per slab; throws if
// the constraints cannot be split that way. Without `seamed`, slabs are strips
// of `slabSize` constraints along the rest shape's longest axis; with it, the
// vertices are cut into strips and each strip's interior constraints form a
// slab (all in one phase) and the constraints across each cut a seam slab.
Schedule schedule_slabs(int numConstraints, int arity, const uint32_t* verts, int numVertices, const f3* rest,
                        int slabSize, int maxVerts, bool seamed = false);

// Tile schedule with tiles of about `tileVerts` own vertices; throws if an
// extended tile would exceed `maxVerts` vertices.
Schedule schedule_tiles(int numConstraints, int arity, const uint32_t* verts, int numVertices, const f3* rest,
                        int tileVerts, int maxVerts);

// Estimated sweep cost of a slab schedule in projection latencies, for a
// workgroup of `threads` threads.
int schedule_cost(const Schedule& s, int threads);

// The least-cost slab schedule among strips and interior/seam splits of
// several sizes that fit maxVerts.
Schedule schedule_slabs_auto(int numConstraints, int arity, const uint32_t* verts, int numVertices,
                             const f3* rest, int maxVerts, int threads);

}  // namespace sc
