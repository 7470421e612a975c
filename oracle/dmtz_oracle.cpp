// DMTz topology-correction oracle: a CPU restatement of the reference path in main_3d.cu.
//
// TEST INFRASTRUCTURE ONLY. Only the test suite, the smoke check and the CPU-baseline leg of the
// benchmark may load this library, and only as the checker. The shipped HIP path never links it.
//
// Structure follows the reference literally: dense per-simplex pairing arrays keyed by the
// reference's global simplex IDs, five gradient passes, XOR mismatch lists, the repair rules,
// the edit buffer and the correction loop. Every function cites the main_3d.cu lines it restates.
// Two gradient implementations exist: LITERAL (the reference's 14^k brute-force neighbour
// searches) and DIRECT (the same candidate sets enumerated through an adjacency table); tests
// check that both produce identical arrays.
//
// Pins (tests/test_oracle.py): closed-form simplex counts, ID bijections and the clique structure
// of the implicit complex, cofacet counts, LITERAL == DIRECT gradients, Euler characteristic and
// matching validity of every gradient, linear fields, invariance under monotone value maps,
// float/double agreement, end-state invariants of the loop, the float32 snap, the stall fixed
// point and the finisher's boxes and tie repair.
// The repair rules, the separatrix checks and the path and wall repairs are further pinned in
// tests/: branch-hit counters (orc_coverage) show that the GPU parity corpus reaches every
// reachable branch (test_oracle_coverage.py), hand-worked known answers fix one mismatched cell
// per top-level branch of the four critical-cell rules (test_oracle_kats.py), and an independent
// serial restatement of the separatrix checks and labels must agree (test_oracle_paths.py).
// Parity unpinned: which edge the deterministic wall walk records first when several qualify
// (walls are timing-dependent in the reference itself).
//
// Field type T is float or double. With T = double all value arithmetic is the reference's.
// With T = float the values are float32, comparisons are exact float32 ordering with index
// tie-break, and edit arithmetic is done in double then rounded to float (see Arith<float>).
//
// Resolutions of reference defects (all documented, none change defined behaviour):
//   * out-of-range pairing-array reads (index -1 or beyond the array) read as -1 ("unpaired");
//   * vertex indices outside [0, N) produced by decoding a wrong-type ID are never edited;
//   * descending/ascending path walks are capped at N+1 steps (the reference has no cap);
//   * ascending connectivity labels are indexed by face id (main_3d.cu:4750/4823 shadowing bug);
//   * walls use a private per-start DFS with a separate index array (main_3d.cu:6706, 5088);
//   * the correction loop stops after max_sweeps sweeps and reports non-convergence;
//   * the correction loop stops when it has reached a fixed point with mismatches left (a sweep
//     that changes no value or counter and repeats an earlier control configuration), where the
//     reference loops forever, and reports it as stalled.
// Extension without a counterpart in the reference (flag finish, off by default): on the k-th
// stall every vertex within Chebyshev distance 2^(k-1) of the zone (the edit sets of the stalled
// sweeps and the vertices of the mismatched critical cells) is snapped (the reference's terminal edit,
// main_3d.cu:1140-1141; a snap of the 14 neighbours sits commented out at main_3d.cu:1144-1151),
// ties the float snap creates are repaired by one-ulp raises, and the box is frozen (counter 7,
// never edited again). Once the boxes would cover the grid, or when the zone is empty, g = f is
// restored everywhere. The run then ends at zero mismatches.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cfloat>
#include <climits>
#include <limits>
#include <vector>
#include <algorithm>
#include <atomic>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

namespace orc {

// ----------------------------------------------------------------------------------------------
// Branch-hit counters of the restated repair rules and separatrix checks (test infrastructure:
// counting is off unless orc_coverage() switches it on). Each entry names the branch and the
// main_3d.cu lines it restates.
// ----------------------------------------------------------------------------------------------
enum CovId {
    C_MIN_F_PAIRED,
    C_MIN_F_CRITICAL,
    C_S1_FP_VERTEX,
    C_S1_FP_VERTEX_G_UNPAIRED,
    C_S1_FP_VERTEX_TARGET,
    C_S1_FP_FACE,
    C_S1_FP_FACE_A,
    C_S1_FP_FACE_B,
    C_S1_FP_FACE_C,
    C_S1_FP_FACE_C_FPAIR,
    C_S1_FP_FACE_C_DIFF,
    C_S1_FP_FACE_D,
    C_S1_FP_FACE_D1,
    C_S1_FP_FACE_D2,
    C_S1_FN_VERTEX,
    C_S1_FN_FACE,
    C_S1_FN_FACE_SAME,
    C_S1_FN_FACE_FVERTEX,
    C_S1_FN_FACE_SEARCH,
    C_S1_FN_FACE_BETWEEN,
    C_S1_FN_FACE_NO_BETWEEN,
    C_S2_FP_EDGE,
    C_S2_FP_EDGE_A,
    C_S2_FP_EDGE_A_TARGET,
    C_S2_FP_EDGE_B,
    C_S2_FP_EDGE_B_GUNPAIRED,
    C_S2_FP_EDGE_B_V2,
    C_S2_FP_EDGE_B_V1,
    C_S2_FP_EDGE_B_V3,
    C_S2_FP_EDGE_C,
    C_S2_FP_TET,
    C_S2_FP_TET_MIXED,
    C_S2_FP_TET_EDGE,
    C_S2_FP_TET_TET,
    C_S2_FP_TET_W,
    C_S2_FP_TET_GUNPAIRED,
    C_S2_FP_TET_ALL_ABOVE,
    C_S2_FN_EDGE,
    C_S2_FN_EDGE_FFACE,
    C_S2_FN_EDGE_FVERTEX,
    C_S2_FN_TET,
    C_S2_FN_TET_ALL,
    C_S2_FN_TET_TRUECELL,
    C_S2_FN_TET_SMALLEST,
    C_S2_FN_TET_NO_SMALLEST,
    C_MAX_FP,
    C_MAX_FP_ALL_BELOW,
    C_MAX_FP_GEDGE,
    C_MAX_FP_GTET,
    C_MAX_FN,
    C_MAX_FN_ALL,
    C_MAX_FN_ALL_FTET,
    C_MAX_FN_FEDGE,
    C_MAX_FN_FTET,
    C_MAX_FN_FCRIT,
    C_ASW_F_CRITICAL,
    C_ASW_DIVERGENT,
    C_ASW_NOT_PAIRED_BACK,
    C_ASW_BOUNDARY,
    C_DSG_RECORD,
    C_DSG_F_MIN,
    C_DSC_FIRST,
    C_DSC_NO_TROUBLE,
    C_DSC_SECOND,
    C_ASC_LABEL_DIFFERS,
    C_ASC_RECORD,
    C_FIXAS_FTET,
    C_FIXAS_FACE,
    C_FIXDS_NO_TROUBLE,
    C_FIXDS_FMIN,
    C_FIXDS_TARGET,
    C_WALL_V3,
    C_WALL_ENDS,
    C_WALL_F_UNPAIRED,
    C_WALL_G_UNPAIRED,
    C_WALLDET_FOUND,
    C_WALLDET_PUSH,
    C_NCOV
};
static const char* const COV_NAMES[C_NCOV] = {
    "MIN_F_PAIRED  :2560-2576 f pairs the vertex: target the paired edge's other end",
    "MIN_F_CRITICAL  :2577-2585 f-critical vertex: target it",
    "S1_FP_VERTEX  :2602-2641 false positive, f partner a vertex",
    "S1_FP_VERTEX_G_UNPAIRED  :2617 g leaves the vertex unpaired: no target",
    "S1_FP_VERTEX_TARGET  :2626-2640 other end above g partner: target it",
    "S1_FP_FACE  :2643-2827 false positive, f partner a face",
    "S1_FP_FACE_A  :2684-2699 both edge ends above the third vertex in g",
    "S1_FP_FACE_B  :2700-2708 both above it in f",
    "S1_FP_FACE_C  :2719-2800 g pairs the face with another edge",
    "S1_FP_FACE_C_FPAIR  :2768-2786 that edge is f-paired with a face: its third vertex",
    "S1_FP_FACE_C_DIFF  :2787-2797 nonzero delta: the edge's far vertex",
    "S1_FP_FACE_D  :2801-2823 g face otherwise paired",
    "S1_FP_FACE_D1  :2802-2812 first end",
    "S1_FP_FACE_D2  :2813-2823 second end",
    "S1_FN_VERTEX  :2839-2850 false negative, g partner a vertex",
    "S1_FN_FACE  :2852-3035 false negative, g partner a face",
    "S1_FN_FACE_SAME  :2891-2915 both ends on one side of the third vertex in g",
    "S1_FN_FACE_FVERTEX  :2919-2937 f pairs the edge with a vertex",
    "S1_FN_FACE_SEARCH  :2940-3003 f-critical edge: search a between vertex",
    "S1_FN_FACE_BETWEEN  :3004-3016 lowest vertex between the ends",
    "S1_FN_FACE_NO_BETWEEN  :3017-3029 none: both ends",
    "S2_FP_EDGE  :3057-3223 false positive, f partner an edge",
    "S2_FP_EDGE_A  :3097-3119 both above the third vertex in g",
    "S2_FP_EDGE_A_TARGET  :3097-3119 third vertex of f's face of the edge",
    "S2_FP_EDGE_B  :3120-3204 sides differ in f and in g",
    "S2_FP_EDGE_B_GUNPAIRED  :3120-3134 edge not face-paired in g: third vertex",
    "S2_FP_EDGE_B_V2  :3154-3169 g face of the edge: end v2",
    "S2_FP_EDGE_B_V1  :3170-3180 g face of the edge: end v1",
    "S2_FP_EDGE_B_V3  :3181-3203 otherwise the third vertex",
    "S2_FP_EDGE_C  :3205-3222 third vertex not below both ends in g",
    "S2_FP_TET  :3226-3477 false positive, f partner a tet",
    "S2_FP_TET_MIXED  :3284-3323 g face of the tet not below in f nor g",
    "S2_FP_TET_EDGE  :3288-3301 that face f-paired with an edge",
    "S2_FP_TET_TET  :3302-3322 that face f-paired with a tet",
    "S2_FP_TET_W  :3324-3337 the g face's vertices",
    "S2_FP_TET_GUNPAIRED  :3338-3350 tet critical in g: fourth vertex",
    "S2_FP_TET_ALL_ABOVE  :3351-3357 fourth vertex below the face in f: target it",
    "S2_FN_EDGE  :3488-3610 false negative, g partner an edge",
    "S2_FN_EDGE_FFACE  :3522-3543 f pairs the edge with a face: its third vertex",
    "S2_FN_EDGE_FVERTEX  :3544-3602 f pairs the edge with a vertex: both ends",
    "S2_FN_TET  :3613-3980 false negative, g partner a tet",
    "S2_FN_TET_ALL  :3714-3930 fourth vertex below the face in g",
    "S2_FN_TET_TRUECELL  :3937-3948 tet f-paired: the face's vertices",
    "S2_FN_TET_SMALLEST  :3949-3961 get_smallest_vertex (:2115-2194)",
    "S2_FN_TET_NO_SMALLEST  :3962-3975 none: the face's vertices",
    "MAX_FP  :3995-4097 false positive",
    "MAX_FP_ALL_BELOW  :4028-4039 face below the fourth vertex in f",
    "MAX_FP_GEDGE  :4049-4063 g pairs the face with an edge: its higher end",
    "MAX_FP_GTET  :4064-4095 g pairs the face with a tet: fourth vertex",
    "MAX_FN  :4098-4298 false negative",
    "MAX_FN_ALL  :4169-4224 face below the fourth vertex in g",
    "MAX_FN_ALL_FTET  :4202-4223 and in f: the f tet's far vertex",
    "MAX_FN_FEDGE  :4231-4243 f pairs the face with an edge: its ends",
    "MAX_FN_FTET  :4244-4258 f pairs the face with a tet: its vertices",
    "MAX_FN_FCRIT  :4259-4296 face f-critical: vertices below",
    "ASW_F_CRITICAL  is_false :4464-4467 walk reaches an f-critical tet",
    "ASW_DIVERGENT  is_false :4469-4479 f and g pairings differ: record",
    "ASW_NOT_PAIRED_BACK  is_false :4486-4489 face not paired back",
    "ASW_BOUNDARY  is_false :4491-4494 walk leaves through the boundary",
    "DSG_RECORD  get_false_ds_vpath :4879-4887 start vertex paired differently: record",
    "DSG_F_MIN  get_false_ds_vpath :4870-4873 endpoint is an f minimum",
    "DSC_FIRST  get_false_ds_vpath_connectivity :5005-5013 first end reaches another minimum",
    "DSC_NO_TROUBLE  get_false_ds_vpath_connectivity :5010 no differing vertex on the path (trouble_maker -1)",
    "DSC_SECOND  get_false_ds_vpath_connectivity :5050-5058 second end reaches another minimum",
    "ASC_LABEL_DIFFERS  get_false_as_vpath_connectivity :4822-4823 g reaches another maximum",
    "ASC_RECORD  get_false_as_vpath_connectivity :4825 and the f path diverges: record",
    "FIXAS_FTET  fix_as_vpath :5188-5217 g face f-paired with a tet: its far vertex",
    "FIXAS_FACE  fix_as_vpath :5218-5243 g face's vertex off the f face",
    "FIXDS_NO_TROUBLE  fix_ds_vpath :5274-5280 recorded vertex -1",
    "FIXDS_FMIN  fix_ds_vpath :5274-5282 recorded vertex is an f minimum",
    "FIXDS_TARGET  fix_ds_vpath :5283-5294 the f edge's other end",
    "WALL_V3  fix_ds_wall :5311-5338 both faces paired back, nonzero delta: third vertex",
    "WALL_ENDS  fix_ds_wall :5339-5353 zero delta: the edge's ends",
    "WALL_F_UNPAIRED  fix_ds_wall :5355-5392 edge not face-paired in f: its ends",
    "WALL_G_UNPAIRED  fix_ds_wall :5393-5411 edge not face-paired in g: f face's third vertex",
    "WALLDET_FOUND  get_false_ds_wall :5114-5121 an edge of the wall paired differently",
    "WALLDET_PUSH  get_false_ds_wall :5125 continue over the f-paired face"
};
static std::atomic<bool> cov_on{false};
static std::atomic<long long> cov_hits[C_NCOV];
#define COV(k)                                                                      \
    do {                                                                            \
        if (cov_on.load(std::memory_order_relaxed))                                  \
            cov_hits[k].fetch_add(1, std::memory_order_relaxed);                     \
    } while (0)

// main_3d.cu:44-51 — the 14-neighbour stencil of the Freudenthal lattice.
static const int DIRS[14][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},
                                {0, 0, -1}, {-1, 1, 0}, {1, -1, 0}, {0, 1, 1},  {0, -1, -1},
                                {-1, 0, 1}, {1, 0, -1}, {-1, 1, 1}, {1, -1, -1}};

// dir_of[(dx+1)*9+(dy+1)*3+(dz+1)] = direction index or -1.
static int dir_of[27];
static bool dir_init = false;
static void init_dirs() {
    if (dir_init) return;
    for (int i = 0; i < 27; i++) dir_of[i] = -1;
    for (int i = 0; i < 14; i++)
        dir_of[(DIRS[i][0] + 1) * 9 + (DIRS[i][1] + 1) * 3 + (DIRS[i][2] + 1)] = i;
    dir_init = true;
}

struct Grid {
    i64 W, H, D, N, E, F, T;
    Grid(i64 w, i64 h, i64 d) : W(w), H(h), D(d) {
        // main_3d.cu:6401-6415
        N = W * H * D;
        E = (W - 1) * H * D + W * (H - 1) * D + W * H * (D - 1) + (W - 1) * (H - 1) * D +
            (W - 1) * H * (D - 1) + W * (H - 1) * (D - 1) + (W - 1) * (H - 1) * (D - 1);
        F = 2 * ((W - 1) * (H - 1) * D + W * (H - 1) * (D - 1) + (W - 1) * H * (D - 1)) +
            6 * (W - 1) * (H - 1) * (D - 1);
        T = 6 * (W - 1) * (H - 1) * (D - 1);
    }
    inline void xyz(i64 v, i64& x, i64& y, i64& z) const {
        x = v % W;
        y = (v / W) % H;
        z = (v / (W * H)) % D;
    }
    inline bool inb(i64 x, i64 y, i64 z) const {
        return x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D;
    }
    inline i64 idx(i64 x, i64 y, i64 z) const { return x + y * W + z * W * H; }
    // Are two in-grid vertices 14-neighbours?
    inline bool adjacent(i64 a, i64 b) const {
        i64 ax, ay, az, bx, by, bz;
        xyz(a, ax, ay, az);
        xyz(b, bx, by, bz);
        i64 dx = bx - ax, dy = by - ay, dz = bz - az;
        if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return false;
        return dir_of[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)] >= 0;
    }
};

// ----------------------------------------------------------------------------------------------
// Implicit complex indexing (main_3d.cu:246-740, 917-1012, 1760-1900, 4363-4430)
// ----------------------------------------------------------------------------------------------

// getEdgeID main_3d.cu:917-1012
static i64 edge_id(const Grid& g, i64 x, i64 y, i64 z, i64 x1, i64 y1, i64 z1) {
    const i64 W = g.W, H = g.H, D = g.D;
    i64 dx = x1 - x, dy = y1 - y, dz = z1 - z;
    const i64 o3 = (W - 1) * H * D + W * (H - 1) * D + W * H * (D - 1);
    const i64 o4 = o3 + (W - 1) * (H - 1) * D;
    const i64 o5 = o4 + W * (H - 1) * (D - 1);
    const i64 o6 = o5 + (W - 1) * H * (D - 1);
    if (dx == 1 && dy == 0 && dz == 0) return z * (W - 1) * H + y * (W - 1) + x;
    if (dx == -1 && dy == 0 && dz == 0) return z * (W - 1) * H + y * (W - 1) + (x - 1);
    if (dx == 0 && dy == 1 && dz == 0) return (W - 1) * H * D + z * W * (H - 1) + x * (H - 1) + y;
    if (dx == 0 && dy == -1 && dz == 0)
        return (W - 1) * H * D + z * W * (H - 1) + x * (H - 1) + y - 1;
    if (dx == 0 && dy == 0 && dz == 1)
        return (W - 1) * H * D + W * (H - 1) * D + y * W * (D - 1) + x * (D - 1) + z;
    if (dx == 0 && dy == 0 && dz == -1)
        return (W - 1) * H * D + W * (H - 1) * D + y * W * (D - 1) + x * (D - 1) + z - 1;
    if (dx == 1 && dy == -1 && dz == 0) return o3 + z * (H - 1) * (W - 1) + (y - 1) * (W - 1) + x;
    if (dx == -1 && dy == 1 && dz == 0) return o3 + z * (H - 1) * (W - 1) + y * (W - 1) + (x - 1);
    if (dx == 0 && dy == 1 && dz == 1) return o4 + x * (H - 1) * (D - 1) + y * (D - 1) + z;
    if (dx == 0 && dy == -1 && dz == -1)
        return o4 + x * (H - 1) * (D - 1) + (y - 1) * (D - 1) + z - 1;
    if (dx == -1 && dy == 0 && dz == 1) return o5 + y * (W - 1) * (D - 1) + z * (W - 1) + x - 1;
    if (dx == 1 && dy == 0 && dz == -1) return o5 + y * (W - 1) * (D - 1) + (z - 1) * (W - 1) + x;
    if (dx != 0 && dy != 0 && dz != 0) {
        i64 cx = std::min(x, x1), cy = std::min(y, y1), cz = std::min(z, z1);
        return o6 + cz * (W - 1) * (H - 1) + cy * (W - 1) + cx;
    }
    return -1;
}

// getVertexIDsFromEdgeID main_3d.cu:1760-1900; returns id1 > id2 (index order).
static void edge_verts(const Grid& g, i64 e, i64& id1, i64& id2) {
    const i64 W = g.W, H = g.H, D = g.D;
    i64 counts[7] = {(W - 1) * H * D,       W * (H - 1) * D,       W * H * (D - 1),
                     (W - 1) * (H - 1) * D, W * (H - 1) * (D - 1), (W - 1) * H * (D - 1),
                     (W - 1) * (H - 1) * (D - 1)};
    i64 acc = 0;
    int dir = -1;
    for (int i = 0; i < 7; ++i) {
        if (e < acc + counts[i]) {
            dir = i;
            break;
        }
        acc += counts[i];
    }
    if (dir == -1) {
        id1 = id2 = -1;
        return;
    }
    i64 r = e - acc;
    i64 x = 0, y = 0, z = 0, x1 = 0, y1 = 0, z1 = 0;
    // C++ integer division truncates toward zero, as in the reference, for the r < 0 case (e<0).
    if (dir == 0) {
        z = r / ((W - 1) * H);
        y = (r % ((W - 1) * H)) / (W - 1);
        x = r % (W - 1);
        x1 = x + 1; y1 = y; z1 = z;
    } else if (dir == 1) {
        z = r / (W * (H - 1));
        x = (r % (W * (H - 1))) / (H - 1);
        y = r % (H - 1);
        y1 = y + 1; x1 = x; z1 = z;
    } else if (dir == 2) {
        y = r / (W * (D - 1));
        x = (r % (W * (D - 1))) / (D - 1);
        z = r % (D - 1);
        y1 = y; x1 = x; z1 = z + 1;
    } else if (dir == 3) {
        i64 z2 = r / ((W - 1) * (H - 1));
        i64 y2 = (r % ((W - 1) * (H - 1))) / (W - 1);
        i64 x2 = r % (W - 1);
        x = x2 + 1; y = y2; z = z2; y1 = y2 + 1; x1 = x2; z1 = z2;
    } else if (dir == 4) {
        i64 x2 = r / ((H - 1) * (D - 1));
        i64 y2 = (r % ((H - 1) * (D - 1))) / (D - 1);
        i64 z2 = r % (D - 1);
        x = x2; y = y2; z = z2; y1 = y2 + 1; x1 = x2; z1 = z2 + 1;
    } else if (dir == 5) {
        i64 y2 = r / ((W - 1) * (D - 1));
        i64 z2 = (r % ((W - 1) * (D - 1))) / (W - 1);
        i64 x2 = r % (W - 1);
        x = x2 + 1; y = y2; z = z2; y1 = y2; x1 = x2; z1 = z2 + 1;
    } else {
        i64 z2 = r / ((W - 1) * (H - 1));
        i64 y2 = (r % ((W - 1) * (H - 1))) / (W - 1);
        i64 x2 = r % (W - 1);
        x = x2 + 1; y = y2; z = z2; y1 = y2 + 1; x1 = x2; z1 = z2 + 1;
    }
    id1 = x + y * W + z * (W * H);
    id2 = x1 + y1 * W + z1 * (W * H);
    if (id1 < id2) std::swap(id1, id2);
}

static inline void sort3_desc(i64& a, i64& b, i64& c) {
    if (a < b) std::swap(a, b);
    if (a < c) std::swap(a, c);
    if (b < c) std::swap(b, c);
}

// getTriangleID1 main_3d.cu:246-363 (inputs already index-sorted descending).
static i64 tri_id1(const Grid& g, i64 v1, i64 v2, i64 v3, int direction) {
    const i64 W = g.W, H = g.H, D = g.D;
    i64 s1 = v1, s2 = v2, s3 = v3;
    sort3_desc(s1, s2, s3);
    i64 x1, y1, z1, x2, y2, z2, x3, y3, z3;
    g.xyz(s1, x1, y1, z1);
    g.xyz(s2, x2, y2, z2);
    g.xyz(s3, x3, y3, z3);
    i64 mx = std::min(x1, std::min(x2, x3));
    i64 my = std::min(y1, std::min(y2, y3));
    i64 mz = std::min(z1, std::min(z2, z3));
    if (direction == 0) {
        i64 cell = my * (W - 1) + mx;
        if (x3 == x1 && x1 == mx && y3 == y2 && y2 == my) return cell * 2 + 2 * mz * (W - 1) * (H - 1);
        return cell * 2 + 1 + 2 * mz * (W - 1) * (H - 1);
    } else if (direction == 1) {
        i64 cell = my * (D - 1) + mz;
        i64 off = 2 * D * (W - 1) * (H - 1) + 2 * mx * (H - 1) * (D - 1);
        if (y2 == y1 && y1 == my + 1 && z3 == z2 && z2 == mz) return cell * 2 + off;
        return cell * 2 + 1 + off;
    } else if (direction == 2) {
        i64 cell = mz * (W - 1) + mx;
        i64 off = 2 * (D * (W - 1) * (H - 1) + W * (H - 1) * (D - 1) + my * (W - 1) * (D - 1));
        if (z2 == z3 && z2 == mz && x1 == x3 && x1 == mx) return cell * 2 + off;
        return cell * 2 + 1 + off;
    } else if (direction == 3) {
        i64 cell = mz * (W - 1) * (H - 1) + my * (W - 1) + mx;
        int cur = -1;
        if (x3 == mx && y3 == my && z3 == mz && x2 == mx + 1 && y2 == my && z2 == mz && x1 == mx &&
            y1 == my + 1 && z1 == mz + 1)
            cur = 0;
        if (x3 == mx + 1 && y3 == my && z3 == mz && x2 == mx && y2 == my + 1 && z2 == mz &&
            x1 == mx && y1 == my + 1 && z1 == mz + 1)
            cur = 1;
        if (x3 == mx + 1 && y3 == my && z3 == mz && x2 == mx + 1 && y2 == my + 1 && z2 == mz &&
            x1 == mx && y1 == my + 1 && z1 == mz + 1)
            cur = 2;
        if (x3 == mx + 1 && y3 == my && z3 == mz && x2 == mx && y2 == my && z2 == mz + 1 &&
            x1 == mx && y1 == my + 1 && z1 == mz + 1)
            cur = 3;
        if (x3 == mx + 1 && y3 == my && z3 == mz && x2 == mx + 1 && y2 == my && z2 == mz + 1 &&
            x1 == mx && y1 == my + 1 && z1 == mz + 1)
            cur = 4;
        if (x3 == mx + 1 && y3 == my && z3 == mz && x2 == mx && y2 == my + 1 && z2 == mz + 1 &&
            x1 == mx + 1 && y1 == my + 1 && z1 == mz + 1)
            cur = 5;
        return cell * 6 + cur + 2 * D * (W - 1) * (H - 1) + 2 * W * (H - 1) * (D - 1) +
               2 * H * (W - 1) * (D - 1);
    }
    return -1;
}

// getTriangleID main_3d.cu:364-423
static i64 tri_id(const Grid& g, i64 v1, i64 v2, i64 v3) {
    i64 s1 = v1, s2 = v2, s3 = v3;
    sort3_desc(s1, s2, s3);
    i64 x1, y1, z1, x2, y2, z2, x3, y3, z3;
    g.xyz(s1, x1, y1, z1);
    g.xyz(s2, x2, y2, z2);
    g.xyz(s3, x3, y3, z3);
    if (z1 == z2 && z2 == z3) return tri_id1(g, s1, s2, s3, 0);
    if (x1 == x2 && x2 == x3) return tri_id1(g, s1, s2, s3, 1);
    if (y1 == y2 && y2 == y3) return tri_id1(g, s1, s2, s3, 2);
    return tri_id1(g, s1, s2, s3, 3);
}

// getVerticesFromTriangleID main_3d.cu:425-565; returns v1 > v2 > v3 (index order).
static void tri_verts(const Grid& g, i64 t, i64& v1, i64& v2, i64& v3) {
    const i64 W = g.W, H = g.H, D = g.D;
    i64 b, mx, my, mz;
    if (t < 2 * D * (W - 1) * (H - 1)) {
        b = t / 2;
        mz = b / ((W - 1) * (H - 1));
        my = (b % ((W - 1) * (H - 1))) / (W - 1);
        mx = b % (W - 1);
        if (t % 2 == 0) {
            v1 = g.idx(mx, my, mz); v2 = g.idx(mx + 1, my, mz); v3 = g.idx(mx, my + 1, mz);
        } else {
            v1 = g.idx(mx + 1, my + 1, mz); v2 = g.idx(mx + 1, my, mz); v3 = g.idx(mx, my + 1, mz);
        }
        sort3_desc(v1, v2, v3);
        return;
    }
    t -= 2 * D * (W - 1) * (H - 1);
    if (t < 2 * W * (H - 1) * (D - 1)) {
        b = t / 2;
        mx = b / ((H - 1) * (D - 1));
        my = (b % ((H - 1) * (D - 1))) / (D - 1);
        mz = b % (D - 1);
        if (t % 2 == 0) {
            v1 = g.idx(mx, my, mz); v2 = g.idx(mx, my + 1, mz); v3 = g.idx(mx, my + 1, mz + 1);
        } else {
            v1 = g.idx(mx, my, mz); v2 = g.idx(mx, my, mz + 1); v3 = g.idx(mx, my + 1, mz + 1);
        }
        sort3_desc(v1, v2, v3);
        return;
    }
    t -= 2 * W * (H - 1) * (D - 1);
    if (t < 2 * H * (W - 1) * (D - 1)) {
        b = t / 2;
        my = b / ((W - 1) * (D - 1));
        mz = (b % ((W - 1) * (D - 1))) / (W - 1);
        mx = b % (W - 1);
        if (t % 2 == 0) {
            v1 = g.idx(mx, my, mz); v2 = g.idx(mx + 1, my, mz); v3 = g.idx(mx, my, mz + 1);
        } else {
            v1 = g.idx(mx + 1, my, mz); v2 = g.idx(mx, my, mz + 1); v3 = g.idx(mx + 1, my, mz + 1);
        }
        sort3_desc(v1, v2, v3);
        return;
    }
    t -= 2 * H * (W - 1) * (D - 1);
    b = t / 6;
    mz = b / ((W - 1) * (H - 1));
    my = (b % ((W - 1) * (H - 1))) / (W - 1);
    mx = b % (W - 1);
    switch (t % 6) {
        case 0: v1 = g.idx(mx, my, mz); v2 = g.idx(mx + 1, my, mz); v3 = g.idx(mx, my + 1, mz + 1); break;
        case 1: v1 = g.idx(mx + 1, my, mz); v2 = g.idx(mx, my + 1, mz); v3 = g.idx(mx, my + 1, mz + 1); break;
        case 2: v1 = g.idx(mx + 1, my, mz); v2 = g.idx(mx + 1, my + 1, mz); v3 = g.idx(mx, my + 1, mz + 1); break;
        case 3: v1 = g.idx(mx + 1, my, mz); v2 = g.idx(mx, my, mz + 1); v3 = g.idx(mx, my + 1, mz + 1); break;
        case 4: v1 = g.idx(mx + 1, my, mz); v2 = g.idx(mx + 1, my, mz + 1); v3 = g.idx(mx, my + 1, mz + 1); break;
        default: v1 = g.idx(mx + 1, my, mz); v2 = g.idx(mx, my + 1, mz + 1); v3 = g.idx(mx + 1, my + 1, mz + 1); break;
    }
    sort3_desc(v1, v2, v3);
}

static inline void sort4_desc(i64& a, i64& b, i64& c, i64& d) {
    // the reference's exchange sequence, main_3d.cu:598-605 / 732-739
    if (a < b) std::swap(a, b);
    if (a < c) std::swap(a, c);
    if (a < d) std::swap(a, d);
    if (b < c) std::swap(b, c);
    if (b < d) std::swap(b, d);
    if (c < d) std::swap(c, d);
}

// getTetrahedraID main_3d.cu:572-670
static i64 tet_id(const Grid& g, i64 v1, i64 v2, i64 v3, i64 v4) {
    const i64 W = g.W, H = g.H;
    i64 x[4], y[4], z[4];
    i64 vs[4] = {v1, v2, v3, v4};
    for (int k = 0; k < 4; k++) g.xyz(vs[k], x[k], y[k], z[k]);
    i64 mx = std::min(std::min(x[0], x[1]), std::min(x[2], x[3]));
    i64 my = std::min(std::min(y[0], y[1]), std::min(y[2], y[3]));
    i64 mz = std::min(std::min(z[0], z[1]), std::min(z[2], z[3]));
    i64 cell = mz * (W - 1) * (H - 1) + my * (W - 1) + mx;
    i64 a = v1, b = v2, c = v3, d = v4;
    sort4_desc(a, b, c, d);
    i64 ax, ay, az, bx, by, bz, cx, cy, cz, dx, dy, dz;
    g.xyz(a, ax, ay, az);
    g.xyz(b, bx, by, bz);
    g.xyz(c, cx, cy, cz);
    g.xyz(d, dx, dy, dz);
    // (d,c,b,a) = sorted ascending: sortedV4, sortedV3, sortedV2, sortedV1
    auto is = [](i64 px, i64 py, i64 pz, i64 qx, i64 qy, i64 qz) {
        return px == qx && py == qy && pz == qz;
    };
    if (is(dx, dy, dz, mx, my, mz) && is(cx, cy, cz, mx + 1, my, mz) && is(bx, by, bz, mx, my + 1, mz) &&
        is(ax, ay, az, mx, my + 1, mz + 1))
        return cell * 6;
    if (is(dx, dy, dz, mx, my, mz) && is(cx, cy, cz, mx + 1, my, mz) && is(bx, by, bz, mx, my, mz + 1) &&
        is(ax, ay, az, mx, my + 1, mz + 1))
        return cell * 6 + 1;
    if (is(dx, dy, dz, mx + 1, my, mz) && is(cx, cy, cz, mx, my + 1, mz) &&
        is(bx, by, bz, mx + 1, my + 1, mz) && is(ax, ay, az, mx, my + 1, mz + 1))
        return cell * 6 + 2;
    if (is(dx, dy, dz, mx + 1, my, mz) && is(cx, cy, cz, mx + 1, my + 1, mz) &&
        is(bx, by, bz, mx, my + 1, mz + 1) && is(ax, ay, az, mx + 1, my + 1, mz + 1))
        return cell * 6 + 3;
    if (is(dx, dy, dz, mx + 1, my, mz) && is(cx, cy, cz, mx, my, mz + 1) &&
        is(bx, by, bz, mx + 1, my, mz + 1) && is(ax, ay, az, mx, my + 1, mz + 1))
        return cell * 6 + 4;
    if (is(dx, dy, dz, mx + 1, my, mz) && is(cx, cy, cz, mx + 1, my, mz + 1) &&
        is(bx, by, bz, mx, my + 1, mz + 1) && is(ax, ay, az, mx + 1, my + 1, mz + 1))
        return cell * 6 + 5;
    return -1;
}

// getVerticesFromTetrahedraID main_3d.cu:672-740; returns v1 > v2 > v3 > v4 (index order).
static void tet_verts(const Grid& g, i64 q, i64& v1, i64& v2, i64& v3, i64& v4) {
    const i64 W = g.W, H = g.H;
    i64 b = q / 6;
    int type = (int)(q % 6);
    i64 cz = b / ((W - 1) * (H - 1));
    i64 cy = (b % ((W - 1) * (H - 1))) / (W - 1);
    i64 cx = b % (W - 1);
    switch (type) {
        case 0: v1 = g.idx(cx, cy, cz); v2 = g.idx(cx + 1, cy, cz); v3 = g.idx(cx, cy + 1, cz); v4 = g.idx(cx, cy + 1, cz + 1); break;
        case 1: v1 = g.idx(cx, cy, cz); v2 = g.idx(cx + 1, cy, cz); v3 = g.idx(cx, cy, cz + 1); v4 = g.idx(cx, cy + 1, cz + 1); break;
        case 2: v1 = g.idx(cx + 1, cy, cz); v2 = g.idx(cx, cy + 1, cz); v3 = g.idx(cx + 1, cy + 1, cz); v4 = g.idx(cx, cy + 1, cz + 1); break;
        case 3: v1 = g.idx(cx + 1, cy, cz); v2 = g.idx(cx + 1, cy + 1, cz); v3 = g.idx(cx, cy + 1, cz + 1); v4 = g.idx(cx + 1, cy + 1, cz + 1); break;
        case 4: v1 = g.idx(cx + 1, cy, cz); v2 = g.idx(cx, cy, cz + 1); v3 = g.idx(cx + 1, cy, cz + 1); v4 = g.idx(cx, cy + 1, cz + 1); break;
        default: v1 = g.idx(cx + 1, cy, cz); v2 = g.idx(cx + 1, cy, cz + 1); v3 = g.idx(cx, cy + 1, cz + 1); v4 = g.idx(cx + 1, cy + 1, cz + 1); break;
    }
    sort4_desc(v1, v2, v3, v4);
}

// Common neighbours of (v1,v2,v3) that complete a tetrahedron, in the reference's search order
// (direction index from v1). LITERAL = the 14^3 search of main_3d.cu:4380-4426 / 2234-2304.
static int tri_cofacet_v4(const Grid& g, i64 v1, i64 v2, i64 v3, bool literal, i64 out[14]) {
    int n = 0;
    i64 x1, y1, z1, x2, y2, z2, x3, y3, z3;
    g.xyz(v1, x1, y1, z1);
    g.xyz(v2, x2, y2, z2);
    g.xyz(v3, x3, y3, z3);
    if (literal) {
        for (int i = 0; i < 14; i++) {
            i64 nx = x1 + DIRS[i][0], ny = y1 + DIRS[i][1], nz = z1 + DIRS[i][2];
            i64 n1 = nx + ny * g.W + nz * g.W * g.H;
            if (nx >= g.W || ny >= g.H || nz >= g.D || nx < 0 || ny < 0 || nz < 0 || n1 >= g.N || n1 < 0) continue;
            for (int j = 0; j < 14; j++) {
                i64 mx = x2 + DIRS[j][0], my = y2 + DIRS[j][1], mz = z2 + DIRS[j][2];
                i64 n2 = mx + my * g.W + mz * g.W * g.H;
                if (mx >= g.W || my >= g.H || mz >= g.D || mx < 0 || my < 0 || mz < 0 || n2 >= g.N || n2 < 0) continue;
                for (int k = 0; k < 14; k++) {
                    i64 px = x3 + DIRS[k][0], py = y3 + DIRS[k][1], pz = z3 + DIRS[k][2];
                    i64 n3 = px + py * g.W + pz * g.W * g.H;
                    if (px >= g.W || py >= g.H || pz >= g.D || px < 0 || py < 0 || pz < 0 || n3 >= g.N || n3 < 0) continue;
                    if (!(n1 == n2 && n2 == n3 && n1 == n3)) continue;
                    if (n1 == v1 || n1 == v2 || n1 == v3) continue;
                    out[n++] = n1;
                }
            }
        }
    } else {
        for (int i = 0; i < 14; i++) {
            i64 nx = x1 + DIRS[i][0], ny = y1 + DIRS[i][1], nz = z1 + DIRS[i][2];
            if (!g.inb(nx, ny, nz)) continue;
            i64 n1 = g.idx(nx, ny, nz);
            if (n1 == v2 || n1 == v3) continue;
            if (g.adjacent(n1, v2) && g.adjacent(n1, v3)) out[n++] = n1;
        }
    }
    return n;
}

// Common neighbours of an edge (v1,v2), main_3d.cu:2027-2040 (14x14 search) or direct.
static int edge_common(const Grid& g, i64 v1, i64 v2, bool literal, i64 out[14]) {
    int n = 0;
    i64 x1, y1, z1, x2, y2, z2;
    g.xyz(v1, x1, y1, z1);
    g.xyz(v2, x2, y2, z2);
    if (literal) {
        for (int i = 0; i < 14; i++) {
            i64 nx = x1 + DIRS[i][0], ny = y1 + DIRS[i][1], nz = z1 + DIRS[i][2];
            for (int j = 0; j < 14; j++) {
                i64 mx = x2 + DIRS[j][0], my = y2 + DIRS[j][1], mz = z2 + DIRS[j][2];
                i64 nb = nx + ny * g.W + nz * (g.H * g.W);
                if (nx == mx && ny == my && nz == mz && nx >= 0 && nx < g.W && ny >= 0 && ny < g.H &&
                    nz >= 0 && nz < g.D && nb < g.N && nb >= 0)
                    out[n++] = nb;
            }
        }
    } else {
        for (int i = 0; i < 14; i++) {
            i64 nx = x1 + DIRS[i][0], ny = y1 + DIRS[i][1], nz = z1 + DIRS[i][2];
            if (!g.inb(nx, ny, nz)) continue;
            i64 nb = g.idx(nx, ny, nz);
            if (nb == v2) continue;
            if (g.adjacent(nb, v2)) out[n++] = nb;
        }
    }
    return n;
}

// get_cell_tetra main_3d.cu:4363-4430: x = first cofacet found, y = second (-1 if none).
static void cell_tetra(const Grid& g, i64 face, i64& x, i64& y, bool literal) {
    i64 v1, v2, v3;
    tri_verts(g, face, v1, v2, v3);
    i64 c[14];
    int n = tri_cofacet_v4(g, v1, v2, v3, literal, c);
    x = -1;
    y = -1;
    for (int k = 0; k < n; k++) {
        i64 t = tet_id(g, v1, v2, v3, c[k]);
        if (k == 0) x = t;
        else y = t;
    }
}

// ----------------------------------------------------------------------------------------------
// Arithmetic of the edit buffer (main_3d.cu:239-244, 1122-1164)
// ----------------------------------------------------------------------------------------------
template <class T> struct Arith;
template <> struct Arith<double> {
    static double delta0(double f, double g, double eps) { return (f - eps - g) / 2.0; }
    // main_3d.cu:1133-1135: accept if |f - (g + d)| <= eps, then g += d.
    static bool step(double f, double& g, double d, double eps) {
        if (std::fabs(f - (g + d)) <= eps) {
            g += d;
            return true;
        }
        return false;
    }
    static double snap(double f, double eps) { return f - eps; }
};
static inline float next_up_f(float x) {
    if (x != x) return x;
    if (x == 0.0f) return std::ldexp(1.0f, -149);  // smallest positive subnormal
    uint32_t b;
    std::memcpy(&b, &x, 4);
    if (x > 0) b += 1;
    else b -= 1;
    float r;
    std::memcpy(&r, &b, 4);
    return r;
}
template <> struct Arith<float> {
    static double delta0(float f, float g, double eps) { return ((double)f - eps - (double)g) / 2.0; }
    // Float32 storage: the candidate is rounded to float first and the eps check is done on the
    // rounded value, so |g' - f| <= eps holds for the stored value.
    static bool step(float f, float& g, double d, double eps) {
        float c = (float)((double)g + d);
        if (std::fabs((double)f - (double)c) <= eps) {
            g = c;
            return true;
        }
        return false;
    }
    // Smallest float x with f - x <= eps (evaluated in double): the float32 form of f - eps.
    static float snap(float f, double eps) {
        float x = (float)((double)f - eps);
        if ((double)f - (double)x > eps) x = next_up_f(x);
        return x;
    }
};

// ----------------------------------------------------------------------------------------------
// Correction state
// ----------------------------------------------------------------------------------------------
struct Flags {
    int32_t preserve_saddles, preserve_vpath, preserve_geometry, filtration, preserve_connectors,
        preserve_type, edit_type, max_sweeps, finish;
};

constexpr int32_t FROZEN = 7;  // counter value of a vertex restored to f by the finisher

enum { F_FIELD = 0, G_FIELD = 1 };

template <class T> struct State {
    Grid g;
    const T* f;
    T* d;  // decompressed field, corrected in place
    double eps, range;
    Flags fl;
    bool literal;
    std::vector<i64> pv[2], pe[2], pf[2], pt[2];
    std::vector<int32_t> counter;
    std::vector<uint8_t> target;
    std::vector<i64> l_min, l_s1, l_s2, l_max, l_as, l_ds, l_dsw;
    std::vector<i64> as_index, ds_index, dsw_index;  // index arrays (2T, 2E, F)
    std::vector<i64> min_label, max_label;           // connectivity-mode f labels (2E, 2F)
    i64 walk_overflow = 0;

    State(i64 W, i64 H, i64 D) : g(W, H, D) {}

    inline T val(int which, i64 v) const { return which == F_FIELD ? f[v] : d[v]; }
    // compare_vertices main_3d.cu:1043-1049: a > b in (value, index) order.
    inline bool cmp(i64 a, i64 b, int which) const {
        T va = val(which, a), vb = val(which, b);
        return va > vb || (va == vb && a > b);
    }
    // Overloaded pairing-array reads with the out-of-range resolution.
    inline i64 PV(int w, i64 i) const { return (i >= 0 && i < g.N) ? pv[w][i] : -1; }
    inline i64 PE(int w, i64 i) const { return (i >= 0 && i < g.E) ? pe[w][i] : -1; }
    inline i64 PF(int w, i64 i) const { return (i >= 0 && i < g.F) ? pf[w][i] : -1; }
    inline i64 PT(int w, i64 i) const { return (i >= 0 && i < g.T) ? pt[w][i] : -1; }

    double delta(i64 v) const {
        // get_delta main_3d.cu:239-244
        if (fl.edit_type == 1) return -std::ldexp(eps, -counter[v]);
        return Arith<T>::delta0(f[v], d[v], eps);
    }
    inline void tgt(i64 v) {
        if (v >= 0 && v < g.N) __atomic_store_n(&target[v], (uint8_t)1, __ATOMIC_RELAXED);
    }

    void alloc() {
        for (int w = 0; w < 2; w++) {
            pv[w].assign(g.N, -1);
            pe[w].assign(g.E, -1);
            pf[w].assign(g.F, -1);
            pt[w].assign(g.T, -1);
        }
        counter.assign(g.N, 1);
        target.assign(g.N, 0);
        as_index.assign(2 * g.T, -1);
        ds_index.assign(2 * g.E, -1);
        dsw_index.assign(g.F, -1);
    }

    // ------------------------------------------------------------------------------------------
    // Discrete gradient: five passes (main_3d.cu:1930-2113, 2197-2316)
    // ------------------------------------------------------------------------------------------
    void reset(int w) {
        // init_vertices/edges/faces/tetras main_3d.cu:4305-4361
        std::fill(pv[w].begin(), pv[w].end(), -1);
        std::fill(pe[w].begin(), pe[w].end(), -1);
        std::fill(pf[w].begin(), pf[w].end(), -1);
        std::fill(pt[w].begin(), pt[w].end(), -1);
    }

    // processVertices main_3d.cu:1930-1988
    void pass_vertices(int w) {
        std::vector<i64>& PVw = pv[w];
        std::vector<i64>& PEw = pe[w];
#pragma omp parallel for schedule(static)
        for (i64 v = 0; v < g.N; v++) {
            i64 x, y, z;
            g.xyz(v, x, y, z);
            bool have = false;
            T best = 0;
            i64 best_id = -1, paired = -1;
            T dv = val(w, v);
            for (int i = 0; i < 14; i++) {
                i64 nx = x + DIRS[i][0], ny = y + DIRS[i][1], nz = z + DIRS[i][2];
                if (!g.inb(nx, ny, nz)) continue;
                i64 n = g.idx(nx, ny, nz);
                T dn = val(w, n);
                bool lower = dn < dv || (dn == dv && n < v);
                // "largest_v" starts at DBL_MAX: any finite value is smaller.
                bool better = !have || dn < best || (dn == best && n < best_id);
                if (lower && better) {
                    have = true;
                    best = dn;
                    best_id = n;
                    paired = edge_id(g, x, y, z, nx, ny, nz);
                }
            }
            if (paired != -1) {
                PVw[v] = paired;
                PEw[paired] = v;
            }
        }
    }

    // edgePairingKernel main_3d.cu:1990-2113
    void pass_edges(int w, int a2) {
        std::vector<i64>& PEw = pe[w];
        std::vector<i64>& PFw = pf[w];
#pragma omp parallel for schedule(static)
        for (i64 e = 0; e < g.E; e++) {
            if (PEw[e] != -1) continue;
            i64 v1, v2;
            edge_verts(g, e, v1, v2);
            i64 cn[14];
            int n = edge_common(g, v1, v2, literal, cn);
            bool have = false;
            T best = 0;
            i64 best_id = INT64_MAX, paired = -1;
            for (int k = 0; k < n; k++) {
                i64 nb = cn[k];
                T dn = val(w, nb);
                i64 tid = tri_id(g, v1, v2, nb);
                if (a2 == 0) {
                    if (cmp(nb, v1, w)) continue;  // main_3d.cu:2057
                    if (cmp(nb, v2, w)) continue;  // main_3d.cu:2058
                    if (!have || dn < best || (dn == best && nb < best_id)) {
                        have = true;
                        best = dn;
                        best_id = nb;
                        paired = tid;
                    }
                } else {
                    bool c1 = cmp(v1, nb, w), c2 = cmp(v2, nb, w);  // main_3d.cu:2075-2078
                    if (c1 != c2) {
                        if (!have || dn < best || (dn == best && nb < best_id)) {
                            have = true;
                            best = dn;
                            best_id = nb;
                            paired = tid;
                        }
                    }
                }
            }
            if (paired != -1 && PFw[paired] == -1) {  // main_3d.cu:2103-2109
                PEw[e] = paired;
                PFw[paired] = e;
            }
        }
    }

    // trianglePairingKernel main_3d.cu:2197-2316
    void pass_triangles(int w, int a2) {
        std::vector<i64>& PFw = pf[w];
        std::vector<i64>& PTw = pt[w];
#pragma omp parallel for schedule(dynamic, 4096)
        for (i64 t = 0; t < g.F; t++) {
            if (PFw[t] != -1) continue;
            i64 v1, v2, v3;
            tri_verts(g, t, v1, v2, v3);
            i64 c[14];
            int n = tri_cofacet_v4(g, v1, v2, v3, literal, c);
            bool have = false;
            T best = 0;
            i64 best_id = INT64_MAX, paired = -1;
            for (int k = 0; k < n; k++) {
                i64 v4 = c[k];
                i64 qid = tet_id(g, v1, v2, v3, v4);
                T d4 = val(w, v4);
                if (a2 == 0) {
                    if (cmp(v4, v1, w) || cmp(v4, v2, w) || cmp(v4, v3, w)) continue;  // :2269-2271
                    if (!have || d4 < best || (d4 == best && v4 < best_id)) {
                        have = true;
                        best = d4;
                        best_id = v4;
                        paired = qid;
                    }
                } else {
                    bool c1 = cmp(v1, v4, w), c2 = cmp(v2, v4, w), c3 = cmp(v3, v4, w);
                    if ((c1 && c2 && !c3) || (c1 && !c2 && c3) || (!c1 && c2 && c3)) {  // :2289
                        if (!have || d4 < best || (d4 == best && v4 < best_id)) {
                            have = true;
                            best = d4;
                            best_id = v4;
                            paired = qid;
                        }
                    }
                }
            }
            if (paired != -1 && PTw[paired] == -1) {  // main_3d.cu:2307-2313
                PFw[t] = paired;
                PTw[paired] = t;
            }
        }
    }

    // The pass sequence of main_3d.cu:6550-6564 (f) and 6864-6874 (g).
    void gradient(int w) {
        reset(w);
        pass_vertices(w);
        pass_edges(w, 0);
        pass_triangles(w, 0);
        pass_edges(w, 1);
        pass_triangles(w, 1);
    }

    // ------------------------------------------------------------------------------------------
    // Mismatch detection (main_3d.cu:2492-2550). Lists are sorted by id (the reference's order is
    // that of atomicAdd, nondeterministic; only the set matters downstream).
    // ------------------------------------------------------------------------------------------
    static void xor_list(const std::vector<i64>& a, const std::vector<i64>& b, std::vector<i64>& out) {
        out.clear();
        for (size_t i = 0; i < a.size(); i++)
            if ((a[i] == -1) != (b[i] == -1)) out.push_back((i64)i);
    }
    void detect_min() { xor_list(pv[0], pv[1], l_min); }
    void detect_s1() { xor_list(pe[0], pe[1], l_s1); }
    void detect_s2() { xor_list(pf[0], pf[1], l_s2); }
    void detect_max() { xor_list(pt[0], pt[1], l_max); }

    // ------------------------------------------------------------------------------------------
    // Repair rules
    // ------------------------------------------------------------------------------------------
    // fix_minimum main_3d.cu:2554-2586
    void fix_minimum(i64 id) {
        if (PV(F_FIELD, id) != -1) {
            i64 vertex, v2;
            edge_verts(g, PV(F_FIELD, id), vertex, v2);
            if (vertex != id) v2 = vertex;
            COV(C_MIN_F_PAIRED);
            tgt(v2);
        } else {
            COV(C_MIN_F_CRITICAL);
            tgt(id);
        }
    }

    // fix_saddle main_3d.cu:2588-3037
    void fix_saddle(i64 id) {
        i64 edge = id;
        i64 p = PE(F_FIELD, id);
        if (p != -1) {
            if (p < g.N && p >= 0 && PV(F_FIELD, p) == id) {  // :2602-2641
                COV(C_S1_FP_VERTEX);
                i64 v1, v2;
                edge_verts(g, edge, v1, v2);
                if (v2 != p) v1 = p;
                else {
                    v2 = v1;
                    v1 = p;
                }
                i64 dv = PV(G_FIELD, v1);
                if (dv == -1 || dv < 0 || dv >= g.E) {
                    COV(C_S1_FP_VERTEX_G_UNPAIRED);
                    return;
                }
                i64 fv1, fv2;
                edge_verts(g, dv, fv1, fv2);
                i64 false_v2 = (v1 == fv1) ? fv2 : fv1;
                if (cmp(v2, false_v2, G_FIELD)) {
                    COV(C_S1_FP_VERTEX_TARGET);
                    tgt(v2);
                }
            } else if (p < g.F && p >= 0 && PF(F_FIELD, p) == id) {  // :2643-2827
                COV(C_S1_FP_FACE);
                i64 face = p;
                i64 v1, v2;
                edge_verts(g, edge, v1, v2);
                i64 a, b, v3;
                tri_verts(g, face, a, b, v3);
                if (a != v1 && a != v2) v3 = a;
                else if (b != v1 && b != v2) v3 = b;
                i64 other1 = !cmp(v1, v3, G_FIELD) ? v3 : v1;    // :2664
                i64 other1_1 = !cmp(v2, v3, G_FIELD) ? v3 : v2;  // :2670
                bool v1_edge = cmp(v1, v3, G_FIELD), v2_edge = cmp(v2, v3, G_FIELD);
                bool o_v1 = cmp(v1, v3, F_FIELD), o_v2 = cmp(v2, v3, F_FIELD);
                if (v1_edge && v2_edge) {
                    COV(C_S1_FP_FACE_A);
                    tgt(v3);  // :2684-2699
                } else if (o_v1 && o_v2) {
                    COV(C_S1_FP_FACE_B);
                    tgt(v3);  // :2700-2708
                } else {
                    i64 q = PF(G_FIELD, p);
                    if (q != -1 && PE(G_FIELD, q) == p) {  // :2719
                        COV(C_S1_FP_FACE_C);
                        i64 false_edge = q;
                        i64 a1, a2;
                        edge_verts(g, false_edge, a1, a2);
                        i64 false_v = (a1 != v1 && a1 != v2) ? a1 : a2;  // :2730
                        double diff = delta(false_v);
                        // :2737-2744: the false edge's endpoints and the face's third vertex
                        i64 w1 = a1, w2 = a2;
                        i64 b1, b2, b3;
                        tri_verts(g, face, b1, b2, b3);
                        if (b1 != w1 && b1 != w2) b3 = b1;
                        else if (b2 != w1 && b2 != w2) b3 = b2;
                        bool ov1 = cmp(w2, b3, F_FIELD), ov2 = cmp(w1, b3, F_FIELD);  // :2765-2766
                        i64 pfe = PE(F_FIELD, false_edge);
                        if (ov1 && ov2 && pfe != -1 && PF(F_FIELD, pfe) == false_edge) {  // :2768
                            i64 c1, c2, c3;
                            tri_verts(g, pfe, c1, c2, c3);
                            if (c1 != w1 && c1 != w2) c3 = c1;
                            else if (c2 != w1 && c2 != w2) c3 = c2;
                            COV(C_S1_FP_FACE_C_FPAIR);
                            tgt(c3);
                        } else if (std::fabs(diff) > 0) {  // :2787
                            COV(C_S1_FP_FACE_C_DIFF);
                            tgt(false_v);
                        }
                    } else {  // :2801-2823
                        COV(C_S1_FP_FACE_D);
                        if (o_v1 && !v1_edge) {
                            COV(C_S1_FP_FACE_D1);
                            tgt(other1);
                        }
                        else {
                            COV(C_S1_FP_FACE_D2);
                            tgt(other1_1);
                        }
                    }
                }
            }
        } else {  // false negative, :2830-3036
            i64 q = PE(G_FIELD, id);
            if (q != -1 && q < g.N && q >= 0 && PV(G_FIELD, q) == id) {  // :2839-2850
                COV(C_S1_FN_VERTEX);
                tgt(q);
            } else if (q != -1 && q < g.F && q >= 0 && PF(G_FIELD, q) == id) {  // :2852
                COV(C_S1_FN_FACE);
                i64 face = q;
                i64 v1, v2;
                edge_verts(g, edge, v1, v2);
                i64 a, b, v3;
                tri_verts(g, face, a, b, v3);
                if (a != v1 && a != v2) v3 = a;
                else if (b != v1 && b != v2) v3 = b;
                bool v1_edge = cmp(v2, v3, G_FIELD), v2_edge = cmp(v1, v3, G_FIELD);  // :2886-2887
                bool or1 = cmp(v2, v3, F_FIELD), or2 = cmp(v1, v3, F_FIELD);
                if (v1_edge == v2_edge) {
                    COV(C_S1_FN_FACE_SAME);
                    if (v1_edge != or1) tgt(v2);
                    if (v2_edge != or2) tgt(v1);
                } else {
                    i64 t = PE(F_FIELD, id);
                    if (t != -1) {
                        if (t >= 0 && t < g.N && PV(F_FIELD, t) == id) {
                            COV(C_S1_FN_FACE_FVERTEX);
                            tgt(v1 == t ? v2 : v1);
                        }
                    } else {
                        COV(C_S1_FN_FACE_SEARCH);
                        // :2942-3003 — lowest f-vertex strictly between the endpoints
                        i64 cn[14];
                        int n = edge_common(g, v1, v2, literal, cn);
                        bool have = false;
                        T best = 0;
                        i64 best_id = INT64_MAX;
                        for (int k = 0; k < n; k++) {
                            i64 nb = cn[k];
                            bool c1 = cmp(v1, nb, F_FIELD), c2 = cmp(v2, nb, F_FIELD);
                            if (c1 != c2) {
                                T dn = f[nb];
                                if (!have || dn < best || (dn == best && nb < best_id)) {
                                    have = true;
                                    best = dn;
                                    best_id = nb;
                                }
                            }
                        }
                        if (have) {
                            COV(C_S1_FN_FACE_BETWEEN);
                            tgt(best_id);
                        }
                        else {
                            COV(C_S1_FN_FACE_NO_BETWEEN);
                            tgt(v1);
                            tgt(v2);
                        }
                    }
                }
            }
        }
    }

    // get_smallest_vertex main_3d.cu:2115-2194
    i64 smallest_vertex(i64 v1, i64 v2, i64 v3) {
        i64 c[14];
        int n = tri_cofacet_v4(g, v1, v2, v3, literal, c);
        bool have = false;
        T best = 0;
        i64 best_id = INT64_MAX;
        for (int k = 0; k < n; k++) {
            i64 v4 = c[k];
            bool c1 = cmp(v1, v4, F_FIELD), c2 = cmp(v2, v4, F_FIELD), c3 = cmp(v3, v4, F_FIELD);
            if ((c1 && c2 && !c3) || (c1 && !c2 && c3) || (!c1 && c2 && c3)) {
                T d4 = f[v4];
                if (!have || d4 < best || (d4 == best && v4 < best_id)) {
                    have = true;
                    best = d4;
                    best_id = v4;
                }
            }
        }
        return have ? best_id : -1;
    }

    static inline i64 other_of_tet(i64 q1, i64 q2, i64 q3, i64 q4, i64 a, i64 b, i64 c) {
        // main_3d.cu:3237-3239 style: first tet vertex not in the face, else the 4th
        i64 v4 = q4;
        if (q1 != a && q1 != b && q1 != c) v4 = q1;
        else if (q2 != a && q2 != b && q2 != c) v4 = q2;
        else if (q3 != a && q3 != b && q3 != c) v4 = q3;
        return v4;
    }

    // fix_2saddle main_3d.cu:3039-3983
    void fix_2saddle(i64 id) {
        i64 face = id;
        i64 p = PF(F_FIELD, id);
        if (p != -1) {
            if (p < g.E && p >= 0 && PE(F_FIELD, p) == id) {  // :3057-3223
                COV(C_S2_FP_EDGE);
                i64 edge = p;
                i64 v1, v2;
                edge_verts(g, edge, v1, v2);
                i64 a, b, v3;
                tri_verts(g, face, a, b, v3);
                if (a != v1 && a != v2) v3 = a;
                else if (b != v1 && b != v2) v3 = b;
                bool v1_edge = cmp(v2, v3, G_FIELD), v2_edge = cmp(v1, v3, G_FIELD);
                bool o_v1 = cmp(v2, v3, F_FIELD), o_v2 = cmp(v1, v3, F_FIELD);
                if (v1_edge && v2_edge) {  // :3097
                    COV(C_S2_FP_EDGE_A);
                    i64 f_id = PE(F_FIELD, edge);
                    if (f_id == -1 || PF(F_FIELD, f_id) != edge) return;
                    i64 c1, c2, c3;
                    tri_verts(g, f_id, c1, c2, c3);
                    if (c1 != v1 && c1 != v2) c3 = c1;
                    else if (c2 != v1 && c2 != v2) c3 = c2;
                    COV(C_S2_FP_EDGE_A_TARGET);
                    tgt(c3);
                } else if (o_v1 != o_v2 && v1_edge != v2_edge) {  // :3120
                    COV(C_S2_FP_EDGE_B);
                    i64 f_id = PE(G_FIELD, edge);
                    if (f_id == -1 || PF(G_FIELD, f_id) != edge) {
                        COV(C_S2_FP_EDGE_B_GUNPAIRED);
                        tgt(v3);
                    } else if (f_id < g.F && f_id >= 0) {  // :3135
                        i64 d1, d2, v3_1;
                        tri_verts(g, f_id, d1, d2, v3_1);
                        if (d1 != v1 && d1 != v2) v3_1 = d1;
                        else if (d2 != v1 && d2 != v2) v3_1 = d2;
                        bool e1 = cmp(v2, v3_1, G_FIELD), e2 = cmp(v1, v3_1, G_FIELD);
                        bool oe1 = cmp(v2, v3_1, F_FIELD), oe2 = cmp(v1, v3_1, F_FIELD);
                        if (e1 && e2 && ((!oe1 && oe2) || (oe1 && !oe2))) {
                            if (!oe1) {
                                COV(C_S2_FP_EDGE_B_V2);
                                tgt(v2);
                            }
                            else {
                                COV(C_S2_FP_EDGE_B_V1);
                                tgt(v1);
                            }
                        } else {
                            COV(C_S2_FP_EDGE_B_V3);
                            tgt(v3);
                        }
                    }
                } else if (!cmp(v1, v3, G_FIELD) || !cmp(v2, v3, G_FIELD)) {  // :3205
                    COV(C_S2_FP_EDGE_C);
                    tgt(v3);
                }
            } else if (p < g.T && p >= 0 && PT(F_FIELD, p) == id) {  // :3226-3477
                COV(C_S2_FP_TET);
                i64 tet = p;
                i64 v1, v2, v3;
                tri_verts(g, id, v1, v2, v3);
                i64 q1, q2, q3, q4;
                tet_verts(g, tet, q1, q2, q3, q4);
                i64 v4 = other_of_tet(q1, q2, q3, q4, v1, v2, v3);
                bool or1 = cmp(v3, v4, F_FIELD), or2 = cmp(v2, v4, F_FIELD), or3 = cmp(v1, v4, F_FIELD);
                if (!or1 || !or2 || !or3) {
                    i64 pff = PT(G_FIELD, tet);
                    if (pff != -1) {
                        i64 w1, w2, w3;
                        tri_verts(g, pff, w1, w2, w3);
                        bool g1 = cmp(w3, v4, G_FIELD), g2 = cmp(w2, v4, G_FIELD), g3 = cmp(w1, v4, G_FIELD);
                        bool h1 = cmp(w3, v4, F_FIELD), h2 = cmp(w2, v4, F_FIELD), h3 = cmp(w1, v4, F_FIELD);
                        if ((!g1 || !g2 || !g3) && (!h1 || !h2 || !h3)) {
                            COV(C_S2_FP_TET_MIXED);
                            i64 idd = PF(F_FIELD, pff);
                            if (idd >= 0 && idd < g.E && PE(F_FIELD, idd) == pff) {  // :3288
                                i64 e1, e2;
                                edge_verts(g, idd, e1, e2);
                                i64 vv3 = w3;
                                if (w1 != e1 && e2 != w1) vv3 = w1;  // :3293
                                else if (w2 != e1 && w2 != e2) vv3 = w2;
                                COV(C_S2_FP_TET_EDGE);
                                tgt(vv3);
                            } else if (idd >= 0 && idd < g.T && PT(F_FIELD, idd) == pff) {  // :3302
                                i64 t1, t2, t3, t4;
                                tet_verts(g, idd, t1, t2, t3, t4);
                                // :3307-3309 assign the face's vertex, not the tet's
                                if (t1 != w1 && t1 != w2 && t1 != w3) t4 = w1;
                                else if (t2 != w1 && t2 != w2 && t2 != w3) t4 = w2;
                                else if (t3 != w1 && t3 != w2 && t3 != w3) t4 = w3;
                                COV(C_S2_FP_TET_TET);
                                tgt(t1);
                                tgt(t2);
                                tgt(t3);
                                tgt(t4);
                            }
                        } else {  // :3324-3334
                            COV(C_S2_FP_TET_W);
                            tgt(w1);
                            tgt(w2);
                            tgt(w3);
                        }
                    } else {
                        COV(C_S2_FP_TET_GUNPAIRED);
                        tgt(v4);  // :3338-3345
                    }
                } else {
                    COV(C_S2_FP_TET_ALL_ABOVE);
                    tgt(v4);  // :3351-3357
                }
            }
        } else {  // false negative :3480-3981
            i64 q = PF(G_FIELD, id);
            if (q != -1 && q < g.E && PE(G_FIELD, q) == id) {  // :3488-3610
                COV(C_S2_FN_EDGE);
                i64 edge_id2 = q;
                i64 v1, v2;
                edge_verts(g, edge_id2, v1, v2);
                i64 s1 = !cmp(v1, v2, G_FIELD) ? v2 : v1;
                i64 s2 = !cmp(v1, v2, G_FIELD) ? v1 : v2;
                i64 tf = PE(F_FIELD, edge_id2);
                if (tf != -1 && PF(F_FIELD, tf) == edge_id2) {  // :3522
                    i64 a, b, v3;
                    tri_verts(g, tf, a, b, v3);
                    if (s1 != a && s2 != a) v3 = a;
                    if (s1 != b && s2 != b) v3 = b;
                    COV(C_S2_FN_EDGE_FFACE);
                    tgt(v3);
                } else if (tf >= 0 && tf < g.N && PV(F_FIELD, tf) == edge_id2) {  // :3544
                    COV(C_S2_FN_EDGE_FVERTEX);
                    tgt(v1);
                    tgt(v2);
                }
            } else if (q != -1 && q < g.T && PT(G_FIELD, q) == id) {  // :3613-3980
                COV(C_S2_FN_TET);
                i64 tet = q;
                i64 v1, v2, v3;
                tri_verts(g, face, v1, v2, v3);
                i64 q1, q2, q3, q4;
                tet_verts(g, tet, q1, q2, q3, q4);
                i64 v4 = other_of_tet(q1, q2, q3, q4, v1, v2, v3);
                bool f1 = cmp(v3, v4, G_FIELD), f2 = cmp(v2, v4, G_FIELD), f3 = cmp(v1, v4, G_FIELD);
                bool o1 = cmp(v3, v4, F_FIELD), o2 = cmp(v2, v4, F_FIELD), o3 = cmp(v1, v4, F_FIELD);
                if (f1 && f2 && f3) {
                    COV(C_S2_FN_TET_ALL);
                    // :3714-3927; the inner "else" branches are unreachable since f1,f2,f3 hold
                    if (f1 != o1) tgt(v3);
                    else if (f2 != o2) tgt(v2);
                    else if (f3 != o3) tgt(v1);
                } else {
                    i64 true_cell = PT(F_FIELD, tet);
                    if (true_cell >= 0 && true_cell < g.T) {  // :3937 (face id compared to T)
                        COV(C_S2_FN_TET_TRUECELL);
                        tgt(v1);
                        tgt(v2);
                        tgt(v3);
                    } else {
                        i64 m = smallest_vertex(v1, v2, v3);
                        if (m != -1) {
                            COV(C_S2_FN_TET_SMALLEST);
                            tgt(m);
                        }
                        else {
                            COV(C_S2_FN_TET_NO_SMALLEST);
                            tgt(v1);
                            tgt(v2);
                            tgt(v3);
                        }
                    }
                }
            }
        }
    }

    // fix_maximum main_3d.cu:3985-4303
    void fix_maximum(i64 id) {
        i64 p = PT(F_FIELD, id);
        if (p != -1) {  // false positive
            COV(C_MAX_FP);
            i64 face1 = p;
            i64 v1, v2, v3;
            tri_verts(g, face1, v1, v2, v3);
            i64 q1, q2, q3, q4;
            tet_verts(g, id, q1, q2, q3, q4);
            i64 v4 = other_of_tet(q1, q2, q3, q4, v1, v2, v3);
            bool o1 = cmp(v1, v4, F_FIELD), o2 = cmp(v2, v4, F_FIELD), o3 = cmp(v3, v4, F_FIELD);
            if (o1 && o2 && o3) {
                COV(C_MAX_FP_ALL_BELOW);
                tgt(v4);
            } else {
                i64 fc = PF(G_FIELD, face1);
                if (fc >= 0) {
                    if (fc < g.E && PE(G_FIELD, fc) == face1) {
                        i64 a, b;
                        edge_verts(g, fc, a, b);
                        COV(C_MAX_FP_GEDGE);
                        tgt(cmp(a, b, G_FIELD) ? a : b);
                    } else if (fc < g.T && PT(G_FIELD, fc) == face1) {
                        COV(C_MAX_FP_GTET);
                        tgt(v4);
                    }
                }
            }
        } else {  // false negative
            COV(C_MAX_FN);
            i64 face1 = PT(G_FIELD, id);
            i64 v1, v2, v3;
            tri_verts(g, face1, v1, v2, v3);
            i64 q1, q2, q3, q4;
            tet_verts(g, id, q1, q2, q3, q4);
            i64 v4 = other_of_tet(q1, q2, q3, q4, v1, v2, v3);
            bool f1 = cmp(v1, v4, G_FIELD), f2 = cmp(v2, v4, G_FIELD), f3 = cmp(v3, v4, G_FIELD);
            bool o1 = cmp(v1, v4, F_FIELD), o2 = cmp(v2, v4, F_FIELD), o3 = cmp(v3, v4, F_FIELD);
            if (f1 && f2 && f3) {
                COV(C_MAX_FN_ALL);
                if (!o1) tgt(v1);
                if (!o2) tgt(v2);
                if (!o3) tgt(v3);
                if (o1 && o2 && o3) {
                    i64 it = PF(F_FIELD, face1);
                    if (it != -1) {
                        i64 t1, t2, t3, t4;
                        tet_verts(g, it, t1, t2, t3, t4);
                        COV(C_MAX_FN_ALL_FTET);
                        tgt(other_of_tet(t1, t2, t3, t4, v1, v2, v3));
                    }
                }
            } else {
                i64 pid = PF(F_FIELD, face1);
                if (pid != -1) {
                    if (pid >= 0 && pid < g.E && PE(F_FIELD, pid) == face1) {
                        i64 a, b;
                        edge_verts(g, pid, a, b);
                        COV(C_MAX_FN_FEDGE);
                        tgt(a);
                        tgt(b);
                    } else {
                        i64 t1, t2, t3, t4;
                        tet_verts(g, pid, t1, t2, t3, t4);
                        COV(C_MAX_FN_FTET);
                        tgt(t1);
                        tgt(t2);
                        tgt(t3);
                        tgt(t4);
                    }
                } else {
                    COV(C_MAX_FN_FCRIT);
                    if (!o1) tgt(v1);
                    if (!o2) tgt(v2);
                    if (!o3) tgt(v3);
                }
            }
        }
    }

    // ------------------------------------------------------------------------------------------
    // Separatrix checks (main_3d.cu:4437-5062) and repairs (main_3d.cu:5135-5295)
    // ------------------------------------------------------------------------------------------
    // is_false main_3d.cu:4437-4497. Returns true and records the divergent tet if found.
    bool is_false(i64 id, i64& divergent) {
        i64 cur = id, old;
        i64 steps = 0;
        do {
            old = cur;
            i64 t = PT(F_FIELD, cur);
            i64 dt = PT(G_FIELD, cur);
            if (t == -1) {
                COV(C_ASW_F_CRITICAL);
                break;
            }
            if (t != dt) {
                divergent = cur;
                COV(C_ASW_DIVERGENT);
                return true;
            }
            i64 pft = PF(F_FIELD, t);
            if (pft == -1 || PT(F_FIELD, pft) != t) {
                COV(C_ASW_NOT_PAIRED_BACK);
                break;
            }
            i64 t1, t2;
            cell_tetra(g, t, t1, t2, literal);
            cur = (t1 == cur) ? t2 : t1;
            if (++steps > g.N) {
                walk_overflow++;
                break;
            }
        } while (cur != old && cur != -1);
        if (cur == -1) COV(C_ASW_BOUNDARY);
        return false;
    }

    // get_paired_id main_3d.cu:4499-4537 (which = field)
    i64 paired_max(i64 id, int which) {
        i64 cur = id, old;
        i64 steps = 0;
        do {
            old = cur;
            i64 t = PT(which, cur);
            if (t == -1) break;
            i64 pft = PF(which, t);
            if (pft == -1 || PT(which, pft) != t) break;
            i64 t1, t2;
            cell_tetra(g, t, t1, t2, literal);
            cur = (t1 == cur) ? t2 : t1;
            if (++steps > g.N) {
                walk_overflow++;
                break;
            }
        } while (cur != old && cur != -1);
        return cur;
    }

    // Cofacet tets of a face in the reference enumeration order (main_3d.cu:4567-4606).
    int face_tets(i64 face, i64 out[2]) {
        i64 v1, v2, v3;
        tri_verts(g, face, v1, v2, v3);
        i64 c[14];
        int n = tri_cofacet_v4(g, v1, v2, v3, literal, c);
        for (int k = 0; k < n && k < 2; k++) out[k] = tet_id(g, v1, v2, v3, c[k]);
        return std::min(n, 2);
    }

    // get_false_as_vpath main_3d.cu:4540-4610
    void detect_as_geometry() {
        l_as.clear();
        std::vector<std::vector<i64>> loc;
#pragma omp parallel
        {
#pragma omp single
            {
#ifdef _OPENMP
                loc.resize(omp_get_num_threads());
#else
                loc.resize(1);
#endif
            }
#ifdef _OPENMP
            std::vector<i64>& my = loc[omp_get_thread_num()];
#else
            std::vector<i64>& my = loc[0];
#endif
#pragma omp for schedule(dynamic, 1024)
            for (i64 i = 0; i < g.F; i++) {
                if (pf[0][i] != -1) continue;
                i64 tets[2];
                int n = face_tets(i, tets);
                for (int k = 0; k < n; k++) {
                    i64 dv;
                    if (is_false(tets[k], dv)) {
                        my.push_back(tets[k]);
                        as_index[2 * tets[k]] = dv;  // same value from every writer
                        as_index[2 * tets[k] + 1] = 3;
                        break;
                    }
                }
            }
        }
        for (auto& v : loc) l_as.insert(l_as.end(), v.begin(), v.end());
        std::sort(l_as.begin(), l_as.end());
    }

    // get_false_ds_vpath main_3d.cu:4837-4959. The reference's dec_current_id (:4849) is never
    // advanced, so the loop leaves at the second iteration through the v != dec_v test (:4861):
    // only the start vertex's f and g pairings are compared. Restated with the same variables.
    bool ds_walk_geometry(i64 start, i64& rec) {
        i64 cur = start, dec_cur = start;
        i64 v = start;
        do {
            v = cur;
            i64 dec_v = dec_cur;
            if (v != dec_v) break;
            i64 ce = PV(F_FIELD, v);
            if (ce == -1) {
                COV(C_DSG_F_MIN);
                break;
            }
            i64 dce = PV(G_FIELD, v);
            if (ce != dce) {
                COV(C_DSG_RECORD);
                rec = v;
                return true;
            }
            if (PE(F_FIELD, ce) == -1) break;
            i64 a, b;
            edge_verts(g, ce, a, b);
            cur = (a == cur) ? b : a;
        } while (PV(F_FIELD, v) != -1 && PV(G_FIELD, v) != -1);
        return false;
    }
    void detect_ds_geometry() {
        l_ds.clear();
        std::vector<std::vector<i64>> loc;
#pragma omp parallel
        {
#pragma omp single
            {
#ifdef _OPENMP
                loc.resize(omp_get_num_threads());
#else
                loc.resize(1);
#endif
            }
#ifdef _OPENMP
            std::vector<i64>& my = loc[omp_get_thread_num()];
#else
            std::vector<i64>& my = loc[0];
#endif
#pragma omp for schedule(dynamic, 1024)
            for (i64 i = 0; i < g.E; i++) {
                if (pe[0][i] != -1) continue;
                i64 v1, v2;
                edge_verts(g, i, v1, v2);
                i64 rec;
                if (ds_walk_geometry(v1, rec) || ds_walk_geometry(v2, rec)) {
                    my.push_back(i);
                    ds_index[2 * i] = rec;
                    ds_index[2 * i + 1] = 0;
                }
            }
        }
        for (auto& v : loc) l_ds.insert(l_ds.end(), v.begin(), v.end());
        std::sort(l_ds.begin(), l_ds.end());
    }

    // f descending labels, get_orginal_saddle_min_pairs main_3d.cu:4613-4687
    i64 f_min_label(i64 start) {
        i64 cur = start, v = start;
        i64 steps = 0;
        while (true) {
            v = cur;
            i64 ce = PV(F_FIELD, v);
            if (ce == -1) break;
            if (PE(F_FIELD, ce) == -1) break;
            i64 a, b;
            edge_verts(g, ce, a, b);
            cur = (a == cur) ? b : a;
            if (++steps > g.N) {
                walk_overflow++;
                break;
            }
        }
        return v;
    }
    void compute_labels() {
        min_label.assign(2 * g.E, -1);
        max_label.assign(2 * g.F, -1);
#pragma omp parallel for schedule(dynamic, 1024)
        for (i64 i = 0; i < g.E; i++) {
            if (pe[0][i] != -1) continue;
            i64 v1, v2;
            edge_verts(g, i, v1, v2);
            min_label[2 * i] = f_min_label(v1);
            min_label[2 * i + 1] = f_min_label(v2);
        }
        // get_orginal_saddle_max_pairs main_3d.cu:4689-4759, indexed by face id (shadowing fix)
#pragma omp parallel for schedule(dynamic, 1024)
        for (i64 i = 0; i < g.F; i++) {
            if (pf[0][i] != -1) continue;
            i64 tets[2];
            int n = face_tets(i, tets);
            for (int k = 0; k < n; k++) max_label[2 * i + k] = paired_max(tets[k], F_FIELD);
        }
    }

    // get_false_ds_vpath_connectivity main_3d.cu:4961-5062
    bool ds_walk_conn(i64 start, i64& trouble, i64& end) {
        i64 cur = start, v = start;
        i64 steps = 0;
        while (true) {
            v = cur;
            i64 dce = PV(G_FIELD, v);
            if (dce == -1) break;
            i64 ce = PV(F_FIELD, v);
            if (ce != dce && trouble == -1) trouble = v;
            if (PE(G_FIELD, dce) == -1) break;
            i64 a, b;
            edge_verts(g, dce, a, b);
            cur = (a == cur) ? b : a;
            if (++steps > g.N) {
                walk_overflow++;
                break;
            }
        }
        end = v;
        return true;
    }
    void detect_ds_connectivity() {
        l_ds.clear();
        for (i64 i = 0; i < g.E; i++) {
            if (pe[0][i] != -1) continue;
            i64 v1, v2;
            edge_verts(g, i, v1, v2);
            i64 trouble = -1, end;
            ds_walk_conn(v1, trouble, end);
            if (end != min_label[2 * i]) {
                COV(C_DSC_FIRST);
                if (trouble == -1) COV(C_DSC_NO_TROUBLE);
                l_ds.push_back(i);
                ds_index[2 * i] = trouble;
                ds_index[2 * i + 1] = 0;
                continue;
            }
            ds_walk_conn(v2, trouble, end);
            if (end != min_label[2 * i + 1]) {
                COV(C_DSC_SECOND);
                l_ds.push_back(i);
                ds_index[2 * i] = trouble;
                ds_index[2 * i + 1] = 0;
            }
        }
    }
    // get_false_as_vpath_connectivity main_3d.cu:4761-4835 (labels indexed by face id)
    void detect_as_connectivity() {
        l_as.clear();
        for (i64 i = 0; i < g.F; i++) {
            if (pf[1][i] != -1) continue;
            i64 tets[2];
            int n = face_tets(i, tets);
            for (int k = 0; k < n; k++) {
                i64 dec_id = paired_max(tets[k], G_FIELD);
                if (dec_id != max_label[2 * i + k]) {
                    COV(C_ASC_LABEL_DIFFERS);
                    i64 dv;
                    if (is_false(tets[k], dv)) {
                        COV(C_ASC_RECORD);
                        l_as.push_back(tets[k]);
                        as_index[2 * tets[k]] = dv;
                        as_index[2 * tets[k] + 1] = 3;
                        break;
                    }
                }
            }
        }
        std::sort(l_as.begin(), l_as.end());
    }

    // Deterministic walls: get_false_ds_wall main_3d.cu:5064-5133 with a private stack and
    // visited set per start face, edges in the reference order (e1,e2,e3), LIFO.
    void detect_ds_wall() {
        l_dsw.clear();
        std::vector<uint8_t> visited(g.F, 0);
        std::vector<i64> touched, stack;
        for (i64 i = 0; i < g.F; i++) {
            if (pf[0][i] != -1) continue;
            for (i64 t : touched) visited[t] = 0;
            touched.clear();
            stack.clear();
            stack.push_back(i);
            bool found = false;
            while (!stack.empty() && !found) {
                i64 tri = stack.back();
                stack.pop_back();
                if (visited[tri]) continue;
                visited[tri] = 1;
                touched.push_back(tri);
                i64 v1, v2, v3;
                tri_verts(g, tri, v1, v2, v3);
                i64 x1, y1, z1, x2, y2, z2, x3, y3, z3;
                g.xyz(v1, x1, y1, z1);
                g.xyz(v2, x2, y2, z2);
                g.xyz(v3, x3, y3, z3);
                i64 es[3] = {edge_id(g, x1, y1, z1, x2, y2, z2), edge_id(g, x1, y1, z1, x3, y3, z3),
                             edge_id(g, x2, y2, z2, x3, y3, z3)};
                for (int k = 0; k < 3; k++) {
                    i64 j = es[k];
                    i64 pj = PE(F_FIELD, j);
                    if (pj == -1) continue;
                    if (pj != PE(G_FIELD, j)) {
                        COV(C_WALLDET_FOUND);
                        l_dsw.push_back(i);
                        dsw_index[i] = j;
                        found = true;
                        break;
                    }
                    if (pj != tri && pj >= 0 && pj < g.F && PF(F_FIELD, pj) == j) {
                        COV(C_WALLDET_PUSH);
                        stack.push_back(pj);
                    }
                }
            }
        }
    }

    // fix_as_vpath main_3d.cu:5135-5244 (only the live else-branch, defect D14)
    void fix_as_vpath(i64 id) {
        i64 t_idx = as_index[2 * id];
        i64 true_edge = PT(F_FIELD, t_idx);
        i64 false_edge = PT(G_FIELD, t_idx);
        if (false_edge == -1 || true_edge == -1) return;  // unreachable with matching critical sets
        i64 pfe = PF(F_FIELD, false_edge);
        if (pfe >= 0 && pfe < g.T && PT(F_FIELD, pfe) == false_edge) {
            COV(C_FIXAS_FTET);
            i64 v1, v2, v3;
            tri_verts(g, false_edge, v1, v2, v3);
            i64 q1, q2, q3, q4;
            tet_verts(g, pfe, q1, q2, q3, q4);
            tgt(other_of_tet(q1, q2, q3, q4, v1, v2, v3));
        } else {
            COV(C_FIXAS_FACE);
            i64 v1, v2, v3;
            tri_verts(g, true_edge, v1, v2, v3);
            i64 a, b, tv3;
            tri_verts(g, false_edge, a, b, tv3);
            if (a != v1 && a != v2 && a != v3) tv3 = a;
            else if (b != v1 && b != v2 && b != v3) tv3 = b;
            tgt(tv3);
        }
    }

    // fix_ds_vpath main_3d.cu:5246-5295 (live else-branch)
    void fix_ds_vpath(i64 id) {
        i64 v_idx = ds_index[2 * id];
        if (v_idx < 0) {  // trouble_maker == -1: the reference reads index -1
            COV(C_FIXDS_NO_TROUBLE);
            return;
        }
        i64 tpe = PV(F_FIELD, v_idx);
        if (tpe == -1) {
            COV(C_FIXDS_FMIN);
            return;
        }
        COV(C_FIXDS_TARGET);
        i64 a, b;
        edge_verts(g, tpe, a, b);
        tgt(v_idx == a ? b : a);
    }

    // fix_ds_wall main_3d.cu:5297-5413
    void fix_ds_wall(i64 i) {
        i64 e = dsw_index[i];
        i64 tf = PE(F_FIELD, e), ff = PE(G_FIELD, e);
        if (tf == ff) return;
        if (tf >= 0 && ff >= 0 && tf < g.F && PF(F_FIELD, tf) == e && ff < g.F && PF(G_FIELD, ff) == e) {
            i64 v1, v2;
            edge_verts(g, e, v1, v2);
            i64 a, b, v3;
            tri_verts(g, tf, a, b, v3);
            if (a != v1 && a != v2) v3 = a;
            if (b != v1 && b != v2) v3 = b;
            double diff = delta(v3);
            if (std::fabs(diff) > 1e-16 * range) {
                COV(C_WALL_V3);
                tgt(v3);
            } else {
                COV(C_WALL_ENDS);
                tgt(v1);
                tgt(v2);
            }
        } else if (tf == -1 || tf >= g.F || (tf < g.F && tf >= 0 && PF(F_FIELD, tf) != e)) {
            COV(C_WALL_F_UNPAIRED);
            i64 v1, v2;
            edge_verts(g, e, v1, v2);
            tgt(v1);
            tgt(v2);
        } else if (ff == -1 || ff >= g.F || (ff < g.F && ff >= 0 && PF(G_FIELD, ff) != e)) {
            COV(C_WALL_G_UNPAIRED);
            i64 v1, v2;
            edge_verts(g, e, v1, v2);
            i64 a, b, v3;
            tri_verts(g, tf, a, b, v3);
            if (a != v1 && a != v2) v3 = a;
            else if (b != v1 && b != v2) v3 = b;
            tgt(v3);
        }
    }

    // ------------------------------------------------------------------------------------------
    // Edit buffer apply (main_3d.cu:1122-1164) and filtration (main_3d.cu:2405-2490)
    // ------------------------------------------------------------------------------------------
    // Returns the number of edit-set vertices whose value or counter changed; the edit set is kept
    // in last_edits. Frozen vertices are never edited.
    std::vector<i64> last_edits;
    i64 apply() {
        last_edits.clear();
        for (i64 v = 0; v < g.N; v++)
            if (target[v]) last_edits.push_back(v);
        const i64 n = (i64)last_edits.size();
        i64 moved = 0;
#pragma omp parallel for schedule(static) reduction(+ : moved)
        for (i64 k = 0; k < n; k++) {
            const i64 v = last_edits[k];
            target[v] = 0;
            if (counter[v] == FROZEN) continue;
            const T g0 = d[v];
            const int32_t k0 = counter[v];
            double dl = delta(v);
            if (fl.preserve_type == 0) {
                T gv = d[v];
                if (std::fabs(dl) > 1e-15 * range && counter[v] < 5 && Arith<T>::step(f[v], gv, dl, eps)) {
                    d[v] = gv;
                    counter[v] += 1;
                } else {
                    counter[v] = 6;
                    d[v] = Arith<T>::snap(f[v], eps);
                }
            } else {
                counter[v] = 6;
                d[v] = Arith<T>::snap(f[v], eps);
            }
            if (!(d[v] == g0) || counter[v] != k0) moved++;
        }
        return moved;
    }

    // Vertices of the first c[k] cells of the four mismatch lists.
    void mismatch_vertices(const i64* c, std::vector<i64>& out) const {
        for (i64 i = 0; i < c[0]; i++) out.push_back(l_min[i]);
        for (i64 i = 0; i < c[1]; i++) {
            i64 a, b;
            edge_verts(g, l_s1[i], a, b);
            out.push_back(a);
            out.push_back(b);
        }
        for (i64 i = 0; i < c[2]; i++) {
            i64 a, b, e;
            tri_verts(g, l_s2[i], a, b, e);
            out.push_back(a);
            out.push_back(b);
            out.push_back(e);
        }
        for (i64 i = 0; i < c[3]; i++) {
            i64 a, b, e, h;
            tet_verts(g, l_max[i], a, b, e, h);
            out.push_back(a);
            out.push_back(b);
            out.push_back(e);
            out.push_back(h);
        }
    }

    // Finisher step (no counterpart in the reference). For r below the longest grid side and a
    // non-empty zone, on the box B of vertices within Chebyshev distance r of a zone vertex:
    //   1. every vertex of B not yet frozen is snapped (g = snap(f)) and B is frozen (counter 7);
    //   2. ties the float snap creates are repaired in Jacobi rounds: a vertex u of B with a
    //      neighbour w such that g(w) == g(u), f(w) < f(u) and w > u (so the index tie-break
    //      orders them against f) is raised to the next float; rounds repeat until none is.
    // Otherwise g = f is restored on every vertex, all frozen. Returns the number of vertices
    // in B (every vertex for a restore).
    i64 finish_step(const std::vector<i64>& zone, int r) {
        const i64 span = std::max(g.W, std::max(g.H, g.D));
        if (zone.empty() || r >= span) {
            for (i64 u = 0; u < g.N; u++) {
                d[u] = f[u];
                counter[u] = FROZEN;
            }
            return g.N;
        }
        std::vector<uint8_t> mark(g.N, 0);
        for (i64 v : zone) {
            const i64 x = v % g.W, y = (v / g.W) % g.H, z = v / (g.W * g.H);
            const i64 x0 = std::max<i64>(x - r, 0), x1 = std::min<i64>(x + r, g.W - 1);
            const i64 y0 = std::max<i64>(y - r, 0), y1 = std::min<i64>(y + r, g.H - 1);
            const i64 z0 = std::max<i64>(z - r, 0), z1 = std::min<i64>(z + r, g.D - 1);
            for (i64 zz = z0; zz <= z1; zz++)
                for (i64 yy = y0; yy <= y1; yy++)
                    for (i64 xx = x0; xx <= x1; xx++) mark[(zz * g.H + yy) * g.W + xx] = 1;
        }
        std::vector<i64> box;
        for (i64 u = 0; u < g.N; u++)
            if (mark[u]) {
                box.push_back(u);
                if (counter[u] != FROZEN) d[u] = Arith<T>::snap(f[u], eps);
                counter[u] = FROZEN;
            }
        for (;;) {
            std::vector<i64> raise;
            for (i64 u : box) {
                const i64 x = u % g.W, y = (u / g.W) % g.H, z = u / (g.W * g.H);
                for (int k = 0; k < 14; k++) {
                    const i64 xx = x + DIRS[k][0], yy = y + DIRS[k][1], zz = z + DIRS[k][2];
                    if (xx < 0 || yy < 0 || zz < 0 || xx >= g.W || yy >= g.H || zz >= g.D) continue;
                    const i64 w = (zz * g.H + yy) * g.W + xx;
                    if (d[w] == d[u] && f[w] < f[u] && w > u) {
                        raise.push_back(u);
                        break;
                    }
                }
            }
            if (raise.empty()) break;
            for (i64 u : raise) d[u] = std::nextafter(d[u], std::numeric_limits<T>::infinity());
        }
        return (i64)box.size();
    }

    void pin(i64 v) {
        d[v] = Arith<T>::snap(f[v], eps);
        counter[v] = 6;
    }
    void filtration() {
        for (i64 v = 0; v < g.N; v++)
            if (pv[0][v] == -1) pin(v);
        for (i64 e = 0; e < g.E; e++)
            if (pe[0][e] == -1) {
                i64 a, b;
                edge_verts(g, e, a, b);
                pin(a);
                pin(b);
            }
        for (i64 t = 0; t < g.F; t++)
            if (pf[0][t] == -1) {
                i64 a, b, c;
                tri_verts(g, t, a, b, c);
                pin(a);
                pin(b);
                pin(c);
            }
        for (i64 q = 0; q < g.T; q++)
            if (pt[0][q] == -1) {
                i64 a, b, c, e;
                tet_verts(g, q, a, b, c, e);
                pin(a);
                pin(b);
                pin(c);
                pin(e);
            }
    }

    template <class Fn> void for_list(const std::vector<i64>& l, Fn fn) {
        i64 n = (i64)l.size();
#pragma omp parallel for schedule(dynamic, 256)
        for (i64 k = 0; k < n; k++) fn(l[k]);
    }

    void clear_targets() { std::fill(target.begin(), target.end(), 0); }

    void detect_paths() {
        if (fl.preserve_geometry == 1) {
            detect_as_geometry();
            detect_ds_geometry();
        } else if (fl.preserve_vpath == 1) {
            detect_as_connectivity();
            detect_ds_connectivity();
        }
        if (fl.preserve_connectors == 1) detect_ds_wall();
        else l_dsw.clear();
    }
};

}  // namespace orc

using namespace orc;

// ================================================================================================
// C ABI for ctypes
// ================================================================================================
typedef struct {
    int32_t preserve_saddles, preserve_vpath, preserve_geometry, filtration, preserve_connectors,
        preserve_type, edit_type, max_sweeps, finish;
} orc_flags;

// Branch-hit counters: enable = 1 clears them and starts counting, 0 stops. out (nullable,
// cap entries) receives the counts; returns the number of counters.
extern "C" int orc_coverage(int enable, long long* out, int cap) {
    if (enable == 1) {
        for (int k = 0; k < C_NCOV; k++) cov_hits[k].store(0);
        cov_on.store(true);
    } else if (enable == 0) {
        cov_on.store(false);
    }
    if (out)
        for (int k = 0; k < C_NCOV && k < cap; k++) out[k] = cov_hits[k].load();
    return C_NCOV;
}
extern "C" const char* orc_coverage_name(int k) { return (k >= 0 && k < C_NCOV) ? COV_NAMES[k] : ""; }

extern "C" int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

extern "C" void orc_counts(i64 W, i64 H, i64 D, i64* out4) {
    Grid g(W, H, D);
    out4[0] = g.N;
    out4[1] = g.E;
    out4[2] = g.F;
    out4[3] = g.T;
}

extern "C" i64 orc_edge_id(i64 W, i64 H, i64 D, i64 a, i64 b) {
    init_dirs();
    Grid g(W, H, D);
    i64 x, y, z, x1, y1, z1;
    g.xyz(a, x, y, z);
    g.xyz(b, x1, y1, z1);
    return edge_id(g, x, y, z, x1, y1, z1);
}
extern "C" void orc_edge_verts(i64 W, i64 H, i64 D, i64 e, i64* out2) {
    Grid g(W, H, D);
    edge_verts(g, e, out2[0], out2[1]);
}
extern "C" i64 orc_tri_id(i64 W, i64 H, i64 D, i64 a, i64 b, i64 c) {
    Grid g(W, H, D);
    return tri_id(g, a, b, c);
}
extern "C" void orc_tri_verts(i64 W, i64 H, i64 D, i64 t, i64* out3) {
    Grid g(W, H, D);
    tri_verts(g, t, out3[0], out3[1], out3[2]);
}
extern "C" i64 orc_tet_id(i64 W, i64 H, i64 D, i64 a, i64 b, i64 c, i64 d) {
    Grid g(W, H, D);
    return tet_id(g, a, b, c, d);
}
extern "C" void orc_tet_verts(i64 W, i64 H, i64 D, i64 q, i64* out4) {
    Grid g(W, H, D);
    tet_verts(g, q, out4[0], out4[1], out4[2], out4[3]);
}
extern "C" void orc_cell_tetra(i64 W, i64 H, i64 D, i64 t, int literal, i64* out2) {
    init_dirs();
    Grid g(W, H, D);
    cell_tetra(g, t, out2[0], out2[1], literal != 0);
}

// Gradient of one field. pv/pe/pf/pt: caller arrays of N/E/F/T int64.
template <class T>
static int gradient_impl(const T* fld, i64 W, i64 H, i64 D, int literal, i64* pv, i64* pe, i64* pf, i64* pt) {
    init_dirs();
    State<T> s(W, H, D);
    s.f = fld;
    s.d = const_cast<T*>(fld);
    s.literal = literal != 0;
    for (int w = 0; w < 2; w++) {
        s.pv[w].assign(s.g.N, -1);
        s.pe[w].assign(s.g.E, -1);
        s.pf[w].assign(s.g.F, -1);
        s.pt[w].assign(s.g.T, -1);
    }
    s.gradient(F_FIELD);
    std::memcpy(pv, s.pv[0].data(), sizeof(i64) * s.g.N);
    std::memcpy(pe, s.pe[0].data(), sizeof(i64) * s.g.E);
    std::memcpy(pf, s.pf[0].data(), sizeof(i64) * s.g.F);
    std::memcpy(pt, s.pt[0].data(), sizeof(i64) * s.g.T);
    return 0;
}
extern "C" int orc_gradient_f32(const float* fld, i64 W, i64 H, i64 D, int literal, i64* pv, i64* pe, i64* pf, i64* pt) {
    return gradient_impl<float>(fld, W, H, D, literal, pv, pe, pf, pt);
}
extern "C" int orc_gradient_f64(const double* fld, i64 W, i64 H, i64 D, int literal, i64* pv, i64* pe, i64* pf, i64* pt) {
    return gradient_impl<double>(fld, W, H, D, literal, pv, pe, pf, pt);
}

// One fix phase on (f, g) without applying: targets[N] receives 0/1, counts4 the mismatch counts.
template <class T>
static int targets_impl(const T* f, const T* gfield, i64 W, i64 H, i64 D, double eps, double range,
                        const orc_flags* fl, uint8_t* targets, i64* counts4) {
    init_dirs();
    State<T> s(W, H, D);
    s.f = f;
    std::vector<T> gcopy(gfield, gfield + s.g.N);
    s.d = gcopy.data();
    s.eps = eps;
    s.range = range;
    std::memcpy(&s.fl, fl, sizeof(orc_flags));
    s.literal = false;
    s.alloc();
    s.gradient(F_FIELD);
    s.gradient(G_FIELD);
    s.detect_min();
    s.detect_s1();
    s.detect_s2();
    s.detect_max();
    counts4[0] = (i64)s.l_min.size();
    counts4[1] = (i64)s.l_s1.size();
    counts4[2] = (i64)s.l_s2.size();
    counts4[3] = (i64)s.l_max.size();
    s.for_list(s.l_min, [&](i64 id) { s.fix_minimum(id); });
    if (fl->preserve_saddles) {
        s.for_list(s.l_s1, [&](i64 id) { s.fix_saddle(id); });
        s.for_list(s.l_s2, [&](i64 id) { s.fix_2saddle(id); });
    }
    s.for_list(s.l_max, [&](i64 id) { s.fix_maximum(id); });
    std::memcpy(targets, s.target.data(), s.g.N);
    return 0;
}
// The repair rule of one mismatched cell on (f, g) as they stand: type 0..3 selects fix_minimum,
// fix_saddle, fix_2saddle or fix_maximum and id the cell (reference simplex ID); targets[N]
// receives its edit set. Returns 1 when the cell is in that type's mismatch list, 0 when not.
static int fix_one_impl(const float* f, const float* gfield, i64 W, i64 H, i64 D, double eps, double range,
                        int edit_type, int type, i64 id, uint8_t* targets) {
    init_dirs();
    State<float> s(W, H, D);
    s.f = f;
    std::vector<float> gcopy(gfield, gfield + s.g.N);
    s.d = gcopy.data();
    s.eps = eps;
    s.range = range;
    std::memset(&s.fl, 0, sizeof(orc_flags));
    s.fl.preserve_saddles = 1;
    s.fl.edit_type = edit_type;
    s.literal = false;
    s.alloc();
    s.gradient(F_FIELD);
    s.gradient(G_FIELD);
    s.detect_min();
    s.detect_s1();
    s.detect_s2();
    s.detect_max();
    const std::vector<i64>* lists[4] = {&s.l_min, &s.l_s1, &s.l_s2, &s.l_max};
    if (type < 0 || type > 3) return -1;
    const std::vector<i64>& l = *lists[type];
    const bool listed = std::find(l.begin(), l.end(), id) != l.end();
    if (listed) {
        if (type == 0) s.fix_minimum(id);
        else if (type == 1) s.fix_saddle(id);
        else if (type == 2) s.fix_2saddle(id);
        else s.fix_maximum(id);
    }
    std::memcpy(targets, s.target.data(), s.g.N);
    return listed ? 1 : 0;
}
extern "C" int orc_fix_one_f32(const float* f, const float* g, i64 W, i64 H, i64 D, double eps, double range,
                               int edit_type, int type, i64 id, uint8_t* targets) {
    return fix_one_impl(f, g, W, H, D, eps, range, edit_type, type, id, targets);
}
extern "C" int orc_targets_f32(const float* f, const float* g, i64 W, i64 H, i64 D, double eps, double range,
                    const orc_flags* fl, uint8_t* targets, i64* counts4) {
    return targets_impl<float>(f, g, W, H, D, eps, range, fl, targets, counts4);
}
extern "C" int orc_targets_f64(const double* f, const double* g, i64 W, i64 H, i64 D, double eps, double range,
                    const orc_flags* fl, uint8_t* targets, i64* counts4) {
    return targets_impl<double>(f, g, W, H, D, eps, range, fl, targets, counts4);
}

// Path mismatch lists on (f, g) as they stand (no fix): out lists sized by caller (E for ds,
// T for as, F for walls); counts3 = {ds, as, ds_wall}; index arrays written as the reference does.
template <class T>
static int paths_impl(const T* f, const T* gfield, i64 W, i64 H, i64 D, const orc_flags* fl,
                      i64* counts3, i64* ds_list, i64* ds_rec, i64* as_list, i64* as_rec, i64* dsw_list,
                      i64* dsw_rec) {
    init_dirs();
    State<T> s(W, H, D);
    s.f = f;
    std::vector<T> gcopy(gfield, gfield + s.g.N);
    s.d = gcopy.data();
    s.eps = 0;
    s.range = 0;
    std::memcpy(&s.fl, fl, sizeof(orc_flags));
    s.literal = false;
    s.alloc();
    s.gradient(F_FIELD);
    s.gradient(G_FIELD);
    if (fl->preserve_vpath == 1 && fl->preserve_geometry == 0) s.compute_labels();
    s.detect_paths();
    counts3[0] = (i64)s.l_ds.size();
    counts3[1] = (i64)s.l_as.size();
    counts3[2] = (i64)s.l_dsw.size();
    for (size_t k = 0; k < s.l_ds.size(); k++) {
        if (ds_list) ds_list[k] = s.l_ds[k];
        if (ds_rec) ds_rec[k] = s.ds_index[2 * s.l_ds[k]];
    }
    for (size_t k = 0; k < s.l_as.size(); k++) {
        if (as_list) as_list[k] = s.l_as[k];
        if (as_rec) as_rec[k] = s.as_index[2 * s.l_as[k]];
    }
    for (size_t k = 0; k < s.l_dsw.size(); k++) {
        if (dsw_list) dsw_list[k] = s.l_dsw[k];
        if (dsw_rec) dsw_rec[k] = s.dsw_index[s.l_dsw[k]];
    }
    return (int)(s.walk_overflow > 0);
}
extern "C" int orc_paths_f32(const float* f, const float* g, i64 W, i64 H, i64 D, const orc_flags* fl, i64* counts3,
                  i64* ds_list, i64* ds_rec, i64* as_list, i64* as_rec, i64* dsw_list, i64* dsw_rec) {
    return paths_impl<float>(f, g, W, H, D, fl, counts3, ds_list, ds_rec, as_list, as_rec, dsw_list, dsw_rec);
}
extern "C" int orc_paths_f64(const double* f, const double* g, i64 W, i64 H, i64 D, const orc_flags* fl, i64* counts3,
                  i64* ds_list, i64* ds_rec, i64* as_list, i64* as_rec, i64* dsw_list, i64* dsw_rec) {
    return paths_impl<double>(f, g, W, H, D, fl, counts3, ds_list, ds_rec, as_list, as_rec, dsw_list, dsw_rec);
}

// Full correction loop, main_3d.cu:6521-7149.
// counts8 = {min, saddle1, saddle2, max, ds_vpath, as_vpath, ds_wall, as_wall} at exit.
// traj (nullable) receives 8 counts per sweep, up to traj_cap sweeps.
// info4 (nullable) receives {stalls, vertices changed by the finisher, finisher radius, 0}.
// Returns 0 converged, 1 not converged within max_sweeps, 2 stalled (a fixed point with
// mismatches left; only without the finisher), -2 a capped walk.
//
// Stall rule: a sweep is a function of (g, counter, counts, sweep kind). When a sweep changes no
// value or counter and its (counts, kind) repeats one seen since the last change, every later
// sweep repeats the same cycle: the reference never leaves it.
template <class T>
static int correct_impl(const T* f, T* gfield, i64 W, i64 H, i64 D, double eps, double range, const orc_flags* fl,
                        i64* counts8, int32_t* counter_out, i64* sweeps_out, i64* traj, i64 traj_cap,
                        double* phase_s, i64* info4) {
    init_dirs();
    State<T> s(W, H, D);
    s.f = f;
    s.d = gfield;
    s.eps = eps;
    s.range = range;
    std::memcpy(&s.fl, fl, sizeof(orc_flags));
    s.literal = false;
    s.alloc();
    double ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // gradient, get_cp, get_path, fix_cp, fix_path
    auto now = []() {
#ifdef _OPENMP
        return omp_get_wtime();
#else
        return 0.0;
#endif
    };
    double t0 = now();
    s.gradient(F_FIELD);
    if (fl->preserve_vpath == 1 && fl->preserve_geometry == 0) s.compute_labels();
    if (fl->filtration == 1) s.filtration();
    std::fill(s.counter.begin(), s.counter.end(), 1);  // init_counter main_3d.cu:6782
    s.gradient(G_FIELD);
    double t1 = now();
    ph[0] += t1 - t0;
    s.detect_min();
    s.detect_s1();
    s.detect_s2();
    s.detect_max();
    ph[1] += now() - t1;
    i64 c[8] = {(i64)s.l_min.size(), (i64)s.l_s1.size(), (i64)s.l_s2.size(), (i64)s.l_max.size(), 1, 1, 1, 0};
    i64 sweeps = 0;
    int status = 0;
    auto record = [&]() {
        if (traj && sweeps - 1 < traj_cap)
            for (int k = 0; k < 8; k++) traj[(sweeps - 1) * 8 + k] = c[k];
    };
    // Re-evaluation after the apply of a critical-point sweep (main_3d.cu:6864-6921).
    auto eval_cp = [&]() {
        double tb = now();
        s.gradient(G_FIELD);
        double tc = now();
        ph[0] += tc - tb;
        s.detect_min();
        if (fl->preserve_saddles == 1) {
            s.detect_s1();
            s.detect_s2();
        } else {
            s.l_s1.clear();
            s.l_s2.clear();
        }
        s.detect_max();
        ph[1] += now() - tc;
        c[0] = (i64)s.l_min.size();
        c[1] = (i64)s.l_s1.size();
        c[2] = (i64)s.l_s2.size();
        c[3] = (i64)s.l_max.size();
    };
    // Re-evaluation after the apply of a separatrix sweep (main_3d.cu:7069-7143).
    auto eval_path = [&]() {
        double tc = now();
        s.gradient(G_FIELD);
        double td = now();
        ph[0] += td - tc;
        s.detect_min();
        s.detect_s1();
        s.detect_s2();
        s.detect_max();
        c[0] = (i64)s.l_min.size();
        c[1] = (i64)s.l_s1.size();
        c[2] = (i64)s.l_s2.size();
        c[3] = (i64)s.l_max.size();
        double te = now();
        ph[1] += te - td;
        c[4] = c[5] = c[6] = c[7] = 0;
        if (c[0] == 0 && c[1] == 0 && c[2] == 0 && c[3] == 0) {
            s.detect_paths();
            c[4] = (i64)s.l_ds.size();
            c[5] = (i64)s.l_as.size();
            c[6] = (i64)s.l_dsw.size();
            ph[2] += now() - te;
        }
    };
    std::vector<std::vector<i64>> seen;  // (counts, kind) since the last change
    std::vector<i64> zone;               // edit sets of the unchanged sweeps since the last change
    i64 stalls = 0, finished = 0;
    int radius = 0;
    // After the re-evaluation of a sweep of the given kind (0 critical points, 1 separatrices):
    // returns true when the run stalls without the finisher.
    auto after_sweep = [&](i64 moved, int kind) -> bool {
        std::vector<i64> key(c, c + 8);
        key.push_back(kind);
        if (moved > 0) {
            seen.assign(1, key);
            zone.clear();
            return false;
        }
        zone.insert(zone.end(), s.last_edits.begin(), s.last_edits.end());
        bool any = false;
        for (int k = 0; k < 8; k++) any = any || c[k] != 0;
        if (!any) return false;
        if (std::find(seen.begin(), seen.end(), key) == seen.end()) {
            seen.push_back(key);
            return false;
        }
        stalls++;
        if (fl->finish != 1) return true;
        radius = radius ? 2 * radius : 1;  // 1, 2, 4, ...: a run that keeps stalling ends soon
        s.mismatch_vertices(c, zone);  // the zone: stalled edit sets + vertices of mismatched cells
        finished += s.finish_step(zone, radius);
        zone.clear();
        if (kind == 0) eval_cp();
        else eval_path();
        key.assign(c, c + 8);
        key.push_back(kind);
        seen.assign(1, key);
        return false;
    };
    while (c[5] > 0 || c[4] > 0 || c[6] > 0 || c[2] > 0 || c[3] > 0 || c[0] > 0 || c[1] > 0) {
        while (c[3] > 0 || c[2] > 0 || c[0] > 0 || c[1] > 0) {
            if (sweeps >= fl->max_sweeps) {
                status = 1;
                goto done;
            }
            sweeps++;
            double ta = now();
            s.clear_targets();
            if (c[0] > 0) s.for_list(s.l_min, [&](i64 id) { s.fix_minimum(id); });
            if (c[1] > 0 && fl->preserve_saddles == 1) s.for_list(s.l_s1, [&](i64 id) { s.fix_saddle(id); });
            if (c[2] > 0 && fl->preserve_saddles == 1) s.for_list(s.l_s2, [&](i64 id) { s.fix_2saddle(id); });
            if (c[3] > 0) s.for_list(s.l_max, [&](i64 id) { s.fix_maximum(id); });
            i64 moved = s.apply();
            ph[3] += now() - ta;
            eval_cp();
            bool stall = after_sweep(moved, 0);
            record();
            if (stall) {
                status = 2;
                goto done;
            }
        }
        if (fl->preserve_vpath == 0) break;
        if (sweeps >= fl->max_sweeps) {
            status = 1;
            goto done;
        }
        sweeps++;
        {
            double ta = now();
            s.detect_paths();
            double tb = now();
            ph[2] += tb - ta;
            i64 nas = (i64)s.l_as.size(), nds = (i64)s.l_ds.size(), ndsw = (i64)s.l_dsw.size();
            s.clear_targets();
            if (nas > 0) s.for_list(s.l_as, [&](i64 id) { s.fix_as_vpath(id); });
            if (nds > 0) s.for_list(s.l_ds, [&](i64 id) { s.fix_ds_vpath(id); });
            if (ndsw > 0 && fl->preserve_connectors == 1) s.for_list(s.l_dsw, [&](i64 id) { s.fix_ds_wall(id); });
            i64 moved = s.apply();
            ph[4] += now() - tb;
            eval_path();
            bool stall = after_sweep(moved, 1);
            record();
            if (stall) {
                status = 2;
                goto done;
            }
        }
    }
done:
    if (fl->preserve_vpath == 0) c[4] = c[5] = c[6] = c[7] = 0;
    for (int k = 0; k < 8; k++) counts8[k] = c[k];
    if (counter_out)
        for (i64 v = 0; v < s.g.N; v++) counter_out[v] = s.counter[v];
    if (sweeps_out) *sweeps_out = sweeps;
    if (phase_s)
        for (int k = 0; k < 5; k++) phase_s[k] = ph[k];
    if (info4) {
        info4[0] = stalls;
        info4[1] = finished;
        info4[2] = radius;
        info4[3] = 0;
    }
    if (s.walk_overflow > 0) return -2;
    return status;
}
extern "C" int orc_correct_f32(const float* f, float* g, i64 W, i64 H, i64 D, double eps, double range, const orc_flags* fl,
                    i64* counts8, int32_t* counter_out, i64* sweeps_out, i64* traj, i64 traj_cap, double* phase_s,
                    i64* info4) {
    return correct_impl<float>(f, g, W, H, D, eps, range, fl, counts8, counter_out, sweeps_out, traj, traj_cap, phase_s,
                               info4);
}
extern "C" int orc_correct_f64(const double* f, double* g, i64 W, i64 H, i64 D, double eps, double range, const orc_flags* fl,
                    i64* counts8, int32_t* counter_out, i64* sweeps_out, i64* traj, i64 traj_cap, double* phase_s,
                    i64* info4) {
    return correct_impl<double>(f, g, W, H, D, eps, range, fl, counts8, counter_out, sweeps_out, traj, traj_cap,
                                phase_s, info4);
}


// Separatrix detection + repair phase on (f, g) without applying: the edit set of
// fix_as_vpath / fix_ds_vpath (main_3d.cu:5135-5295) over the lists of detect_paths.
template <class T>
static int path_targets_impl(const T* f, const T* gfield, i64 W, i64 H, i64 D, double eps, double range,
                             const orc_flags* fl, uint8_t* targets, i64* counts2) {
    init_dirs();
    State<T> s(W, H, D);
    s.f = f;
    std::vector<T> gcopy(gfield, gfield + s.g.N);
    s.d = gcopy.data();
    s.eps = eps;
    s.range = range;
    std::memcpy(&s.fl, fl, sizeof(orc_flags));
    s.literal = false;
    s.alloc();
    s.gradient(F_FIELD);
    s.gradient(G_FIELD);
    if (fl->preserve_vpath == 1 && fl->preserve_geometry == 0) s.compute_labels();
    s.detect_paths();
    counts2[0] = (i64)s.l_ds.size();
    counts2[1] = (i64)s.l_as.size();
    s.for_list(s.l_as, [&](i64 id) { s.fix_as_vpath(id); });
    s.for_list(s.l_ds, [&](i64 id) { s.fix_ds_vpath(id); });
    std::memcpy(targets, s.target.data(), s.g.N);
    return 0;
}
extern "C" int orc_path_targets_f32(const float* f, const float* g, i64 W, i64 H, i64 D, double eps, double range,
                                    const orc_flags* fl, uint8_t* targets, i64* counts2) {
    return path_targets_impl<float>(f, g, W, H, D, eps, range, fl, targets, counts2);
}

// Wall list on (f, g) as they stand and the edit set of fix_ds_wall over it (the wall step of the
// path sweep, main_3d.cu:7010-7016 with preserve_connectors set).
template <class T>
static int wall_targets_impl(const T* f, const T* gfield, i64 W, i64 H, i64 D, double eps, double range,
                             const orc_flags* fl, uint8_t* targets, i64* count) {
    init_dirs();
    State<T> s(W, H, D);
    s.f = f;
    std::vector<T> gcopy(gfield, gfield + s.g.N);
    s.d = gcopy.data();
    s.eps = eps;
    s.range = range;
    std::memcpy(&s.fl, fl, sizeof(orc_flags));
    s.fl.preserve_connectors = 1;
    s.literal = false;
    s.alloc();
    s.gradient(F_FIELD);
    s.gradient(G_FIELD);
    s.detect_ds_wall();
    *count = (i64)s.l_dsw.size();
    s.for_list(s.l_dsw, [&](i64 id) { s.fix_ds_wall(id); });
    std::memcpy(targets, s.target.data(), s.g.N);
    return 0;
}
extern "C" int orc_wall_targets_f32(const float* f, const float* g, i64 W, i64 H, i64 D, double eps, double range,
                                    const orc_flags* fl, uint8_t* targets, i64* count) {
    return wall_targets_impl<float>(f, g, W, H, D, eps, range, fl, targets, count);
}
extern "C" int orc_wall_targets_f64(const double* f, const double* g, i64 W, i64 H, i64 D, double eps, double range,
                                    const orc_flags* fl, uint8_t* targets, i64* count) {
    return wall_targets_impl<double>(f, g, W, H, D, eps, range, fl, targets, count);
}

// ---------------------------------------------------------------------------------------------
// Edit streams of the reference's cost reports, restated byte by byte (main_3d.cu:1399-1757).

// encodeTo3BitBitmap main_3d.cu:1446-1476: values 0..6, three bits each, least significant bit
// first; a byte is pushed when 8 bits are filled, the spilled high bits of the value start the
// next byte, and a partial last byte is pushed. Returns the number of bytes, or -1 when a value
// lies outside 0..6 (the reference prints an error and returns an empty bitmap), or -2 when out
// is too small.
extern "C" i64 orc_encode3(const int32_t* data, i64 n, uint8_t* out, i64 cap) {
    i64 len = 0;
    uint8_t cur = 0;
    int bit = 0;
    for (i64 i = 0; i < n; i++) {
        const int v = data[i];
        if (v < 0 || v > 6) return -1;
        cur = (uint8_t)(cur | (uint8_t)(v << bit));  // the uint8 store drops bits past 7 (:1459)
        bit += 3;
        if (bit >= 8) {
            if (len >= cap) return -2;
            out[len++] = cur;
            bit -= 8;
            cur = (uint8_t)(v >> (3 - bit));  // :1466
        }
    }
    if (bit > 0) {
        if (len >= cap) return -2;
        out[len++] = cur;
    }
    return len;
}

// runLengthEncodeBitmap main_3d.cu:1399-1426: the bits of every byte from the most significant
// down, as (bit, run length) pairs; an empty bitmap gives no pairs. Writes up to cap pairs as
// int32 (bit, count) and returns the number of runs.
extern "C" i64 orc_rle_bitmap(const uint8_t* bm, i64 nbytes, int32_t* pairs, i64 cap) {
    if (nbytes <= 0) return 0;
    i64 runs = 0;
    int cur = (bm[0] >> 7) & 1;
    i64 count = 0;
    auto emit = [&](int b, i64 c) {
        if (runs < cap) {
            pairs[2 * runs] = b;
            pairs[2 * runs + 1] = (int32_t)c;
        }
        runs++;
    };
    for (i64 j = 0; j < nbytes; j++)
        for (int i = 7; i >= 0; --i) {
            const int b = (bm[j] >> i) & 1;
            if (b == cur) {
                count++;
            } else {
                emit(cur, count);
                cur = b;
                count = 1;
            }
        }
    emit(cur, count);
    return runs;
}

// The streams of cost main_3d.cu:1494-1510 on float64 fields: index[i] = counter[i] where g1
// differs from g0 (0 elsewhere), and one value per index 6, in vertex order,
// -(g0 - (f - bound)). Returns the number of changed vertices; *n_edits = values written (at most
// edits_cap).
extern "C" i64 orc_cost_streams(const double* f, const double* g0, const double* g1, const uint8_t* counter, i64 n,
                                double bound, int32_t* index, double* edits, i64 edits_cap, i64* n_edits) {
    i64 cnt = 0, m = 0;
    for (i64 i = 0; i < n; i++) {
        index[i] = 0;
        if (g0[i] != g1[i]) {
            index[i] = counter[i];
            if (index[i] == 6) {
                if (m < edits_cap) edits[m] = -(g0[i] - (f[i] - bound));
                m++;
            }
            cnt++;
        }
    }
    *n_edits = m;
    return cnt;
}

// The streams of original_cost main_3d.cu:1649-1671: the changed vertices' indices as int32
// differences (the first index, then the gap to the one before) and g1 - g0 per changed vertex.
// Returns the number of changed vertices (entries written: at most cap).
extern "C" i64 orc_original_cost_streams(const double* g0, const double* g1, i64 n, int32_t* diffs, double* edits,
                                         i64 cap) {
    i64 m = 0;
    double prev = 0;
    for (i64 i = 0; i < n; i++) {
        if (g0[i] != g1[i]) {
            // the reference keeps the indices in a vector<double> and narrows each difference to int
            const double idx = (double)i;
            if (m < cap) {
                diffs[m] = (int32_t)(m == 0 ? idx : idx - prev);
                edits[m] = g1[i] - g0[i];
            }
            prev = idx;
            m++;
        }
    }
    return m;
}
