// nb_internal.h — shared declarations of the gfx950 N-body library.
//
// Device data layout (one context, N bodies):
//   bodies   pos double2[N], vel double2[N], mass double[N], orig u32[N], two
//            copies (ping-pong). Barnes-Hut keeps them in Morton order across
//            steps, as BarnesHutCuda does (BarnesHutCuda.cu:701-708);
//            orig[i] is the caller's index of sorted body i.
//   sort     keys u64[N] x2, vals u32[N] x2, per-tile digit histograms.
//   tree     NodeRec[S + 4*I] slots (S = systems, I = internal nodes): slot s
//            is the root of system s, and internal node p (numbered in
//            pre-order) owns the four child slots S+4p .. S+4p+3, the layout
//            of the reference's contiguous children (BarnesHutCuda.cu:212-221)
//            with a deterministic numbering. One system: root 0, children
//            1+4p .. 4+4p.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/nbody.h"
#include "nb_mlp.h"

namespace nb {

constexpr int kWave = 64;
constexpr int kMaxLevel = 30;          // level-30 nodes are leaves (BarnesHutCuda.cu:182)
constexpr int kComMargin = 2;          // levels past the deepest non-empty one kept as launches
constexpr int kComRefit = 2048;        // a deeper level of more nodes has the cut picked again
constexpr int kStackDepth = 128;       // >= 3*30+1, the deepest masked DFS stack
constexpr int kSortBits = 8;           // radix digit width
constexpr int kSortBins = 1 << kSortBits;
constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;         // keys per thread per tile
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kSortMaxPasses = 10;     // 8 key passes + 2 system-number passes
// sort_glob words: per-pass digit histograms, per-pass tile counters, epoch
constexpr int kSortGlobTile = kSortMaxPasses * kSortBins;
constexpr int kSortGlobEpoch = kSortGlobTile + 16;
constexpr int kSortGlobScanEpoch = kSortGlobEpoch + 1;  // tag of launch_scan_u32's look-back
constexpr int kSortGlobWords = kSortGlobEpoch + 16;
constexpr int kBoxStride = 8;          // doubles of box data per system
constexpr int kBboxThreads = 256;      // threads of the box partial kernels (and the box-writing drift)
constexpr int kRefBoxTile = 2048;      // bodies per block of the REF_CPU box scan
// Group traversal: list entries after which a group is split and walked again
// (nb_params.group_budget = 0). Splits restart walks from the root, so a low
// budget costs more than it saves once a simulation has evolved: on a
// 512-system ensemble 16384 ran 2x faster per step than 8192 and 6.6x faster
// than 4096, and on a 1M-body disk the budget made no measurable difference.
constexpr int kGroupBudget = 16384;

// Meta words kept in device memory (u32 each). Words that every wavefront
// of a kernel updates or polls sit on 128-byte L2 lines of their own: atomics
// on a line serialise in L2, and would stall the reads of its other words.
enum Meta : int {
    META_INTERNAL = 0,   // internal nodes of the last build
    META_FLAGS = 1,      // FLAG_* bits of the current step
    META_STEPS_DONE = 2, // completed steps since the host last cleared it
    META_MULTI_LEAVES = 3, // multi-body leaves queued by the current build
    META_CELLS_NEEDED = 4, // most accepted cells of one group (learned far field)
    META_CELLS_TOTAL = 32, // accepted cells of all groups (learned far field), u64 in 32..33
    META_QUAD_TERMS = 34,  // near-cell quadrupole terms applied per body, u64 in 34..35
    META_COM_HI = 40,        // 1 + deepest non-empty tree level, the most of the steps since cleared
    META_COM_DEEP_MAX = 41,  // most nodes of one level the one-workgroup COM launch took
    META_QHEAD = 64,       // traversal work queue: entries pushed
    META_QTAIL = 96,       // traversal work queue: entries taken
    META_COUNT = 128
};
enum Flag : uint32_t {
    FLAG_TREE_CAPACITY = 1u,  // internal nodes exceed the allocated capacity
    FLAG_STACK = 2u,          // traversal stack overflow
    FLAG_NONFINITE = 4u,      // non-finite bounding box
    FLAG_CELL_CAPACITY = 8u,  // a group's accepted cells exceed its region
    FLAG_SORT = 16u,          // a radix-sort tile waited too long for its predecessors
};
// Flags that stop the rest of a step (the host grows the buffer and re-runs,
// or reports the error).
constexpr uint32_t kHaltFlags =
    FLAG_TREE_CAPACITY | FLAG_NONFINITE | FLAG_CELL_CAPACITY | FLAG_SORT;

// One tree slot, 64 bytes so one wave-uniform node fetch is a single 64-byte
// scalar load.
struct __attribute__((aligned(64))) NodeRec {
    double cx, cy;   // centre of mass
    double m;        // mass
    double wsq;      // squared cell side (BarnesHutCuda.cu:216-217)
    int32_t c0;      // first child slot (internal), -1 leaf (empty slots are not written)
    int32_t mask;    // internal: bit q set = child c0+q holds bodies
    int32_t bs, be;  // body range [bs, be] in sorted order
    float ixx, ixy, iyy;  // second moments about the COM
    float mf;        // mass as FP32 for the pair arithmetic
};
static_assert(sizeof(NodeRec) == 64, "NodeRec must be 64 bytes");

// A finished subgroup of a traversal group (learned far field): its box
// centre, its range of accepted cells in cell_slot (absolute indices) and
// the squared distance from the centre beyond which a cell's quadrupole term
// is taken at the centre (with its Jacobian) instead of at every body.
struct SubRec {
    double ox, oy;
    uint32_t start, count;
    double thr;
};

// Device-side constants of a step, read by the kernels so set_dt needs no
// graph rebuild.
struct StepConsts {
    double dt;
    double G;
    double G_dt;      // G * dt, as BarnesHutCuda.cu:814
    double theta_sq;  // theta * theta (BarnesHutCuda.cpp:8)
    double eps2;      // softening^2
};

struct Buffers {
    double2* pos[2];
    double2* vel[2];
    double* mass[2];
    uint32_t* orig[2];
    uint64_t* keys[2];
    uint32_t* vals[2];
    uint32_t* sort_glob;   // [kSortGlobWords] digit histogram of every pass, the
                           //     passes' tile counters, the sort epoch
    uint64_t* sort_status; // [tiles + groups][bins] digit counts / prefixes of the
                           //     tiles and tile groups published for the
                           //     look-back (tagged with epoch and pass)
    uint32_t* sort_part;   // [Morton workgroups][kSortMaxPasses * bins] digit counts
    double* box;           // per system [8]: x_min, x_max, y_min, y_max, root size^2 of
                           //     the opening test (side^2, or w^2 + h^2 for DIAG)
    uint32_t* sys_of;      // [N] system of the body at position i (positions of a
                           //     system never change: sorting permutes inside it)
    uint32_t* sys_off;     // [nsys + 1] first position of each system
    uint32_t* grp_first;   // [ngroups] first body of traversal group g
    uint32_t* grp_info;    // [ngroups] (system << 8) | bodies (1..64)
    double* bbox_part;     // [blocks][4]
    double2* ref_pos;      // [N] positions in the caller's order (REF_CPU cells' box)
    double4* ref_part;     // [ref blocks] {min x, min y, max x, max y} per block of it
    int8_t* delta;         // [N] common-prefix digit count of pair (i, i+1)
    uint32_t* cnt;         // [N] internal nodes owned by pair i, then its offset
    uint64_t* scan_status; // published tile totals / prefixes of launch_scan_u32
    uint32_t* inode_slot;  // [icap] slot of internal node p
    uint8_t* inode_level;  // [icap]
    uint8_t* inode_mask;   // [icap]
    NodeRec* nodes;        // [1 + 4*icap]
    float4* mom3;          // [1 + 4*icap] third moments about the COM {T_xxx, T_xxy,
                           //     T_xyy, T_yyy} of each slot (learned far field only)
    uint32_t* node_a;      // [icap] first body of internal node p
    uint4* inode_info;     // [icap] internal node p: {first body, last body, its wsq as two words}
    uint32_t* multi_list;  // [N/2 + 1] slots of multi-body leaves
    uint32_t* lvl_cnt;     // [kMaxLevel * ceil(icap/256)] level counts, then offsets
    uint4* by_level;       // [icap] internal nodes sorted by level: {id, slot, mask, 0}
    double2* acc;          // [N] accelerations (nb_compute_forces) / partial sums
    uint32_t* inter;       // [N] force terms per body
    double2* ap_part;      // all-pairs partial sums [split][N]
    uint32_t* meta;        // [META_COUNT]
    StepConsts* consts;
    // far-field network (allocated on first use)
    float* mlp_w;          // [4866] FP32 master parameters, canonical order
    float* mlp_m;          // [4866] Adam first moments
    float* mlp_v;          // [4866] Adam second moments
    void* mlp_tiles;       // BF16 A tiles in the per-lane MFMA layouts (kMlpTileBytes)
    float* mlp_bias;       // [144] b1, b2, b3, zero padding
    float* mlp_gpart;      // [parts][4866] per-workgroup partial gradients
    float* mlp_loss;       // [parts]
    uint32_t* pair_body;   // [pair_cap] training pairs: sample body
    uint32_t* pair_slot;   // [pair_cap] training pairs: cell slot
    float* pair_x;         // [pair_cap][8] features
    float* pair_t;         // [pair_cap][2] targets
    uint32_t* sample_off;  // [N + 2] per-sample pair counts, then offsets; [N + 1] total
    // learned far field in the step: per group of 64 bodies
    uint32_t* cell_slot;   // [groups][cell_cap] tree slots of the accepted cells
    uint32_t* cell_cnt;    // [groups] finished subgroups
    SubRec* subs;          // [groups][64]
    float4* body_w;        // [N] {offset from the subgroup centre (x, y), subgroup, -}
    uint64_t* queue;       // [queue_cap] traversal work queue (all zero between kernels)
};

}  // namespace nb

struct nb_ctx {
    nb_params p;
    uint64_t n = 0;
    uint32_t nsys = 1;           // independent systems (contiguous body ranges)
    uint32_t ngroups = 0;        // traversal groups: <= 64 bodies, inside one system
    int device = 0;
    uint32_t cus = 256;          // compute units of the device
    hipStream_t stream = nullptr;
    nb::Buffers b{};
    int cur = 0;                 // which copy holds the state
    uint64_t icap = 0;           // internal-node capacity
    uint32_t bbox_blocks = 0;
    uint32_t ref_blocks = 0;     // blocks of the REF_CPU box scan (kRefBoxTile bodies each)
    uint32_t sort_tiles = 0;
    uint32_t ap_split = 1;
    uint32_t ap_rows = 4;        // all-pairs bodies per thread (2, 4 or 8)
    uint32_t ap_split_max = 1;   // j splits the partial buffer holds
    uint64_t device_bytes = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};   // one step from copy v
    hipGraphExec_t graph4[2] = {nullptr, nullptr};  // four steps from copy v (ending on v)
    // COM reduction: levels >= com_deep in one single-workgroup launch, the
    // others one launch each (fixed in the captured steps)
    int com_deep = nb::kMaxLevel;
    // The drift also writes the box partials (bbox_part) of the positions it
    // makes, so a step's box needs only the final reduction. box_parts_ok: the
    // partials describe pos[cur] (a gather permutes the same positions, and a
    // halted drift leaves them unchanged). capturing: steps are being recorded
    // into graphs, whose replays are preceded by ensure_box_parts.
    bool box_parts_ok = false;
    bool capturing = false;
    uint64_t regrowths = 0;       // buffer regrowths after a halted step
    uint64_t graph_captures = 0;  // build_graphs calls that captured steps
    bool timing = false;
    hipEvent_t ev[2 * NB_PHASE_COUNT] = {};  // start/stop per phase
    uint32_t marked = 0;                     // phases recorded in the current step
    double phase_ms[NB_PHASE_COUNT] = {};
    uint64_t steps_timed = 0;
    uint64_t iteration = 0;
    double sim_time = 0.0;
    double dt = 0.0;
    uint64_t last_internal = 0;
    uint32_t* host_meta = nullptr;  // pinned copy of meta
    uint64_t pair_cap = 0;          // training-pair capacity
    uint64_t pair_count = 0;        // pairs held
    uint32_t train_parts = 0;       // partial-gradient rows allocated
    uint32_t train_step = 0;        // Adam step counter
    uint32_t cell_cap = 0;          // accepted-cell region per group
    uint32_t queue_cap = 0;         // traversal work-queue entries
    uint64_t last_cells = 0;        // accepted cells of the last nb_compute_forces
    uint64_t last_quad = 0;         // near-cell quadrupole terms of the same
    // asynchronous snapshots: two slots, used by tickets alternately
    struct Snapshot {
        double2* dev = nullptr;     // [2n] positions then velocities, original order
        double* host = nullptr;     // pinned [4n] copy of dev
        hipEvent_t ready = nullptr; // un-permute done (context stream)
        hipEvent_t done = nullptr;  // device-to-host copy done (copy stream)
        uint64_t ticket = 0;        // 0: never used
        uint64_t iteration = 0;     // steps taken when the snapshot was begun
    } snap[2];
    hipStream_t copy_stream = nullptr;
    uint64_t snap_next = 1;
    // A failed regrowth leaves device buffers missing: the context then only
    // serves its state (nb_get_state, snapshots) and every call that would
    // run a step returns this error.
    bool dead = false;
    std::string dead_msg;
    bool keep_mom3 = false;         // the tree keeps third moments (b.mom3)
    std::vector<uint32_t> sys_off_h;  // first body of each system, and n
    uint32_t fail_allocs = 0;       // test hook: allocations left to fail
    std::string err;
};

namespace nb {

// Group walk settings of a context (nb_params; 0 = the default).
inline int walk_stack(const nb_ctx* c) { return c->p.walk_stack > 0 ? c->p.walk_stack : 2048; }
// Default bodies per walk wavefront: 32 below kWalk32Below bodies with the
// monopole far field, where 64-body groups leave too few wavefronts per CU to
// hide the walk's dependent node loads; 64 above, and for the learned far
// field (whose subgroup bookkeeping is laid out per 64 bodies by default).
constexpr uint64_t kWalk32Below = 1ull << 19;
inline int walk_bodies(const nb_ctx* c) {
    if (c->p.walk_bodies > 0) return c->p.walk_bodies;
    return c->n < kWalk32Below && c->p.far_field != NB_FAR_LEARNED ? 32 : 64;
}
inline int list_flush(const nb_ctx* c) { return c->p.list_flush > 0 ? c->p.list_flush : 128; }
// Squared near radius of the learned walk's per-body quadrupole terms, in
// squared subgroup half-diagonals (< 0: every cell is near).
inline double quad_near_sq(const nb_ctx* c) {
    const int k = c->p.quad_near == 0 ? 8 : c->p.quad_near;
    return k < 0 ? -1.0 : (double)k * k;
}

// ---- launchers (nb_sort.hip) ----
void launch_bbox(nb_ctx* c, int src);
void launch_morton(nb_ctx* c, int src);
void launch_radix_sort(nb_ctx* c);
void launch_gather(nb_ctx* c, int src);
void launch_unpermute(nb_ctx* c, double2* out);
void launch_unpermute2(nb_ctx* c, double2* pos_out, double2* vel_out);
void launch_scan_u32(nb_ctx* c, uint32_t* data, uint32_t count, uint32_t* total_out);

// ---- launchers (nb_tree.hip) ----
void launch_tree_build(nb_ctx* c, int src);

// ---- launchers (nb_force.hip) ----
void launch_traverse(nb_ctx* c, int src, bool kick, bool count_step);
void launch_allpairs(nb_ctx* c, int src, bool kick, bool count_step);
void launch_drift(nb_ctx* c, int src, bool check, bool count_step);
// Whether the drift writes the box partials (one system, Morton cells).
bool box_from_drift(const nb_ctx* c);
// Computes the box partials of pos[src] (the state) unless they are current.
void ensure_box_parts(nb_ctx* c, int src);
// ---- far-field network (nb_mlp.hip) ----
void free_mlp(nb_ctx* c);
int prepare_mlp(nb_ctx* c);
int alloc_cells(nb_ctx* c, uint32_t cap);
int regroup_cells(nb_ctx* c, uint64_t old_groups);
void launch_mlp_apply(nb_ctx* c, int src, bool kick, bool count_step);
// Builds the tree of the current state (nb_capi.hip).
int build_tree(nb_ctx* c);
// Makes the tree keep third moments from the next build on (nb_capi.hip).
int ensure_mom3(nb_ctx* c);

hipError_t direct_sum_f64(uint64_t n, const double2* pos, const double* mass, double eps2,
                          uint64_t m, const int64_t* targets, double2* part, uint32_t split,
                          hipStream_t s);
hipError_t direct_sum_f64_finish(const double2* part, uint32_t split, uint64_t m, uint64_t n,
                                 const int64_t* targets, double G, double2* acc, hipStream_t s);

// Number of leading equal base-4 digits of two keys (<= kMaxLevel).
__device__ __forceinline__ int pair_delta(uint64_t a, uint64_t b) {
    const uint64_t x = a ^ b;
    if (x == 0) return kMaxLevel;
    const int d = __clzll(x) >> 1;
    return d < kMaxLevel ? d : kMaxLevel;
}

// Pair i of the sorted keys (bodies i and i + 1, i + 1 < n): delta_i (-1
// between two systems) and the number of internal nodes the pair owns, the
// levels (delta_{i-1}, min(delta_i, 29)] (see nb_tree.hip).
// HBM traffic: 8 B keys read (x3 from cache), 5 B written per pair.
__device__ __forceinline__ void tree_pair(uint32_t i, const uint64_t* __restrict__ keys,
                                          const uint32_t* __restrict__ sys_of,
                                          int8_t* __restrict__ delta,
                                          uint32_t* __restrict__ cnt) {
    const uint64_t k0 = keys[i], k1 = keys[i + 1];
    const uint32_t s0 = sys_of[i];
    const int d = s0 == sys_of[i + 1] ? pair_delta(k0, k1) : -1;
    const int prev = (i && sys_of[i - 1] == s0) ? pair_delta(keys[i - 1], k0) : -1;
    const int hi = d < kMaxLevel - 1 ? d : kMaxLevel - 1;
    delta[i] = (int8_t)d;
    cnt[i] = hi > prev ? (uint32_t)(hi - prev) : 0u;
}

}  // namespace nb
