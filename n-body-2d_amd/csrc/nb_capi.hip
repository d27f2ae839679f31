// nb_capi.hip — the C ABI (include/nbody.h): context lifetime, step
// sequencing with hipGraph replay, readback in original order, error codes.
//
// Step sequence (Barnes-Hut, kick-drift), everything on the context stream:
//   bbox -> Morton -> 8 radix passes -> gather -> tree delta/scan/nodes ->
//   30 COM levels -> traversal + kick -> drift
// (BarnesHutCuda::simulate, BarnesHutCuda.cpp:30-40, without the per-level
// host round trips or the per-step device-to-host copy of :632-646).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "nb_internal.h"

using namespace nb;

namespace {

thread_local std::string g_err;

// Default accepted-cell region per group of 64 bodies (learned far field);
// a step that needs more halts, and the region grows to twice the need.
constexpr uint32_t kInitialCells = 4096;

int fail(nb_ctx* c, int code, const std::string& msg) {
    if (c)
        c->err = msg;
    else
        g_err = msg;
    return code;
}

#define NB_HIP(c, expr)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((c), e_ == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP, \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                \
    } while (0)

template <class T>
hipError_t dalloc(nb_ctx* c, T** p, uint64_t count) {
    uint64_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 256;
    if (c->fail_allocs > 0) {  // nb_debug_fail_alloc
        c->fail_allocs--;
        return hipErrorOutOfMemory;
    }
    const hipError_t e = hipMalloc((void**)p, bytes);
    if (e == hipSuccess) c->device_bytes += bytes;
    return e;
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

bool is_bh(const nb_ctx* c) { return c->p.algorithm == NB_ALGO_BARNES_HUT; }
bool is_dk(const nb_ctx* c) { return c->p.integrator == NB_INTEG_DK; }

uint64_t tree_bytes(uint64_t icap, uint64_t nsys, bool mom3) {
    const uint64_t ng = (icap + 255) / 256;
    return (nsys + 4 * icap) * (sizeof(NodeRec) + (mom3 ? sizeof(float4) : 0)) +
           icap * (2 * sizeof(uint32_t) + 2 + 2 * sizeof(uint4)) + kMaxLevel * ng * sizeof(uint32_t);
}

// Third moments are kept for the learned far field's inputs only (contexts
// with far_field = LEARNED, and any context once it generated training pairs).
bool wants_mom3(const nb_ctx* c) { return c->keep_mom3; }

void free_tree(nb_ctx* c) {
    if (c->icap) c->device_bytes -= tree_bytes(c->icap, c->nsys, wants_mom3(c));
    dfree(c->b.nodes);
    dfree(c->b.mom3);
    dfree(c->b.inode_slot);
    dfree(c->b.inode_level);
    dfree(c->b.inode_mask);
    dfree(c->b.node_a);
    dfree(c->b.inode_info);
    dfree(c->b.by_level);
    dfree(c->b.lvl_cnt);
    c->icap = 0;
}

int alloc_tree(nb_ctx* c, uint64_t icap) {
    free_tree(c);
    const uint64_t before = c->device_bytes;
    NB_HIP(c, dalloc(c, &c->b.nodes, c->nsys + 4 * icap));
    if (wants_mom3(c)) NB_HIP(c, dalloc(c, &c->b.mom3, c->nsys + 4 * icap));
    NB_HIP(c, dalloc(c, &c->b.inode_slot, icap));
    NB_HIP(c, dalloc(c, &c->b.inode_level, icap));
    NB_HIP(c, dalloc(c, &c->b.inode_mask, icap));
    NB_HIP(c, dalloc(c, &c->b.node_a, icap));
    NB_HIP(c, dalloc(c, &c->b.inode_info, icap));
    NB_HIP(c, dalloc(c, &c->b.by_level, icap));
    NB_HIP(c, dalloc(c, &c->b.lvl_cnt, kMaxLevel * ((icap + 255) / 256)));
    c->device_bytes = before + tree_bytes(icap, c->nsys, wants_mom3(c));
    c->icap = icap;
    return NB_OK;
}

void destroy_graphs(nb_ctx* c) {
    for (auto& g : c->graph) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    for (auto& g : c->graph4) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

void free_snapshots(nb_ctx* c) {
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    for (auto& s : c->snap) {
        dfree(s.dev);
        if (s.host) (void)hipHostFree(s.host);
        if (s.ready) (void)hipEventDestroy(s.ready);
        if (s.done) (void)hipEventDestroy(s.done);
        s = nb_ctx::Snapshot{};
    }
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    c->copy_stream = nullptr;
}

// Snapshot slots, allocated on the first nb_snapshot_begin: all or nothing.
hipError_t alloc_snapshots(nb_ctx* c) {
    hipError_t e = hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking);
    for (auto& s : c->snap) {
        if (!e) e = hipEventCreateWithFlags(&s.ready, hipEventDisableTiming);
        if (!e) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        if (!e) e = dalloc(c, &s.dev, 2 * c->n);
        if (!e) e = hipHostMalloc((void**)&s.host, std::max<uint64_t>(32 * c->n, 256));
    }
    if (e) free_snapshots(c);
    return e;
}

void free_all(nb_ctx* c) {
    free_snapshots(c);
    destroy_graphs(c);
    free_tree(c);
    free_mlp(c);
    Buffers& b = c->b;
    for (int k = 0; k < 2; k++) {
        dfree(b.pos[k]);
        dfree(b.vel[k]);
        dfree(b.mass[k]);
        dfree(b.orig[k]);
        dfree(b.keys[k]);
        dfree(b.vals[k]);
    }
    dfree(b.sort_glob);
    dfree(b.sort_status);
    dfree(b.sort_part);
    dfree(b.box);
    dfree(b.bbox_part);
    dfree(b.ref_pos);
    dfree(b.ref_part);
    dfree(b.delta);
    dfree(b.cnt);
    dfree(b.scan_status);
    dfree(b.multi_list);
    dfree(b.queue);
    dfree(b.sys_of);
    dfree(b.sys_off);
    dfree(b.grp_first);
    dfree(b.grp_info);
    dfree(b.acc);
    dfree(b.inter);
    dfree(b.ap_part);
    dfree(b.meta);
    dfree(b.consts);
    for (auto& e : c->ev) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    if (c->host_meta) (void)hipHostFree(c->host_meta);
    c->host_meta = nullptr;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
}

int check_params(const nb_params* p, std::string* why) {
    if (!p) return *why = "params is NULL", NB_ERR_INVALID_ARG;
    if (!std::isfinite(p->G)) return *why = "G must be finite", NB_ERR_INVALID_ARG;
    if (!std::isfinite(p->dt) || p->dt <= 0.0)
        return *why = "dt must be finite and > 0", NB_ERR_INVALID_ARG;
    if (!std::isfinite(p->theta) || p->theta < 0.0)
        return *why = "theta must be finite and >= 0", NB_ERR_INVALID_ARG;
    if (!std::isfinite(p->softening) || p->softening < 0.0)
        return *why = "softening must be finite and >= 0", NB_ERR_INVALID_ARG;
    if (p->algorithm != NB_ALGO_BARNES_HUT && p->algorithm != NB_ALGO_ALL_PAIRS)
        return *why = "unknown algorithm", NB_ERR_INVALID_ARG;
    if (p->integrator != NB_INTEG_KD && p->integrator != NB_INTEG_DK)
        return *why = "unknown integrator", NB_ERR_INVALID_ARG;
    if (p->traversal != NB_TRAV_EXACT && p->traversal != NB_TRAV_GROUP)
        return *why = "unknown traversal", NB_ERR_INVALID_ARG;
    if (p->group_budget < 0) return *why = "group_budget must be >= 0", NB_ERR_INVALID_ARG;
    if (p->cell_capacity < 0) return *why = "cell_capacity must be >= 0", NB_ERR_INVALID_ARG;
    if (p->walk_stack != 0 && p->walk_stack != 1024 && p->walk_stack != 2048)
        return *why = "walk_stack must be 0, 1024 or 2048", NB_ERR_INVALID_ARG;
    if (p->walk_bodies != 0 && p->walk_bodies != 32 && p->walk_bodies != 64)
        return *why = "walk_bodies must be 0, 32 or 64", NB_ERR_INVALID_ARG;
    if (p->list_flush != 0 && (p->list_flush < 64 || p->list_flush > 192))
        return *why = "list_flush must be 0 or in 64 .. 192", NB_ERR_INVALID_ARG;
    if (p->quad_near < -1 || p->quad_near > 1024)
        return *why = "quad_near must be -1, 0 or in 1 .. 1024", NB_ERR_INVALID_ARG;
    if (p->far_field != NB_FAR_MONOPOLE && p->far_field != NB_FAR_LEARNED)
        return *why = "unknown far_field", NB_ERR_INVALID_ARG;
    if (p->force_kernel != NB_FORCE_PLUMMER && p->force_kernel != NB_FORCE_REF_CPU)
        return *why = "unknown force_kernel", NB_ERR_INVALID_ARG;
    if (p->mac != NB_MAC_SIDE && p->mac != NB_MAC_DIAG)
        return *why = "unknown mac", NB_ERR_INVALID_ARG;
    if (p->cells != NB_CELLS_SQUARE && p->cells != NB_CELLS_REF_CPU)
        return *why = "unknown cells", NB_ERR_INVALID_ARG;
    if (p->cells == NB_CELLS_REF_CPU && p->mac != NB_MAC_DIAG)
        return *why = "cells=REF_CPU needs mac=DIAG (its cells are rectangles)",
               NB_ERR_NOT_SUPPORTED;
    if (p->far_field == NB_FAR_LEARNED &&
        (p->algorithm != NB_ALGO_BARNES_HUT || p->traversal != NB_TRAV_GROUP))
        return *why = "far_field=LEARNED needs algorithm=BARNES_HUT and traversal=GROUP",
               NB_ERR_NOT_SUPPORTED;
    if (p->far_field == NB_FAR_LEARNED &&
        (p->force_kernel != NB_FORCE_PLUMMER || p->mac != NB_MAC_SIDE ||
         p->cells != NB_CELLS_SQUARE))
        return *why = "far_field=LEARNED needs the Plummer force, the SIDE test and SQUARE cells",
               NB_ERR_NOT_SUPPORTED;
    return NB_OK;
}

int upload_consts(nb_ctx* c) {
    StepConsts k;
    k.dt = c->dt;
    k.G = c->p.G;
    k.G_dt = c->p.G * c->dt;
    k.theta_sq = c->p.theta * c->p.theta;
    k.eps2 = c->p.softening * c->p.softening;
    NB_HIP(c, hipMemcpyAsync(c->b.consts, &k, sizeof(k), hipMemcpyHostToDevice, c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    return NB_OK;
}

// ---- phase marks for eager timing ----
struct Marks {
    nb_ctx* c;
    void begin(int ph) {
        if (!c->timing) return;
        (void)hipEventRecord(c->ev[2 * ph], c->stream);
        c->marked |= 1u << ph;
    }
    void end(int ph) {
        if (c->timing) (void)hipEventRecord(c->ev[2 * ph + 1], c->stream);
    }
};

// bbox -> Morton -> sort -> gather -> tree -> traversal (or all-pairs).
// Returns the buffer that holds the state afterwards.
int enqueue_force(nb_ctx* c, int src, bool kick, bool count_step) {
    Marks mk{c};
    int s = src;
    if (is_bh(c)) {
        mk.begin(NB_PHASE_BBOX_MORTON);
        launch_bbox(c, s);
        launch_morton(c, s);
        mk.end(NB_PHASE_BBOX_MORTON);
        mk.begin(NB_PHASE_SORT);
        launch_radix_sort(c);
        launch_gather(c, s);
        mk.end(NB_PHASE_SORT);
        s ^= 1;
        mk.begin(NB_PHASE_BUILD);
        launch_tree_build(c, s);
        mk.end(NB_PHASE_BUILD);
        mk.begin(NB_PHASE_FORCE);
        launch_traverse(c, s, kick, count_step);
        mk.end(NB_PHASE_FORCE);
        if (c->p.far_field == NB_FAR_LEARNED) {
            mk.begin(NB_PHASE_FAR_FIELD);
            launch_mlp_apply(c, s, kick, count_step);
            mk.end(NB_PHASE_FAR_FIELD);
        }
    } else {
        mk.begin(NB_PHASE_FORCE);
        launch_allpairs(c, s, kick, count_step);
        mk.end(NB_PHASE_FORCE);
    }
    return s;
}

int enqueue_step(nb_ctx* c, int src) {
    Marks mk{c};
    int s = src;
    if (is_dk(c)) {
        mk.begin(NB_PHASE_DRIFT);
        launch_drift(c, s, true, false);
        mk.end(NB_PHASE_DRIFT);
        s = enqueue_force(c, s, true, true);
    } else {
        s = enqueue_force(c, s, true, false);
        mk.begin(NB_PHASE_DRIFT);
        launch_drift(c, s, true, true);
        mk.end(NB_PHASE_DRIFT);
    }
    return s;
}

// Steps are replayed from captured graphs: one step from either copy of the
// body arrays, and four steps in one graph (which ends on the copy it started
// from), so that a multi-step call pays the gap between two graph replays
// (~8 us on MI355X) once per four steps.
constexpr int kGraphSteps = 4;

int build_graphs(nb_ctx* c) {
    destroy_graphs(c);
    if (!c->p.use_graph || c->n == 0) return NB_OK;
    c->graph_captures++;
    const bool timing = c->timing;
    c->timing = false;
    const int variants = is_bh(c) ? 2 : 1;
    for (int v = 0; v < variants; v++) {
        for (int multi = 0; multi < 2; multi++) {
            hipGraph_t g = nullptr;
            NB_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            c->capturing = true;
            int s = v;
            for (int k = 0; k < (multi ? kGraphSteps : 1); k++) s = enqueue_step(c, s);
            const hipError_t le = hipGetLastError();
            c->capturing = false;
            NB_HIP(c, hipStreamEndCapture(c->stream, &g));
            NB_HIP(c, le);
            hipGraphExec_t* dst = multi ? &c->graph4[v] : &c->graph[v];
            const hipError_t ie = hipGraphInstantiate(dst, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            NB_HIP(c, ie);
            // upload now, not at the first launch of the step
            NB_HIP(c, hipGraphUpload(*dst, c->stream));
        }
    }
    c->timing = timing;
    return NB_OK;
}

int clear_meta(nb_ctx* c) {
    NB_HIP(c, hipMemsetAsync(c->b.meta, 0, META_COUNT * sizeof(uint32_t), c->stream));
    return NB_OK;
}

// The learned walk's accepted-cell count, a 64-bit counter in two meta words.
uint64_t cells_total(const nb_ctx* c) {
    return (uint64_t)c->host_meta[META_CELLS_TOTAL] |
           (uint64_t)c->host_meta[META_CELLS_TOTAL + 1] << 32;
}

// The learned walk's near-cell quadrupole terms (over all bodies), a 64-bit
// counter in two meta words.
uint64_t quad_total(const nb_ctx* c) {
    return (uint64_t)c->host_meta[META_QUAD_TERMS] |
           (uint64_t)c->host_meta[META_QUAD_TERMS + 1] << 32;
}

int read_meta(nb_ctx* c) {
    NB_HIP(c, hipMemcpyAsync(c->host_meta, c->b.meta, META_COUNT * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    return NB_OK;
}

// The first level of the COM reduction's one-workgroup launch: kComMargin
// levels past the deepest non-empty level of the trees read back (those and
// the levels above stay one launch each). Picked from the first tree, and
// again (the steps captured again, ~ms) only when that launch met a level of
// more than kComRefit nodes, i.e. the trees grew deeper than the cut.
int fit_com_deep(nb_ctx* c, bool first) {
    if (!is_bh(c) || (c->host_meta[META_FLAGS] & kHaltFlags)) return NB_OK;
    if (!first && c->host_meta[META_COM_DEEP_MAX] <= (uint32_t)kComRefit) return NB_OK;
    const int deep = (int)std::min<uint32_t>(c->host_meta[META_COM_HI] + kComMargin, kMaxLevel);
    if (deep == c->com_deep) return NB_OK;
    c->com_deep = deep;
    return c->graph[0] ? build_graphs(c) : NB_OK;
}

// A context whose regrowth failed: the step graphs are gone and buffers
// they used are freed or missing, so nothing may launch on it any more.
int mark_dead(nb_ctx* c, int rc) {
    c->dead = true;
    c->dead_msg = "context unusable after a failed buffer regrowth: " + c->err;
    return rc;
}

int check_alive(nb_ctx* c) {
    return c->dead ? fail(c, NB_ERR_OUT_OF_MEMORY, c->dead_msg) : NB_OK;
}

// Grows whichever buffer the halted step overflowed. The captured graphs
// point at the buffers about to be freed, so they go first: a failure below
// must not leave a graph that a later step could replay.
int grow_for(nb_ctx* c, uint32_t flags) {
    destroy_graphs(c);
    c->regrowths++;
    if (flags & FLAG_TREE_CAPACITY) {
        const uint64_t icap = c->last_internal + c->last_internal / 4 + 1024;
        int rc = alloc_tree(c, icap);
        if (rc) return mark_dead(c, rc);
    }
    if (flags & FLAG_CELL_CAPACITY) {
        // at least double: the regions must grow for the step to converge
        const uint64_t need = c->host_meta[META_CELLS_NEEDED];
        const uint64_t cap = std::max<uint64_t>(2 * need + 64, 2ull * c->cell_cap);
        if (cap > 0xffffffffull)
            return mark_dead(c, fail(c, NB_ERR_OUT_OF_MEMORY, "accepted-cell region"));
        int rc = alloc_cells(c, (uint32_t)cap);
        if (rc) return mark_dead(c, rc);
    }
    return build_graphs(c);
}

// Builds the tree of the current state (moving it into Morton order) and,
// if requested, evaluates the forces without a kick. Grows the node pool
// until the tree fits.
int build_and_force(nb_ctx* c, bool forces) {
    if (int rc = check_alive(c)) return rc;
    for (int attempt = 0; attempt < 4; attempt++) {
        int rc = clear_meta(c);
        if (rc) return rc;
        if (forces) {
            c->cur = enqueue_force(c, c->cur, false, false);
        } else {
            Marks mk{c};
            (void)mk;
            launch_bbox(c, c->cur);
            launch_morton(c, c->cur);
            launch_radix_sort(c);
            launch_gather(c, c->cur);
            c->cur ^= 1;
            launch_tree_build(c, c->cur);
        }
        NB_HIP(c, hipGetLastError());
        rc = read_meta(c);
        if (rc) return rc;
        rc = fit_com_deep(c, false);
        if (rc) return rc;
        const uint32_t flags = c->host_meta[META_FLAGS];
        c->last_internal = c->host_meta[META_INTERNAL];
        if (forces) {
            c->last_cells = cells_total(c);
            c->last_quad = quad_total(c);
        }
        if (flags & FLAG_NONFINITE) return fail(c, NB_ERR_NUMERIC, "non-finite body positions");
        if (flags & FLAG_STACK) return fail(c, NB_ERR_OVERFLOW, "traversal stack overflow");
        if (flags & FLAG_SORT) return mark_dead(c, fail(c, NB_ERR_HIP, "radix sort stalled"));
        if (!(flags & (FLAG_TREE_CAPACITY | FLAG_CELL_CAPACITY))) return NB_OK;
        rc = grow_for(c, flags);
        if (rc) return rc;
    }
    return fail(c, NB_ERR_HIP, "tree capacity did not converge");
}

}  // namespace

namespace nb {
int build_tree(nb_ctx* c) { return build_and_force(c, false); }

int ensure_mom3(nb_ctx* c) {
    if (c->keep_mom3) return NB_OK;
    // Captured steps hold the tree buffers' addresses (no third moments):
    // they are recaptured with the new buffer.
    destroy_graphs(c);
    if (c->icap) {
        NB_HIP(c, dalloc(c, &c->b.mom3, c->nsys + 4 * c->icap));
        c->device_bytes += (c->nsys + 4 * c->icap) * sizeof(float4);
    }
    c->keep_mom3 = true;
    return NB_OK;
}
}  // namespace nb

extern "C" {

int nb_params_default(nb_params* p) {
    if (!p) return NB_ERR_INVALID_ARG;
    std::memset(p, 0, sizeof(*p));
    p->G = 1.0;
    p->dt = 1e-3;
    p->theta = 0.5;
    p->softening = 0.01;
    p->algorithm = NB_ALGO_BARNES_HUT;
    p->integrator = NB_INTEG_KD;
    p->self_guard = 0;
    p->far_field = NB_FAR_MONOPOLE;
    p->use_graph = 1;
    p->traversal = NB_TRAV_EXACT;
    return NB_OK;
}

}  // extern "C"

namespace {

// Traversal groups of walk_bodies(c) consecutive bodies, never across two
// systems, and the group walk's work queue (zero between kernels).
hipError_t alloc_groups(nb_ctx* c) {
    Buffers& b = c->b;
    std::vector<uint32_t> grp_first, grp_info;
    const uint32_t bpw = (uint32_t)walk_bodies(c);
    const auto& off = c->sys_off_h;
    for (uint32_t s = 0; s + 1 < (uint32_t)off.size(); s++) {
        for (uint32_t g = off[s]; g < off[s + 1]; g += bpw) {
            grp_first.push_back(g);
            grp_info.push_back((s << 8) | std::min<uint32_t>(bpw, off[s + 1] - g));
        }
    }
    if (b.grp_first) c->device_bytes -= 2ull * c->ngroups * sizeof(uint32_t);
    if (b.queue) c->device_bytes -= (uint64_t)c->queue_cap * sizeof(uint64_t);
    dfree(b.grp_first);
    dfree(b.grp_info);
    dfree(b.queue);
    c->ngroups = (uint32_t)grp_first.size();
    hipError_t e = dalloc(c, &b.grp_first, c->ngroups);
    if (!e) e = dalloc(c, &b.grp_info, c->ngroups);
    if (!e && c->ngroups) {
        e = hipMemcpy(b.grp_first, grp_first.data(), c->ngroups * sizeof(uint32_t),
                      hipMemcpyHostToDevice);
        if (!e)
            e = hipMemcpy(b.grp_info, grp_info.data(), c->ngroups * sizeof(uint32_t),
                          hipMemcpyHostToDevice);
    }
    if (!e && is_bh(c)) {
        c->queue_cap = (uint32_t)std::min<uint64_t>(4ull * c->ngroups + 64, 0xffffffffull);
        e = dalloc(c, &b.queue, c->queue_cap);
        if (!e) e = hipMemsetAsync(b.queue, 0, c->queue_cap * sizeof(uint64_t), c->stream);
        if (!e) e = hipStreamSynchronize(c->stream);
    }
    return e;
}

// One context for `nsys` independent systems whose bodies are the contiguous
// ranges [offsets[s], offsets[s+1]) (nsys = 1: the whole array).
int create_impl(const nb_params* p, uint64_t n, uint32_t nsys, const uint64_t* offsets,
                const double* mass, const double* pos_xy, const double* vel_xy,
                nb_ctx** ctx_out) {
    if (!ctx_out) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx_out is NULL");
    *ctx_out = nullptr;
    std::string why;
    int rc = check_params(p, &why);
    if (rc) return fail(nullptr, rc, why);
    if (n > 0x7fffffffull) return fail(nullptr, NB_ERR_INVALID_ARG, "n exceeds 2^31-1 bodies");
    if (n && (!mass || !pos_xy || !vel_xy))
        return fail(nullptr, NB_ERR_INVALID_ARG, "NULL body array");
    if (nsys > 1) {
        if (p->algorithm != NB_ALGO_BARNES_HUT || p->traversal != NB_TRAV_GROUP)
            return fail(nullptr, NB_ERR_NOT_SUPPORTED,
                        "several systems need algorithm=BARNES_HUT and traversal=GROUP");
        if (p->cells == NB_CELLS_REF_CPU)
            return fail(nullptr, NB_ERR_NOT_SUPPORTED, "cells=REF_CPU holds one system");
        if (nsys > 65536) return fail(nullptr, NB_ERR_INVALID_ARG, "at most 65536 systems");
    }
    std::vector<uint32_t> off(nsys + 1);
    for (uint32_t s = 0; s <= nsys; s++) {
        const uint64_t o = offsets ? offsets[s] : (s ? n : 0);
        if ((s == 0 && o != 0) || (s == nsys && o != n) || (s && n && o <= off[s - 1]) ||
            o > n)
            return fail(nullptr, NB_ERR_INVALID_ARG,
                        "offsets must rise strictly from 0 to n (no empty system)");
        off[s] = (uint32_t)o;
    }
    if (n == 0 && nsys > 1) return fail(nullptr, NB_ERR_INVALID_ARG, "no bodies");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, NB_ERR_NO_DEVICE, "no HIP device");
    nb_ctx* c = new (std::nothrow) nb_ctx();
    if (!c) return fail(nullptr, NB_ERR_OUT_OF_MEMORY, "host allocation failed");
    c->p = *p;
    c->n = n;
    c->nsys = nsys;
    c->dt = p->dt;
    c->keep_mom3 = p->far_field == NB_FAR_LEARNED;
    auto bail = [&](int code) {
        g_err = c->err;
        free_all(c);
        delete c;
        return code;
    };
#define NB_TRY(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            c->err = std::string(#expr) + ": " + hipGetErrorString(e_);                    \
            return bail(e_ == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP);    \
        }                                                                                  \
    } while (0)
    NB_TRY(hipGetDevice(&c->device));
    {
        int cus = 0;
        NB_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        if (cus > 0) c->cus = (uint32_t)cus;
    }
    NB_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) NB_TRY(hipEventCreate(&e));
    NB_TRY(hipHostMalloc((void**)&c->host_meta, META_COUNT * sizeof(uint32_t)));
    Buffers& b = c->b;
    const bool bh = is_bh(c);
    for (int k = 0; k < 2; k++) {
        const uint64_t nk = (k == 0 || bh) ? n : 0;
        NB_TRY(dalloc(c, &b.pos[k], nk));
        NB_TRY(dalloc(c, &b.vel[k], nk));
        NB_TRY(dalloc(c, &b.mass[k], nk));
        NB_TRY(dalloc(c, &b.orig[k], nk));
        NB_TRY(dalloc(c, &b.keys[k], bh ? n : 0));
        NB_TRY(dalloc(c, &b.vals[k], bh ? n : 0));
    }
    c->sort_tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
    c->bbox_blocks = (uint32_t)std::min<uint64_t>(std::max<uint64_t>((n + 2047) / 2048, 1), 1024);
    // published words of the widest scan (one per 1024 values, more than its
    // tiles): the level counts of the largest node pool (internal nodes are
    // fewer than 30 per body pair); zeroed once, as the sort's
    const uint64_t scan_words = std::max<uint64_t>((n + 1023) / 1024,
                                                   (uint64_t)kMaxLevel * (30 * n / 256 + 2) / 1024) + 2;
    // radix sort state: zeroed once, so no published word can match a sort's tag
    // (tiles + tile groups of 16) published words, and one histogram row per
    // Morton workgroup (kSortTile keys each)
    const uint64_t sort_words = (uint64_t)kSortBins * (c->sort_tiles + (c->sort_tiles + 15) / 16);
    NB_TRY(dalloc(c, &b.sort_glob, kSortGlobWords));
    NB_TRY(dalloc(c, &b.sort_status, bh ? sort_words : 0));
    NB_TRY(dalloc(c, &b.sort_part, bh ? (uint64_t)c->sort_tiles * kSortMaxPasses * kSortBins : 0));
    NB_TRY(hipMemsetAsync(b.sort_glob, 0, kSortGlobWords * sizeof(uint32_t), c->stream));
    if (bh && c->sort_tiles)
        NB_TRY(hipMemsetAsync(b.sort_status, 0, sort_words * sizeof(uint64_t), c->stream));
    NB_TRY(dalloc(c, &b.box, (uint64_t)kBoxStride * nsys));
    // systems and traversal groups (<= 64 bodies, never across two systems)
    std::vector<uint32_t> sys_of(n);
    for (uint32_t s = 0; s < nsys; s++)
        for (uint32_t i = off[s]; i < off[s + 1]; i++) sys_of[i] = s;
    c->sys_off_h = off;
    NB_TRY(dalloc(c, &b.sys_of, n));
    NB_TRY(dalloc(c, &b.sys_off, nsys + 1));
    if (n)
        NB_TRY(hipMemcpy(b.sys_of, sys_of.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    NB_TRY(alloc_groups(c));
    NB_TRY(hipMemcpy(b.sys_off, off.data(), (nsys + 1) * sizeof(uint32_t),
                     hipMemcpyHostToDevice));
    NB_TRY(dalloc(c, &b.bbox_part, 4ull * c->bbox_blocks));
    if (bh && p->cells == NB_CELLS_REF_CPU) {
        c->ref_blocks = (uint32_t)std::max<uint64_t>((n + kRefBoxTile - 1) / kRefBoxTile, 1);
        NB_TRY(dalloc(c, &b.ref_pos, n));
        NB_TRY(dalloc(c, &b.ref_part, c->ref_blocks));
    }
    NB_TRY(dalloc(c, &b.delta, bh ? n : 0));
    NB_TRY(dalloc(c, &b.cnt, bh ? n : 0));
    NB_TRY(dalloc(c, &b.scan_status, scan_words));
    NB_TRY(hipMemsetAsync(b.scan_status, 0, scan_words * sizeof(uint64_t), c->stream));
    NB_TRY(dalloc(c, &b.multi_list, bh ? n / 2 + 1 : 0));
    NB_TRY(dalloc(c, &b.acc, n));
    NB_TRY(dalloc(c, &b.inter, n));
    NB_TRY(dalloc(c, &b.meta, META_COUNT));
    NB_TRY(dalloc(c, &b.consts, 1));
    if (!bh && n) {
        const uint32_t ib = (uint32_t)((n + 1023) / 1024);
        uint32_t split = (2048 + ib - 1) / ib;
        split = std::max<uint32_t>(1, std::min<uint32_t>(split, 64));
        while (split > 1 && (n + split - 1) / split < 256) split--;
        c->ap_split = split;
        // room for the largest split the autotuner tries
        uint32_t smax = 64;
        while (smax > split && (n + smax - 1) / smax < 256) smax >>= 1;
        c->ap_split_max = std::max(smax, split);
        NB_TRY(dalloc(c, &b.ap_part, (uint64_t)c->ap_split_max * n));
    }
    if (n) {
        NB_TRY(hipMemcpyAsync(b.pos[0], pos_xy, n * sizeof(double2), hipMemcpyHostToDevice,
                              c->stream));
        NB_TRY(hipMemcpyAsync(b.vel[0], vel_xy, n * sizeof(double2), hipMemcpyHostToDevice,
                              c->stream));
        NB_TRY(hipMemcpyAsync(b.mass[0], mass, n * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        std::vector<uint32_t> iota(n);
        for (uint64_t i = 0; i < n; i++) iota[i] = (uint32_t)i;
        NB_TRY(hipMemcpyAsync(b.orig[0], iota.data(), n * sizeof(uint32_t),
                              hipMemcpyHostToDevice, c->stream));
        NB_TRY(hipStreamSynchronize(c->stream));
    }
#undef NB_TRY
    rc = upload_consts(c);
    if (rc) return bail(rc);
    if (c->p.far_field == NB_FAR_LEARNED && n) {
        // the graphs capture the network's buffers: allocate them (zero
        // weights, i.e. no correction) before the first launch
        rc = prepare_mlp(c);
        if (rc) return bail(rc);
        rc = alloc_cells(c, c->p.cell_capacity > 0 ? (uint32_t)c->p.cell_capacity
                                                     : kInitialCells);
        if (rc) return bail(rc);
    }
    if (bh && n) {
        rc = alloc_tree(c, std::max<uint64_t>(n, 1024));
        if (rc) return bail(rc);
        // Sizes the node pool to the first tree and, with the learned far
        // field, the accepted-cell regions to four times the first walk's
        // busiest group (the cells a walk accepts do not depend on the
        // weights): once bodies scatter, wavefronts also walk queued pieces
        // of split groups into their regions, and on a 1M disk the busiest
        // region then holds about 3x the first walk's. A step that still
        // overflows halts, and the regions and graphs are rebuilt (~4 ms).
        const bool learned = c->p.far_field == NB_FAR_LEARNED;
        rc = build_and_force(c, learned);
        if (rc) return bail(rc);
        rc = fit_com_deep(c, true);  // from the first tree (no steps captured yet)
        if (rc) return bail(rc);
        if (c->last_internal + c->last_internal / 4 > c->icap) {
            rc = alloc_tree(c, c->last_internal + c->last_internal / 4 + 1024);
            if (rc) return bail(rc);
        }
        const uint64_t cells = 4ull * c->host_meta[META_CELLS_NEEDED] + 64;
        if (learned && c->p.cell_capacity == 0 && cells > c->cell_cap && cells <= 0xffffffffull) {
            rc = alloc_cells(c, (uint32_t)cells);
            if (rc) return bail(rc);
        }
    }
    rc = build_graphs(c);
    if (rc) return bail(rc);
    *ctx_out = c;
    return NB_OK;
}

}  // namespace

extern "C" {

int nb_create(const nb_params* p, uint64_t n, const double* mass, const double* pos_xy,
              const double* vel_xy, nb_ctx** ctx_out) {
    return create_impl(p, n, 1, nullptr, mass, pos_xy, vel_xy, ctx_out);
}

int nb_create_batch(const nb_params* p, uint32_t nsys, const uint64_t* offsets,
                    const double* mass, const double* pos_xy, const double* vel_xy,
                    nb_ctx** ctx_out) {
    if (nsys == 0 || !offsets) return fail(nullptr, NB_ERR_INVALID_ARG, "no systems");
    return create_impl(p, offsets[nsys], nsys, offsets, mass, pos_xy, vel_xy, ctx_out);
}

int nb_step(nb_ctx* c, uint64_t k) {
    if (!c) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx is NULL");
    if (int rc = check_alive(c)) return rc;
    if (c->n == 0) {
        c->iteration += k;
        c->sim_time += (double)k * c->dt;
        return NB_OK;
    }
    uint64_t remaining = k;
    int regrowths = 0;  // each grows a buffer past the need that halted a step
    while (remaining > 0) {
        int rc = clear_meta(c);
        if (rc) return rc;
        for (uint64_t s = 0; s < remaining; s++) {
            if (!c->timing && c->graph4[c->cur] && remaining - s >= (uint64_t)kGraphSteps) {
                ensure_box_parts(c, c->cur);
                NB_HIP(c, hipGraphLaunch(c->graph4[c->cur], c->stream));  // ends on c->cur
                s += kGraphSteps - 1;
            } else if (c->timing) {
                c->marked = 0;
                const int dst = enqueue_step(c, c->cur);
                NB_HIP(c, hipGetLastError());
                NB_HIP(c, hipStreamSynchronize(c->stream));
                for (int ph = 0; ph < NB_PHASE_COUNT; ph++) {
                    if (!((c->marked >> ph) & 1u)) continue;  // phase absent from this step
                    float ms = 0.f;
                    NB_HIP(c, hipEventElapsedTime(&ms, c->ev[2 * ph], c->ev[2 * ph + 1]));
                    c->phase_ms[ph] += ms;
                }
                c->steps_timed++;
                c->cur = dst;
            } else if (c->graph[c->cur]) {
                ensure_box_parts(c, c->cur);
                NB_HIP(c, hipGraphLaunch(c->graph[c->cur], c->stream));
                if (is_bh(c)) c->cur ^= 1;
            } else {
                c->cur = enqueue_step(c, c->cur);
                NB_HIP(c, hipGetLastError());
            }
        }
        rc = read_meta(c);
        if (rc) return rc;
        rc = fit_com_deep(c, false);
        if (rc) return rc;
        const uint32_t flags = c->host_meta[META_FLAGS];
        const uint64_t done = std::min<uint64_t>(c->host_meta[META_STEPS_DONE], remaining);
        c->last_internal = c->host_meta[META_INTERNAL];
        if (done) {
            c->last_cells = cells_total(c) / done;
            c->last_quad = quad_total(c) / done;
        }
        c->iteration += done;
        c->sim_time += (double)done * c->dt;
        remaining -= done;
        if (flags & FLAG_NONFINITE) return fail(c, NB_ERR_NUMERIC, "non-finite body positions");
        if (flags & FLAG_STACK) return fail(c, NB_ERR_OVERFLOW, "traversal stack overflow");
        if (flags & FLAG_SORT) return mark_dead(c, fail(c, NB_ERR_HIP, "radix sort stalled"));
        if (flags & (FLAG_TREE_CAPACITY | FLAG_CELL_CAPACITY)) {
            if (++regrowths > 32) return fail(c, NB_ERR_HIP, "buffer capacity did not converge");
            rc = grow_for(c, flags);
            if (rc) return rc;
            if (is_dk(c) && remaining > 0) {
                // The halted step drifted but did not kick: finish it.
                rc = clear_meta(c);
                if (rc) return rc;
                c->cur = enqueue_force(c, c->cur, true, true);
                NB_HIP(c, hipGetLastError());
                rc = read_meta(c);
                if (rc) return rc;
                if (c->host_meta[META_FLAGS] & (FLAG_TREE_CAPACITY | FLAG_CELL_CAPACITY))
                    return fail(c, NB_ERR_HIP, "buffer capacity did not converge");
                c->iteration += 1;
                c->sim_time += c->dt;
                remaining -= 1;
            }
            continue;
        }
        if (remaining > 0) return fail(c, NB_ERR_HIP, "steps did not complete");
    }
    return NB_OK;
}

int nb_get_state(nb_ctx* c, double* pos_xy, double* vel_xy) {
    if (!c) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx is NULL");
    const uint64_t n = c->n;
    if (n == 0) return NB_OK;
    const double2* ps = c->b.pos[c->cur];
    const double2* vs = c->b.vel[c->cur];
    if (is_bh(c)) {
        // Morton order on the device: un-permute on the GPU into the idle copy
        // of the body arrays (the next step's gather overwrites it anyway),
        // then copy straight into the caller's arrays
        double2* pt = c->b.pos[c->cur ^ 1];
        double2* vt = c->b.vel[c->cur ^ 1];
        launch_unpermute2(c, pt, vt);
        NB_HIP(c, hipGetLastError());
        ps = pt;
        vs = vt;
    }
    // (all-pairs contexts never reorder their bodies)
    if (pos_xy)
        NB_HIP(c, hipMemcpyAsync(pos_xy, ps, n * sizeof(double2), hipMemcpyDeviceToHost,
                                 c->stream));
    if (vel_xy)
        NB_HIP(c, hipMemcpyAsync(vel_xy, vs, n * sizeof(double2), hipMemcpyDeviceToHost,
                                 c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    return NB_OK;
}

int nb_snapshot_begin(nb_ctx* c, uint64_t* ticket) {
    if (!c || !ticket) return fail(c, NB_ERR_INVALID_ARG, "ctx or ticket is NULL");
    const uint64_t n = c->n;
    const uint64_t t = c->snap_next;
    nb_ctx::Snapshot& s = c->snap[t & 1];
    if (!c->copy_stream) NB_HIP(c, alloc_snapshots(c));
    if (n > 0) {
        // the slot's previous copy must have left the device buffer before the
        // un-permute overwrites it
        if (s.ticket) NB_HIP(c, hipStreamWaitEvent(c->stream, s.done, 0));
        launch_unpermute(c, s.dev);
        NB_HIP(c, hipGetLastError());
        NB_HIP(c, hipEventRecord(s.ready, c->stream));
        NB_HIP(c, hipStreamWaitEvent(c->copy_stream, s.ready, 0));
        NB_HIP(c, hipMemcpyAsync(s.host, s.dev, 32 * n, hipMemcpyDeviceToHost, c->copy_stream));
    }
    NB_HIP(c, hipEventRecord(s.done, c->copy_stream));
    s.ticket = t;
    s.iteration = c->iteration;
    c->snap_next = t + 1;
    *ticket = t;
    return NB_OK;
}

static nb_ctx::Snapshot* find_snapshot(nb_ctx* c, uint64_t ticket) {
    if (!c || ticket == 0) return nullptr;
    nb_ctx::Snapshot& s = c->snap[ticket & 1];
    return s.ticket == ticket ? &s : nullptr;
}

int nb_snapshot_query(nb_ctx* c, uint64_t ticket, int* ready) {
    if (!c || !ready) return fail(c, NB_ERR_INVALID_ARG, "ctx or ready is NULL");
    nb_ctx::Snapshot* s = find_snapshot(c, ticket);
    if (!s) return fail(c, NB_ERR_INVALID_ARG, "no snapshot holds this ticket");
    const hipError_t e = hipEventQuery(s->done);
    if (e == hipErrorNotReady) {
        *ready = 0;
        return NB_OK;
    }
    NB_HIP(c, e);
    *ready = 1;
    return NB_OK;
}

int nb_snapshot_wait(nb_ctx* c, uint64_t ticket, const double** pos_xy, const double** vel_xy,
                     uint64_t* iteration) {
    if (!c) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx is NULL");
    nb_ctx::Snapshot* s = find_snapshot(c, ticket);
    if (!s) return fail(c, NB_ERR_INVALID_ARG, "no snapshot holds this ticket");
    NB_HIP(c, hipEventSynchronize(s->done));
    if (pos_xy) *pos_xy = s->host;
    if (vel_xy) *vel_xy = s->host + 2 * c->n;
    if (iteration) *iteration = s->iteration;
    return NB_OK;
}

int nb_compute_forces(nb_ctx* c, double* acc_xy, uint32_t* interactions, uint64_t* total) {
    if (!c) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx is NULL");
    const uint64_t n = c->n;
    if (total) *total = 0;
    if (n == 0) return NB_OK;
    int rc = build_and_force(c, true);
    if (rc) return rc;
    std::vector<double> a(2 * n);
    std::vector<uint32_t> k(n), o(n);
    NB_HIP(c, hipMemcpyAsync(a.data(), c->b.acc, n * sizeof(double2), hipMemcpyDeviceToHost,
                             c->stream));
    NB_HIP(c, hipMemcpyAsync(k.data(), c->b.inter, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             c->stream));
    NB_HIP(c, hipMemcpyAsync(o.data(), c->b.orig[c->cur], n * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, c->stream));
    NB_HIP(c, hipStreamSynchronize(c->stream));
    uint64_t t = 0;
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t j = o[i];
        if (acc_xy) {
            acc_xy[2 * j] = a[2 * i];
            acc_xy[2 * j + 1] = a[2 * i + 1];
        }
        if (interactions) interactions[j] = k[i];
        t += k[i];
    }
    if (total) *total = t;
    return NB_OK;
}

int nb_debug_fail_alloc(nb_ctx* c, uint32_t count) {
    if (!c) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx is NULL");
    c->fail_allocs = count;
    return NB_OK;
}

int nb_debug_com_deep(nb_ctx* c, int32_t set, int32_t* current) {
    if (!c) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx is NULL");
    if (set < -1 || set > kMaxLevel) return fail(c, NB_ERR_INVALID_ARG, "level out of range");
    if (set >= 0 && set != c->com_deep) {
        if (int rc = check_alive(c)) return rc;
        c->com_deep = set;
        if (c->graph[0]) {
            if (int rc = build_graphs(c)) return rc;
        }
    }
    if (current) *current = c->com_deep;
    return NB_OK;
}

int nb_set_dt(nb_ctx* c, double dt) {
    if (!c) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx is NULL");
    if (!std::isfinite(dt) || dt <= 0.0) return fail(c, NB_ERR_INVALID_ARG, "dt must be > 0");
    c->dt = dt;
    return upload_consts(c);
}

int nb_get_info(nb_ctx* c, nb_info* info) {
    if (!c || !info) return fail(c, NB_ERR_INVALID_ARG, "NULL argument");
    info->n = c->n;
    info->iteration = c->iteration;
    info->simulated_time = c->sim_time;
    info->internal_nodes = c->last_internal;
    info->tree_nodes = c->n ? c->nsys + 4 * c->last_internal : 0;
    info->node_capacity = c->icap;
    info->device_bytes = c->device_bytes;
    info->far_field_cells = c->last_cells;
    info->quad_terms = c->last_quad;
    info->regrowths = c->regrowths;
    info->graph_captures = c->graph_captures;
    return NB_OK;
}

int nb_set_timing(nb_ctx* c, int enable) {
    if (!c) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx is NULL");
    c->timing = enable != 0;
    for (double& m : c->phase_ms) m = 0.0;
    c->steps_timed = 0;
    return NB_OK;
}

int nb_get_timing(nb_ctx* c, double* phase_ms, int nphase, uint64_t* steps_timed) {
    if (!c) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx is NULL");
    for (int k = 0; k < nphase && k < NB_PHASE_COUNT; k++) phase_ms[k] = c->phase_ms[k];
    if (steps_timed) *steps_timed = c->steps_timed;
    return NB_OK;
}

int nb_autotune(nb_ctx* c, uint32_t repeats, nb_tuning* out) {
    if (!c) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx is NULL");
    if (int rc = check_alive(c)) return rc;
    if (out) std::memset(out, 0, sizeof(*out));
    if (c->n == 0) return NB_OK;
    if (repeats == 0) repeats = 1;
    const bool timing = c->timing;
    // Force evaluation time (ms, mean of `repeats`) with the current settings.
    auto measure = [&](double* ms_out) -> int {
        double sum = 0.0;
        for (uint32_t r = 0; r < repeats;) {
            int rc = clear_meta(c);
            if (rc) return rc;
            c->timing = true;
            c->marked = 0;
            c->cur = enqueue_force(c, c->cur, false, false);
            c->timing = timing;
            NB_HIP(c, hipGetLastError());
            rc = read_meta(c);
            if (rc) return rc;
            const uint32_t flags = c->host_meta[META_FLAGS];
            c->last_internal = c->host_meta[META_INTERNAL];
            if (flags & FLAG_NONFINITE) return fail(c, NB_ERR_NUMERIC, "non-finite positions");
            if (flags & FLAG_STACK) return fail(c, NB_ERR_OVERFLOW, "traversal stack overflow");
            if (flags & FLAG_SORT) return mark_dead(c, fail(c, NB_ERR_HIP, "radix sort stalled"));
            if (flags & (FLAG_TREE_CAPACITY | FLAG_CELL_CAPACITY)) {
                rc = grow_for(c, flags);  // then time this setting again
                if (rc) return rc;
                continue;
            }
            double ms = 0.0;
            for (int ph : {(int)NB_PHASE_FORCE, (int)NB_PHASE_FAR_FIELD}) {
                if (!((c->marked >> ph) & 1u)) continue;
                float t = 0.f;
                NB_HIP(c, hipEventElapsedTime(&t, c->ev[2 * ph], c->ev[2 * ph + 1]));
                ms += t;
            }
            sum += ms;
            r++;
        }
        *ms_out = sum / repeats;
        return NB_OK;
    };
    // A setting replaces the current one only when it is clearly faster:
    // walks vary from step to step, and a setting that wins a close race on
    // one state (a small list budget, say) can lose on the next steps.
    constexpr double kClearly = 0.97;
    nb_tuning t{};
    double base = 0.0;
    int rc = measure(&base);
    if (rc) return rc;
    double best = base;
    t.candidates = 1;
    if (is_bh(c) && c->p.traversal == NB_TRAV_GROUP) {
        // One tunable at a time, each from the best setting so far. Changing
        // the bodies per wavefront rebuilds the traversal groups (and the
        // learned far field's per-group buffers).
        auto set_bodies = [&](int32_t v) -> int {
            const uint64_t old_groups = c->ngroups;
            c->p.walk_bodies = v;
            destroy_graphs(c);
            NB_HIP(c, alloc_groups(c));
            return regroup_cells(c, old_groups);
        };
        auto tune = [&](int32_t* field, std::initializer_list<int32_t> values, int32_t deflt,
                        bool regroup) -> int {
            const int32_t start = *field > 0 ? *field : deflt;
            int32_t pick = start;
            for (int32_t v : values) {
                if (v == start) continue;
                int rc2 = regroup ? set_bodies(v) : (*field = v, NB_OK);
                if (rc2) return rc2;
                double ms = 0.0;
                rc2 = measure(&ms);
                if (rc2) return rc2;
                t.candidates++;
                if (ms < best * kClearly) {
                    best = ms;
                    pick = v;
                }
            }
            if (regroup && c->p.walk_bodies != pick) return set_bodies(pick);
            *field = pick;
            return NB_OK;
        };
        if ((rc = tune(&c->p.walk_bodies, {64, 32}, 64, true))) return rc;
        if ((rc = tune(&c->p.walk_stack, {2048, 1024}, 2048, false))) return rc;
        if ((rc = tune(&c->p.list_flush, {64, 128, 192}, 128, false))) return rc;
        // The list budget is not tuned: its effect changes from step to step.
        // On a 1M disk at theta 0.15, a budget of 2048 timed 5 % faster on
        // the tuned state and ran 20 % slower over the next steps (every
        // split restarts walks, and which groups split moves with the bodies).
    } else if (!is_bh(c)) {
        const uint32_t r0 = c->ap_rows, s0 = c->ap_split;
        uint32_t pr = r0, ps = s0;
        for (uint32_t rows : {2u, 4u, 8u}) {
            for (uint32_t split = 1; split <= c->ap_split_max; split <<= 1) {
                if ((rows == r0 && split == s0) || (c->n + split - 1) / split < 256) continue;
                c->ap_rows = rows;
                c->ap_split = split;
                double ms = 0.0;
                rc = measure(&ms);
                if (rc) return rc;
                t.candidates++;
                if (ms < best * kClearly) {
                    best = ms;
                    pr = rows;
                    ps = split;
                }
            }
        }
        c->ap_rows = pr;
        c->ap_split = ps;
    }
    rc = build_graphs(c);
    if (rc) return rc;
    t.group_budget = c->p.group_budget > 0 ? c->p.group_budget : kGroupBudget;
    t.walk_bodies = walk_bodies(c);
    t.walk_stack = walk_stack(c);
    t.list_flush = list_flush(c);
    t.ap_rows = (int32_t)c->ap_rows;
    t.ap_split = (int32_t)c->ap_split;
    t.best_ms = best;
    t.default_ms = base;
    if (out) *out = t;
    return NB_OK;
}

void nb_destroy(nb_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
}

const char* nb_last_error(const nb_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

const char* nb_status_string(int s) {
    switch (s) {
    case NB_OK: return "ok";
    case NB_ERR_INVALID_ARG: return "invalid argument";
    case NB_ERR_HIP: return "HIP runtime error";
    case NB_ERR_OUT_OF_MEMORY: return "out of device memory";
    case NB_ERR_OVERFLOW: return "traversal stack overflow";
    case NB_ERR_NOT_SUPPORTED: return "not supported";
    case NB_ERR_NO_DEVICE: return "no HIP device";
    case NB_ERR_NUMERIC: return "non-finite state";
    default: return "unknown status";
    }
}

int nb_debug_build(nb_ctx* c, double* box4, uint64_t* sorted_keys, uint32_t* orig_index,
                   uint64_t* node_count) {
    if (!c) return fail(nullptr, NB_ERR_INVALID_ARG, "ctx is NULL");
    if (!is_bh(c)) return fail(c, NB_ERR_NOT_SUPPORTED, "no tree for all-pairs");
    const uint64_t n = c->n;
    if (n == 0) {
        if (node_count) *node_count = 0;
        return NB_OK;
    }
    int rc = build_and_force(c, false);
    if (rc) return rc;
    if (box4) NB_HIP(c, hipMemcpy(box4, c->b.box, 4 * sizeof(double), hipMemcpyDeviceToHost));
    if (sorted_keys)
        NB_HIP(c, hipMemcpy(sorted_keys, c->b.keys[0], n * sizeof(uint64_t),
                            hipMemcpyDeviceToHost));
    if (orig_index)
        NB_HIP(c, hipMemcpy(orig_index, c->b.orig[c->cur], n * sizeof(uint32_t),
                            hipMemcpyDeviceToHost));
    if (node_count) *node_count = c->nsys + 4 * c->last_internal;
    return NB_OK;
}

int nb_debug_tree(nb_ctx* c, nb_node* nodes, uint64_t cap) {
    if (!c || !nodes) return fail(c, NB_ERR_INVALID_ARG, "NULL argument");
    if (c->n == 0) return NB_OK;
    const uint64_t count = c->nsys + 4 * c->last_internal;
    if (cap < count) return fail(c, NB_ERR_INVALID_ARG, "cap smaller than the node count");
    std::vector<NodeRec> h(count);
    NB_HIP(c, hipMemcpy(h.data(), c->b.nodes, count * sizeof(NodeRec), hipMemcpyDeviceToHost));
    std::vector<float4> t3(c->b.mom3 ? count : 0);
    if (c->b.mom3)
        NB_HIP(c, hipMemcpy(t3.data(), c->b.mom3, count * sizeof(float4), hipMemcpyDeviceToHost));
    // Empty child slots are not written on the device (their parent's mask
    // marks them); they are reported here as empty nodes. A parent's slot
    // precedes its children's, so one forward pass finds them.
    std::vector<int32_t> level(count, 0);
    std::vector<uint8_t> live(count, 0);
    std::vector<double> empty_wsq(count, 0.0);
    for (uint64_t s = 0; s < c->nsys && s < count; s++) live[s] = 1;
    for (uint64_t s = 0; s < count; s++) {
        nb_node& o = nodes[s];
        if (!live[s]) {
            o = nb_node{};
            o.level = level[s];
            o.bstart = -1;
            o.bend = -1;
            o.c0 = -2;
            o.wsq = empty_wsq[s];
            continue;
        }
        const NodeRec& r = h[s];
        if (r.c0 >= 1 && (uint64_t)r.c0 + 3 < count)
            for (int q = 0; q < 4; q++) {
                level[r.c0 + q] = level[s] + 1;
                live[r.c0 + q] = (uint8_t)((r.mask >> q) & 1);
                empty_wsq[r.c0 + q] = r.wsq * 0.25;
            }
        o.level = level[s];
        o.bstart = r.bs;
        o.bend = r.be;
        o.c0 = r.c0;
        o.mass = r.m;
        o.cx = r.cx;
        o.cy = r.cy;
        o.wsq = r.wsq;
        o.ixx = r.ixx;
        o.ixy = r.ixy;
        o.iyy = r.iyy;
        o.child_mask = r.c0 >= 1 ? r.mask : 0;
        if (!t3.empty()) {
            o.t3[0] = t3[s].x;
            o.t3[1] = t3[s].y;
            o.t3[2] = t3[s].z;
            o.t3[3] = t3[s].w;
        }
    }
    return NB_OK;
}

}  // extern "C"

namespace {

// Partial sums of nb_direct_sum_f64_noalloc: [split][m] double2 entries in a
// fixed device array (16 MiB on every GPU the library runs on, held until
// the process ends). Host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device.
constexpr uint64_t kNoAllocScratchEntries = (uint64_t)NB_NOALLOC_SCRATCH_BYTES / sizeof(double2);
__device__ __align__(256) double2 scratch[kNoAllocScratchEntries];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mtx;
double2* g_scratch_ptr[kMaxDevices] = {};

// The scratch array's address on the current device, looked up once per
// device with that device current.
hipError_t scratch_on_current_device(double2** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mtx);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double2*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

// The j split of the direct sum for m targets over n sources.
uint32_t direct_sum_split(uint64_t n, uint64_t m) {
    const uint32_t gx = (uint32_t)((m + 255) / 256);
    uint32_t split = std::max<uint32_t>(1, std::min<uint32_t>((2048 + gx - 1) / gx, 256));
    while (split > 1 && (n + split - 1) / split < 256) split--;
    return split;
}

}  // namespace

extern "C" {

int nb_direct_sum_f64_noalloc(uint64_t n, const double* pos_xy, const double* mass, double G,
                              double softening, uint64_t m, const int64_t* targets,
                              double* acc_xy, void* stream) {
    if ((n && (!pos_xy || !mass)) || (m && (!targets || !acc_xy)))
        return fail(nullptr, NB_ERR_INVALID_ARG, "NULL argument");
    if (n > 0x7fffffffull || m > 0x7fffffffull)
        return fail(nullptr, NB_ERR_INVALID_ARG, "size exceeds 2^31-1");
    if (m > kNoAllocScratchEntries)
        return fail(nullptr, NB_ERR_INVALID_ARG,
                    "m exceeds the scratch buffer (NB_NOALLOC_SCRATCH_BYTES / 16 targets)");
    if (m == 0) return NB_OK;
    double2* part = nullptr;
    hipError_t e = scratch_on_current_device(&part);
    if (e != hipSuccess)
        return fail(nullptr, NB_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    uint32_t split = direct_sum_split(n, m);
    while ((uint64_t)split * m > kNoAllocScratchEntries) split--;  // keeps split >= 1
    hipStream_t s = static_cast<hipStream_t>(stream);
    e = direct_sum_f64(n, reinterpret_cast<const double2*>(pos_xy), mass,
                       softening * softening, m, targets, part, split, s);
    if (e == hipSuccess)
        e = direct_sum_f64_finish(part, split, m, n, targets, G,
                                  reinterpret_cast<double2*>(acc_xy), s);
    if (e != hipSuccess)
        return fail(nullptr, NB_ERR_HIP, std::string("direct sum: ") + hipGetErrorString(e));
    return NB_OK;
}

int nb_direct_sum_f64(uint64_t n, const double* pos_xy, const double* mass, double G,
                      double softening, uint64_t m, const int64_t* targets, double* acc_xy) {
    if ((n && (!pos_xy || !mass)) || (m && (!targets || !acc_xy)))
        return fail(nullptr, NB_ERR_INVALID_ARG, "NULL argument");
    if (n > 0x7fffffffull || m > 0x7fffffffull)
        return fail(nullptr, NB_ERR_INVALID_ARG, "size exceeds 2^31-1");
    for (uint64_t t = 0; t < m; t++)
        if (targets[t] < 0 || (uint64_t)targets[t] >= n)
            return fail(nullptr, NB_ERR_INVALID_ARG, "target index out of range");
    if (m == 0) return NB_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, NB_ERR_NO_DEVICE, "no HIP device");
    const uint32_t split = direct_sum_split(n, m);
    double2 *dpos = nullptr, *dpart = nullptr;
    double* dmass = nullptr;
    int64_t* dt = nullptr;
    hipStream_t s = nullptr;
    int rc = NB_OK;
    std::vector<double2> part((uint64_t)split * m);
    auto run = [&]() -> hipError_t {
        hipError_t e;
        if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking))) return e;
        if ((e = hipMalloc((void**)&dpos, n * sizeof(double2)))) return e;
        if ((e = hipMalloc((void**)&dmass, n * sizeof(double)))) return e;
        if ((e = hipMalloc((void**)&dt, m * sizeof(int64_t)))) return e;
        if ((e = hipMalloc((void**)&dpart, (uint64_t)split * m * sizeof(double2)))) return e;
        if ((e = hipMemcpyAsync(dpos, pos_xy, n * sizeof(double2), hipMemcpyHostToDevice, s)))
            return e;
        if ((e = hipMemcpyAsync(dmass, mass, n * sizeof(double), hipMemcpyHostToDevice, s)))
            return e;
        if ((e = hipMemcpyAsync(dt, targets, m * sizeof(int64_t), hipMemcpyHostToDevice, s)))
            return e;
        if ((e = direct_sum_f64(n, dpos, dmass, softening * softening, m, dt, dpart, split, s)))
            return e;
        if ((e = hipMemcpyAsync(part.data(), dpart, (uint64_t)split * m * sizeof(double2),
                                hipMemcpyDeviceToHost, s)))
            return e;
        return hipStreamSynchronize(s);
    };
    const hipError_t e = run();
    if (e != hipSuccess)
        rc = fail(nullptr, e == hipErrorOutOfMemory ? NB_ERR_OUT_OF_MEMORY : NB_ERR_HIP,
                  std::string("direct sum: ") + hipGetErrorString(e));
    if (dpos) (void)hipFree(dpos);
    if (dmass) (void)hipFree(dmass);
    if (dt) (void)hipFree(dt);
    if (dpart) (void)hipFree(dpart);
    if (s) (void)hipStreamDestroy(s);
    if (rc) return rc;
    for (uint64_t t = 0; t < m; t++) {
        double ax = 0.0, ay = 0.0;
        for (uint32_t k = 0; k < split; k++) {
            ax += part[(uint64_t)k * m + t].x;
            ay += part[(uint64_t)k * m + t].y;
        }
        acc_xy[2 * t] = ax * G;
        acc_xy[2 * t + 1] = ay * G;
    }
    return NB_OK;
}

}  // extern "C"
