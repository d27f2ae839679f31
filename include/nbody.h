/*
 * nbody.h — C ABI of the MI355X (gfx950) 2-D gravitational N-body step.
 *
 * Drop-in boundary for the reference's `Simulation` step interface
 * (src/Simulation/include/Simulation/Simulation.hpp:13-67): a context owns the
 * bodies on the GPU, a step advances them, and the state can be read back in
 * the caller's original body order at any time. The C++ adapter
 * `nbody2d::GpuSimulation` (n-body-2d_amd/host/) wraps these calls behind the
 * reference's run/pause/get_stats/set_timestep API.
 *
 * Conventions
 *   - Every function returns NB_OK (0) or an NB_ERR_* code; nothing throws or
 *     exits across the ABI (the reference's CUDA_CHECK calls ::exit,
 *     BarnesHutCuda.cu:12-19). nb_last_error() gives the message.
 *   - Positions/velocities/accelerations are interleaved x,y FP64 arrays,
 *     byte-compatible with the reference's std::vector<sf::Vector2<double>>
 *     (Body.hpp:34-37) and `Vector2 {double x, y}` (BarnesHutCuda.hpp:5-7).
 *   - Host pointers are borrowed for the duration of a call only.
 *   - A context is used by one thread at a time.
 */
#ifndef NBODY_H
#define NBODY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NB_ABI_VERSION 1

typedef struct nb_ctx nb_ctx; /* opaque context */

enum nb_status {
    NB_OK = 0,
    NB_ERR_INVALID_ARG = 1,   /* bad parameter (range checks of Config.cpp:224-253) */
    NB_ERR_HIP = 2,           /* a HIP runtime call failed */
    NB_ERR_OUT_OF_MEMORY = 3,
    NB_ERR_OVERFLOW = 4,      /* traversal stack overflow reported by the kernel */
    NB_ERR_NOT_SUPPORTED = 5,
    NB_ERR_NO_DEVICE = 6,
    NB_ERR_NUMERIC = 7        /* non-finite positions found while building the tree */
};

/* Force evaluation. BARNES_HUT is a quadtree walk (the cells, opening test
 * and force law below pick the reference path it reproduces); ALL_PAIRS is
 * the direct sum of AllPairsSim (AllPairs.cpp:38-49). */
enum nb_algorithm { NB_ALGO_BARNES_HUT = 0, NB_ALGO_ALL_PAIRS = 1 };

/* Pair force law. PLUMMER: G m M d / (d^2 + eps^2)^{3/2}, the CUDA kernel's
 * (BarnesHutCuda.cu:427-431). REF_CPU: G m M / (d^2 + eps^2) along d / |d|,
 * Simulation::force and body_to_quad_force of the CPU paths
 * (Simulation.cpp:74-80, BarnesHut.cpp:130-135). The two agree at eps = 0. */
enum nb_force_kernel { NB_FORCE_PLUMMER = 0, NB_FORCE_REF_CPU = 1 };

/* Opening test of an internal cell at distance d from its centre of mass.
 * SIDE: side^2 < theta^2 d^2 with d^2 = fma(dx, dx, dy * dy)
 * (BarnesHutCuda.cu:413-415, 461-462). DIAG: (w^2 + h^2) / d^2 < theta^2 with
 * d^2 = dx * dx + dy * dy, the cell's diagonal (BarnesHut.cpp:113-115); on a
 * square cell DIAG at theta equals SIDE at theta / sqrt(2). */
enum nb_mac { NB_MAC_SIDE = 0, NB_MAC_DIAG = 1 };

/* Quadtree cells. SQUARE: the true bounding box squared about its shorter
 * axis, split on 64-bit Morton digits (BarnesHutCuda.cu:54-132, 648-680).
 * REF_CPU: the reference CPU quadtree (Quadtree.cpp:9-152): the box of
 * get_universe_boundaries, whose `else if` lets body 0 never raise the
 * maximum, kept rectangular, each cell split at its centre computed as the
 * reference does (position + size / 2, halved sizes), centres of mass summed
 * in the reference's order; leaves below level 30 hold one body, as in the
 * reference (it has no depth cap: bodies closer than 2^-30 of the box stay
 * together in one level-30 leaf here). Needs mac = DIAG and one system. */
enum nb_cells { NB_CELLS_SQUARE = 0, NB_CELLS_REF_CPU = 1 };

/* Update order. KD = kick then drift (BarnesHut.cpp:78-79, BarnesHutCuda.cpp:35-36);
 * DK = drift then kick (AllPairs.cpp:25-26). */
enum nb_integrator { NB_INTEG_KD = 0, NB_INTEG_DK = 1 };

/* Far field of an accepted cell: the reference's monopole, or the monopole
 * plus the learned correction (nb_mlp_*; GROUP traversal only). */
enum nb_far_field { NB_FAR_MONOPOLE = 0, NB_FAR_LEARNED = 1 };

/* Barnes-Hut traversal. EXACT: every body applies the reference's opening
 * test to every cell it reaches (per-body MAC of BarnesHutCuda.cu:461-471), so
 * each body gets exactly the reference's force terms; a wavefront of 64
 * Morton-adjacent bodies walks the tree together with one shared stack of
 * (cell, lane mask) entries. GROUP: the same wavefront accepts a cell only
 * when the test holds for the whole group's bounding box (a conservative
 * superset of every member's test), collects the accepted cells and leaf
 * bodies in an LDS interaction list, then every lane sums the whole list; a
 * group whose list grows past group_budget (or whose walk runs unusually
 * long) is split into pieces of 8 bodies, and a piece into single bodies,
 * shared with other wavefronts and walked again; a single body's group test
 * is exactly its own opening test.
 * GROUP applies more (never fewer) force terms per body than EXACT at the
 * same theta, so its force error is no larger. */
enum nb_traversal { NB_TRAV_EXACT = 0, NB_TRAV_GROUP = 1 };

typedef struct nb_params {
    double G;           /* gravitational constant; the reference hard-codes
                           6.67430e-11 (Constants.hpp:17) */
    double dt;          /* "timestep" (Config.cpp:30) */
    double theta;       /* "theta" (Config.cpp:33); opening test: enum nb_mac */
    double softening;   /* absolute Plummer epsilon; the reference derives it from
                           "softening_factor" (Simulation.cpp:82-126) */
    int32_t algorithm;  /* enum nb_algorithm */
    int32_t integrator; /* enum nb_integrator */
    int32_t self_guard; /* 1: never accept a cell that contains the body itself
                           (the reference can, for theta > 1/sqrt(2)) */
    int32_t far_field;  /* enum nb_far_field */
    int32_t use_graph;  /* 1: steps replay a captured hipGraph (default) */
    int32_t traversal;  /* enum nb_traversal */
    int32_t group_budget; /* GROUP: interaction-list entries after which a group
                             is split into pieces and walked again (0 = 16384);
                             for theta < 0.5 it is scaled by 0.5/theta, as
                             lists grow */
    int32_t cell_capacity; /* LEARNED: initial accepted-cell entries per group of
                              64 bodies (0 = four times the busiest wavefront
                              of a first walk at creation, at least 4096); a
                              step that needs more halts, the regions grow (at
                              least double) and the step runs again */
    int32_t force_kernel; /* enum nb_force_kernel */
    int32_t mac;          /* enum nb_mac */
    int32_t cells;        /* enum nb_cells */
    int32_t walk_stack;   /* GROUP: LDS stack entries of a traversal wavefront,
                             1024 or 2048 (0 = 2048); fewer let more wavefronts
                             share a CU, more let a wavefront take up to 64
                             cells per step at any depth */
    int32_t walk_bodies;  /* GROUP: bodies per traversal wavefront, 64 (one per
                             lane) or 32 (two lanes per body, each summing half
                             of the interaction list; smaller groups accept
                             cells sooner) (0 = 32 below 524,288 bodies
                             with the monopole far field, else 64) */
    int32_t list_flush;   /* GROUP: interaction-list entries held in LDS before
                             they are summed, 64 .. 192 (0 = 128) */
    int32_t quad_near;    /* LEARNED: an accepted cell whose centre of mass lies
                             within quad_near half-diagonals of a traversal
                             subgroup's box from the box centre adds its
                             quadrupole term at every body's own position; a
                             farther cell adds it once at the centre, moved to
                             each body by its Jacobian (0 = 8; -1 = every cell
                             at every body) */
} nb_params;

typedef struct nb_info {
    uint64_t n;               /* bodies */
    uint64_t iteration;       /* completed steps (Simulation::Stats::iteration) */
    double simulated_time;    /* sum of dt over completed steps */
    uint64_t tree_nodes;      /* node slots of the last tree (1 + 4 * internal) */
    uint64_t internal_nodes;  /* internal nodes of the last tree */
    uint64_t node_capacity;   /* internal-node capacity currently allocated */
    uint64_t device_bytes;    /* device memory held by the context */
    uint64_t far_field_cells; /* accepted cells the network corrected per force
                                 evaluation (far_field = LEARNED): the last
                                 nb_compute_forces, or the mean over the last
                                 nb_step call; the network runs once per cell
                                 and traversal subgroup */
    uint64_t quad_terms;      /* of the same force evaluations (far_field =
                                 LEARNED): body terms that also added the cell's
                                 softened quadrupole term at the body (the
                                 cells near each traversal subgroup) */
    uint64_t regrowths;       /* buffers regrown after a step overflowed one
                                 (the step halted and ran again), since creation */
    uint64_t graph_captures;  /* times the steps were captured as graphs (at
                                 creation, after a regrowth, a tuning, or a new
                                 cut of the tree's deepest levels; ~ms each) */
} nb_info;

/* Timing phases reported by nb_get_timing (milliseconds, summed over the
 * steps run since timing was enabled). */
enum nb_phase {
    NB_PHASE_BBOX_MORTON = 0, /* bbox + Morton keys */
    NB_PHASE_SORT = 1,        /* LSD radix sort + gather */
    NB_PHASE_BUILD = 2,       /* quadtree topology + mass/COM */
    NB_PHASE_FORCE = 3,       /* traversal (or all-pairs) + kick */
    NB_PHASE_DRIFT = 4,       /* drift */
    NB_PHASE_FAR_FIELD = 5,   /* learned far-field network + kick (LEARNED only) */
    NB_PHASE_COUNT = 6
};

/* Exported tree slot for parity checks (canonical topology + COM). */
typedef struct nb_node {
    int32_t level;
    int32_t bstart, bend; /* body range in sorted order, -1/-1 = empty slot */
    int32_t c0;           /* first child slot, -1 = leaf, -2 = empty */
    double mass, cx, cy, wsq;
    float ixx, ixy, iyy;  /* second moments about the COM */
    int32_t child_mask;   /* bit q set = child slot c0+q holds bodies */
    float t3[4];          /* third moments about the COM {xxx, xxy, xyy, yyy}: kept
                             for the learned far field (else zero) */
} nb_node;

/* Fills *p with the defaults: G=1, dt=1e-3, theta=0.5, softening=0.01,
 * Barnes-Hut, KD, no self guard, monopole, graphs on, EXACT traversal,
 * Plummer force, SIDE test, SQUARE cells (the reference CUDA path). */
int nb_params_default(nb_params* p);

/* Replaces: the BarnesHutCuda constructor + init_device_resources
 * (BarnesHutCuda.cpp:6-10, BarnesHutCuda.cu:550-606). Copies the host arrays
 * to the GPU; the context owns every device buffer. n may be 0. */
int nb_create(const nb_params* p, uint64_t n, const double* mass, const double* pos_xy,
              const double* vel_xy, nb_ctx** ctx_out);

/* Batched ensemble: one context for `nsys` independent systems (no force
 * between systems), system s being the bodies [offsets[s], offsets[s+1]) of
 * the arrays (offsets rise strictly from 0 to n = offsets[nsys]; nsys <=
 * 65,536). Each step runs every system through one set of launches: one
 * sort (two extra passes on the system number), one tree build (one root per
 * system) and one traversal launch over all traversal groups; each system's
 * result equals a context of that system alone, bit for bit. No counterpart
 * in the reference, which simulates one universe (main.cpp:26-38). Needs
 * algorithm = BARNES_HUT and traversal = GROUP. nb_get_state and
 * nb_compute_forces use the input order of all n bodies. */
int nb_create_batch(const nb_params* p, uint32_t nsys, const uint64_t* offsets,
                    const double* mass, const double* pos_xy, const double* vel_xy,
                    nb_ctx** ctx_out);

/* Replaces: k iterations of BarnesHutCuda::simulate's loop body
 * (BarnesHutCuda.cpp:30-40) / AllPairsSim::simulate (AllPairs.cpp:23-29) /
 * BarnesHut::simulate (BarnesHut.cpp:48-69). Returns after the GPU finished. */
int nb_step(nb_ctx* ctx, uint64_t k);

/* Replaces: sync_from_gpu_bodies (BarnesHutCuda.cu:632-646): the state in the
 * caller's original body order. Either pointer may be NULL. */
int nb_get_state(nb_ctx* ctx, double* pos_xy, double* vel_xy);

/* Asynchronous, double-buffered state snapshot for rendering and output, in
 * place of the reference's blocking per-step scatter + copy
 * (scatter_to_original_order + sync_from_gpu_bodies, BarnesHutCuda.cu:498-506,
 * 632-646). nb_snapshot_begin un-permutes the current state into the
 * caller's body order on the GPU (ordered after every step already issued)
 * and copies it to pinned host memory on a second stream, then returns at
 * once with a ticket (1, 2, 3, ...); later nb_step calls run while the copy
 * is in flight. Two slots alternate: ticket t stays readable until ticket
 * t + 2 is begun. The first call allocates the slots (64 B per body on the
 * GPU, 64 B per body pinned on the host). */
int nb_snapshot_begin(nb_ctx* ctx, uint64_t* ticket);
/* *ready = 1 when the snapshot's copy has finished, else 0. */
int nb_snapshot_query(nb_ctx* ctx, uint64_t ticket, int* ready);
/* Waits for the copy and points *pos_xy / *vel_xy at the snapshot (n
 * interleaved x,y FP64 each, original order; either may be NULL), valid
 * until ticket + 2 is begun or the context is destroyed; *iteration (may be
 * NULL) is the step count the snapshot was taken at. A ticket that was never
 * issued or whose slot was reused is NB_ERR_INVALID_ARG. */
int nb_snapshot_wait(nb_ctx* ctx, uint64_t ticket, const double** pos_xy,
                     const double** vel_xy, uint64_t* iteration);

/* Accelerations (G included) of the current state in original order, and the
 * number of force terms per body (may be NULL) and in total (may be NULL).
 * Replaces the force half of compute_force (BarnesHutCuda.cu:382-486) without
 * the kick. */
int nb_compute_forces(nb_ctx* ctx, double* acc_xy, uint32_t* interactions, uint64_t* total);

/* Replaces: Simulation::set_timestep (Simulation.cpp:50-52); the new dt is
 * used from the next step on (the latch of Simulation.cpp:70). */
int nb_set_dt(nb_ctx* ctx, double dt);

int nb_get_info(nb_ctx* ctx, nb_info* info);

/* Per-phase GPU timing with hipEvents on the context's stream; enabling it
 * launches the step eagerly (no graph) so every phase can be bracketed. */
int nb_set_timing(nb_ctx* ctx, int enable);
int nb_get_timing(nb_ctx* ctx, double* phase_ms, int nphase, uint64_t* steps_timed);

/* Replaces: ~BarnesHutCuda + release_device_resources (BarnesHutCuda.cu:608-630). */
void nb_destroy(nb_ctx* ctx);

const char* nb_last_error(const nb_ctx* ctx);
const char* nb_status_string(int status);

/* ---- parity hooks (the stages of one step, exposed for tests) ---- */

/* Runs bbox -> Morton -> sort -> gather -> tree on the current state without
 * moving any body, then copies out the squared box {x_min, x_max, y_min,
 * y_max}, the sorted keys, the original index of each sorted body, and the
 * tree slot count. Any output pointer may be NULL. */
int nb_debug_build(nb_ctx* ctx, double* box4, uint64_t* sorted_keys, uint32_t* orig_index,
                   uint64_t* node_count);

/* Test hook: the context's next `count` device allocations fail as if the
 * device were out of memory (regrowths of the node pool go through them).
 * After a failed regrowth the context keeps its state for nb_get_state and
 * snapshots, and every call that would run a step returns
 * NB_ERR_OUT_OF_MEMORY. */
int nb_debug_fail_alloc(nb_ctx* ctx, uint32_t count);

/* Test hook: the tree's mass/COM reduction takes the levels from `com_deep`
 * to 29 in one single-workgroup launch and the levels above one launch each.
 * A context picks com_deep from its first tree (3 + the deepest level holding
 * internal nodes) and picks it again, capturing its steps again, when the
 * single-workgroup launch met a level of more than 2,048 internal nodes.
 * `set` in 0..30 replaces it (-1 leaves it); any value gives the same tree.
 * *current (if not NULL) receives the value in use afterwards. */
int nb_debug_com_deep(nb_ctx* ctx, int32_t set, int32_t* current);

/* Copies the tree built by the last step or nb_debug_build. */
int nb_debug_tree(nb_ctx* ctx, nb_node* nodes, uint64_t cap);

/* Stateless FP64 direct sum on the GPU, the reference for force errors:
 * acc[t] = G * sum_{j != targets[t]} m_j (r_j - r_t) / (|r|^2 + eps^2)^{3/2}. */
int nb_direct_sum_f64(uint64_t n, const double* pos_xy, const double* mass, double G,
                      double softening, uint64_t m, const int64_t* targets, double* acc_xy);

/* No counterpart in the reference. The FP64 direct sum of nb_direct_sum_f64
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): pos_xy,
 * mass, targets and acc_xy are DEVICE arrays of the current device, the work
 * is enqueued on `stream` (a hipStream_t; NULL = the default stream) and the
 * call returns without waiting. It allocates nothing: the partial sums live
 * in a fixed device array of NB_NOALLOC_SCRATCH_BYTES per GPU (held until the
 * process ends), so m may be at most NB_NOALLOC_SCRATCH_BYTES / 16; larger m
 * is NB_ERR_INVALID_ARG. Results equal nb_direct_sum_f64's bit for bit (the
 * same j split, kernels and summation order; the split's partials always fit
 * the array). A target outside [0, n) yields NaN. Calls that overlap on one
 * GPU share the array and race: serialise them (one stream, or events). */
#define NB_NOALLOC_SCRATCH_BYTES (16u << 20)
int nb_direct_sum_f64_noalloc(uint64_t n, const double* pos_xy, const double* mass, double G,
                              double softening, uint64_t m, const int64_t* targets,
                              double* acc_xy, void* stream);

/* Autotuner: times one force evaluation of the current state (sort, tree and
 * traversal [+ far field], or the all-pairs kernel; mean of `repeats`) for
 * each setting of the context's launch tunables and keeps the fastest -- a
 * setting replaces the current one only when it is at least 3 % faster. GROUP
 * traversal, one tunable at a time in this order: bodies per wavefront (64,
 * 32), LDS stack depth (2048, 1024), list flush size (64, 128, 192); the list
 * budget (nb_params.group_budget) is left as set, since the state it is timed
 * on does not predict the steps that follow. All-pairs: bodies per thread
 * (2, 4, 8) x the split of the j range (1 .. 64). The state is not advanced;
 * the captured step graphs are rebuilt. Tune a state a few steps evolved:
 * the walk of an initial state is not that of later steps. No counterpart in
 * the reference. */
typedef struct nb_tuning {
    int32_t group_budget; /* list budget in use (GROUP traversal; not tuned) */
    int32_t ap_rows;      /* chosen all-pairs bodies per thread */
    int32_t ap_split;     /* chosen all-pairs j split */
    int32_t candidates;   /* settings timed */
    double best_ms;       /* force evaluation with the chosen setting (mean) */
    double default_ms;    /* force evaluation with the setting before tuning */
    int32_t walk_bodies;  /* chosen bodies per traversal wavefront */
    int32_t walk_stack;   /* chosen LDS stack entries */
    int32_t list_flush;   /* chosen list flush size */
    int32_t reserved;
} nb_tuning;

int nb_autotune(nb_ctx* ctx, uint32_t repeats, nb_tuning* out);

/* ---------------------------------------------------------------------------
 * Learned far-field correction. No counterpart in the reference: its force
 * is the monopole of every accepted cell (BarnesHutCuda.cu:461-471). The
 * correction of an accepted cell, a_exact - a_monopole, is approximated by
 * the cell's softened quadrupole term, computed from its second moments, plus
 * s/R times a network (12 -> 64 -> 64 -> 2, ReLU, BF16 weights and
 * activations on MFMA, FP32 accumulation) trained on the rest over s/R, from
 * the cell's offset, size/distance, mass fraction, second moments and third
 * moments (the octupole the rest is led by); a context with far_field =
 * NB_FAR_LEARNED adds it for every accepted internal cell. Parameters are
 * FP32 in the order W1[64][12], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]
 * (NB_MLP_PARAMS values); a new context holds zeros (the quadrupole term
 * alone).
 * ------------------------------------------------------------------------- */
#define NB_MLP_PARAMS 5122
#define NB_MLP_INPUTS 12

int nb_mlp_set_weights(nb_ctx* ctx, const float* params, uint64_t count);
int nb_mlp_get_weights(nb_ctx* ctx, float* params, uint64_t count);

/* He-normal initial weights from `seed`, zero biases, Adam state cleared. */
int nb_mlp_init(nb_ctx* ctx, uint64_t seed);

/* Training pairs from the current state (Barnes-Hut contexts): builds the
 * tree, walks it from `sample_bodies` seeded bodies with the per-body opening
 * test at `theta` (cells holding the body are never accepted), and keeps
 * every accepted internal cell of at most 8,192 bodies, up to `max_pairs`.
 * Each pair gets its NB_MLP_INPUTS inputs and its target, the exact FP64 Plummer
 * acceleration of the cell's bodies minus the cell's monopole and minus its
 * softened quadrupole term, in the network's output units, over s/R (the
 * input f2).
 * *pairs_out = pairs held. */
int nb_mlp_generate(nb_ctx* ctx, double theta, uint32_t sample_bodies, uint64_t max_pairs,
                    uint64_t seed, uint64_t* pairs_out);

/* Copies up to `cap` held pairs: features [k][12], targets [k][2], the sample
 * body (index in the tree's sorted order) and the cell's tree slot. Any
 * output may be NULL; *count = pairs held. */
int nb_mlp_get_pairs(nb_ctx* ctx, float* features, float* targets, uint32_t* body,
                     uint32_t* cell, uint64_t cap, uint64_t* count);

/* Replaces the held pairs with caller data. */
int nb_mlp_set_pairs(nb_ctx* ctx, const float* features, const float* targets, uint64_t count);

/* `steps` Adam steps (beta 0.9/0.999) on seeded minibatches of `batch` held
 * pairs, loss = mean over the batch of the squared error summed over the two
 * outputs. *last_loss = loss of the last minibatch. */
int nb_mlp_train(nb_ctx* ctx, uint32_t steps, uint32_t batch, float lr, uint64_t seed,
                 float* last_loss);

/* Gradient of that loss over exactly the given pairs (which become the held
 * pairs) at the current weights, from the training kernels. */
int nb_mlp_grad(nb_ctx* ctx, const float* features, const float* targets, uint64_t count,
                float* grad, float* loss);

/* Network outputs for `count` input rows on MFMA: out [count][2]. */
int nb_mlp_forward(nb_ctx* ctx, const float* features, uint64_t count, float* out);

/* Bookkeeping of the learned far field in the last force evaluation, for
 * tests. Groups are up to 64 consecutive bodies of the tree's sorted order,
 * never across two systems (*groups_out of them; ceil(n/64) with one
 * system). Per group: nsub[g] finished subgroups; per subgroup
 * slot (64 per group): its box centre sub_centre[2k..2k+1], its range
 * {start, count} in cell_slots, the tree slots of all regions (groups x
 * *cell_cap_out entries, indexed from 0), and sub_thr[k], the squared
 * distance from the centre beyond which a cell's quadrupole term is taken at
 * the centre (nb_params.quad_near); per sorted body: its subgroup and its
 * offset from that subgroup's centre. Any output may be NULL. */
int nb_debug_far_field(nb_ctx* ctx, uint32_t* nsub, double* sub_centre, uint32_t* sub_range,
                       double* sub_thr, uint32_t* cell_slots, int32_t* body_sub,
                       float* body_off, uint32_t* cell_cap_out, uint32_t* groups_out);

#ifdef __cplusplus
}
#endif
#endif
