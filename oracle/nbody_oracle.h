/*
 * nbody_oracle.h — CPU restatement of the reference's 2-D N-body step.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (the HIP library, the C++
 * host adapter) links or calls this code. It is used by tests/ as the oracle
 * and by bench.py's cpu_baseline leg as the timed CPU baseline.
 *
 * Three parts:
 *   A. Reference CPU semantics: Quadtree (non-square bbox with the `else if`
 *      quirk, rect split, DFS node numbering), BarnesHut (per-body DFS, diagonal
 *      MAC, position-equality self skip, master/worker chunking), AllPairsSim
 *      (drift then symmetric kick), Simulation::force.
 *   B. Reference CUDA semantics: squared bbox, 64-bit Morton codes, stable sort,
 *      level-synchronous quadtree topology (all four child slots materialised,
 *      level cap 30), bottom-up mass/COM, per-body DFS force with the side MAC,
 *      kick then drift.
 *   C. FP64 direct sum with compensated (Neumaier) accumulation, the
 *      reference for every force error.
 * Parity status of each function is listed in the header comment of
 * nbody_oracle.c.
 */
#ifndef NBODY_ORACLE_H
#define NBODY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- Part A: reference CPU semantics ---------------- */

/* Force kernels for the CPU restatement.
 * OR_FORCE_REF_CPU: G*m_a*m_b/(d^2+eps^2) along the unit vector (Simulation.cpp:74-80).
 * OR_FORCE_PLUMMER: G*m_a*m_b*d/(d^2+eps^2)^{3/2} (BarnesHutCuda.cu:427-431). */
enum { OR_FORCE_REF_CPU = 0, OR_FORCE_PLUMMER = 1 };

/* Run `steps` threaded Barnes-Hut steps in place (BarnesHut.cpp:48-88).
 * pos/vel interleaved x,y. Returns 0, or -1 on invalid threads (BarnesHut.cpp:10-13).
 * *seconds_out (may be NULL) receives the steady-clock wall time of the steps. */
int or_refcpu_bh_run(uint64_t n, double* pos, double* vel, const double* mass, double G,
                     double dt, double theta, double eps2, int force_kernel, int threads,
                     uint64_t steps, double* seconds_out);

/* Accelerations (force / m_i, G included) the reference CPU BH would apply to
 * every body of the current state (one tree build + one traversal, no update).
 * interactions (may be NULL) receives the per-body count of force terms. */
int or_refcpu_bh_accel(uint64_t n, const double* pos, const double* mass, double G, double theta,
                       double eps2, int force_kernel, double* acc, uint64_t* interactions);

/* Reference quadtree (Quadtree.cpp:9-152) exported as flat arrays for tests:
 * returns the node count; arrays of length >= returned count (call with cap=0
 * to get the count). rect = {x, y, w, h}; com = {x, y}. */
int64_t or_refcpu_quadtree(uint64_t n, const double* pos, const double* mass, int64_t cap,
                           uint32_t* top_left_idx, uint32_t* body_count, double* rect,
                           double* com, double* total_mass);

/* Non-square bbox with the reference's `else if` quirk (Quadtree.cpp:9-31): {x, y, w, h}. */
void or_refcpu_bbox(uint64_t n, const double* pos, double rect_out[4]);

/* AllPairsSim steps in place: drift then symmetric kick (AllPairs.cpp:23-49). */
void or_refcpu_allpairs_run(uint64_t n, double* pos, double* vel, const double* mass, double G,
                            double dt, double eps2, int force_kernel, uint64_t steps);

/* compute_plummer_softening exact branch (Simulation.cpp:82-102): factor * sum_{i<j} d_ij / n.
 * Only the deterministic branch (n*n <= max_samples) is restated; returns -1 otherwise. */
double or_plummer_softening_exact(uint64_t n, const double* pos, double factor, double max_samples);

/* ---------------- Part B: reference CUDA semantics ---------------- */

typedef struct or_node {
    int32_t level;
    int32_t bstart, bend;     /* -1,-1 for an empty slot */
    int32_t c0;               /* first child slot, -1 for a leaf or empty slot */
    uint64_t prefix;          /* key >> (64 - 2*level), 0 for the root */
    double mass, cx, cy, wsq;
    double ixx, ixy, iyy;     /* second moments about the COM */
} or_node;

/* True min/max then squared about the midpoint of the shorter axis
 * (BarnesHutCuda.cu:78-114, 648-680). box = {x_min, x_max, y_min, y_max}. */
void or_bbox_square(uint64_t n, const double* pos, double box[4]);

/* 64-bit Morton keys, x in odd bits (BarnesHutCuda.cu:54-69, 117-132). */
void or_morton(uint64_t n, const double* pos, const double box[4], uint64_t* keys);
uint64_t or_morton_one(double x, double y, const double box[4]);
uint64_t or_spread_bits(uint32_t v);

/* Stable ascending sort by key; perm[i] = source index of the i-th key
 * (semantics of cub::DeviceRadixSort::SortPairs, BarnesHutCuda.cu:691-693). */
void or_stable_sort(uint64_t n, const uint64_t* keys, uint32_t* perm);

/* Level-synchronous topology + bottom-up COM over sorted keys/bodies
 * (BarnesHutCuda.cu:153-344, 713-810). Nodes are numbered breadth-first; every
 * internal node owns four consecutive child slots. Returns the node count
 * (call with cap=0 to size). */
int64_t or_tree_build(uint64_t n, const uint64_t* skeys, const double* spos, const double* smass,
                      double root_wsq, int64_t cap, or_node* nodes);

/* Per-body DFS force over an or_tree_build tree (BarnesHutCuda.cu:382-486), FP64.
 * acc excludes G. self_guard=1 refuses to accept an internal cell whose body
 * range holds the body itself. interactions may be NULL. */
void or_tree_forces(uint64_t n, const or_node* nodes, const double* spos, const double* smass,
                    double theta, double eps2, int self_guard, double* acc,
                    uint64_t* interactions);

/* One full CUDA-semantics step in place on state kept in sorted order:
 * bbox -> morton -> stable sort -> gather (pos, vel, mass, orig) -> tree ->
 * force+kick -> drift. orig[i] is the original index of sorted body i. */
int or_model_step(uint64_t n, double* pos, double* vel, double* mass, uint32_t* orig, double G,
                  double dt, double theta, double eps2, int self_guard);

/* ---------------- Part C: FP64 direct sum ---------------- */

/* acc[t] = sum_{j != targets[t]} m_j (r_j - r_t) / (|r|^2 + eps2)^{3/2}, G excluded,
 * Neumaier-compensated FP64 sums. threads >= 1. */
void or_direct_sum(uint64_t n, const double* pos, const double* mass, double eps2, uint64_t m,
                   const int64_t* targets, int threads, double* acc);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
/* ---------------- Part D: far-field MLP (12 -> 64 -> 64 -> 2, ReLU) ----------------
 * Parameters in the canonical order W1[64][12], b1[64], W2[64][64], b2[64],
 * W3[2][64], b3[2] (5,122 values). Loss = mean over the batch of the squared
 * error summed over the 2 outputs. bf16 = 1 rounds weights and hidden
 * activations to bfloat16 (round to nearest even) as the GPU does. */
#define OR_MLP_PARAMS 5122
double or_mlp_forward(const double* params, uint64_t count, const float* x, float* y, int bf16);
double or_mlp_grad(const double* params, uint64_t count, const float* x, const float* t,
                   double* grad, int bf16);
/* Softened quadrupole term of a cell (second moments about its COM) at the
 * point at offset d = com - x: (3/(2R^5))[-tr(I)d - 2Id + 5(d.Id)d/R^2],
 * R^2 = |d|^2 + eps2. */
void or_quadrupole(double dx, double dy, double ixx, double ixy, double iyy, double eps2,
                   double* ax, double* ay);
/* Jacobian J = {dQx/ddx, dQx/ddy, dQy/ddx, dQy/ddy} of or_quadrupole. */
void or_quadrupole_jacobian(double dx, double dy, double ixx, double ixy, double iyy,
                            double eps2, double* J);
/* Network inputs f[12] of a (cell, point) pair (t3 = third moments {T_xxx,
 * T_xxy, T_xyy, T_yyy} about the COM); returns the output scale. */
double or_mlp_features(double dx, double dy, double s, double M, double ixx, double ixy,
                       double iyy, const double* t3, double eps2, double m_root, double ell,
                       double* f);
/* The softened quadrupole term along (u, w) in output units, from the inputs
 * f[8] and rho = r^2 / R^2. */
void or_mlp_quad_units(const double* f, double rho, double* q);
/* Exact minus monopole acceleration of bodies [bs, be] at (x, y). */
void or_cell_correction(const double* pos, const double* mass, int64_t bs, int64_t be, double x,
                        double y, double eps2, double* out);
#ifdef __cplusplus
}
#endif
#endif
