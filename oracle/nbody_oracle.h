/*
 * nbody_oracle.h -- CPU restatement of the reference N-body stage pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product library (libnbody.so) never links or calls it.
 *
 * Every function restates the arithmetic of one reference kernel (OpenCL C in
 * the reference's kernels/ directory) under one of two rule sets:
 *   ORC_RULES_REFERENCE  - the tree shape and opening rule the reference
 *                          kernels define (root hop into slot 0, y half of a
 *                          child box chosen by the x comparison, opening
 *                          threshold (s/2)^2/theta + eps per level);
 *   ORC_RULES_STANDARD   - cubic root cell, consistent y halving, opening
 *                          criterion s/d < theta.
 * Arithmetic that both the oracle and the GPU perform identically (box
 * halving, octant choice, centre-of-mass sums with explicit fmaf, opening
 * thresholds, squared distances) is written so that results are bit-equal;
 * only the reciprocal square root differs (hardware rsqrt on the GPU).
 */
#ifndef NBODY_ORACLE_H
#define NBODY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_RULES_REFERENCE 0
#define ORC_RULES_STANDARD 1
#define ORC_LAW_PLUMMER 0
#define ORC_LAW_REFERENCE_UNIT 1

/* Deepest cell level the tree may hold (root = level 0).  Follows the force
 * kernel's MAX_DEPTH 64 (kernels/calculate_forces.hcl:1). */
#define ORC_MAX_LEVEL 64
#define ORC_DQ_LEN (ORC_MAX_LEVEL + 1)
/* Bytes reserved per body for its root-to-leaf octant path. */
#define ORC_PATH_STRIDE 72

/* Status codes, same values as the product's nb_status. */
#define ORC_OK 0
#define ORC_E_TREE_OVERFLOW 5
#define ORC_E_DEPTH_OVERFLOW 6
#define ORC_E_NONFINITE 7

typedef struct orc_tree orc_tree;

orc_tree* orc_tree_new(int n, int cap);
void orc_tree_free(orc_tree* t);
int orc_tree_ncells(const orc_tree* t);
int orc_tree_max_depth(const orc_tree* t);

int orc_bbox(int n, const float* pos, float* box6);
void orc_root_box(int rules, const float* box6, float* lo3, float* hi3);
int orc_path_of(int rules, const float* p3, const float* box6, int maxlen, uint8_t* digits);
int orc_build(orc_tree* t, int rules, const float* pos, const float* box6);
void orc_summarize(orc_tree* t, const float* pos, const float* mass);
void orc_sort(const orc_tree* t, int32_t* sorted);
/* The force kernel's vote-group order: STANDARD rules, depth-first with children along a
 * 3-D Hilbert curve; REFERENCE rules, the tree order of orc_sort. */
void orc_force_order(const orc_tree* t, int32_t* order);
int orc_export(const orc_tree* t, int32_t* sorted, uint8_t* paths, int32_t* path_len,
               int32_t* cell_level, int32_t* cell_slot, int32_t* cell_count, float* cell_com);
void orc_dq_table(int rules, const float* box6, float theta, float eps, float* dq);
void orc_force_bh(const orc_tree* t, const float* pos, const float* mass, const int32_t* sorted,
                  const float* dq, float eps, float G, int vote, int kahan,
                  const int32_t* groups, int ngroups, float* acc, double* absacc, int64_t* ninter);
void orc_force_bf(int law, float eps, float G, int kahan, int n, const float* pos, const float* mass,
                  const int32_t* targets, int ntargets, float* acc, double* absacc);
void orc_direct_f64(int law, double eps, double G, int n, const float* pos, const float* mass,
                    const int32_t* targets, int ntargets, double* acc);
void orc_euler(int n, float dt, const float* acc, float* vel, float* pos);
void orc_kick(int n, float h, const float* acc, float* vel);
void orc_drift(int n, float dt, const float* vel, float* pos);
/* Collision stage: reflects velocities of bodies closer than 2 x 0.015 (in place). */
void orc_collision(int n, const float* pos, float* vel);
/* FP64 conservation quantities; out[9] = kinetic, potential, px, py, pz, Lx, Ly, Lz, total mass.
 * No counterpart in the reference (it computes no diagnostics). */
void orc_energy_momentum_f64(int law, double eps, double G, int n, const float* pos, const float* vel,
                             const float* mass, double* out);
int orc_bh_step(int rules, int nsteps, int n, int cap, float theta, float eps, float G, float dt,
                int vote, int kahan, float* pos, float* vel, const float* mass, float* acc);
int orc_threads(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
