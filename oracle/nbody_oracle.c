/*
 * nbody_oracle.c -- CPU restatement of the reference N-body stages.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See nbody_oracle.h.  Citations are
 * to the reference's own files (the kernels/ OpenCL sources and src/MyApp.cpp).
 *
 * Pinning (what checks each function against something other than itself) is
 * in tests/test_oracle.py: closed forms, brute-force path enumeration on tiny
 * inputs, FP64 sums, symmetry, invariants; orc_force_order against Skilling's
 * Hilbert index computed independently in the test.
 */
#include "nbody_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct orc_tree {
    int n, cap, ncells, max_depth, status, rules;
    int32_t* child;  /* 8*cap: -1 empty, [0,n) body, n+ci cell */
    int32_t* parent; /* cap */
    int32_t* level;  /* cap */
    int32_t* slot;   /* cap: slot of this cell in its parent */
    int32_t* count;  /* cap: bodies below the cell */
    float* com;      /* 4*cap: centre of mass x,y,z and mass */
    int32_t* depth;  /* n: level of the cell holding the body */
    float root_lo[3], root_hi[3];
};

int orc_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Threads of the parallel loops that follow (the CPU baseline's one-thread figure). */
void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

orc_tree* orc_tree_new(int n, int cap) {
    orc_tree* t = (orc_tree*)calloc(1, sizeof(orc_tree));
    t->n = n;
    t->cap = cap < 1 ? 1 : cap;
    t->child = (int32_t*)malloc(sizeof(int32_t) * 8 * (size_t)t->cap);
    t->parent = (int32_t*)malloc(sizeof(int32_t) * (size_t)t->cap);
    t->level = (int32_t*)malloc(sizeof(int32_t) * (size_t)t->cap);
    t->slot = (int32_t*)malloc(sizeof(int32_t) * (size_t)t->cap);
    t->count = (int32_t*)malloc(sizeof(int32_t) * (size_t)t->cap);
    t->com = (float*)malloc(sizeof(float) * 4 * (size_t)t->cap);
    t->depth = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    return t;
}

void orc_tree_free(orc_tree* t) {
    if (!t) return;
    free(t->child); free(t->parent); free(t->level); free(t->slot);
    free(t->count); free(t->com); free(t->depth); free(t);
}

int orc_tree_ncells(const orc_tree* t) { return t->ncells; }
int orc_tree_max_depth(const orc_tree* t) { return t->max_depth; }

/* Root bounding box: exact componentwise min/max over all bodies
 * (intent of hybrid_reduce_root/parallel_reduce_root, kernels/boundary_reduce.cl:6-159,
 * taken over all N bodies rather than the first 2*WGS the host wiring feeds it,
 * src/MyApp.cpp:282).  -0 is canonicalised to +0 by adding +0.0f. */
int orc_bbox(int n, const float* pos, float* box6) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    int bad = 0;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            float v = pos[3 * i + k];
            if (!isfinite(v)) { bad = 1; continue; }
            if (v < lo[k]) lo[k] = v;
            if (v > hi[k]) hi[k] = v;
        }
    }
    for (int k = 0; k < 3; ++k) { box6[k] = lo[k] + 0.0f; box6[3 + k] = hi[k] + 0.0f; }
    return bad ? ORC_E_NONFINITE : ORC_OK;
}

/* Root cell box.  REFERENCE: the bounding box itself (kernels/build_tree.cl:53-54).
 * STANDARD: the cube of half-edge r = half the largest extent about the box centre. */
void orc_root_box(int rules, const float* box6, float* lo, float* hi) {
    if (rules == ORC_RULES_REFERENCE) {
        for (int k = 0; k < 3; ++k) { lo[k] = box6[k]; hi[k] = box6[3 + k]; }
        return;
    }
    float ex = box6[3] - box6[0], ey = box6[4] - box6[1], ez = box6[5] - box6[2];
    float r = 0.5f * fmaxf(fmaxf(ex, ey), ez);
    for (int k = 0; k < 3; ++k) {
        float c = (box6[k] + box6[3 + k]) * 0.5f;
        lo[k] = c - r;
        hi[k] = c + r;
    }
}

/* One octant choice plus the box halving that follows it.
 * Octant = 4*[y<=cy] + 2*[x<=cx] + [z<=cz] (calculateOctantIdx,
 * kernels/build_tree.hcl:14-29; an exact tie, undefined there, goes to the
 * lower half so that the slot agrees with the box halving).  The halving is
 * adjustBoundaryValues (kernels/build_tree.hcl:31-54): under REFERENCE rules the
 * y half follows the x comparison (build_tree.hcl:42). */
static inline int orc_digit(int rules, const float* p, float* lo, float* hi) {
    float cx = (lo[0] + hi[0]) * 0.5f;
    float cy = (lo[1] + hi[1]) * 0.5f;
    float cz = (lo[2] + hi[2]) * 0.5f;
    int gx = p[0] > cx, gy = p[1] > cy, gz = p[2] > cz;
    int d = (gy ? 0 : 4) + (gx ? 0 : 2) + (gz ? 0 : 1);
    if (gx) lo[0] = cx; else hi[0] = cx;
    int ysel = rules == ORC_RULES_REFERENCE ? gx : gy;
    if (ysel) lo[1] = cy; else hi[1] = cy;
    if (gz) lo[2] = cz; else hi[2] = cz;
    return d;
}

/* Digit taken at cell level `lvl`.  REFERENCE rules: the root's only used slot
 * is 0 (build_tree.cl:57-58,65 start at child_path = 0 and do not consult the
 * octant at the root), and the box is not halved for that hop. */
static inline int orc_level_digit(int rules, int lvl, const float* p, float* lo, float* hi) {
    if (rules == ORC_RULES_REFERENCE && lvl == 0) return 0;
    return orc_digit(rules, p, lo, hi);
}

/* The body's octant digits from the root, computed without any tree. */
int orc_path_of(int rules, const float* p3, const float* box6, int maxlen, uint8_t* digits) {
    float lo[3], hi[3];
    orc_root_box(rules, box6, lo, hi);
    for (int l = 0; l < maxlen; ++l) digits[l] = (uint8_t)orc_level_digit(rules, l, p3, lo, hi);
    return maxlen;
}

/* Sequential insertion, same tree as the lock-free build_tree kernel
 * (kernels/build_tree.cl:6-165): descend through cells, place the body in an
 * empty slot, or split an occupied slot by a chain of new cells until the two
 * bodies take different slots (build_tree.cl:105-141).  The resulting topology
 * does not depend on insertion order: a cell exists for every path prefix that
 * at least two bodies share.  Cells may sit at levels 0..ORC_MAX_LEVEL. */
int orc_build(orc_tree* t, int rules, const float* pos, const float* box6) {
    const int n = t->n;
    t->rules = rules;
    t->status = ORC_OK;
    t->ncells = 1;
    t->max_depth = 1; /* build_tree.cl:28 local_max_depth starts at 1 */
    for (int k = 0; k < 8; ++k) t->child[k] = -1;
    t->parent[0] = -1; t->level[0] = 0; t->slot[0] = -1;
    orc_root_box(rules, box6, t->root_lo, t->root_hi);

    for (int b = 0; b < n; ++b) {
        const float* p = pos + 3 * b;
        float lo[3] = {t->root_lo[0], t->root_lo[1], t->root_lo[2]};
        float hi[3] = {t->root_hi[0], t->root_hi[1], t->root_hi[2]};
        int node = 0, lvl = 0;
        int d = orc_level_digit(rules, lvl, p, lo, hi);
        for (;;) {
            int v = t->child[8 * node + d];
            if (v == -1) {
                t->child[8 * node + d] = b;
                t->depth[b] = lvl;
                break;
            }
            if (v >= n) {
                node = v - n;
                ++lvl;
                d = orc_level_digit(rules, lvl, p, lo, hi);
                continue;
            }
            /* split: body q already occupies the slot; it shares b's box so far */
            int q = v;
            const float* pq = pos + 3 * q;
            float qlo[3] = {lo[0], lo[1], lo[2]}, qhi[3] = {hi[0], hi[1], hi[2]};
            int cur = node, cd = d;
            for (;;) {
                if (lvl + 1 > ORC_MAX_LEVEL) { t->status = ORC_E_DEPTH_OVERFLOW; return t->status; }
                if (t->ncells >= t->cap) { t->status = ORC_E_TREE_OVERFLOW; return t->status; }
                int c = t->ncells++;
                for (int k = 0; k < 8; ++k) t->child[8 * c + k] = -1;
                t->parent[c] = cur; t->level[c] = lvl + 1; t->slot[c] = cd;
                t->child[8 * cur + cd] = n + c;
                ++lvl;
                int db = orc_level_digit(rules, lvl, p, lo, hi);
                int dq = orc_level_digit(rules, lvl, pq, qlo, qhi);
                cur = c;
                if (db != dq) {
                    t->child[8 * c + db] = b;
                    t->child[8 * c + dq] = q;
                    t->depth[b] = lvl;
                    t->depth[q] = lvl;
                    break;
                }
                cd = db;
            }
            break;
        }
    }
    int md = 1;
    for (int b = 0; b < n; ++b) if (t->depth[b] > md) md = t->depth[b];
    t->max_depth = md;
    return ORC_OK;
}

/* Bottom-up centre of mass, mass and body count (saturate_tree,
 * kernels/saturate_tree.cl:6-146), summing children in slot order once all are
 * ready (the reference sums late children in LIFO order, saturate_tree.cl:96-120).
 * Children always have larger cell numbers than their parent, so a reverse
 * sweep sees every child first.  COM = (sum fmaf(x_k, m_k, .)) * (1/mass). */
void orc_summarize(orc_tree* t, const float* pos, const float* mass) {
    const int n = t->n;
    for (int c = t->ncells - 1; c >= 0; --c) {
        float m = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
        int cnt = 0;
        for (int k = 0; k < 8; ++k) {
            int v = t->child[8 * c + k];
            if (v < 0) continue;
            float x, y, z, w;
            if (v < n) {
                x = pos[3 * v]; y = pos[3 * v + 1]; z = pos[3 * v + 2]; w = mass[v];
                cnt += 1;
            } else {
                const float* cc = t->com + 4 * (size_t)(v - n);
                x = cc[0]; y = cc[1]; z = cc[2]; w = cc[3];
                cnt += t->count[v - n];
            }
            m = m + w;
            cx = fmaf(x, w, cx);
            cy = fmaf(y, w, cy);
            cz = fmaf(z, w, cz);
        }
        float inv = 1.0f / m;
        float* o = t->com + 4 * (size_t)c;
        o[0] = cx * inv; o[1] = cy * inv; o[2] = cz * inv; o[3] = m;
        t->count[c] = cnt;
    }
}

/* Tree-order body permutation: depth-first, children in slot order.  The
 * reference's sort stage is empty (kernels/sort.cl is 0 bytes; the intent is
 * noted at kernels/calculate_forces.cl:148); this is Burtscher & Pingali's
 * in-order placement driven by the per-cell counts saturate_tree computes. */
/* Force order: the bodies in depth-first order with each cell's children visited
 * along a 3-D Hilbert curve (no counterpart in the reference, whose work-group
 * vote groups hold unsorted bodies, calculate_forces.cl:84-148).  The engine's
 * force kernel forms its vote groups from consecutive bodies of this order:
 * consecutive Hilbert cells share a face, so a group spans less space than a
 * group of the lexicographic (Z-order) tree order.  State machine: 24
 * orientations; ORC_H_ORDER[s] lists the octants (4[y<c]+2[x<c]+[z<c]) in visiting
 * order, ORC_H_NEXT[s][octant] is the child's orientation; the root is state 0.
 * The tables follow the Hilbert index of Skilling ("Programming the Hilbert
 * curve", AIP Conf. Proc. 707, 2004) on the octree's cells. */
static const uint8_t ORC_H_ORDER[24][8] = {
    {7, 6, 2, 3, 1, 0, 4, 5},
    {7, 5, 1, 3, 2, 0, 4, 6},
    {4, 6, 2, 0, 1, 3, 7, 5},
    {1, 0, 2, 3, 7, 6, 4, 5},
    {4, 0, 2, 6, 7, 3, 1, 5},
    {7, 6, 4, 5, 1, 0, 2, 3},
    {7, 3, 1, 5, 4, 0, 2, 6},
    {4, 0, 1, 5, 7, 3, 2, 6},
    {7, 3, 2, 6, 4, 0, 1, 5},
    {2, 3, 1, 0, 4, 5, 7, 6},
    {2, 0, 1, 3, 7, 5, 4, 6},
    {7, 5, 4, 6, 2, 0, 1, 3},
    {1, 0, 4, 5, 7, 6, 2, 3},
    {2, 0, 4, 6, 7, 5, 1, 3},
    {2, 6, 4, 0, 1, 5, 7, 3},
    {4, 5, 1, 0, 2, 3, 7, 6},
    {4, 6, 7, 5, 1, 3, 2, 0},
    {1, 3, 2, 0, 4, 6, 7, 5},
    {4, 5, 7, 6, 2, 3, 1, 0},
    {1, 3, 7, 5, 4, 6, 2, 0},
    {1, 5, 7, 3, 2, 6, 4, 0},
    {1, 5, 4, 0, 2, 6, 7, 3},
    {2, 3, 7, 6, 4, 5, 1, 0},
    {2, 6, 7, 3, 1, 5, 4, 0}
};
static const uint8_t ORC_H_NEXT[24][8] = {
    {0, 1, 0, 2, 3, 4, 5, 6},
    {1, 1, 0, 15, 10, 11, 7, 8},
    {0, 15, 2, 2, 7, 8, 16, 17},
    {12, 19, 3, 4, 0, 2, 3, 20},
    {7, 8, 4, 4, 16, 17, 3, 18},
    {5, 6, 12, 13, 5, 14, 0, 1},
    {6, 6, 7, 8, 5, 9, 10, 11},
    {4, 7, 6, 7, 18, 10, 9, 16},
    {8, 4, 8, 6, 11, 3, 17, 5},
    {6, 9, 13, 22, 14, 9, 1, 15},
    {13, 10, 22, 7, 1, 10, 15, 23},
    {11, 13, 8, 12, 11, 1, 21, 0},
    {3, 20, 5, 14, 12, 13, 12, 19},
    {10, 11, 23, 21, 13, 13, 12, 22},
    {5, 9, 10, 11, 14, 14, 23, 21},
    {1, 15, 2, 15, 4, 18, 6, 9},
    {22, 7, 19, 16, 15, 23, 2, 16},
    {8, 12, 17, 19, 21, 0, 17, 2},
    {19, 22, 4, 18, 2, 15, 20, 18},
    {23, 21, 16, 17, 12, 22, 19, 19},
    {16, 17, 3, 18, 23, 21, 20, 20},
    {11, 3, 17, 5, 21, 20, 21, 14},
    {20, 18, 14, 9, 13, 22, 19, 22},
    {18, 10, 9, 16, 20, 23, 14, 23}
};

void orc_force_order(const orc_tree* t, int32_t* order) {
    const int n = t->n;
    if (n == 0) return;
    if (t->rules == ORC_RULES_REFERENCE) {  /* REFERENCE cells are not boxes in y: tree order */
        orc_sort(t, order);
        return;
    }
    int stk_node[ORC_MAX_LEVEL + 2], stk_next[ORC_MAX_LEVEL + 2], stk_state[ORC_MAX_LEVEL + 2];
    int sp = 0, nb = 0;
    stk_node[0] = 0; stk_next[0] = 0; stk_state[0] = 0;
    while (sp >= 0) {
        const int node = stk_node[sp], i = stk_next[sp], s = stk_state[sp];
        if (i >= 8) { --sp; continue; }
        stk_next[sp] = i + 1;
        const int k = ORC_H_ORDER[s][i];
        const int v = t->child[8 * node + k];
        if (v < 0) continue;
        if (v < n) {
            order[nb++] = v;
        } else {
            ++sp;
            stk_node[sp] = v - n;
            stk_next[sp] = 0;
            stk_state[sp] = ORC_H_NEXT[s][k];
        }
    }
}

void orc_sort(const orc_tree* t, int32_t* sorted) {
    orc_export(t, sorted, NULL, NULL, NULL, NULL, NULL, NULL);
}

/* Canonical export by depth-first preorder: bodies in tree order with their
 * octant paths, cells in preorder with (level, slot in parent, count, COM). */
int orc_export(const orc_tree* t, int32_t* sorted, uint8_t* paths, int32_t* path_len,
               int32_t* cell_level, int32_t* cell_slot, int32_t* cell_count, float* cell_com) {
    const int n = t->n;
    if (n == 0) return 0;
    int* stk_node = (int*)malloc(sizeof(int) * (ORC_MAX_LEVEL + 2));
    int* stk_next = (int*)malloc(sizeof(int) * (ORC_MAX_LEVEL + 2));
    uint8_t digits[ORC_MAX_LEVEL + 2];
    int sp = 0, nb = 0, nc = 0;
    stk_node[0] = 0; stk_next[0] = 0;
    if (cell_level) { cell_level[0] = 0; cell_slot[0] = -1; cell_count[0] = t->count[0];
                      memcpy(cell_com, t->com, 4 * sizeof(float)); }
    nc = 1;
    while (sp >= 0) {
        int node = stk_node[sp];
        int k = stk_next[sp];
        if (k >= 8) { --sp; continue; }
        stk_next[sp] = k + 1;
        int v = t->child[8 * node + k];
        if (v < 0) continue;
        digits[sp] = (uint8_t)k;
        if (v < n) {
            if (sorted) sorted[nb] = v;
            if (paths) {
                memset(paths + (size_t)v * ORC_PATH_STRIDE, 0xff, ORC_PATH_STRIDE);
                memcpy(paths + (size_t)v * ORC_PATH_STRIDE, digits, (size_t)sp + 1);
                path_len[v] = sp + 1;
            }
            ++nb;
        } else {
            int c = v - n;
            if (cell_level) {
                cell_level[nc] = t->level[c];
                cell_slot[nc] = k;
                cell_count[nc] = t->count[c];
                memcpy(cell_com + 4 * (size_t)nc, t->com + 4 * (size_t)c, 4 * sizeof(float));
            }
            ++nc;
            ++sp;
            stk_node[sp] = c;
            stk_next[sp] = 0;
        }
    }
    free(stk_node); free(stk_next);
    return nc;
}

/* Per-level opening thresholds.  REFERENCE (kernels/calculate_forces.cl:45-62):
 * r = half the largest box extent, q0 = r*r/theta (r*r when theta <= 0),
 * q_l = 0.25*q_{l-1}, dq_l = q_l + eps.  STANDARD: q0 = (r/theta)^2 with r the
 * edge of the root's children, dq_l = q_l + eps^2 (accept iff s/d < theta);
 * theta <= 0 opens every cell. */
void orc_dq_table(int rules, const float* box6, float theta, float eps, float* dq) {
    float r = 0.5f * fmaxf(fmaxf(box6[3] - box6[0], box6[4] - box6[1]), box6[5] - box6[2]);
    float q;
    if (rules == ORC_RULES_REFERENCE) {
        q = theta > 0.0f ? r * r / theta : r * r;
        for (int l = 0; l < ORC_DQ_LEN; ++l) {
            if (l) q = 0.25f * q;
            dq[l] = q + eps;
        }
    } else {
        if (!(theta > 0.0f)) {
            for (int l = 0; l < ORC_DQ_LEN; ++l) dq[l] = INFINITY;
            return;
        }
        float s = r / theta;
        float e2 = eps * eps;
        q = s * s;
        for (int l = 0; l < ORC_DQ_LEN; ++l) {
            if (l) q = 0.25f * q;
            dq[l] = q + e2;
        }
    }
}

static inline float orc_d2(float dx, float dy, float dz, float e2) {
    return fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, e2)));
}

/* Barnes-Hut force for vote groups of `vote` consecutive bodies of `sorted` (the caller
 * passes the force order, orc_force_order, to emulate the engine's groups)
 * (calculate_force_local, kernels/calculate_forces.cl:84-148).  A body child is
 * always taken; a cell child at parent level l is taken for the whole group iff
 * every member has eps^2+|d|^2 >= dq[l] (the reference votes with
 * work_group_all, calculate_forces.cl:122), otherwise the group descends.
 * Interaction: acc += d * m * rsqrt(eps^2+|d|^2)^3 (calculate_forces.cl:116-129),
 * finally acc *= G (calculate_forces.cl:148).  Depth-first, slot order, as the
 * reference walks.  groups==NULL evaluates every group.  absacc (optional) gets
 * sum |term| per member in double, ninter (optional) the interactions per member. */
void orc_force_bh(const orc_tree* t, const float* pos, const float* mass, const int32_t* sorted,
                  const float* dq, float eps, float G, int vote, int kahan,
                  const int32_t* groups, int ngroups, float* acc, double* absacc, int64_t* ninter) {
    const int n = t->n;
    if (n == 0) return;
    if (vote < 1) vote = 1;
    const int total_groups = (n + vote - 1) / vote;
    const int ng = groups ? ngroups : total_groups;
    const float e2 = eps * eps;
#pragma omp parallel
    {
        float* ax = (float*)malloc(sizeof(float) * 6 * (size_t)vote);
        double* aa = (double*)malloc(sizeof(double) * (size_t)vote);
        int64_t* ni = (int64_t*)malloc(sizeof(int64_t) * (size_t)vote);
        float* d2v = (float*)malloc(sizeof(float) * (size_t)vote);
        int stk_node[ORC_MAX_LEVEL + 2], stk_next[ORC_MAX_LEVEL + 2];
#pragma omp for schedule(dynamic, 4)
        for (int gi = 0; gi < ng; ++gi) {
            int g = groups ? groups[gi] : gi;
            int first = g * vote;
            int cnt = n - first < vote ? n - first : vote;
            for (int i = 0; i < 6 * cnt; ++i) ax[i] = 0.0f;
            for (int i = 0; i < cnt; ++i) { aa[i] = 0.0; ni[i] = 0; }
            int sp = 0;
            stk_node[0] = 0; stk_next[0] = 0;
            while (sp >= 0) {
                int node = stk_node[sp];
                int k = stk_next[sp];
                if (k >= 8) { --sp; continue; }
                stk_next[sp] = k + 1;
                int v = t->child[8 * node + k];
                if (v < 0) continue;
                float cxv, cyv, czv, cm;
                if (v < n) { cxv = pos[3 * v]; cyv = pos[3 * v + 1]; czv = pos[3 * v + 2]; cm = mass[v]; }
                else { const float* cc = t->com + 4 * (size_t)(v - n); cxv = cc[0]; cyv = cc[1]; czv = cc[2]; cm = cc[3]; }
                int take = 1;
                for (int i = 0; i < cnt; ++i) {
                    int b = sorted[first + i];
                    float dx = cxv - pos[3 * b], dy = cyv - pos[3 * b + 1], dz = czv - pos[3 * b + 2];
                    d2v[i] = orc_d2(dx, dy, dz, e2);
                    if (v >= n && !(d2v[i] >= dq[sp])) take = 0;
                }
                if (take) {
                    for (int i = 0; i < cnt; ++i) {
                        int b = sorted[first + i];
                        float dx = cxv - pos[3 * b], dy = cyv - pos[3 * b + 1], dz = czv - pos[3 * b + 2];
                        float ri = 1.0f / sqrtf(d2v[i]);
                        float F = cm * (ri * ri * ri);
                        float tx = dx * F, ty = dy * F, tz = dz * F;
                        float* a = ax + 6 * i;
                        if (kahan) {
                            float y, s;
                            y = tx - a[3]; s = a[0] + y; a[3] = (s - a[0]) - y; a[0] = s;
                            y = ty - a[4]; s = a[1] + y; a[4] = (s - a[1]) - y; a[1] = s;
                            y = tz - a[5]; s = a[2] + y; a[5] = (s - a[2]) - y; a[2] = s;
                        } else {
                            a[0] = fmaf(dx, F, a[0]);
                            a[1] = fmaf(dy, F, a[1]);
                            a[2] = fmaf(dz, F, a[2]);
                        }
                        if (absacc) aa[i] += fabs((double)F) * sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
                        ni[i] += 1;
                    }
                } else {
                    ++sp;
                    stk_node[sp] = v - n;
                    stk_next[sp] = 0;
                }
            }
            for (int i = 0; i < cnt; ++i) {
                int b = sorted[first + i];
                float* a = ax + 6 * i;
                float sx = kahan ? a[0] - a[3] : a[0];
                float sy = kahan ? a[1] - a[4] : a[1];
                float sz = kahan ? a[2] - a[5] : a[2];
                acc[3 * b] = sx * G; acc[3 * b + 1] = sy * G; acc[3 * b + 2] = sz * G;
                if (absacc) absacc[b] = aa[i] * fabs((double)G);
                if (ninter) ninter[b] = ni[i];
            }
        }
        free(ax); free(aa); free(ni); free(d2v);
    }
}

/* All-pairs force, j ascending, self skipped.
 * PLUMMER: acc += d * m_j * rsqrt(|d|^2+eps^2)^3 (the physical law; same as the
 *          Barnes-Hut leaf interaction, kernels/calculate_forces.cl:116-129).
 * REFERENCE_UNIT: the law the reference's brute-force kernels compute
 *          (update, kernels/brute_force_update.cl:43-62; update_local via
 *          calc_body_body :108-124): n = normalize(p_j - p_i),
 *          acc += n * m_j / (|n|^2 + eps^2) where pow(x, 3/2) is pow(x, 1)
 *          (integer division) and eps is the kernel's constant (0.01 in update,
 *          0.1 in update_local).  normalize(0) = 0. */
void orc_force_bf(int law, float eps, float G, int kahan, int n, const float* pos, const float* mass,
                  const int32_t* targets, int ntargets, float* acc, double* absacc) {
    const int nt = targets ? ntargets : n;
    const float e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 16)
    for (int ti = 0; ti < nt; ++ti) {
        int i = targets ? targets[ti] : ti;
        float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        float a[6] = {0, 0, 0, 0, 0, 0};
        double aa = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            float dx = pos[3 * j] - px, dy = pos[3 * j + 1] - py, dz = pos[3 * j + 2] - pz;
            float fx, fy, fz, tx, ty, tz;
            if (law == ORC_LAW_PLUMMER) {
                float d2 = orc_d2(dx, dy, dz, e2);
                float ri = 1.0f / sqrtf(d2);
                float F = mass[j] * (ri * ri * ri);
                fx = dx; fy = dy; fz = dz;
                tx = dx * F; ty = dy * F; tz = dz * F;
                aa += fabs((double)F) * sqrt((double)dx * dx + (double)dy * dy + (double)dz * dz);
                if (!kahan) { a[0] = fmaf(fx, F, a[0]); a[1] = fmaf(fy, F, a[1]); a[2] = fmaf(fz, F, a[2]); }
            } else {
                float l2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                float nx = 0.0f, ny = 0.0f, nz = 0.0f;
                if (l2 > 0.0f) { float ri = 1.0f / sqrtf(l2); nx = dx * ri; ny = dy * ri; nz = dz * ri; }
                float d2 = fmaf(nz, nz, fmaf(ny, ny, fmaf(nx, nx, e2)));
                float s = mass[j] / d2;
                tx = nx * s; ty = ny * s; tz = nz * s;
                aa += fabs((double)s);
                if (!kahan) { a[0] = fmaf(nx, s, a[0]); a[1] = fmaf(ny, s, a[1]); a[2] = fmaf(nz, s, a[2]); }
            }
            if (kahan) {
                float y, s2;
                y = tx - a[3]; s2 = a[0] + y; a[3] = (s2 - a[0]) - y; a[0] = s2;
                y = ty - a[4]; s2 = a[1] + y; a[4] = (s2 - a[1]) - y; a[1] = s2;
                y = tz - a[5]; s2 = a[2] + y; a[5] = (s2 - a[2]) - y; a[2] = s2;
            }
        }
        float sx = kahan ? a[0] - a[3] : a[0];
        float sy = kahan ? a[1] - a[4] : a[1];
        float sz = kahan ? a[2] - a[5] : a[2];
        acc[3 * i] = sx * G; acc[3 * i + 1] = sy * G; acc[3 * i + 2] = sz * G;
        if (absacc) absacc[i] = aa * fabs((double)G);
    }
}

/* FP64 direct sum (the error yardstick).  Same laws as orc_force_bf. */
void orc_direct_f64(int law, double eps, double G, int n, const float* pos, const float* mass,
                    const int32_t* targets, int ntargets, double* acc) {
    const int nt = targets ? ntargets : n;
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 16)
    for (int ti = 0; ti < nt; ++ti) {
        int i = targets ? targets[ti] : ti;
        double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        double ax = 0, ay = 0, az = 0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double dx = pos[3 * j] - px, dy = pos[3 * j + 1] - py, dz = pos[3 * j + 2] - pz;
            if (law == ORC_LAW_PLUMMER) {
                double d2 = dx * dx + dy * dy + dz * dz + e2;
                double F = mass[j] / (d2 * sqrt(d2));
                ax += dx * F; ay += dy * F; az += dz * F;
            } else {
                double l = sqrt(dx * dx + dy * dy + dz * dz);
                if (l > 0) { dx /= l; dy /= l; dz /= l; } else { dx = dy = dz = 0; }
                double s = mass[j] / (dx * dx + dy * dy + dz * dz + e2);
                ax += dx * s; ay += dy * s; az += dz * s;
            }
        }
        acc[3 * ti] = ax * G; acc[3 * ti + 1] = ay * G; acc[3 * ti + 2] = az * G;
    }
}

/* Semi-implicit Euler as in the reference's EULER branch
 * (kernels/calculate_forces.cl:169-182, brute_force_update.cl:78-86):
 * a <- acc; v <- v + a*dt; x <- x + v*dt (fused multiply-add). */
void orc_euler(int n, float dt, const float* acc, float* vel, float* pos) {
    for (int i = 0; i < 3 * n; ++i) {
        vel[i] = fmaf(acc[i], dt, vel[i]);
        pos[i] = fmaf(vel[i], dt, pos[i]);
    }
}

/* Kick-drift-kick leapfrog pieces (the reference's LEAPFROG branch,
 * calculate_forces.cl:153-167, never stores a; this is the intended scheme). */
void orc_kick(int n, float h, const float* acc, float* vel) {
    for (int i = 0; i < 3 * n; ++i) vel[i] = fmaf(acc[i], h, vel[i]);
}
void orc_drift(int n, float dt, const float* vel, float* pos) {
    for (int i = 0; i < 3 * n; ++i) pos[i] = fmaf(vel[i], dt, pos[i]);
}

/* FP64 conservation diagnostics: out = {kinetic, potential, px, py, pz}. */
/* Collision stage (update_collision_local + calc_collision_traj, kernels/collision_update.cl:3-83),
 * as intended: spheres of radius 0.015; for each body i and every other body j in index order,
 * if distance(p_i, p_j) <= r_i + r_j then v_i -= 2 (n.v_i) n with n = (p_j - p_i)/|p_j - p_i|,
 * using v_i as already reflected by earlier j (collision_update.cl:62-64).  The reference tile loop
 * (:48-70) re-reads tile 0 and tests the wrong partner for self; this sweeps all pairs once.
 * A coincident pair has no direction and leaves v_i unchanged. */
void orc_collision(int n, const float* pos, float* vel) {
    const float r = 0.015f, rsum = r + r;
    float* out = (float*)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        float vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            float dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1], dz = pos[3 * j + 2] - pos[3 * i + 2];
            float l2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            if (l2 == 0.0f) continue;
            float dist = sqrtf(l2);
            if (dist > rsum) continue;
            float nx = dx / dist, ny = dy / dist, nz = dz / dist;
            float k2 = 2.0f * fmaf(nz, vz, fmaf(ny, vy, nx * vx));
            vx = fmaf(-k2, nx, vx); vy = fmaf(-k2, ny, vy); vz = fmaf(-k2, nz, vz);
        }
        out[3 * i] = vx; out[3 * i + 1] = vy; out[3 * i + 2] = vz;
    }
    memcpy(vel, out, sizeof(float) * 3 * (size_t)n);
    free(out);
}

void orc_energy_momentum_f64(int law, double eps, double G, int n, const float* pos, const float* vel,
                             const float* mass, double* out) {
    /* out: kinetic, potential, px, py, pz, Lx, Ly, Lz, total mass.
     * PLUMMER pair potential -G m_i m_j / sqrt(r^2 + eps^2) (the force law's potential);
     * REFERENCE_UNIT (constant-magnitude pull G m_j / (1 + eps^2) along the unit vector,
     * brute_force_update.cl:43-62) has pair potential +G m_i m_j r / (1 + eps^2). */
    double ke = 0, pe = 0, px = 0, py = 0, pz = 0, lx = 0, ly = 0, lz = 0, mt = 0;
    for (int i = 0; i < n; ++i) {
        double m = mass[i], vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        ke += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        px += m * vx; py += m * vy; pz += m * vz;
        lx += m * (y * vz - z * vy); ly += m * (z * vx - x * vz); lz += m * (x * vy - y * vx);
        mt += m;
    }
#pragma omp parallel for reduction(+ : pe) schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int j = i + 1; j < n; ++j) {
            double dx = (double)pos[3 * j] - pos[3 * i], dy = (double)pos[3 * j + 1] - pos[3 * i + 1],
                   dz = (double)pos[3 * j + 2] - pos[3 * i + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            s += law == ORC_LAW_PLUMMER ? -(double)mass[j] / sqrt(r2 + eps * eps)
                                        : (double)mass[j] * sqrt(r2) / (1.0 + eps * eps);
        }
        pe += G * mass[i] * s;
    }
    out[0] = ke; out[1] = pe; out[2] = px; out[3] = py; out[4] = pz;
    out[5] = lx; out[6] = ly; out[7] = lz; out[8] = mt;
}

/* Full Barnes-Hut Euler step sequence (CMyApp::Update, src/MyApp.cpp:583-598):
 * bounds -> reset -> build -> summarize -> sort -> force -> update; the force pass
 * groups bodies in the force order, as the engine does. */
int orc_bh_step(int rules, int nsteps, int n, int cap, float theta, float eps, float G, float dt,
                int vote, int kahan, float* pos, float* vel, const float* mass, float* acc) {
    orc_tree* t = orc_tree_new(n, cap);
    int32_t* sorted = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    float box[6], dq[ORC_DQ_LEN];
    int st = ORC_OK;
    for (int s = 0; s < nsteps && st == ORC_OK; ++s) {
        st = orc_bbox(n, pos, box);
        if (st) break;
        st = orc_build(t, rules, pos, box);
        if (st) break;
        orc_summarize(t, pos, mass);
        orc_force_order(t, sorted);  /* the engine's vote groups */
        orc_dq_table(rules, box, theta, eps, dq);
        orc_force_bh(t, pos, mass, sorted, dq, eps, G, vote, kahan, NULL, 0, acc, NULL, NULL);
        orc_euler(n, dt, acc, vel, pos);
    }
    free(sorted);
    orc_tree_free(t);
    return st;
}
