/* Drives every function of the oracle on small seeded inputs, including the
 * failure cases (coincident bodies, a pool too small, zero bodies), so that a build
 * with -fsanitize=address,undefined checks the restatement's memory accesses and
 * arithmetic.  Exit status 0 = every call returned what the case expects. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nbody_oracle.h"

static uint64_t rng = 0x9e3779b97f4a7c15ull;
static float unif(void) {  /* splitmix64 -> [0,1) */
    uint64_t z = (rng += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return (float)((z ^ (z >> 31)) >> 40) / (float)(1ull << 24);
}

static int run_tree(int rules, int n, int cap, const float* pos, const float* mass, int expect) {
    float box[6], dq[ORC_DQ_LEN];
    if (orc_bbox(n, pos, box) != ORC_OK) return n > 0;
    orc_tree* t = orc_tree_new(n, cap);
    int st = orc_build(t, rules, pos, box);
    if (st != expect) { fprintf(stderr, "build: got %d want %d\n", st, expect); orc_tree_free(t); return 1; }
    if (st == ORC_OK && n > 0) {
        orc_summarize(t, pos, mass);
        int32_t* sorted = malloc(sizeof(int32_t) * n);
        orc_sort(t, sorted);
        const int nc = orc_tree_ncells(t);
        uint8_t* paths = malloc((size_t)n * ORC_PATH_STRIDE);
        int32_t* plen = malloc(sizeof(int32_t) * n);
        int32_t* lv = malloc(sizeof(int32_t) * nc), *sl = malloc(sizeof(int32_t) * nc), *ct = malloc(sizeof(int32_t) * nc);
        float* com = malloc(sizeof(float) * 4 * nc);
        orc_export(t, sorted, paths, plen, lv, sl, ct, com);
        orc_dq_table(rules, box, 0.5f, 0.01f, dq);
        float* acc = malloc(sizeof(float) * 3 * n);
        double* absacc = malloc(sizeof(double) * n);
        int64_t* ni = malloc(sizeof(int64_t) * n);
        int32_t groups[2] = {0, (n - 1) / 64};
        orc_force_bh(t, pos, mass, sorted, dq, 0.01f, 1.0f, 64, 0, NULL, 0, acc, absacc, ni);
        orc_force_bh(t, pos, mass, sorted, dq, 0.01f, 1.0f, 1, 1, NULL, 0, acc, absacc, ni);
        orc_force_bh(t, pos, mass, sorted, dq, 0.01f, 1.0f, 64, 0, groups, 2, acc, NULL, NULL);
        uint8_t digits[ORC_MAX_LEVEL + 2];
        orc_path_of(rules, pos, box, ORC_MAX_LEVEL + 1, digits);
        free(sorted); free(paths); free(plen); free(lv); free(sl); free(ct); free(com);
        free(acc); free(absacc); free(ni);
    }
    orc_tree_free(t);
    return 0;
}

int main(void) {
    int bad = 0;
    for (int rules = 0; rules < 2; ++rules) {
        for (int n = 0; n <= 700; n = n ? n * 3 : 1) {
            float* pos = malloc(sizeof(float) * 3 * (n + 1));
            float* mass = malloc(sizeof(float) * (n + 1));
            for (int i = 0; i < 3 * n; ++i) pos[i] = 2.0f * unif() - 1.0f;
            for (int i = 0; i < n; ++i) mass[i] = 0.5f + unif();
            bad |= run_tree(rules, n, 2 * n + 16, pos, mass, ORC_OK);
            if (n >= 9) {  /* nine bodies in eight root slots: at least one new cell */
                bad |= run_tree(rules, n, 1, pos, mass, ORC_E_TREE_OVERFLOW);
                memcpy(pos + 3, pos, sizeof(float) * 3);  /* two coincident bodies: a chain past level 64 */
                bad |= run_tree(rules, n, 2 * n + 16 + 2 * ORC_MAX_LEVEL, pos, mass, ORC_E_DEPTH_OVERFLOW);
            }
            if (n > 0) {
                float* acc = malloc(sizeof(float) * 3 * n);
                float* vel = calloc(3 * (size_t)n, sizeof(float));
                double* absacc = malloc(sizeof(double) * n);
                double* a64 = malloc(sizeof(double) * 3 * n);
                double e9[9];
                int32_t tg[1] = {n - 1};
                for (int law = 0; law < 2; ++law) {
                    orc_force_bf(law, 0.01f, 1.0f, 0, n, pos, mass, NULL, 0, acc, absacc);
                    orc_force_bf(law, 0.01f, 1.0f, 1, n, pos, mass, tg, 1, acc, absacc);
                    orc_direct_f64(law, 0.01, 1.0, n, pos, mass, NULL, 0, a64);
                    orc_energy_momentum_f64(law, 0.01, 1.0, n, pos, vel, mass, e9);
                }
                orc_euler(n, 1e-3f, acc, vel, pos);
                orc_kick(n, 5e-4f, acc, vel);
                orc_drift(n, 1e-3f, vel, pos);
                orc_collision(n, pos, vel);
                free(acc); free(vel); free(absacc); free(a64);
            }
            free(pos);
            free(mass);
        }
        /* whole steps */
        const int n = 300;
        float* pos = malloc(sizeof(float) * 3 * n), *vel = calloc(3 * n, sizeof(float));
        float* mass = malloc(sizeof(float) * n), *acc = malloc(sizeof(float) * 3 * n);
        for (int i = 0; i < 3 * n; ++i) pos[i] = unif();
        for (int i = 0; i < n; ++i) mass[i] = 1.0f / n;
        if (orc_bh_step(rules, 3, n, 2 * n + 16, 0.5f, 0.01f, 1.0f, 1e-3f, 64, 0, pos, vel, mass, acc) != ORC_OK) bad = 1;
        free(pos); free(vel); free(mass); free(acc);
    }
    printf(bad ? "FAILED\n" : "ok\n");
    return bad;
}
