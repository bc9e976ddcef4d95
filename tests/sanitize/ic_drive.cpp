// Drives the host-side initial-condition generators (csrc/nb_ic.cpp) over every kind
// and reference mode, including 0 and 1 bodies and more massive objects than bodies,
// so that a build with -fsanitize=address,undefined checks them.  Exit status 0 =
// every call returned NB_OK and every value is finite.
#include <cmath>
#include <cstdio>
#include <vector>

#include "nbody/nbody_ic.h"

static bool finite_all(const std::vector<float>& v) {
    for (float x : v)
        if (!std::isfinite(x)) return false;
    return true;
}

int main() {
    int bad = 0;
    for (int n : {0, 1, 2, 7, 1000}) {
        std::vector<float> pos(3 * (size_t)n + 3), vel(3 * (size_t)n + 3), mass((size_t)n + 1);
        for (int kind = NB_IC_NORMAL; kind <= NB_IC_COLLISION; ++kind) {
            nb_ic_params p;
            nb_ic_params_default(&p, kind);
            if (nb_ic_generate(&p, n, 42u + kind, pos.data(), vel.data(), mass.data()) != NB_OK) bad = 1;
            if (!finite_all(pos) || !finite_all(vel) || !finite_all(mass)) bad = 1;
        }
        for (int pc = 0; pc < 2; ++pc)
            for (int vc = 0; vc < 5; ++vc)
                for (int massive : {0, 3, n + 2}) {
                    nb_ic_reference_params r;
                    nb_ic_reference_default(&r);
                    r.pos_config = pc == 0 ? NB_SPHERE_POS : NB_UNIFORM_POS;
                    r.vel_config = vc;  // NB_RANDOM_VEL .. NB_TANGENT_XZ_VEL
                    r.number_of_massive_objects = massive;
                    nb_status st = nb_ic_reference(&r, n, 7u, pos.data(), vel.data(), mass.data());
                    if (st != NB_OK && st != NB_E_INVALID_ARG) bad = 1;
                    if (st == NB_OK && (!finite_all(pos) || !finite_all(vel) || !finite_all(mass))) bad = 1;
                }
    }
    std::printf(bad ? "FAILED\n" : "ok\n");
    return bad;
}
