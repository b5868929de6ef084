/*
 * spmd_io.hip -- the reference's restart and configuration files, host code
 * on top of the public C ABI (include/spmd.h):
 *
 *   spmd_write_fluid    MPC::print_fluid_complete   MPC.cpp:570-579
 *   spmd_read_fluid     MPC::initialize(file, old)  MPC.cpp:57-113
 *   spmd_write_pdb      Box::print_PDB_with_velocity Box.cpp:563-602
 *   spmd_add_star_pdb   Box::add_star(file, ...)    Box.cpp:48-54 +
 *                       Molecule::star_from_file    Molecule.cpp:178-224
 *
 * The formats are the reference's, byte for byte where it prints them with
 * printf.  Device data are FP32 (solvent) and FP64 (monomers); the text
 * files carry %f / %7.3f precision, as the reference's do.
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spmd.h"
#include "spmd_math.h"

namespace {

constexpr uint32_t TAG_RESTART = 7;  /* Philox tag of the particles a restart adds */

struct File {
    FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

/* plain wrap MPC.cpp:473-478 or MPC::LEBC MPC.cpp:421-429 (MPC::wrap) */
void wrap_host(const spmd::Box& b, double* x, double* v) {
    spmd::le_wrap(b, x[0], x[1], x[2], v[0]);
}

spmd::Box host_box(const spmd_params& p, double delrx) {
    spmd::Box b{};
    for (int c = 0; c < 3; ++c) {
        const int L = p.L[c] + (p.L[c] % 2);  /* Box.cpp:14-19 */
        b.L[c] = L;
        b.invL[c] = 1.0 / L;
        b.half[c] = 0.5 * L;
        b.inner[c] = 0.5 * L * (1.0 - 1e-9);
        b.iL[c] = L;
    }
    b.delrx = delrx;
    b.shear = p.shear;
    b.shear_on = p.shear != 0.0;
    return b;
}

}  // namespace

extern "C" {

int spmd_write_fluid(spmd_sys* s, uint32_t r, const char* path) {
    if (!s || !path) return SPMD_ERR_ARG;
    uint64_t n = 0;
    if (int rc = spmd_num_fluid(s, r, &n)) return rc;
    std::vector<float> pos(3 * n), vel(3 * n);
    if (int rc = spmd_get_fluid(s, r, pos.data(), vel.data(), nullptr)) return rc;
    File out(path, "w");
    if (!out.f) return SPMD_ERR_ARG;
    for (uint64_t i = 0; i < n; ++i)
        std::fprintf(out.f, "%f %f %f %f %f %f \n", (double)pos[3 * i], (double)pos[3 * i + 1],
                     (double)pos[3 * i + 2], (double)vel[3 * i], (double)vel[3 * i + 1], (double)vel[3 * i + 2]);
    return std::ferror(out.f) ? SPMD_ERR_ARG : SPMD_OK;
}

int spmd_read_fluid(spmd_sys* s, uint32_t r, const char* path, const int32_t* old_box, uint64_t* n_from_file) {
    if (!s || !path || !old_box) return SPMD_ERR_ARG;
    uint32_t nrep = 0;
    if (int rc = spmd_num_replicas(s, &nrep)) return rc;
    if (r >= nrep) return SPMD_ERR_ARG;
    spmd_params p{};
    if (int rc = spmd_get_params(s, r, &p)) return rc;
    double delrx = 0.0;
    if (int rc = spmd_get_delrx(s, r, &delrx)) return rc;
    const spmd::Box b = host_box(p, delrx);
    uint64_t n = 0;
    if (int rc = spmd_num_fluid(s, r, &n)) return rc;
    File in(path, "r");
    if (!in.f) return SPMD_ERR_ARG;
    std::vector<float> pos(3 * n), vel(3 * n);
    double cmv[3] = {0, 0, 0};
    uint64_t from_file = 0;
    bool more = true;
    for (uint64_t i = 0; i < n; ++i) {
        double x[3], v[3];
        bool have = false;
        if (more) {
            have = std::fscanf(in.f, "%lf %lf %lf %lf %lf %lf", &x[0], &x[1], &x[2], &v[0], &v[1], &v[2]) == 6;
            more = have;
        }
        /* a particle outside the new box, or beyond the end of the file, is
         * placed at random in the region the box gained (anywhere along an
         * axis that did not grow) with U(-1/2, 1/2) velocities (MPC.cpp:73-101;
         * Philox draws keyed by the particle instead of the global engine) */
        bool inside = have;
        for (int c = 0; c < 3 && have; ++c) inside = inside && std::fabs(std::round(x[c] / b.L[c])) < 1.0;
        if (inside) {
            ++from_file;
        } else {
            uint32_t w0[4], w1[4];
            spmd::draw4(p.seed, r, TAG_RESTART, (uint32_t)i, (uint32_t)(i >> 32), 0, w0);
            spmd::draw4(p.seed, r, TAG_RESTART, (uint32_t)i, (uint32_t)(i >> 32), 1, w1);
            for (int c = 0; c < 3; ++c) {
                double u = spmd::u01(w0[c]);
                const int delta = (int)b.L[c] - old_box[c];
                if (delta > 0) u = u < 0.5 ? -b.L[c] * 0.5 + delta * u : old_box[c] * 0.5 + delta * (u - 0.5);
                else u = b.L[c] * (u - 0.5);
                x[c] = u;
                v[c] = spmd::u01(w1[c]) - 0.5;
            }
        }
        wrap_host(b, x, v);
        for (int c = 0; c < 3; ++c) {
            pos[3 * i + c] = (float)x[c];
            vel[3 * i + c] = (float)v[c];
            cmv[c] += v[c];
        }
    }
    for (int c = 0; c < 3; ++c) cmv[c] /= (double)(n ? n : 1);
    for (uint64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) vel[3 * i + c] = (float)((double)vel[3 * i + c] - cmv[c]);
    if (n_from_file) *n_from_file = from_file;
    return spmd_set_fluid(s, r, pos.data(), vel.data(), n);
}

int spmd_write_pdb(spmd_sys* s, uint32_t r, const char* path, int64_t step, int append) {
    if (!s || !path) return SPMD_ERR_ARG;
    uint32_t n = 0, arms = 0, A = 0, B = 0;
    if (int rc = spmd_star_layout(s, r, &n, &arms, &A, &B)) return rc;
    if (n == 0) return SPMD_ERR_STATE;
    std::vector<double> pos(3ull * n), vel(3ull * n);
    if (int rc = spmd_get_monomers(s, r, pos.data(), vel.data(), nullptr)) return rc;
    File out(path, append ? "a" : "w");
    if (!out.f) return SPMD_ERR_ARG;
    const unsigned per = A + B;
    std::fprintf(out.f, "MODEL     %d \n", (int)step);
    for (unsigned i = 0; i < n; i++) {
        /* bead k > A of an arm (index i = a*per + k) is solvophobic */
        const bool bead_b = i > 0 && ((i - 1) % per) + 1 > A;
        std::fprintf(out.f, "ATOM %6d  %s   GLY    %2d     %7.3f %7.3f %7.3f %7.3f %7.3f %7.3f \n", (int)i,
                     bead_b ? "C" : "O", 1, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], vel[3 * i], vel[3 * i + 1],
                     vel[3 * i + 2]);
    }
    std::fprintf(out.f, "TER \n");
    unsigned count = 1, arm = 0;
    while (arm < arms) {
        std::fprintf(out.f, "CONECT %4i ", 1);
        while (count % 5) {
            std::fprintf(out.f, "%4i ", (int)(arm * per + 2));
            arm++;
            count++;
            if (arm >= arms) break;
        }
        count = 1;
        std::fprintf(out.f, "\n");
    }
    for (unsigned i = 0; i < arms; i++) {
        std::fprintf(out.f, "CONECT %4u %4u \n", i * per + 2, 1u);
        for (unsigned j = 1; j < per; j++) {
            const unsigned number = i * per + j;
            std::fprintf(out.f, "CONECT %4u %4u %4u \n", number + 1, number, number + 2);
        }
        std::fprintf(out.f, "CONECT %4u %4u \n", (i + 1) * per + 1, (i + 1) * per);
    }
    std::fprintf(out.f, "ENDMDL \n");
    return std::ferror(out.f) ? SPMD_ERR_ARG : SPMD_OK;
}

int spmd_add_star_pdb(spmd_sys* s, uint32_t r, const char* path, uint32_t A, uint32_t B, uint32_t arms,
                      double mass, int set_zero) {
    if (!s || !path) return SPMD_ERR_ARG;
    /* the topology is the builder's (Molecule.cpp:199-212); the
     * coordinates then come from the file */
    if (int rc = spmd_add_star(s, r, A, B, arms, mass, 1.01, 2.0)) return rc;
    const uint32_t n = (A + B) * arms + 1;
    File in(path, "r");
    if (!in.f) return SPMD_ERR_ARG;
    std::vector<double> pos(3ull * n), vel(3ull * n);
    char line[512];
    if (!std::fgets(line, sizeof line, in.f)) return SPMD_ERR_ARG;  /* the MODEL line */
    for (uint32_t i = 0; i < n; ++i) {
        if (!std::fgets(line, sizeof line, in.f) || std::strlen(line) < 78) return SPMD_ERR_ARG;
        /* fixed columns of the writer's %7.3f fields (Molecule.cpp:190-195) */
        const int col[6] = {31, 39, 47, 55, 63, 71};
        for (int c = 0; c < 6; ++c) {
            char fld[8];
            std::memcpy(fld, line + col[c], 7);
            fld[7] = '\0';
            (c < 3 ? pos : vel)[3 * i + (c % 3)] = std::strtod(fld, nullptr);
        }
    }
    if (set_zero) {  /* Molecule.cpp:213-221: COM and COM velocity to zero */
        double cp[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
        for (uint32_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) { cp[c] += pos[3 * i + c]; cv[c] += vel[3 * i + c]; }
        for (int c = 0; c < 3; ++c) { cp[c] /= n; cv[c] /= n; }
        for (uint32_t i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) { pos[3 * i + c] -= cp[c]; vel[3 * i + c] -= cv[c]; }
    }
    return spmd_set_monomers(s, r, pos.data(), vel.data());
}

}  // extern "C"
