/*
 * spmd_facade.h -- the reference's C++ hot-path classes, re-declared on top
 * of the C ABI of libspmd.so (include/spmd.h).
 *
 * Names, constructor arguments, virtual interfaces and call order follow
 *   Box              star-polymers/lib/Box.h:30-121      (Box.cpp:4-46)
 *   Hydrodynamics    star-polymers/lib/Hydrodynamics/Hydrodynamics.h:7-22
 *   MPC              star-polymers/lib/Hydrodynamics/MPC.h:25-122
 *   Thermostat       star-polymers/lib/Thermostats/Thermostat.h:7-22
 *   Thermostat_None  star-polymers/lib/Thermostats/Thermostat_None.h:9-21
 * so a driver written like star-polymers/src/star-polymers.cpp:351-536
 * compiles against it, base-class pointers included:
 *
 *   Box box(Lx, Ly, Lz, Temperature, Lambda);
 *   MPC hydrodynamics{box, Temperature, N_c, Shear, 100, false};
 *   box.add_star(A, B, Arms, Mass);
 *   Thermostat* thermostat = new Thermostat_None{box, dt};
 *   hydrodynamics.initialize();
 *   for (long n = 0; n <= Steps; ++n) { thermostat->propagate(); hydrodynamics.step(n, 100 * dt); }
 *
 * Differences, each forced by the device owning the state:
 *  - Box::Molecules and MPC::Fluid are host mirrors of the device state,
 *    read from the device on the first access after the device changed it
 *    and written back before the next device operation when the host took
 *    a non-const reference into them (Mirror below).  A write of MPC::Fluid
 *    replaces the solvent (spmd_set_fluid): the next collision sorts anew.
 *  - Errors the reference turns into raise(SIGSEGV) (MPC.cpp:216-219,
 *    Box.cpp:133-142) throw spmd::Error with the library's status code.
 *  - The random stream is Philox4x32-10 keyed by Box::Seed instead of the
 *    global mt19937_64(1) of Rand.cpp:6.
 *  - Vector3d is a small value type with the Eigen members the drivers and
 *    functors use (operator(), transpose(), norm(), printing), not Eigen.
 */
#ifndef SPMD_FACADE_H
#define SPMD_FACADE_H

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <functional>
#include <iomanip>
#include <map>
#include <memory>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include <stdlib.h>
#include <unistd.h>

#include "../../include/spmd.h"

namespace spmd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const spmd_sys* h) {
    if (rc == SPMD_OK) return;
    std::string m = spmd_error_string(rc);
    if (h) m += std::string(": ") + spmd_last_error(h);
    throw Error(rc, m);
}

/* Coefficients printed as Eigen's default IOFormat does (Eigen/src/Core/
 * IO.h): each right-aligned to the widest of them in the stream's format,
 * separated by sep. */
inline void print_coeffs(std::ostream& os, const double* c, int n, const char* sep) {
    std::string t[3];
    std::size_t w = 0;
    for (int i = 0; i < n; ++i) {
        std::ostringstream q;
        q.copyfmt(os);
        q.width(0);
        q << c[i];
        t[i] = q.str();
        w = std::max(w, t[i].size());
    }
    for (int i = 0; i < n; ++i) {
        if (i) os << sep;
        os << std::string(w - t[i].size(), ' ') << t[i];
    }
}

/* The members of Eigen::Vector3d the reference's drivers and functors use. */
struct Vector3d {
    double v[3] = {0.0, 0.0, 0.0};
    Vector3d() = default;
    Vector3d(double a, double b, double c) : v{a, b, c} {}
    static Vector3d Zero() { return Vector3d{}; }
    double& operator[](int i) { return v[i]; }
    const double& operator[](int i) const { return v[i]; }
    double& operator()(int i) { return v[i]; }
    const double& operator()(int i) const { return v[i]; }
    double* data() { return v; }
    const double* data() const { return v; }
    static constexpr int size() { return 3; }
    double* begin() { return v; }
    double* end() { return v + 3; }
    const double* begin() const { return v; }
    const double* end() const { return v + 3; }
    Vector3d& operator+=(const Vector3d& o) { for (int i = 0; i < 3; ++i) v[i] += o.v[i]; return *this; }
    Vector3d& operator-=(const Vector3d& o) { for (int i = 0; i < 3; ++i) v[i] -= o.v[i]; return *this; }
    Vector3d& operator*=(double a) { for (double& x : v) x *= a; return *this; }
    Vector3d& operator/=(double a) { for (double& x : v) x /= a; return *this; }
    friend Vector3d operator+(Vector3d a, const Vector3d& b) { return a += b; }
    friend Vector3d operator-(Vector3d a, const Vector3d& b) { return a -= b; }
    friend Vector3d operator-(Vector3d a) { return a *= -1.0; }
    friend Vector3d operator*(Vector3d a, double s) { return a *= s; }
    friend Vector3d operator*(double s, Vector3d a) { return a *= s; }
    friend Vector3d operator/(Vector3d a, double s) { return a /= s; }
    double dot(const Vector3d& o) const { return (v[0] * o.v[0] + v[1] * o.v[1]) + v[2] * o.v[2]; }
    double squaredNorm() const { return dot(*this); }
    double norm() const { return std::sqrt(squaredNorm()); }
    /* the row view Eigen prints on one line */
    struct Row {
        const Vector3d& x;
    };
    Row transpose() const { return Row{*this}; }
};
inline std::ostream& operator<<(std::ostream& os, const Vector3d::Row& r) {
    print_coeffs(os, r.x.v, 3, " ");
    return os;
}
inline std::ostream& operator<<(std::ostream& os, const Vector3d& x) {
    print_coeffs(os, x.v, 3, "\n");
    return os;
}

/* Eigen's Matrix3d is column-major; data()[i] of the reference's printing
 * (star-polymers.cpp:477-479) is element i of this array */
using Matrix3d = std::array<double, 9>;

/* Particle.h:17-82 */
struct MDParticle {
    Vector3d Position{}, Velocity{}, Force{};
    double Mass{};
};
struct MPCParticle {
    Vector3d Position{}, Velocity{};
};
struct Molecule {
    std::vector<MDParticle> Monomers;
    unsigned NumberOfMonomers{};
    /* Molecule::calculate_center_of_mass(_velocity) Molecule.cpp:76-92:
     * plain means in monomer order */
    Vector3d calculate_center_of_mass() const {
        Vector3d c{};
        for (auto& m : Monomers) c += m.Position;
        return c /= NumberOfMonomers;
    }
    Vector3d calculate_center_of_mass_velocity() const {
        Vector3d c{};
        for (auto& m : Monomers) c += m.Velocity;
        return c /= NumberOfMonomers;
    }
    /* Molecule::print Molecule.cpp:226-231 */
    std::ostream& print(std::ostream& os) const {
        for (auto& m : Monomers)
            os << m.Position.transpose() << " ; " << m.Velocity.transpose() << " ; " << m.Force.transpose() << '\n';
        return os;
    }
};
inline std::ostream& operator<<(std::ostream& os, const Molecule& m) { return m.print(os); }

/* A host mirror of device state: read from the device on the first access
 * after invalidate(); a non-const access marks it written, and flush()
 * (called before every device operation) then sends it back. */
template <class T>
class Mirror {
    std::vector<T> v_;
    std::function<void(std::vector<T>&)> pull_;
    std::function<void(const std::vector<T>&)> push_;
    bool stale_ = true, dirty_ = false;
    void refresh() {
        if (stale_ && pull_) {
            pull_(v_);
            stale_ = false;
        }
    }
    Mirror& self() const { return const_cast<Mirror&>(*this); }

public:
    using value_type = T;
    using iterator = typename std::vector<T>::iterator;
    using const_iterator = typename std::vector<T>::const_iterator;
    void bind(std::function<void(std::vector<T>&)> pull, std::function<void(const std::vector<T>&)> push) {
        pull_ = std::move(pull);
        push_ = std::move(push);
        stale_ = true;
        dirty_ = false;
    }
    /* the device changed the state */
    void invalidate() { stale_ = true; }
    /* host writes to the device (before a device operation) */
    void flush() {
        if (dirty_ && push_) push_(v_);
        dirty_ = false;
    }
    bool written() const { return dirty_; }
    /* read again now */
    void sync() {
        stale_ = true;
        refresh();
    }
    std::size_t size() const { self().refresh(); return v_.size(); }
    bool empty() const { return size() == 0; }
    const T& operator[](std::size_t i) const { self().refresh(); return v_[i]; }
    T& operator[](std::size_t i) { refresh(); dirty_ = true; return v_[i]; }
    const T& at(std::size_t i) const { self().refresh(); return v_.at(i); }
    T& at(std::size_t i) { refresh(); dirty_ = true; return v_.at(i); }
    const_iterator begin() const { self().refresh(); return v_.cbegin(); }
    const_iterator end() const { self().refresh(); return v_.cend(); }
    const_iterator cbegin() const { return begin(); }
    const_iterator cend() const { return end(); }
    iterator begin() { refresh(); dirty_ = true; return v_.begin(); }
    iterator end() { refresh(); return v_.end(); }
    /* the host copy as it stands (no device read) */
    const std::vector<T>& host() const { return v_; }
};

class MPC;

class Box {
    friend class MPC;
    friend class Thermostat;
    friend class Thermostat_None;
    std::shared_ptr<spmd_sys> h_;
    Mirror<MPCParticle>* fluid_ = nullptr;  /* the MPC's solvent mirror */
    spmd_sys* handle() const {
        if (!h_) throw Error(SPMD_ERR_STATE, "construct the MPC on this Box first (star-polymers.cpp:353)");
        return h_.get();
    }
    void bind_molecules() {
        Molecules.bind(
            [this](std::vector<Molecule>& out) {
                const unsigned n = numberOfMonomers();
                out.assign(n ? 1 : 0, Molecule{});
                if (!n) return;
                std::vector<double> p(3 * n), v(3 * n), f(3 * n);
                check(spmd_get_monomers(handle(), 0, p.data(), v.data(), f.data()), h_.get());
                Molecule& m = out[0];
                m.NumberOfMonomers = n;
                m.Monomers.resize(n);
                for (unsigned i = 0; i < n; ++i)
                    for (int c = 0; c < 3; ++c) {
                        m.Monomers[i].Position[c] = p[3 * i + c];
                        m.Monomers[i].Velocity[c] = v[3 * i + c];
                        m.Monomers[i].Force[c] = f[3 * i + c];
                    }
                std::vector<double> mass(n);
                if (spmd_get_monomer_masses(handle(), 0, mass.data()) == SPMD_OK)
                    for (unsigned i = 0; i < n; ++i) m.Monomers[i].Mass = mass[i];
            },
            [this](const std::vector<Molecule>& in) {
                if (in.empty()) return;
                const Molecule& m = in[0];
                std::vector<double> p(3 * m.Monomers.size()), v(3 * m.Monomers.size());
                for (std::size_t i = 0; i < m.Monomers.size(); ++i)
                    for (int c = 0; c < 3; ++c) {
                        p[3 * i + c] = m.Monomers[i].Position[c];
                        v[3 * i + c] = m.Monomers[i].Velocity[c];
                    }
                check(spmd_set_monomers(handle(), 0, p.data(), v.data()), h_.get());
            });
    }

public:
    double Temperature;
    double Lambda;
    std::array<int, 3> BoxSize;
    uint64_t Seed = 1;
    double DeltaT = 0.01;
    Mirror<Molecule> Molecules;

    /* Box::Box(Lx, Ly, Lz, T, Lambda) Box.cpp:4-27 (odd sides are rounded
     * up to even by the library, Box.cpp:14-16) */
    Box(int Lx, int Ly, int Lz, double temperature, double lambda)
        : Temperature(temperature), Lambda(lambda), BoxSize{Lx + (Lx & 1), Ly + (Ly & 1), Lz + (Lz & 1)} {}
    Box(const Box&) = delete;
    Box& operator=(const Box&) = delete;

    /* before a device operation: host writes of the mirrors go down */
    void before_device() {
        Molecules.flush();
        if (fluid_) fluid_->flush();
    }
    /* after a device operation that may change the state */
    void after_device() {
        Molecules.invalidate();
        if (fluid_) fluid_->invalidate();
    }

    /* Box::add_star Box.cpp:39-46 */
    void add_star(unsigned A, unsigned B, unsigned Arms, double Mass, double Bond = 1.01, double AnchorBond = 2.0) {
        before_device();
        check(spmd_add_star(handle(), 0, A, B, Arms, Mass, Bond, AnchorBond), h_.get());
        after_device();
    }
    /* seeded random-coil arms (no reference counterpart) */
    void add_star_coil(unsigned A, unsigned B, unsigned Arms, double Mass, double Bond = 1.01,
                       double AnchorBond = 2.0) {
        before_device();
        check(spmd_add_star_coil(handle(), 0, A, B, Arms, Mass, Bond, AnchorBond), h_.get());
        after_device();
    }
    /* Box::add_star(filename, A, B, Arms, Mass, set_zero) Box.cpp:48-54 */
    void add_star(const std::string& filename, unsigned A, unsigned B, unsigned Arms, double Mass, bool set_zero) {
        before_device();
        check(spmd_add_star_pdb(handle(), 0, filename.c_str(), A, B, Arms, Mass, set_zero ? 1 : 0), h_.get());
        after_device();
    }
    unsigned numberOfMonomers() const {
        uint32_t n = 0;
        check(spmd_num_monomers(handle(), 0, &n), h_.get());
        return n;
    }
    /* Molecules read from the device now / host edits written now */
    void sync_molecules() {
        before_device();
        Molecules.sync();
    }
    void upload_molecules() { Molecules.flush(); }
    /* Molecule::Epot after the last propagate(true) */
    double epot() {
        before_device();
        double e = 0.0;
        check(spmd_epot(handle(), 0, &e), h_.get());
        return e;
    }

    /* ---- analysis of Box.cpp:296-507, 605-660 on the device ---- */
    Vector3d COM_Pos{}, COM_Vel{};
    /* Box::center_of_mass_reference_frame Box.cpp:81-102: COM_Pos / COM_Vel
     * (the device observers compute the frame themselves) */
    void center_of_mass_reference_frame() {
        COM_Pos = COM_Vel = Vector3d{};
        const Mirror<Molecule>& mols = Molecules;
        if (mols.empty()) return;
        COM_Pos = mols[0].calculate_center_of_mass();
        COM_Vel = mols[0].calculate_center_of_mass_velocity();
    }
    /* Box::print_center_of_mass Box.cpp:654-660 */
    std::ostream& print_center_of_mass(std::ostream& os) const {
        for (const Molecule& mol : Molecules)
            os << mol.calculate_center_of_mass().transpose() << " " << mol.calculate_center_of_mass_velocity().transpose()
               << '\n';
        return os;
    }
    /* Box::print_molecules Box.cpp:517-522 */
    std::ostream& print_molecules(std::ostream& os) const {
        for (const Molecule& mol : Molecules) os << mol << std::endl;
        return os;
    }
    /* the 18 values of a statistics row (spmd_statistics) */
    std::array<double, 18> statistics() {
        before_device();
        std::array<double, 18> row{};
        check(spmd_statistics(handle(), 0, row.data()), h_.get());
        return row;
    }
    /* Box::calculate_patches_new Box.cpp:434-507 */
    std::vector<double> calculate_patches_new() {
        before_device();
        std::vector<double> r(2, 0.0);
        check(spmd_patches(handle(), 0, &r[0], &r[1]), h_.get());
        return r;
    }
    /* Box::calculate_gyration_tensor Box.cpp:330-352: (R_gyr, G) */
    std::tuple<double, Matrix3d> calculate_gyration_tensor() {
        before_device();
        Matrix3d g{};
        check(spmd_gyration_tensor(handle(), 0, g.data()), h_.get());
        return std::make_tuple(std::sqrt((g[0] + g[4]) + g[8]), g);
    }
    /* Box::calculate_rotation_frequency Box.cpp:354-379 */
    Vector3d calculate_rotation_frequency() {
        before_device();
        Vector3d w{};
        check(spmd_rotation_frequency(handle(), 0, w.data()), h_.get());
        return w;
    }
    /* Box::print_Epot Box.cpp:605-613 */
    std::ostream& print_Epot(std::ostream& os) { return os << statistics()[0] << " "; }
    /* Box::print_Ekin_and_Temperature Box.cpp:635-646 */
    std::ostream& print_Ekin_and_Temperature(std::ostream& os) {
        const auto r = statistics();
        return os << r[1] << " " << r[2] << " ";
    }
    /* Box::print_PDB_with_velocity Box.cpp:563-602, to a file (appended) */
    void print_PDB_with_velocity(const std::string& path, int step, bool append = true) {
        before_device();
        check(spmd_write_pdb(handle(), 0, path.c_str(), step, append ? 1 : 0), h_.get());
    }
    /* the same to an open stream, as the reference's signature takes it:
     * the library writes the model to a temporary file, copied here */
    void print_PDB_with_velocity(FILE* pdb, int step) {
        char tmpl[] = "/tmp/spmd_pdb_XXXXXX";
        const int fd = mkstemp(tmpl);
        if (fd < 0) throw Error(SPMD_ERR_ARG, "cannot create a temporary file for the PDB model");
        close(fd);
        struct Unlink {
            const char* p;
            ~Unlink() { std::remove(p); }
        } unlink_tmp{tmpl};
        print_PDB_with_velocity(std::string(tmpl), step, false);
        FILE* in = std::fopen(tmpl, "rb");
        if (!in) throw Error(SPMD_ERR_ARG, "cannot read back the PDB model");
        char buf[1 << 14];
        std::size_t n;
        while ((n = std::fread(buf, 1, sizeof buf, in)) > 0) std::fwrite(buf, 1, n, pdb);
        std::fclose(in);
    }
};

/* Hydrodynamics.h:7-22: the solvent interface the driver holds */
class Hydrodynamics {
protected:
    Box& SimBox;
    ~Hydrodynamics() = default;

public:
    explicit Hydrodynamics(Box& box) : SimBox(box) {}
    virtual void initialize() {}
    virtual void step(const long int& /*t*/, const double& /*dt*/) {}
    virtual double calculateCurrentTemperature() { return 0.0; }
    virtual void print_fluid(FILE*, int, int, int) {}
    virtual void print_fluid_with_coordinates(FILE*, int, int, int) {}
};

/* MPC.h:25-122 */
class MPC : public Hydrodynamics {
    spmd_sys* h() const { return SimBox.handle(); }
    /* observers with a device pass (sample(spmd_sys*)) run it; any other
     * functor is applied to every particle of the solvent mirror */
    template <class F>
    auto apply(F& f, int) -> decltype(f.sample(static_cast<spmd_sys*>(nullptr)), void()) {
        SimBox.before_device();
        f.sample(h());
    }
    template <class F>
    void apply(F& f, long) {
        const Mirror<MPCParticle>& fl = Fluid;
        for (auto it = fl.cbegin(), last = fl.cend(); it != last; ++it) f(*it);
    }

public:
    Mirror<MPCParticle> Fluid;

    /* MPC::MPC(Box&, T, N_c, aShear, step, angular_mom) MPC.h:44 */
    MPC(Box& box, double T, int N_c, double aShear = 0., unsigned step = 1, bool angular_mom = false)
        : Hydrodynamics(box) {
        spmd_params p{};
        for (int c = 0; c < 3; ++c) p.L[c] = box.BoxSize[c];
        p.kT = T;
        p.n_c = N_c;
        p.shear = aShear;
        p.step_update = step;
        p.alpha_deg = 130.0;
        p.lambda = box.Lambda;
        p.md_dt = box.DeltaT;
        p.seed = box.Seed;
        p.angular_momentum = angular_mom ? 1 : 0;
        spmd_sys* hh = nullptr;
        check(spmd_create(&p, &hh), nullptr);
        box.h_ = std::shared_ptr<spmd_sys>(hh, [](spmd_sys* q) { spmd_destroy(q); });
        box.bind_molecules();
        box.fluid_ = &Fluid;
        Fluid.bind(
            [this](std::vector<MPCParticle>& out) {
                const unsigned n = NumberOfParticles();
                std::vector<float> pp(3ull * n), vv(3ull * n);
                check(spmd_get_fluid(h(), 0, pp.data(), vv.data(), nullptr), h());
                out.resize(n);
                for (unsigned i = 0; i < n; ++i)
                    for (int c = 0; c < 3; ++c) {
                        out[i].Position[c] = pp[3ull * i + c];
                        out[i].Velocity[c] = vv[3ull * i + c];
                    }
            },
            [this](const std::vector<MPCParticle>& in) {
                std::vector<float> pp(3 * in.size()), vv(3 * in.size());
                for (std::size_t i = 0; i < in.size(); ++i)
                    for (int c = 0; c < 3; ++c) {
                        pp[3 * i + c] = (float)in[i].Position[c];
                        vv[3 * i + c] = (float)in[i].Velocity[c];
                    }
                check(spmd_set_fluid(h(), 0, pp.data(), vv.data(), in.size()), h());
            });
    }
    MPC(const MPC&) = delete;
    MPC& operator=(const MPC&) = delete;
    ~MPC() {
        if (SimBox.fluid_ == &Fluid) SimBox.fluid_ = nullptr;
    }
    /* MPC::initialize() MPC.cpp:34-55 */
    void initialize() override {
        SimBox.before_device();
        check(spmd_fluid_initialize(h()), h());
        SimBox.after_device();
    }
    /* MPC::initialize(filename, Box_old) MPC.cpp:57-113 */
    void initialize(const std::string& filename, std::array<int, 3> Box_old) {
        SimBox.before_device();
        const int32_t ob[3] = {Box_old[0], Box_old[1], Box_old[2]};
        check(spmd_read_fluid(h(), 0, filename.c_str(), ob, nullptr), h());
        SimBox.after_device();
    }
    /* MPC::step(t, dt) MPC.cpp:116-142 */
    void step(const long int& t, const double& dt) override {
        SimBox.before_device();
        check(spmd_mpc_step(h(), t, dt), h());
        SimBox.after_device();
    }
    /* MPC::calculateCurrentTemperature MPC.cpp:405-409.  Before the first
     * collision after the fluid was set (a continued run, MPC.cpp:57-113,
     * has not sorted yet) the reference's cell lists are empty and the
     * result is 0; the library reports that state as SPMD_ERR_STATE. */
    double calculateCurrentTemperature() override {
        SimBox.before_device();
        double T = 0.0;
        const int rc = spmd_temperature(h(), 0, &T);
        if (rc == SPMD_ERR_STATE) return 0.0;
        check(rc, h());
        return T;
    }
    /* MPC::print_fluid MPC.cpp:533-553: per z plane the cell centres, then
     * the mean solvent velocity (x, y components) of every cell of the last
     * collision, one row per x */
    void print_fluid(FILE* fluid_file, int step, int z_start, int z_stop) override {
        const int Lx = SimBox.BoxSize[0], Ly = SimBox.BoxSize[1];
        std::vector<double> cv = cell_velocities(z_start, z_stop);
        std::fprintf(fluid_file, "TIME %d \n", step);
        for (int z = z_start; z <= z_stop; z++) {
            std::fprintf(fluid_file, "Z %d \n", z);
            for (int x = 0; x < Lx; x++) std::fprintf(fluid_file, "%f ", x + 0.5);
            std::fprintf(fluid_file, "\n");
            for (int y = 0; y < Ly; y++) std::fprintf(fluid_file, "%f ", y + 0.5);
            std::fprintf(fluid_file, "\n");
            for (int x = 0; x < Lx; x++) {
                for (int y = 0; y < Ly; y++) {
                    const double* c = &cv[3 * ((std::size_t)x + (std::size_t)Lx * (y + (std::size_t)Ly * (z - z_start)))];
                    std::fprintf(fluid_file, "%f %f ", c[0], c[1]);
                }
                std::fprintf(fluid_file, "\n");
            }
            std::fprintf(fluid_file, "\n");
        }
        std::fprintf(fluid_file, "\n");
    }
    /* MPC::print_fluid_with_coordinates MPC.cpp:555-568 (z_stop exclusive) */
    void print_fluid_with_coordinates(FILE* fluid_file, int step, int z_start, int z_stop) override {
        const int Lx = SimBox.BoxSize[0], Ly = SimBox.BoxSize[1];
        std::fprintf(fluid_file, "TIME     %d \n", step);
        if (z_stop > z_start) {
            std::vector<double> cv = cell_velocities(z_start, z_stop - 1);
            for (int z = z_start; z < z_stop; z++)
                for (int x = 0; x < Lx; x++)
                    for (int y = 0; y < Ly; y++) {
                        const double* c =
                            &cv[3 * ((std::size_t)x + (std::size_t)Lx * (y + (std::size_t)Ly * (z - z_start)))];
                        std::fprintf(fluid_file, "%f %f %f %f %f %f \n", x + 0.5, y + 0.5, z + 0.5, c[0], c[1], c[2]);
                    }
        }
        std::fprintf(fluid_file, "\n");
    }
    /* MPC::calculateFluidVelocity (MPC.cpp:334-342) of the cells with
     * z_start <= z <= z_stop, 3 per cell in cell-index order */
    std::vector<double> cell_velocities(int z_start, int z_stop) {
        SimBox.before_device();
        std::vector<double> cv(3ull * SimBox.BoxSize[0] * SimBox.BoxSize[1] * (z_stop - z_start + 1));
        check(spmd_cell_fluid_velocity(h(), 0, z_start, z_stop, cv.data()), h());
        return cv;
    }
    /* MPC::print_fluid_complete MPC.cpp:570-579 */
    void print_fluid_complete(FILE* fluid_file) {
        for (const MPCParticle& part : static_cast<const Mirror<MPCParticle>&>(Fluid))
            std::fprintf(fluid_file, "%f %f %f %f %f %f \n", part.Position(0), part.Position(1), part.Position(2),
                         part.Velocity(0), part.Velocity(1), part.Velocity(2));
    }
    /* the same, written by the library to a file */
    void print_fluid_complete(const std::string& path) {
        SimBox.before_device();
        check(spmd_write_fluid(h(), 0, path.c_str()), h());
    }
    /* MPC::operator()(UnaryFunc&) MPC.h:110-116: the functor on every
     * solvent particle (device observers: their own pass) */
    template <class UnaryFunc>
    void operator()(UnaryFunc& ufunc) {
        apply(ufunc, 0);
    }
    /* MPC::operator()(UnaryFunc&, BinaryFunc&) MPC.h:98-108 */
    template <class UnaryFunc, class BinaryFunc>
    void operator()(UnaryFunc& ufunc, BinaryFunc& bfunc) {
        const Mirror<MPCParticle>& fl = Fluid;
        for (auto first = fl.cbegin(), last = fl.cend(); first != last; ++first) {
            ufunc(*first);
            for (auto second = first + 1; second != last; ++second) bfunc(*first, *second);
        }
    }
    /* MPC::unitary MPC.h:93-96 */
    template <class UnitaryFunc>
    UnitaryFunc unitary(UnitaryFunc& func) {
        const Mirror<MPCParticle>& fl = Fluid;
        return std::for_each(fl.cbegin(), fl.cend(), func);
    }
    /* MPC::NumberOfParticles MPC.h:91 */
    unsigned NumberOfParticles() const {
        uint64_t n = 0;
        check(spmd_num_fluid(h(), 0, &n), h());
        return (unsigned)n;
    }
    /* Fluid read from the device now / host edits written now */
    void sync_fluid() {
        SimBox.before_device();
        Fluid.sync();
    }
    void upload_fluid() { Fluid.flush(); }
    /* the star-polymers.cpp:412-536 loop body n times from t0, replayed from a HIP graph */
    void run(long t0, long n, double dt) {
        SimBox.before_device();
        check(spmd_run(h(), t0, n, dt), h());
        SimBox.after_device();
    }
    void sync() { check(spmd_sync(h()), h()); }
};

/* VelocityX (Velocity_x.h:22-148): mean solvent velocity relative to the
 * star's centre of mass, binned by x and by y; applied as
 * hydrodynamics(velocity_average_x) like the reference's functor. */
class VelocityX {
    double width;
    std::map<double, std::array<double, 3>> maps[2];  /* key -> sum vx, sum vy, count */

public:
    VelocityX(Box&, MPC&, double a_width = 0.5) : width(a_width) {}
    void sample(spmd_sys* h) {
        spmd_params p{};
        check(spmd_get_params(h, 0, &p), h);
        uint32_t nmax = 0;
        for (int a = 0; a < 2; ++a) {
            const double half = 0.5 * (p.L[a] + (p.L[a] & 1));
            nmax = std::max<uint32_t>(nmax, (uint32_t)(std::floor(half / width) - std::floor(-half / width) + 1));
        }
        int32_t b0[2];
        uint32_t nb[2];
        std::vector<double> sums(2ull * nmax * 3);
        check(spmd_velocity_profile(h, 0, width, nmax, b0, nb, sums.data()), h);
        for (int a = 0; a < 2; ++a)
            for (uint32_t b = 0; b < nb[a]; ++b) {
                const double* q = &sums[3ull * (a * nmax + b)];
                if (q[2] == 0) continue;
                auto& e = maps[a][(b0[a] + (int)b) * width];
                for (int c = 0; c < 3; ++c) e[c] += q[c];
            }
    }
    /* VelocityX::print_result Velocity_x.h:98-144 */
    std::ostream& print_result(std::ostream& os) const {
        char buf[128];
        for (int a = 0; a < 2; ++a) {
            for (auto& kv : maps[a]) {
                std::snprintf(buf, sizeof buf, "%.8e %.8e %.8e %.8e\n", kv.first, kv.second[0] / kv.second[2],
                              kv.second[1] / kv.second[2], kv.second[2]);
                os << buf;
            }
            if (a == 0) os << '\n';
        }
        return os;
    }
};

/* VelocityDistribution (Vel_distribution.h:17-57): speeds binned by
 * floor(|v| / width) (the device histogram holds 1024 bins and counts the
 * speeds beyond them, which print_result leaves out but normalises by). */
class VelocityDistribution {
    static constexpr uint32_t kBins = 1024;
    double width;
    std::vector<uint64_t> counts;

public:
    explicit VelocityDistribution(double a_width = 0.5) : width(a_width), counts(kBins + 1, 0) {}
    void sample(spmd_sys* h) {
        std::vector<uint64_t> c(kBins + 1);
        check(spmd_speed_histogram(h, 0, width, kBins, c.data()), h);
        for (uint32_t b = 0; b <= kBins; ++b) counts[b] += c[b];
    }
    /* VelocityDistribution::print_result Vel_distribution.h:43-56 */
    std::ostream& print_result(std::ostream& os) {
        uint64_t total = 0;
        for (uint64_t c : counts) total += c;
        os.precision(8);
        os << std::scientific;
        for (uint32_t b = 0; b < kBins; ++b)
            if (counts[b]) os << b * width << ' ' << (double)counts[b] / ((double)total * width) << '\n';
        os << '\n';
        return os;
    }
};

/* Fourier_Autocorr (Fourier_Autocorr.h:19-127): the transverse velocity
 * modes V_j at k = 2 pi / L e_j (one device pass per sample,
 * spmd_fourier_modes) in a window of Entries samples; from the window's
 * filling on, every sample adds Re(V_j(oldest)^H V_j(oldest + i)) to lag i.
 * Sums in double (the reference's are long double). */
class Fourier_Autocorr {
    using Modes = std::array<std::complex<double>, 9>;
    unsigned Entries;
    double DeltaT;
    unsigned EntryCounter = 0, Counter = 0;
    std::deque<Modes> Velocities;
    std::array<std::vector<double>, 3> Autocorrelation;

public:
    Fourier_Autocorr(unsigned /*N_p*/, unsigned N, double dt, double /*L*/)
        : Entries(N), DeltaT(dt), Velocities(N, Modes{}) {
        for (auto& a : Autocorrelation) a.assign(N, 0.0);
    }
    void sample(spmd_sys* h) {
        double m[18];
        check(spmd_fourier_modes(h, 0, m), h);
        Modes v;
        for (int q = 0; q < 9; ++q) v[q] = std::complex<double>(m[2 * q], m[2 * q + 1]);
        Velocities.pop_front();
        Velocities.push_back(v);
        if (EntryCounter < Entries) EntryCounter++;
        if (EntryCounter == Entries) {  /* calculate(), Fourier_Autocorr.h:80-93 */
            const Modes& first = Velocities.front();
            for (unsigned i = 0; i < 3; ++i)
                for (unsigned j = 0; j < Entries; ++j) {
                    double d = 0.0;
                    for (int c = 0; c < 3; ++c) d += (std::conj(first[3 * i + c]) * Velocities[j][3 * i + c]).real();
                    Autocorrelation[i][j] += d;
                }
            Counter++;
        }
    }
    /* Fourier_Autocorr::print_result Fourier_Autocorr.h:113-119 */
    std::ostream& print_result(std::ostream& os) const {
        for (unsigned i = 0; i < Entries; ++i)
            os << i * DeltaT << " " << Autocorrelation[0][i] / Counter << " " << Autocorrelation[1][i] / Counter << " "
               << Autocorrelation[2][i] / Counter << "\n";
        return os;
    }
};

/* Vel_Autocorr_Fluid (Vel_Autocorr_Fluid.h:19-116): every solvent
 * particle's window of Entries velocities and the lag sums live on the
 * device (spmd_fluid_vacf_*); particles are followed by id, so the first
 * sample turns tracking on. */
class Vel_Autocorr_Fluid {
    unsigned Entries, NumberOfParticles;
    double DeltaT;
    spmd_sys* h_ = nullptr;

public:
    Vel_Autocorr_Fluid(unsigned N_p, unsigned N, double dt) : Entries(N), NumberOfParticles(N_p), DeltaT(dt) {}
    void sample(spmd_sys* h) {
        if (!h_) check(spmd_set_tracking(h, 1), h);
        h_ = h;
        check(spmd_fluid_vacf_sample(h, 0, Entries), h);
    }
    /* calculate_average (:80-88), value (:90-97), print_result (:99-106) */
    std::ostream& print_result(std::ostream& os) const {
        std::vector<double> sums(Entries, 0.0), avg(Entries, 0.0);
        uint64_t counter = 0;
        if (h_) check(spmd_fluid_vacf_read(h_, 0, sums.data(), &counter), h_);
        for (unsigned i = 0; i < Entries; ++i) avg[i] = sums[i] / (double)counter / (double)NumberOfParticles;
        double diffusion = 0.0;
        for (double a : avg) diffusion += a * DeltaT / (double)counter;
        diffusion /= 3.;
        os << "#Diffusion coefficient = " << diffusion << "\n";
        for (unsigned i = 0; i < Entries; ++i) os << i * DeltaT << " " << avg[i] << "\n";
        return os;
    }
};

/* statistics*.dat header of star-polymers.cpp:307 */
inline const char* statistics_header() {
    return "Step   Epot     Ekin     Temp   n_p s_p  R_gyr   G_xx      G_xy      G_xz       G_yx       G_yy      "
           "G_yz      G_zx     Gzy     Gzz      wx      wy     wz\n";
}

/* ---- run parameters encoded in file names ----
 * The production driver names every output file by a parameter string
 * (star-polymers.cpp:270-288), e.g.
 *   _Star_A20_B20_Arms30_Mass10_Lx64_Ly64_Lz64_Lambda1.00_T1.00_run10000_MPCON_Shear0.05000
 * (integers as such, Lambda and T with 2 decimals, the run's total step
 * count, the shear rate with 5 decimals or _MPCOFF), writes the restart
 * files ./results/end_config<tag>.pdb and end_fluid<tag>.dat, and reads
 * the parameters of a continued run back from the end_config name
 * (:118-133, :199-216) with find_parameter (:74-83). */
struct RunParams {
    bool chain = false;
    int A = 0, B = 0, Arms = 0, Mass = 0, Lx = 0, Ly = 0, Lz = 0;
    double Lambda = 0.0, T = 0.0, Shear = 0.0;
    long Steps = 0;  /* "run": the step count the name carries */
    bool MPC_on = false;
};

inline std::string run_tag(const RunParams& p) {
    std::ostringstream s;
    s << (p.chain ? "_Chain" : "_Star") << "_A" << p.A << "_B" << p.B;
    if (!p.chain) s << "_Arms" << p.Arms;
    s << "_Mass" << p.Mass << "_Lx" << p.Lx << "_Ly" << p.Ly << "_Lz" << p.Lz;
    s << std::fixed << std::setprecision(2) << "_Lambda" << p.Lambda << "_T" << p.T;
    s << "_run" << p.Steps;
    if (p.MPC_on) s << std::setprecision(5) << "_MPCON_Shear" << p.Shear;
    else s << "_MPCOFF";
    return s.str();
}

/* the text after the first occurrence of name up to the next '_', "0.0"
 * when name is absent (star-polymers.cpp:74-83) */
inline std::string find_parameter(const std::string& tag, const std::string& name) {
    const std::string::size_type i = tag.find(name);
    if (i == std::string::npos) return "0.0";
    std::string v = tag.substr(i + name.size());
    const std::string::size_type j = v.find('_');
    if (j != std::string::npos) v.erase(j);
    return v;
}

/* the parameters of a restart file name ".../end_config<tag>.pdb"
 * (star-polymers.cpp:118-133, 199-216) */
inline RunParams parse_restart_name(const std::string& path) {
    std::string s = path;
    const std::string::size_type i = s.find("end_config");
    if (i == std::string::npos) throw Error(SPMD_ERR_ARG, "not an end_config file name: " + path);
    s.erase(0, i + std::string("end_config").size());
    const std::string::size_type k = s.find(".pdb");
    if (k != std::string::npos) s.erase(k, 4);
    RunParams p;
    try {
        p.chain = s.compare(0, 6, "_Chain") == 0;
        p.Lx = std::stoi(find_parameter(s, "Lx"));
        p.Ly = std::stoi(find_parameter(s, "Ly"));
        p.Lz = std::stoi(find_parameter(s, "Lz"));
        p.A = std::stoi(find_parameter(s, "A"));
        p.B = std::stoi(find_parameter(s, "B"));
        p.Arms = std::stoi(find_parameter(s, "Arms"));
        p.Mass = std::stoi(find_parameter(s, "Mass"));
        p.Steps = std::stol(find_parameter(s, "run"));
        p.T = std::stod(find_parameter(s, "T"));
        p.Shear = std::stod(find_parameter(s, "Shear"));
        p.Lambda = std::stod(find_parameter(s, "Lambda"));
        p.MPC_on = find_parameter(s, "MPC") == "ON";
    } catch (const std::logic_error&) {
        throw Error(SPMD_ERR_ARG, "unreadable parameters in file name: " + path);
    }
    return p;
}

/* Thermostat.h:7-22: the MD integrator interface the driver holds
 * (star-polymers.cpp:101 Thermostat* thermostat) */
class Thermostat {
protected:
    Box& SimBox;
    double DeltaT;
    ~Thermostat() = default;

public:
    Thermostat(Box& box, double dt) : SimBox(box), DeltaT(dt) {}
    double dtime() const { return DeltaT; }
    virtual void dtime(double& new_dtime) { DeltaT = new_dtime; }
    virtual void update_temp() = 0;
    virtual void propagate(bool calc_epot = false) = 0;
    virtual void propagate_gaussian(bool calc_epot = false) = 0;
    virtual std::string name() const = 0;
    virtual std::string info() const = 0;
};

/* Thermostat_None (Thermostat_None.h:9-21, Thermostat_None.cpp:5-38):
 * velocity Verlet of the star on the device; the solvent is MPC's */
class Thermostat_None final : public Thermostat {
public:
    Thermostat_None(Box& box, double dt, bool gaussian = false) : Thermostat(box, dt) {
        if (gaussian) throw Error(SPMD_ERR_UNSUPPORTED, "Gaussian chains are not on this path");
        update_temp();
        Thermostat_None::dtime(dt);
        box.before_device();
        check(spmd_md_init(box.handle()), box.handle());
        box.after_device();
    }
    using Thermostat::dtime;
    void update_temp() override {}
    void dtime(double& delta_time) override {
        Thermostat::dtime(delta_time);
        check(spmd_set_md_dt(SimBox.handle(), delta_time), SimBox.handle());
    }
    void propagate(bool calc_epot = false) override {
        SimBox.before_device();
        check(spmd_md_propagate(SimBox.handle(), calc_epot ? 1 : 0), SimBox.handle());
        SimBox.after_device();
    }
    void propagate_gaussian(bool /*calc_epot*/ = false) override {
        throw Error(SPMD_ERR_UNSUPPORTED, "Gaussian chains are not on this path");
    }
    std::string name() const override { return "None"; }
    std::string info() const override { return "velocity Verlet without thermostat (solvent: MPC)"; }
};

}  // namespace spmd

#endif
