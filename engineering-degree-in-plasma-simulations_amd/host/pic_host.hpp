// pic_host.hpp -- the reference's C++ class interfaces (ch4/v3: World,
// Species, PotentialSolver, Interaction / MC_MEX_Ionization) re-implemented
// as host mirrors over the C ABI of libpic_mi355.so (include/pic_abi.h).
//
// Particles and fields live in device memory.  The public field members the
// reference exposes (world.phi, world.rho, world.ef, sp.den, ...) are host
// mirrors that pull from the device lazily on first access after device work,
// and push back when they were written through a non-const accessor.  The
// method names, arguments and error behaviour follow the reference:
// constructors throw std::invalid_argument, solvers return a bool converged
// flag, Interaction::apply is noexcept.
#pragma once
#include <array>
#include <chrono>
#include <cmath>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pic_abi.h"

using type_calc = double;  // all.h:11

namespace Const {  // all.h:13-23
const double eps_0 = 8.85418782e-12;
const double q_e = 1.602176565e-19;
const double amu = 1.660538921e-27;
const double m_e = 9.10938215e-31;
const double k = 1.380648e-23;
const double pi = 3.141592653;
const double eV_to_K = q_e / k;
const double N_a = 6.02214076e23;
}  // namespace Const

// Vec3.h: 3-vector value type
struct type_calc3 {
    double v[3] = {0, 0, 0};
    type_calc3() = default;
    type_calc3(double a, double b, double c) : v{a, b, c} {}
    explicit type_calc3(double a) : v{a, a, a} {}
    double& operator[](int i) { return v[i]; }
    const double& operator[](int i) const { return v[i]; }
    type_calc3 operator+(const type_calc3& o) const { return {v[0] + o[0], v[1] + o[1], v[2] + o[2]}; }
    type_calc3 operator-(const type_calc3& o) const { return {v[0] - o[0], v[1] - o[1], v[2] - o[2]}; }
    type_calc3 operator*(double s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
    type_calc3 operator/(double s) const { return *this * (1.0 / s); }
    double operator*(const type_calc3& o) const { return v[0] * o[0] + v[1] * o[1] + v[2] * o[2]; }
    double length() const { return std::sqrt(*this * *this); }
};

// Species.h:12-29 (AoS, 7 doubles): the ABI's particle layout
struct Particle {
    type_calc3 pos, vel;
    type_calc macro_weight;
    Particle(type_calc3 p, type_calc3 v, type_calc w) : pos(p), vel(v), macro_weight(w) {}
};
static_assert(sizeof(Particle) == 7 * sizeof(double), "Particle must be 7 packed doubles");

class World;

// Field.h (node data, [i][j][k] access) mirrored from the device on demand
template <class T>
class Field {
  public:
    const int ni, nj, nk;
    Field(int ni_, int nj_, int nk_) : ni(ni_), nj(nj_), nk(nk_), data((size_t)ni_ * nj_ * nk_) {}
    // reference access pattern field[i][j][k]
    struct Row {
        Field* f;
        int i, j;
        T& operator[](int k) { return f->data[((size_t)k * f->nj + j) * f->ni + i]; }
    };
    struct Plane {
        Field* f;
        int i;
        Row operator[](int j) { return Row{f, i, j}; }
    };
    Plane operator[](int i) {
        pull();
        host_dirty = true;
        return Plane{this, i};
    }
    const T& at(int i, int j, int k) {
        pull();
        return data[((size_t)k * nj + j) * ni + i];
    }
    std::vector<T>& raw() {  // i-fastest flat data (the solver index u)
        pull();
        return data;
    }
    // synchronisation hooks set by the owner
    std::function<void(std::vector<T>&)> puller;
    std::function<void(const std::vector<T>&)> pusher;
    bool device_newer = false;
    bool host_dirty = false;
    void pull() {
        if (device_newer && puller) {
            puller(data);
            device_newer = false;
        }
    }
    void push() {
        if (host_dirty && pusher) pusher(data);
        host_dirty = false;
    }

  private:
    std::vector<T> data;
};

class Species;

class World {
  public:
    const int nn[3];
    const int ni, nj, nk;
    const int ni_1, nj_1, nk_1;
    const int nv;
    const int num_cells;
    Field<type_calc> phi, rho, node_vol;
    Field<type_calc3> ef;

    World(int ni, int nj, int nk, type_calc3 x0, type_calc3 xm);  // World.cpp:22-26
    ~World();
    World(const World&) = delete;
    World& operator=(const World&) = delete;

    type_calc3 getX0() const { return x0; }
    type_calc3 getDx() const { return dx; }
    type_calc3 getXm() const { return xm; }
    type_calc3 getXc() const { return (xm + x0) * 0.5; }
    type_calc3 getL() const { return {dx[0] * ni_1, dx[1] * nj_1, dx[2] * nk_1}; }
    type_calc getCellVolume() const { return dx[0] * dx[1] * dx[2]; }
    int getNumCells() const { return num_cells; }
    type_calc getPE();                                          // World.cpp:108-118
    void computeChargeDensity(std::vector<Species>& species);  // World.cpp:193-200
    void dirichletBoundries();                                  // World.cpp:263-273
    void computeObjectID() {}  // objects are rasterised when added (World.cpp:276-292)
    template <class T, class... Args>
    void addObject(Args&&... args) {  // World.h:110-121
        T obj(std::forward<Args>(args)...);
        add_rectangle(obj.pos, obj.phi, obj.sides);
    }
    void setTime(type_calc dt_, int num_ts_) {  // World.cpp:321-324
        dt = dt_;
        num_ts = num_ts_;
        pic_set_dt(ctx, dt);
    }
    type_calc getDt() const { return dt; }
    int getTs() const { return ts; }
    type_calc getTime() const { return time; }
    bool isLastTimeStep() const { return ts == num_ts - 1; }
    bool advanceTime() {  // World.cpp:337-341
        time += dt;
        ts++;
        return ts <= num_ts;
    }
    type_calc getWallTime() const {
        return std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    }

    // device plumbing
    pic_ctx* ctx = nullptr;
    void check(int rc, const char* what) const;
    void device_fields_changed();  // phi/rho/ef on the device are newer
    void flush_host_fields();      // host writes go to the device first

  private:
    type_calc3 x0, xm, dx;
    type_calc dt = 1e-4;
    int num_ts = 0, ts = -1;
    type_calc time = 0;
    std::chrono::high_resolution_clock::time_point t0 = std::chrono::high_resolution_clock::now();
    void add_rectangle(type_calc3 pos, type_calc phi, type_calc3 sides);
};

// Object.h Rectangle(pos, phi, sides) (Object.cpp:166-172)
struct Rectangle {
    type_calc3 pos;
    type_calc phi;
    type_calc3 sides;
    Rectangle(type_calc3 p, type_calc ph, type_calc3 s) : pos(p), phi(ph), sides(s) {}
};

class Species {
  public:
    const std::string name;
    const type_calc mass, charge, mpw0;
    const type_calc E_ion = -666;
    Field<type_calc> den, T;
    Field<type_calc> den_avg;  // running mean of den over up to 20 samples (Field.h:246-261)
    Field<type_calc3> vel;
    std::vector<type_calc> macro_part_count;  // per cell, c = (k*(nj-1)+j)*(ni-1)+i

    // Species.cpp:30-39; capacity bounds the device arrays (births included)
    Species(std::string name, type_calc mass, type_calc charge, World& world, type_calc mpw0,
            type_calc E_ion = -666, size_t capacity = 1 << 22);
    size_t getNumParticles() const;
    void advanceElectrons(type_calc dt);                                       // :258-317
    void advanceNonElectron(Species& neutrals, Species& spherium, type_calc dt);  // :47-77
    void computeNumberDensity();                                               // :401-416
    void sampleMoments();                                                      // :767-776
    void computeGasProperties();                                               // :777-804
    void clearSamples();                                                       // :805-812
    void computeMacroParticlesCount();                                         // :813-819
    void loadParticleBoxThermal(type_calc3 x0, type_calc3 sides, type_calc num_den,
                                type_calc T);                                  // :564-599
    void sortIndexes();                                                        // :905-929
    void merge();                                                              // :987-1145
    void updateAverages();                                                     // :764-766
    std::vector<Particle>& getPartRef();              // host mirror, pushed back on next use
    const std::vector<Particle>& getConstPartRef();
    type_calc getMicroCount();       // :732-738
    type_calc3 getMomentum();        // :739-745
    type_calc getKE();               // :746-752
    type_calc getPE();               // :753-763
    int sid = -1;
    World& world;
    uint32_t load_seed = 1;
    uint32_t surface_seed = 1;  // draws of object reflections / emissions
    uint32_t merge_seed = 1;    // draws of merge's position picks
    pic_merge_stats last_merge{};

  private:
    std::vector<Particle> host;
    bool host_valid = false, host_dirty = false;
    int avg_samples = 0;
    void sync_to_device();
    void diag(double out[6]);
};

// Outputs.h / Outputs.cpp:9-205: the reference's files (results/ under the
// working directory), written through libpic_io (include/pic_io.h)
namespace Output {
void fieldsOutput(World& world, std::vector<Species>& species, std::string name1 = "");
void diagOutput(World& world, std::vector<Species>& species);
void particlesOutput(World& world, std::vector<Species>& species, int num_parts_to_output_base,
                     std::string name1 = "");
void screenOutput(World& world, std::vector<Species>& species);
}  // namespace Output

enum SolverType { GS, PCG, QN };  // PotentialSolver.h:24

class PotentialSolver {
  public:
    // PotentialSolver.cpp:44-52: GS sweeps 20 x max_it for the PCG type
    PotentialSolver(World& world, unsigned max_solver_it, type_calc tolerance,
                    SolverType solver_type);
    bool solve();                                                // :55-67
    void computeEF();                                            // :354-408
    void setReferenceValues(type_calc phi0, type_calc n0, type_calc Te0);  // :409-413
    unsigned get_GS_max_it() const { return GS_max_solver_it; }
    unsigned get_PCG_max_it() const { return PCG_max_solver_it; }
    // build-side options (not in the reference)
    double omega = 1.4;        // SOR weight (PotentialSolver.h:38)
    int tol_kind = PIC_TOL_REF_RMS;
    int pcg_method = PIC_SOLVER_PCG;
    pic_solver_stats last{};

  private:
    World& world;
    const SolverType solver_type;
    unsigned PCG_max_solver_it, GS_max_solver_it;
    type_calc tolerance, phi0 = 0, n0 = 0, Te0 = 1;
};

class Interaction {  // Interactions.h:9-13
  public:
    virtual ~Interaction() = default;
    virtual void apply(type_calc dt) noexcept = 0;
};

class MC_MEX_Ionization : public Interaction {  // Interactions.h:99-143
  public:
    MC_MEX_Ionization(Species& neutrals, Species& ions, Species& electrons, World& world,
                      std::string coll_data_path = "data/Oxygen_momentum_transfer.txt",
                      int freq_should_use_map = 10);
    void apply(type_calc dt) noexcept override;
    pic_mcc_stats last{};
    uint32_t seed = 1;
    // initial W*sigma*v_rel maximum; 0 keeps the reference's 1e-14 * max(mpw0)
    // (Interactions.cpp:129, :534).  Read when the first apply() creates the
    // device object.
    double wsv_max_init = 0;
    int last_status = PIC_OK;

  private:
    Species &neutrals, &ions, &electrons;
    World& world;
    int id = -1;
    std::vector<double> tE, tV;
    pic_mcc_desc desc{};
    uint32_t step = 0;
};

// DSMC_MEX (Interactions.h:42-67): equal-weight DSMC with Bird's VHS cross
// section inside one species or between two of the same mpw0
class DSMC_MEX : public Interaction {
  public:
    DSMC_MEX(Species& species, World& world) noexcept;                 // Interactions.cpp:143-155
    DSMC_MEX(Species& species1, Species& species2, World& world);      // :156-172
    void apply(type_calc dt) noexcept override;                       // :179-181
    pic_dsmc_stats last{};
    uint32_t seed = 1;
    int last_status = PIC_OK;

  private:
    World& world;
    int id = -1;
    uint32_t step = 0;
};

// Source.h: particles injected through one face of the box every sample()
class Source {
  public:
    virtual ~Source() = default;
    virtual void sample() const noexcept;  // Source.cpp:101-104, :178-181
    uint32_t seed = 1;
    mutable size_t last_added = 0;
    mutable int last_status = PIC_OK;

  protected:
    // Source.cpp:3-28; area_frac is accepted and unused, as in the reference
    Source(Species& species, World& world, type_calc v_drift, type_calc den, type_calc T,
           std::string inlet_face, type_calc area_frac) noexcept;
    World& world;
    int id = -1;
};
class ColdBeamSource : public Source {  // Source.cpp:32-104
  public:
    ColdBeamSource(Species& species, World& world, type_calc v_drift, type_calc den,
                   std::string inlet_face = "-z", type_calc area_frac = 1.0) noexcept;
};
class WarmBeamSource : public Source {  // Source.cpp:105-181
  public:
    WarmBeamSource(Species& species, World& world, type_calc v_drift, type_calc den, type_calc T,
                   std::string inlet_face = "-z", type_calc area_frac = 1.0) noexcept;
};
