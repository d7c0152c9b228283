// pic_host.cpp -- host mirrors of World / Species / PotentialSolver /
// MC_MEX_Ionization over the C ABI.  Every compute call goes to the gfx950
// library; the host only keeps mirrors for the reference's public fields.
#include "pic_host.hpp"

#include <cctype>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <sstream>

#include "../../include/pic_io.h"

// ---------------------------------------------------------------- World
World::World(int ni_, int nj_, int nk_, type_calc3 x0_, type_calc3 xm_)
    : nn{ni_, nj_, nk_},
      ni(ni_),
      nj(nj_),
      nk(nk_),
      ni_1(ni_ - 1),
      nj_1(nj_ - 1),
      nk_1(nk_ - 1),
      nv(ni_ * nj_ * nk_),
      num_cells((ni_ - 1) * (nj_ - 1) * (nk_ - 1)),
      phi(ni_, nj_, nk_),
      rho(ni_, nj_, nk_),
      node_vol(ni_, nj_, nk_),
      ef(ni_, nj_, nk_),
      x0(x0_),
      xm(xm_) {
    pic_grid_desc g{ni_, nj_, nk_, {x0_[0], x0_[1], x0_[2]}, {xm_[0], xm_[1], xm_[2]}};
    int rc = pic_create(&g, &ctx);
    if (rc != PIC_OK) throw std::invalid_argument("World: no gfx950 device or invalid grid");
    for (int d = 0; d < 3; d++) dx[d] = (xm[d] - x0[d]) / (nn[d] - 1);  // World.cpp:72-76
    phi.puller = [this](std::vector<double>& h) { check(pic_field_download(ctx, PIC_FIELD_PHI, h.data()), "phi"); };
    phi.pusher = [this](const std::vector<double>& h) { check(pic_field_upload(ctx, PIC_FIELD_PHI, h.data()), "phi"); };
    rho.puller = [this](std::vector<double>& h) { check(pic_field_download(ctx, PIC_FIELD_RHO, h.data()), "rho"); };
    rho.pusher = [this](const std::vector<double>& h) { check(pic_field_upload(ctx, PIC_FIELD_RHO, h.data()), "rho"); };
    node_vol.puller = [this](std::vector<double>& h) {
        check(pic_field_download(ctx, PIC_FIELD_NODE_VOL, h.data()), "node_vol");
    };
    ef.puller = [this](std::vector<type_calc3>& h) {
        check(pic_field_download(ctx, PIC_FIELD_EF, &h[0][0]), "ef");
    };
    ef.pusher = [this](const std::vector<type_calc3>& h) {
        check(pic_field_upload(ctx, PIC_FIELD_EF, &h[0][0]), "ef");
    };
    node_vol.device_newer = true;
    device_fields_changed();
}

World::~World() { pic_destroy(ctx); }

void World::check(int rc, const char* what) const {
    if (rc < 0) throw std::runtime_error(std::string(what) + ": " + pic_last_error(ctx));
}
void World::device_fields_changed() {
    phi.device_newer = rho.device_newer = ef.device_newer = true;
}
void World::flush_host_fields() {
    phi.push();
    rho.push();
    ef.push();
}
void World::dirichletBoundries() { check(pic_dirichlet_walls(ctx), "dirichletBoundries"); }
void World::add_rectangle(type_calc3 pos, type_calc ph, type_calc3 sides) {
    flush_host_fields();
    check(pic_add_rectangle(ctx, pos.v, sides.v, ph), "addObject");
    device_fields_changed();
}
type_calc World::getPE() {
    flush_host_fields();
    double v = 0;
    check(pic_field_energy(ctx, &v), "getPE");
    return v;
}
void World::computeChargeDensity(std::vector<Species>& species) {
    (void)species;  // the device holds every species' density
    flush_host_fields();
    check(pic_charge_density(ctx), "computeChargeDensity");
    rho.device_newer = true;
}

// -------------------------------------------------------------- Species
Species::Species(std::string name_, type_calc mass_, type_calc charge_, World& w, type_calc mpw0_,
                 type_calc E_ion_, size_t capacity)
    : name(std::move(name_)),
      mass(mass_),
      charge(charge_),
      mpw0(mpw0_),
      E_ion(E_ion_),
      den(w.ni, w.nj, w.nk),
      T(w.ni, w.nj, w.nk),
      den_avg(w.ni, w.nj, w.nk),
      vel(w.ni, w.nj, w.nk),
      world(w) {
    w.check(pic_species_add(w.ctx, mass, charge, mpw0, capacity, &sid), "Species");
    // captures the World and the species id, not this: Species objects may be
    // moved when a std::vector<Species> grows, as in the reference's main
    World* wp = &w;
    int id = sid;
    den.puller = [wp, id](std::vector<double>& h) {
        wp->check(pic_species_field_download(wp->ctx, id, PIC_SFIELD_DEN, h.data()), "den");
    };
}

void Species::sync_to_device() {
    if (host_dirty) {
        world.check(pic_particles_upload(world.ctx, sid, reinterpret_cast<const double*>(host.data()),
                                         host.size()),
                    "particles");
        host_dirty = false;
    }
}
std::vector<Particle>& Species::getPartRef() {
    if (!host_valid) {
        size_t n = 0;
        world.check(pic_particles_download(world.ctx, sid, nullptr, 0, &n), "getPartRef");
        host.assign(n, Particle({}, {}, 0));
        world.check(pic_particles_download(world.ctx, sid, reinterpret_cast<double*>(host.data()), n, &n),
                    "getPartRef");
        host_valid = true;
    }
    host_dirty = true;  // a mutable reference may be written
    return host;
}
const std::vector<Particle>& Species::getConstPartRef() {
    std::vector<Particle>& h = getPartRef();
    host_dirty = false;
    return h;
}
size_t Species::getNumParticles() const {
    if (host_valid && host_dirty) return host.size();
    size_t n = 0;
    world.check(pic_particles_count(world.ctx, sid, &n), "getNumParticles");
    return n;
}
void Species::advanceElectrons(type_calc dt) {  // objects absorb (Species.cpp:333-338)
    sync_to_device();
    world.flush_host_fields();
    world.check(pic_species_set_surface(world.ctx, sid, PIC_SURF_ABSORB, -1, surface_seed),
                "advanceElectrons");
    world.check(pic_push_sort(world.ctx, sid, dt), "advanceElectrons");
    host_valid = false;
}
// advanceNoSputteringSerial (Species.cpp:170-256): neutrals reflect at
// objects, ions are recycled into `neutrals`.  Sputtering into `spherium`
// (advanceSputteringSerial, :79-168) is off in the reference's runs
// (Config SPUTTERING) and not built.
void Species::advanceNonElectron(Species& neutrals, Species& spherium, type_calc dt) {
    (void)spherium;
    sync_to_device();
    neutrals.sync_to_device();
    world.flush_host_fields();
    const bool neutral = charge == 0;
    // an ion species passed as its own neutrals (a run without neutrals)
    // has nowhere to recycle into: its object hits absorb
    const int mode = neutral ? PIC_SURF_REFLECT : (neutrals.sid == sid ? PIC_SURF_ABSORB : PIC_SURF_RECYCLE);
    world.check(pic_species_set_surface(world.ctx, sid, mode, mode == PIC_SURF_RECYCLE ? neutrals.sid : -1,
                                        surface_seed),
                "advanceNonElectron");
    world.check(pic_push_sort(world.ctx, sid, dt), "advanceNonElectron");
    host_valid = false;
    if (!neutral) neutrals.host_valid = false;
}
void Species::merge() {
    sync_to_device();
    world.flush_host_fields();
    world.check(pic_merge(world.ctx, sid, merge_seed, &last_merge), "merge");
    host_valid = false;
}
void Species::computeNumberDensity() {
    sync_to_device();
    world.check(pic_deposit(world.ctx, sid, PIC_DEPOSIT_DENSITY), "computeNumberDensity");
    den.device_newer = true;
}
void Species::sampleMoments() {
    sync_to_device();
    world.check(pic_deposit(world.ctx, sid, PIC_DEPOSIT_MOMENTS), "sampleMoments");
}
void Species::clearSamples() { world.check(pic_clear_samples(world.ctx, sid), "clearSamples"); }
void Species::computeGasProperties() {
    size_t n = (size_t)world.nv;
    std::vector<double> ns(n), nv(3 * n), uu(n), vv(n), ww(n);
    world.check(pic_species_field_download(world.ctx, sid, PIC_SFIELD_NSUM, ns.data()), "moments");
    world.check(pic_species_field_download(world.ctx, sid, PIC_SFIELD_NVSUM, nv.data()), "moments");
    world.check(pic_species_field_download(world.ctx, sid, PIC_SFIELD_NUU, uu.data()), "moments");
    world.check(pic_species_field_download(world.ctx, sid, PIC_SFIELD_NVV, vv.data()), "moments");
    world.check(pic_species_field_download(world.ctx, sid, PIC_SFIELD_NWW, ww.data()), "moments");
    std::vector<type_calc3>& V = vel.raw();
    std::vector<double>& TT = T.raw();
    for (size_t u = 0; u < n; u++) {  // Species.cpp:777-804
        double c = ns[u];
        type_calc3 m(nv[3 * u], nv[3 * u + 1], nv[3 * u + 2]);
        V[u] = c != 0 ? m / c : type_calc3();
        if (c <= 0) {
            TT[u] = 0;
            continue;
        }
        double a = uu[u] / c - V[u][0] * V[u][0], b = vv[u] / c - V[u][1] * V[u][1],
               d = ww[u] / c - V[u][2] * V[u][2];
        TT[u] = mass / (2 * Const::k) * (a + b + d);
    }
}
// Species::updateAverages -> Field::updateMovingAverage (Field.h:246-261):
// the mean of up to 20 samples, restarted after the 20th
void Species::updateAverages() {
    std::vector<double>& a = den_avg.raw();
    const std::vector<double>& d = den.raw();
    if (++avg_samples > 20) {
        std::fill(a.begin(), a.end(), 0.0);
        avg_samples = 1;
    }
    for (size_t u = 0; u < a.size(); u++) a[u] = (d[u] + a[u] * (avg_samples - 1.0)) / avg_samples;
}
void Species::computeMacroParticlesCount() {
    sync_to_device();
    macro_part_count.assign((size_t)world.num_cells, 0.0);
    world.check(pic_species_field_download(world.ctx, sid, PIC_SFIELD_MACRO_COUNT,
                                           macro_part_count.data()),
                "computeMacroParticlesCount");
}
void Species::loadParticleBoxThermal(type_calc3 x0, type_calc3 sides, type_calc num_den,
                                     type_calc temp) {
    sync_to_device();
    world.flush_host_fields();
    size_t n = 0;
    int rc = pic_load_box_thermal(world.ctx, sid, x0.v, sides.v, num_den, temp, load_seed++, 0, &n);
    if (rc == PIC_ERR_ARG)
        throw std::invalid_argument("number of macroparticles less than 1, change initial values\n");
    world.check(rc, "loadParticleBoxThermal");
    world.check(pic_sort(world.ctx, sid), "loadParticleBoxThermal");
    host_valid = false;
}
void Species::sortIndexes() {
    sync_to_device();
    world.check(pic_sort(world.ctx, sid), "sortIndexes");
    host_valid = false;
}
void Species::diag(double out[6]) {
    sync_to_device();
    world.flush_host_fields();
    world.check(pic_diag_sums(world.ctx, sid, out), "diagnostics");
}
type_calc Species::getMicroCount() {
    double d[6];
    diag(d);
    return d[0];
}
type_calc3 Species::getMomentum() {
    double d[6];
    diag(d);
    return {d[1], d[2], d[3]};
}
type_calc Species::getKE() {
    double d[6];
    diag(d);
    return d[4];
}
type_calc Species::getPE() {
    double d[6];
    diag(d);
    return d[5];
}

// ------------------------------------------------------ PotentialSolver
PotentialSolver::PotentialSolver(World& w, unsigned max_solver_it, type_calc tol, SolverType type)
    : world(w), solver_type(type), PCG_max_solver_it(max_solver_it), tolerance(tol) {
    GS_max_solver_it = type == GS ? max_solver_it : 20 * max_solver_it;
}
void PotentialSolver::setReferenceValues(type_calc p0, type_calc n0_, type_calc te0) {
    phi0 = p0;
    n0 = n0_;
    Te0 = te0;
}
bool PotentialSolver::solve() {
    if (solver_type == QN) {  // PotentialSolver.cpp:349-352
        std::cerr << "quasi neutral not implemented yet\n";
        return false;
    }
    world.flush_host_fields();
    pic_solver_params p{};
    p.method = solver_type == GS ? PIC_SOLVER_SOR_RB : pcg_method;
    p.tol_kind = tol_kind;
    p.tol = tolerance;
    p.max_it = (int)(solver_type == GS ? GS_max_solver_it : PCG_max_solver_it);
    p.omega = omega;
    p.check_every = solver_type == GS ? 25 : 8;
    p.phi0 = phi0;
    p.n0 = n0;
    p.Te0 = Te0;
    int rc = pic_poisson(world.ctx, &p, &last);
    if (rc < 0) world.check(rc, "solve");
    world.device_fields_changed();
    if (rc == PIC_NOT_CONVERGED && solver_type == PCG) {
        // fall back to Gauss-Seidel (:222-224): the reference runs up to
        // GS_max_solver_it sweeps in each of its up to 20 Newton steps
        // (:179, :315-347); here one red-black SOR solve of up to that many
        // sweeps in all, from the phi PCG left, to the same criterion
        p.method = PIC_SOLVER_SOR_RB;
        p.max_it = (int)(20 * GS_max_solver_it);
        p.check_every = 25;
        rc = pic_poisson(world.ctx, &p, &last);
        if (rc < 0) world.check(rc, "solve");
    }
    if (rc != PIC_OK)
        std::cerr << "solver failed to converge, residual = " << last.ref_l2
                  << " tolerance = " << tolerance << " time step = " << world.getTs() << std::endl;
    return rc == PIC_OK;
}
void PotentialSolver::computeEF() {
    world.flush_host_fields();
    world.check(pic_efield(world.ctx), "computeEF");
    world.ef.device_newer = true;
}

// ---------------------------------------------------- MC_MEX_Ionization
MC_MEX_Ionization::MC_MEX_Ionization(Species& n, Species& i, Species& e, World& w,
                                     std::string coll_data_path, int freq_should_use_map)
    : neutrals(n), ions(i), electrons(e), world(w) {
    (void)freq_should_use_map;
    // Interactions.cpp:479-489 guards
    if (n.mpw0 < 1e2 * e.mpw0)
        throw std::invalid_argument("neutrals species should have mpw0 at least 100 times greater than electrons");
    if (!n.E_ion || n.E_ion < 0)
        throw std::invalid_argument("neutral species must have proper ionization energy E_ion (in Joules)");
    if (n.mpw0 < i.mpw0) throw std::invalid_argument("neutrals mpw0 must be >= ions mpw0");
    // Interactions.cpp:518-533: (energy eV, sigma m^2) pairs, map-ordered
    std::ifstream in(coll_data_path);
    if (!in.is_open())
        throw std::invalid_argument("Couldn't open file " + coll_data_path +
                                    " for crosssection data of collisions");
    std::map<double, double> table;
    double E, s;
    while (in >> E >> s) table[E] = s;
    for (auto& kv : table) {
        tE.push_back(kv.first);
        tV.push_back(kv.second);
    }
    if (tE.empty()) throw std::invalid_argument("empty cross-section table " + coll_data_path);
    desc.E_ion_J = n.E_ion;
    desc.c0 = 1.015e-18;  // Interactions.cpp:506-508
    desc.c1 = 9.793;
    desc.c2 = 61.81;
    desc.B_inc = 10.0;  // Interactions.cpp:511
}
void MC_MEX_Ionization::apply(type_calc dt) noexcept {
    if (id < 0) {
        desc.wsv_max_init = wsv_max_init;
        last_status = pic_mcc_create(world.ctx, neutrals.sid, ions.sid, electrons.sid, &desc,
                                     tE.data(), tV.data(), (int)tE.size(), &id);
        if (last_status != PIC_OK) {
            std::cerr << "MC_MEX_Ionization: " << pic_last_error(world.ctx) << "\n";
            return;
        }
    }
    last_status = pic_mcc_apply(world.ctx, id, dt, seed, step++, &last);
    if (last_status != PIC_OK) std::cerr << "MC_MEX_Ionization: " << pic_last_error(world.ctx) << "\n";
}

// --------------------------------------------------------------- DSMC_MEX
DSMC_MEX::DSMC_MEX(Species& species, World& w) noexcept : world(w) {
    if (pic_dsmc_create(world.ctx, species.sid, species.sid, &id) != PIC_OK) {
        std::cerr << "DSMC_MEX: " << pic_last_error(world.ctx) << "\n";
        id = -1;
    }
}
DSMC_MEX::DSMC_MEX(Species& species1, Species& species2, World& w) : world(w) {
    if (species1.mpw0 != species2.mpw0)  // Interactions.cpp:160-162
        throw std::invalid_argument(
            "species must have the same macroparticle weight for this algorithm to work properly.");
    world.check(pic_dsmc_create(world.ctx, species1.sid, species2.sid, &id), "DSMC_MEX");
}
void DSMC_MEX::apply(type_calc dt) noexcept {
    if (id < 0) {
        last_status = PIC_ERR_STATE;
        return;
    }
    last_status = pic_dsmc_apply(world.ctx, id, dt, seed, step++, &last);
    if (last_status != PIC_OK) std::cerr << "DSMC_MEX: " << pic_last_error(world.ctx) << "\n";
}

// ----------------------------------------------------------------- Source
// inletName2Index (World.cpp:369-390); an unknown name throws, which a
// noexcept constructor turns into std::terminate as in the reference
static int inlet_index(std::string face) {
    for (char& ch : face) ch = (char)std::tolower((unsigned char)ch);
    static const std::map<std::string, int> faces{{"x-", 0}, {"x+", 1}, {"y-", 2}, {"y+", 3},
                                                  {"z-", 4}, {"z+", 5}, {"-x", 0}, {"+x", 1},
                                                  {"-y", 2}, {"+y", 3}, {"-z", 4}, {"+z", 5}};
    auto it = faces.find(face);
    if (it == faces.end()) throw std::invalid_argument("Passed wrong face name.");
    return it->second;
}
Source::Source(Species& species, World& w, type_calc v_drift, type_calc den, type_calc T,
               std::string inlet_face, type_calc area_frac) noexcept
    : world(w) {
    (void)area_frac;
    const int face = inlet_index(inlet_face);
    world.flush_host_fields();
    if (pic_source_create(world.ctx, species.sid, face, v_drift, den, T, &id) != PIC_OK) {
        std::cerr << "Source: " << pic_last_error(world.ctx) << "\n";
        id = -1;
    }
    world.device_fields_changed();  // the inlet face is Dirichlet at phi = 0
}
void Source::sample() const noexcept {
    if (id < 0) {
        last_status = PIC_ERR_STATE;
        return;
    }
    last_status = pic_source_sample(world.ctx, id, seed, &last_added);
    if (last_status != PIC_OK) std::cerr << "Source: " << pic_last_error(world.ctx) << "\n";
}
ColdBeamSource::ColdBeamSource(Species& species, World& w, type_calc v_drift, type_calc den,
                               std::string inlet_face, type_calc area_frac) noexcept
    : Source(species, w, v_drift, den, 0.0, std::move(inlet_face), area_frac) {}
WarmBeamSource::WarmBeamSource(Species& species, World& w, type_calc v_drift, type_calc den,
                               type_calc T, std::string inlet_face, type_calc area_frac) noexcept
    : Source(species, w, v_drift, den, T, std::move(inlet_face), area_frac) {}

// ----------------------------------------------------------------- Output
namespace Output {

static std::string step_name(const std::string& stem, int ts, const char* ext) {
    std::stringstream name;  // Outputs.cpp:17, :220
    name << stem << std::setfill('0') << std::setw(5) << ts << ext;
    return name.str();
}

void fieldsOutput(World& world, std::vector<Species>& species, std::string name1) {
    for (Species& sp : species) {
        sp.computeGasProperties();
        sp.computeMacroParticlesCount();
    }
    const std::string path = step_name("results/" + name1 + "fields_", world.getTs(), ".vti");
    const size_t nn = (size_t)world.nv;
    std::vector<int32_t> oid(nn), nt(nn);
    world.check(pic_get_node_flags(world.ctx, oid.data(), nt.data()), "fieldsOutput");
    std::vector<double> ef(3 * nn);
    const std::vector<type_calc3>& e = world.ef.raw();
    for (size_t u = 0; u < nn; u++)
        for (int c = 0; c < 3; c++) ef[3 * u + c] = e[u][c];
    std::vector<std::vector<double>> vel(species.size());
    std::vector<pic_io_species> sp(species.size());
    for (size_t s = 0; s < species.size(); s++) {
        Species& x = species[s];
        vel[s].resize(3 * nn);
        const std::vector<type_calc3>& v = x.vel.raw();
        for (size_t u = 0; u < nn; u++)
            for (int c = 0; c < 3; c++) vel[s][3 * u + c] = v[u][c];
        sp[s] = pic_io_species{x.name.c_str(), x.den.raw().data(), x.den_avg.raw().data(), vel[s].data(),
                               x.T.raw().data(), x.macro_part_count.data()};
    }
    const type_calc3 x0 = world.getX0(), dx = world.getDx();
    pic_io_fields f{world.ni, world.nj, world.nk, {x0[0], x0[1], x0[2]}, {dx[0], dx[1], dx[2]},
                    world.node_vol.raw().data(), oid.data(), nt.data(), world.phi.raw().data(),
                    world.rho.raw().data(), ef.data(), (int)sp.size(), sp.data()};
    if (pic_io_write_vti(path.c_str(), &f) != 0) {
        std::cerr << "Could not open " << path << std::endl;
        return;
    }
    for (Species& x : species) x.clearSamples();  // Outputs.cpp:119-121 (no steady-state test here)
}

void particlesOutput(World& world, std::vector<Species>& species, int num_parts_to_output_base,
                     std::string name1) {
    for (Species& sp : species) {
        const std::string path =
            step_name("results/" + name1 + "parts_" + sp.name + "_", world.getTs(), ".vtp");
        const std::vector<Particle>& p = sp.getConstPartRef();
        if (pic_io_write_vtp(path.c_str(), sp.name.c_str(), reinterpret_cast<const double*>(p.data()), p.size(),
                             num_parts_to_output_base) != 0) {
            std::cerr << "Could not open " << path << std::endl;
            return;
        }
    }
}

static bool diag_started = false;

void diagOutput(World& world, std::vector<Species>& species) {
    std::vector<const char*> names;
    std::vector<double> d;
    for (Species& sp : species) {
        names.push_back(sp.name.c_str());
        type_calc3 m = sp.getMomentum();
        d.insert(d.end(), {(double)sp.getNumParticles(), sp.getMicroCount(), m[0], m[1], m[2], sp.getKE()});
    }
    const double pe = world.getPE();
    if (pic_io_diag_row("results/runtime_diags.csv", diag_started ? 0 : 1, (int)names.size(), names.data(),
                        world.getTs(), world.getTime(), world.getWallTime(), d.data(), pe) != 0) {
        std::cerr << "Could not open results/runtime_diags.csv" << std::endl;
        return;
    }
    diag_started = true;
}

void screenOutput(World& world, std::vector<Species>& species) {  // Outputs.cpp:127-136
    std::cout << "\r                                                                  \rts: " << world.getTs();
    for (Species& sp : species)
        std::cout << std::setprecision(3) << "\t " << sp.name << ":" << sp.getNumParticles() << " ";
    std::cout << std::flush;
}

}  // namespace Output
