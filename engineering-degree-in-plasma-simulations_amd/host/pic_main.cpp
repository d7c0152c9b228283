// pic_main.cpp -- driver over the host classes, following the reference's
// main loop (ch4/v3/src/main.cpp:177-288) with the per-step field solve of
// ch4/v2/main.cpp:114-136.
//
//   pic_driver --config plain     [--nodes N] [--particles P] [--num_ts K]
//              --config discharge [--nodes N] [--particles P] [--num_ts K]
//   common:    [--tol T] [--tol-rel] [--solver GS|PCG] [--dump PREFIX]
//              [--diag FILE.csv] [--output N] [--seed S]
//
// plain:     Dirichlet phi = 0 box, dx = 1e-2 m, kinetic e- (2 eV) and O+
//            (16 amu, 0.1 eV), n0 = 1e12 m^-3, dt = 1e-10 s, no subcycling,
//            solve + E-field every step.
// discharge: parallel-plate electrodes (Rectangle slabs of 0.1 L at the z
//            ends, +phi / 0 V), O2 background as variable-weight neutrals,
//            electron seed near the cathode, MC_MEX_Ionization every step,
//            the reference's subcycling (main.cpp:211-237).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/pic_io.h"
#include "pic_host.hpp"

static const char* arg(int argc, char** argv, const char* key, const char* def) {
    for (int i = 1; i + 1 < argc; i++)
        if (!std::strcmp(argv[i], key)) return argv[i + 1];
    return def;
}
static bool flag(int argc, char** argv, const char* key) {
    for (int i = 1; i < argc; i++)
        if (!std::strcmp(argv[i], key)) return true;
    return false;
}

static void dump(const std::string& path, Species& sp) {
    const std::vector<Particle>& p = sp.getConstPartRef();
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    if (!p.empty()) std::fwrite(p.data(), sizeof(Particle), p.size(), f);
    std::fclose(f);
}

// runtime diagnostics in the columns of Output::diagOutput
// (Outputs.cpp:143-179), to a file of the caller's choosing
static void diag_line(const std::string& path, bool header, World& w, std::vector<Species>& sps) {
    std::vector<const char*> names;
    std::vector<double> d;
    for (Species& s : sps) {
        names.push_back(s.name.c_str());
        type_calc3 m = s.getMomentum();
        d.insert(d.end(), {(double)s.getNumParticles(), s.getMicroCount(), m[0], m[1], m[2], s.getKE()});
    }
    if (pic_io_diag_row(path.c_str(), header ? 1 : 0, (int)names.size(), names.data(), w.getTs(), w.getTime(),
                        w.getWallTime(), d.data(), w.getPE()) != 0)
        throw std::runtime_error("cannot write " + path);
}

int main(int argc, char** argv) {
    std::string config = arg(argc, argv, "--config", "plain");
    int n = std::atoi(arg(argc, argv, "--nodes", "32"));
    long np = std::atol(arg(argc, argv, "--particles", "200000"));
    int num_ts = std::atoi(arg(argc, argv, "--num_ts", "200"));
    double tol = std::atof(arg(argc, argv, "--tol", "1e-6"));
    std::string solver = arg(argc, argv, "--solver", "GS");
    std::string dump_prefix = arg(argc, argv, "--dump", "");
    std::string diag_path = arg(argc, argv, "--diag", "");
    uint32_t seed = (uint32_t)std::atoi(arg(argc, argv, "--seed", "1"));
    // --output N: the reference's files under results/ (Outputs.cpp): the
    // fields and 1000 sampled particles per species every N steps, the
    // runtime diagnostics every 10
    const int output_every = std::atoi(arg(argc, argv, "--output", "0"));
    if (output_every > 0) std::system("mkdir -p results");
    bool diag_header = true;

    if (config == "plain") {
        double dx = 1e-2, L = dx * (n - 1), n0 = 1e12, dt = 1e-10;
        World world(n, n, n, {0, 0, 0}, {L, L, L});
        world.setTime(dt, num_ts);
        world.dirichletBoundries();
        double w = n0 * L * L * L / (double)np;
        std::vector<Species> species;
        species.emplace_back("e-", Const::m_e, -Const::q_e, world, w, -666, (size_t)(np * 1.05) + 64);
        species.emplace_back("O+", 16 * Const::amu, Const::q_e, world, w, -666, (size_t)(np * 1.05) + 64);
        species[0].load_seed = seed;
        species[1].load_seed = seed + 1000;
        species[0].loadParticleBoxThermal(world.getXc(), world.getL(), n0, 2.0 * Const::eV_to_K);
        species[1].loadParticleBoxThermal(world.getXc(), world.getL(), n0, 0.1 * Const::eV_to_K);
        PotentialSolver solver_obj(world, solver == "GS" ? 20000 : 5000, tol,
                                   solver == "GS" ? GS : PCG);
        solver_obj.omega = 2.0 / (1.0 + std::sin(Const::pi / (n - 1)));
        if (flag(argc, argv, "--tol-rel")) solver_obj.tol_kind = PIC_TOL_REL;
        for (Species& sp : species) sp.computeNumberDensity();
        world.computeChargeDensity(species);
        solver_obj.solve();
        solver_obj.computeEF();
        if (!dump_prefix.empty()) {
            dump(dump_prefix + "_e0.bin", species[0]);
            dump(dump_prefix + "_i0.bin", species[1]);
        }
        while (world.advanceTime()) {
            if (world.getTs() >= num_ts) break;
            species[0].advanceElectrons(world.getDt());
            species[1].advanceNonElectron(species[1], species[1], world.getDt());
            for (Species& sp : species) sp.computeNumberDensity();
            world.computeChargeDensity(species);
            if (!solver_obj.solve()) return 3;
            solver_obj.computeEF();
            if (!diag_path.empty() && (world.getTs() % 10 == 0 || world.isLastTimeStep())) {
                diag_line(diag_path, diag_header, world, species);
                diag_header = false;
            }
            if (output_every > 0) {
                if (world.getTs() % 10 == 0) Output::diagOutput(world, species);
                if (world.getTs() % output_every == 0) {
                    for (Species& sp : species) {
                        sp.sampleMoments();
                        sp.updateAverages();
                    }
                    Output::fieldsOutput(world, species);
                    Output::particlesOutput(world, species, 1000);
                }
            }
        }
        if (!dump_prefix.empty()) {
            dump(dump_prefix + "_e1.bin", species[0]);
            dump(dump_prefix + "_i1.bin", species[1]);
            std::vector<double>& phi = world.phi.raw();
            FILE* f = std::fopen((dump_prefix + "_phi.bin").c_str(), "wb");
            std::fwrite(phi.data(), sizeof(double), phi.size(), f);
            std::fclose(f);
        }
        std::printf("{\"config\": \"plain\", \"steps\": %d, \"electrons\": %zu, \"ions\": %zu, "
                    "\"last_solve_iterations\": %d, \"last_rel_residual\": %.3e, \"wall_s\": %.3f}\n",
                    num_ts, species[0].getNumParticles(), species[1].getNumParticles(),
                    solver_obj.last.iterations, solver_obj.last.rel_residual, world.getWallTime());
        return 0;
    }

    if (config == "discharge") {
        // plates of 0.1 L at both z ends (main.cpp:94-98), O2 at 10 Pa / 300 K
        int nz = 4 * (n - 1) + 1;
        double dx = 1e-4;
        World world(n, n, nz, {0, 0, 0}, {dx * (n - 1), dx * (n - 1), dx * (nz - 1)});
        double dt = std::atof(arg(argc, argv, "--dt", "1e-12"));
        world.setTime(dt, num_ts);
        type_calc3 L = world.getL(), c = world.getXc();
        double phi_plate = std::atof(arg(argc, argv, "--phi", "300"));
        world.addObject<Rectangle>(type_calc3(c[0], c[1], world.getX0()[2]), phi_plate,
                                   type_calc3(L[0], L[1], L[2] * 0.1));
        world.addObject<Rectangle>(type_calc3(c[0], c[1], world.getXm()[2]), 0.0,
                                   type_calc3(L[0], L[1], L[2] * 0.1));
        double n_gas = 10.0 / (Const::k * 300.0);
        double gas_vol = L[0] * L[1] * L[2] * 0.8;
        double w_n = n_gas * gas_vol / (double)np;
        double w_e = w_n / 1000.0;
        double E_ion_O2 = 12.07 * Const::q_e;
        std::vector<Species> species;
        species.reserve(3);
        species.emplace_back("O2", 32 * Const::amu, 0, world, w_n, E_ion_O2, (size_t)(np * 1.3) + 1024);
        species.emplace_back("O2+", 32 * Const::amu, Const::q_e, world, w_e, -666, (size_t)np / 2 + 1024);
        species.emplace_back("e-", Const::m_e, -Const::q_e, world, w_e, -666, (size_t)np / 2 + 1024);
        Species &neu = species[0], &ion = species[1], &ele = species[2];
        neu.load_seed = seed;
        ele.load_seed = seed + 7;
        neu.loadParticleBoxThermal(c, type_calc3(L[0], L[1], L[2] * 0.79), n_gas, 300.0);
        ele.loadParticleBoxThermal(type_calc3(c[0], c[1], c[2] - 0.3 * L[2]),
                                   type_calc3(L[0] * 0.5, L[1] * 0.5, L[2] * 0.05),
                                   n_gas * std::atof(arg(argc, argv, "--ne-frac", "1e-3")),
                                   std::atof(arg(argc, argv, "--te", "10")) * Const::eV_to_K);
        std::string table = arg(argc, argv, "--sigma", "data/sigma_el_O.dat");
        MC_MEX_Ionization mcc(neu, ion, ele, world, table);
        mcc.seed = seed;
        mcc.wsv_max_init = std::atof(arg(argc, argv, "--wsv0", "0"));
        PotentialSolver solver_obj(world, 20000, tol, GS);
        solver_obj.omega = 1.9;
        solver_obj.setReferenceValues(0, 0, 1e20);  // main.cpp:138
        solver_obj.solve();                          // solve once (main.cpp:172-173)
        solver_obj.computeEF();
        long long iz = 0, coll = 0;
        while (world.advanceTime()) {
            if (world.getTs() >= num_ts) break;
            mcc.apply(world.getDt());
            if (mcc.last_status != PIC_OK) return 4;
            iz += mcc.last.ionizations;
            coll += mcc.last.collisions;
            // subcycling (main.cpp:211-237): e- every step, ions every 10th
            // step with 10 dt, neutrals every 100th with 50 dt
            ele.advanceElectrons(world.getDt());
            ele.computeNumberDensity();
            if (world.getTs() % 10 == 0) {
                ion.advanceNonElectron(neu, neu, world.getDt() * 10);
                ion.computeNumberDensity();
            }
            if (world.getTs() % 100 == 0) {
                neu.advanceNonElectron(neu, neu, world.getDt() * 50);
                neu.computeNumberDensity();
            }
            world.computeChargeDensity(species);
            if (!diag_path.empty() && (world.getTs() % 10 == 0 || world.isLastTimeStep())) {
                diag_line(diag_path, diag_header, world, species);
                diag_header = false;
            }
            if (output_every > 0) {
                if (world.getTs() % 10 == 0) Output::diagOutput(world, species);
                if (world.getTs() % output_every == 0) {
                    for (Species& sp : species) {
                        sp.sampleMoments();
                        sp.updateAverages();
                    }
                    Output::fieldsOutput(world, species);
                    Output::particlesOutput(world, species, 1000);
                }
            }
        }
        std::printf("{\"config\": \"discharge\", \"steps\": %d, \"neutrals\": %zu, \"ions\": %zu, "
                    "\"electrons\": %zu, \"collisions\": %lld, \"ionizations\": %lld, "
                    "\"wall_s\": %.3f}\n",
                    num_ts, neu.getNumParticles(), ion.getNumParticles(), ele.getNumParticles(),
                    coll, iz, world.getWallTime());
        return 0;
    }
    std::fprintf(stderr, "unknown --config %s\n", config.c_str());
    return 2;
}
