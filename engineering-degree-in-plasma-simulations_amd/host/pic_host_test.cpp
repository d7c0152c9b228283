// pic_host_test.cpp -- exercises the C++ host classes (the reference's
// interfaces over the C ABI) on inputs given as raw float64 files, for the
// GPU tests:
//   pic_host_test fallback <grid.bin> <rho.bin> <phi_out.bin>
//       World + dirichletBoundries, PotentialSolver(world, 2, 1e-6, PCG):
//       the inner PCG stops after 2 iterations, so solve() takes the
//       reference's Gauss-Seidel fallback path (PotentialSolver.cpp:222-224)
//   pic_host_test gas <grid.bin> <parts.bin> <mass> <vel_out.bin> <T_out.bin>
//       Species + getPartRef, sampleMoments, computeGasProperties
//       (Species.cpp:767-804)
//   pic_host_test collide <grid.bin> <parts.bin> <mass> <mpw0> <dt> <steps>
//       Species + getPartRef, DSMC_MEX(species, world).apply(dt) steps times
//       (Interactions.cpp:143-217), then a ColdBeamSource on "+z" sampled
//       once (Source.cpp:32-104); prints per-apply groups and collisions and
//       the species' momentum and kinetic energy before and after
// grid.bin: ni, nj, nk, x0[3], xm[3] as doubles.  Prints one JSON line.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "pic_host.hpp"

static std::vector<double> read_all(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot read ") + path);
    std::vector<double> v;
    double x;
    while (std::fread(&x, sizeof(double), 1, f) == 1) v.push_back(x);
    std::fclose(f);
    return v;
}
static void write_all(const char* path, const double* p, size_t n) {
    FILE* f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    std::fwrite(p, sizeof(double), n, f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    std::vector<double> gr = read_all(argv[2]);
    World world((int)gr[0], (int)gr[1], (int)gr[2], type_calc3(gr[3], gr[4], gr[5]),
                type_calc3(gr[6], gr[7], gr[8]));
    world.setTime(2e-9, 10);
    if (mode == "fallback" && argc >= 5) {
        world.dirichletBoundries();
        std::vector<double> rho = read_all(argv[3]);
        world.rho.raw() = rho;
        world.rho[0][0][0] = rho[0];  // marks the host copy as written
        PotentialSolver ps(world, 2, 1e-6, PCG);
        ps.setReferenceValues(0, 0, 1e20);
        bool ok = ps.solve();
        std::vector<double>& phi = world.phi.raw();
        write_all(argv[4], phi.data(), phi.size());
        std::printf("{\"converged\": %d, \"method\": %d, \"iterations\": %d, \"ref_l2\": %.6e, "
                    "\"rel_residual\": %.6e}\n",
                    (int)ok, ps.pcg_method, ps.last.iterations, ps.last.ref_l2, ps.last.rel_residual);
        return 0;
    }
    if (mode == "gas" && argc >= 7) {
        std::vector<double> p = read_all(argv[3]);
        Species sp("s", std::atof(argv[4]), -Const::q_e, world, 1.0, -666, p.size() / 7 + 64);
        std::vector<Particle>& h = sp.getPartRef();
        h.clear();
        for (size_t n = 0; n < p.size() / 7; n++)
            h.emplace_back(type_calc3(p[7 * n], p[7 * n + 1], p[7 * n + 2]),
                           type_calc3(p[7 * n + 3], p[7 * n + 4], p[7 * n + 5]), p[7 * n + 6]);
        sp.sampleMoments();
        sp.computeGasProperties();
        std::vector<type_calc3>& v = sp.vel.raw();
        write_all(argv[5], &v[0][0], 3 * v.size());
        std::vector<double>& T = sp.T.raw();
        write_all(argv[6], T.data(), T.size());
        std::printf("{\"particles\": %zu}\n", sp.getNumParticles());
        return 0;
    }
    if (mode == "collide" && argc >= 8) {
        std::vector<double> p = read_all(argv[3]);
        const double mass = std::atof(argv[4]), mpw0 = std::atof(argv[5]), dt = std::atof(argv[6]);
        const int steps = std::atoi(argv[7]);
        world.setTime(dt, steps + 1);
        Species sp("n", mass, 0.0, world, mpw0, -666, p.size() / 7 + 200000);
        std::vector<Particle>& h = sp.getPartRef();
        h.clear();
        for (size_t n = 0; n < p.size() / 7; n++)
            h.emplace_back(type_calc3(p[7 * n], p[7 * n + 1], p[7 * n + 2]),
                           type_calc3(p[7 * n + 3], p[7 * n + 4], p[7 * n + 5]), p[7 * n + 6]);
        type_calc3 m0 = sp.getMomentum();
        const double ke0 = sp.getKE();
        DSMC_MEX dsmc(sp, world);
        dsmc.seed = 55;
        std::string groups = "[", colls = "[";
        for (int s = 0; s < steps; s++) {
            dsmc.apply(dt);
            if (dsmc.last_status != PIC_OK) return 3;
            groups += (s ? ", " : "") + std::to_string(dsmc.last.groups);
            colls += (s ? ", " : "") + std::to_string(dsmc.last.collisions);
        }
        type_calc3 m1 = sp.getMomentum();
        const double ke1 = sp.getKE();
        const size_t n_before = sp.getNumParticles();
        ColdBeamSource src(sp, world, 1e5, 2e19, "+z");
        src.seed = 77;
        src.sample();
        if (src.last_status != PIC_OK) return 4;
        std::printf("{\"groups\": %s], \"collisions\": %s], \"momentum0\": [%.17g, %.17g, %.17g], "
                    "\"momentum1\": [%.17g, %.17g, %.17g], \"ke0\": %.17g, \"ke1\": %.17g, "
                    "\"source_added\": %zu, \"particles_after_source\": %zu, \"particles_before_source\": %zu}\n",
                    groups.c_str(), colls.c_str(), m0[0], m0[1], m0[2], m1[0], m1[1], m1[2], ke0, ke1,
                    src.last_added, sp.getNumParticles(), n_before);
        return 0;
    }
    return 2;
}
