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
    return 2;
}
