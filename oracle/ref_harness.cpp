// ref_harness.cpp -- test infrastructure only.
//
// A main() of our own, compiled against the reference's ch4/v3 sources where
// they lie (see oracle/Makefile, target "ref").  It drives the reference's
// own World / Species / PotentialSolver / MC_MEX_Ionization classes to
//   golden <dir>   write small input/output vectors (raw float64 files plus
//                  manifest.txt) that pin the CPU oracle, and
//   bench ...      time the reference's multithreaded CPU path (threaded
//                  electron push, Species.cpp:258-355) as the reported CPU
//                  baseline, and
//   bench_mcc ...  time its MC_MEX_Ionization::apply (the DSMC step).
// Nothing here is reference source; it only calls the reference's API.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sys/stat.h>
#include <unistd.h>
#include <string>
#include <vector>

#include "Config.h"
#include "Interactions.h"
#include "Outputs.h"
#include "PotentialSolver.h"
#include "Rnd.h"
#include "Source.h"
#include "Species.h"
#include "World.h"

static std::string g_dir;
static std::ofstream g_manifest;

static void dump(const std::string& name, const std::vector<double>& v) {
    std::string path = g_dir + "/" + name + ".bin";
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(2);
    }
    if (!v.empty()) fwrite(v.data(), sizeof(double), v.size(), f);
    fclose(f);
    g_manifest << name << " " << v.size() << "\n";
}

// node field -> flat i-fastest vector
static std::vector<double> flat(const Field<double>& f) {
    std::vector<double> v((size_t)f.ni * f.nj * f.nk);
    for (int i = 0; i < f.ni; i++)
        for (int j = 0; j < f.nj; j++)
            for (int k = 0; k < f.nk; k++) v[((size_t)k * f.nj + j) * f.ni + i] = f[i][j][k];
    return v;
}
static std::vector<double> flat3(const Field<type_calc3>& f) {
    std::vector<double> v((size_t)f.ni * f.nj * f.nk * 3);
    for (int i = 0; i < f.ni; i++)
        for (int j = 0; j < f.nj; j++)
            for (int k = 0; k < f.nk; k++)
                for (int c = 0; c < 3; c++)
                    v[3 * (((size_t)k * f.nj + j) * f.ni + i) + c] = f[i][j][k][c];
    return v;
}
static std::vector<double> parts(const Species& sp) {
    const std::vector<Particle>& p = sp.getConstPartRef();
    std::vector<double> v(p.size() * 7);
    for (size_t n = 0; n < p.size(); n++) {
        for (int c = 0; c < 3; c++) {
            v[7 * n + c] = p[n].pos[c];
            v[7 * n + 3 + c] = p[n].vel[c];
        }
        v[7 * n + 6] = p[n].macro_weight;
    }
    return v;
}
static void set_parts(Species& sp, const std::vector<double>& v) {
    std::vector<Particle>& p = sp.getPartRef();
    p.clear();
    for (size_t n = 0; n < v.size() / 7; n++)
        p.emplace_back(v[7 * n], v[7 * n + 1], v[7 * n + 2], v[7 * n + 3], v[7 * n + 4],
                       v[7 * n + 5], v[7 * n + 6]);
}

// exposes the protected running maximum of W*sigma*v_rel
struct MccProbe : public MC_MEX_Ionization {
    using MC_MEX_Ionization::MC_MEX_Ionization;
    double wsv() const { return W_sigma_v_rel_max; }
};

static void golden(const std::string& dir, const std::string& table_path) {
    g_dir = dir;
    g_manifest.open(dir + "/manifest.txt");
    Config& cfg = Config::getInstance();
    cfg.setMULTITHREADING(false);  // deterministic FP64 order
    cfg.setSPUTTERING(false);

    // ---- grid with distinct extents per axis (catches transposed axes) ----
    const int NI = 9, NJ = 8, NK = 7;
    type_calc3 x0(-0.01, 0.02, 0.0), xm(0.07, 0.104, 0.054);
    World world(NI, NJ, NK, x0, xm);
    world.setTime(2e-9, 10);
    dump("grid", {(double)NI, (double)NJ, (double)NK, x0[0], x0[1], x0[2], xm[0], xm[1], xm[2],
                  2e-9});
    dump("node_vol", flat(world.node_vol));

    // ---- analytic potential -> E (computeEF) ----
    for (int i = 0; i < NI; i++)
        for (int j = 0; j < NJ; j++)
            for (int k = 0; k < NK; k++) {
                type_calc3 x = world.LtoX(i, j, k);
                world.phi[i][j][k] = 40.0 * std::sin(31.0 * x[0] + 0.3) * std::cos(23.0 * x[1]) +
                                     900.0 * x[2] * x[2] - 150.0 * x[0] * x[1] + 12.0 * x[2];
            }
    dump("phi_analytic", flat(world.phi));
    {
        PotentialSolver ps(world, 10, 1e-3, GS);
        ps.computeEF();
    }
    dump("ef_analytic", flat3(world.ef));

    // ---- loader (mt19937 stream, seeded) ----
    rnd = Rnd(12345u);
    Species ele("e-", Const::m_e, -Const::q_e, world, 1.5e5);
    Species ion("O+", 16 * Const::amu, Const::q_e, world, 2e5);
    type_calc3 xc = world.getXc(), L = world.getL();
    ele.loadParticleBoxThermal(xc, type_calc3(L[0] * 0.98, L[1] * 0.98, L[2] * 0.98), 1e12,
                               2.0 * 11604.52);
    dump("ele_loaded", parts(ele));
    rnd = Rnd(777u);
    ion.loadParticleBoxThermal(xc, type_calc3(L[0] * 0.9, L[1] * 0.9, L[2] * 0.9), 1e12, 0.1 * 11604.52);
    dump("ion_loaded", parts(ion));

    // ---- push (serial electron push, serial non-electron push) ----
    double dt = 2e-9;
    ele.advance_time_multi = 1e30;  // force the serial path
    ele.advance_time_serial = 0;
    ele.advanceElectrons(dt);
    dump("ele_pushed", parts(ele));
    ion.advanceNonElectron(ion, ion, dt);
    dump("ion_pushed", parts(ion));
    // a second, long electron step so that some electrons leave the box
    std::vector<double> before = parts(ele);
    dump("ele_pre_long", before);
    ele.advanceElectrons(4e-8);
    dump("ele_post_long", parts(ele));

    // ---- density, moments, counts, rho ----
    ele.computeNumberDensity();
    ion.computeNumberDensity();
    dump("ele_den", flat(ele.den));
    dump("ion_den", flat(ion.den));
    ele.sampleMoments();
    ele.computeGasProperties();
    dump("ele_vel", flat3(ele.vel));
    dump("ele_T", flat(ele.T));
    ele.computeMacroParticlesCount();
    {
        std::vector<double> mc((size_t)(NI - 1) * (NJ - 1) * (NK - 1));
        for (int i = 0; i < NI - 1; i++)
            for (int j = 0; j < NJ - 1; j++)
                for (int k = 0; k < NK - 1; k++)
                    mc[((size_t)k * (NJ - 1) + j) * (NI - 1) + i] = ele.macro_part_count[i][j][k];
        dump("ele_macro_count", mc);
    }
    std::vector<Species> sps;
    sps.push_back(std::move(ele));
    sps.push_back(std::move(ion));
    world.computeChargeDensity(sps);
    dump("rho", flat(world.rho));

    // ---- diagnostics (phi_analytic is still world.phi) ----
    {
        std::vector<double> d;
        for (Species& sp : sps) {
            type_calc3 m = sp.getMomentum();
            d.push_back(sp.getMicroCount());
            d.push_back(m[0]);
            d.push_back(m[1]);
            d.push_back(m[2]);
            d.push_back(sp.getKE());
            d.push_back(sp.getPE());
        }
        d.push_back(world.getPE());
        dump("diag", d);
    }

    // ---- v3 GS-SOR with Neumann faces and one rectangle object ----
    {
        World w2(NI, NJ, NK, x0, xm);
        w2.addObject<Rectangle>(type_calc3(0.03, 0.06, 0.027), 5.0, type_calc3(0.02, 0.02, 0.018));
        w2.computeObjectID();
        w2.rho = world.rho;
        dump("gs_object_id", [&] {
            std::vector<double> v((size_t)NI * NJ * NK);
            for (int i = 0; i < NI; i++)
                for (int j = 0; j < NJ; j++)
                    for (int k = 0; k < NK; k++) v[((size_t)k * NJ + j) * NI + i] = w2.object_id[i][j][k];
            return v;
        }());
        dump("gs_phi_in", flat(w2.phi));
        PotentialSolver gs(w2, 4000, 1e-4, GS);
        gs.setReferenceValues(0, 0, 1e20);
        bool ok = gs.solve();
        dump("gs_phi_out", flat(w2.phi));
        dump("gs_ok", {(double)ok});
    }
    // ---- PCG (Newton) with Dirichlet walls, linear (n0 = 0) ----
    {
        World w3(NI, NJ, NK, x0, xm);
        w3.dirichletBoundries();
        w3.rho = world.rho;
        PotentialSolver pcg(w3, 500, 1e-6, PCG);
        pcg.setReferenceValues(0, 0, 1e20);
        bool ok = pcg.solve();
        dump("pcg_phi_out", flat(w3.phi));
        dump("pcg_ok", {(double)ok});
        pcg.computeEF();
        dump("pcg_ef", flat3(w3.ef));
    }
    // ---- PCG (Newton) with Dirichlet walls and Boltzmann electrons ----
    {
        World w4(NI, NJ, NK, x0, xm);
        w4.dirichletBoundries();
        for (int i = 0; i < NI; i++)
            for (int j = 0; j < NJ; j++)
                for (int k = 0; k < NK; k++) w4.rho[i][j][k] = Const::q_e * 1.5e12;
        PotentialSolver nr(w4, 500, 1e-8, PCG);
        nr.setReferenceValues(0.0, 1.0e12, 3.0);
        bool ok = nr.solve();
        dump("nr_phi_out", flat(w4.phi));
        dump("nr_ok", {(double)ok});
    }

    // ---- MC_MEX_Ionization on a small, dense box ----
    {
        type_calc3 y0(0.0, 0.0, 0.0), ym(8e-4, 7e-4, 6e-4);
        World wd(9, 8, 7, y0, ym);
        double ddt = 1e-10;
        wd.setTime(ddt, 10);
        double E_ion_O = 1313.9 * 1000 / Const::N_a;
        Species neu("O", 16 * Const::amu, 0, wd, 5e11, E_ion_O);
        Species oi("O+", 16 * Const::amu, Const::q_e, wd, 100);
        Species el("e-", Const::m_e, -Const::q_e, wd, 100);
        rnd = Rnd(4242u);
        type_calc3 c = wd.getXc(), Ld = wd.getL();
        // 336 cells; ~30 neutrals and ~10 electrons per cell
        type_calc3 box = Ld * 0.999;
        double volb = box[0] * box[1] * box[2];
        neu.loadParticleBoxThermal(c, box, 336.0 * 30 * 5e11 / volb, 300.0);
        el.loadParticleBoxThermal(c, box, 336.0 * 10 * 100 / volb, 60.0 * 11604.52);
        dump("mcc_grid", {9, 8, 7, y0[0], y0[1], y0[2], ym[0], ym[1], ym[2], ddt});
        dump("mcc_neu_in", parts(neu));
        dump("mcc_ele_in", parts(el));
        MccProbe mcc(neu, oi, el, wd, table_path);
        dump("mcc_wsv0", {mcc.wsv()});
        rnd = Rnd(99u);
        for (int call = 0; call < 3; call++) {
            mcc.apply(ddt);
            std::string s = std::to_string(call);
            dump("mcc_neu_" + s, parts(neu));
            dump("mcc_ion_" + s, parts(oi));
            dump("mcc_ele_" + s, parts(el));
            dump("mcc_wsv_" + s, {mcc.wsv()});
        }
    }

    // ---- threaded electron push (ThreadPool path, Species.cpp:263-310) ----
    {
        cfg.setMULTITHREADING(true);
        cfg.setNUM_THREADS(3);
        World wt(NI, NJ, NK, x0, xm);
        wt.setTime(2e-9, 10);
        wt.ef = world.ef;
        Species et("e-", Const::m_e, -Const::q_e, wt, 1.5e5);
        set_parts(et, before);
        et.advance_time_multi = -1;  // force the threaded path
        et.advance_time_serial = 1e30;
        et.advanceElectrons(4e-8);
        dump("thr_post_long", parts(et));
        cfg.setMULTITHREADING(false);
    }

    // ---- object surfaces: electrons absorbed (Species.cpp:356-399), ions
    // recycled into neutrals and neutrals reflected (Species.cpp:170-256) ----
    {
        World ws(NI, NJ, NK, x0, xm);
        type_calc3 c = ws.getXc(), L = ws.getL();
        // plates at both z ends (as main.cpp:94-98) and a block in the middle
        // (x and y faces, edges and corners)
        std::vector<std::pair<type_calc3, type_calc3>> rects = {
            {type_calc3(c[0], c[1], ws.getX0()[2]), type_calc3(L[0], L[1], L[2] * 0.2)},
            {type_calc3(c[0], c[1], ws.getXm()[2]), type_calc3(L[0], L[1], L[2] * 0.2)},
            {type_calc3(c[0] + 0.1 * L[0], c[1] - 0.05 * L[1], c[2]),
             type_calc3(L[0] * 0.3, L[1] * 0.25, L[2] * 0.2)}};
        std::vector<double> rd;
        for (auto& r : rects) {
            ws.addObject<Rectangle>(r.first, 0.0, r.second);
            for (int d = 0; d < 3; d++) rd.push_back(r.first[d]);
            for (int d = 0; d < 3; d++) rd.push_back(r.second[d] * 0.5);
        }
        ws.computeObjectID();
        dump("sf_rects", rd);
        for (int i = 0; i < NI; i++)
            for (int j = 0; j < NJ; j++)
                for (int k = 0; k < NK; k++) ws.ef[i][j][k] = world.ef[i][j][k] * 1e-3;
        dump("sf_ef", flat3(ws.ef));
        const double dt_e = 2e-9, dt_i = 5e-6, dt_n = 4e-5;
        dump("sf_dt", {dt_e, dt_i, dt_n});
        ws.setTime(dt_e, 10);
        rnd = Rnd(555u);
        Species sn("O", 16 * Const::amu, 0, ws, 1e10);
        Species si("O+", 16 * Const::amu, Const::q_e, ws, 2.5e10);
        Species se("e-", Const::m_e, -Const::q_e, ws, 2e4);
        sn.loadParticleBoxThermal(c, L * 0.99, 1e10 * 3000 / (L[0] * L[1] * L[2]), 300.0);
        si.loadParticleBoxThermal(c, L * 0.99, 2.5e10 * 2000 / (L[0] * L[1] * L[2]), 1e4);
        se.loadParticleBoxThermal(c, L * 0.99, 2e4 * 2000 / (L[0] * L[1] * L[2]), 2.0 * 11604.52);
        dump("sf_neu_in", parts(sn));
        dump("sf_ion_in", parts(si));
        dump("sf_ele_in", parts(se));
        se.advance_time_multi = 1e30;  // serial path
        se.advance_time_serial = 0;
        se.advanceElectrons(dt_e);
        dump("sf_ele_out", parts(se));
        rnd = Rnd(556u);
        si.advanceNonElectron(sn, sn, dt_i);
        dump("sf_ion_out", parts(si));
        dump("sf_neu_mid", parts(sn));
        rnd = Rnd(557u);
        sn.advanceNonElectron(sn, sn, dt_n);
        dump("sf_neu_out", parts(sn));
    }

    // ---- Newton-PCG whose PCG gives up: the GS fallback (:222-224, :315-347) ----
    {
        World w5(NI, NJ, NK, x0, xm);
        w5.dirichletBoundries();
        w5.rho = world.rho;
        PotentialSolver fb(w5, 2, 1e-6, PCG);
        fb.setReferenceValues(0, 0, 1e20);
        bool ok = fb.solve();
        dump("fb_phi_out", flat(w5.phi));
        dump("fb_ok", {(double)ok});
    }

    // ---- Species::merge (Species.cpp:987-1145) on a neutral and an
    // electron species: a few crowded cells, velocities from a small set
    // (elastic-split copies) plus scattered ones ----
    {
        World wm(NI, NJ, NK, x0, xm);
        wm.setTime(2e-9, 10);
        wm.ef = world.ef;
        std::mt19937 gen(2718u);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        type_calc3 dxm = wm.getDx();
        auto make = [&](double vs, double w0) {
            std::vector<double> v;
            // particle 0 first: a far-away cell with an extreme velocity
            // (the source starts every cell's velocity range from it)
            v.insert(v.end(), {x0[0] + 0.5 * dxm[0], x0[1] + 0.5 * dxm[1], x0[2] + 0.5 * dxm[2],
                               2.5 * vs, -1.7 * vs, 0.3 * vs, w0});
            const int cells[5][3] = {{3, 2, 2}, {4, 4, 3}, {1, 5, 1}, {6, 1, 4}, {2, 3, 5}};
            for (int ci = 0; ci < 5; ci++) {
                double base[4][3];
                for (auto& b : base)
                    for (double& x : b) x = vs * (2 * U(gen) - 1);
                int np = 14 + 12 * ci;
                for (int p = 0; p < np; p++) {
                    double pos[3];
                    for (int d = 0; d < 3; d++) pos[d] = x0[d] + (cells[ci][d] + 0.02 + 0.96 * U(gen)) * dxm[d];
                    double vel[3];
                    int pick = p % 6;
                    for (int d = 0; d < 3; d++)
                        vel[d] = pick < 4 ? base[pick][d] * (p % 5 == 4 ? 1.0 + 1e-4 * U(gen) : 1.0)
                                          : vs * (2 * U(gen) - 1);
                    double w = w0 * (0.5 + U(gen));
                    v.insert(v.end(), {pos[0], pos[1], pos[2], vel[0], vel[1], vel[2], w});
                }
            }
            return v;
        };
        std::vector<double> nin = make(600.0, 1e10), ein = make(8e5, 1e4);
        Species mn("O", 16 * Const::amu, 0, wm, 1e10);
        Species me("e-", Const::m_e, -Const::q_e, wm, 1e4);
        set_parts(mn, nin);
        set_parts(me, ein);
        dump("mg_neu_in", nin);
        dump("mg_ele_in", ein);
        rnd = Rnd(31415u);
        mn.merge();
        dump("mg_neu_out", parts(mn));
        rnd = Rnd(27182u);
        me.merge();
        dump("mg_ele_out", parts(me));
    }
    // ---- DSMC_MEX (Interactions.cpp:143-284): one species and two species
    // of equal weight on a dense box ----
    {
        type_calc3 y0(0.0, 0.0, 0.0), ym(8e-4, 7e-4, 6e-4);
        World wd(9, 8, 7, y0, ym);
        const double ddt = 1e-7;
        wd.setTime(ddt, 10);
        dump("dm_grid", {9, 8, 7, y0[0], y0[1], y0[2], ym[0], ym[1], ym[2], ddt});
        Species a("O", 16 * Const::amu, 0, wd, 1e9);
        Species b("Ar", 40 * Const::amu, 0, wd, 1e9);
        rnd = Rnd(8080u);
        type_calc3 c = wd.getXc(), box = wd.getL() * 0.999;
        const double vol = box[0] * box[1] * box[2];
        a.loadParticleBoxThermal(c, box, 336.0 * 12 * 1e9 / vol, 3000.0);
        b.loadParticleBoxThermal(c, box, 336.0 * 6 * 1e9 / vol, 2000.0);
        dump("dm_a_in", parts(a));
        dump("dm_b_in", parts(b));
        DSMC_MEX one(a, wd);
        DSMC_MEX two(a, b, wd);
        rnd = Rnd(8081u);
        for (int call = 0; call < 2; call++) {
            one.apply(ddt);
            dump("dm_one_" + std::to_string(call), parts(a));
        }
        rnd = Rnd(8082u);
        for (int call = 0; call < 2; call++) {
            two.apply(ddt);
            dump("dm_two_a_" + std::to_string(call), parts(a));
            dump("dm_two_b_" + std::to_string(call), parts(b));
        }
    }

    // ---- sources (Source.cpp:3-185): a cold beam through x- and a warm
    // beam through z+, with a block in front of part of the z+ face ----
    {
        World wsrc(NI, NJ, NK, x0, xm);
        wsrc.setTime(2e-9, 10);
        type_calc3 c = wsrc.getXc(), L = wsrc.getL();
        wsrc.addObject<Rectangle>(type_calc3(c[0], c[1], wsrc.getXm()[2]), 0.0,
                                  type_calc3(L[0] * 0.4, L[1] * 0.3, L[2] * 0.2));
        wsrc.computeObjectID();
        for (int i = 0; i < NI; i++)
            for (int j = 0; j < NJ; j++)
                for (int k = 0; k < NK; k++) wsrc.ef[i][j][k] = world.ef[i][j][k];
        Species cold("e-", Const::m_e, -Const::q_e, wsrc, 1e4);
        Species warm("O+", 16 * Const::amu, Const::q_e, wsrc, 5e3);
        ColdBeamSource s1(cold, wsrc, 2e5, 1e12, "x-");
        WarmBeamSource s2(warm, wsrc, 3e3, 2e12, 5000.0, "z+");
        rnd = Rnd(1234u);
        s1.sample();
        s1.sample();
        rnd = Rnd(4321u);
        s2.sample();
        dump("src_cold", parts(cold));
        dump("src_warm", parts(warm));
    }

    // ---- Output writers (Outputs.cpp:9-205): the reference writes its own
    // files for a small world into <dir>/results; the inputs they print are
    // dumped beside them ----
    {
        char cwd[4096];
        if (!getcwd(cwd, sizeof(cwd))) exit(2);
        if (chdir(dir.c_str()) != 0) exit(2);
        mkdir("results", 0755);
        const int OI = 5, OJ = 4, OK = 3;
        type_calc3 y0(-0.01, 0.0, 0.02), ym(0.03, 0.045, 0.05);
        World wo(OI, OJ, OK, y0, ym);
        wo.setTime(1e-9, 10);
        for (int t = 0; t < 3; t++) wo.advanceTime();
        wo.addObject<Rectangle>(type_calc3(0.01, 0.0225, 0.035), 2.0, type_calc3(0.012, 0.016, 0.014));
        wo.computeObjectID();
        wo.dirichletBoundries();
        for (int i = 0; i < OI; i++)
            for (int j = 0; j < OJ; j++)
                for (int k = 0; k < OK; k++) {
                    type_calc3 x = wo.LtoX(i, j, k);
                    wo.phi[i][j][k] = 3.0 * std::sin(80.0 * x[0]) + 1e-7 * j - 0.25 * k;
                    wo.rho[i][j][k] = 1e-9 * (i - 2.0 * j + 0.5 * k) / 3.0;
                }
        {
            PotentialSolver ps(wo, 10, 1e-3, GS);
            ps.computeEF();
        }
        std::mt19937 gen(99u);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        type_calc3 L = wo.getL();
        auto cloud = [&](int n, double vs, double w) {
            std::vector<double> v;
            for (int q = 0; q < n; q++) {
                for (int d = 0; d < 3; d++) v.push_back(y0[d] + (0.01 + 0.98 * U(gen)) * L[d]);
                for (int d = 0; d < 3; d++) v.push_back(vs * (2 * U(gen) - 1));
                v.push_back(w * (0.5 + U(gen)));
            }
            return v;
        };
        std::vector<Species> sps;
        sps.reserve(2);
        sps.emplace_back("e-", Const::m_e, -Const::q_e, wo, 10.0);
        sps.emplace_back("O+", 16 * Const::amu, Const::q_e, wo, 20.0);
        set_parts(sps[0], cloud(30, 1e6, 10.0));
        set_parts(sps[1], cloud(41, 2e3, 20.0));
        for (Species& sp : sps) {
            sp.computeNumberDensity();
            sp.sampleMoments();
            sp.updateAverages();
            sp.computeNumberDensity();
            sp.updateAverages();
            sp.computeMacroParticlesCount();
        }
        Output::fieldsOutput(wo, sps);
        Output::particlesOutput(wo, sps, 7);
        std::vector<double> drows;
        for (int row = 0; row < 2; row++) {
            Output::diagOutput(wo, sps);
            drows.push_back(wo.getTs());
            drows.push_back(wo.getTime());
            for (Species& sp : sps) {
                type_calc3 m = sp.getMomentum();
                drows.insert(drows.end(), {(double)sp.getNumParticles(), sp.getMicroCount(), m[0], m[1], m[2],
                                           sp.getKE()});
            }
            drows.push_back(wo.getPE());
            wo.advanceTime();
        }
        if (chdir(cwd) != 0) exit(2);
        dump("out_grid", {(double)OI, (double)OJ, (double)OK, y0[0], y0[1], y0[2], wo.getDx()[0], wo.getDx()[1],
                          wo.getDx()[2], 2.0});
        dump("out_node_vol", flat(wo.node_vol));
        std::vector<double> oid, nty;
        for (int k = 0; k < OK; k++)
            for (int j = 0; j < OJ; j++)
                for (int i = 0; i < OI; i++) {
                    oid.push_back(wo.object_id[i][j][k]);
                    nty.push_back(wo.node_type[i][j][k]);
                }
        dump("out_object_id", oid);
        dump("out_node_type", nty);
        dump("out_phi", flat(wo.phi));
        dump("out_rho", flat(wo.rho));
        dump("out_ef", flat3(wo.ef));
        for (int s = 0; s < 2; s++) {
            Species& sp = sps[s];
            std::string t = std::to_string(s);
            dump("out_den_" + t, flat(sp.den));
            dump("out_den_avg_" + t, flat(sp.den_avg));
            dump("out_vel_" + t, flat3(sp.vel));
            dump("out_T_" + t, flat(sp.T));
            std::vector<double> mc;
            for (int k = 0; k < OK - 1; k++)
                for (int j = 0; j < OJ - 1; j++)
                    for (int i = 0; i < OI - 1; i++) mc.push_back(sp.macro_part_count[i][j][k]);
            dump("out_mpc_" + t, mc);
            dump("out_parts_" + t, parts(sp));
        }
        dump("out_diag_rows", drows);
    }
    g_manifest.close();
}

// reported CPU baseline: the reference's threaded electron push on a
// uniform box (and, optionally, its PCG solve), medians over steps.
static void bench(int n, long np, int steps, int threads, int do_solve) {
    Config& cfg = Config::getInstance();
    cfg.setMULTITHREADING(true);
    cfg.setNUM_THREADS(threads);
    type_calc3 x0(0, 0, 0);
    double dx = 1e-2;
    type_calc3 xm(dx * (n - 1), dx * (n - 1), dx * (n - 1));
    World world(n, n, n, x0, xm);
    double dt = 1e-10;
    world.setTime(dt, steps);
    world.dirichletBoundries();
    Species ele("e-", Const::m_e, -Const::q_e, world, 1.0);
    rnd = Rnd(2024u);
    type_calc3 L = world.getL();
    double vol = L[0] * L[1] * L[2];
    // 0.98 L: a particle within rounding of xm gets lc == n-1 and the
    // reference gather (Field.h:204-221) would read past the last node
    auto l0 = std::chrono::steady_clock::now();
    ele.loadParticleBoxThermal(world.getXc(), L * 0.98, (double)np / (vol * 0.98 * 0.98 * 0.98),
                               2.0 * 11604.52);
    auto l1 = std::chrono::steady_clock::now();
    const double t_load = std::chrono::duration<double>(l1 - l0).count() / (double)ele.getNumParticles();
    // a smooth field so the gather does real work
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++) {
                world.ef[i][j][k][0] = 10.0 * std::sin(0.1 * i);
                world.ef[i][j][k][1] = 5.0 * std::cos(0.07 * j);
                world.ef[i][j][k][2] = 1.0;
            }
    std::vector<double> t_push, t_den, t_solve;
    PotentialSolver* ps = nullptr;
    if (do_solve) ps = new PotentialSolver(world, 20000, 1e-6, PCG);
    for (int s = 0; s < steps; s++) {
        size_t nbefore = ele.getNumParticles();
        ele.advance_time_multi = -1;
        ele.advance_time_serial = 1e30;
        auto a = std::chrono::steady_clock::now();
        ele.advanceElectrons(dt);
        auto b = std::chrono::steady_clock::now();
        ele.computeNumberDensity();
        auto c = std::chrono::steady_clock::now();
        t_push.push_back(std::chrono::duration<double>(b - a).count() / (double)nbefore);
        t_den.push_back(std::chrono::duration<double>(c - b).count() / (double)nbefore);
        if (ps) {
            world.rho = ele.den;
            world.rho *= -Const::q_e;
            auto d = std::chrono::steady_clock::now();
            ps->solve();
            auto e = std::chrono::steady_clock::now();
            t_solve.push_back(std::chrono::duration<double>(e - d).count());
        }
    }
    auto med = [](std::vector<double> v) {
        if (v.empty()) return 0.0;
        std::sort(v.begin(), v.end());
        return v[v.size() / 2];
    };
    // moments (n, nv, nuu, nvv, nww) and the diagnostic sums on the final state
    const double npart = (double)ele.getNumParticles();
    auto m0 = std::chrono::steady_clock::now();
    ele.sampleMoments();
    auto m1 = std::chrono::steady_clock::now();
    volatile double sink = ele.getMicroCount() + ele.getKE() + ele.getMomentum()[0];
    (void)sink;
    auto m2 = std::chrono::steady_clock::now();
    printf("{\"n\": %d, \"particles\": %zu, \"steps\": %d, \"threads\": %d, "
           "\"push_s_per_particle\": %.6e, \"deposit_s_per_particle\": %.6e, "
           "\"pushes_per_s\": %.6e, \"solve_s\": %.6e, \"load_s_per_particle\": %.6e, "
           "\"moments_s_per_particle\": %.6e, \"diag_s_per_particle\": %.6e}\n",
           n, ele.getNumParticles(), steps, threads, med(t_push), med(t_den), 1.0 / med(t_push),
           med(t_solve), t_load, std::chrono::duration<double>(m1 - m0).count() / npart,
           std::chrono::duration<double>(m2 - m1).count() / npart);
    delete ps;
}

// reported CPU baseline of the DSMC step: the reference's
// MC_MEX_Ionization::apply (its per-call cell sort and the serial NTC pair
// loop, Interactions.cpp:567-762) on an O2 / electron box of the discharge
// configuration's cell densities; medians over calls.
static void bench_mcc(int nxy, int nz, long n_neu, long n_ele, int steps, double dt,
                      const std::string& table_path) {
    Config& cfg = Config::getInstance();
    cfg.setMULTITHREADING(false);
    const double dx = 1e-4;
    type_calc3 x0(0, 0, 0), xm(dx * (nxy - 1), dx * (nxy - 1), dx * (nz - 1));
    World world(nxy, nxy, nz, x0, xm);
    world.setTime(dt, steps);
    type_calc3 L = world.getL(), box = L * 0.98;
    const double vol = box[0] * box[1] * box[2];
    const double n_gas = 10.0 / (Const::k * 300.0);  // O2 at 10 Pa, 300 K
    const double w_n = n_gas * vol / (double)n_neu, w_e = w_n / 1000.0;
    Species neu("O2", 32 * Const::amu, 0, world, w_n, 12.07 * Const::q_e);
    Species ion("O2+", 32 * Const::amu, Const::q_e, world, w_e);
    Species ele("e-", Const::m_e, -Const::q_e, world, w_e);
    rnd = Rnd(31u);
    neu.loadParticleBoxThermal(world.getXc(), box, n_gas, 300.0);
    ele.loadParticleBoxThermal(world.getXc(), box, (double)n_ele * w_e / vol, 10.0 * 11604.52);
    MccProbe mcc(neu, ion, ele, world, table_path);
    // per call: the neutrals and electrons it starts with over its time
    // (elastic splits grow the neutral count from call to call)
    std::vector<double> t, rate;
    for (int s = 0; s < steps; s++) {
        const double np = (double)(neu.getNumParticles() + ele.getNumParticles());
        auto a = std::chrono::steady_clock::now();
        mcc.apply(dt);
        auto b = std::chrono::steady_clock::now();
        t.push_back(std::chrono::duration<double>(b - a).count());
        rate.push_back(np / t.back());
    }
    auto med = [](std::vector<double> v) {
        std::sort(v.begin(), v.end());
        return v[v.size() / 2];
    };
    printf("{\"cells\": %d, \"neutrals\": %zu, \"electrons\": %zu, \"ions\": %zu, \"steps\": %d, "
           "\"apply_s\": %.6e, \"particles_per_s\": %.6e, \"wsv_max\": %.6e}\n",
           (nxy - 1) * (nxy - 1) * (nz - 1), neu.getNumParticles(), ele.getNumParticles(),
           ion.getNumParticles(), steps, med(t), med(rate), mcc.wsv());
}

// reported CPU baseline of the Newton-Boltzmann solve (PotentialSolver::
// solveNRPCG, PotentialSolver.cpp:178-241): n^3 box with Dirichlet walls, a
// uniform ion charge density n0 q (1.5 n0), Boltzmann electrons n0, Te0 = 2 eV
static void bench_newton(int n, int threads, double tol) {
    Config& cfg = Config::getInstance();
    cfg.setMULTITHREADING(threads > 1);
    cfg.setNUM_THREADS(threads);
    const double dx = 1e-2;
    World world(n, n, n, type_calc3(0, 0, 0), type_calc3(dx * (n - 1), dx * (n - 1), dx * (n - 1)));
    world.setTime(1e-10, 1);
    world.dirichletBoundries();
    const double n0 = 1e12;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++) world.rho[i][j][k] = Const::q_e * 1.5 * n0;
    PotentialSolver ps(world, 20000, tol, PCG);
    ps.setReferenceValues(0.0, n0, 2.0);
    auto a = std::chrono::steady_clock::now();
    bool ok = ps.solve();
    auto b = std::chrono::steady_clock::now();
    printf("{\"n\": %d, \"threads\": %d, \"tol\": %.1e, \"converged\": %d, \"solve_s\": %.6e}\n", n,
           threads, tol, (int)ok, std::chrono::duration<double>(b - a).count());
}

int main(int argc, char** argv) {
    if (argc >= 4 && std::string(argv[1]) == "golden") {
        golden(argv[2], argv[3]);
        return 0;
    }
    if (argc >= 7 && std::string(argv[1]) == "bench") {
        bench(atoi(argv[2]), atol(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6]));
        return 0;
    }
    if (argc >= 5 && std::string(argv[1]) == "bench_newton") {
        bench_newton(atoi(argv[2]), atoi(argv[3]), atof(argv[4]));
        return 0;
    }
    if (argc >= 9 && std::string(argv[1]) == "bench_mcc") {
        bench_mcc(atoi(argv[2]), atoi(argv[3]), atol(argv[4]), atol(argv[5]), atoi(argv[6]), atof(argv[7]),
                  argv[8]);
        return 0;
    }
    fprintf(stderr,
            "usage: ref_harness golden <outdir> <sigma_el_table>\n"
            "       ref_harness bench <nodes_per_axis> <particles> <steps> <threads> <solve0|1>\n"
            "       ref_harness bench_mcc <nodes_xy> <nodes_z> <neutrals> <electrons> <steps> <dt> "
            "<sigma_el_table>\n"
            "       ref_harness bench_newton <nodes_per_axis> <threads> <tol>\n");
    return 1;
}
