// C++ parity test of the host mirror (pis_amd/host/pis.hpp) over the C ABI, written the
// way the reference's own code uses its manager.  Needs a GPU; run by
// tests/test_gpu_cpp.py.  Checks:
//   1. LJGpuManager::compute_potential vs the CPU oracle's LJVOffsetManager restatement:
//      forces within 1e-10 normwise, PE within the serial-sum bound;
//   2. the device-resident verlet_step_nve vs the trait's default host-side integrator
//      (which calls compute_potential each step) over 10 steps;
//   3. Simulation::run prints the reference's thermo lines;
//   4. error behaviour: empty table -> MissingPotential, no atoms -> NoAtomsDefined;
//   5. Simulation::run with NVT tracks the oracle's orc_run_nvt step by step;
//   6. Simulation::run with NPT tracks orc_run_npt (thermo and box), moves atoms.sim_box,
//      and writes trajectory blocks at steps 0, 10, 20, 30 in the reference's format.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>

#include "../../oracle/pis_oracle.h"
#include "../../pis_amd/host/pis.hpp"

static int failures = 0;
#define EXPECT(cond, ...)                            \
    do {                                             \
        if (!(cond)) {                               \
            std::fprintf(stderr, "FAIL: " __VA_ARGS__); \
            std::fprintf(stderr, "\n");              \
            ++failures;                              \
        }                                            \
    } while (0)

static pis::Atoms make_liquid(int m, double density, unsigned seed)
{
    const double a = std::cbrt(4.0 / density), L = m * a;
    const double basis[4][3] = {{0, 0, 0}, {0, .5, .5}, {.5, 0, .5}, {.5, .5, 0}};
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> jit(-0.05, 0.05);
    std::normal_distribution<double> gauss(0.0, 0.85);
    pis::Atoms at;
    for (int ix = 0; ix < m; ++ix)
        for (int iy = 0; iy < m; ++iy)
            for (int iz = 0; iz < m; ++iz)
                for (auto& b : basis) {
                    const double p[3] = {a * (ix + b[0]) + jit(rng), a * (iy + b[1]) + jit(rng), a * (iz + b[2]) + jit(rng)};
                    for (double q : p) at.positions.push_back(q - L * std::floor(q / L));
                    for (int d = 0; d < 3; ++d) at.velocities.push_back(gauss(rng));
                }
    at.n_atoms = at.positions.size() / 3;
    at.type_ids.assign(at.n_atoms, 1);
    at.masses = {1.0};
    at.forces.assign(3 * at.n_atoms, 0.0);
    at.sim_box = pis::SimulationBox::from_lammps_data(0, L, 0, L, 0, L, 0, 0, 0);
    return at;
}

// The oracle's view of a single-type system (eps = sigma = 1, rcut 2.5, shift on).
struct OrcView {
    std::vector<double> pos, vel, force;
    std::vector<int32_t> types;
    int32_t present = 1;
    double eps = 1.0, sig = 1.0, rc = 2.5, mass = 1.0;
    orc_sys s{};
    explicit OrcView(const pis::Atoms& a)
        : pos(a.positions), vel(a.velocities), force(3 * a.n_atoms, 0.0), types(a.n_atoms, 1)
    {
        s.n = (int64_t)a.n_atoms;
        s.pos = pos.data();
        s.vel = vel.data();
        s.force = force.data();
        s.type = types.data();
        s.mass = &mass;
        std::memcpy(s.box.h, a.sim_box.h.data(), sizeof(s.box.h));
        std::memcpy(s.box.hinv, a.sim_box.h_inv.data(), sizeof(s.box.hinv));
        s.box.pbc[0] = s.box.pbc[1] = s.box.pbc[2] = 1;
        s.table.ntypes = 1;
        s.table.present = &present;
        s.table.eps = &eps;
        s.table.sigma = &sig;
        s.table.rcut = &rc;
        s.table.shift = 1;
    }
};

static void expect_tracks(const char* what, const std::vector<pis_thermo>& th, const std::vector<double>& ref,
                          double tol)
{
    double scale = 0.0, err = 0.0;
    for (size_t t = 0; t < th.size(); ++t) {
        scale = std::max(scale, std::fabs(ref[5 * t]));
        const double got[3] = {th[t].pe, th[t].ke, th[t].etot};
        for (int c = 0; c < 3; ++c) err = std::max(err, std::fabs(got[c] - ref[5 * t + c]));
    }
    std::printf("%s: %zu steps, max thermo error %.3e of %.3e\n", what, th.size(), err, scale);
    EXPECT(!th.empty() && err <= tol * scale, "%s thermo drifts from the oracle", what);
}

int main()
{
    pis::Atoms atoms = make_liquid(8, 0.8442, 42);
    const size_t n = atoms.n_atoms;
    pis::LJGpuManager mgr(0.3, 1.0, PIS_PRECISION_FP64);
    mgr.insert({1, 1}, pis::LennardJones(1.0, 1.0, 2.5, true));
    EXPECT(mgr.max_rcut() == 2.5, "max_rcut");

    // 1. compute_potential vs the oracle
    std::vector<double> f_ref(3 * n, 0.0), pos = atoms.positions, vel = atoms.velocities;
    int32_t type1[1] = {1};
    std::vector<int32_t> types(n, 1);
    int32_t present = 1;
    double eps = 1.0, sig = 1.0, rc = 2.5, mass = 1.0;
    orc_sys s{};
    s.n = (int64_t)n;
    s.pos = pos.data();
    s.vel = vel.data();
    s.force = f_ref.data();
    s.type = types.data();
    s.mass = &mass;
    std::memcpy(s.box.h, atoms.sim_box.h.data(), sizeof(s.box.h));
    std::memcpy(s.box.hinv, atoms.sim_box.h_inv.data(), sizeof(s.box.hinv));
    s.box.pbc[0] = s.box.pbc[1] = s.box.pbc[2] = 1;
    s.table.ntypes = 1;
    s.table.present = &present;
    s.table.eps = &eps;
    s.table.sigma = &sig;
    s.table.rcut = &rc;
    s.table.shift = 1;
    (void)type1;
    const double pe_ref = orc_compute_offset(&s);
    const double pe = mgr.compute_potential(atoms);  // atoms.forces started at zero
    double rms = 0.0, maxd = 0.0;
    for (size_t k = 0; k < 3 * n; ++k) {
        rms += f_ref[k] * f_ref[k];
        maxd = std::max(maxd, std::fabs(atoms.forces[k] - f_ref[k]));
    }
    rms = std::sqrt(rms / (double)n);
    std::printf("compute_potential: PE gpu %.15g oracle %.15g, max|dF|/rms %.3e\n", pe, pe_ref, maxd / rms);
    EXPECT(maxd <= 1e-10 * rms, "forces differ from the oracle");
    EXPECT(std::fabs(pe - pe_ref) <= 1e-11 * std::fabs(pe_ref), "PE differs from the oracle");

    // 2. device-resident steps vs the trait default (host integration, GPU forces)
    pis::Atoms a1 = atoms, a2 = atoms;
    std::vector<pis_thermo> th;
    mgr.run_nve(a1, 0.005, 10, &th);
    for (int t = 0; t < 10; ++t) {
        const double p = mgr.pis::PotentialManager::verlet_step_nve(a2, 0.005);
        EXPECT(std::fabs(p - th[t].pe) <= 1e-11 * std::fabs(p), "step %d PE", t);
    }
    double dx = 0.0;
    for (size_t k = 0; k < 3 * n; ++k) dx = std::max(dx, std::fabs(a1.positions[k] - a2.positions[k]));
    std::printf("10 steps: device vs trait-default integrator max|dx| %.3e\n", dx);
    EXPECT(dx <= 1e-11, "trajectories diverge");

    // 3. Simulation::run thermo lines
    pis::Atoms a3 = atoms;
    auto rec = pis::Simulation::run(mgr, a3, 0.005, 3, stdout);
    EXPECT(rec.size() == 3 && std::isfinite(rec[2].etot), "Simulation::run");

    // 4. error behaviour
    pis::LJGpuManager empty(0.3, 1.0);
    try {
        pis::Atoms a4 = atoms;
        empty.compute_potential(a4);
        EXPECT(false, "empty table must throw");
    } catch (const pis::PisError& e) {
        EXPECT(e.status == PIS_ERR_MISSING_POTENTIAL, "status %d", (int)e.status);
    }
    try {
        pis::Atoms none;
        mgr.compute_potential(none);
        EXPECT(false, "no atoms must throw");
    } catch (const pis::PisError& e) {
        EXPECT(e.status == PIS_ERR_NO_ATOMS, "status %d", (int)e.status);
    }
    // 5. NVT through Simulation::run against orc_run_nvt
    {
        pis::Atoms a5 = atoms;
        OrcView o(a5);
        std::vector<double> ref(5 * 30);
        EXPECT(orc_run_nvt(&o.s, 0.004, 30, ORC_MGR_OFFSET, 1.0, 0.9, 1.1, 0.5, 3, ref.data()) == 0, "orc_run_nvt");
        pis::LJGpuManager m5(0.3, 1.0, PIS_PRECISION_FP64);
        m5.insert({1, 1}, pis::LennardJones(1.0, 1.0, 2.5, true));
        const pis::NHChainArgs chain{0.9, 1.1, 0.5, 3};
        auto th = pis::Simulation::run(m5, a5, 0.004, 30, nullptr, pis::Ensemble::NVT, chain);
        expect_tracks("NVT", th, ref, 1e-9);
    }

    // 6. NPT with a trajectory through Simulation::run against orc_run_npt
    {
        pis::Atoms a6 = atoms;
        OrcView o(a6);
        std::vector<double> ref(5 * 30), hb(9 * 30);
        EXPECT(orc_run_npt(&o.s, 0.004, 30, ORC_MGR_OFFSET, 1.0, 0.9, 0.9, 0.5, 3, 1.4, 20.0, ref.data(), hb.data()) == 0,
               "orc_run_npt");
        pis::LJGpuManager m6(0.3, 1.0, PIS_PRECISION_FP64);
        m6.insert({1, 1}, pis::LennardJones(1.0, 1.0, 2.5, true));
        const pis::NHChainArgs chain{0.9, 0.9, 0.5, 3};
        const pis::MTKBarostatArgs baro{1.4, 20.0};
        pis::DumpArgs dump;
        dump.dump_step = 10;
        dump.file_name = "/tmp/pis_test_manager_npt.lammpstrj";
        auto th = pis::Simulation::run(m6, a6, 0.004, 30, nullptr, pis::Ensemble::NPT, chain, baro, &dump);
        expect_tracks("NPT", th, ref, 1e-8);
        double dh = 0.0, hmax = 0.0;
        for (int k = 0; k < 9; ++k) {
            dh = std::max(dh, std::fabs(a6.sim_box.h[k] - hb[9 * 29 + k]));
            hmax = std::max(hmax, std::fabs(hb[9 * 29 + k]));
        }
        std::printf("NPT box: max |h - h_oracle| %.3e, h00 %.10f (start %.10f)\n", dh, a6.sim_box.h[0], atoms.sim_box.h[0]);
        EXPECT(dh <= 1e-8 * hmax, "NPT box differs from the oracle");
        EXPECT(a6.sim_box.h[0] != atoms.sim_box.h[0], "the box did not move");
        std::ifstream in(dump.file_name);
        std::stringstream ss;
        ss << in.rdbuf();
        const std::string text = ss.str();
        size_t frames = 0, pos = 0, last = 0;
        while ((pos = text.find("ITEM: TIMESTEP\n", pos)) != std::string::npos) {
            last = pos;
            ++frames;
            ++pos;
        }
        EXPECT(frames == 4, "expected 4 trajectory blocks, found %zu", frames);
        const std::string head = "ITEM: TIMESTEP\n30\nITEM: NUMBER OF ATOMS\n" + std::to_string(n) +
                                 "\nITEM: BOX BOUNDS pp pp pp\n0 " + pis::rust_display(a6.sim_box.h[0]) + "\n";
        EXPECT(text.compare(last, head.size(), head) == 0, "last block header");
        const std::string atom1 = "ITEM: ATOMS id type x y z\n1 1 " + pis::rust_display(a6.positions[0]) + " " +
                                  pis::rust_display(a6.positions[1]) + " " + pis::rust_display(a6.positions[2]) + "\n";
        EXPECT(text.find(atom1, last) != std::string::npos, "last block's first atom line");
        std::remove(dump.file_name.c_str());
    }

    std::printf(failures ? "FAILED (%d)\n" : "OK\n", failures);
    return failures ? 1 : 0;
}
