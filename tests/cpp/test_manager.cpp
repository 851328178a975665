// C++ parity test of the host mirror (pis_amd/host/pis.hpp) over the C ABI, written the
// way the reference's own code uses its manager.  Needs a GPU; run by
// tests/test_gpu_cpp.py.  Checks:
//   1. LJGpuManager::compute_potential vs the CPU oracle's LJVOffsetManager restatement:
//      forces within 1e-10 normwise, PE within the serial-sum bound;
//   2. the device-resident verlet_step_nve vs the trait's default host-side integrator
//      (which calls compute_potential each step) over 10 steps;
//   3. Simulation::run prints the reference's thermo lines;
//   4. error behaviour: empty table -> MissingPotential, no atoms -> NoAtomsDefined.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

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
    std::printf(failures ? "FAILED (%d)\n" : "OK\n", failures);
    return failures ? 1 : 0;
}
