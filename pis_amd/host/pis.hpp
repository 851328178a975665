// pis.hpp -- C++ host mirror of the reference's manager API over the C ABI
// (include/pis_hip.h).  Header-only, C++17, no GPU headers: it links against
// libpis_hip.so.
//
// Names, argument meaning and error behaviour follow the Rust reference so code written
// against it reads the same:
//   SimulationBox::new_/from_lammps_data/volume   src/simulation_box.rs:5-69
//   Atoms (AoS xyz, 1-based type_ids)            src/atoms/new.rs:9-17, properties.rs:9-65
//   LennardJones                                 src/potentials/lennard_jones.rs:14-60
//   PotentialManager::compute_potential          src/potentials/potential.rs:12-13
//   PotentialManager::verlet_step_nve (default)  src/potentials/potential.rs:15-33
//   PairPotentialManager table API               src/potentials/potential.rs:148-193
//   PotentialManager::verlet_step_nvt_nhc/_npt_mtk (device-resident runs)
//                                                src/potentials/potential.rs:35-58, 112-135
//   NHChainArgs / MTKBarostatArgs                src/ensemble/nvt.rs:95-108, npt.rs:66-87
//   DumpTraj::write_step                         src/writers/dump_traj.rs:33-74
//   Simulation::run (NVE, NVT, NPT)              src/simulation.rs:8-116
//   PisError                                     src/errors.rs:5-100
#pragma once

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pis_hip.h"

namespace pis {

constexpr double KB_KJPERMOLEKELVIN = 0.0083144621;  // src/constants.rs:3

class PisError : public std::runtime_error {
public:
    PisError(pis_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
    pis_status status;
};

inline void check(pis_status s)
{
    if (s != PIS_OK) throw PisError(s, std::string(pis_status_string(s)) + ": " + pis_last_error());
}

// nalgebra Matrix3 column-major: element (r, c) at c*3 + r.
struct SimulationBox {
    std::array<double, 9> h{};
    std::array<double, 9> h_inv{};
    std::array<bool, 3> pbc{{true, true, true}};

    // SimulationBox::new (simulation_box.rs:12-15): h_inv by nalgebra's 3x3 try_inverse
    // (cofactors, first-row determinant, each entry divided by it).
    static SimulationBox new_(const std::array<double, 9>& h, std::array<bool, 3> pbc)
    {
        SimulationBox b;
        b.h = h;
        b.pbc = pbc;
        auto m = [&](int r, int c) { return h[c * 3 + r]; };
        const double m11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
        const double m12 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
        const double m13 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
        const double det = m(0, 0) * m11 - m(0, 1) * m12 + m(0, 2) * m13;
        if (det == 0.0) throw std::invalid_argument("Box matrix should be invertible");
        auto put = [&](int r, int c, double v) { b.h_inv[c * 3 + r] = v; };
        put(0, 0, m11 / det);
        put(0, 1, -(m(0, 1) * m(2, 2) - m(2, 1) * m(0, 2)) / det);
        put(0, 2, (m(0, 1) * m(1, 2) - m(1, 1) * m(0, 2)) / det);
        put(1, 0, -m12 / det);
        put(1, 1, (m(0, 0) * m(2, 2) - m(2, 0) * m(0, 2)) / det);
        put(1, 2, -(m(0, 0) * m(1, 2) - m(1, 0) * m(0, 2)) / det);
        put(2, 0, m13 / det);
        put(2, 1, -(m(0, 0) * m(2, 1) - m(2, 0) * m(0, 1)) / det);
        put(2, 2, (m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1)) / det);
        return b;
    }

    // SimulationBox::from_lammps_data (simulation_box.rs:44-65)
    static SimulationBox from_lammps_data(double xlo, double xhi, double ylo, double yhi, double zlo,
                                          double zhi, double xy, double xz, double yz)
    {
        std::array<double, 9> h{};
        h[0] = xhi - xlo;
        h[3] = xy;
        h[4] = yhi - ylo;
        h[6] = xz;
        h[7] = yz;
        h[8] = zhi - zlo;
        return new_(h, {true, true, true});
    }

    double volume() const
    {
        auto m = [&](int r, int c) { return h[c * 3 + r]; };
        const double m11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
        const double m12 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
        const double m13 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
        return std::fabs(m(0, 0) * m11 - m(0, 1) * m12 + m(0, 2) * m13);
    }
};

struct Atoms {
    size_t n_atoms = 0;
    std::vector<uint32_t> type_ids;  // 1-based
    std::vector<double> masses;      // per type
    std::vector<double> positions;   // 3 * n_atoms, xyz per atom
    std::vector<double> velocities;
    std::vector<double> forces;
    SimulationBox sim_box;

    double mass_i(size_t i) const { return masses.at(type_ids.at(i) - 1); }

    // kinetic_energy (properties.rs:17-23)
    double kinetic_energy() const
    {
        double ek = 0.0;
        for (size_t i = 0; i < n_atoms; ++i) {
            const double* v = &velocities[3 * i];
            ek += 0.5 * mass_i(i) * ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        }
        return ek;
    }

    double degress_of_freedom() const { return 3.0 * (double)n_atoms; }

    double temerature(double kinetic_energy, double kb = KB_KJPERMOLEKELVIN) const
    {
        return (2.0 * kinetic_energy) / (degress_of_freedom() * kb);
    }

    // trace of virial_tensor (properties.rs:49-51) and pressure (properties.rs:61-65)
    double virial_trace() const
    {
        double w = 0.0;
        for (size_t k = 0; k < 3 * n_atoms; ++k) w += positions[k] * forces[k];
        return w;
    }

    double pressure(double kinetic_energy) const
    {
        return (2.0 * kinetic_energy + virial_trace()) / (3.0 * sim_box.volume());
    }
};

struct LennardJones {
    double epsilon, sigma, rcut;
    bool shift;
    LennardJones(double e, double s, double rc, bool sh) : epsilon(e), sigma(s), rcut(rc), shift(sh) {}
    double get_rcut() const { return rcut; }
};

// The reference's PotentialManager trait: compute_potential is required, verlet_step_nve
// has the trait's default body (host-side integration calling compute_potential).
class PotentialManager {
public:
    virtual ~PotentialManager() = default;
    virtual double compute_potential(Atoms& atoms) = 0;

    // potential.rs:15-33, op for op
    virtual double verlet_step_nve(Atoms& atoms, double dt)
    {
        const size_t n = atoms.n_atoms;
        std::vector<double> a_t(3 * n);
        for (size_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) a_t[3 * i + d] = atoms.forces[3 * i + d] / atoms.mass_i(i);
        const double dt2 = dt * dt;
        for (size_t k = 0; k < 3 * n; ++k) atoms.positions[k] += atoms.velocities[k] * dt + a_t[k] * 0.5 * dt2;
        const auto& b = atoms.sim_box;
        for (size_t i = 0; i < n; ++i) {  // apply_boundary_conditions_pos
            double* r = &atoms.positions[3 * i];
            double s[3];
            for (int q = 0; q < 3; ++q) s[q] = b.h_inv[q] * r[0];
            for (int c = 1; c < 3; ++c)
                for (int q = 0; q < 3; ++q) s[q] = b.h_inv[c * 3 + q] * r[c] + s[q];
            for (int q = 0; q < 3; ++q)
                if (b.pbc[q]) s[q] -= std::floor(s[q]);
            for (int q = 0; q < 3; ++q) r[q] = b.h[q] * s[0];
            for (int c = 1; c < 3; ++c)
                for (int q = 0; q < 3; ++q) r[q] = b.h[c * 3 + q] * s[c] + r[q];
        }
        std::fill(atoms.forces.begin(), atoms.forces.end(), 0.0);
        const double pe = compute_potential(atoms);
        for (size_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) {
                const double a = atoms.forces[3 * i + d] / atoms.mass_i(i);
                atoms.velocities[3 * i + d] += (a_t[3 * i + d] + a) * 0.5 * dt;
            }
        return pe;
    }
};

// Arguments of NHThermostatChain::new_from_args (nvt.rs:95-108) and of
// MTKBarostat::new_from_args (npt.rs:66-87), as `fix nvt` / `fix npt` give them.
struct NHChainArgs {
    double start_temperature = 0.0, end_temperature = 0.0, tau = 0.0;
    int32_t chain = 3;
};
struct MTKBarostatArgs {
    double start_pressure = 0.0, tau = 0.0;
};
enum class Ensemble { NVE, NVT, NPT };

// DumpArgs (simulation_context.rs:78-101): a frame every dump_step steps.
struct DumpArgs {
    size_t dump_step = 1;
    std::string file_name = "dump.lammpstrj";
};

// f64 as Rust's `{}` prints it (shortest round-trip digits, positional).
inline std::string rust_display(double x)
{
    char buf[400];
    size_t n = 0;
    check(pis_format_f64(x, buf, sizeof buf, &n));
    return std::string(buf, n);
}

// GPU LJ manager: the pair table of PairPotentialManager and compute_potential /
// verlet_step_nve through the C ABI.  A context is created on first use for the atoms'
// shape and reused while the shape, box and types do not change.
class LJGpuManager : public PotentialManager {
public:
    using AtomPair = std::pair<size_t, size_t>;

    explicit LJGpuManager(double skin = 0.0, double kb = KB_KJPERMOLEKELVIN,
                          pis_precision precision = PIS_PRECISION_FP64)
        : skin_(skin), kb_(kb), precision_(precision)
    {
    }
    ~LJGpuManager() override { release(); }
    LJGpuManager(const LJGpuManager&) = delete;
    LJGpuManager& operator=(const LJGpuManager&) = delete;

    // PairPotentialManager::insert / get / is_empty / max_rcut (potential.rs:157-179)
    void insert(AtomPair key, LennardJones lj)
    {
        table_.erase(key);
        table_.emplace(key, lj);
        release();
    }
    const LennardJones* get(AtomPair key) const
    {
        auto it = table_.find(key);
        return it == table_.end() ? nullptr : &it->second;
    }
    bool is_empty() const { return table_.empty(); }
    double max_rcut() const
    {
        double m = 0.0;
        for (const auto& kv : table_) m = std::max(m, kv.second.get_rcut());
        return m;
    }

    // PotentialManager::compute_potential: forces ADDED into atoms.forces, PE returned.
    double compute_potential(Atoms& atoms) override
    {
        ensure(atoms);
        check(pis_set_state(ctx_, atoms.positions.data(), atoms.velocities.data(), nullptr));
        double pe = 0.0;
        check(pis_compute_potential(ctx_, atoms.forces.data(), &pe));
        return pe;
    }

    // verlet_step_nve with the state kept on the device for the step.
    double verlet_step_nve(Atoms& atoms, double dt) override
    {
        std::vector<pis_thermo> th(1);
        run_nve(atoms, dt, 1, &th);
        return th[0].pe;
    }

    // nsteps device-resident steps; thermo per step if requested.
    void run_nve(Atoms& atoms, double dt, uint64_t nsteps, std::vector<pis_thermo>* thermo = nullptr)
    {
        ensure(atoms);
        check(pis_set_state(ctx_, atoms.positions.data(), atoms.velocities.data(), atoms.forces.data()));
        if (thermo) thermo->resize(nsteps);
        check(pis_verlet_step_nve(ctx_, dt, nsteps, thermo ? thermo->data() : nullptr));
        check(pis_get_state(ctx_, atoms.positions.data(), atoms.velocities.data(), atoms.forces.data()));
    }

    // NHThermostatChain::new_from_args on the device (xi = eta = G = 0, target = start).
    void nvt_init(const Atoms& atoms, const NHChainArgs& c)
    {
        ensure(atoms);
        const pis_nhc_params p{c.start_temperature, c.end_temperature, c.tau, c.chain, 0};
        check(pis_nvt_init(ctx_, &p));
    }

    // The chain as above plus MTKBarostat::new_from_args (target p0 I, zero momentum).
    void npt_init(const Atoms& atoms, const NHChainArgs& c, const MTKBarostatArgs& b)
    {
        ensure(atoms);
        const pis_nhc_params p{c.start_temperature, c.end_temperature, c.tau, c.chain, 0};
        const pis_mtk_params m{b.start_pressure, b.tau, {0.0, 0.0}};
        check(pis_npt_init(ctx_, &p, &m));
    }

    // nsteps of verlet_step_nvt_nhc with Simulation::step's target ramp continued from
    // step ramp_step0 of a ramp_total-step run (simulation.rs:53-57).
    void run_nvt(Atoms& atoms, double dt, uint64_t nsteps, uint64_t ramp_step0, uint64_t ramp_total,
                 std::vector<pis_thermo>* thermo = nullptr)
    {
        ensure(atoms);
        check(pis_set_state(ctx_, atoms.positions.data(), atoms.velocities.data(), atoms.forces.data()));
        if (thermo) thermo->resize(nsteps);
        check(pis_batch_verlet_step_nvt_ramp(ctx_, &dt, nsteps, ramp_step0, ramp_total,
                                             thermo ? thermo->data() : nullptr));
        check(pis_get_state(ctx_, atoms.positions.data(), atoms.velocities.data(), atoms.forces.data()));
    }

    // nsteps of verlet_step_npt_mtk; the barostat moved the box on the device, so
    // atoms.sim_box takes the new h (h_inv by SimulationBox::new, as scale_box does).
    void run_npt(Atoms& atoms, double dt, uint64_t nsteps, uint64_t ramp_step0, uint64_t ramp_total,
                 std::vector<pis_thermo>* thermo = nullptr)
    {
        ensure(atoms);
        check(pis_set_state(ctx_, atoms.positions.data(), atoms.velocities.data(), atoms.forces.data()));
        if (thermo) thermo->resize(nsteps);
        check(pis_batch_verlet_step_npt(ctx_, &dt, nsteps, ramp_step0, ramp_total,
                                        thermo ? thermo->data() : nullptr));
        check(pis_get_state(ctx_, atoms.positions.data(), atoms.velocities.data(), atoms.forces.data()));
        std::array<double, 9> h{};
        check(pis_npt_get_box(ctx_, 0, h.data(), nullptr));
        atoms.sim_box = SimulationBox::new_(h, atoms.sim_box.pbc);
        h_ = h;  // the context follows the box: no new context for the next call
    }

private:
    void release()
    {
        if (ctx_) pis_ctx_destroy(ctx_);
        ctx_ = nullptr;
    }

    void ensure(const Atoms& atoms)
    {
        if (atoms.n_atoms == 0) throw PisError(PIS_ERR_NO_ATOMS, "No atoms defined in input file");
        if (table_.empty())
            throw PisError(PIS_ERR_MISSING_POTENTIAL,
                           "Potential manager not initialized - missing pair_style or pair_coeff commands");
        if (ctx_ && n_ == atoms.n_atoms && h_ == atoms.sim_box.h && types_ == atoms.type_ids &&
            masses_ == atoms.masses)
            return;
        release();
        std::vector<pis_pair_coeff> pairs;
        bool shift = true;
        for (const auto& kv : table_) {
            pairs.push_back({(uint32_t)kv.first.first, (uint32_t)kv.first.second, kv.second.epsilon,
                             kv.second.sigma, kv.second.rcut});
            shift = shift && kv.second.shift;
        }
        pis_system_desc d{};
        d.n_atoms = atoms.n_atoms;
        d.n_types = (uint32_t)atoms.masses.size();
        d.n_pairs = (uint32_t)pairs.size();
        d.type_ids = atoms.type_ids.data();
        d.masses = atoms.masses.data();
        d.pairs = pairs.data();
        for (int k = 0; k < 9; ++k) {
            d.h[k] = atoms.sim_box.h[k];
            d.h_inv[k] = atoms.sim_box.h_inv[k];
        }
        for (int k = 0; k < 3; ++k) d.pbc[k] = atoms.sim_box.pbc[k] ? 1 : 0;
        d.shift = shift ? 1 : 0;
        d.precision = (uint8_t)precision_;
        d.skin = skin_;
        d.kb = kb_;
        check(pis_ctx_create(&d, &ctx_));
        n_ = atoms.n_atoms;
        h_ = atoms.sim_box.h;
        types_ = atoms.type_ids;
        masses_ = atoms.masses;
    }

    std::map<AtomPair, LennardJones> table_;
    double skin_, kb_;
    pis_precision precision_;
    pis_ctx* ctx_ = nullptr;
    size_t n_ = 0;
    std::array<double, 9> h_{};
    std::vector<uint32_t> types_;
    std::vector<double> masses_;
};

// DumpTraj (dump_traj.rs:12-75): File::create, then one LAMMPS text block per write_step,
// formatted by the library's writer.
class DumpTraj {
public:
    explicit DumpTraj(const DumpArgs& args) : path_(args.file_name)
    {
        fd_ = ::open(path_.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
        if (fd_ < 0) throw PisError(PIS_ERR_IO, "Failed to create trajectory file '" + path_ + "'");
    }
    ~DumpTraj()
    {
        if (fd_ >= 0) ::close(fd_);
    }
    DumpTraj(const DumpTraj&) = delete;
    DumpTraj& operator=(const DumpTraj&) = delete;

    void write_step(const Atoms& atoms, uint64_t step)
    {
        check(pis_dump_write_block(fd_, step, atoms.n_atoms, atoms.positions.data(), atoms.type_ids.data(),
                                   atoms.sim_box.h.data(), 0));
    }

private:
    std::string path_;
    int fd_ = -1;
};

// Simulation::run (simulation.rs:8-116): the step-0 frame, one compute_potential from zero
// forces ("0 PE" printed as Rust's {}), then the steps on the device in chunks that end at
// every dump_step, one thermo line "step PE KE H T P" ({:.3}) per step and a frame at the
// end of each chunk.  H carries the chain's (NVT) and the barostat's (NPT) energies.
// Without dump arguments no trajectory is written.
struct Simulation {
    static std::vector<pis_thermo> run(LJGpuManager& mgr, Atoms& atoms, double dt, uint64_t steps,
                                       FILE* out = nullptr, Ensemble ensemble = Ensemble::NVE,
                                       const NHChainArgs& chain = {}, const MTKBarostatArgs& barostat = {},
                                       const DumpArgs* dump = nullptr)
    {
        std::unique_ptr<DumpTraj> dumper;
        if (dump) {
            if (dump->dump_step == 0) throw PisError(PIS_ERR_INVALID_ARGUMENT, "dump_step must be positive");
            dumper.reset(new DumpTraj(*dump));
            dumper->write_step(atoms, 0);
        }
        std::fill(atoms.forces.begin(), atoms.forces.end(), 0.0);
        const double first = mgr.compute_potential(atoms);
        if (out) std::fprintf(out, "0 %s\n", rust_display(first).c_str());
        if (ensemble == Ensemble::NVT) mgr.nvt_init(atoms, chain);
        if (ensemble == Ensemble::NPT) mgr.npt_init(atoms, chain, barostat);
        const uint64_t cadence = dump ? dump->dump_step : (steps ? steps : 1);
        std::vector<pis_thermo> all, th;
        uint64_t done = 0;
        while (done < steps) {
            const uint64_t k = std::min<uint64_t>(cadence - done % cadence, steps - done);
            if (ensemble == Ensemble::NVT)
                mgr.run_nvt(atoms, dt, k, done, steps, &th);
            else if (ensemble == Ensemble::NPT)
                mgr.run_npt(atoms, dt, k, done, steps, &th);
            else
                mgr.run_nve(atoms, dt, k, &th);
            if (out)
                for (uint64_t j = 0; j < k; ++j)
                    std::fprintf(out, "%llu %.3f %.3f %.3f %.3f %.3f\n", (unsigned long long)(done + j + 1),
                                 th[j].pe, th[j].ke, th[j].etot, th[j].temperature, th[j].pressure);
            all.insert(all.end(), th.begin(), th.end());
            done += k;
            if (dumper && done % cadence == 0) dumper->write_step(atoms, done);
        }
        return all;
    }
};

}  // namespace pis
