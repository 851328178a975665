/*
 * pis_oracle.h -- CPU FP64 restatement of the reference's Lennard-Jones force and
 * velocity-Verlet path (vivekadishankara/pis, Rust).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the smoke
 * entry point and bench.py's CPU-baseline leg may load it.  The product path
 * (pis_amd/libpis_hip.so) never links, loads or calls it.
 *
 * Built with gcc -O2 -ffp-contract=off (Rust never contracts a*b+c into an FMA, so
 * neither may the restatement).  Every function cites the reference file:line it
 * follows.  Matrices are 3x3 column-major doubles, like nalgebra's Matrix3
 * (element (r,c) at index c*3+r); positions/velocities/forces are AoS xyz (3*n
 * doubles), like nalgebra's Matrix3xX.  Atom types are 1-based.
 *
 * Parity status per function (what pins it):
 *   orc_box_inverse      -- nalgebra 0.34.1 3x3 try_inverse, restated from its
 *                           published source from memory; pinned only by
 *                           h*h_inv = I and 1-ulp agreement with 1/L for diagonal
 *                           boxes.  Exact bits for triclinic boxes: parity unpinned.
 *   orc_gemv             -- nalgebra gemv column order; exact for orthorhombic boxes
 *                           (off-diagonal terms are signed zeros), unpinned for tilt.
 *   orc_mat3_exp         -- this oracle's own method (nalgebra's Pade exp() is not
 *                           vendored): checked against scipy.linalg.expm; NPT parity
 *                           with the reference is tolerance-based.
 *   orc_kinetic_tensor,
 *   orc_virial_tensor    -- sums in ascending atom order (matrixmultiply's blocked
 *                           order is not restated): tolerance-based.
 *   everything else      -- pinned by closed forms, brute force and invariants in
 *                           tests/test_oracle.py (listed there per function).
 */
#ifndef PIS_ORACLE_H
#define PIS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double h[9];    /* box vectors as columns */
    double hinv[9];
    int32_t pbc[3];
} orc_box;

/* Dense symmetric pair table, (ti-1)*ntypes + (tj-1); present==0 means the
 * reference's HashMap lookup misses (potential.rs:181-192) and the pair is skipped. */
typedef struct {
    int32_t ntypes;
    const int32_t* present;
    const double* eps;
    const double* sigma;
    const double* rcut;
    int32_t shift; /* LennardJones::shift (always true on the shipped path) */
} orc_table;

typedef struct {
    int64_t n;
    double* pos;          /* 3n */
    double* vel;          /* 3n */
    double* force;        /* 3n */
    const int32_t* type;  /* n, 1-based */
    const double* mass;   /* ntypes */
    orc_box box;
    orc_table table;
} orc_sys;

/* CSR neighbour list (full list, i != j). */
typedef struct {
    int64_t n;
    int64_t* start; /* n+1 */
    int64_t* idx;
    int64_t cap;
} orc_list;

/* Manager selector for orc_compute / orc_verlet_step_nve. */
enum {
    ORC_MGR_OFFSET = 0,    /* LJVOffsetManager, the shipped one (lennard_jones.rs:182-244) */
    ORC_MGR_BRUTE = 1,     /* LJManager O(N^2) (lennard_jones.rs:62-99) */
    ORC_MGR_VERLET27 = 2,  /* LJVerletManager (lennard_jones.rs:103-178) */
    ORC_MGR_FULLLIST = 3   /* LJVPBuildListManager (lennard_jones.rs:340-456) */
};

int orc_box_inverse(const double* h, double* hinv);
void orc_box_from_lammps(double xlo, double xhi, double ylo, double yhi, double zlo,
                         double zhi, double xy, double xz, double yz, orc_box* out);
double orc_box_volume(const orc_box* b);
void orc_gemv(const double* m, const double* x, double* y);
void orc_min_image(const orc_box* b, double* d);
void orc_wrap(const orc_box* b, double* r);
void orc_divide_into_cells(const orc_box* b, double rcut, int64_t* ncell);
int64_t orc_cell_of(const orc_box* b, const int64_t* ncell, const double* r);
double orc_norm(const double* v);
double orc_max_rcut(const orc_table* t);

void orc_lj_pair(double eps, double sigma, double rcut, int32_t shift, const double* rij,
                 double* u, double* f);

double orc_compute(orc_sys* s, int32_t manager);
double orc_compute_offset(orc_sys* s);
double orc_compute_brute(orc_sys* s);
double orc_compute_verlet27(orc_sys* s);

int orc_build_list(const orc_sys* s, double rcut_add, orc_list* out);
void orc_list_free(orc_list* l);
int orc_brute_list(const orc_sys* s, double rcut_add, orc_list* out);
double orc_compute_fulllist(orc_sys* s, const orc_list* l, int32_t check_rcut,
                            double* pe_atom);

/* splitmix64's output function, the entry hash of orc_list_digest. */
uint64_t orc_mix64(uint64_t x);
/* Per-atom row length and order-independent row hash (sum mod 2^64 of orc_mix64(j)) of
 * the list orc_build_list(s, rcut_add) returns, computed without storing it. */
int orc_list_digest(const orc_sys* s, double rcut_add, uint32_t* count, uint64_t* hash);
/* Forces (3m, AoS) and unhalved pair energies (m, may be NULL) of the atoms idx[0..m),
 * bit-identical to what orc_compute_fulllist over orc_build_list(s, 0) gives them. */
int orc_forces_subset(const orc_sys* s, const int64_t* idx, int64_t m, double* f_out, double* pe_out);

/* Periodic pair virial sum_{i<j} r_ij . F_ij of the pairs within rcut. */
double orc_pair_virial(const orc_sys* s);

double orc_verlet_step_nve(orc_sys* s, double dt, int32_t manager);
double orc_kinetic_energy(const orc_sys* s);
double orc_temperature(const orc_sys* s, double ke, double kb);
double orc_virial_trace(const orc_sys* s);
double orc_pressure(const orc_sys* s, double ke);

/* Simulation::run NVE loop: initial compute, then nsteps of verlet_step_nve; per step
 * writes PE, KE, H=PE+KE, T, P into thermo[5*k]. */
int orc_run_nve(orc_sys* s, double dt, int64_t nsteps, int32_t manager, double kb,
                double* thermo);

/* Nose-Hoover chain (ensemble/nvt.rs:4-119), restated with its quirks: the half-step
 * propagation *sets* xi from G (it does not add to it), eta never moves, and the chain
 * masses keep the target temperature of construction. */
#define ORC_NHC_MAX 8
typedef struct {
    int32_t chain;
    double t_start, t_end, target, kb;
    double xi[ORC_NHC_MAX], eta[ORC_NHC_MAX], g[ORC_NHC_MAX], q[ORC_NHC_MAX];
} orc_nhc;

/* NHThermostatChain::new via new_from_args (nvt.rs:22-57, 95-108): target = start. */
int orc_nhc_new(orc_nhc* c, double t_start, double t_end, double tau, int32_t chain, double kb);
void orc_nhc_compute_forces(orc_nhc* c, double ke, int64_t n);  /* nvt.rs:60-71 */
void orc_nhc_propagate_half_step(orc_nhc* c, double dt);        /* nvt.rs:74-82 */
double orc_nhc_kinetic_energy(const orc_nhc* c);                /* nvt.rs:84-90 */
double orc_nhc_potential_energy(const orc_nhc* c, int64_t n);   /* nvt.rs:92-99 */
void orc_nhc_target(orc_nhc* c, int64_t step, int64_t total);   /* nvt.rs:109-118 */
/* verlet_step_nvt_nhc (potential.rs:35-58); returns PE(t+dt). */
double orc_verlet_step_nvt_nhc(orc_sys* s, double dt, orc_nhc* c, int32_t manager);
/* Simulation::run with Ensemble::NVT (simulation.rs:8-64, 90-116): per step writes PE,
 * KE, H = PE + KE + chain KE + chain PE, T, P into thermo[5*k]. */
int orc_run_nvt(orc_sys* s, double dt, int64_t nsteps, int32_t manager, double kb, double t_start,
                double t_end, double tau, int32_t chain, double* thermo);

/* MTK barostat (ensemble/npt.rs:10-88) and the NPT step (potential.rs:112-135).
 * Tensors are 3x3 column-major.  The matrix exponential is this oracle's own method
 * (scaling and squaring of a Taylor series; nalgebra 0.34.1's Pade-based exp() is not
 * vendored in the reference): parity of exp() and of the tensor sums is tolerance-based. */
typedef struct {
    double target[9];   /* target pressure tensor: p0 * I for `fix ... iso p0 p1 tau` */
    double momentum[9]; /* barostat momentum */
    double w;           /* barostat mass */
} orc_mtk;

/* MTKBarostat::new via new_from_args (npt.rs:24-42, 66-87): w = 3n KB T tau^2 with T the
 * chain's start temperature (300 without a chain) and KB = 0.0083144621. */
void orc_mtk_new(orc_mtk* m, double p0, double tau, int64_t n, double t_target);
void orc_kinetic_tensor(const orc_sys* s, double* out);   /* properties.rs:41-43, no masses */
void orc_virial_tensor(const orc_sys* s, double* out);    /* properties.rs:45-47 */
void orc_pressure_tensor(const orc_sys* s, double* out);  /* properties.rs:53-59 */
void orc_mtk_delta_momentum(const orc_mtk* m, const orc_sys* s, double dt, double* out); /* npt.rs:44-49 */
void orc_mtk_scale(const orc_mtk* m, double dt, int32_t velocity_scaling, double* out);  /* npt.rs:51-58 */
double orc_mtk_kinetic_energy(const orc_mtk* m);                 /* npt.rs:60-62 */
double orc_mtk_potential_energy(const orc_mtk* m, const double* h); /* npt.rs:64-66 */
void orc_mat3_mul(const double* a, const double* b, double* out);   /* nalgebra gemm order */
void orc_mat3_exp(const double* a, double* out);
void orc_scale_box(orc_sys* s, const double* scale);            /* transformations.rs:6-15 */
double orc_verlet_step_npt_mtk(orc_sys* s, double dt, orc_mtk* m, orc_nhc* c, int32_t manager);
/* Simulation::run with Ensemble::NPT (simulation.rs:8-64, 90-116): per step writes PE, KE,
 * H = PE + KE + chain KE + chain PE + barostat KE + barostat PE, T, P into thermo[5*k] and,
 * when hbox is given, the box matrix after the step into hbox[9*k]. */
int orc_run_npt(orc_sys* s, double dt, int64_t nsteps, int32_t manager, double kb, double t_start,
                double t_end, double tau_t, int32_t chain, double p0, double tau_p, double* thermo,
                double* hbox);

/* Idealised parallel CPU baseline (not a restatement): full list + OpenMP gather. */
int orc_ideal_run_nve(orc_sys* s, double dt, int64_t nsteps, double skin, int32_t nthreads,
                      double* pe_out);

#ifdef __cplusplus
}
#endif
#endif
