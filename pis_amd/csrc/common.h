// common.h -- device data layout shared by the gfx950 kernels and the host engine.
//
// One simulated system ("sys") owns:
//   * two buffer sets of per-atom SoA state in cell-sorted order (double buffered so a
//     rebuild can permute without a copy-back; the live set is chosen on the device by
//     ctl[CTL_CUR] so captured launch sequences never need host decisions);
//   * reference positions (r at the last list build) for the skin trigger;
//   * a cell grid whose edge is >= max(rcut)+skin, cells numbered tile-major so that the
//     atoms of one tile are contiguous;
//   * a full neighbour list of u16 LDS byte offsets, 8 per uint4, laid out
//     [group][atom] so a wavefront's loads of one group are one coalesced 1 KiB read.
#pragma once
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace pis {

constexpr int kTileThreads = 256;  // workgroup of the tile kernels (4 wave64)
constexpr int kAtomThreads = 256;  // workgroup of the per-atom kernels
constexpr int kMaxHalo = 256;      // halo-box cells a tile may stage
constexpr int kMaxTypes = 4;
constexpr int kMaxList = 4096;     // neighbour-list capacity limit per atom (list memory: 2 B per entry)
constexpr int kMaxLanes = 4;       // streams a context spreads its staging classes over

// ctl[] slots (device ints per system)
enum {
    CTL_CUR = 0,       // live buffer set (0/1)
    CTL_REBUILD = 1,   // rebuild requested for this step
    CTL_OVF_NB = 2,    // neighbour capacity exceeded at the last build
    CTL_OVF_STAGE = 3, // staging capacity exceeded
    CTL_NAN = 4,       // non-finite position or energy seen
    CTL_MAXNB = 5,     // largest list length at the last build
    CTL_NREBUILD = 6,  // builds performed
    CTL_STEP = 7,      // thermo record index
    CTL_MAXSTAGE = 8,  // most atoms staged by one tile since the state was set
    CTL_LEFT = 9,      // NVE steps still to run: < 0 unlimited, 0 frozen (steps skip it)
    CTL_REMAIN = 10,   // steps left in the current call (the force kernel advances the
                       // positions of the next step while more steps follow)
    CTL_REBUILD_NEXT = 11,  // skin check of those advanced positions
    CTL_BOX = 12,      // the box (NPT) became too thin for the planned cell grid
    CTL_BADINPUT = 13, // device-resident input rejected: bit 0 atom type, bit 1 position
    CTL_COUNT = 14
};

// Device-resident box of a system (SysDev::box): h and h^-1 column-major, |det h|.  The
// host's bits on upload; the barostat updates them on the device.
enum { BOX_H = 0, BOX_HINV = 9, BOX_VOL = 18, BOX_COUNT = 20 };

// MTK barostat state per system (SysDev::npt, ensemble/npt.rs:10-20), device-updated.
enum {
    NPT_TARGET = 0,     // target pressure tensor (column-major)
    NPT_MOM = 9,        // barostat momentum
    NPT_W = 18,         // barostat mass
    NPT_SCALE = 19,     // velocity scaling matrix of the current step
    NPT_HINV_OLD = 28,  // h^-1 before the current step's box scaling
    NPT_RCL = 37,       // cell edge the grid needs (rcut + skin with margin)
    NPT_HREF_INV = 40,  // h^-1 of the box the current neighbour list was built in
    NPT_DEF = 49,       // deformation since that build, h h_ref^-1, of the current step
    NPT_SKIN2 = 58,     // bound on |r - D r_ref|^2 that keeps the list complete this step
    NPT_RCMAX = 59,     // largest cutoff
    NPT_LIST_OK = 60,   // 1 once NPT_HREF_INV describes the current list
    NPT_COUNT = 64
};
// per-block partial sums of the barostat kernels (SysDev::tens): v v^T, r F^T, (1/2) m v^2
enum { TENS_K = 0, TENS_W = 9, TENS_KE = 18, TENS_COUNT = 20 };

enum { MODE_TILE_REL = 0, MODE_RAW_MI = 1, MODE_WHOLE = 2 };
// Kinds of systems a launch covers, by box shape and staging mode: tile-relative staging of
// an orthorhombic (KIND_ORTHO) or general box (KIND_TRI), and raw staging with a per-pair
// image (KIND_ORTHO_RAW, KIND_TRI_RAW: boxes too small for tile-relative staging, whole-box
// systems).  The build kernel comes in an orthorhombic and a general-box instantiation; the
// force kernel has one per kind, so the common tile-relative kinds keep their registers and
// LDS (FP32 staging in mixed precision) free of the raw kinds' needs.  Tiles are grouped by
// kind, and each launch takes only its own kind's tiles.
enum { KIND_ORTHO = 1, KIND_TRI = 2, KIND_TRI_RAW = 4, KIND_ORTHO_RAW = 8 };

struct PairParam {
    double r2_force;  // largest r^2 with sqrt_rn(r^2) <= rcut (-1: pair absent)
    double r2_list;   // largest r^2 with sqrt_rn(r^2) <= rcut + skin (-1: absent)
    // Rounding band around each threshold: an r^2 computed from staged coordinates is
    // within band of the reference's r^2, so below *_lo is surely inside, above *_hi
    // surely outside, and only the band needs the bit-exact reference predicate.
    double f_lo, f_hi, l_lo, l_hi;
    double sig2;      // sigma^2
    double c4;        // 4 eps
    double c24;       // 24 eps
    double c48;       // 48 eps
    double ucut;      // energy shift U(rcut) (0 without shift)
    double isig;      // 1 / sigma: the single-species mixed loop stages coordinates in sigma units
    float sig2f, c4f, c24f, c48f, ucutf;
    float r2f;        // FP32 cutoff of the mixed pair loop (rcut^2 rounded to FP32)
    float r2s32;      // the same cutoff in sigma units, (rcut / sigma)^2 rounded to FP32
    // list-band edges for r^2 from the list build's FP32 staging (wider: the staged
    // coordinates carry FP32 rounding); -1 for an absent pair
    float l_lo32, l_hi32;
};

// Nose-Hoover chain state per system, in device memory (slots of SysDev::nhc).
enum {
    NHC_XI = 0,      // thermostat velocities xi[0..7]
    NHC_Q = 8,       // masses q[0..7]
    NHC_G = 16,      // forces G[0..7]
    NHC_T = 24,      // target temperature
    NHC_T0 = 25,     // start temperature
    NHC_T1 = 26,     // end temperature
    NHC_KB = 27,     // Boltzmann constant of the chain
    NHC_SCALE = 28,  // exp(-dt xi[0] / 2) of the current step
    NHC_STEP = 29,   // step index inside the current run (for the target ramp)
    NHC_TOTAL = 30,  // steps of the current run
    NHC_CHAIN = 31,  // chain length
    NHC_COUNT = 40,
    NHC_MAX_CHAIN = 8
};

struct State {
    double *x, *y, *z;
    double *vx, *vy, *vz;
    double *fx, *fy, *fz;
    uint8_t* type;  // 0-based
    uint32_t* orig; // original atom index
};

struct ThermoDev {
    double pe, ke, etot, temperature, pressure, virial_abs, virial_pair;
};

struct SysDev {
    State st[2];
    int32_t* ctl;
    double *rx, *ry, *rz;  // positions at the last build (sorted order)
    int32_t* cell_count;   // [ncell] atoms per cell while sorting; zero between rebuilds
    int32_t* cell_start;   // [ncell+1], sorted (tile-major) cell ids
    int32_t* nat2sorted;   // [ncell]: natural (cz*ny+cy)*nx+cx -> sorted id
    int32_t* tile_cell0;   // [ntile+1]
    int32_t* atom_cell;    // [n] sorted cell id of atom (live order) at binning
    int32_t* scat_src;     // [n] scatter slot -> source atom
    uint32_t* scat_key;    // [n] scatter slot -> orig index (sort key, ties)
    double* scat_x;        // [n] scatter slot -> x (sort key)
    uint4* nbr;            // [(maxnb/8) * npad]
    uint16_t* nnb;         // [npad]
    double* partial;       // [ntile*4]: pe, ke, virial_abs, virial_pair
    ThermoDev* thermo;     // [thermo_cap]
    const PairParam* pp;   // [ntypes*ntypes]
    const double* mass;    // [ntypes]
    double L[3], Linv[3];  // h diagonal and h_inv diagonal (caller's bits)
    double cellw[3];
    double half_skin2;     // (skin/2)^2
    double wslack;         // x-window slack of the list build: FP32 staging error + margin
    double dt, dt2;
    double kb, volume;
    int32_t n, npad, ntypes, maxnb;
    int32_t nc[3], ncell;
    int32_t tile[3], nt[3], ntile;
    int32_t mode;
    int32_t thermo_cap;
    int32_t nhc_on;        // NVT: apply the chain's velocity scaling in the step kernels
    double* nhc;           // [NHC_COUNT] Nose-Hoover chain state (device-updated)
    int32_t tri;           // general (tilted) box: minimum image, wrap and binning use box[]
    int32_t npt_on;        // NPT: the chain's start-of-step KE comes from the barostat's sums
    double* box;           // [BOX_COUNT] h, h^-1, volume (device-updated under NPT)
    double* npt;           // [NPT_COUNT] barostat state
    double* tens;          // [nblk * TENS_COUNT] per-atom-block partial sums
    int32_t nblk;          // 256-atom blocks (npad / 256, rounded up)
    int32_t list_shift;    // list entry = (LDS slot << list_shift) | partner type: 3 when the
                           // force kernel stages FP64 coordinates (8-byte slots), 2 for FP32
};

// Cell numbering.  Cells are numbered tile-major: tiles in x-fastest order, and inside a
// tile its cells in x-fastest order, so the atoms of a tile are contiguous after the sort.
// First sorted cell id of tile (tx, ty, tz): the cells of every earlier tile (full z-slabs,
// then full y-rows of this slab, then the tiles before it in its row).
__host__ __device__ inline int32_t tile_first_cell(const int32_t* nc, const int32_t* tile, int tx, int ty, int tz)
{
    const int sz = min(tile[2], nc[2] - tz * tile[2]), sy = min(tile[1], nc[1] - ty * tile[1]);
    return nc[0] * nc[1] * (tz * tile[2]) + nc[0] * (ty * tile[1]) * sz + (tx * tile[0]) * sy * sz;
}

// Sorted (tile-major) id of natural cell (x, y, z).
__host__ __device__ inline int32_t sorted_cell_id(const int32_t* nc, const int32_t* tile, int x, int y, int z)
{
    const int tx = x / tile[0], ty = y / tile[1], tz = z / tile[2];
    const int sx = min(tile[0], nc[0] - tx * tile[0]), sy = min(tile[1], nc[1] - ty * tile[1]);
    return tile_first_cell(nc, tile, tx, ty, tz) + ((z - tz * tile[2]) * sy + (y - ty * tile[1])) * sx +
           (x - tx * tile[0]);
}

// Work items of batched launches.
struct WorkItem {
    int32_t sys;
    int32_t idx;  // first atom of a 256-atom chunk, or tile id
};

// Everything the allocation-free entry point puts into its static scratch array besides
// zeros, passed by value to one set-up kernel (no host-to-device copy): the descriptor
// (pointers into the scratch), where it goes, the work tables and the small tables.
struct NoAllocSetup {
    SysDev dev;
    SysDev* dev_slot;
    WorkItem* atom_work;
    WorkItem* tile_work;
    int32_t n_atom_work, n_tile_work;
    PairParam pp[kMaxTypes * kMaxTypes];
    double mass[kMaxTypes];
    double box[BOX_COUNT];
};

}  // namespace pis
