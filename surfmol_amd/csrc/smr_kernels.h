// smr_kernels.h -- launch interface between capi.cpp and relax_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/surfmol_relax.h"
#include "../../include/surfmol_raff.h"
#include "smr_device.h"
#include "sweep_schedules.h"

namespace smr {

enum { KMODE_FIRE = 0, KMODE_EVAL = 1, KMODE_MD = 2, KMODE_LAYER = 3, KMODE_LAYER_P = 4 };

// Cell list of a coupled layer (atoms sorted by cell, by atom index inside a
// cell), rebuilt every step.  Cells are at least as large as the cutoff, so
// the 27 neighbouring cells hold every partner within it.
struct CellGrid {
    const float4 *spos;      // [natoms] local xyz of the atom, w = its system (int bits)
    const float4 *sreq;      // [natoms] REQ rows in the same order
    const int *start;        // [ncells + 1]
    const float4 *part;      // [K][natoms] intermolecular force of each atom (by sorted slot) per neighbour cell
    int nox, noy, noz;       // neighbour offsets per axis (3, or 1 for a single-cell axis); K = nox*noy*noz
    const int *slot_of;      // [natoms] sorted slot of each atom
    int natoms;
    int ncx, ncy, ncz;
    float lox, loy, loz;     // lower corner in the origin-grid frame
    float inv_h;             // 1 / cell size (z; x and y unless periodic)
    float inv_hx, inv_hy;    // 1 / cell size along x, y
    // lateral periodic box (pbc_x/pbc_y: 0 open, 1 periodic): lengths, their
    // inverses and their length in origin-grid steps (an exact integer), and
    // the wrap count of each atom by sorted slot
    int pbc_x, pbc_y;
    float Lx, Ly, inv_Lx, inv_Ly;
    int nLx, nLy;
    const int2 *swrap;
    float rc2;               // cutoff squared (force test, FP32)
    double rc2_64;           // cutoff squared (energy test, FP64)
    float ron, ron2, inv_w;  // switch: start, start squared, 1 / width (ron2 = rc2: hard cutoff)
    double ron64, w64;
};

// FIRE state of a coupled layer (one for all its systems) and the decision
// of the last step for the move kernel.
struct LayerState {
    double dt, alpha;
    int npos, steps, status, nsys;
    double fmax;
    float mix_a, mix_b, dt_move;
    int reset;
    int pending_move;        // the decision of this step still has to be applied
    int cells_dirty;         // an atom changed cell in this step's move: the cell list is rebuilt
    int rebuilds;            // neighbour-list builds of the one-launch path (a statistic)
};

// A coupled layer in one cooperative launch (KMODE_LAYER_P): every molecule
// keeps one workgroup for all steps of the launch; the workgroups exchange
// per step, through device-coherent (sc1) stores and loads, each atom's
// position, velocity and force and each molecule's FIRE record, separated by
// a grid barrier; intermolecular partners come from per-atom Verlet lists
// over neighbour atoms staged in LDS.
constexpr int kLpMaxInst = 64;     // neighbour (molecule, box image) instances per molecule
constexpr int kLpGroups = 8;       // arrival counters of the grid barrier (128 B apart)
constexpr int kLpBarWords = (kLpGroups + 2) * 32;   // counters, top counter, error word
constexpr int kLpRec = 8;          // doubles per molecule record
enum { LP_OK = 0, LP_ERR_BARRIER = 1, LP_ERR_INST = 2, LP_ERR_STAGE = 3, LP_ERR_LIST = 4 };
struct LayerPersist {
    float4 *xbuf;            // [2][3][natoms] position (before the move), velocity (after v += F dt), force; by step parity
    double *xrec;            // [2][nsys][kLpRec] v.F, v.v, F.F, max F^2, max v^2, max displacement^2, sphere (4 floats)
    unsigned *bar;           // barrier words: 8 group counters (128 B apart), the top counter, the error word
    int maxl, maxs;          // list entries per atom (lists in LDS, entry-major), staged atoms per molecule
    int lds_off;             // byte offset of the staging area in dynamic LDS
    int natoms;
    float skin;              // list radius = cutoff + skin; lists are rebuilt when an atom may have moved skin/2
};
// the lists are evaluated in chunks of kLpChunk entries spread over the wave
constexpr int kLpChunk = 16;
__host__ __device__ inline int lp_chunks(int maxl) { return (maxl + kLpChunk - 1) / kLpChunk; }
// dynamic LDS of the staging area and the lists (lds_atoms atoms of maxl
// entries, their chunk partial sums and chunk owners)
inline size_t lp_lds_bytes(int maxs, int lds_atoms, int maxl) {
    const size_t items = (size_t)lds_atoms * lp_chunks(maxl);
    const size_t lists = ((size_t)lds_atoms * maxl * 2 + 15) & ~(size_t)15;
    return (size_t)maxs * (16 + 16 + 4) + (size_t)kLpMaxInst * (16 + 16 + 4 + 4) + 16 + 64 * 8 + items * 16 +
           ((items + 15) & ~(size_t)15) + lists;
}

struct KernelArgs {
    DevTables tab;
    FireConst fire;
    float4 *pos;             // [total atoms] local coordinates
    float4 *vel;
    float4 *force;           // written in KMODE_EVAL
    SysState *state;         // [nsys]
    const int *work_list;    // systems to process
    int nwork;
    int *work_counter;       // atomic fetch counter (zeroed before launch)
    float *fint_scratch;     // [grid * 3 * nf_max] when pieces do not fit LDS
    float4 *stencil;         // [total atoms * 16] stored surface stencils (NULL: sample the grid directly)
    int nf_max;
    int lds_atoms;           // atoms reserved per LDS slot (relax_lds_atoms: >= 16 per wave)
    int lds_na, lds_nd, lds_ni;  // interaction slots reserved per LDS slot
    int steps_per_launch;
    int mode;
    int pair_mode;           // 1: unordered-pair sweep, per-wave reaction buffers; 0: ordered pairs
    int *stencil_key;        // [total atoms] cell of each stored stencil (-1: none), kept across launches
    CellGrid cells;          // coupled layer: intermolecular partners (spos == NULL: none)
    LayerState *layer;       // coupled layer: shared FIRE state (KMODE_LAYER)
    double4 *partials;       // coupled layer: per-system (v.F, v.v, F.F, max F^2)
    LayerPersist lp;         // coupled layer in one cooperative launch (KMODE_LAYER_P)
};

struct LaunchShape {
    int block = 0, apt = 0, grid = 0;
    bool fint_lds = false;
    bool stencil_cache = true;    // per-atom stored surface stencil (re-read on a cell change)
    bool pair_mode = true;        // unordered-pair sweep (needs a reaction buffer per wave)
    size_t lds_bytes = 0;
};

// atoms reserved per LDS slot: at least 16 per wave, so that the hneigh area
// (4 float4 per atom) also holds each wave's surface-stencil staging slots
// (4 atoms x 16 float4, used while hneigh is idle)
inline int relax_lds_atoms(int block, int max_atoms) { return max_atoms > 16 * (block / 64) ? max_atoms : 16 * (block / 64); }

// bytes of dynamic LDS for a launch shape: positions, REQ rows, hneigh,
// pair reaction forces (one buffer per wave in the unordered-pair sweep),
// reduction scratch, packed interaction slots and piece destinations (not
// for 4 atoms per thread: those tables stay in global memory), and the force
// pieces (optional)
// floats after a force-piece area that the gather may read (not use)
constexpr int kPieceTail = 8;

// Rows of the position and REQ arrays: one wave with one atom per thread
// (the register-rotation pair sweep) keeps copies of atoms after its own
// rows, so that the partner row of round t of lane v is row v + t without a
// wrap: for up to kSchedMaxAtoms atoms the rows 0..63+T of its schedule
// (sweep_schedules.h), for more atoms rows 0..N/2 again after row N-1; its
// reaction forces stay in registers, so it has no per-wave reaction buffer.
__host__ __device__ inline bool relax_rot_shape(int block, int apt) { return block == 64 && apt == 1; }
__host__ __device__ inline int relax_lds_rows(int block, int apt, int lds_atoms) {
    if (!relax_rot_shape(block, apt)) return lds_atoms;
    const int sched = 64 + kSchedTmax[lds_atoms < kSchedMaxAtoms ? lds_atoms : kSchedMaxAtoms];
    const int wrap = lds_atoms + lds_atoms / 2 + 1;
    return sched > wrap ? sched : wrap;
}

inline size_t relax_lds_bytes(int block, int apt, int lds_atoms, int nf_max, int na, int nd, int ni, bool fint_lds,
                              bool pair_mode) {
    const int nw = block / 64;
    const int rows = relax_lds_rows(block, apt, lds_atoms);
    const int fp = relax_rot_shape(block, apt) ? 0 : (pair_mode ? nw : 1);
    size_t b = (size_t)16 * (2 * rows + (size_t)lds_atoms * (4 + fp)) + (size_t)nw * 8 * 8 + 16;
    const bool tables = apt < 4;
    // slots (8, 8, 4 B) and piece destinations (8 B) per dihedral, inversion,
    // angle (one wave with one atom per thread reads the inversions'
    // destinations through the caches: its LDS holds the sweep's rows)
    if (tables) b += (size_t)nd * 16 + (size_t)ni * (relax_rot_shape(block, apt) ? 8 : 16) + (size_t)na * 12;
    // force pieces (x, y, z arrays of nf_max) and the 8-float tail that the
    // gather's rounds of eight may read past the last piece
    if (fint_lds) b += (size_t)nf_max * 12 + kPieceTail * sizeof(float);
    return b;
}

struct PlainArgs {
    DevTables tab;
    int topo;
    int iux, iuy;
    float4 *pos, *vel, *force;   // this system's slices
    float4 *hneigh;              // [4N]
    float4 *fint;                // [nf] (xyz, energy in piece 0 of each term)
    float *eatom;                // [N] bond energy per atom
    double *epair;               // [N]
    float *esurf;                // [N]
    double *cvf;                 // [10] reduction output
};

hipError_t launch_relax(const KernelArgs &A, const LaunchShape &L, hipStream_t stream);
// the coupled layer in one cooperative launch (64 threads per molecule, one
// atom per thread); hipErrorCooperativeLaunchTooLarge when the molecules do
// not fit the chip at once
hipError_t launch_relax_layer_persistent(const KernelArgs &A, int grid, size_t lds, hipStream_t stream);
const void *relax_layer_persistent_ptr();
const void *relax_kernel_ptr(int block, int apt, bool fint_lds);
hipError_t launch_zero_i32(int *p, int n, hipStream_t st);
// device-to-device copies described by a list (serving loop)
struct CopyDesc {
    unsigned long long src, dst, bytes;   // device addresses; bytes even, 2-byte aligned
};
hipError_t launch_copy_desc(const CopyDesc *d, int n, hipStream_t st);
// FP64 energies (state[s].E) of the listed systems at their current positions
hipError_t launch_energy64(const DevTables &T, const float4 *pos, SysState *state, const int *list, int n,
                           hipStream_t st);
// coupled layer: cell-list build, move, inter-molecular FP64 energies
struct LayerArgs {
    DevTables tab;
    float4 *pos, *vel;
    const float4 *force;
    const int *atom_sys;     // [natoms]
    const float4 *req_atom;  // [natoms]
    int natoms, nsys;
    int *count, *start, *sidx;   // [ncells] (counts / fill cursors beyond the LDS build), [ncells+1], [natoms]
    float4 *spos, *sreq;
    int *atom_cell;          // [natoms]
    int2 *atom_wrap;         // [natoms] box wraps of each atom (periodic layer)
    int2 *swrap;             // [natoms] the same by sorted slot
    float4 *part;            // [K][natoms]
    int *slot_of;            // [natoms]
    CellGrid cells;
    LayerState *layer;
    int *work_counters;      // [nwork_counters] reset by the cell build for the next force launch (or NULL)
    int nwork_counters;
    const double4 *partials; // [nsys] per-molecule FIRE sums of the step's force launch
    FireConst fire;          // FIRE parameters of a relaxation (k_layer_decide)
};
hipError_t launch_layer_cells(const LayerArgs &L, bool move, hipStream_t st);
hipError_t launch_layer_decide(const LayerArgs &L, hipStream_t st);
hipError_t launch_energy64_inter(const LayerArgs &L, SysState *state, const int *list, int n, hipStream_t st);
hipError_t launch_plain_step_forces(const PlainArgs &P, int natoms, int nang, int ndih, int ninv, hipStream_t st);
hipError_t launch_plain_reduce(const PlainArgs &P, double dt, double flim, hipStream_t st);
hipError_t launch_plain_apply(const PlainArgs &P, int natoms, float a, float b, int reset, float dt, hipStream_t st);

// multi-replica rigid-atom port relaxation (raff_kernels.hip)
struct RaffArgs {
    int natoms, nnode, nsys, nsteps;
    // shared topology (local atom index)
    const float *radius;       // [natoms]
    const int4 *neighs, *excl1, *excl2, *bk_slots;   // [natoms]
    const int *fixmask;        // [natoms] or NULL
    const float4 *port_local;  // [nnode*4]
    const float *kflat;        // [nnode*4]
    // per-replica state [nsys*natoms] and correction buffers
    float4 *pos, *quat, *dpos_mom, *dquat_mom, *dpos_coll;
    float4 *dpos_node, *drot_node;   // [nsys*nnode]
    float4 *dpos_neigh;              // [nsys*nnode*4]
    float *metric;                   // [nsys] max |component| of the last step's corrections
    int *steps, *status;             // [nsys] relax mode
    float dt, rot_mass_scale, relaxation, beta, tol;
    int collide, relax, max_steps;
};
hipError_t launch_raff(const RaffArgs &A, int seg, hipStream_t st);

// GridFF build / fit (gridff_kernels.hip)
struct GridBuildArgs {
    int nx, ny, nz;
    double p0[3];
    double L[3];             // cell lengths (nx*step, ...)
    int npbc;
    double alpha, r2damp;
    int nsub;
    const double *sub_pos;   // [nsub*3]
    const double *sub_req;   // [nsub*4]
    double *out;             // [nx*ny*nz*4] z-fastest
};
hipError_t launch_gridff_build(const GridBuildArgs &G, hipStream_t st);
hipError_t launch_bspline_fit(double *E, int nx, int ny, int nz, double *work, hipStream_t st);
hipError_t launch_pack_f32(const double *C, float *out, size_t n4, hipStream_t st);

}  // namespace smr
