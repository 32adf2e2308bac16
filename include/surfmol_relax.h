/*
 * surfmol_relax.h -- C ABI of the MI355X (gfx950) multi-system relaxation
 * library libsurfmol_relax.so.
 *
 * This is the drop-in boundary for SurfMol's relaxation path: FIRE / damped
 * MD driving UFF bonded forces, intra-molecular LJ + Coulomb pairs and the
 * tricubic B-spline GridFF surface.  Every entry point names the reference
 * interface it replaces (paths relative to the reference repository root).
 *
 * Conventions
 *  - Return value: SMR_OK (0) or an SMR_ERR_* code; smr_last_error() returns
 *    a thread-local message with the offending values.  Nothing aborts
 *    across the ABI (the reference's ffi_uff.rs panics instead:
 *    crates/libs/molff/src/ffi_uff.rs:22-37).
 *  - Ownership: the caller owns every host array passed in or out (plain
 *    pointers + sizes, C-contiguous, double unless stated); the library owns
 *    device state behind opaque handles.
 *  - Threading: a context, and everything created from it, is used from one
 *    thread at a time.  The serving loop (smr_server_*) owns an internal
 *    worker thread; its entry points may be called from any thread.
 *  - Units: eV, Angstrom, elementary charge, amu (all masses 1), time unit
 *    sqrt(amu*A^2/eV) = 10.1805 fs.
 *  - Numerics: forces in FP32 with compensated (Kahan) per-atom sums;
 *    energies, FIRE reductions and convergence tests in FP64; GridFF build
 *    and B-spline fit in FP64.
 */
#ifndef SURFMOL_RELAX_H
#define SURFMOL_RELAX_H

#ifdef __cplusplus
extern "C" {
#endif

#define SMR_OK 0
#define SMR_ERR_INVALID 1     /* bad argument (message says which)            */
#define SMR_ERR_HIP 2         /* HIP runtime error                            */
#define SMR_ERR_NOMEM 3       /* device or host allocation failed             */
#define SMR_ERR_UNSUPPORTED 4 /* valid request outside what is built          */
#define SMR_ERR_STATE 5       /* call not valid in the current object state   */

/* per-system status codes reported by relax / serve */
#define SMR_SYS_RUNNING 0
#define SMR_SYS_CONVERGED 1
#define SMR_SYS_MAXSTEPS 2    /* step cap reached before Fmax < fconv          */
#define SMR_SYS_NONFINITE 3   /* a force or position became NaN/Inf            */

typedef struct smr_ctx smr_ctx;
typedef struct smr_gridff smr_gridff;
typedef struct smr_batch smr_batch;
typedef struct smr_server smr_server;

/* Per-thread message of the last failing call ("" if none). */
const char *smr_last_error(void);
/* Library version string. */
const char *smr_version(void);

/* ---------------------------------------------------------------- context */
/* Replaces oclff::nvidia_proque + UffOcl::new
 * (crates/libs/oclff/src/lib.rs:15-23, crates/libs/oclff/src/uff.rs:18-20). */
int smr_ctx_create(int device, smr_ctx **out);
int smr_ctx_destroy(smr_ctx *ctx);
/* Device facts: out[5] = (free HBM bytes, total HBM bytes, compute units,
 * LDS bytes per workgroup, wavefront size).  No counterpart in the
 * reference. */
int smr_ctx_info(smr_ctx *ctx, double *out);

/* ----------------------------------------------------------------- GridFF */
typedef struct {
    int nx, ny, nz;          /* nodes; x,y periodic, z open                  */
    double p0[3];            /* position of node (0,0,0)                      */
    double step[3];          /* node spacing (cell = n*step)                  */
    double alpha_morse;      /* GFFParams.y: Morse decay [1/A]                */
    double rdamp;            /* GFFParams.x: Coulomb damping radius [A]       */
    int keep_reference;      /* smr_gridff_build: keep the FP64 reference grid */
} smr_grid_desc;

/* Upload fitted B-spline coefficients (FP32, z-fastest, 4 channels
 * Pauli/London/Coulomb/0 per node: index iz + nz*(iy + ny*ix)), the layout of
 * GridFF.h pack_Bspline_d consumed by fe3d_pbc_comb
 * (opencl/gridff_eval.cl:234-286).  Replaces the BsplinePLQ buffer argument
 * of getNonBond_GridFF_Bspline (opencl/UFF.cl:1524-1541). */
int smr_gridff_upload(smr_ctx *ctx, const smr_grid_desc *desc, const float *coefs, smr_gridff **out);

/* Build the reference grid on the device in FP64 and fit the B-spline
 * coefficients in FP64, then store them as FP32.  Reference grid:
 * make_GridFF (opencl/gridff_build.cl:908-1062): node (ix,iy,iz) sits at
 * p0 + (ix*Lx/nx, iy*Ly/ny, iz*Lz/nz) with the cell L = n*step, and takes
 * the sum over substrate atoms and their images shifted by (jx*Lx, jy*Ly, 0),
 * |jx|,|jy| <= npbc (the reference's fixed image set), of
 *     Pauli  += sqrtE*exp(-2a(r-R)),  London += -2 sqrtE*exp(-a(r-R)),
 *     Coul   += kC*Q/sqrt(r^2+rdamp^2).
 * Fit: (Sx (x) Sy (x) Sz) C = E with the stencil [1/6,2/3,1/6], periodic in
 * x,y and truncated in z (BsplineConv3D, opencl/gridff_build.cl:8-61).
 * sub_reqs[4*i] = (R, sqrt(E), Q, unused). */
int smr_gridff_build(smr_ctx *ctx, const smr_grid_desc *desc, int nsub, const double *sub_pos,
                     const double *sub_reqs, int npbc, smr_gridff **out);
/* Copy the FP32 coefficients back (nx*ny*nz*4 floats). */
int smr_gridff_download(const smr_gridff *g, float *coefs);
/* Copy the FP64 reference-grid values (before the fit) back (nx*ny*nz*4);
 * available only for grids made by smr_gridff_build with keep_reference. */
int smr_gridff_download_reference(const smr_gridff *g, double *values);
int smr_gridff_destroy(smr_gridff *g);

/* On-disk coefficient cache: the reference's lazy file cache of fitted grids
 * (GridFF.h tryLoad_new reads Bspline_PLQd.npy and refits only when it is
 * missing; notes/designs/2026-08-29_gridff_faf_porting_notes.md:393-394).
 * Files are .npy, float32, shape (nx, ny, nz, 4) in C order -- the node
 * layout above -- so numpy.load reads them.
 * smr_gridff_save writes the grid's FP32 coefficients (to path.tmp, then
 * renamed).  smr_gridff_load reads such a file into a new grid; the file's
 * shape must match desc.  smr_gridff_build_cached keys a file in cache_dir
 * by a hash of every build input (desc, npbc, substrate positions and REQ
 * rows): it loads the file when present and valid (*hit = 1), else builds
 * as smr_gridff_build, writes the file and sets *hit = 0 (hit may be NULL).
 * A grid with keep_reference is always built. */
int smr_gridff_save(const smr_gridff *g, const char *path);
int smr_gridff_load(smr_ctx *ctx, const smr_grid_desc *desc, const char *path, smr_gridff **out);
int smr_gridff_build_cached(smr_ctx *ctx, const smr_grid_desc *desc, int nsub, const double *sub_pos,
                            const double *sub_reqs, int npbc, const char *cache_dir, int *hit, smr_gridff **out);

/* ----------------------------------------------------------- systems */
/* One molecule's force-field tables: exactly the public SoA fields of the
 * reference's molff::uff::Uff (crates/libs/molff/src/uff.rs:44-85), the
 * neighbour tables of moltopo::molecular::Atoms (molecular.rs:7-13) and
 * NonBondedFF.reqs (crates/libs/molff/src/nonbonded.rs:48).  Indices are
 * 0-based and local to the molecule; *_ngs are hneigh slot indices
 * (4*atom + slot) as produced by Uff::bake_*_neighs (uff.rs:250-315). */
typedef struct {
    int natoms, nbonds, nangles, ndih, ninv;
    const int *neighs;        /* [4*natoms], -1 padded                        */
    const int *neigh_bs;      /* [4*natoms], bond index per slot              */
    const double *bon_params; /* [2*nbonds]  (k, l0)                          */
    const int *ang_atoms;     /* [3*nangles] (i, j, k), j central             */
    const int *ang_ngs;       /* [2*nangles]                                  */
    const double *ang_params; /* [5*nangles] (k, c0, c1, c2, c3)              */
    const int *dih_atoms;     /* [4*ndih]                                     */
    const int *dih_ngs;       /* [3*ndih]                                     */
    const double *dih_params; /* [3*ndih]    (V, d, n)                        */
    const int *inv_atoms;     /* [4*ninv]    i central                        */
    const int *inv_ngs;       /* [3*ninv]                                     */
    const double *inv_params; /* [4*ninv]    (K, C0, C1, C2)                  */
    const double *reqs;       /* [4*natoms]  (RvdW, sqrt(EvdW), Q, Hb)        */
} smr_topology;

typedef struct {
    double rdamp;             /* Coulomb damping radius of the pair term [A]   */
    int nonbonded;            /* 1: intra-molecular LJ+Coulomb pairs (getNonBond,
                                 opencl/UFF.cl:1023-1204: 1-2 exclusions,
                                 E0 = sqrtEi*sqrtEj); 0: off                   */
    int surface;              /* 1: GridFF surface term; 0: off                */
    double inter_cutoff;      /* 0: the systems are independent molecules.  > 0:
                                 the batch is one coupled layer -- pairs of atoms
                                 of different systems interact (LJ + Coulomb, no
                                 exclusions) when closer than inter_cutoff [A],
                                 found through a device cell list, and one FIRE
                                 state drives all systems (the broad-phase culling
                                 of NonBondedFF::eval_broad, molff/src/nonbonded.rs:
                                 251-324; spacc/src/uniform_grid.rs:87-116).      */
    double inter_switch;      /* coupled layer: 0 = hard cutoff (the reference's
                                 culling); w > 0 = each intermolecular pair energy
                                 is multiplied by S = 1 - 3x^2 + 2x^3,
                                 x = (r - (inter_cutoff - w)) / w, over the last w
                                 below the cutoff, so energy and forces stay
                                 continuous when pairs cross it                    */
    double pbc_lx, pbc_ly;    /* coupled layer: lateral periodic box [A] (0: open
                                 along that axis).  Every atom pair interacts in
                                 every periodic image within the cutoff, a
                                 molecule with its own images included; only an
                                 atom pair of one molecule in the same image is
                                 the intramolecular pair (the 3x3 image loop of
                                 getNonBond's PBC branch, opencl/UFF.cl:1154-1182;
                                 NonBondedFF::eval_pbc, molff/src/nonbonded.rs:
                                 376-418, culled at the cutoff).  A box length
                                 must be >= 3 cutoffs, a whole number of GridFF
                                 node steps (1 A without a surface) and of GridFF
                                 periods when there is a surface.                 */
} smr_ff_options;

/* Create a batch of nsys systems.  Systems reference one of ntopo
 * topologies (sys_topo[s] in [0,ntopo)), so replicas share device tables
 * (the per-replica/shared split of oclff::RRsp3Multi,
 * crates/libs/oclff/src/rrsp3.rs:742-839).  apos holds the concatenated
 * initial positions [sum_s natoms(s), 3].  grid may be NULL when
 * opts->surface == 0.  Replaces the per-call buffer set-up of
 * UffOcl::eval_bonds (crates/libs/oclff/src/uff.rs:30-58). */
int smr_batch_create(smr_ctx *ctx, const smr_gridff *grid, int ntopo, const smr_topology *topos, int nsys,
                     const int *sys_topo, const double *apos, const smr_ff_options *opts, smr_batch **out);
int smr_batch_destroy(smr_batch *b);
/* Total atoms in the batch; per-system atom offsets (nsys+1 ints). */
int smr_batch_natoms(const smr_batch *b, int *total_atoms);
int smr_batch_atom_offsets(const smr_batch *b, int *offsets);
int smr_batch_set_positions(smr_batch *b, const double *apos);
/* Restore the positions given at smr_batch_create (kept on the device) and
 * zero the velocities, without any host transfer.  No counterpart in the
 * reference; used to repeat a relaxation from the same inputs. */
int smr_batch_reset(smr_batch *b);
int smr_batch_get_positions(const smr_batch *b, double *apos);
/* Per-step parity hook: forces [sum N,3] and per-system energy terms
 * [nsys,6] = (bond, angle, dihedral, inversion, nonbonded, surface), with the
 * reference's bookkeeping (bond and pair energies counted at both ends,
 * angle energy without the c1*cos(theta) term: uff.rs:462,482-493,
 * nonbonded.rs:249).  Replaces MolWorld::eval_forces
 * (crates/libs/surfmol/src/mol_world.rs:60-86) and the force half of the
 * OpenCL launch sequence (opencl/UFF.cl:141-880,1524-1718). */
int smr_batch_eval_forces(smr_batch *b, double *forces, double *energies);

typedef struct {
    double dt0, dt_max;       /* FireState::new(dt0, dt_max) (molff/src/raff.rs:1057) */
    double alpha0;            /* 0.1  */
    int n_min;                /* 5    */
    double f_inc, f_dec, f_alpha; /* 1.1, 0.5, 0.99 */
    double flim;              /* <= 0: no clamp; > 0: clamp |F| used by the step (raff.rs:1093) */
} smr_fire_params;

typedef struct {
    int steps;                /* force evaluations done                        */
    int status;               /* SMR_SYS_*                                     */
    double energy;            /* total energy at the last force evaluation     */
    double fmax;              /* max_i |F_i| at the last force evaluation      */
    double terms[6];          /* energy terms, as smr_batch_eval_forces        */
} smr_relax_result;

/* Batched relaxation: every system of the batch runs the FIRE loop of
 * surfmol/tests/relax_pentacene_uff.rs:142-193 (FireState semantics,
 * raff.rs:1040-1062) until max_i|F_i| < fconv or max_steps, all systems in
 * one persistent kernel (one workgroup per system).  Replaces FireCore's
 * multi-relax call over relax_multi.cl (opencl/relax_multi.cl:1-45,
 * 1089-1413, 1916-2086).  results: [nsys]; positions are left on the device
 * (smr_batch_get_positions). */
int smr_relax_batched(smr_batch *b, const smr_fire_params *fire, int max_steps, double fconv,
                      smr_relax_result *results);

/* Plain method: relax one system of the batch with one launch sequence per
 * FIRE step (bonds -> angles/dihedrals/inversions -> assemble + pairs +
 * surface -> per-system (v.F, |v|^2, |F|^2, Fmax) reduction) and the FIRE
 * decision taken on the host after reading the reductions back each step:
 * the reference's per-step shape (opencl/UFF.cl:141-1011 kernels driven by a
 * host loop reading cvf, opencl/relax_multi.cl:1384). */
int smr_relax_plain(smr_batch *b, int isys, const smr_fire_params *fire, int max_steps, double fconv,
                    smr_relax_result *result);

/* smr_relax_plain without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during the
 * call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same kernels, results and errors as smr_relax_plain; the
 * kernels' scratch (hneigh, force pieces, per-atom energies, reduction output)
 * is a fixed device array of 600 KiB per GPU, kept until the process ends,
 * instead of buffers allocated on the batch.  A system with more than 1024
 * atoms or more than 32768 force pieces (4 per dihedral and inversion, 3 per
 * angle) is rejected with SMR_ERR_UNSUPPORTED.  Calls on one GPU share that
 * array: calls that overlap on one GPU race. */
int smr_relax_plain_noalloc(smr_batch *b, int isys, const smr_fire_params *fire, int max_steps, double fconv,
                            smr_relax_result *result);

/* Launch variant of the fused kernel for this batch: threads per workgroup
 * (64/128/256/512; 0 = automatic by system size; atoms per thread follow),
 * the stored per-atom surface stencils (1 on, 0 off, -1 automatic: with one
 * atom per thread the stencil lives in registers during a launch and the
 * stored copy carries it across launches; with more atoms per thread the
 * stored copy is the cache) and the pair loop (1 unordered-pair sweep, 0
 * every ordered pair, -1 automatic).  All variants compute the same forces; results differ
 * only by FP32 summation order.  No counterpart in the reference (tuning
 * knob of this implementation). */
int smr_batch_set_launch(smr_batch *b, int block, int stencil_cache, int pair_sweep);

/* Launch shape and device time of the last fused launch on this batch:
 * shape4 = (threads per workgroup, atoms per thread, workgroups, LDS bytes
 * with bit 30 set when the stencil cache was on);
 * kernel_ms = HIP-event time of that launch on the context's stream (only
 * measured by smr_relax_batched).  No counterpart in the reference. */
int smr_batch_last_launch(const smr_batch *b, int *shape4, float *kernel_ms);

/* Persistent launches of the batch and their shapes.  Systems are grouped
 * by atom count into size classes (<= 64, 128, 256, 512, 1024), each with
 * its own launch shape; adjacent classes whose shapes agree share one
 * launch when it keeps the resident workgroups of each.  *nclass receives
 * the number of launches (at most 5) and info [8 * nclass] per launch:
 * (largest class id in it, systems, max atoms, threads per workgroup, atoms
 * per thread, workgroups, LDS bytes, flags: 1 force pieces in LDS,
 * 2 unordered-pair sweep, 4 stencil cache).  No counterpart in the
 * reference (its multi-system ABI has one fixed work-group shape,
 * opencl/relax_multi.cl:425-443). */
int smr_batch_plan(const smr_batch *b, int *nclass, int *info);

/* Autotuner: time every launch shape that can hold each size class of this
 * batch on a FIRE run of trial_steps steps and keep the fastest per class in
 * the context (used by later batches and servers of the context).  The
 * batch's positions, velocities and FIRE state are restored.  Results of a
 * relaxation depend on the launch shape only through FP32 summation order.
 * No counterpart in the reference. */
int smr_batch_autotune(smr_batch *b, const smr_fire_params *fire, int trial_steps);

/* 1 (default): capture the launch sequence of a relaxation (counter reset +
 * one persistent launch per size class on forked streams; for a coupled
 * layer, 16 steps of force launches and cell-list kernels) into a HIP graph
 * once and replay it while nothing changes; 0: issue the launches directly.
 * Results are the same either way.  No counterpart in the reference. */
int smr_batch_set_graph(smr_batch *b, int use_graph);

/* Coupled layers: 1 (default) run the whole relaxation as one cooperative
 * launch when the layer fits the chip at once (one 64-thread workgroup per
 * molecule of <= 64 atoms; per step the workgroups exchange positions,
 * velocities, forces and FIRE sums through device-coherent memory between
 * grid barriers, and take intermolecular partners from Verlet lists of
 * radius cutoff + skin_A, rebuilt when an atom may have moved skin_A / 2);
 * 0 use the step launches (force launch, decision, cell list).  Results
 * agree to FP32 summation order.  A layer that does not fit, or whose
 * neighbourhoods overflow the lists, runs the step launches.  skin_A <= 0
 * keeps the current skin (default 1 A).  No counterpart in the reference. */
int smr_batch_set_layer_persistent(smr_batch *b, int on, double skin_A);

/* Which path the last coupled-layer relaxation took: *persistent 1 for the
 * one-launch path; *fallback the reason it was not (0 none or switched off,
 * 1 a grid barrier timed out, 2/3/4 neighbour instances / staged atoms /
 * list entries overflowed, 5 the layer does not fit the chip at once, 6 its
 * molecules need another launch shape); *rebuilds the neighbour-list builds.
 * Each pointer may be NULL.  No counterpart in the reference. */
int smr_batch_layer_path(const smr_batch *b, int *persistent, int *fallback, int *rebuilds);

/* Order in which smr_relax_batched starts the systems of a size class:
 * 1 (default) larger systems first and, among systems of one size, highest
 * FP64 energy per atom of the starting positions first
 * (one energy evaluation before the launch; systems far from a minimum need
 * the most FIRE steps, and starting them first shortens the launch's tail);
 * 0 batch order.  Results do not depend on it.  No counterpart in the
 * reference. */
int smr_batch_set_order(smr_batch *b, int energy_order);

/* Re-ranking of a relaxation's schedule (with the energy order on, and where
 * a size class has more systems than resident workgroups): smr_relax_batched
 * runs two launches.  Every system first takes up to `steps` FIRE steps; the
 * unfinished ones are then ranked again by the FP64 energy per atom of their
 * current positions and continue from their persisted state (positions,
 * velocities, FIRE state, stored stencils).  The energy after a few hundred
 * steps ranks the remaining step counts better than the starting energy, so
 * long relaxations start earlier in the continuation.  -1 (default): after
 * max_steps / 5 steps (one launch when that is below 50); 0: one launch.
 * Results do not depend on it.  No counterpart in the reference. */
int smr_batch_set_rerank(smr_batch *b, int steps);

/* Damped MD of MolWorld::run_md (crates/libs/surfmol/src/mol_world.rs:141-164,
 * moltopo/src/molecular.rs:118-137) for every system: clamp |F| to flim,
 * v = (1-damping) v + F dt, x += v dt, v = 0 when sum v.F < 0, stop when
 * sum|F|^2 < fconv^2.  steps[nsys] receives the steps taken. */
int smr_run_md(smr_batch *b, int niter, double dt, double fconv, double flim, double damping, int *steps);

/* ----------------------------------------------------------- serving */
typedef struct {
    int max_resident;         /* slots: systems on the device at once           */
    int max_atoms;            /* largest request accepted (<= 1024)             */
    double hbm_fraction;      /* device memory for resident requests, as a
                                 fraction of the HBM free at creation (0,1]     */
    int chunk_steps;          /* FIRE steps per persistent launch of requests of
                                 <= 64 atoms; size class c (64, 128, 256, 512,
                                 1024 atoms) takes chunk_steps*2/(2+c)          */
    int max_steps;            /* step cap per request                           */
    double fconv;             /* convergence threshold [eV/A]                   */
    smr_fire_params fire;
    smr_ff_options ff;
} smr_server_config;

/* Continuous-batching relax server.  Requests of any size (up to
 * max_atoms) are queued by smr_server_submit; a worker thread admits them
 * into free slots while the device arenas (sized from hbm_fraction of free
 * HBM) hold their topology and state, relaxes every resident request by
 * chunked persistent launches (one launch per size class; the larger
 * classes take fewer steps per chunk), and retires each request as soon as it converges (or hits
 * max_steps), with FP64 energies of its final positions.  Requests that
 * arrive while others run are admitted at the next chunk.  A request's
 * result does not depend on what else is resident (its size class fixes
 * its launch shape).  No counterpart in the reference (it has no serving
 * loop); the relaxation itself is smr_relax_batched's. */
int smr_server_create(smr_ctx *ctx, const smr_gridff *grid, const smr_server_config *cfg, smr_server **out);
/* Submit one request: a topology (copied; the caller's arrays may be freed
 * after the call) and initial positions [natoms,3].  *id receives the
 * request id.  Thread-safe.  SMR_ERR_NOMEM if the request can never fit the
 * arenas. */
int smr_server_submit(smr_server *s, const smr_topology *topo, const double *apos, long long *id);
/* Poll a request.  *done = 1 when it has retired: result, final positions
 * ([natoms,3], may be NULL) and the submit / retire times (seconds on the
 * server's steady clock) are filled and the request is forgotten (a later
 * poll of the id fails).  *done = 0 while it is queued or resident. */
int smr_server_poll(smr_server *s, long long id, int *done, smr_relax_result *result, double *apos_out,
                    double *t_submit, double *t_done);
/* Block until every submitted request has retired. */
int smr_server_drain(smr_server *s);
/* Counters: out[10] = (submitted, retired, resident, queued, chunks,
 * peak resident, arena bytes, arena bytes in use, atoms capacity, seconds
 * of device time in relax launches). */
int smr_server_stats(smr_server *s, double *out);
/* Stop the worker after its current chunk and free everything; requests
 * still queued or resident are dropped. */
int smr_server_destroy(smr_server *s);

#ifdef __cplusplus
}
#endif
#endif /* SURFMOL_RELAX_H */
