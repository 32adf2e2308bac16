/*
 * surfmol_raff.h -- C ABI of the multi-replica rigid-atom port relaxation
 * (RAFF "RRsp3" cluster step) in libsurfmol_relax.so.
 *
 * Independent replicas of one molecule relax by position-based (XPBD) port
 * constraints: node atoms carry an orientation quaternion and four port arms
 * (the directions to their bonded neighbours at the reference geometry), the
 * neighbours are pulled onto the port tips, and optional sphere collisions
 * push non-bonded atoms apart.  This is the drop-in for oclff::RRsp3Multi
 * (crates/libs/oclff/src/rrsp3.rs:699-1088, kernels opencl/RRsp3.cl:1774-
 * 2065): every entry point below names the method it replaces, takes the
 * same arrays in the same layout (float32, as the reference uploads them)
 * and runs the same step.  All replicas share one topology; the per-replica
 * state lives on the device.
 *
 * Conventions are those of surfmol_relax.h: SMR_OK or an SMR_ERR_* code,
 * smr_last_error() for the message, caller-owned host arrays, opaque
 * handles, one thread per handle.
 */
#ifndef SURFMOL_RAFF_H
#define SURFMOL_RAFF_H

#include "surfmol_relax.h"

#ifdef __cplusplus
extern "C" {
#endif

/* per-replica status of smr_raff_relax */
#define SMR_RAFF_RUNNING 0
#define SMR_RAFF_CONVERGED 1
#define SMR_RAFF_MAXSTEPS 2

#define SMR_RAFF_MAX_ATOMS 256   /* atoms per replica (padding included) */

typedef struct smr_raff smr_raff;

/* StepConfig (crates/libs/oclff/src/rrsp3.rs:54-73).  The multi-replica
 * step uses dt, relaxation, momentum_beta and rot_mass_scale; k_coll,
 * bbox_margin, n_rot_substeps, rot_eps, theta_max and damp belong to the
 * reference's other step variants and are carried unused, as there. */
typedef struct {
    float dt, k_coll, relaxation, bbox_margin, momentum_beta, rot_mass_scale;
    int n_rot_substeps;
    float rot_eps, theta_max, damp;
} smr_raff_step_config;

/* RRsp3Multi::new(natoms_per_sys, group_size, nsys, _) (rrsp3.rs:742-839).
 * natoms_per_sys counts the replica's atoms as packed (nodes first, then
 * caps, then padding atoms of inverse mass 0); at most SMR_RAFF_MAX_ATOMS.
 * The reference's group_size is a launch detail here (chosen by the
 * library).  Collision, momentum and fix-mask buffers start at zero, the fix
 * mask at 0, orientations at the identity. */
int smr_raff_create(smr_ctx *ctx, int natoms_per_sys, int nsys, smr_raff **out);
int smr_raff_destroy(smr_raff *r);

/* Shared topology uploads (rrsp3.rs:858-934), one row per atom of one
 * replica: radius [n] (>= 0; all zero disables collisions), neighs / excl1 /
 * excl2 / bk_slots [4n] (-1 padded, local indices; bk_slots index node ports
 * inode*4 + k), fixmask [n] (bits 1,2,4 freeze x,y,z and the node's
 * rotation, 8 pins z to 0), port_local_atoms [16n] (float4 per port, .xyz
 * used) and k_atoms [4n] of which the first nnode_per_group atoms' rows are
 * used. */
int smr_raff_upload_radius(smr_raff *r, const float *radius);
int smr_raff_upload_neighs_and_exclusions(smr_raff *r, const int *neighs, const int *excl1, const int *excl2);
int smr_raff_upload_fixmask(smr_raff *r, const int *fixmask);
int smr_raff_upload_cluster_ports(smr_raff *r, const float *port_local_atoms, const float *k_atoms,
                                  int nnode_per_group);
int smr_raff_upload_bk_slots(smr_raff *r, const int *bk_slots);

/* RRsp3Multi::upload_state_multi (rrsp3.rs:936-974): positions
 * pos3 [nsys*n*3], inverse masses [n] (shared; padding atoms 0, their
 * positions become NaN as in the reference) and orientations [nsys*n*4]
 * (x,y,z,w; NULL = identity).  Zeroes both heavy-ball momenta. */
int smr_raff_upload_state(smr_raff *r, const float *pos3, const float *inv_mass, const float *quat);

/* nsteps consecutive calls of RRsp3Multi::step_cluster_multi with the
 * Current port kernel (rrsp3.rs:1028-1050), as one persistent launch. */
int smr_raff_step(smr_raff *r, const smr_raff_step_config *cfg, int nsteps);

/* Relaxation with per-replica retirement (no counterpart in the reference,
 * whose benchmark loop steps every replica and samples the corrections of
 * replica 0, crates/libs/oclff/src/bin/raff_ocl_benchmark.rs:153-170): each
 * replica steps until the largest |component| of its step's collision and
 * node corrections is below tol, or until max_steps steps.  steps[nsys],
 * status[nsys] (SMR_RAFF_*) and metric[nsys] (the last step's largest
 * |component|) may be NULL. */
int smr_raff_relax(smr_raff *r, const smr_raff_step_config *cfg, int max_steps, float tol, int *steps, int *status,
                   float *metric);

/* Downloads (rrsp3.rs:976-1025,1051-1083): positions of all replicas
 * [nsys*n*3] or of one [n*3]; orientations of one replica [n*4]; the last
 * step's collision corrections [n*3] and node translations [nnode*3]; the
 * largest |component| of each replica's last step corrections [nsys]. */
int smr_raff_download_pos(const smr_raff *r, float *pos3);
int smr_raff_download_pos_replica(const smr_raff *r, int isys, float *pos3);
int smr_raff_download_quat_replica(const smr_raff *r, int isys, float *quat4);
int smr_raff_download_dpos_coll_replica(const smr_raff *r, int isys, float *d3);
int smr_raff_download_dpos_node_replica(const smr_raff *r, int isys, float *d3);
int smr_raff_download_metric(const smr_raff *r, float *metric);

/* ----------------------------------------------------------- serving */
typedef struct smr_raff_server smr_raff_server;
typedef struct {
    int capacity;              /* replicas resident at once (slots)              */
    int chunk_steps;           /* steps per persistent launch                     */
    int max_steps;             /* step cap per request                            */
    float tol;                 /* retire when the step's largest |correction
                                  component| < tol (as smr_raff_relax)            */
    smr_raff_step_config step;
} smr_raff_server_config;

/* Continuous-batching relax server for replicas of one molecule (no
 * counterpart in the reference, whose multi-replica harness steps a fixed
 * set).  The template supplies the shared topology and the inverse masses
 * (its first replica's): its uploads must be complete.  A worker thread
 * admits queued requests into free slots, relaxes every resident replica by
 * chunked persistent launches and retires each one when it converges or
 * reaches max_steps; requests that arrive meanwhile join at the next chunk.
 * A request's result equals smr_raff_relax's on a one-replica set. */
int smr_raff_server_create(const smr_raff *tmpl, const smr_raff_server_config *cfg, smr_raff_server **out);
/* One replica's positions [n*3] (float32, packed order) and orientations
 * [n*4] (NULL = identity).  Thread-safe. */
int smr_raff_server_submit(smr_raff_server *s, const float *pos3, const float *quat, long long *id);
/* *done = 1 when the request has retired: steps, status (SMR_RAFF_*), the
 * last step's largest |correction component|, final positions [n*3] and the
 * submit / retire times (seconds on the server's steady clock) are filled
 * (each may be NULL) and the request is forgotten. */
int smr_raff_server_poll(smr_raff_server *s, long long id, int *done, int *steps, int *status, float *metric,
                         float *pos3, double *t_submit, double *t_done);
/* Block until every submitted request has retired. */
int smr_raff_server_drain(smr_raff_server *s);
/* out[7] = (submitted, retired, resident, queued, chunks, peak resident,
 * seconds of device time in chunks). */
int smr_raff_server_stats(smr_raff_server *s, double *out);
int smr_raff_server_destroy(smr_raff_server *s);

#ifdef __cplusplus
}
#endif
#endif /* SURFMOL_RAFF_H */
