/*
 * surfmol_host.h -- host-side set-up of the relaxation path (part of
 * libsurfmol_relax.so).
 *
 * C++ mirror of the reference's Rust set-up API, used by callers that do not
 * have the Rust crates: topology enumeration, neighbour tables, UFF typing,
 * UFF parameter assignment and REQH lookup.  The result is an smr_topology
 * (surfmol_relax.h) filled exactly as the reference's Uff/NonBondedFF fields
 * would be.  A Rust caller keeps its own crates and fills smr_topology from
 * them directly.
 */
#ifndef SURFMOL_HOST_H
#define SURFMOL_HOST_H

#include "surfmol_relax.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SMH_OK 0
#define SMH_ERR 1

typedef struct smh_molecule smh_molecule;

const char *smh_last_error(void);

/* Build one molecule.  param_dir holds ElementTypes.dat and AtomTypes.dat
 * (the format of the reference's data/ElementTypes.dat and data/AtomTypes.dat,
 * moltopo/src/params.rs:130-258).
 * Bonds: explicit (nbonds > 0, [2*nbonds]) or, when nbonds == 0, every pair
 * closer than rcut_bond (build_bonds_by_cutoff, moltopo/src/topology.rs:101).
 * Then angles/dihedrals/inversions (topology.rs:113-182), neighbour tables
 * (molecular.rs:34-78), UFF types (assign_uff.rs:8-118), parameters
 * (molff/src/uff.rs:699-887), hneigh slots (uff.rs:250-315) and REQH
 * (params.rs:595-604).  charges (may be NULL) overrides REQH.Q per atom.
 * Fails (no abort) on more than 4 bonds per atom. */
int smh_molecule_build(const char *param_dir, int natoms, const char *const *elems, const double *apos,
                       const double *charges, int nbonds, const int *bonds, double rcut_bond,
                       smh_molecule **out);
/* (natoms, nbonds, nangles, ndih, ninv) */
int smh_molecule_counts(const smh_molecule *m, int *counts5);
/* Fill an smr_topology whose pointers stay valid while m lives. */
int smh_molecule_topology(const smh_molecule *m, smr_topology *t);
/* Bond list [2*nbonds]. */
int smh_molecule_bonds(const smh_molecule *m, int *bonds);
/* Space-separated UFF type names. */
const char *smh_molecule_types(const smh_molecule *m);
void smh_molecule_free(smh_molecule *m);

/* ---- the reference's on-disk topology formats ----
 * Both readers take the file's bonds, angles, dihedrals and inversions as
 * given (Uff::from_topology, molff/src/uff.rs:238-240, keeps the topology's
 * lists and order), check that every term is made of bonds, then build the
 * neighbour tables, parameters, hneigh slots and REQH as smh_molecule_build
 * does.  charges (may be NULL) overrides REQH.Q per atom: neither format
 * carries charges.  Errors (no abort): unreadable file, bad magic/version,
 * counts that disagree with the payload, indices out of range, a term that
 * is not made of bonds, an unknown UFF type. */
/* UFFTOPO binary written by buildff --bin (crates/apps/buildff/src/main.rs:
 * 10-30 format, 125-182 writer): "UFFTOPO", u8 version 1, i32 natoms nbonds
 * nangles ndihedrals ninversions ntypes flags, then f64 apos[3N], i32 type
 * index[N], i32 bonds[2nb], angles[3na], dihedrals[4nd], inversions[4ni] and
 * the type table (8-byte names padded with spaces); little endian.  UFF types
 * come from the file's table (elements from AtomTypes.dat). */
int smh_molecule_read_ufftopo(const char *param_dir, const char *path, const double *charges, smh_molecule **out);
/* TopologyData JSON (moltopo/src/export.rs:12-26; buildff --json, read by
 * import_json :57-76 / surfmol::import::load_topology_from_json): natoms,
 * elements, positions, bonds, angles, dihedrals, inversions; the *_params
 * arrays are ignored (the reference writes them empty).  UFF types are
 * assigned from the bonds (assign_uff.rs:8-118). */
int smh_molecule_read_json(const char *param_dir, const char *path, const double *charges, smh_molecule **out);
/* Writers of the same two formats (the reference's export_json and
 * buildff --bin layouts), e.g. to hand a topology built here to the
 * reference's tools. */
int smh_molecule_write_ufftopo(const smh_molecule *m, const char *path);
int smh_molecule_write_json(const smh_molecule *m, const char *path);
/* Positions the molecule was built from or read with, [3*natoms]. */
int smh_molecule_positions(const smh_molecule *m, double *apos);

/* ---- one replica of the multi-replica rigid-atom port relaxation ----
 * (surfmol_raff.h).  smh_raff_build restates, for one molecule, the set-up
 * of the reference's RAFF benchmark (crates/libs/oclff/src/bin/
 * raff_ocl_benchmark.rs:78-97) with oclff::pack (crates/libs/oclff/src/
 * pack.rs): pack_molecules (nodes = atoms of bond degree > 1 first, then
 * caps, then padding atoms "X" up to group_size; nnode < 0 infers the node
 * count, otherwise it must equal it), masses_from_elems (inverse masses,
 * padding 0), build_neighs_from_bonds, make_exclusions_1st_2nd,
 * make_ports_from_neighs (port = initial neighbour displacement, stiffness
 * k_stiff) and make_bk_slots_clustered.  group_size 0 means no padding.
 * pos3 [3*natoms], bonds [2*nbonds] in the caller's atom order; the tables
 * are in packed order (perm[k] = caller index of packed atom k).  Errors
 * (no abort): bad indices, more than 4 bonds or back slots per atom, a node
 * count that disagrees with the bonds. */
typedef struct smh_raff_mol smh_raff_mol;
typedef struct {
    int natoms, nreal, nnode;
    const float *pos3, *inv_mass;          /* [3*natoms], [natoms]          */
    const int *neighs, *excl1, *excl2;     /* [4*natoms]                    */
    const int *bk_slots;                   /* [4*natoms]                    */
    const float *port_local, *kflat;       /* [16*natoms], [4*natoms]       */
    const int *is_padding, *perm;          /* [natoms], [nreal]             */
    const char *elems;                     /* space-separated, packed order */
} smh_raff_tables;
const char *smh_raff_last_error(void);
int smh_raff_build(int natoms, const char *const *elems, const float *pos3, int nbonds, const int *bonds, int nnode,
                   int group_size, float k_stiff, smh_raff_mol **out);
int smh_raff_tables_of(const smh_raff_mol *m, smh_raff_tables *t);
void smh_raff_free(smh_raff_mol *m);

#ifdef __cplusplus
}
#endif
#endif /* SURFMOL_HOST_H */
