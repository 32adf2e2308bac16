// host_raff.cpp -- host set-up of one replica for the multi-replica rigid-
// atom port relaxation (include/surfmol_host.h, smh_raff_*): the packing and
// table builders of the reference's oclff::pack (crates/libs/oclff/src/
// pack.rs:54-316) for one molecule, restated in C++ for callers without the
// Rust crates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/surfmol_host.h"

struct smh_raff_mol {
    int natoms = 0, nreal = 0, nnode = 0;
    std::vector<std::string> elems;       // packed order; "X" for padding
    std::vector<float> pos3, inv_mass, port_local, kflat;
    std::vector<int> neighs, excl1, excl2, bk_slots, is_padding, perm, bonds;
    std::string types_joined;
};

namespace {
thread_local std::string g_raff_err;

int raff_fail(const std::string &m) {
    g_raff_err = m;
    return SMH_ERR;
}

// element_mass (pack.rs:132-143)
float element_mass(const std::string &e) {
    static const struct { const char *el; float m; } tab[] = {
        {"H", 1.008f}, {"C", 12.011f}, {"N", 14.007f}, {"O", 15.999f}, {"F", 18.998f}, {"P", 30.974f},
        {"S", 32.06f}, {"Cl", 35.45f}, {"Br", 79.904f}, {"I", 126.904f}, {"Na", 22.990f}, {"K", 39.098f},
        {"Mg", 24.305f}, {"Ca", 40.078f}, {"Fe", 55.845f}, {"Cu", 63.546f}, {"Zn", 65.38f}, {"Al", 26.982f},
        {"Si", 28.085f}, {"B", 10.81f}, {"Li", 6.94f}, {"Be", 9.012f}, {"Ne", 20.180f}, {"Ar", 39.948f}};
    for (const auto &t : tab)
        if (e == t.el) return t.m;
    return 12.0f;
}
}  // namespace

extern "C" const char *smh_raff_last_error(void) { return g_raff_err.c_str(); }

extern "C" int smh_raff_build(int natoms, const char *const *elems, const float *pos3, int nbonds, const int *bonds,
                              int nnode, int group_size, float k_stiff, smh_raff_mol **out) {
    if (!out || !elems || !pos3 || (nbonds > 0 && !bonds)) return raff_fail("smh_raff_build: NULL argument");
    *out = nullptr;
    if (natoms <= 0) return raff_fail("smh_raff_build: natoms must be > 0");
    if (group_size == 0) group_size = natoms;
    if (natoms > group_size)
        return raff_fail("smh_raff_build: " + std::to_string(natoms) + " atoms > group_size " + std::to_string(group_size));
    if (!std::isfinite(k_stiff)) return raff_fail("smh_raff_build: k_stiff is not finite");
    // pack_molecules (pack.rs:54-116): nodes (bond degree > 1) first, then caps
    std::vector<int> deg(natoms, 0);
    for (int b = 0; b < nbonds; b++) {
        const int i = bonds[2 * b], j = bonds[2 * b + 1];
        if (i < 0 || i >= natoms || j < 0 || j >= natoms)
            return raff_fail("smh_raff_build: bond " + std::to_string(b) + " (" + std::to_string(i) + "," +
                             std::to_string(j) + ") out of range");
        if (i == j) continue;
        deg[i]++;
        deg[j]++;
    }
    int ninf = 0;
    for (int i = 0; i < natoms; i++) ninf += deg[i] > 1;
    if (nnode < 0) nnode = ninf;
    if (nnode != ninf)
        return raff_fail("smh_raff_build: nnode = " + std::to_string(nnode) + " but " + std::to_string(ninf) +
                         " atoms have bond degree > 1");
    auto *m = new smh_raff_mol();
    m->natoms = group_size;
    m->nreal = natoms;
    m->nnode = nnode;
    for (int i = 0; i < natoms; i++)
        if (deg[i] > 1) m->perm.push_back(i);
    for (int i = 0; i < natoms; i++)
        if (deg[i] <= 1) m->perm.push_back(i);
    std::vector<int> inv(natoms);
    for (int k = 0; k < natoms; k++) inv[m->perm[k]] = k;
    const int N = group_size;
    m->elems.assign(N, "X");
    m->pos3.assign(3 * (size_t)N, 0.f);
    m->inv_mass.assign(N, 0.f);
    m->is_padding.assign(N, 1);
    for (int k = 0; k < natoms; k++) {
        const int p = m->perm[k];
        m->elems[k] = elems[p];
        for (int d = 0; d < 3; d++) m->pos3[3 * k + d] = pos3[3 * p + d];
        // masses_from_elems (pack.rs:145-150)
        m->inv_mass[k] = 1.0f / element_mass(m->elems[k]);
        m->is_padding[k] = 0;
    }
    for (int b = 0; b < nbonds; b++) {
        m->bonds.push_back(inv[bonds[2 * b]]);
        m->bonds.push_back(inv[bonds[2 * b + 1]]);
    }
    // build_neighs_from_bonds (pack.rs:155-170)
    m->neighs.assign(4 * (size_t)N, -1);
    std::vector<int> d(N, 0);
    for (size_t b = 0; b < m->bonds.size() / 2; b++) {
        const int i = m->bonds[2 * b], j = m->bonds[2 * b + 1];
        if (d[i] >= 4 || d[j] >= 4) {
            delete m;
            return raff_fail("smh_raff_build: atom " + std::to_string(d[i] >= 4 ? i : j) + " has more than 4 bonds");
        }
        m->neighs[4 * i + d[i]++] = j;
        m->neighs[4 * j + d[j]++] = i;
    }
    // make_exclusions_1st_2nd (pack.rs:176-211)
    m->excl1.assign(4 * (size_t)N, -1);
    m->excl2.assign(4 * (size_t)N, -1);
    for (int i = 0; i < N; i++) {
        std::vector<int> n1;
        for (int k = 0; k < 4; k++)
            if (m->neighs[4 * i + k] >= 0) n1.push_back(m->neighs[4 * i + k]);
        for (size_t k = 0; k < n1.size(); k++) m->excl1[4 * i + k] = n1[k];
        std::vector<int> s2;
        for (int j : n1) {
            for (int k = 0; k < 4 && s2.size() < 4; k++) {
                const int t = m->neighs[4 * j + k];
                if (t < 0 || t == i) continue;
                if (std::find(n1.begin(), n1.end(), t) != n1.end()) continue;
                if (std::find(s2.begin(), s2.end(), t) != s2.end()) continue;
                s2.push_back(t);
            }
            if (s2.size() >= 4) break;
        }
        for (size_t k = 0; k < s2.size(); k++) m->excl2[4 * i + k] = s2[k];
    }
    // make_ports_from_neighs (pack.rs:293-316): port = pos[j] - pos[i], K
    m->port_local.assign(16 * (size_t)N, 0.f);
    m->kflat.assign(4 * (size_t)N, 0.f);
    for (int i = 0; i < N; i++)
        for (int k = 0; k < 4; k++) {
            const int j = m->neighs[4 * i + k];
            if (j < 0) continue;
            for (int c = 0; c < 3; c++) m->port_local[(4 * i + k) * 4 + c] = m->pos3[3 * j + c] - m->pos3[3 * i + c];
            m->kflat[4 * i + k] = k_stiff;
        }
    // make_bk_slots_clustered (pack.rs:218-244), one group
    m->bk_slots.assign(4 * (size_t)N, -1);
    std::vector<int> cnt(N, 0);
    for (int il = 0; il < nnode; il++)
        for (int k = 0; k < 4; k++) {
            const int ja = m->neighs[4 * il + k];
            if (ja < 0) continue;
            if (cnt[ja] >= 4) {
                delete m;
                return raff_fail("smh_raff_build: atom " + std::to_string(ja) + " has more than 4 back slots");
            }
            m->bk_slots[4 * ja + cnt[ja]++] = 4 * il + k;
        }
    for (int k = 0; k < N; k++) m->types_joined += (k ? " " : "") + m->elems[k];
    *out = m;
    return SMH_OK;
}

extern "C" int smh_raff_tables_of(const smh_raff_mol *m, smh_raff_tables *t) {
    if (!m || !t) return raff_fail("smh_raff_tables_of: NULL argument");
    t->natoms = m->natoms;
    t->nreal = m->nreal;
    t->nnode = m->nnode;
    t->pos3 = m->pos3.data();
    t->inv_mass = m->inv_mass.data();
    t->neighs = m->neighs.data();
    t->excl1 = m->excl1.data();
    t->excl2 = m->excl2.data();
    t->bk_slots = m->bk_slots.data();
    t->port_local = m->port_local.data();
    t->kflat = m->kflat.data();
    t->is_padding = m->is_padding.data();
    t->perm = m->perm.data();
    t->elems = m->types_joined.c_str();
    return SMH_OK;
}

extern "C" void smh_raff_free(smh_raff_mol *m) { delete m; }
