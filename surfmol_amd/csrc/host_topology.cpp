// host_topology.cpp -- host-side set-up of the relaxation path (C++ mirror of
// the reference's moltopo + molff set-up API).
//
// Builds, for one molecule, exactly the tables the reference's Uff and
// NonBondedFF structs hold, from element symbols + positions (+ optional
// explicit bond list):
//   topology   moltopo/src/topology.rs:101-182  (dihedrals sorted: the
//              reference's HashSet order carries no meaning)
//   neighbours moltopo/src/molecular.rs:34-78
//   UFF types  moltopo/src/assign_uff.rs:8-123
//   params     moltopo/src/params.rs:120-291,595-657; molff/src/uff.rs:699-887
//   hneigh     molff/src/uff.rs:250-315
//   REQH       moltopo/src/params.rs:595-604
// The tables are returned through the C ABI in surfmol_host.h and feed
// smr_batch_create() (surfmol_relax.h) unchanged.  Topologies can also be
// read from (and written to) the reference's two on-disk formats: buildff's
// UFFTOPO binary (crates/apps/buildff/src/main.rs:10-30,125-182) and the
// TopologyData JSON (moltopo/src/export.rs:12-76); their angle, dihedral and
// inversion lists are taken as given, as Uff::from_topology does.
#include "surfmol_host.h"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <fstream>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

namespace {

thread_local std::string g_host_err;

int host_fail(const std::string &msg) {
    g_host_err = msg;
    return SMH_ERR;
}

constexpr double kKcalToEv = 4.1840 / 60.2214076 / 1.602176634;  // params.rs:607

struct ElementRow {
    double r_vdw = 0, e_vdw = 0, q_uff = 0, u_uff = 0, v_uff = 0, e_aff = 0;
};
struct AtomTypeRow {
    std::string element;
    double r_uff = 0, r_vdw = 0, e_vdw = 0, q_base = 0, h_b = 0, a_ss = 0;
};

struct ParamTables {
    std::map<std::string, ElementRow> elements;
    std::map<std::string, AtomTypeRow> atom_types;
    const ElementRow *element_of(const std::string &t) const {
        auto it = atom_types.find(t);
        if (it == atom_types.end()) return nullptr;
        auto e = elements.find(it->second.element);
        return e == elements.end() ? nullptr : &e->second;
    }
    const AtomTypeRow *atype(const std::string &t) const {
        auto it = atom_types.find(t);
        return it == atom_types.end() ? nullptr : &it->second;
    }
};

std::vector<std::string> split_ws(const std::string &line) {
    std::istringstream is(line);
    std::vector<std::string> out;
    std::string w;
    while (is >> w) out.push_back(w);
    return out;
}

std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

// Params::load_element_types / load_atom_types (params.rs:130-258)
bool load_params(const std::string &dir, ParamTables &P, std::string &err) {
    std::ifstream fe(dir + "/ElementTypes.dat");
    if (!fe) { err = "cannot read " + dir + "/ElementTypes.dat"; return false; }
    std::string line;
    while (std::getline(fe, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        auto p = split_ws(line);
        if (p.size() < 10) continue;
        ElementRow r;
        r.r_vdw = std::stod(p[7]);
        r.e_vdw = std::stod(p[8]);
        r.q_uff = std::stod(p[9]);
        if (p.size() >= 11) r.u_uff = std::stod(p[10]);
        if (p.size() >= 12) r.v_uff = std::stod(p[11]);
        if (p.size() >= 16) r.e_aff = std::stod(p[12]);
        P.elements[p[0]] = r;
    }
    std::ifstream fa(dir + "/AtomTypes.dat");
    if (!fa) { err = "cannot read " + dir + "/AtomTypes.dat"; return false; }
    while (std::getline(fa, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        auto p = split_ws(line);
        if (p.size() < 13) continue;
        AtomTypeRow r;
        r.element = p[2];
        r.r_uff = std::stod(p[8]);
        r.r_vdw = std::stod(p[9]);
        r.e_vdw = std::stod(p[10]);
        r.q_base = std::stod(p[11]);
        r.h_b = std::stod(p[12]);
        if (p.size() >= 19) r.a_ss = std::stod(p[13]);
        P.atom_types[p[0]] = r;
    }
    return true;
}

std::mutex g_params_mu;
std::map<std::string, ParamTables> g_params_cache;

const ParamTables *get_params(const char *dir, std::string &err) {
    std::lock_guard<std::mutex> lk(g_params_mu);
    auto it = g_params_cache.find(dir);
    if (it != g_params_cache.end()) return &it->second;
    ParamTables P;
    if (!load_params(dir, P, err)) return nullptr;
    return &(g_params_cache[dir] = std::move(P));
}

// ---- UFF formulas (params.rs:617-647) ----
double bond_order_from_types(const std::string &a, const std::string &b) {
    auto ends = [](const std::string &t, const char *suf) {
        size_t n = std::strlen(suf);
        return t.size() >= n && t.compare(t.size() - n, n, suf) == 0;
    };
    if (ends(a, "_1") || ends(b, "_1")) return 3.0;
    if (ends(a, "_2") && ends(b, "_2")) return 2.0;
    return 1.0;
}

double calc_rij(const ParamTables &P, const std::string &ta, const std::string &tb, double bo) {
    const AtomTypeRow *ti = P.atype(ta), *tj = P.atype(tb);
    const ElementRow *ei = P.element_of(ta), *ej = P.element_of(tb);
    if (!ti || !tj || !ei || !ej) return 1.5;
    double r_bo = -0.1332 * (ti->r_uff + tj->r_uff) * std::log(bo);
    double r_en = 0.0;
    if (ei->e_aff < 0.0 && ej->e_aff < 0.0) {
        double s = std::sqrt(-ei->e_aff) - std::sqrt(-ej->e_aff);
        r_en = ti->r_uff * tj->r_uff * s * s / (-ei->e_aff * ti->r_uff - ej->e_aff * tj->r_uff);
    }
    return ti->r_uff + tj->r_uff + r_bo - r_en;
}

std::string elem_sym(const std::string &t) {
    size_t i = t.find('_');
    return i == std::string::npos ? t : t.substr(0, i);
}
char hyb_suffix(const std::string &t) {
    size_t i = t.rfind('_');
    if (i != std::string::npos && i + 1 < t.size()) return t[i + 1];
    return '3';
}
bool is_sp2(char h) { return h == '2' || h == 'R'; }

int ne_pairs(const std::string &el) {
    static const std::map<std::string, int> m = {{"H", 0}, {"C", 0}, {"N", 1}, {"O", 2}, {"F", 3},
                                                 {"Si", 0}, {"P", 1}, {"S", 2}, {"Cl", 3}};
    auto it = m.find(el);
    return it == m.end() ? 0 : it->second;
}
// topology.rs:28-34
int hybridization(const std::string &el, int nn) {
    if (el == "H") return 1;
    int npi = 4 - ne_pairs(el) - nn;
    if (npi < 0) return 1;
    return npi == 0 ? 3 : (npi == 1 ? 2 : 1);
}
bool type_exists(const std::string &t) {
    static const std::set<std::string> k = {"H_", "C_3", "C_R", "C_2", "C_1", "N_3", "N_R", "N_2", "N_1",
                                            "O_3", "O_R", "O_2", "O_1", "F", "Cl", "Si3", "P", "S"};
    return k.count(t) > 0;
}

struct Built {
    std::vector<std::string> types;
    std::vector<std::array<int, 2>> bonds;
    std::vector<std::array<int, 3>> angles;
    std::vector<std::array<int, 4>> dih, inv;
    std::vector<std::array<int, 4>> neighs, neigh_bs;
};

}  // namespace

struct smh_molecule {
    int natoms = 0;
    std::vector<int> neighs, neigh_bs;         // [4N]
    std::vector<int> bonds;                    // [2nb]
    std::vector<double> bon_params;            // [2nb]
    std::vector<int> ang_atoms, ang_ngs;       // [3na], [2na]
    std::vector<double> ang_params;            // [5na]
    std::vector<int> dih_atoms, dih_ngs;       // [4nd], [3nd]
    std::vector<double> dih_params;            // [3nd]
    std::vector<int> inv_atoms, inv_ngs;       // [4ni], [3ni]
    std::vector<double> inv_params;            // [4ni]
    std::vector<double> reqs;                  // [4N]
    std::vector<std::string> types;
    std::string types_joined;
    std::vector<std::string> elems;            // element symbols
    std::vector<double> apos;                  // [3N] positions as built or read
};

extern "C" const char *smh_last_error(void) { return g_host_err.c_str(); }

namespace {

// Term lists and UFF types a caller supplies instead of the enumerated /
// assigned ones (a topology read from a file: Uff::from_topology,
// molff/src/uff.rs:238-240, takes the topology's lists as given).
struct GivenTerms {
    const std::vector<std::array<int, 3>> *angles = nullptr;
    const std::vector<std::array<int, 4>> *dihedrals = nullptr;
    const std::vector<std::array<int, 4>> *inversions = nullptr;
    const std::vector<std::string> *types = nullptr;
};

int build_core(const ParamTables &P, const std::vector<std::string> &el, const double *apos, const double *charges,
               const std::vector<std::array<int, 2>> &bonds, const GivenTerms &given, const std::string &who,
               smh_molecule **out);

}  // namespace

extern "C" int smh_molecule_build(const char *param_dir, int natoms, const char *const *elems, const double *apos,
                                  const double *charges, int nbonds_in, const int *bonds_in, double rcut_bond,
                                  smh_molecule **out) {
    if (!out) return host_fail("smh_molecule_build: out == NULL");
    *out = nullptr;
    if (natoms <= 0 || !elems || !apos) return host_fail("smh_molecule_build: natoms must be > 0 with elems/apos");
    std::string err;
    const ParamTables *Pp = get_params(param_dir ? param_dir : ".", err);
    if (!Pp) return host_fail("smh_molecule_build: " + err);
    const ParamTables &P = *Pp;
    const int n = natoms;
    std::vector<std::string> el(n);
    for (int i = 0; i < n; i++) el[i] = elems[i] ? elems[i] : "";

    // ---- bonds: explicit, or build_bonds_by_cutoff (topology.rs:101-111) ----
    std::vector<std::array<int, 2>> bonds;
    if (nbonds_in > 0) {
        for (int b = 0; b < nbonds_in; b++) {
            int a = bonds_in[2 * b], c = bonds_in[2 * b + 1];
            if (a < 0 || c < 0 || a >= n || c >= n || a == c)
                return host_fail("smh_molecule_build: bond " + std::to_string(b) + " has invalid atoms (" +
                                 std::to_string(a) + "," + std::to_string(c) + ")");
            bonds.push_back({a, c});
        }
    } else {
        double r2cut = rcut_bond * rcut_bond;
        for (int i = 0; i < n; i++)
            for (int j = i + 1; j < n; j++) {
                double dx = apos[3 * j] - apos[3 * i], dy = apos[3 * j + 1] - apos[3 * i + 1],
                       dz = apos[3 * j + 2] - apos[3 * i + 2];
                if (dx * dx + dy * dy + dz * dz < r2cut) bonds.push_back({i, j});
            }
    }
    return build_core(P, el, apos, charges, bonds, GivenTerms(), "smh_molecule_build", out);
}

namespace {

int build_core(const ParamTables &P, const std::vector<std::string> &el, const double *apos, const double *charges,
               const std::vector<std::array<int, 2>> &bonds, const GivenTerms &given, const std::string &who,
               smh_molecule **out) {
    const int n = (int)el.size();
    std::vector<std::vector<int>> adj(n);
    for (auto &b : bonds) {
        adj[b[0]].push_back(b[1]);
        adj[b[1]].push_back(b[0]);
    }
    for (int i = 0; i < n; i++)
        if (adj[i].size() > 4)
            return host_fail(who + ": atom " + std::to_string(i) + " (" + el[i] + ") has " +
                             std::to_string(adj[i].size()) + " bonds; neighbour tables hold at most 4");
    auto bonded = [&](int a, int b) { return std::find(adj[a].begin(), adj[a].end(), b) != adj[a].end(); };

    // ---- angles (topology.rs:113-129), or as given: i-j-k with j bonded to both ----
    std::vector<std::array<int, 3>> angles;
    if (given.angles) {
        for (size_t q = 0; q < given.angles->size(); q++) {
            const auto &a = (*given.angles)[q];
            for (int v : a)
                if (v < 0 || v >= n) return host_fail(who + ": angle " + std::to_string(q) + " has an atom out of range");
            if (!bonded(a[0], a[1]) || !bonded(a[2], a[1]))
                return host_fail(who + ": angle " + std::to_string(q) + " is not two bonds at atom " + std::to_string(a[1]));
        }
        angles = *given.angles;
    } else {
        for (int j = 0; j < n; j++)
            for (size_t a = 0; a < adj[j].size(); a++)
                for (size_t b = a + 1; b < adj[j].size(); b++) angles.push_back({adj[j][a], j, adj[j][b]});
    }
    // ---- dihedrals (topology.rs:131-159), as a sorted set, or as given: a bonded chain i-j-k-l ----
    std::vector<std::array<int, 4>> dih;
    if (given.dihedrals) {
        for (size_t q = 0; q < given.dihedrals->size(); q++) {
            const auto &d = (*given.dihedrals)[q];
            for (int v : d)
                if (v < 0 || v >= n) return host_fail(who + ": dihedral " + std::to_string(q) + " has an atom out of range");
            if (!bonded(d[0], d[1]) || !bonded(d[1], d[2]) || !bonded(d[2], d[3]))
                return host_fail(who + ": dihedral " + std::to_string(q) + " is not a chain of three bonds");
        }
        dih = *given.dihedrals;
    } else {
        std::set<std::array<int, 4>> dset;
        for (int j = 0; j < n; j++)
            for (int k : adj[j])
                for (int i : adj[j]) {
                    if (i == k) continue;
                    for (int l : adj[k]) {
                        if (l == j) continue;
                        if (l > i) dset.insert({i, j, k, l});
                    }
                }
        dih.assign(dset.begin(), dset.end());
    }
    // ---- inversions (topology.rs:161-182), or as given: centre i bonded to j, k, l ----
    std::vector<std::array<int, 4>> inv;
    if (given.inversions) {
        for (size_t q = 0; q < given.inversions->size(); q++) {
            const auto &v4 = (*given.inversions)[q];
            for (int v : v4)
                if (v < 0 || v >= n) return host_fail(who + ": inversion " + std::to_string(q) + " has an atom out of range");
            if (!bonded(v4[0], v4[1]) || !bonded(v4[0], v4[2]) || !bonded(v4[0], v4[3]))
                return host_fail(who + ": inversion " + std::to_string(q) + " is not a centre with three bonds");
        }
        inv = *given.inversions;
    } else {
        for (int i = 0; i < n; i++)
            if (adj[i].size() == 3) {
                const auto &s = adj[i];
                inv.push_back({i, s[0], s[1], s[2]});
                inv.push_back({i, s[2], s[0], s[1]});
                inv.push_back({i, s[1], s[2], s[0]});
            }
    }
    // ---- neighbour tables (molecular.rs:34-78) ----
    std::vector<std::array<int, 4>> ng(n, {-1, -1, -1, -1}), ngb(n, {-1, -1, -1, -1});
    for (size_t ib = 0; ib < bonds.size(); ib++) {
        int ia = bonds[ib][0], ja = bonds[ib][1];
        for (int s = 0; s < 4; s++) if (ng[ia][s] < 0) { ng[ia][s] = ja; ngb[ia][s] = (int)ib; break; }
        for (int s = 0; s < 4; s++) if (ng[ja][s] < 0) { ng[ja][s] = ia; ngb[ja][s] = (int)ib; break; }
    }
    auto nneigh = [&](int i) { int c = 0; for (int s = 0; s < 4; s++) { if (ng[i][s] < 0) break; c++; } return c; };

    // ---- UFF typing (assign_uff.rs:8-118), or as given ----
    std::vector<std::string> types(el);
    std::vector<char> set_atom(n, 0);
    if (given.types) {
        types = *given.types;
        for (int ia = 0; ia < n; ia++) {
            if (!P.atype(types[ia]))
                return host_fail(who + ": atom " + std::to_string(ia) + " has UFF type '" + types[ia] +
                                 "', not in AtomTypes.dat");
            set_atom[ia] = 1;
        }
    }
    for (int ia = 0; ia < n; ia++) {
        if (set_atom[ia]) continue;
        const std::string &name = el[ia];
        int nb = nneigh(ia);
        int h = hybridization(name, nb);
        if (name == "H") { types[ia] = "H_"; set_atom[ia] = 1; continue; }
        if (name == "N" && nb == 3) {
            int n_o = 0;
            for (int k = 0; k < 3; k++) if (ng[ia][k] >= 0 && el[ng[ia][k]] == "O") n_o++;
            if (n_o == 2) {
                types[ia] = "N_R"; set_atom[ia] = 1;
                for (int k = 0; k < 3; k++) {
                    int ja = ng[ia][k];
                    if (ja >= 0 && el[ja] == "O") { types[ja] = "O_R"; set_atom[ja] = 1; }
                }
                continue;
            }
        }
        if (name == "O" && nb == 1) { types[ia] = "O_2"; set_atom[ia] = 1; continue; }
        if (name == "C" && nb == 2) { types[ia] = "C_1"; set_atom[ia] = 1; continue; }
        const char *suf = "3";
        if (h == 2) suf = (name == "C" || name == "N" || name == "O") ? "R" : "2";
        else if (h == 1) suf = "1";
        std::string tname = name + "_" + suf;
        if (type_exists(tname)) { types[ia] = tname; set_atom[ia] = 1; }
        else if (type_exists(name)) { types[ia] = name; set_atom[ia] = 1; }
    }

    auto *m = new smh_molecule();
    m->natoms = n;
    m->types = types;
    m->elems = el;
    m->apos.assign(apos, apos + 3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        m->types_joined += types[i];
        if (i + 1 < n) m->types_joined += " ";
    }
    for (int i = 0; i < n; i++)
        for (int s = 0; s < 4; s++) { m->neighs.push_back(ng[i][s]); m->neigh_bs.push_back(ngb[i][s]); }

    // ---- bond params (uff.rs:704-717) ----
    for (auto &b : bonds) {
        const std::string &ta = types[b[0]], &tb = types[b[1]];
        double bo = bond_order_from_types(ta, tb);
        double l0 = calc_rij(P, ta, tb, bo);
        const ElementRow *ei = P.element_of(ta), *ej = P.element_of(tb);
        double k = (ei && ej) ? 0.5 * 28.7989689090648 * ei->q_uff * ej->q_uff / (l0 * l0 * l0) : 100.0;
        m->bonds.push_back(b[0]); m->bonds.push_back(b[1]);
        m->bon_params.push_back(k); m->bon_params.push_back(l0);
    }
    // ---- angle params (uff.rs:719-763) ----
    const double deg2rad = M_PI / 180.0;
    for (auto &a : angles) {
        int i = a[0], j = a[1], k = a[2];
        const std::string &tj = types[j], &ti = types[i], &tk = types[k];
        std::array<double, 5> row{0, 0, 0, 0, 0};
        const AtomTypeRow *atj = P.atype(tj);
        const ElementRow *ei = P.element_of(ti), *ek = P.element_of(tk);
        if (elem_sym(tj) != "H" && atj && ei && ek) {
            double ct = std::cos(atj->a_ss * deg2rad);
            double st2 = std::pow(std::sin(atj->a_ss * deg2rad), 2);
            double rij = calc_rij(P, ti, tj, bond_order_from_types(ti, tj));
            double rjk = calc_rij(P, tj, tk, bond_order_from_types(tj, tk));
            double rik = std::sqrt(rij * rij + rjk * rjk - 2.0 * rij * rjk * ct);
            double kappa = 28.7989689090648 * ei->q_uff * ek->q_uff / std::pow(rik, 5) *
                           (3.0 * rij * rjk * st2 - rik * rik * ct);
            char hy = hyb_suffix(tj);
            if (hy == '1') row = {kappa, 1.0, 1.0, 0.0, 0.0};
            else if (is_sp2(hy)) row = {kappa / 9.0, 1.0, 0.0, 0.0, -1.0};
            else if (hy == '3') {
                double c2 = 1.0 / (4.0 * st2), c1 = -4.0 * c2 * ct, c0 = c2 * (2.0 * ct * ct + 1.0);
                row = {kappa, c0, c1, c2, 0.0};
            } else row = {kappa, 1.0, 0.0, 0.0, -1.0};
        }
        for (int q = 0; q < 3; q++) m->ang_atoms.push_back(a[q]);
        for (double v : row) m->ang_params.push_back(v);
        // bake_angle_neighs (uff.rs:250-267)
        int ji = -1, jk = -1;
        for (int s = 0; s < 4; s++) {
            int ing = ng[j][s];
            if (ing < 0) break;
            if (ing == i) ji = j * 4 + s;
            else if (ing == k) jk = j * 4 + s;
        }
        m->ang_ngs.push_back(ji); m->ang_ngs.push_back(jk);
    }
    // ---- dihedral params (uff.rs:765-844) ----
    for (auto &d : dih) {
        int i2 = d[1], i3 = d[2];
        const std::string &t2 = types[i2], &t3 = types[i3];
        double v = 0.0, dsg = 1.0, nn = 3.0;
        char h2 = hyb_suffix(t2), h3 = hyb_suffix(t3);
        if (elem_sym(t2) == "H" || elem_sym(t3) == "H" || h2 == '1' || h3 == '1') {
            v = 0.0; dsg = 1.0; nn = 3.0;
        } else {
            bool s2 = h2 == '3', s3 = h3 == '3';
            const ElementRow *e2 = P.element_of(t2), *e3 = P.element_of(t3);
            double bo23 = bond_order_from_types(t2, t3);
            double u2 = e2 ? e2->u_uff : 0.0, u3 = e3 ? e3->u_uff : 0.0;
            if (s2 && s3) {
                std::string el2 = elem_sym(t2), el3 = elem_sym(t3);
                if ((el2 == "O" || el2 == "S") && (el3 == "O" || el3 == "S")) {
                    double kk = kKcalToEv;
                    kk *= (el2 == "O") ? 2.0 : 6.8;
                    kk *= (el3 == "O") ? 2.0 : 6.8;
                    v = std::sqrt(kk); dsg = 1.0; nn = 2.0;
                } else {
                    v = std::sqrt((e2 ? e2->v_uff : 0.0) * (e3 ? e3->v_uff : 0.0)); dsg = 1.0; nn = 3.0;
                }
            } else if ((s2 && is_sp2(h3)) || (is_sp2(h2) && s3)) {
                v = kKcalToEv; dsg = -1.0; nn = 6.0;
                std::string el_sp3 = s2 ? elem_sym(t2) : elem_sym(t3);
                if (el_sp3 == "O" || el_sp3 == "S") {
                    v = 5.0 * std::sqrt(u2 * u3) * (1.0 + 4.18 * std::log(bo23)); dsg = 1.0; nn = 2.0;
                }
                int sp3_idx = s2 ? i2 : i3, sp2_idx = s2 ? i3 : i2;
                bool has = false;
                for (int s = 0; s < 4; s++) {
                    int q = ng[sp2_idx][s];
                    if (q < 0) break;
                    if (q == sp3_idx) continue;
                    if (is_sp2(hyb_suffix(types[q]))) { has = true; break; }
                }
                if (has) { v = 2.0 * kKcalToEv; dsg = 1.0; nn = 3.0; }
            } else if (is_sp2(h2) && is_sp2(h3)) {
                v = 5.0 * std::sqrt(u2 * u3) * (1.0 + 4.18 * std::log(bo23)); dsg = -1.0; nn = 2.0;
            } else {
                v = 0.0; dsg = 1.0; nn = 3.0;
            }
            double nb2 = nneigh(i2), nb3 = nneigh(i3);
            double den = 0.5 * (nb2 - 1.0) * (nb3 - 1.0);
            if (den > 0.0) v /= den;
        }
        for (int q = 0; q < 4; q++) m->dih_atoms.push_back(d[q]);
        m->dih_params.push_back(v); m->dih_params.push_back(dsg); m->dih_params.push_back(nn);
        // bake_dihedral_neighs (uff.rs:269-294)
        int i = d[0], j = d[1], k = d[2], l = d[3];
        int ji = -1, jk = -1, kl = -1;
        for (int s = 0; s < 4; s++) {
            int ing = ng[j][s];
            if (ing < 0) break;
            if (ing == i) ji = j * 4 + s;
            else if (ing == k) jk = j * 4 + s;
        }
        for (int s = 0; s < 4; s++) {
            int ing = ng[k][s];
            if (ing < 0) break;
            if (ing == l) kl = k * 4 + s;
        }
        m->dih_ngs.push_back(ji); m->dih_ngs.push_back(jk); m->dih_ngs.push_back(kl);
    }
    // ---- inversion params (uff.rs:846-886) ----
    for (auto &v4 : inv) {
        int i1 = v4[0];
        const std::string &t1 = types[i1];
        std::string el1 = elem_sym(t1);
        char h1 = hyb_suffix(t1);
        std::array<double, 4> row{0, 0, 0, 0};
        if (el1 == "C" && is_sp2(h1)) {
            bool carbonyl = false;
            for (int s = 0; s < 4; s++) {
                int q = ng[i1][s];
                if (q < 0) break;
                if (elem_sym(types[q]) == "O" && hyb_suffix(types[q]) == '2') { carbonyl = true; break; }
            }
            row = {(carbonyl ? 50.0 : 6.0) * kKcalToEv, 1.0, -1.0, 0.0};
        } else if (el1 == "N" && is_sp2(h1)) {
            row = {6.0 * kKcalToEv, 1.0, -1.0, 0.0};
        } else if (el1 == "P") {
            double om = 84.4339 * deg2rad;
            double c0 = 4.0 * std::pow(std::cos(om), 2) - std::cos(2.0 * om), c1 = -4.0 * std::cos(om), c2 = 1.0;
            row = {22.0 * kKcalToEv / (c0 + c1 + c2), c0, c1, c2};
        }
        for (int q = 0; q < 4; q++) m->inv_atoms.push_back(v4[q]);
        m->inv_params.push_back(row[0] / 3.0);
        m->inv_params.push_back(row[1]); m->inv_params.push_back(row[2]); m->inv_params.push_back(row[3]);
        // bake_inversion_neighs (uff.rs:296-315): slots 0..2 only
        int i = v4[0], j = v4[1], k = v4[2], l = v4[3];
        int ij = -1, ik = -1, il = -1;
        for (int s = 0; s < 3; s++) {
            int ing = ng[i][s];
            if (ing == j) ij = i * 4 + s;
            else if (ing == k) ik = i * 4 + s;
            else if (ing == l) il = i * 4 + s;
        }
        m->inv_ngs.push_back(ij); m->inv_ngs.push_back(ik); m->inv_ngs.push_back(il);
    }
    // ---- REQH (params.rs:595-604), charges overridden when given ----
    for (int i = 0; i < n; i++) {
        std::array<double, 4> r{1.5, 0.0, 0.0, 0.0};
        if (const AtomTypeRow *at = P.atype(types[i])) r = {at->r_vdw, std::sqrt(at->e_vdw), at->q_base, at->h_b};
        else {
            auto e = P.elements.find(types[i]);
            if (e != P.elements.end()) r = {e->second.r_vdw, std::sqrt(e->second.e_vdw), e->second.q_uff, 0.0};
        }
        if (charges) r[2] = charges[i];
        for (double v : r) m->reqs.push_back(v);
    }
    *out = m;
    return SMH_OK;
}

}  // namespace

extern "C" int smh_molecule_counts(const smh_molecule *m, int *counts5) {
    if (!m || !counts5) return host_fail("smh_molecule_counts: NULL argument");
    counts5[0] = m->natoms;
    counts5[1] = (int)m->bonds.size() / 2;
    counts5[2] = (int)m->ang_atoms.size() / 3;
    counts5[3] = (int)m->dih_atoms.size() / 4;
    counts5[4] = (int)m->inv_atoms.size() / 4;
    return SMH_OK;
}

extern "C" int smh_molecule_topology(const smh_molecule *m, smr_topology *t) {
    if (!m || !t) return host_fail("smh_molecule_topology: NULL argument");
    t->natoms = m->natoms;
    t->nbonds = (int)m->bonds.size() / 2;
    t->nangles = (int)m->ang_atoms.size() / 3;
    t->ndih = (int)m->dih_atoms.size() / 4;
    t->ninv = (int)m->inv_atoms.size() / 4;
    t->neighs = m->neighs.data();
    t->neigh_bs = m->neigh_bs.data();
    t->bon_params = m->bon_params.data();
    t->ang_atoms = m->ang_atoms.data();
    t->ang_ngs = m->ang_ngs.data();
    t->ang_params = m->ang_params.data();
    t->dih_atoms = m->dih_atoms.data();
    t->dih_ngs = m->dih_ngs.data();
    t->dih_params = m->dih_params.data();
    t->inv_atoms = m->inv_atoms.data();
    t->inv_ngs = m->inv_ngs.data();
    t->inv_params = m->inv_params.data();
    t->reqs = m->reqs.data();
    return SMH_OK;
}

extern "C" int smh_molecule_bonds(const smh_molecule *m, int *bonds) {
    if (!m || !bonds) return host_fail("smh_molecule_bonds: NULL argument");
    std::copy(m->bonds.begin(), m->bonds.end(), bonds);
    return SMH_OK;
}

extern "C" const char *smh_molecule_types(const smh_molecule *m) { return m ? m->types_joined.c_str() : ""; }

extern "C" void smh_molecule_free(smh_molecule *m) { delete m; }

// ===========================================================================
// On-disk topology formats of the reference
// ===========================================================================
namespace {

// ---- UFFTOPO binary (crates/apps/buildff/src/main.rs:10-30, writer :125-182):
//      "UFFTOPO" u8 version=1, i32 natoms nbonds nangles ndihedrals
//      ninversions ntypes flags (little endian), then apos f64[3N], type index
//      i32[N], bonds i32[2nb], angles i32[3na], dihedrals i32[4nd], inversions
//      i32[4ni], type table ntypes x 8 bytes (names padded with spaces) ----
constexpr char kUffTopoMagic[7] = {'U', 'F', 'F', 'T', 'O', 'P', 'O'};

struct ByteReader {
    const std::vector<unsigned char> &b;
    size_t p = 0;
    bool ok = true;
    explicit ByteReader(const std::vector<unsigned char> &buf) : b(buf) {}
    bool take(void *dst, size_t n) {
        if (!ok || p + n > b.size()) { ok = false; return false; }
        std::memcpy(dst, b.data() + p, n);
        p += n;
        return true;
    }
    int32_t i32() {
        unsigned char x[4] = {0, 0, 0, 0};
        take(x, 4);
        return (int32_t)((uint32_t)x[0] | ((uint32_t)x[1] << 8) | ((uint32_t)x[2] << 16) | ((uint32_t)x[3] << 24));
    }
    double f64() {
        unsigned char x[8] = {0};
        take(x, 8);
        uint64_t u = 0;
        for (int k = 7; k >= 0; k--) u = (u << 8) | x[k];
        double d;
        std::memcpy(&d, &u, 8);
        return d;
    }
};

void put_i32(std::vector<unsigned char> &o, int32_t v) {
    const uint32_t u = (uint32_t)v;
    for (int k = 0; k < 4; k++) o.push_back((unsigned char)((u >> (8 * k)) & 0xffu));
}
void put_f64(std::vector<unsigned char> &o, double d) {
    uint64_t u;
    std::memcpy(&u, &d, 8);
    for (int k = 0; k < 8; k++) o.push_back((unsigned char)((u >> (8 * k)) & 0xffu));
}

bool read_file(const char *path, std::vector<unsigned char> &buf, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { err = std::string("cannot open ") + path; return false; }
    buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

// ---- a small JSON reader, enough for TopologyData (moltopo/src/export.rs:12-26):
//      objects, arrays, strings, numbers, true/false/null ----
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal *get(const std::string &k) const {
        for (auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const std::string &t;
    size_t p = 0;
    std::string err;
    int depth = 0;
    explicit JParser(const std::string &text) : t(text) {}
    void ws() { while (p < t.size() && std::isspace((unsigned char)t[p])) p++; }
    bool fail(const std::string &m) {
        if (err.empty()) err = m + " at byte " + std::to_string(p);
        return false;
    }
    bool str(std::string &out) {
        if (p >= t.size() || t[p] != '"') return fail("expected a string");
        p++;
        out.clear();
        while (p < t.size() && t[p] != '"') {
            char c = t[p++];
            if (c == '\\') {
                if (p >= t.size()) return fail("bad escape");
                char e = t[p++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (p + 4 > t.size()) return fail("bad \\u escape");
                        unsigned cp = (unsigned)std::stoul(t.substr(p, 4), nullptr, 16);
                        p += 4;
                        out += cp < 0x80 ? (char)cp : '?';   // names are ASCII
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p >= t.size()) return fail("unterminated string");
        p++;
        return true;
    }
    bool value(JVal &v) {
        if (++depth > 64) return fail("nesting too deep");
        ws();
        if (p >= t.size()) return fail("unexpected end");
        const char c = t[p];
        bool ok = true;
        if (c == '{') {
            v.kind = JVal::OBJ;
            p++;
            ws();
            if (p < t.size() && t[p] == '}') { p++; depth--; return true; }
            while (ok) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (p >= t.size() || t[p] != ':') return fail("expected ':'");
                p++;
                v.obj.emplace_back(k, JVal());
                if (!value(v.obj.back().second)) return false;
                ws();
                if (p < t.size() && t[p] == ',') { p++; continue; }
                if (p < t.size() && t[p] == '}') { p++; break; }
                return fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = JVal::ARR;
            p++;
            ws();
            if (p < t.size() && t[p] == ']') { p++; depth--; return true; }
            while (ok) {
                v.arr.emplace_back();
                if (!value(v.arr.back())) return false;
                ws();
                if (p < t.size() && t[p] == ',') { p++; continue; }
                if (p < t.size() && t[p] == ']') { p++; break; }
                return fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = JVal::STR;
            if (!str(v.str)) return false;
        } else if (t.compare(p, 4, "true") == 0) {
            v.kind = JVal::BOOL; v.num = 1; p += 4;
        } else if (t.compare(p, 5, "false") == 0) {
            v.kind = JVal::BOOL; v.num = 0; p += 5;
        } else if (t.compare(p, 4, "null") == 0) {
            v.kind = JVal::NUL; p += 4;
        } else {
            const char *b = t.c_str() + p;
            char *e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) return fail("bad value");
            v.kind = JVal::NUM;
            p += (size_t)(e - b);
        }
        depth--;
        return true;
    }
};

// rows of an array of k-integer arrays
bool int_rows(const JVal *a, int k, std::vector<int> &out, std::string &err, const char *name) {
    out.clear();
    if (!a) return true;   // absent: empty
    if (a->kind != JVal::ARR) { err = std::string(name) + " is not an array"; return false; }
    for (size_t q = 0; q < a->arr.size(); q++) {
        const JVal &r = a->arr[q];
        if (r.kind != JVal::ARR || (int)r.arr.size() != k) {
            err = std::string(name) + "[" + std::to_string(q) + "] is not an array of " + std::to_string(k);
            return false;
        }
        for (const JVal &x : r.arr) {
            if (x.kind != JVal::NUM || x.num != std::floor(x.num)) {
                err = std::string(name) + "[" + std::to_string(q) + "] holds a non-integer";
                return false;
            }
            out.push_back((int)x.num);
        }
    }
    return true;
}

// shared tail of both readers: bonds + given terms -> molecule
int build_read(const char *param_dir, const std::vector<std::string> &el, const std::vector<double> &apos,
               const double *charges, const std::vector<int> &b2, const std::vector<int> &a3, const std::vector<int> &d4,
               const std::vector<int> &i4, const std::vector<std::string> *types, const std::string &who,
               smh_molecule **out) {
    std::string err;
    const ParamTables *Pp = get_params(param_dir ? param_dir : ".", err);
    if (!Pp) return host_fail(who + ": " + err);
    const int n = (int)el.size();
    std::vector<std::array<int, 2>> bonds;
    for (size_t q = 0; q < b2.size() / 2; q++) {
        const int a = b2[2 * q], c = b2[2 * q + 1];
        if (a < 0 || c < 0 || a >= n || c >= n || a == c)
            return host_fail(who + ": bond " + std::to_string(q) + " has invalid atoms (" + std::to_string(a) + "," +
                             std::to_string(c) + ")");
        bonds.push_back({a, c});
    }
    std::vector<std::array<int, 3>> ang(a3.size() / 3);
    for (size_t q = 0; q < ang.size(); q++) ang[q] = {a3[3 * q], a3[3 * q + 1], a3[3 * q + 2]};
    std::vector<std::array<int, 4>> dih(d4.size() / 4), inv(i4.size() / 4);
    for (size_t q = 0; q < dih.size(); q++) dih[q] = {d4[4 * q], d4[4 * q + 1], d4[4 * q + 2], d4[4 * q + 3]};
    for (size_t q = 0; q < inv.size(); q++) inv[q] = {i4[4 * q], i4[4 * q + 1], i4[4 * q + 2], i4[4 * q + 3]};
    GivenTerms g;
    g.angles = &ang;
    g.dihedrals = &dih;
    g.inversions = &inv;
    g.types = types;
    return build_core(*Pp, el, apos.data(), charges, bonds, g, who, out);
}

}  // namespace

extern "C" int smh_molecule_read_ufftopo(const char *param_dir, const char *path, const double *charges,
                                         smh_molecule **out) {
    const std::string who = "smh_molecule_read_ufftopo";
    if (!out || !path) return host_fail(who + ": NULL argument");
    *out = nullptr;
    std::vector<unsigned char> buf;
    std::string err;
    if (!read_file(path, buf, err)) return host_fail(who + ": " + err);
    ByteReader r(buf);
    char magic[7];
    unsigned char version = 0;
    if (!r.take(magic, 7) || std::memcmp(magic, kUffTopoMagic, 7) != 0)
        return host_fail(who + ": " + path + " does not start with \"UFFTOPO\"");
    r.take(&version, 1);
    if (version != 1) return host_fail(who + ": version " + std::to_string(version) + " (only 1 is defined)");
    const int n = r.i32(), nb = r.i32(), na = r.i32(), nd = r.i32(), ni = r.i32(), nt = r.i32();
    (void)r.i32();   // flags (reserved)
    if (!r.ok || n <= 0 || nb < 0 || na < 0 || nd < 0 || ni < 0 || nt <= 0)
        return host_fail(who + ": bad header counts");
    const size_t need = (size_t)n * 24 + (size_t)n * 4 + (size_t)nb * 8 + (size_t)na * 12 + (size_t)nd * 16 +
                        (size_t)ni * 16 + (size_t)nt * 8;
    if (buf.size() - r.p != need)
        return host_fail(who + ": " + std::to_string(buf.size() - r.p) + " payload bytes, the header implies " +
                         std::to_string(need));
    std::vector<double> apos(3 * (size_t)n);
    for (auto &x : apos) x = r.f64();
    std::vector<int> ti(n), b2(2 * (size_t)nb), a3(3 * (size_t)na), d4(4 * (size_t)nd), i4(4 * (size_t)ni);
    for (auto &x : ti) x = r.i32();
    for (auto &x : b2) x = r.i32();
    for (auto &x : a3) x = r.i32();
    for (auto &x : d4) x = r.i32();
    for (auto &x : i4) x = r.i32();
    std::vector<std::string> table(nt);
    for (auto &t : table) {
        char c[8];
        r.take(c, 8);
        std::string name(c, 8);
        while (!name.empty() && (name.back() == ' ' || name.back() == '\0')) name.pop_back();
        t = name;
    }
    if (!r.ok) return host_fail(who + ": truncated file");
    std::string perr;
    const ParamTables *Pp = get_params(param_dir ? param_dir : ".", perr);
    if (!Pp) return host_fail(who + ": " + perr);
    std::vector<std::string> types(n), el(n);
    for (int i = 0; i < n; i++) {
        if (ti[i] < 0 || ti[i] >= nt) return host_fail(who + ": atom " + std::to_string(i) + " type index out of range");
        types[i] = table[ti[i]];
        const AtomTypeRow *at = Pp->atype(types[i]);
        if (!at) return host_fail(who + ": UFF type '" + types[i] + "' is not in AtomTypes.dat");
        el[i] = at->element;
    }
    return build_read(param_dir, el, apos, charges, b2, a3, d4, i4, &types, who, out);
}

extern "C" int smh_molecule_read_json(const char *param_dir, const char *path, const double *charges,
                                      smh_molecule **out) {
    const std::string who = "smh_molecule_read_json";
    if (!out || !path) return host_fail(who + ": NULL argument");
    *out = nullptr;
    std::vector<unsigned char> buf;
    std::string err;
    if (!read_file(path, buf, err)) return host_fail(who + ": " + err);
    const std::string text(buf.begin(), buf.end());
    JParser jp(text);
    JVal root;
    if (!jp.value(root)) return host_fail(who + ": " + path + ": " + jp.err);
    if (root.kind != JVal::OBJ) return host_fail(who + ": the top level is not an object");
    const JVal *na = root.get("natoms"), *je = root.get("elements"), *jp_ = root.get("positions");
    if (!na || na->kind != JVal::NUM || !je || je->kind != JVal::ARR || !jp_ || jp_->kind != JVal::ARR)
        return host_fail(who + ": natoms, elements and positions are required");
    const int n = (int)na->num;
    if (n <= 0 || (int)je->arr.size() != n || (int)jp_->arr.size() != n)
        return host_fail(who + ": natoms = " + std::to_string(n) + " but " + std::to_string(je->arr.size()) +
                         " elements and " + std::to_string(jp_->arr.size()) + " positions");
    std::vector<std::string> el(n);
    std::vector<double> apos(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        if (je->arr[i].kind != JVal::STR) return host_fail(who + ": elements[" + std::to_string(i) + "] is not a string");
        el[i] = je->arr[i].str;
        const JVal &q = jp_->arr[i];
        if (q.kind != JVal::ARR || q.arr.size() != 3) return host_fail(who + ": positions[" + std::to_string(i) + "] is not [x,y,z]");
        for (int k = 0; k < 3; k++) {
            if (q.arr[k].kind != JVal::NUM) return host_fail(who + ": positions[" + std::to_string(i) + "] holds a non-number");
            apos[3 * (size_t)i + k] = q.arr[k].num;
        }
    }
    std::vector<int> b2, a3, d4, i4;
    if (!int_rows(root.get("bonds"), 2, b2, err, "bonds") || !int_rows(root.get("angles"), 3, a3, err, "angles") ||
        !int_rows(root.get("dihedrals"), 4, d4, err, "dihedrals") ||
        !int_rows(root.get("inversions"), 4, i4, err, "inversions"))
        return host_fail(who + ": " + err);
    // the *_params arrays carry nothing the parameters are computed from
    // (export.rs:37-42 writes them empty); UFF types follow from the bonds
    return build_read(param_dir, el, apos, charges, b2, a3, d4, i4, nullptr, who, out);
}

extern "C" int smh_molecule_positions(const smh_molecule *m, double *apos) {
    if (!m || !apos) return host_fail("smh_molecule_positions: NULL argument");
    std::copy(m->apos.begin(), m->apos.end(), apos);
    return SMH_OK;
}

extern "C" int smh_molecule_write_ufftopo(const smh_molecule *m, const char *path) {
    if (!m || !path) return host_fail("smh_molecule_write_ufftopo: NULL argument");
    // type table: sorted, deduplicated names, as buildff writes it
    std::vector<std::string> table(m->types);
    std::sort(table.begin(), table.end());
    table.erase(std::unique(table.begin(), table.end()), table.end());
    for (auto &t : table)
        if (t.size() > 8) return host_fail("smh_molecule_write_ufftopo: type name '" + t + "' longer than 8 bytes");
    std::vector<unsigned char> o(kUffTopoMagic, kUffTopoMagic + 7);
    o.push_back(1);
    const int n = m->natoms;
    put_i32(o, n);
    put_i32(o, (int)m->bonds.size() / 2);
    put_i32(o, (int)m->ang_atoms.size() / 3);
    put_i32(o, (int)m->dih_atoms.size() / 4);
    put_i32(o, (int)m->inv_atoms.size() / 4);
    put_i32(o, (int)table.size());
    put_i32(o, 0);
    for (double x : m->apos) put_f64(o, x);
    for (int i = 0; i < n; i++)
        put_i32(o, (int)(std::lower_bound(table.begin(), table.end(), m->types[i]) - table.begin()));
    for (int v : m->bonds) put_i32(o, v);
    for (int v : m->ang_atoms) put_i32(o, v);
    for (int v : m->dih_atoms) put_i32(o, v);
    for (int v : m->inv_atoms) put_i32(o, v);
    for (auto &t : table) {
        for (size_t k = 0; k < 8; k++) o.push_back(k < t.size() ? (unsigned char)t[k] : (unsigned char)' ');
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) return host_fail(std::string("smh_molecule_write_ufftopo: cannot open ") + path);
    f.write(reinterpret_cast<const char *>(o.data()), (std::streamsize)o.size());
    if (!f) return host_fail(std::string("smh_molecule_write_ufftopo: write failed: ") + path);
    return SMH_OK;
}

extern "C" int smh_molecule_write_json(const smh_molecule *m, const char *path) {
    if (!m || !path) return host_fail("smh_molecule_write_json: NULL argument");
    std::ostringstream s;
    s.precision(17);
    auto rows = [&](const char *name, const std::vector<int> &v, int k) {
        s << "  \"" << name << "\": [";
        for (size_t q = 0; q < v.size() / k; q++) {
            s << (q ? ", [" : "[");
            for (int c = 0; c < k; c++) s << (c ? ", " : "") << v[q * k + c];
            s << "]";
        }
        s << "],\n";
    };
    s << "{\n  \"natoms\": " << m->natoms << ",\n  \"elements\": [";
    for (int i = 0; i < m->natoms; i++) s << (i ? ", " : "") << "\"" << m->elems[i] << "\"";
    s << "],\n  \"positions\": [";
    for (int i = 0; i < m->natoms; i++)
        s << (i ? ", [" : "[") << m->apos[3 * i] << ", " << m->apos[3 * i + 1] << ", " << m->apos[3 * i + 2] << "]";
    s << "],\n";
    rows("bonds", m->bonds, 2);
    rows("angles", m->ang_atoms, 3);
    rows("dihedrals", m->dih_atoms, 4);
    rows("inversions", m->inv_atoms, 4);
    s << "  \"bond_params\": [],\n  \"angle_params\": [],\n  \"dihedral_params\": [],\n"
         "  \"inversion_params\": [],\n  \"atom_params\": []\n}\n";
    std::ofstream f(path);
    if (!f) return host_fail(std::string("smh_molecule_write_json: cannot open ") + path);
    f << s.str();
    if (!f) return host_fail(std::string("smh_molecule_write_json: write failed: ") + path);
    return SMH_OK;
}
