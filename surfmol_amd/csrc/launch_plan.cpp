// launch_plan.cpp -- size classes, launch shapes and the launch sequence of
// the fused relax kernel (host side).
//
// Systems of uneven size are grouped into classes by atom count (<= 64, 128,
// 256, 512, 1024).  Each class gets its own launch shape (threads per
// workgroup, atoms per thread, force pieces in LDS or in global scratch,
// pair sweep) and one persistent launch that drains the class's slice of the
// work list through an atomic counter.  The launches of one chunk are issued
// on up to four streams (fork/join on the context stream) so that the tail
// of one class overlaps the next, and the whole sequence is captured once
// into a HIP graph and replayed while nothing in it changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "smr_internal.h"

namespace smr {

int size_class(int natoms) {
    for (int k = 0; k < NCLASS; k++)
        if (natoms <= kClassMaxAtoms[k]) return k;
    return -1;
}

LaunchShape make_shape(const smr_ctx *c, const ClassPlan &cp, int block, int apt, bool fint_lds, bool pair_mode,
                       bool stencil_cache, bool have_grid) {
    LaunchShape L;
    L.block = 0;
    if (block * apt < cp.max_atoms) return L;
    const void *k = relax_kernel_ptr(block, apt, fint_lds);
    if (!k) return L;
    const size_t lds = relax_lds_bytes(block, apt, relax_lds_atoms(block, cp.max_atoms), cp.nf_max, cp.na_max, cp.nd_max, cp.ni_max, fint_lds, pair_mode);
    if (lds > c->max_lds) return L;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, block, lds) != hipSuccess || occ < 1) occ = 1;
    const long long cap = (long long)occ * c->num_cu;
    L.block = block;
    L.apt = apt;
    L.fint_lds = fint_lds;
    L.pair_mode = pair_mode;
    L.stencil_cache = stencil_cache && have_grid;
    L.lds_bytes = lds;
    L.grid = (int)std::max<long long>(1, std::min<long long>(cap, cp.count));
    return L;
}

// preference among the variants of one (block, apt): unordered-pair sweep
// when its per-wave buffers fit (unless pairs = 0 forbids it, or 1 demands
// it), force pieces in LDS when the total stays within 64 KB (two or more
// workgroups per CU)
static LaunchShape best_variant(const smr_ctx *c, const ClassPlan &cp, int block, int apt, bool cache, bool have_grid,
                                int pairs) {
    const size_t lds_cap = std::min<size_t>(c->max_lds, 64 * 1024);
    for (int pm = 1; pm >= 0; pm--) {
        if ((pairs == 0 && pm == 1) || (pairs == 1 && pm == 0)) continue;
        const size_t with = relax_lds_bytes(block, apt, relax_lds_atoms(block, cp.max_atoms), cp.nf_max, cp.na_max, cp.nd_max, cp.ni_max, true, pm);
        const bool fl = with <= lds_cap;
        LaunchShape L = make_shape(c, cp, block, apt, fl, pm, cache, have_grid);
        if (L.block) return L;
        if (fl) {
            L = make_shape(c, cp, block, apt, false, pm, cache, have_grid);
            if (L.block) return L;
        }
    }
    LaunchShape none;
    none.block = 0;
    return none;
}

LaunchShape class_shape(const smr_ctx *c, const ClassPlan &cp, bool have_grid, int tune_block, int tune_cache,
                        int tune_pairs) {
    // one atom per thread up to 512 atoms: a large system gets 8 waves (two
    // per SIMD of its CU) rather than 4 waves with 2 atoms per thread
    static const int def_block[NCLASS] = {64, 128, 256, 512, 256};
    static const int def_apt[NCLASS] = {1, 1, 1, 1, 4};
    // stored surface stencils: on unless switched off
    const bool cache = (tune_cache != 0) && have_grid;
    if (tune_block > 0) {
        int apt = (cp.max_atoms + tune_block - 1) / tune_block;
        apt = apt <= 1 ? 1 : apt <= 2 ? 2 : 4;
        LaunchShape L = best_variant(c, cp, tune_block, apt, cache, have_grid, tune_pairs);
        if (L.block) return L;
    }
    const TunedShape &t = c->tuned[cp.cls];
    if (t.valid && tune_block <= 0 && (tune_pairs < 0 || tune_pairs == (int)t.pair_mode)) {
        LaunchShape L = make_shape(c, cp, t.block, t.apt, t.fint_lds, t.pair_mode, cache, have_grid);
        if (L.block) return L;
    }
    const int b = def_block[cp.cls], a = def_apt[cp.cls];
    LaunchShape L = best_variant(c, cp, b, a, cache, have_grid, tune_pairs);
    if (L.block) return L;
    // the default does not fit this class's LDS needs: any other shape that
    // holds its atoms, fewest atoms per thread first
    static const int fb[][2] = {{512, 1}, {256, 1}, {512, 2}, {256, 2}, {128, 2}, {256, 4}, {128, 4}, {64, 4}};
    for (const auto &f : fb) {
        if (f[0] * f[1] < cp.max_atoms) continue;
        L = best_variant(c, cp, f[0], f[1], cache, have_grid, tune_pairs);
        if (L.block) return L;
    }
    return L;
}

void build_plans(std::vector<SysSize> sys, std::vector<ClassPlan> &plans, std::vector<int> &work) {
    std::stable_sort(sys.begin(), sys.end(), [](const SysSize &a, const SysSize &b) {
        if (a.natoms != b.natoms) return a.natoms > b.natoms;
        return a.id < b.id;
    });
    plans.clear();
    work.clear();
    work.reserve(sys.size());
    for (size_t q = 0; q < sys.size(); q++) {
        const SysSize &z = sys[q];
        const int k = size_class(z.natoms);
        if (plans.empty() || plans.back().cls != k) {
            ClassPlan cp;
            cp.cls = k;
            cp.first = (int)q;
            plans.push_back(cp);
        }
        ClassPlan &cp = plans.back();
        cp.count++;
        cp.max_atoms = std::max(cp.max_atoms, z.natoms);
        cp.nf_max = std::max(cp.nf_max, z.nf);
        cp.na_max = std::max(cp.na_max, z.na);
        cp.nd_max = std::max(cp.nd_max, z.nd);
        cp.ni_max = std::max(cp.ni_max, z.ni);
        work.push_back(z.id);
    }
}

hipError_t StreamSet::create() {
    hipError_t e = hipStreamCreateWithFlags(&main, hipStreamNonBlocking);
    for (int a = 0; a < 3 && e == hipSuccess; a++) e = hipStreamCreateWithFlags(&aux[a], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    for (int a = 0; a < 3 && e == hipSuccess; a++) e = hipEventCreateWithFlags(&join[a], hipEventDisableTiming);
    if (e != hipSuccess) destroy();
    return e;
}

void StreamSet::destroy() {
    if (main) (void)hipStreamSynchronize(main);
    for (int a = 0; a < 3; a++) {
        if (aux[a]) { (void)hipStreamSynchronize(aux[a]); (void)hipStreamDestroy(aux[a]); }
        if (join[a]) (void)hipEventDestroy(join[a]);
        aux[a] = nullptr;
        join[a] = nullptr;
    }
    if (fork) (void)hipEventDestroy(fork);
    if (main) (void)hipStreamDestroy(main);
    fork = nullptr;
    main = nullptr;
}

LaunchSet::~LaunchSet() { drop_graph(); }

void LaunchSet::drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    key.clear();
}

namespace {
template <class T>
void put(std::vector<unsigned char> &k, const T &v) {
    const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
    k.insert(k.end(), p, p + sizeof(T));
}
}  // namespace

int prepare_plans(LaunchSet &ls, const std::vector<ClassPlan> &plans, const KernelArgs &base,
                  std::vector<KernelArgs> &args) {
    hipError_t e;
    // force-piece scratch of the classes whose pieces do not fit LDS
    std::vector<size_t> off(plans.size(), 0);
    size_t need = 0;
    for (size_t q = 0; q < plans.size(); q++) {
        const ClassPlan &p = plans[q];
        if (p.shape.block == 0)
            return fail(SMR_ERR_UNSUPPORTED, "k_relax: no launch shape allowed by the launch settings fits " +
                                                 std::to_string(p.max_atoms) + " atoms with " + std::to_string(p.nf_max) +
                                                 " force pieces in the device's LDS (smr_batch_set_launch with "
                                                 "pair_sweep = 1 demands per-wave pair buffers)");
        off[q] = need;
        if (!p.shape.fint_lds) need += (size_t)3 * std::max(1, p.nf_max) * p.shape.grid;
    }
    if (need) need += kPieceTail;   // the gather's reads past the last class's last piece
    if (need && ls.scratch.bytes < need * sizeof(float)) {
        ls.drop_graph();
        if ((e = ls.scratch.alloc(need * sizeof(float))) != hipSuccess) return hip_fail(e, "force-piece scratch");
    }
    if (!ls.counters.ptr) {
        ls.drop_graph();
        if ((e = ls.counters.alloc(sizeof(int) * NCLASS)) != hipSuccess) return hip_fail(e, "work counters");
    }
    // byte copies keep the caller's zeroed padding, so a graph key compares
    // only meaningful bytes
    args.assign(plans.size(), KernelArgs());
    for (size_t q = 0; q < plans.size(); q++) {
        const ClassPlan &p = plans[q];
        KernelArgs &A = args[q];
        memcpy(&A, &base, sizeof(KernelArgs));
        A.work_list = base.work_list + p.first;
        A.nwork = p.count;
        if (p.steps > 0) A.steps_per_launch = p.steps;
        A.work_counter = (int *)ls.counters.ptr + p.cls;
        A.fint_scratch = (float *)ls.scratch.ptr + off[q];
        A.nf_max = std::max(1, p.nf_max);
        A.lds_atoms = relax_lds_atoms(p.shape.block, p.max_atoms);
        A.lds_na = p.na_max;
        A.lds_nd = p.nd_max;
        A.lds_ni = p.ni_max;
        A.pair_mode = p.shape.pair_mode ? 1 : 0;
        A.stencil = p.shape.stencil_cache ? base.stencil : nullptr;
    }
    return SMR_OK;
}

hipError_t issue_plans(const StreamSet &ss, LaunchSet &ls, const std::vector<ClassPlan> &plans,
                       const std::vector<KernelArgs> &args, bool zero_counters) {
    const int naux = std::min<int>(3, (int)plans.size() - 1);
    hipError_t r = hipSuccess;
    // the work counters start at zero (a coupled layer's cell build zeroes
    // them for the next step instead)
    if (zero_counters && (r = launch_zero_i32((int *)ls.counters.ptr, NCLASS, ss.main)) != hipSuccess) return r;
    if (naux > 0) {
        if ((r = hipEventRecord(ss.fork, ss.main)) != hipSuccess) return r;
        for (int a = 0; a < naux; a++)
            if ((r = hipStreamWaitEvent(ss.aux[a], ss.fork, 0)) != hipSuccess) return r;
    }
    for (size_t q = 0; q < plans.size(); q++) {
        hipStream_t st = q == 0 ? ss.main : ss.aux[(q - 1) % 3];
        if ((r = launch_relax(args[q], plans[q].shape, st)) != hipSuccess) return r;
    }
    for (int a = 0; a < naux; a++) {
        if ((r = hipEventRecord(ss.join[a], ss.aux[a])) != hipSuccess) return r;
        if ((r = hipStreamWaitEvent(ss.main, ss.join[a], 0)) != hipSuccess) return r;
    }
    return hipSuccess;
}

int launch_plans_async(const StreamSet &ss, LaunchSet &ls, const std::vector<ClassPlan> &plans, const KernelArgs &base,
                       hipEvent_t ev_begin, hipEvent_t ev_end) {
    hipError_t e;
    std::vector<KernelArgs> args;
    int rc = prepare_plans(ls, plans, base, args);
    if (rc) return rc;
    auto issue = [&]() -> hipError_t { return issue_plans(ss, ls, plans, args); };
    if (ls.use_graph) {
        std::vector<unsigned char> key;
        for (size_t q = 0; q < plans.size(); q++) {
            put(key, args[q]);
            const LaunchShape &L = plans[q].shape;
            const int sk[6] = {L.block, L.apt, L.grid, (int)L.fint_lds, (int)L.stencil_cache, (int)L.pair_mode};
            put(key, sk);
            put(key, L.lds_bytes);
        }
        if (!ls.gexec || key != ls.key) {
            ls.drop_graph();
            if ((e = hipStreamBeginCapture(ss.main, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                return hip_fail(e, "hipStreamBeginCapture");
            hipError_t ei = issue();
            e = hipStreamEndCapture(ss.main, &ls.graph);
            if (ei != hipSuccess) { ls.drop_graph(); return hip_fail(ei, "capture of the relax launches"); }
            if (e != hipSuccess) { ls.drop_graph(); return hip_fail(e, "hipStreamEndCapture"); }
            if ((e = hipGraphInstantiate(&ls.gexec, ls.graph, nullptr, nullptr, 0)) != hipSuccess) {
                ls.drop_graph();
                return hip_fail(e, "hipGraphInstantiate");
            }
            ls.key = key;
        }
    }
    e = ev_begin ? hipEventRecord(ev_begin, ss.main) : hipSuccess;
    if (e == hipSuccess) e = ls.use_graph ? hipGraphLaunch(ls.gexec, ss.main) : issue();
    if (e == hipSuccess && ev_end) e = hipEventRecord(ev_end, ss.main);
    if (e != hipSuccess) return hip_fail(e, "k_relax launches");
    return SMR_OK;
}

int launch_plans(const StreamSet &ss, LaunchSet &ls, const std::vector<ClassPlan> &plans, const KernelArgs &base,
                 float *kernel_ms) {
    if (plans.empty()) {
        if (kernel_ms) *kernel_ms = 0.f;
        return SMR_OK;
    }
    hipError_t e;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (kernel_ms) {
        if ((e = hipEventCreate(&e0)) != hipSuccess) return hip_fail(e, "hipEventCreate");
        if ((e = hipEventCreate(&e1)) != hipSuccess) { (void)hipEventDestroy(e0); return hip_fail(e, "hipEventCreate"); }
    }
    int rc = launch_plans_async(ss, ls, plans, base, e0, e1);
    e = hipSuccess;
    if (rc == SMR_OK) e = hipStreamSynchronize(ss.main);
    if (rc == SMR_OK && e == hipSuccess && kernel_ms) e = hipEventElapsedTime(kernel_ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "k_relax launches");
    return SMR_OK;
}

}  // namespace smr
