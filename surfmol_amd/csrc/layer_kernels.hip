// layer_kernels.hip -- coupled layer of molecules: cell list, FIRE move and
// intermolecular FP64 energies.
//
// A coupled layer is a batch whose systems interact: pairs of atoms of
// different molecules closer than the cutoff add LJ + Coulomb (the broad-
// phase culling of NonBondedFF::eval_broad, molff/src/nonbonded.rs:251-324,
// with the uniform cell grid of spacc/src/uniform_grid.rs:87-116), and one
// FIRE state drives every atom.  One step is
//   k_relax (KMODE_LAYER)  forces of every molecule, v += F dt, per-molecule
//                          FIRE sums
//   k_layer_decide         reduce the sums, take the FIRE decision
//   k_layer_move_count     apply the decision (mix v, x += v dt), bin atoms,
//                          write the new position into the atom's sorted slot
//                          and flag the step if any atom changed cell
//   k_cells_build          one workgroup: count (LDS atomics), cell starts
//                          (exclusive scan), atoms into their cells; clears
//                          the decision
//   k_cells_sort           order each cell by atom index, gather the sorted
//                          positions / REQ rows
// (build and sort do nothing more on a step where no atom changed cell:
// cells at least the cutoff wide keep their atoms for many steps)
//   k_layer_inter          intermolecular forces: one workgroup per (cell,
//                          neighbour cell), one to four lanes per atom of the
//                          cell; one partial force per atom and neighbour
//                          cell that the next force launch adds in a fixed
//                          order
// All kernels are launched every step (no host decision); once the layer has
// converged the force launch returns at once and the move is a no-op.
#include "smr_device.h"
#include "smr_kernels.h"

namespace smr {

namespace {

__device__ __forceinline__ int cell_of(const CellGrid &C, float gx, float gy, float gz) {
    int cx = (int)floorf((gx - C.lox) * C.inv_hx);
    int cy = (int)floorf((gy - C.loy) * C.inv_hy);
    int cz = (int)floorf((gz - C.loz) * C.inv_h);
    cx = min(max(cx, 0), C.ncx - 1);
    cy = min(max(cy, 0), C.ncy - 1);
    cz = min(max(cz, 0), C.ncz - 1);
    return (cz * C.ncy + cy) * C.ncx + cx;
}

}  // namespace

// The shared FIRE decision of a step (one wave): reduce the per-molecule sums
// of the force launch in a fixed order, then the FIRE update of
// relax_pentacene_uff.rs:162-181 (as in k_relax) and the status; the move
// is applied by k_layer_move_count.  A layer that has stopped takes no
// decision (its force launch returned at once).  A launch of its own: a
// last-finisher ticket in the force launch serialised its 512 workgroups
// on one device-scope atomic (half the step).
__global__ __launch_bounds__(64) void k_layer_decide(LayerArgs L) {
    LayerState *S = L.layer;
    if (S->status != SMR_SYS_RUNNING) return;
    const int lane = threadIdx.x;
    const int nsys = S->nsys;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    // molecules lane, lane+64, ... in order; eight loads in flight
    for (int q0 = lane; q0 < nsys; q0 += 8 * 64) {
        double4 p[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (q0 + 64 * u < nsys) p[u] = L.partials[q0 + 64 * u];
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (q0 + 64 * u < nsys) { a0 += p[u].x; a1 += p[u].y; a2 += p[u].z; a3 = fmax(a3, p[u].w); }
    }
    const double vf = wave_sum(a0), vv = wave_sum(a1), ff = wave_sum(a2), fm = wave_max(a3);
    if (lane != 0) return;
    const FireConst &P = L.fire;
    const double dt_old = S->dt;
    const bool nonfinite = !(isfinite(vf) && isfinite(vv) && isfinite(ff));
    if (vf > 0.0 && ff > 1e-30 && vv > 1e-30) {
        S->mix_a = (float)(1.0 - S->alpha);
        S->mix_b = (float)(S->alpha * sqrt(vv) / sqrt(ff));
        S->reset = 0;
        S->npos++;
        if (S->npos > P.n_min) {
            S->dt = fmin(S->dt * P.f_inc, P.dt_max);
            S->alpha *= P.f_alpha;
        }
    } else {
        S->npos = 0;
        S->dt *= P.f_dec;
        S->alpha = P.alpha0;
        S->mix_a = S->mix_b = 0.f;
        S->reset = 1;
    }
    S->dt_move = (float)dt_old;
    S->steps++;
    S->fmax = sqrt(fm);
    if (nonfinite) S->status = SMR_SYS_NONFINITE;
    else if (S->fmax < P.fconv) S->status = SMR_SYS_CONVERGED;
    else if (S->steps >= P.max_steps) S->status = SMR_SYS_MAXSTEPS;
    S->pending_move = 1;
}

// Apply the FIRE decision of the step (if one is pending), then bin the atom.
__global__ void k_layer_move_count(LayerArgs L, int move) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= L.natoms) return;
    float4 p = L.pos[a];
    if (move && L.layer->pending_move) {
        const LayerState &S = *L.layer;
        float3 v;
        if (S.reset) {
            v = make_float3(0.f, 0.f, 0.f);
        } else {
            const float4 v4 = L.vel[a], F = L.force[a];
            v = make_float3(v4.x * S.mix_a + F.x * S.mix_b, v4.y * S.mix_a + F.y * S.mix_b, v4.z * S.mix_a + F.z * S.mix_b);
        }
        L.vel[a] = make_float4(v.x, v.y, v.z, 0.f);
        p.x += v.x * S.dt_move;
        p.y += v.y * S.dt_move;
        p.z += v.z * S.dt_move;
        L.pos[a] = p;
    }
    const SysDesc sd = L.tab.sys[L.atom_sys[a]];
    const CellGrid &C = L.cells;
    float gx = p.x + (float)sd.gx * L.tab.ostep_x, gy = p.y + (float)sd.gy * L.tab.ostep_y;
    // periodic axes: the atom's image in the box [0, L) and how many box
    // lengths it was moved by (an exact integer used by the pair kernels)
    int2 w = make_int2(0, 0);
    if (C.pbc_x) {
        w.x = (int)floorf(gx * C.inv_Lx);
        gx -= (float)w.x * C.Lx;
    }
    if (C.pbc_y) {
        w.y = (int)floorf(gy * C.inv_Ly);
        gy -= (float)w.y * C.Ly;
    }
    L.atom_wrap[a] = w;
    const int c = cell_of(C, gx, gy, p.z + (float)sd.gz * L.tab.ostep_z);
    // without a move (set-up, evaluation) the list is always rebuilt
    if (!move || c != L.atom_cell[a]) L.layer->cells_dirty = 1;
    L.atom_cell[a] = c;
    if (move) {
        // the sorted slot of the last rebuild (rewritten by the sort when the
        // list is rebuilt this step)
        const int q = L.slot_of[a];
        L.spos[q] = make_float4(p.x, p.y, p.z, __int_as_float(L.atom_sys[a]));
        L.swrap[q] = w;
    }
}

// The cell list of the atoms' cells (one workgroup of 1024 threads): count
// the atoms of each cell, exclusive scan into the cell starts, scatter the
// atom indices into their cells (in no particular order inside a cell:
// k_cells_sort orders them).  Counts and fill cursors live in LDS when the
// grid has at most kCellsLds cells, else in the global count array.  Runs
// only on a step where an atom changed cell; every step it clears the
// applied decision and resets the next force launch's work counters.
// (Counting here rather than in the move keeps the per-atom device-scope
// atomics off the steps whose cells do not change.)
constexpr int kCellsLds = 12288;

template <bool LDS>
__device__ void build_cells(const LayerArgs &L, int n, int *cnt, int *part) {
    const int t = threadIdx.x;
    const int N = L.natoms;
    for (int c = t; c < n; c += 1024) cnt[c] = 0;
    if (!LDS) __threadfence();
    __syncthreads();
    // count: eight cell indices in flight per thread
    for (int a0 = t; a0 < N; a0 += 8 * 1024) {
        int cl[8];
#pragma unroll
        for (int u = 0; u < 8; u++) cl[u] = a0 + u * 1024 < N ? L.atom_cell[a0 + u * 1024] : -1;
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (cl[u] >= 0) atomicAdd(&cnt[cl[u]], 1);
    }
    if (!LDS) __threadfence();
    __syncthreads();
    // exclusive scan: each thread a contiguous chunk, chunk totals in LDS
    const int chunk = (n + 1023) / 1024;
    const int c0 = min(n, t * chunk), c1 = min(n, c0 + chunk);
    int sum = 0;
    for (int c = c0; c < c1; c++) sum += cnt[c];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;   // exclusive prefix of this chunk
    for (int c = c0; c < c1; c++) {
        const int k = cnt[c];
        L.start[c] = run;
        cnt[c] = run;          // from here on: the cell's fill cursor
        run += k;
    }
    if (t == 1023) L.start[n] = part[1023];
    if (!LDS) __threadfence();
    __syncthreads();
    // scatter
    for (int a0 = t; a0 < N; a0 += 8 * 1024) {
        int cl[8];
#pragma unroll
        for (int u = 0; u < 8; u++) cl[u] = a0 + u * 1024 < N ? L.atom_cell[a0 + u * 1024] : -1;
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (cl[u] >= 0) L.sidx[atomicAdd(&cnt[cl[u]], 1)] = a0 + u * 1024;
    }
}

__global__ __launch_bounds__(1024) void k_cells_build(LayerArgs L) {
    __shared__ int part[1024];
    __shared__ int cnt[kCellsLds];
    const int t = threadIdx.x;
    if (t == 0) L.layer->pending_move = 0;
    if (L.work_counters && t < L.nwork_counters) L.work_counters[t] = 0;
    if (!L.layer->cells_dirty) return;   // no atom changed cell: starts and order stand
    const int n = L.cells.ncx * L.cells.ncy * L.cells.ncz;
    if (n <= kCellsLds) build_cells<true>(L, n, cnt, part);
    else build_cells<false>(L, n, L.count, part);
}

// One workgroup per cell: order the cell's atoms by index (bitonic network
// in LDS over the next power of two >= the cell's atoms, at most 256; a
// longer cell is sorted serially), then gather positions (w = system) and
// REQ rows in that order.
__global__ __launch_bounds__(256) void k_cells_sort(LayerArgs L) {
    __shared__ int key[256];
    const int c = blockIdx.x, t = threadIdx.x;
    if (!L.layer->cells_dirty) return;
    const int st = L.start[c], n = L.start[c + 1] - st;
    if (n == 0) return;
    if (n <= 256) {
        int w = 2;
        while (w < n) w <<= 1;
        key[t] = t < n ? L.sidx[st + t] : 0x7fffffff;
        __syncthreads();
        for (int k = 2; k <= w; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                const int o = t ^ j;
                if (o > t) {
                    const int a = key[t], b = key[o];
                    if ((a > b) == ((t & k) == 0)) {
                        key[t] = b;
                        key[o] = a;
                    }
                }
                __syncthreads();
            }
        if (t < n) {
            const int a = key[t];
            L.sidx[st + t] = a;
            L.slot_of[a] = st + t;
            const float4 p = L.pos[a];
            L.spos[st + t] = make_float4(p.x, p.y, p.z, __int_as_float(L.atom_sys[a]));
            L.sreq[st + t] = L.req_atom[a];
            L.swrap[st + t] = L.atom_wrap[a];
        }
    } else if (t == 0) {
        // a crowded cell: insertion sort and gather by one thread
        for (int q = st + 1; q < st + n; q++) {
            const int v = L.sidx[q];
            int r = q - 1;
            while (r >= st && L.sidx[r] > v) {
                L.sidx[r + 1] = L.sidx[r];
                r--;
            }
            L.sidx[r + 1] = v;
        }
        for (int q = st; q < st + n; q++) {
            const int a = L.sidx[q];
            L.slot_of[a] = q;
            const float4 p = L.pos[a];
            L.spos[q] = make_float4(p.x, p.y, p.z, __int_as_float(L.atom_sys[a]));
            L.sreq[q] = L.req_atom[a];
            L.swrap[q] = L.atom_wrap[a];
        }
    }
}

// Intermolecular forces on the atoms of cell c from the atoms of its
// neighbour cell number o (0..K-1, K = 27, or 9 / 3 / 1 when the grid is a
// single cell thick along some axis): partners of other molecules within
// the cutoff, Kahan-summed in slot order by each of the atom's 1, 2 or 4
// lanes (every 1st, 2nd or 4th partner), the lanes' sums then added
// pairwise.  Writes part[o][slot] for every atom of the cell (zero when the
// neighbour cell lies outside the grid).
// Along a periodic axis the neighbour cell wraps around the box and carries
// an image shift; the displacement of a pair is then
//   (pj - pi) + ostep * ((gj - gi) + (image - wj + wi) * nL)
// with g the systems' origin offsets and w the atoms' box wraps, all exact
// integers, and a molecule meets its own atoms in every other image.
__global__ __launch_bounds__(64) void k_layer_inter(LayerArgs L) {
    const CellGrid &C = L.cells;
    const DevTables &T = L.tab;
    const int K = C.nox * C.noy * C.noz;
    const int c = blockIdx.x / K, o = blockIdx.x - K * c;
    // the move, scan, scatter and sort of this step have read the flag
    if (blockIdx.x == 0 && threadIdx.x == 0) L.layer->cells_dirty = 0;
    const int s0 = C.start[c], n = C.start[c + 1] - s0;
    if (n == 0) return;
    const int cx = c % C.ncx, cy = (c / C.ncx) % C.ncy, cz = c / (C.ncx * C.ncy);
    // neighbour offsets only along axes with more than one cell
    int x = cx + (C.nox == 3 ? o % 3 - 1 : 0);
    int y = cy + (C.noy == 3 ? (o / C.nox) % 3 - 1 : 0);
    const int z = cz + (C.noz == 3 ? o / (C.nox * C.noy) - 1 : 0);
    int imx = 0, imy = 0;
    if (C.pbc_x) {
        imx = x < 0 ? -1 : (x >= C.ncx ? 1 : 0);
        x -= imx * C.ncx;
    }
    if (C.pbc_y) {
        imy = y < 0 ? -1 : (y >= C.ncy ? 1 : 0);
        y -= imy * C.ncy;
    }
    const bool inside = x >= 0 && x < C.ncx && y >= 0 && y < C.ncy && z >= 0 && z < C.ncz;
    const int cc = inside ? (z * C.ncy + y) * C.ncx + x : 0;
    const int q0 = inside ? C.start[cc] : 0, q1 = inside ? C.start[cc + 1] : 0;
    float4 *out = L.part + (size_t)o * L.natoms;
    // nl adjacent lanes per atom when the cell is small (cells are about the
    // cutoff wide and hold ~30 atoms): sub-lane k takes partners k, k+nl, ...
    // of the neighbour cell, so the wave's lanes stay busy
    const int nl = n <= 16 ? 4 : (n <= 32 ? 2 : 1);
    const int k = threadIdx.x & (nl - 1);
    for (int t0 = 0; t0 < n; t0 += 64 / nl) {
        const int t = t0 + (int)threadIdx.x / nl;
        float3 v = make_float3(0.f, 0.f, 0.f);
        if (t < n) {
            const float4 pi = C.spos[s0 + t];
            const float4 ri = C.sreq[s0 + t];
            const int si = __float_as_int(pi.w);
            const SysDesc di = T.sys[si];
            const int2 wi = C.swrap[s0 + t];
            KahanF3 acc;
            acc.init(make_float3(0.f, 0.f, 0.f));
            int last = -1, ogx = 0, ogy = 0, ogz = 0;
            for (int q = q0 + k; q < q1; q += nl) {
                const float4 pj = C.spos[q];
                const int sj = __float_as_int(pj.w);
                const int2 wj = C.swrap[q];
                // box lengths between the images (0 on an open axis)
                const int bx = imx - wj.x + wi.x, by = imy - wj.y + wi.y;
                if (sj == si && bx == 0 && by == 0) continue;   // the molecule itself: intramolecular
                if (sj != last) {
                    const SysDesc dj = T.sys[sj];
                    ogx = dj.gx - di.gx;
                    ogy = dj.gy - di.gy;
                    ogz = dj.gz - di.gz;
                    last = sj;
                }
                const float dgx = (float)(ogx + bx * C.nLx) * T.ostep_x;
                const float dgy = (float)(ogy + by * C.nLy) * T.ostep_y;
                const float dgz = (float)ogz * T.ostep_z;
                const float3 dp = make_float3((pj.x - pi.x) + dgx, (pj.y - pi.y) + dgy, (pj.z - pi.z) + dgz);
                const float r2 = dot3(dp, dp);
                if (r2 >= C.rc2) continue;
                float3 f;
                const float e = ljq_pair(dp, ri, C.sreq[q], T.r2damp, f);
                if (r2 > C.ron2) {
                    // switching shell: E S, force (fr S + E S'/r) dp (1/r from
                    // one reciprocal square root: no divide, no sqrt)
                    const float ir = frsq(r2), r = r2 * ir, x = (r - C.ron) * C.inv_w;
                    const float S = 1.f - x * x * (3.f - 2.f * x), dS_r = 6.f * x * (x - 1.f) * C.inv_w * ir;
                    f = f * S + dp * (e * dS_r);
                }
                acc.add(f);
            }
            v = acc.value();
        }
        // the atom's nl partial sums, added pairwise in a fixed order
        for (int m = 1; m < nl; m <<= 1)
            v = v + make_float3(__shfl_xor(v.x, m), __shfl_xor(v.y, m), __shfl_xor(v.z, m));
        if (t < n && k == 0) out[s0 + t] = make_float4(v.x, v.y, v.z, 0.f);
    }
}

// FP64 intermolecular pair energies of the listed systems (added to E[4]):
// every partner of another system within the cutoff, counted at this
// system's end (each pair appears once in each of its two systems).
__global__ __launch_bounds__(256) void k_energy64_inter(LayerArgs L, SysState *state, const int *list) {
    const int s = list[blockIdx.x];
    const DevTables &T = L.tab;
    const SysDesc sd = T.sys[s];
    const TopoDesc td = T.topo[sd.topo];
    const CellGrid &C = L.cells;
    const double rc2 = C.rc2_64;
    double e = 0.0;
    for (int i = threadIdx.x; i < td.natoms; i += blockDim.x) {
        const int a = sd.atom_base + i;
        const float4 pi = L.pos[a];
        const double4 ri = T.req64[td.atom_off + i];
        const int c = L.atom_cell[a];
        const int2 wi = L.atom_wrap[a];
        const int cx = c % C.ncx, cy = (c / C.ncx) % C.ncy, cz = c / (C.ncx * C.ncy);
        // neighbour cells: clipped on open axes, wrapped with an image on periodic ones
        const int x0 = C.pbc_x ? cx - 1 : max(cx - 1, 0), x1 = C.pbc_x ? cx + 1 : min(cx + 1, C.ncx - 1);
        const int y0 = C.pbc_y ? cy - 1 : max(cy - 1, 0), y1 = C.pbc_y ? cy + 1 : min(cy + 1, C.ncy - 1);
        for (int z = max(cz - 1, 0); z <= min(cz + 1, C.ncz - 1); z++)
            for (int yy = y0; yy <= y1; yy++)
                for (int xx = x0; xx <= x1; xx++) {
                    const int imx = xx < 0 ? -1 : (xx >= C.ncx ? 1 : 0), imy = yy < 0 ? -1 : (yy >= C.ncy ? 1 : 0);
                    const int x = xx - imx * C.ncx, y = yy - imy * C.ncy;
                    const int cc = (z * C.ncy + y) * C.ncx + x;
                    for (int q = C.start[cc]; q < C.start[cc + 1]; q++) {
                        const float4 pj = C.spos[q];
                        const int sj = __float_as_int(pj.w);
                        const int2 wj = C.swrap[q];
                        const int bx = imx - wj.x + wi.x, by = imy - wj.y + wi.y;
                        if (sj == s && bx == 0 && by == 0) continue;
                        const SysDesc dj = T.sys[sj];
                        const double dx = ((double)pj.x - (double)pi.x) +
                                          (double)((dj.gx - sd.gx) + bx * C.nLx) * T.ostep64_x;
                        const double dy = ((double)pj.y - (double)pi.y) +
                                          (double)((dj.gy - sd.gy) + by * C.nLy) * T.ostep64_y;
                        const double dz = ((double)pj.z - (double)pi.z) + (double)(dj.gz - sd.gz) * T.ostep64_z;
                        const double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 >= rc2) continue;
                        const int aj = L.sidx[q];
                        const int lj = aj - T.sys[sj].atom_base;
                        const double4 rj = T.req64[T.topo[dj.topo].atom_off + lj];
                        const double R0 = ri.x + rj.x, E0 = ri.y * rj.y, kQ = ri.z * rj.z;
                        const double u2 = R0 * R0 / r2;
                        const double u6 = u2 * u2 * u2;
                        double eij = (u6 - 2.0) * u6 * E0 + kQ / sqrt(r2 + T.r2damp64);
                        if (C.w64 > 0.0 && r2 > C.ron64 * C.ron64) {
                            const double x = (sqrt(r2) - C.ron64) / C.w64;
                            eij *= 1.0 - x * x * (3.0 - 2.0 * x);
                        }
                        e += eij;
                    }
                }
    }
    __shared__ double red[4];
    const double r = wave_sum(e);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) acc += red[w];
        state[s].E[4] += acc;
    }
}

hipError_t launch_layer_cells(const LayerArgs &L, bool move, hipStream_t st) {
    const int nb = (L.natoms + 255) / 256;
    const int ncells = L.cells.ncx * L.cells.ncy * L.cells.ncz;
    hipLaunchKernelGGL(k_layer_move_count, dim3(nb), dim3(256), 0, st, L, move ? 1 : 0);
    hipLaunchKernelGGL(k_cells_build, dim3(1), dim3(1024), 0, st, L);
    hipLaunchKernelGGL(k_cells_sort, dim3(ncells), dim3(256), 0, st, L);
    hipLaunchKernelGGL(k_layer_inter, dim3(ncells * L.cells.nox * L.cells.noy * L.cells.noz), dim3(64), 0, st, L);
    return hipGetLastError();
}

hipError_t launch_layer_decide(const LayerArgs &L, hipStream_t st) {
    hipLaunchKernelGGL(k_layer_decide, dim3(1), dim3(64), 0, st, L);
    return hipGetLastError();
}

hipError_t launch_energy64_inter(const LayerArgs &L, SysState *state, const int *list, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_energy64_inter, dim3(n), dim3(256), 0, st, L, state, list);
    return hipGetLastError();
}

}  // namespace smr
