// engine_kernels.hip - __global__ kernels of the FibreNet step (gfx950).
//
// Granular kernels reproduce one reference kernel each (ExecuteKernel by
// name); the fused kernel does the whole step of one bead - forces, closing
// kick and the next step's drift - so that a step is one pass over HBM.
#include "kernels.h"
#include "engine.h"

namespace fnk {

// Logical block index spread so that consecutive bead ranges stay on one XCD
// (blocks are dealt round-robin over the 8 XCDs, each with its own L2).
__device__ __forceinline__ int xcd_block()
{
    const int nb = gridDim.x, b = blockIdx.x;
    const int per = nb / 8;
    if (per == 0 || b >= per * 8) return b;
    return (b & 7) * per + (b >> 3);
}

template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_advance_position(Args<T, P> a, int n_active)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_active) return;
    const T fx = a.fx[i], fy = a.fy[i], fz = a.fz[i];
    P px = a.px[i], py = a.py[i], pz = a.pz[i];
    P ix = a.ix[i], iy = a.iy[i], iz = a.iz[i];
    V3<T> vp;
    drift(a.drift, fx, fy, fz, a.wx[i], a.wy[i], a.wz[i], a.lam[i], a.hlf2[i], px, py, pz, ix, iy, iz, vp);
    const V3<T> f = base_force<T, P>(a.drift, vp, px, py, pz);
    a.px[i] = px; a.py[i] = py; a.pz[i] = pz;
    if (a.ix != a.px) { a.ix[i] = ix; a.iy[i] = iy; a.iz[i] = iz; }
    a.vx[i] = vp.x; a.vy[i] = vp.y; a.vz[i] = vp.z;
    a.gx[i] = fx; a.gy[i] = fy; a.gz[i] = fz;        // force_old = force
    a.fx[i] = f.x; a.fy[i] = f.y; a.fz[i] = f.z;
}

template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_resolve_springs(Args<T, P> a)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    if (a.sp_start[i] == a.sp_start[i + 1]) return;
    V3<T> f{a.fx[i], a.fy[i], a.fz[i]};
    gather_springs(a, i, f);
    a.fx[i] = f.x; a.fy[i] = f.y; a.fz[i] = f.z;
}

template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_resolve_angles(Args<T, P> a)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    if (a.an_start[i] == a.an_start[i + 1]) return;
    V3<T> f{a.fx[i], a.fy[i], a.fz[i]};
    gather_angles<T, P, false>(a, i, f);
    a.fx[i] = f.x; a.fy[i] = f.y; a.fz[i] = f.z;
}

template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_resolve_contacts(Args<T, P> a)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    if (a.nb_start[i] == a.nb_start[i + 1]) return;
    V3<T> f{a.fx[i], a.fy[i], a.fz[i]};
    gather_contacts(a, i, f, a.pairs);
    a.fx[i] = f.x; a.fy[i] = f.y; a.fz[i] = f.z;
}

// AdvanceVelocity (kernel.cl:444-464); also counts the step.
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_advance_velocity(Args<T, P> a, int n_active)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i == 0) *a.step += 1;
    if (i >= n_active) return;
    const T h = a.hlf[i];
    T sx = a.fx[i], sy = a.fy[i], sz = a.fz[i];
    sx += a.gx[i];
    sy += a.gy[i];
    sz += a.gz[i];
    a.wx[i] = tfma(sx, h, a.wx[i]);
    a.wy[i] = tfma(sy, h, a.wy[i]);
    a.wz[i] = tfma(sz, h, a.wz[i]);
}

// ResetBeadEnergy (kernel.cl:468-483).
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_reset_energy(Args<T, P> a, int n_active)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_active) return;
    a.fx[i] = a.fy[i] = a.fz[i] = (T)0;
    a.gx[i] = a.gy[i] = a.gz[i] = (T)0;
    a.vx[i] = a.vy[i] = a.vz[i] = (T)0;
    a.wx[i] = a.wy[i] = a.wz[i] = (T)0;
}

// First drift of a fused chunk: positions of the step into the other buffer.
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_drift_first(Args<T, P> a)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    P px = a.px[i], py = a.py[i], pz = a.pz[i];
    P ix = a.ix[i], iy = a.iy[i], iz = a.iz[i];
    V3<T> vp;
    drift(a.drift, a.fx[i], a.fy[i], a.fz[i], a.wx[i], a.wy[i], a.wz[i], a.lam[i], a.hlf2[i], px, py, pz, ix, iy,
          iz, vp);
    a.qx[i] = px; a.qy[i] = py; a.qz[i] = pz;
    if (a.jx != a.qx) { a.jx[i] = ix; a.jy[i] = iy; a.jz[i] = iz; }
}

// LDS slots for the bond terms of one block's beads (entries of the bead
// incidence lists); a block whose lists are longer evaluates in-thread.
constexpr int kSpCap = 640;
constexpr int kAnCap = 800;
constexpr int kPairCap = 32;

// One whole step of one bead: base force, springs, angles, contacts, closing
// kick; with NEXT the drift of the following step, with LAST the predicted
// velocity of this step for read-back.  Reads f of the previous step from
// f*, writes this step's force to g*.
//
// Each bond is evaluated once per block by its owner bead (spring: p1, angle:
// b2) into LDS; the other beads of the bond read the term there when the
// owner is in the same block and evaluate it themselves otherwise.  The terms
// are the same values either way and are added in incidence-list order, so
// the result equals the per-kernel path bit for bit.
template <typename T, typename P, bool SPR, bool ANG, bool NEXT, bool LAST>
__global__ void __launch_bounds__(kBlock) k_fused_step(Args<T, P> a)
{
    __shared__ SpTerm<T> sps[SPR ? kSpCap : 1];
    __shared__ AnTerm<T> ans[ANG ? kAnCap : 1];
    __shared__ PairC prs[kPairCap];
    const int b0 = xcd_block() * kBlock;
    const int i = b0 + threadIdx.x;
    const bool live = i < a.n;
    const int bend = min(b0 + kBlock, a.n);
    for (int k = threadIdx.x; k < a.n_pairs; k += kBlock) prs[k] = a.pairs[k];
    int se0 = 0, se1 = 0, ae0 = 0, ae1 = 0;
    bool sp_lds = false, an_lds = false;
    if (SPR) {
        se0 = a.sp_start[b0];
        se1 = a.sp_start[bend];
        sp_lds = se1 - se0 <= kSpCap;
    }
    if (ANG) {
        ae0 = a.an_start[b0];
        ae1 = a.an_start[bend];
        an_lds = ae1 - ae0 <= kAnCap;
    }
    if (live) {
        if (SPR && sp_lds)
            for (int e = a.sp_start[i], e1 = a.sp_start[i + 1]; e < e1; ++e) {
                const int ent = a.sp_ent[e];
                const int s = ent >> 1;
                if (!(ent & 1) && s < a.n_springs) sps[e - se0] = spring_owner(a, s);
            }
        if (ANG && an_lds)
            for (int e = a.an_start[i], e1 = a.an_start[i + 1]; e < e1; ++e) {
                const int ent = a.an_ent[e];
                const int q = ent >> 2;
                if ((ent & 3) == 1 && q < a.n_angles) ans[e - ae0] = angle_term<T, P, true>(a, q, true);
            }
    }
    __syncthreads();
    if (!live) return;
    const T fpx = a.fx[i], fpy = a.fy[i], fpz = a.fz[i];
    const T wx = a.wx[i], wy = a.wy[i], wz = a.wz[i];
    const T lam = a.lam[i];
    V3<T> vp;
    vp.x = tfma(fpx, lam, wx);
    vp.y = tfma(fpy, lam, wy);
    vp.z = tfma(fpz, lam, wz);
    P px = a.px[i], py = a.py[i], pz = a.pz[i];
    V3<T> f = base_force<T, P>(a.drift, vp, px, py, pz);
    if (SPR)
        for (int e = a.sp_start[i], e1 = a.sp_start[i + 1]; e < e1; ++e) {
            const int ent = a.sp_ent[e];
            const int s = ent >> 1;
            if (s >= a.n_springs) continue;
            const bool p2 = ent & 1;
            const int o = p2 ? a.sp_twin[e] : e;
            const SpTerm<T> t = (sp_lds && o >= se0 && o < se1) ? sps[o - se0]
                                                                  : (p2 ? spring_partner(a, s) : spring_owner(a, s));
            add_spring(f, t, p2);
        }
    if (ANG)
        for (int e = a.an_start[i], e1 = a.an_start[i + 1]; e < e1; ++e) {
            const int ent = a.an_ent[e];
            const int q = ent >> 2;
            const int role = ent & 3;
            if (q >= a.n_angles) continue;
            const int o = role == 1 ? e : a.an_owner[e];
            const AnTerm<T> t = (an_lds && o >= ae0 && o < ae1) ? ans[o - ae0]
                                                                  : angle_term<T, P, true>(a, q, role == 1);
            add_angle(f, t, role);
        }
    gather_contacts(a, i, f, prs);
    a.gx[i] = f.x; a.gy[i] = f.y; a.gz[i] = f.z;
    const T h = a.hlf[i];
    T sx = f.x, sy = f.y, sz = f.z;
    sx += fpx;
    sy += fpy;
    sz += fpz;
    const T nwx = tfma(sx, h, wx), nwy = tfma(sy, h, wy), nwz = tfma(sz, h, wz);
    a.wx[i] = nwx; a.wy[i] = nwy; a.wz[i] = nwz;
    if (LAST) { a.vx[i] = vp.x; a.vy[i] = vp.y; a.vz[i] = vp.z; }
    if (NEXT) {
        P ix = a.ix[i], iy = a.iy[i], iz = a.iz[i];
        V3<T> vn;
        drift(a.drift, f.x, f.y, f.z, nwx, nwy, nwz, lam, a.hlf2[i], px, py, pz, ix, iy, iz, vn);
        a.qx[i] = px; a.qy[i] = py; a.qz[i] = pz;
        if (a.jx != a.qx) { a.jx[i] = ix; a.jy[i] = iy; a.jz[i] = iz; }
    }
}

__global__ void k_add_int(int *p, int v) { *p += v; }

__global__ void __launch_bounds__(kBlock) k_fill_int(int *p, int v, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(kBlock) k_copy_int(const int *src, int *dst, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ---- grid rebuild ----------------------------------------------------------------

// FillGrid's cell of a bead (kernel.cl:492-504) in fp64; -1 outside the grid.
template <typename P>
__global__ void __launch_bounds__(kBlock) k_hash(const P *px, const P *py, const P *pz, int n, Box b, int *key,
                                                 int *count)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int cx = (int)floor((double)px[i] / b.sx);
    const int cy = (int)floor((double)py[i] / b.sy);
    const int cz = (int)floor((double)pz[i] / b.sz);
    int c = -1;
    if (cx >= 0 && cx < b.cx && cy >= 0 && cy < b.cy && cz >= 0 && cz < b.cz) {
        c = cx + cy * b.cx + cz * b.cxy;
        atomicAdd(&count[c], 1);
    }
    key[i] = c;
}

__global__ void __launch_bounds__(kBlock) k_scatter(const int *key, int n, int *cursor, int *sorted)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int c = key[i];
    if (c >= 0) sorted[atomicAdd(&cursor[c], 1)] = i;
}

// Wave-wide max and sum (all 64 lanes take part), so that statistics cost one
// atomic per wave instead of one per thread on a single address.
__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Bead data in cell order, 36 B per bead: fp64 position and tag, loc, type.
template <typename P>
__global__ void __launch_bounds__(kBlock) k_gather_sorted(const int *sorted, const int *n_hashed, const P *px,
                                                          const P *py, const P *pz, const int *type, const int *tag,
                                                          const int *loc, Sorted so)
{
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= *n_hashed) return;
    const int i = sorted[s];
    so.x[s] = (double)px[i];
    so.y[s] = (double)py[i];
    so.z[s] = (double)pz[i];
    so.tag[s] = tag[i];
    so.loc[s] = loc[i];
    so.type[s] = type[i];
    so.slot_of[i] = s;
}

// One candidate range of cell-sorted slots [lo, hi) for the bead in slot s:
// BuildContactList's predicate (kernel.cl:574-659) in fp64.
__device__ __forceinline__ void scan_range(int lo, int hi, int s, double xi, double yi, double zi, int ti, int gi,
                                           int li, double ri, double si, const Sorted &so, const int *sorted,
                                           const double *radius, const double *skin, const PairC *pairs,
                                           const Box &b, int *row, int K, int &k)
{
    for (int t = lo; t < hi; ++t) {
        if (t == s) continue;
        if (gi == so.tag[t] && abs(li - so.loc[t]) <= 1) continue;
        double rx = xi - so.x[t], ry = yi - so.y[t], rz = zi - so.z[t];
        double len = rx * rx + ry * ry + rz * rz;
        if (b.bc == 0 && len > b.hbs) {
            if (rx * rx > b.hbs) rx += (rx > 0.0) ? -b.lx : b.lx;
            if (ry * ry > b.hbs) ry += (ry > 0.0) ? -b.ly : b.ly;
            if (rz * rz > b.hbs) rz += (rz > 0.0) ? -b.lz : b.lz;
            len = rx * rx + ry * ry + rz * rz;
        }
        const int pid = pair_index(ti, so.type[t], b.types);
        const int pt = pairs[pid].type;
        bool take = false;
        if (pt == 1 || pt == 3) take = len < b.cutoff_sq;
        if (pt == 2) {
            const int j = sorted[t];
            const double rj = radius[j], sj = skin[j];
            const double min_skin = (si > sj) ? sj : si;
            const double max_r = (ri > rj) ? ri : rj;
            const double th = max_r + min_skin;
            take = len < th * th;
        }
        if (!take) continue;
        if (k < K) row[k] = sorted[t] | (pid << kPidShift);
        ++k;
    }
}

// Contact candidates of the bead in slot s: its own cell and the 26
// neighbours - the reference's own cell + 13-cell half shell
// (SimulationManager.cpp:2210-2301) seen from both cells of a pair.  The three
// cells of an x-row are one contiguous slot range unless the row wraps or is
// clipped.  Up to K entries go to the slot's scratch row; cnt[bead] = count.
__global__ void __launch_bounds__(kBlock) k_search(const int *sorted, const int *cell_start, int n_cells,
                                                   const int *key, Sorted so, const double *radius, const double *skin,
                                                   const PairC *pairs, Box b, int *cnt, int *scratch, int K,
                                                   int *stats)
{
    const int s = blockIdx.x * kBlock + threadIdx.x;
    int k = 0, occ = 0;
    if (s < cell_start[n_cells]) {
        const int i = sorted[s];
        const int c = key[i];
        if (cell_start[c] == s) occ = cell_start[c + 1] - s;   // the cell's first slot reports its occupancy
        const int cz = c / b.cxy;
        const int cy = (c - cz * b.cxy) / b.cx;
        const int cx = c - cz * b.cxy - cy * b.cx;
        const double xi = so.x[s], yi = so.y[s], zi = so.z[s];
        const int ti = so.type[s], gi = so.tag[s], li = so.loc[s];
        const double ri = radius[i], si = skin[i];
        int *row = scratch + (size_t)s * K;
        for (int dz = -1; dz <= 1; ++dz) {
            int z = cz + dz;
            if (b.bc == 0) z = z < 0 ? b.cz - 1 : (z >= b.cz ? 0 : z);
            else if (z < 0 || z >= b.cz) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                int y = cy + dy;
                if (b.bc == 0) y = y < 0 ? b.cy - 1 : (y >= b.cy ? 0 : y);
                else if (y < 0 || y >= b.cy) continue;
                const int rbase = y * b.cx + z * b.cxy;
                if (cx >= 1 && cx + 1 < b.cx) {
                    scan_range(cell_start[rbase + cx - 1], cell_start[rbase + cx + 2], s, xi, yi, zi, ti, gi, li, ri,
                               si, so, sorted, radius, skin, pairs, b, row, K, k);
                } else {
                    for (int dx = -1; dx <= 1; ++dx) {
                        int x = cx + dx;
                        if (b.bc == 0) x = x < 0 ? b.cx - 1 : (x >= b.cx ? 0 : x);
                        else if (x < 0 || x >= b.cx) continue;
                        scan_range(cell_start[rbase + x], cell_start[rbase + x + 1], s, xi, yi, zi, ti, gi, li, ri,
                                   si, so, sorted, radius, skin, pairs, b, row, K, k);
                    }
                }
            }
        }
        cnt[i] = k;
    }
    const int mx = wave_max(k);
    const int over = wave_sum(k > K ? 1 : 0);
    const int mocc = wave_max(occ);
    const int big = wave_sum(occ > 64 ? 1 : 0);
    if ((threadIdx.x & 63) == 0) {
        if (mx > 0) atomicMax(&stats[kStatMaxList], mx);
        if (over > 0) atomicAdd(&stats[kStatListOver], over);
        if (mocc > 0) atomicMax(&stats[kStatMaxCell], mocc);
        if (big > 0) atomicAdd(&stats[kStatCellsOver], big);
    }
}

// Scratch rows (by slot) -> the bead-indexed CSR list the force kernels read,
// each bead's entries in ascending order.  The order of beads inside a cell
// follows the atomic scatter and is arbitrary; ordering the entries makes the
// list, and so the contact-force summation order, deterministic.
__global__ void __launch_bounds__(kBlock) k_compact(const int *cnt, const int *nb_start, const int *slot_of,
                                                    const int *scratch, int K, int n, int *nb_list, int cap)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int m = min(cnt[i], K);
    if (m == 0) return;
    const int base = nb_start[i];
    if (base + m > cap) return;   // capacity exceeded: the host reports it
    const int *row = scratch + (size_t)slot_of[i] * K;
    int *out = nb_list + base;
    for (int k = 0; k < m; ++k) {
        const int v = row[k];
        int j = k - 1;
        while (j >= 0 && out[j] > v) {
            out[j + 1] = out[j];
            --j;
        }
        out[j + 1] = v;
    }
}

// ---- device-wide exclusive scan (n ints -> n+1 offsets) ----------------------------

constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ int block_excl_scan(int v, int *lds, int &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    int off = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        const int s = lds[w];
        if (w < wave) off += s;
        total += s;
    }
    __syncthreads();
    return off + x - v;
}

__global__ void __launch_bounds__(kBlock) k_scan_reduce(const int *in, int n, int *bsum)
{
    __shared__ int lds[kBlock / 64];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += (base + k < n) ? in[base + k] : 0;
    int total;
    block_excl_scan(s, lds, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Exclusive scan of the block sums by one block; bsum[nb] receives the total.
__global__ void __launch_bounds__(kBlock) k_scan_blocks(int *bsum, int nb)
{
    __shared__ int lds[kBlock / 64];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kBlock) {
        const int k = b0 + threadIdx.x;
        const int v = k < nb ? bsum[k] : 0;
        int total;
        const int ex = block_excl_scan(v, lds, total);
        if (k < nb) bsum[k] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(const int *in, int n, const int *bsum, int nb, int *out)
{
    __shared__ int lds[kBlock / 64];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int total;
    int run = block_excl_scan(s, lds, total) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = bsum[nb];
}

// ---- explicit instantiation through launchers ----------------------------------------

static inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

template <typename T, typename P>
void Launch<T, P>::advance_position(const Args<T, P> &a, int n_active, hipStream_t s)
{
    if (n_active > 0) k_advance_position<T, P><<<blocks(n_active), kBlock, 0, s>>>(a, n_active);
}
template <typename T, typename P>
void Launch<T, P>::resolve_springs(const Args<T, P> &a, hipStream_t s)
{
    k_resolve_springs<T, P><<<blocks(a.n), kBlock, 0, s>>>(a);
}
template <typename T, typename P>
void Launch<T, P>::resolve_angles(const Args<T, P> &a, hipStream_t s)
{
    k_resolve_angles<T, P><<<blocks(a.n), kBlock, 0, s>>>(a);
}
template <typename T, typename P>
void Launch<T, P>::resolve_contacts(const Args<T, P> &a, hipStream_t s)
{
    k_resolve_contacts<T, P><<<blocks(a.n), kBlock, 0, s>>>(a);
}
template <typename T, typename P>
void Launch<T, P>::advance_velocity(const Args<T, P> &a, int n_active, hipStream_t s)
{
    k_advance_velocity<T, P><<<blocks(n_active > 0 ? n_active : 1), kBlock, 0, s>>>(a, n_active);
}
template <typename T, typename P>
void Launch<T, P>::reset_energy(const Args<T, P> &a, int n_active, hipStream_t s)
{
    if (n_active > 0) k_reset_energy<T, P><<<blocks(n_active), kBlock, 0, s>>>(a, n_active);
}
template <typename T, typename P>
void Launch<T, P>::drift_first(const Args<T, P> &a, hipStream_t s)
{
    k_drift_first<T, P><<<blocks(a.n), kBlock, 0, s>>>(a);
}
template <typename T, typename P>
void Launch<T, P>::fused(const Args<T, P> &a, bool spr, bool ang, bool next, bool last, hipStream_t s)
{
    const int g = blocks(a.n);
#define FN_FUSED(S_, A_, N_, L_) k_fused_step<T, P, S_, A_, N_, L_><<<g, kBlock, 0, s>>>(a)
#define FN_FUSED_NL(S_, A_)                 \
    if (next && last) FN_FUSED(S_, A_, true, true);       \
    else if (next) FN_FUSED(S_, A_, true, false);         \
    else if (last) FN_FUSED(S_, A_, false, true);         \
    else FN_FUSED(S_, A_, false, false);
    if (spr && ang) { FN_FUSED_NL(true, true) }
    else if (spr) { FN_FUSED_NL(true, false) }
    else if (ang) { FN_FUSED_NL(false, true) }
    else { FN_FUSED_NL(false, false) }
#undef FN_FUSED_NL
#undef FN_FUSED
}
template <typename T, typename P>
void Launch<T, P>::hash(const P *px, const P *py, const P *pz, int n, const Box &b, int *key, int *count,
                        hipStream_t s)
{
    k_hash<P><<<blocks(n), kBlock, 0, s>>>(px, py, pz, n, b, key, count);
}
template <typename T, typename P>
void Launch<T, P>::gather_sorted(const int *sorted, const int *n_hashed, const P *px, const P *py, const P *pz,
                                 const int *type, const int *tag, const int *loc, const Sorted &so, int n,
                                 hipStream_t s)
{
    k_gather_sorted<P><<<blocks(n), kBlock, 0, s>>>(sorted, n_hashed, px, py, pz, type, tag, loc, so);
}

template struct Launch<double, double>;
template struct Launch<float, double>;

void fill_int(int *p, int v, int n, hipStream_t s)
{
    if (n > 0) k_fill_int<<<blocks(n), kBlock, 0, s>>>(p, v, n);
}
void copy_int(const int *src, int *dst, int n, hipStream_t s)
{
    if (n > 0) k_copy_int<<<blocks(n), kBlock, 0, s>>>(src, dst, n);
}
void add_int(int *p, int v, hipStream_t s) { k_add_int<<<1, 1, 0, s>>>(p, v); }
void search(const int *sorted, const int *cell_start, int n_cells, const int *key, const Sorted &so,
            const double *radius, const double *skin, const PairC *pairs, const Box &b, int *cnt, int *scratch, int K,
            int *stats, int n, hipStream_t s)
{
    k_search<<<blocks(n), kBlock, 0, s>>>(sorted, cell_start, n_cells, key, so, radius, skin, pairs, b, cnt, scratch,
                                          K, stats);
}
void compact(const int *cnt, const int *nb_start, const int *slot_of, const int *scratch, int K, int n, int *nb_list,
             int cap, hipStream_t s)
{
    k_compact<<<blocks(n), kBlock, 0, s>>>(cnt, nb_start, slot_of, scratch, K, n, nb_list, cap);
}
void scatter(const int *key, int n, int *cursor, int *sorted, hipStream_t s)
{
    k_scatter<<<blocks(n), kBlock, 0, s>>>(key, n, cursor, sorted);
}
int scan_blocks_needed(int n) { return (n + kScanTile - 1) / kScanTile; }
void exclusive_scan(const int *in, int n, int *out, int *tmp, hipStream_t s)
{
    const int nb = scan_blocks_needed(n > 0 ? n : 1);
    k_scan_reduce<<<nb, kBlock, 0, s>>>(in, n, tmp);
    k_scan_blocks<<<1, kBlock, 0, s>>>(tmp, nb);
    k_scan_apply<<<nb, kBlock, 0, s>>>(in, n, tmp, nb, out);
}

}  // namespace fnk
