// engine_kernels.hip - __global__ kernels of the FibreNet step (gfx950).
//
// Granular kernels reproduce one reference kernel each (ExecuteKernel by
// name).  The fast path runs a step as three launches: spring terms (one
// thread per spring), angle terms (one thread per angle) and the bead kernel,
// which sums the terms, adds contacts, closes the kick and drifts the next
// step, so that the bead state makes one pass over HBM per step.  Each launch
// is a short chain of independent loads: the bond kernels read their beads
// through the bond arrays, the bead kernel its incidence lists.
#include <utility>

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

// Spring terms of one step (kernel.cl:371-417), one thread per spring: the
// force on p1 (zero for an inactive spring) and, for an HR_FL spring whose
// stress exceeds its threshold, the status change and the break step.
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_spring_terms(Args<T, P> a)
{
    const int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= a.n_springs) return;
    V3<T> F;
    bool broke;
    const bool on = spring_eval(a, s, F, broke);
    if (broke) {
        a.sp_status[s] = 0;
        a.sp_break[s] = *a.step + a.step_off;
    }
    a.sfx[s] = on ? F.x : (T)0;
    a.sfy[s] = on ? F.y : (T)0;
    a.sfz[s] = on ? F.z : (T)0;
}

// Angle terms of one step (kernel.cl:274-348), one thread per angle, after
// the spring terms: an HR_FL angle dies with either of its springs; a live
// angle writes its forces on b1 and b3 (b2 receives their negated sum).
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_angle_terms(Args<T, P> a)
{
    const int q = xcd_block() * kBlock + threadIdx.x;
    if (q >= a.n_angles) return;
    int st = a.an_status[q];
    if (st == 1 && a.an_type[q] == 1 && (a.sp_status[a.an_s1[q]] == 0 || a.sp_status[a.an_s2[q]] == 0)) {
        st = 0;
        a.an_status[q] = 0;
    }
    V3<T> pa{(T)0, (T)0, (T)0}, pc{(T)0, (T)0, (T)0};
    if (st == 1) angle_eval(a, q, pa, pc);
    a.apx[q] = pa.x; a.apy[q] = pa.y; a.apz[q] = pa.z;
    a.acx[q] = pc.x; a.acy[q] = pc.y; a.acz[q] = pc.z;
}

// ResolveSprings' accumulation seen from each bead: adds the spring terms.
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_gather_springs(Args<T, P> a)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    if (a.sp_start[i] == a.sp_start[i + 1]) return;
    V3<T> f{a.fx[i], a.fy[i], a.fz[i]};
    gather_spring_terms(a, i, f);
    a.fx[i] = f.x; a.fy[i] = f.y; a.fz[i] = f.z;
}

// ResolveAngles' accumulation seen from each bead: adds the angle terms.
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_gather_angles(Args<T, P> a)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    if (a.an_start[i] == a.an_start[i + 1]) return;
    V3<T> f{a.fx[i], a.fy[i], a.fz[i]};
    gather_angle_terms(a, i, f);
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

// Pair-table entries kept in LDS by the bead kernel (the table of T types has
// T(T+1)/2 entries; contact-list entries address at most 32).
constexpr int kPairCap = 32;

// One whole step of one bead after the bond terms: base force, the spring and
// angle terms in incidence-list order, contacts, closing kick; with NEXT the
// drift of the following step, with LAST the predicted velocity of this step
// for read-back.  Reads f of the previous step from f*, writes this step's
// force to g*.  The terms and their summation order are those of the
// per-kernel path, so both paths agree bit for bit.
template <typename T, typename P, bool SPR, bool ANG, bool NEXT, bool LAST>
__global__ void __launch_bounds__(kBlock) k_bead_step(Args<T, P> a)
{
    __shared__ PairC prs[kPairCap];
    const int i = xcd_block() * kBlock + threadIdx.x;
    for (int k = threadIdx.x; k < a.n_pairs; k += kBlock) prs[k] = a.pairs[k];
    __syncthreads();
    if (i >= a.n) return;
    const T fpx = a.fx[i], fpy = a.fy[i], fpz = a.fz[i];
    const T wx = a.wx[i], wy = a.wy[i], wz = a.wz[i];
    const T lam = a.lam[i];
    V3<T> vp;
    vp.x = tfma(fpx, lam, wx);
    vp.y = tfma(fpy, lam, wy);
    vp.z = tfma(fpz, lam, wz);
    P px = a.px[i], py = a.py[i], pz = a.pz[i];
    V3<T> f = base_force<T, P>(a.drift, vp, px, py, pz);
    if (SPR) gather_spring_terms(a, i, f);
    if (ANG) gather_angle_terms(a, i, f);
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

// Bead data in cell order, 36 B per bead: fp64 position, tag, loc, type.
template <typename P>
__global__ void __launch_bounds__(kBlock) k_gather_sorted(const int *sorted, const int *n_hashed, const P *px,
                                                          const P *py, const P *pz, const int *type, const int *tag,
                                                          const int *loc, const int *key, Sorted so)
{
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= *n_hashed) return;
    const int i = sorted[s];
    SortedBead b;
    b.x = (double)px[i];
    b.y = (double)py[i];
    b.z = (double)pz[i];
    b.tag = tag[i];
    b.loc = loc[i];
    so.bead[s] = b;
    so.type[s] = type[i];
    so.cell[s] = key[i];
    so.slot_of[i] = s;
}

// Scan of one candidate range of cell-sorted slots [lo, hi) for the bead in
// slot s: BuildContactList's predicate (kernel.cl:574-659) in fp64.  The
// distance is tested first against the largest listing radius of any pair
// type (reach2), so that most candidates cost two loads.
__device__ __forceinline__ void scan_range(int lo, int hi, int s, const SortedBead &me, int ti, double ri, double si,
                                           const Sorted &so, const int *sorted, const double *radius,
                                           const double *skin, const PairC *pairs, const Box &b, double reach2,
                                           int *row, int K, int &k)
{
    for (int t = lo; t < hi; ++t) {
        const SortedBead o = so.bead[t];
        double rx = me.x - o.x, ry = me.y - o.y, rz = me.z - o.z;
        double len = rx * rx + ry * ry + rz * rz;
        if (b.bc == 0 && len > b.hbs) {
            if (rx * rx > b.hbs) rx += (rx > 0.0) ? -b.lx : b.lx;
            if (ry * ry > b.hbs) ry += (ry > 0.0) ? -b.ly : b.ly;
            if (rz * rz > b.hbs) rz += (rz > 0.0) ? -b.lz : b.lz;
            len = rx * rx + ry * ry + rz * rz;
        }
        if (!(len < reach2) || t == s) continue;
        if (me.tag == o.tag && abs(me.loc - o.loc) <= 1) continue;
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
__global__ void __launch_bounds__(kBlock) k_search(const int *sorted, const int *cell_start, int n_cells, Sorted so,
                                                   const double *radius, const double *skin, const PairC *pairs,
                                                   Box b, double reach2, int *cnt, int *scratch, int K, int *stats)
{
    const int s = blockIdx.x * kBlock + threadIdx.x;
    int k = 0, occ = 0;
    if (s < cell_start[n_cells]) {
        const int i = sorted[s];
        const int c = so.cell[s];
        if (cell_start[c] == s) occ = cell_start[c + 1] - s;   // the cell's first slot reports its occupancy
        const int cz = c / b.cxy;
        const int cy = (c - cz * b.cxy) / b.cx;
        const int cx = c - cz * b.cxy - cy * b.cx;
        const SortedBead me = so.bead[s];
        const int ti = so.type[s];
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
                    scan_range(cell_start[rbase + cx - 1], cell_start[rbase + cx + 2], s, me, ti, ri, si, so, sorted,
                               radius, skin, pairs, b, reach2, row, K, k);
                } else {
                    for (int dx = -1; dx <= 1; ++dx) {
                        int x = cx + dx;
                        if (b.bc == 0) x = x < 0 ? b.cx - 1 : (x >= b.cx ? 0 : x);
                        else if (x < 0 || x >= b.cx) continue;
                        scan_range(cell_start[rbase + x], cell_start[rbase + x + 1], s, me, ti, ri, si, so, sorted,
                                   radius, skin, pairs, b, reach2, row, K, k);
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
    if (m <= 8) {
        int v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = k < m ? row[k] : 0x7fffffff;
#pragma unroll
        for (int p = 0; p < 8; ++p)          // odd-even transposition sort of 8 registers
#pragma unroll
            for (int k = p & 1; k + 1 < 8; k += 2)
                if (v[k] > v[k + 1]) {
                    const int t = v[k];
                    v[k] = v[k + 1];
                    v[k + 1] = t;
                }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < m) out[k] = v[k];
        return;
    }
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
void Launch<T, P>::spring_terms(const Args<T, P> &a, hipStream_t s)
{
    if (a.n_springs > 0) k_spring_terms<T, P><<<blocks(a.n_springs), kBlock, 0, s>>>(a);
}
template <typename T, typename P>
void Launch<T, P>::angle_terms(const Args<T, P> &a, hipStream_t s)
{
    if (a.n_angles > 0) k_angle_terms<T, P><<<blocks(a.n_angles), kBlock, 0, s>>>(a);
}
template <typename T, typename P>
void Launch<T, P>::resolve_springs(const Args<T, P> &a, hipStream_t s)
{
    spring_terms(a, s);
    k_gather_springs<T, P><<<blocks(a.n), kBlock, 0, s>>>(a);
}
template <typename T, typename P>
void Launch<T, P>::resolve_angles(const Args<T, P> &a, hipStream_t s)
{
    angle_terms(a, s);
    k_gather_angles<T, P><<<blocks(a.n), kBlock, 0, s>>>(a);
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
template <typename T, typename P, int M>
static void launch_fused_mask(const Args<T, P> &a, hipStream_t s)
{
    k_bead_step<T, P, (M & 1) != 0, (M & 2) != 0, (M & 4) != 0, (M & 8) != 0><<<blocks(a.n), kBlock, 0, s>>>(a);
}

template <typename T, typename P, int... M>
static void dispatch_fused(const Args<T, P> &a, int mask, hipStream_t s, std::integer_sequence<int, M...>)
{
    using Fn = void (*)(const Args<T, P> &, hipStream_t);
    static constexpr Fn table[] = {&launch_fused_mask<T, P, M>...};
    table[mask](a, s);
}

template <typename T, typename P>
void Launch<T, P>::bead_step(const Args<T, P> &a, bool spr, bool ang, bool next, bool last, hipStream_t s)
{
    const int mask = (spr ? 1 : 0) | (ang ? 2 : 0) | (next ? 4 : 0) | (last ? 8 : 0);
    dispatch_fused<T, P>(a, mask, s, std::make_integer_sequence<int, 16>{});
}
template <typename T, typename P>
void Launch<T, P>::hash(const P *px, const P *py, const P *pz, int n, const Box &b, int *key, int *count,
                        hipStream_t s)
{
    k_hash<P><<<blocks(n), kBlock, 0, s>>>(px, py, pz, n, b, key, count);
}
template <typename T, typename P>
void Launch<T, P>::gather_sorted(const int *sorted, const int *n_hashed, const P *px, const P *py, const P *pz,
                                 const int *type, const int *tag, const int *loc, const int *key, const Sorted &so,
                                 int n, hipStream_t s)
{
    k_gather_sorted<P><<<blocks(n), kBlock, 0, s>>>(sorted, n_hashed, px, py, pz, type, tag, loc, key, so);
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
void search(const int *sorted, const int *cell_start, int n_cells, const Sorted &so, const double *radius,
            const double *skin, const PairC *pairs, const Box &b, double reach2, int *cnt, int *scratch, int K,
            int *stats, int n, hipStream_t s)
{
    k_search<<<blocks(n), kBlock, 0, s>>>(sorted, cell_start, n_cells, so, radius, skin, pairs, b, reach2, cnt,
                                          scratch, K, stats);
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
