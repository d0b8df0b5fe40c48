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
    gather_contacts(a, i, f);
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

// One whole step of one bead: base force, springs, angles, contacts, closing
// kick; with NEXT the drift of the following step, with LAST the predicted
// velocity of this step for read-back.  Reads f of the previous step from
// f*, writes this step's force to g*.
template <typename T, typename P, bool SPR, bool ANG, bool NEXT, bool LAST>
__global__ void __launch_bounds__(kBlock) k_fused_step(Args<T, P> a)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
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
    if (SPR) gather_springs(a, i, f);
    if (ANG) gather_angles<T, P, true>(a, i, f);
    gather_contacts(a, i, f);
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

// Orders each cell's beads by id (the atomic scatter order is arbitrary) and
// records occupancy statistics.
__global__ void __launch_bounds__(kBlock) k_sort_cells(const int *start, int *sorted, int n_cells, int *stats)
{
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= n_cells) return;
    const int s0 = start[c], s1 = start[c + 1];
    const int m = s1 - s0;
    if (m <= 0) return;
    for (int k = s0 + 1; k < s1; ++k) {
        const int v = sorted[k];
        int j = k - 1;
        while (j >= s0 && sorted[j] > v) {
            sorted[j + 1] = sorted[j];
            --j;
        }
        sorted[j + 1] = v;
    }
    atomicMax(&stats[kStatMaxCell], m);
    if (m > 64) atomicAdd(&stats[kStatCellsOver], 1);
}

// Candidate pairs of bead i: its own cell and the 26 neighbours, which is the
// reference's own cell + 13-cell half shell (SimulationManager.cpp:2210-2301)
// seen from both cells of a pair.  The predicate is BuildContactList's
// (kernel.cl:574-659) in fp64.  FILL=false counts, FILL=true writes.
template <typename P, bool FILL>
__global__ void __launch_bounds__(kBlock) k_pairs(const P *px, const P *py, const P *pz, const int *key,
                                                  const int *sorted, const int *cell_start, int n_cells,
                                                  const int *type, const int *tag, const int *loc,
                                                  const double *radius, const double *skin, const PairC *pairs,
                                                  Box b, int *cnt, const int *nb_start, int *nb_list, int cap,
                                                  int *stats)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= cell_start[n_cells]) return;
    const int i = sorted[t];
    const int c = key[i];
    const int cz = c / b.cxy;
    const int cy = (c - cz * b.cxy) / b.cx;
    const int cx = c - cz * b.cxy - cy * b.cx;
    const double xi = (double)px[i], yi = (double)py[i], zi = (double)pz[i];
    const int ti = type[i], gi = tag[i], li = loc[i];
    const double ri = radius[i], si = skin[i];
    int k = 0;
    int base = 0, end = 0;
    if (FILL) {
        base = nb_start[i];
        end = nb_start[i + 1];
        if (end > cap) return;   // capacity exceeded: the host reports it
    }
    for (int dz = -1; dz <= 1; ++dz) {
        int z = cz + dz;
        if (b.bc == 0) z = z < 0 ? b.cz - 1 : (z >= b.cz ? 0 : z);
        else if (z < 0 || z >= b.cz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            int y = cy + dy;
            if (b.bc == 0) y = y < 0 ? b.cy - 1 : (y >= b.cy ? 0 : y);
            else if (y < 0 || y >= b.cy) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                int x = cx + dx;
                if (b.bc == 0) x = x < 0 ? b.cx - 1 : (x >= b.cx ? 0 : x);
                else if (x < 0 || x >= b.cx) continue;
                const int nc = x + y * b.cx + z * b.cxy;
                const int s1 = cell_start[nc + 1];
                for (int s = cell_start[nc]; s < s1; ++s) {
                    const int j = sorted[s];
                    if (j == i) continue;
                    const int gj = tag[j];
                    if (gi == gj && abs(li - loc[j]) <= 1) continue;
                    double rx = xi - (double)px[j], ry = yi - (double)py[j], rz = zi - (double)pz[j];
                    double len = rx * rx + ry * ry + rz * rz;
                    if (b.bc == 0 && len > b.hbs) {
                        if (rx * rx > b.hbs) rx += (rx > 0.0) ? -b.lx : b.lx;
                        if (ry * ry > b.hbs) ry += (ry > 0.0) ? -b.ly : b.ly;
                        if (rz * rz > b.hbs) rz += (rz > 0.0) ? -b.lz : b.lz;
                        len = rx * rx + ry * ry + rz * rz;
                    }
                    const int pid = pair_index(ti, type[j], b.types);
                    const int pt = pairs[pid].type;
                    bool take = false;
                    if (pt == 1 || pt == 3) take = len < b.cutoff_sq;
                    if (pt == 2) {
                        const double rj = radius[j], sj = skin[j];
                        const double min_skin = (si > sj) ? sj : si;
                        const double max_r = (ri > rj) ? ri : rj;
                        const double th = max_r + min_skin;
                        take = len < th * th;
                    }
                    if (!take) continue;
                    if (FILL) nb_list[base + k] = j | (pid << kPidShift);
                    ++k;
                }
            }
        }
    }
    if (!FILL) {
        cnt[i] = k;
        atomicMax(&stats[kStatMaxList], k);
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
void Launch<T, P>::pairs(bool fill, const P *px, const P *py, const P *pz, const int *key, const int *sorted,
                         const int *cell_start, int n_cells, int n, const int *type, const int *tag,
                         const int *loc, const double *radius, const double *skin, const PairC *pairs_,
                         const Box &b, int *cnt, const int *nb_start, int *nb_list, int cap, int *stats,
                         hipStream_t s)
{
    if (fill)
        k_pairs<P, true><<<blocks(n), kBlock, 0, s>>>(px, py, pz, key, sorted, cell_start, n_cells, type, tag, loc,
                                                      radius, skin, pairs_, b, cnt, nb_start, nb_list, cap, stats);
    else
        k_pairs<P, false><<<blocks(n), kBlock, 0, s>>>(px, py, pz, key, sorted, cell_start, n_cells, type, tag,
                                                       loc, radius, skin, pairs_, b, cnt, nb_start, nb_list, cap,
                                                       stats);
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
void scatter(const int *key, int n, int *cursor, int *sorted, hipStream_t s)
{
    k_scatter<<<blocks(n), kBlock, 0, s>>>(key, n, cursor, sorted);
}
void sort_cells(const int *start, int *sorted, int n_cells, int *stats, hipStream_t s)
{
    k_sort_cells<<<blocks(n_cells), kBlock, 0, s>>>(start, sorted, n_cells, stats);
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
