// engine_kernels.hip - __global__ kernels of the FibreNet step (gfx950).
//
// Granular kernels reproduce one reference kernel each (ExecuteKernel by
// name).  The fast path runs a step as three launches: spring terms (one
// thread per spring), angle terms (one thread per angle) and the bead kernel,
// which sums the terms, adds contacts, closes the kick and drifts the next
// step, so that the bead state makes one pass over HBM per step.  Each launch
// is a short chain of independent loads: the bond kernels read their beads
// through the bond arrays, the bead kernel its incidence lists.
#include <algorithm>
#include <cstdint>
#include <stdexcept>
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

// Spring terms of one step (kernel.cl:371-417), one thread per spring of the
// `total` uploaded: +F at the p1 slot, -F at the p2 slot; for an HR_FL
// spring whose stress exceeds its threshold, the status change and the break
// step.  Springs that do not act this step (inactive, or beyond the
// launched range a.n_springs) leave zero terms.
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_spring_terms(Args<T, P> a, int total)
{
    const int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= total) return;
    V3<T> F{(T)0, (T)0, (T)0};
    if (s < a.n_springs) {
        bool broke;
        if (!spring_eval(a, s, F, broke)) F = V3<T>{(T)0, (T)0, (T)0};
        if (broke) {
            a.sp_status[s] = 0;
            a.sp_break[s] = *a.step + a.step_off;
        }
    }
    a.sp_term[a.sp_slot[2 * s]] = F;
    a.sp_term[a.sp_slot[2 * s + 1]] = V3<T>{F.x * (T)-1, F.y * (T)-1, F.z * (T)-1};
}

// Angle terms of one step (kernel.cl:274-360), one thread per angle of the
// `total` uploaded, after the spring terms: an HR_FL angle dies with either
// of its springs; a live angle puts pa at the b1 slot, pc at the b3 slot and
// -pa-pc at the b2 slot.
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_angle_terms(Args<T, P> a, int total)
{
    const int q = xcd_block() * kBlock + threadIdx.x;
    if (q >= total) return;
    V3<T> pa{(T)0, (T)0, (T)0}, pc{(T)0, (T)0, (T)0};
    if (q < a.n_angles) {
        int st = a.an_status[q];
        if (st == 1 && a.an_type[q] == 1 && (a.sp_status[a.an_s1[q]] == 0 || a.sp_status[a.an_s2[q]] == 0)) {
            st = 0;
            a.an_status[q] = 0;
        }
        if (st == 1) angle_eval(a, q, pa, pc);
    }
    a.an_term[a.an_slot[3 * q]] = pa;
    a.an_term[a.an_slot[3 * q + 1]] = pc;
    a.an_term[a.an_slot[3 * q + 2]] = V3<T>{-pa.x - pc.x, -pa.y - pc.y, -pa.z - pc.z};
}

// ResolveSprings' / ResolveAngles' accumulation seen from each bead: adds
// the terms at the bead's slots of `start` to f.
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_gather_terms(Args<T, P> a, const int *start, const V3<T> *term)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const int e0 = start[i], e1 = start[i + 1];
    if (e0 == e1) return;
    V3<T> f{a.fx[i], a.fy[i], a.fz[i]};
    add_terms(term, e0, e1, f);
    a.fx[i] = f.x; a.fy[i] = f.y; a.fz[i] = f.z;
}

template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_resolve_contacts(Args<T, P> a)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    if (a.nb_cnt[i] == 0) return;
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

// The Box of bead i's network (ENS) or the single network's.
template <bool ENS, typename T, typename P>
__device__ __forceinline__ const Box &drift_box(const Args<T, P> &a, int i)
{
    if constexpr (ENS) return a.boxes[a.net[i]];
    else return a.drift;
}

// First drift of a fused chunk: positions of the step into the other buffer.
template <typename T, typename P, bool ENS>
__global__ void __launch_bounds__(kBlock) k_drift_first(Args<T, P> a)
{
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    P px = a.px[i], py = a.py[i], pz = a.pz[i];
    P ix = a.ix[i], iy = a.iy[i], iz = a.iz[i];
    V3<T> vp;
    drift(drift_box<ENS>(a, i), a.fx[i], a.fy[i], a.fz[i], a.wx[i], a.wy[i], a.wz[i], a.lam[i], a.hlf2[i], px, py,
          pz, ix, iy, iz, vp);
    a.qx[i] = px; a.qy[i] = py; a.qz[i] = pz;
    if (a.jx != a.qx) { a.jx[i] = ix; a.jy[i] = iy; a.jz[i] = iz; }
}

// Pair-table entries kept in LDS by the bead kernel (the table of T types has
// T(T+1)/2 entries; contact-list entries address at most 32).
constexpr int kPairCap = 32;

// Bead state read at the start of the bead kernel.
template <typename T, typename P>
struct BeadState {
    T fpx, fpy, fpz, wx, wy, wz, lam;
    V3<T> vp;
    P px, py, pz;
    int ms;                 // the bead's mass set (with a mass-set table), else -1
};

// Loads bead i and returns its base force: the predicted velocity of this
// step (kernel.cl:200-205) gives damping, plus gravity and walls.  With ms
// (the mass-set table in LDS), lambda*dt/m comes from the bead's set.
template <typename T, typename P>
__device__ __forceinline__ V3<T> bead_begin(const Args<T, P> &a, const Box &bx, int i, BeadState<T, P> &b,
                                            const MassSet<T> *ms = nullptr)
{
    b.fpx = a.fx[i]; b.fpy = a.fy[i]; b.fpz = a.fz[i];
    b.wx = a.wx[i]; b.wy = a.wy[i]; b.wz = a.wz[i];
    if (ms) {
        b.ms = a.ms_code[i];
        b.lam = ms[b.ms].lam;
    } else {
        b.ms = -1;
        b.lam = a.lam[i];
    }
    b.vp.x = tfma(b.fpx, b.lam, b.wx);
    b.vp.y = tfma(b.fpy, b.lam, b.wy);
    b.vp.z = tfma(b.fpz, b.lam, b.wz);
    b.px = a.px[i]; b.py = a.py[i]; b.pz = a.pz[i];
    return base_force<T, P>(bx, b.vp, b.px, b.py, b.pz);
}

// Closing kick of the step (kernel.cl:444-464) and, with NEXT, the drift of
// the next step from image (ix, iy, iz); with LAST the predicted velocity of
// this step for read-back.
template <typename T, typename P, bool NEXT, bool LAST>
__device__ __forceinline__ void bead_end(const Args<T, P> &a, const Box &bx, int i, const BeadState<T, P> &b,
                                         const V3<T> &f, P ix, P iy, P iz, const MassSet<T> *ms = nullptr)
{
    a.gx[i] = f.x; a.gy[i] = f.y; a.gz[i] = f.z;
    const T h = ms ? ms[b.ms].hlf : a.hlf[i];
    T sx = f.x, sy = f.y, sz = f.z;
    sx += b.fpx;
    sy += b.fpy;
    sz += b.fpz;
    const T nwx = tfma(sx, h, b.wx), nwy = tfma(sy, h, b.wy), nwz = tfma(sz, h, b.wz);
    a.wx[i] = nwx; a.wy[i] = nwy; a.wz[i] = nwz;
    if (LAST) { a.vx[i] = b.vp.x; a.vy[i] = b.vp.y; a.vz[i] = b.vp.z; }
    if (NEXT) {
        P px = b.px, py = b.py, pz = b.pz;
        V3<T> vn;
        drift(bx, f.x, f.y, f.z, nwx, nwy, nwz, ms ? ms[b.ms].lam : b.lam, ms ? ms[b.ms].hlf2 : a.hlf2[i], px, py,
              pz, ix, iy, iz, vn);
        a.qx[i] = px; a.qy[i] = py; a.qz[i] = pz;
        if (a.jx != a.qx) { a.jx[i] = ix; a.jy[i] = iy; a.jz[i] = iz; }
    }
}

// One whole step of one bead after the bond kernels: base force, the spring
// and angle terms at its incidence slots, contacts, closing kick and next
// drift.  Reads f of the previous step from f*, writes this step's force to
// g*.  The terms and their summation order are those of the per-kernel
// path, so both paths agree bit for bit.
template <typename T, typename P, bool SPR, bool ANG, bool NEXT, bool LAST, bool ENS, int BLK>
__global__ void __launch_bounds__(BLK) k_bead_step(Args<T, P> a)
{
    __shared__ PairTab<T> prs[kPairCap];
    __shared__ WaveContactLds<T> cl[BLK / 64];
    const int i = xcd_block() * BLK + threadIdx.x;
    for (int k = threadIdx.x; k < min(a.n_pairs, kPairCap); k += BLK) prs[k] = to_tab<T>(a.pairs[k]);
    __syncthreads();
    const bool lj_uni = !ENS && lj_uniform_wave(prs, min(min(a.n_pairs, kPairCap), a.force.types * (a.force.types + 1) / 2));
    const bool live = i < a.n;
    const int ic = live ? i : a.n - 1;      // lanes past the end follow along for the wave-wide contacts
    const int e0 = a.nb_start[ic], cnt = live ? a.nb_cnt[ic] : 0;
    WaveContactLds<T> &w = cl[threadIdx.x >> 6];
    const ContactAhead ca = contacts_ahead(a, e0, cnt, w);
    const T r1 = a.radius[ic];
    const Box &bx = drift_box<ENS>(a, ic);
    BeadState<T, P> b;
    V3<T> f = bead_begin(a, bx, ic, b);
    if (SPR) add_terms(a.sp_term, a.sp_start[ic], a.sp_start[ic + 1], f);
    if (ANG) add_terms(a.an_term, a.an_start[ic], a.an_start[ic + 1], f);
    contacts_finish<T, P, ENS>(a, ca, e0, cnt, r1, b.px, b.py, b.pz, f, prs, w, ENS ? a.net[ic] : 0, lj_uni);
    if (!live) return;
    if (NEXT) bead_end<T, P, NEXT, LAST>(a, bx, i, b, f, a.ix[i], a.iy[i], a.iz[i]);
    else bead_end<T, P, NEXT, LAST>(a, bx, i, b, f, (P)0, (P)0, (P)0);
}

// Chain layout: a wave covers 64 consecutive beads and owns the middle 60;
// two halo lanes on each side supply the images, spring terms and angle
// terms that the owned beads need from their chain neighbours.

// The bead kernel for networks whose bonds all follow the chain layout
// (spring (i, i+1), angle (i-1, i, i+1), numbered along the chain).  Each
// lane evaluates the spring to its right and the angle centred on it once;
// the neighbours' images and the terms they need travel between adjacent
// lanes by wave shuffles instead of through memory.  The terms, their zero
// values for inactive bonds and their summation order are those of the
// per-kernel path, so both paths agree bit for bit.  Only owned lanes write
// bond statuses and bead state.
// Waves per SIMD the chain kernel is compiled for: five for the fp32 build
// (95 registers, no spills), the compiler's own choice otherwise (the fp64
// and ensemble forms would spill).
template <typename T, bool ENS>
constexpr int kChainWaves = (sizeof(T) == sizeof(float) && !ENS) ? 5 : 1;

template <typename T, typename P, bool SPR, bool ANG, bool NEXT, bool LAST, bool ENS, int BLK>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(kChainWaves<T, ENS>, 8))) k_bead_step_chain(Args<T, P> a)
{
    __shared__ PairTab<T> prs[kPairCap];
    __shared__ WaveContactLds<T> cl[BLK / 64];
    // the mass sets in the launch's dynamic LDS, a.n_ms of them (sized at
    // launch, so that a network with few sets keeps its LDS small)
    extern __shared__ __align__(16) unsigned char chain_dyn_lds[];
    MassSet<T> *mss = reinterpret_cast<MassSet<T> *>(chain_dyn_lds);
    for (int k = threadIdx.x; k < min(a.n_pairs, kPairCap); k += BLK) prs[k] = to_tab<T>(a.pairs[k]);
    const MassSet<T> *ms = nullptr;
    if (a.ms_code) {                     // block-uniform
        for (int k = threadIdx.x; k < a.n_ms; k += BLK) mss[k] = a.ms_sets[k];
        ms = mss;
    }
    __syncthreads();
    const bool lj_uni = !ENS && lj_uniform_wave(prs, min(min(a.n_pairs, kPairCap), a.force.types * (a.force.types + 1) / 2));
    const int lane = threadIdx.x & 63;
    // a wave's beads depend on its index over the whole grid only, so every
    // block size covers each bead with the same wave and the same lanes
    const int i = (xcd_block() * (BLK / 64) + (int)(threadIdx.x >> 6)) * kChainOwned + lane - kHalo;
    const bool in = i >= 0 && i < a.n;
    const bool own = in && lane >= kHalo && lane < 64 - kHalo;
    const int ic = in ? i : (i < 0 ? 0 : a.n - 1);
    // every load that depends on nothing else first: the bead's image, state,
    // list offsets, and the indices of its chain bonds; where one array holds
    // both positions and images (a finite box), the image is the position
    const bool alias = a.ix == a.px;     // grid-uniform
    P x = (P)0, y = (P)0, z = (P)0;
    if (in && !alias) {
        x = a.ix[i];
        y = a.iy[i];
        z = a.iz[i];
    }
    const Box &bx = drift_box<ENS>(a, ic);
    BeadState<T, P> b;
    V3<T> f = bead_begin(a, bx, ic, b, ms);
    if (in && alias) {
        x = b.px;
        y = b.py;
        z = b.pz;
    }
    const int e0 = a.nb_start[ic], cnt = own ? a.nb_cnt[ic] : 0;
    const T r1 = ms ? ms[b.ms].r : a.radius[ic];
    // the chain bonds of bead i: spring (i, i+1) and the angle centred on i,
    // their parameters by bead (parameter sets) or through the bond indices
    bool hs = false, hq = false;
    int sr = -1, q = -1;
    SpringP<T> spp{};
    AngleP<T> anp{};
    unsigned code = kNoSet | (kNoSet << 16);
    if (a.bond_code) {
        if (in && (SPR || ANG)) {
            code = a.bond_code[i];
            const unsigned sc = code & kNoSet;
            if (sc != kNoSet) {
                const SpringSet ss = a.sp_sets[sc];
                hs = true;
                spp = SpringP<T>{ss.ty, (code & kSpringOn) ? 1 : 0, ss.l0, ss.c1, ss.c2, (T)ss.k};
            }
        }
        if (in && ANG) {
            const unsigned ac = (code >> 16) & kNoSet;
            if (ac != kNoSet) {
                const AngleSet as = a.an_sets[ac];
                hq = true;
                anp = AngleP<T>{(code & kAngleOn) ? 1 : 0, as.ty, (T)as.th0, (T)as.k};
            }
        }
    } else {
        sr = (in && (SPR || ANG)) ? a.ch_spring[i] : -1;
        q = (in && ANG) ? a.ch_angle[i] : -1;
        hs = sr >= 0;
        hq = q >= 0;
        if (hs) spp = load_spring(a, sr);
        if (hq) anp = load_angle(a, q);
    }
    // and the first list entries of the wave's contacts
    WaveContactLds<T> &w = cl[threadIdx.x >> 6];
    const ContactAhead ca = contacts_ahead(a, e0, cnt, w);

    // images of beads i + 1 and i - 1 (meaningless in the outermost lanes,
    // whose results no owned lane uses)
    const P xn = lane_next(x), yn = lane_next(y), zn = lane_next(z);
    const P xp = lane_prev(x), yp = lane_prev(y), zp = lane_prev(z);

    // spring (i, i+1): its term and its status after this step's spring pass
    V3<T> fr{(T)0, (T)0, (T)0};
    int str = 1;
    if (hs) {
        str = spp.st;
        if (SPR) {
            bool broke;
            if (!spring_force(spp, x - xn, y - yn, z - zn, fr, broke)) fr = V3<T>{(T)0, (T)0, (T)0};
            if (broke) {
                str = 0;
                if (own) {
                    const int s = a.bond_code ? a.ch_spring[i] : sr;
                    a.sp_status[s] = 0;
                    a.sp_break[s] = *a.step + a.step_off;
                    if (a.bond_code) a.bond_code[i] = (code &= ~kSpringOn);
                }
            }
        }
    }
    const int sl = lane_prev(hs ? 1 : 0);
    const int stl = lane_prev(str);
    const V3<T> fl{lane_prev(fr.x), lane_prev(fr.y), lane_prev(fr.z)};

    // angle (i-1, i, i+1) with this step's HR_FL rule (its springs are the
    // two chain springs of bead i)
    V3<T> pa{(T)0, (T)0, (T)0}, pc{(T)0, (T)0, (T)0};
    if (hq) {
        int st = anp.st;
        if (st == 1 && anp.ty == 1 && (stl == 0 || str == 0)) {
            st = 0;
            if (own) {
                a.an_status[a.bond_code ? a.ch_angle[i] : q] = 0;
                if (a.bond_code) a.bond_code[i] = (code &= ~kAngleOn);
            }
        }
        if (st == 1) angle_force(anp, xp, yp, zp, x, y, z, xn, yn, zn, pa, pc);
    }
    const int hql = lane_prev(hq ? 1 : 0), hqn = lane_next(hq ? 1 : 0);
    const V3<T> pcl{lane_prev(pc.x), lane_prev(pc.y), lane_prev(pc.z)};
    const V3<T> pan{lane_next(pa.x), lane_next(pa.y), lane_next(pa.z)};

    if (SPR) {
        if (sl) {               // bead i is p2 of spring (i-1, i)
            f.x += fl.x * (T)-1;
            f.y += fl.y * (T)-1;
            f.z += fl.z * (T)-1;
        }
        if (hs) {               // p1 of spring (i, i+1)
            f.x += fr.x;
            f.y += fr.y;
            f.z += fr.z;
        }
    }
    if (ANG) {
        if (hql) {              // b3 of angle (i-2, i-1, i)
            f.x += pcl.x;
            f.y += pcl.y;
            f.z += pcl.z;
        }
        if (hq) {               // b2 of angle (i-1, i, i+1)
            f.x += -pa.x - pc.x;
            f.y += -pa.y - pc.y;
            f.z += -pa.z - pc.z;
        }
        if (hqn) {              // b1 of angle (i, i+1, i+2)
            f.x += pan.x;
            f.y += pan.y;
            f.z += pan.z;
        }
    }
    contacts_finish<T, P, ENS>(a, ca, e0, cnt, r1, b.px, b.py, b.pz, f, prs, w, ENS ? a.net[ic] : 0, lj_uni);
    if (!own) return;
    // the bead's position and image again (cache hits) rather than holding
    // twelve registers through the contact passes
    b.px = a.px[i]; b.py = a.py[i]; b.pz = a.pz[i];
    if (alias) bead_end<T, P, NEXT, LAST>(a, bx, i, b, f, b.px, b.py, b.pz, ms);
    else bead_end<T, P, NEXT, LAST>(a, bx, i, b, f, a.ix[i], a.iy[i], a.iz[i], ms);
}

__global__ void k_add_int(int *p, int v) { *p += v; }

// Copies src[0 .. min(*count, cap)) to dst: the used part of a list whose
// length is known on the device only.
__global__ void __launch_bounds__(kBlock) k_copy_bounded(const int *src, int *dst, const int *count, int cap)
{
    const int m = min(*count, cap);
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < m; k += gridDim.x * kBlock) dst[k] = src[k];
}

// The status bits of the chain layout's bond words from the bond-ordered
// statuses (set fields kept).
__global__ void __launch_bounds__(kBlock) k_chain_status(const int *ch_spring, const int *ch_angle,
                                                         const int *sp_status, const int *an_status,
                                                         unsigned *bond_code, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int s = ch_spring[i], q = ch_angle[i];
    unsigned c = bond_code[i] & ~(kSpringOn | kAngleOn);
    if (s >= 0 && sp_status[s] == 1) c |= kSpringOn;
    if (q >= 0 && an_status[q] == 1) c |= kAngleOn;
    bond_code[i] = c;
}

__global__ void __launch_bounds__(kBlock) k_fill_int(int *p, int v, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}


// ---- grid rebuild ----------------------------------------------------------------
//
// A rebuild lists, for every bead, the partners BuildContactList would list
// (kernel.cl:532-789): both beads inside the grid, in adjacent cells
// (SimulationManager.cpp:2193-2301, seen from both cells), not bonded
// neighbours of one fibre, and within the pair type's listing distance.
// The engine keeps a candidate list - every pair within reach + skin, found by
// a counting sort of the beads into the grid's cells and a search of the
// surrounding cells - and filters it at each rebuild with that exact rule on
// the current positions.  The candidates are rebuilt only when some bead has
// moved more than skin/2 since they were built (any pair within reach now was
// within reach + skin then), so the filtered list is the reference's list.
// The candidate kernels are gated by a device flag so that the sequence can
// be captured once and replayed.

__device__ __forceinline__ bool gated_off(const int *gate) { return gate && *gate == 0; }

// The Box of bead i's network in an ensemble, else the single network's.
template <bool ENS>
__device__ __forceinline__ const Box &net_box(const Box &single, const Box *boxes, const int *net, int i)
{
    if constexpr (ENS) return boxes[net[i]];
    else return single;
}

__global__ void __launch_bounds__(kBlock) k_fill_gated(int *p, int v, int n, const int *gate)
{
    if (gated_off(gate)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_set_int(int *p, int v, const int *gate)
{
    if (gated_off(gate)) return;
    *p = v;
}

// FillGrid's cell coordinate floor(v / step) (kernel.cl:493-495) for a grid
// of n cells on the axis: the cell itself inside [0, n), -1 below the grid
// or for NaN, n above it; converted to int only in range.
__device__ __forceinline__ int cell_coord(double v, double step, int n)
{
    const double c = floor(v / step);
    if (c >= 0.0 && c < (double)n) return (int)c;
    return c >= (double)n ? n : -1;
}

// FillGrid's cell of a bead (kernel.cl:492-504) in fp64 (-1 outside the grid)
// into key; the cell of its position clamped into the box into ckey, with the
// bead's rank among the beads of that cell (first pass of a counting sort).
// In an ensemble the cells are those of the bead's network, numbered from its
// Box's cell_base.
template <bool ENS>
__global__ void __launch_bounds__(kBlock) k_hash(const double *px, const double *py, const double *pz, int n, Box bs,
                                                 const int *net, const Box *boxes, int *key, int *ckey, int *rank,
                                                 int *count, const int *gate)
{
    if (gated_off(gate)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int c = -1;                                   // -1: no bead on this lane
    if (i < n) {
        const Box &b = net_box<ENS>(bs, boxes, net, i);
        const int cx = cell_coord(px[i], b.sx, b.cx);
        const int cy = cell_coord(py[i], b.sy, b.cy);
        const int cz = cell_coord(pz[i], b.sz, b.cz);
        const bool in = cx >= 0 && cx < b.cx && cy >= 0 && cy < b.cy && cz >= 0 && cz < b.cz;
        key[i] = in ? b.cell_base + cx + cy * b.cx + cz * b.cxy : -1;
        const int qx = min(max(cx, 0), b.cx - 1), qy = min(max(cy, 0), b.cy - 1), qz = min(max(cz, 0), b.cz - 1);
        c = b.cell_base + qx + qy * b.cx + qz * b.cxy;
        ckey[i] = c;
    }
    // consecutive beads (along a fibre) often share a cell: one atomic per
    // run of equal cells among adjacent lanes, ranks within the run in lane order
    const int prev = __shfl_up(c, 1, 64);
    const bool start = lane == 0 || prev != c;
    const unsigned long long starts = __ballot(start);
    const unsigned long long below = starts & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
    const int lead = 63 - __builtin_clzll(below);                 // the run's first lane
    const unsigned long long above = lane == 63 ? 0ull : (starts >> (lane + 1));
    const int len = above ? __builtin_ctzll(above) + 1 : 64 - lane;  // run length, read at its first lane
    int r0 = 0;
    if (start && c >= 0) r0 = atomicAdd(&count[c], len);
    const int base = __shfl(r0, lead, 64);
    if (i < n) rank[i] = base + (lane - lead);
}

// Second pass of the counting sort: each bead goes to slot cell_start[cell] +
// rank as one record (fp64 position, tag, loc, type, cell, bead index), read
// in bead order.  The slot's cell is stored as c for a bead inside the grid
// and as ~c for one placed by its clamped position.
__global__ void __launch_bounds__(kBlock) k_place(const int *key, const int *ckey, const int *rank,
                                                  const int *cell_start, const double *px, const double *py,
                                                  const double *pz, const int *type, const int *tag, const int *loc,
                                                  int n, Sorted so, int *count, int n_count, const int *gate)
{
    if (gated_off(gate)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    // the cell counts are spent (the offsets are scanned): zero them for the
    // next sort, whose hash adds into them
    for (int c = i; c < n_count; c += gridDim.x * kBlock) count[c] = 0;
    if (i >= n) return;
    const int c = ckey[i];
    const int s = cell_start[c] + rank[i];
    SortedBead b;
    b.x = px[i];
    b.y = py[i];
    b.z = pz[i];
    b.tag = tag[i];
    b.loc = loc[i];
    b.type = type[i];
    b.cell = key[i] >= 0 ? c : ~c;
    b.bead = i;
    b.pad = 0;
    so.bead[s] = b;
}

// Cell of a sorted slot (see k_place) and whether its bead is inside the grid.
__device__ __forceinline__ int slot_cell(int code) { return code >= 0 ? code : ~code; }

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

// Raises *p to v; the plain read first skips the atomic (which serialises on
// one address across the whole grid) once the maximum has settled.
__device__ __forceinline__ void raise_max(int *p, int v)
{
    if (v > *(volatile int *)p) atomicMax(p, v);
}

// The search rows of n beads in one scratch array of n * K entries: the
// first kRowHead entries of every bead packed densely (32 B per bead, read
// coalesced by the compaction; most lists are that short), the rest in a
// tail of K - kRowHead entries per bead after them.
constexpr int kRowHead = 8;
__device__ __forceinline__ size_t row_at(int i, int p, int n, int K)
{
    return p < kRowHead ? (size_t)i * kRowHead + p
                        : (size_t)n * kRowHead + (size_t)i * (K - kRowHead) + (size_t)(p - kRowHead);
}
struct SearchRow {
    int *scratch;
    int i, n, K;
    __device__ void put(int p, int v) const
    {
        if (p < K) scratch[row_at(i, p, n, K)] = v;
    }
};

// BuildContactList's distance (kernel.cl:600-628): minimum image per
// component, only in a periodic box and only beyond halfBoxSq.
__device__ __forceinline__ double list_len2(double rx, double ry, double rz, const Box &b)
{
    double len = rx * rx + ry * ry + rz * rz;
    if (b.bc == 0 && len > b.hbs) {
        if (rx * rx > b.hbs) rx += (rx > 0.0) ? -b.lx : b.lx;
        if (ry * ry > b.hbs) ry += (ry > 0.0) ? -b.ly : b.ly;
        if (rz * rz > b.hbs) rz += (rz > 0.0) ? -b.lz : b.lz;
        len = rx * rx + ry * ry + rz * rz;
    }
    return len;
}

// Partners of the bead in slot s among slots [lo, hi): not a bonded
// neighbour on the same fibre (kernel.cl:586-590), of a pair type that can
// be listed, and within r2 (candidates) or - EXACT - within the pair type's
// listing distance (kernel.cl:632-659; r2 is then the largest of them, a
// cheap first rejection).  Slot t's record is src[t - shift] (global memory
// or a staged copy in LDS).
template <bool EXACT>
__device__ __forceinline__ void scan_range_from(int lo, int hi, SortedBead nxt, int s, const SortedBead &me,
                                                double ri, double si, const SortedBead *src, int shift,
                                                const PairC *pairs, const double *radius, const double *skin,
                                                const Box &b, double r2, const SearchRow &row, int &k)
{
    for (int t = lo; t < hi; ++t) {
        const SortedBead o = nxt;                  // the next record is in flight while this one is tested
        if (t + 1 < hi) nxt = src[t + 1 - shift];
        const double len = list_len2(me.x - o.x, me.y - o.y, me.z - o.z, b);
        if (!(len < r2) || t == s) continue;
        if (me.tag == o.tag && abs(me.loc - o.loc) <= 1) continue;
        const int pid = pair_index(me.type, o.type, b.types);
        const int pt = pairs[pid].type;
        if (pt < 1 || pt > 3) continue;
        if (EXACT) {
            if (o.cell < 0) continue;   // FillGrid does not hash a bead outside the grid
            bool take = false;
            if (pt == 1 || pt == 3) take = len < b.cutoff_sq;
            if (pt == 2) {
                const int j = o.bead;
                const double rj = radius[j], sj = skin[j];
                const double min_skin = (si > sj) ? sj : si;
                const double max_r = (ri > rj) ? ri : rj;
                const double th = max_r + min_skin;
                take = len < th * th;
            }
            if (!take) continue;
        }
        row.put(k, o.bead | (pid << kPidShift));
        ++k;
    }
}

// The same over slots [lo, hi), loading the first record itself.
template <bool EXACT>
__device__ __forceinline__ void scan_range(int lo, int hi, int s, const SortedBead &me, double ri, double si,
                                           const SortedBead *src, int shift, const PairC *pairs,
                                           const double *radius, const double *skin, const Box &b, double r2,
                                           const SearchRow &row, int &k)
{
    if (lo >= hi) return;
    scan_range_from<EXACT>(lo, hi, src[lo - shift], s, me, ri, si, src, shift, pairs, radius, skin, b, r2, row, k);
}

// Cells visited along one axis around cell c: c-w..c+w, wrapped (periodic) or
// clipped (finite), or every cell once when `all`.
struct AxisRange {
    int c, n, w, all, periodic;
    __device__ int first() const { return all ? 0 : -w; }
    __device__ int last() const { return all ? n - 1 : w; }
    // cell for step d, -1 when clipped
    __device__ int at(int d) const
    {
        if (all) return d;
        int x = c + d;
        if (periodic) {
            x = x < 0 ? x + n : (x >= n ? x - n : x);
            return x;
        }
        return (x < 0 || x >= n) ? -1 : x;
    }
};

// Slot ranges of stencil row q (dz = q / 3 - 1, dy = q % 3 - 1) around cell
// (cx, cy, cz) for the 3x3x3 stencil with no axis visited whole: the cells
// cx - 1 .. cx + 1 of that row as one slot range, or two where a periodic
// row wraps (a finite box clips the row and skips rows outside the grid).
__device__ __forceinline__ void stencil_row(const int *cell_start, const Box &b, int cx, int cy, int cz, int q,
                                            int &lo0, int &hi0, int &lo1, int &hi1)
{
    lo0 = hi0 = lo1 = hi1 = 0;
    const bool per = b.bc == 0;
    int z = cz + q / 3 - 1, y = cy + q % 3 - 1;
    if (per) {
        z = z < 0 ? z + b.cz : (z >= b.cz ? z - b.cz : z);
        y = y < 0 ? y + b.cy : (y >= b.cy ? y - b.cy : y);
    } else if (z < 0 || z >= b.cz || y < 0 || y >= b.cy) {
        return;
    }
    const int rb = b.cell_base + y * b.cx + z * b.cxy;
    if (!per || (cx >= 1 && cx + 1 < b.cx)) {
        const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx + 1 < b.cx ? cx + 1 : b.cx - 1;
        lo0 = cell_start[rb + x0];
        hi0 = cell_start[rb + x1 + 1];
    } else if (cx == 0) {                      // cells cx - 1 = b.cx - 1, then 0 and 1
        lo0 = cell_start[rb + b.cx - 1];
        hi0 = cell_start[rb + b.cx];
        lo1 = cell_start[rb];
        hi1 = cell_start[rb + 2];
    } else {                                   // cx = b.cx - 1: cells b.cx - 2, b.cx - 1, then 0
        lo0 = cell_start[rb + b.cx - 2];
        hi0 = cell_start[rb + b.cx];
        lo1 = cell_start[rb];
        hi1 = cell_start[rb + 1];
    }
}

// Candidate search of the bead in slot s over the stencil's cells; the cells
// of one x-row are one contiguous slot range unless the row wraps or is
// clipped.  Up to K entries go to the bead's scratch row; cnt[bead] = count.
template <bool EXACT, bool ENS>
__global__ void __launch_bounds__(kBlock) k_search(const int *cell_start, int n_cells, Sorted so,
                                                   const PairC *pairs_all, const double *radius, const double *skin,
                                                   Box bs, const int *net, const Box *boxes, Stencil st, double r2,
                                                   int *cnt, int *scratch, int K, int *stats, const int *gate)
{
    if (gated_off(gate)) return;
    const int s = xcd_block() * kBlock + threadIdx.x;   // neighbouring slot ranges share one XCD's L2
    int k = 0, occ = 0;
    bool dense = false;
    if (s < cell_start[n_cells]) {
        const SortedBead me = so.bead[s];
        const int i = me.bead;
        // an ensemble's network Box by value: the scans read its fields for
        // every candidate, which through a reference would be loads that the
        // row stores keep in the loop
        const Box b = net_box<ENS>(bs, boxes, net, i);
        const PairC *pairs = pairs_all + b.pair_base;
        const int code = me.cell;
        const int cg = slot_cell(code);
        if (cell_start[cg] == s) occ = cell_start[cg + 1] - s;   // the cell's first slot reports its occupancy
        dense = cell_start[cg + 1] - cell_start[cg] > kMaxCellBeads;
        const int c = cg - b.cell_base;                          // cell within the bead's network
        const int cz = c / b.cxy;
        const int cy = (c - cz * b.cxy) / b.cx;
        const int cx = c - cz * b.cxy - cy * b.cx;
        const double ri = EXACT ? radius[i] : 0.0, si = EXACT ? skin[i] : 0.0;
        const SearchRow row{scratch, i, cell_start[n_cells], K};
        const int per = b.bc == 0 ? 1 : 0;
        // an ensemble's networks have grids of their own: a periodic axis
        // shorter than the stencil is visited whole (once per cell)
        const int all0 = st.all[0] | ((ENS && per && b.cx < 2 * st.w[0] + 1) ? 1 : 0);
        const int all1 = st.all[1] | ((ENS && per && b.cy < 2 * st.w[1] + 1) ? 1 : 0);
        const int all2 = st.all[2] | ((ENS && per && b.cz < 2 * st.w[2] + 1) ? 1 : 0);
        const AxisRange ax{cx, b.cx, st.w[0], all0, per}, ay{cy, b.cy, st.w[1], all1, per},
            az{cz, b.cz, st.w[2], all2, per};
        const bool row_run = !all0 && cx - st.w[0] >= 0 && cx + st.w[0] < b.cx;
        if (dense) {
            // refused: the host reports the run invalid
        } else if (EXACT && code >= 0 && st.w[0] == 1 && st.w[1] == 1 && st.w[2] == 1 && !all0 && !all1 && !all2) {
            // the reference's 3x3x3 stencil: the nine rows' slot ranges (two
            // where a periodic row wraps) in one round trip, then the rows
            int lo[9], hi[9], lo2[9], hi2[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) stencil_row(cell_start, b, cx, cy, cz, q, lo[q], hi[q], lo2[q], hi2[q]);
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                scan_range<EXACT>(lo[q], hi[q], s, me, ri, si, so.bead, 0, pairs, radius, skin, b, r2, row, k);
                scan_range<EXACT>(lo2[q], hi2[q], s, me, ri, si, so.bead, 0, pairs, radius, skin, b, r2, row, k);
            }
        } else
        for (int dz = az.first(); dz <= az.last() && !(EXACT && code < 0); ++dz) {
            const int z = az.at(dz);
            if (z < 0) continue;
            for (int dy = ay.first(); dy <= ay.last(); ++dy) {
                const int y = ay.at(dy);
                if (y < 0) continue;
                const int rbase = b.cell_base + y * b.cx + z * b.cxy;
                if (row_run) {
                    scan_range<EXACT>(cell_start[rbase + cx - st.w[0]], cell_start[rbase + cx + st.w[0] + 1], s, me,
                                      ri, si, so.bead, 0, pairs, radius, skin, b, r2, row, k);
                } else {
                    for (int dx = ax.first(); dx <= ax.last(); ++dx) {
                        const int x = ax.at(dx);
                        if (x < 0) continue;
                        scan_range<EXACT>(cell_start[rbase + x], cell_start[rbase + x + 1], s, me, ri, si, so.bead,
                                          0, pairs, radius, skin, b, r2, row, k);
                    }
                }
            }
        }
        cnt[i] = min(k, K);   // the row holds at most K; the count past it goes to the statistics
    }
    const int mx = wave_max(k);
    const int over = wave_sum(k > K ? 1 : 0);
    const int mocc = wave_max(occ);
    const int big = wave_sum(occ > 64 ? 1 : 0);
    const int nd = wave_sum(dense ? 1 : 0);
    if ((threadIdx.x & 63) == 0) {
        raise_max(&stats[EXACT ? kStatMaxList : kStatMaxCand], mx);
        if (over > 0) atomicAdd(&stats[kStatListOver], over);
        raise_max(&stats[kStatMaxCell], mocc);
        if (big > 0) atomicAdd(&stats[kStatCellsOver], big);
        if (nd > 0) atomicAdd(&stats[kStatDense], nd);
    }
}

// Candidates of a wave staged per pass, and the owner records; pair-table
// entries whose types the flat search keeps in LDS.
constexpr int kFlatChunk = 768;
constexpr int kFlatPairCap = 64;
struct FlatOwner {
    double x, y, z;
    int tag, loc, type, slot, bead, pad;
};
struct WaveSearchLds {
    int cand[kFlatChunk];      // candidate slot | owner lane << kPidShift
    int cnt[64];               // entries each lane's bead has collected
    FlatOwner own[64];
};

// The same search as k_search for the 3x3x3 stencil, flattened over the
// wave: the wave's beads lay their candidate slots end to end in LDS, every
// lane tests every 64th of them against its owner (so a wave costs its total
// candidates / 64 rounds, and the loads of a round are independent and
// mostly consecutive), and an accepted entry takes the next place in its
// owner's row through an LDS counter.  The rows hold the same entries as
// k_search's, in another order, which the compaction's per-bead sort
// removes.
template <bool EXACT>
__global__ void __launch_bounds__(kBlock) k_search_flat(const int *cell_start, int n_cells, Sorted so,
                                                        const PairC *pairs_all, const double *radius,
                                                        const double *skin, Box b, double r2, int *cnt, int *scratch,
                                                        int K, int *stats, const int *gate)
{
    if (gated_off(gate)) return;
    __shared__ WaveSearchLds wl[kBlock / 64];
    __shared__ int ptab[kFlatPairCap];
    WaveSearchLds &w = wl[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int s = xcd_block() * kBlock + threadIdx.x;
    const PairC *pairs = pairs_all + b.pair_base;
    // pair types of the table in LDS (the table of a handful of bead types)
    const int np = b.types * (b.types + 1) / 2;
    const bool lds_pairs = np <= kFlatPairCap;
    if (lds_pairs)
        for (int k = threadIdx.x; k < np; k += kBlock) ptab[k] = pairs[k].type;
    __syncthreads();
    const int n_slots = cell_start[n_cells];
    const bool live = s < n_slots;
    int i = 0, code = -1, occ = 0, cx = 0, cy = 0, cz = 0;
    bool elig = false, dense = false;
    if (live) {
        const SortedBead me = so.bead[s];
        i = me.bead;
        code = me.cell;
        const int cg = slot_cell(code);
        if (cell_start[cg] == s) occ = cell_start[cg + 1] - s;
        dense = cell_start[cg + 1] - cell_start[cg] > kMaxCellBeads;
        const int c = cg - b.cell_base;
        cz = c / b.cxy;
        cy = (c - cz * b.cxy) / b.cx;
        cx = c - cz * b.cxy - cy * b.cx;
        elig = !(EXACT && code < 0) && !dense;
        w.own[lane] = FlatOwner{me.x, me.y, me.z, me.tag, me.loc, me.type, s, i, 0};
    }
    w.cnt[lane] = 0;
    int nb = 0;
    if (elig) {
        for (int q = 0; q < 9; ++q) {
            int lo0, hi0, lo1, hi1;
            stencil_row(cell_start, b, cx, cy, cz, q, lo0, hi0, lo1, hi1);
            nb += (hi0 - lo0) + (hi1 - lo1);
        }
        if (nb > (1 << 24)) {          // keeps the wave's total below 2^30
            nb = 0;
            elig = false;
            dense = true;
        }
    }
    int incl = nb;
    incl = wave_incl_scan(incl);
    const int off = incl - nb;
    const int total = __shfl(incl, 63, 64);
    for (int base = 0; base < total; base += kFlatChunk) {          // wave-uniform
        if (elig && off < base + kFlatChunk && off + nb > base) {
            int pos = off;
            for (int q = 0; q < 9; ++q) {
                int lo[2], hi[2];
                stencil_row(cell_start, b, cx, cy, cz, q, lo[0], hi[0], lo[1], hi[1]);
                for (int h = 0; h < 2; ++h) {
                    const int len = hi[h] - lo[h];
                    const int a0 = pos > base ? pos : base;
                    const int a1 = (pos + len < base + kFlatChunk) ? pos + len : base + kFlatChunk;
                    for (int k = a0; k < a1; ++k)
                        w.cand[k - base] = (int)((unsigned)(lo[h] + (k - pos)) | ((unsigned)lane << kPidShift));
                    pos += len;
                }
            }
        }
        wave_sync_lds();
        const int n = (total - base) < kFlatChunk ? total - base : kFlatChunk;
        for (int r0 = 0; r0 < n; r0 += 64) {
            const int k = r0 + lane;
            if (k < n) {
                const int e = w.cand[k];
                const int t = e & kIdxMask, L = (int)((unsigned)e >> kPidShift);
                const SortedBead o = so.bead[t];
                const int jt = o.bead;
                const int ct = EXACT ? o.cell : 0;
                const FlatOwner me = w.own[L];
                const double len = list_len2(me.x - o.x, me.y - o.y, me.z - o.z, b);
                bool take = len < r2 && t != me.slot && !(me.tag == o.tag && abs(me.loc - o.loc) <= 1);
                int pid = 0;
                if (take) {
                    pid = pair_index(me.type, o.type, b.types);
                    const int pt = lds_pairs ? ptab[pid] : pairs[pid].type;
                    take = pt >= 1 && pt <= 3;
                    if (EXACT && take) {
                        take = ct >= 0;               // FillGrid does not hash a bead outside the grid
                        if (take && (pt == 1 || pt == 3)) take = len < b.cutoff_sq;
                        if (take && pt == 2) {
                            const double ri = radius[me.bead], si = skin[me.bead];
                            const double rj = radius[jt], sj = skin[jt];
                            const double min_skin = (si > sj) ? sj : si;
                            const double max_r = (ri > rj) ? ri : rj;
                            const double th = max_r + min_skin;
                            take = len < th * th;
                        }
                    }
                }
                if (take) {
                    const int p = atomicAdd(&w.cnt[L], 1);
                    if (p < K) scratch[row_at(me.bead, p, n_slots, K)] = jt | (pid << kPidShift);
                }
            }
        }
        wave_sync_lds();
    }
    const int kc = w.cnt[lane];
    if (live) cnt[i] = min(kc, K);
    const int mx = wave_max(kc);
    const int over = wave_sum(kc > K ? 1 : 0);
    const int mocc = wave_max(occ);
    const int big = wave_sum(occ > 64 ? 1 : 0);
    const int nd = wave_sum(dense ? 1 : 0);
    if (lane == 0) {
        raise_max(&stats[EXACT ? kStatMaxList : kStatMaxCand], mx);
        if (over > 0) atomicAdd(&stats[kStatListOver], over);
        raise_max(&stats[kStatMaxCell], mocc);
        if (big > 0) atomicAdd(&stats[kStatCellsOver], big);
        if (nd > 0) atomicAdd(&stats[kStatDense], nd);
    }
}

// The exact 3x3x3 search of an ensemble, flattened over the wave: each lane
// takes its bead's nine stencil rows (one or two slot ranges each, wrapped
// by its own network's periodic grid) into registers once, the wave lays
// all their candidate slots end to end in LDS, and every lane tests every
// 64th candidate against its owner with the owner's network Box.  In the
// sparse grids of an ensemble a bead has about one candidate in 18 ranges,
// so the per-bead search ran each range's loop for a few active lanes.
// Every network's grid has at least three cells per axis (the host checks).
template <bool EXACT>
__global__ void __launch_bounds__(kBlock) k_search_ens_wave(const int *cell_start, int n_cells, Sorted so,
                                                            const PairC *pairs_all, const double *radius,
                                                            const double *skin, const int *net, const Box *boxes,
                                                            double r2, int *cnt, int *scratch, int K, int *stats,
                                                            const int *gate)
{
    if (gated_off(gate)) return;
    __shared__ WaveSearchLds wl[kBlock / 64];
    WaveSearchLds &w = wl[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int s = xcd_block() * kBlock + threadIdx.x;
    const int n_slots = cell_start[n_cells];
    const bool live = s < n_slots;
    int i = 0, occ = 0, nb = 0;
    bool dense = false;
    int lo[9], hi[9], lo2[9], hi2[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) lo[q] = hi[q] = lo2[q] = hi2[q] = 0;
    if (live) {
        const SortedBead me = so.bead[s];
        i = me.bead;
        const int nw = net[i];
        const Box b = boxes[nw];
        const int code = me.cell;
        const int cg = slot_cell(code);
        if (cell_start[cg] == s) occ = cell_start[cg + 1] - s;
        dense = cell_start[cg + 1] - cell_start[cg] > kMaxCellBeads;
        const int c = cg - b.cell_base;
        const int cz = c / b.cxy;
        const int cy = (c - cz * b.cxy) / b.cx;
        const int cx = c - cz * b.cxy - cy * b.cx;
        w.own[lane] = FlatOwner{me.x, me.y, me.z, me.tag, me.loc, me.type, s, i, nw};
        if (!(EXACT && code < 0) && !dense) {
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                stencil_row(cell_start, b, cx, cy, cz, q, lo[q], hi[q], lo2[q], hi2[q]);
                nb += (hi[q] - lo[q]) + (hi2[q] - lo2[q]);
            }
            if (nb > (1 << 24)) {          // keeps the wave's total below 2^30
                nb = 0;
                dense = true;
#pragma unroll
                for (int q = 0; q < 9; ++q) lo[q] = hi[q] = lo2[q] = hi2[q] = 0;
            }
        }
    }
    w.cnt[lane] = 0;
    int incl = nb;
    incl = wave_incl_scan(incl);
    const int off = incl - nb;
    const int total = __shfl(incl, 63, 64);
    for (int base = 0; base < total; base += kFlatChunk) {          // wave-uniform
        if (nb > 0 && off < base + kFlatChunk && off + nb > base) {
            int pos = off;
#pragma unroll
            for (int q = 0; q < 18; ++q) {
                const int a = q < 9 ? lo[q] : lo2[q - 9], e = q < 9 ? hi[q] : hi2[q - 9];
                const int len = e - a;
                const int a0 = pos > base ? pos : base;
                const int a1 = (pos + len < base + kFlatChunk) ? pos + len : base + kFlatChunk;
                for (int k = a0; k < a1; ++k)
                    w.cand[k - base] = (int)((unsigned)(a + (k - pos)) | ((unsigned)lane << kPidShift));
                pos += len;
            }
        }
        wave_sync_lds();
        const int n = (total - base) < kFlatChunk ? total - base : kFlatChunk;
        for (int r0 = 0; r0 < n; r0 += 64) {
            const int k = r0 + lane;
            if (k < n) {
                const int e = w.cand[k];
                const int t = e & kIdxMask, L = (int)((unsigned)e >> kPidShift);
                const SortedBead o = so.bead[t];
                const FlatOwner me = w.own[L];
                const Box &b = boxes[me.pad];             // the owner's network
                const double len = list_len2(me.x - o.x, me.y - o.y, me.z - o.z, b);
                bool take = len < r2 && t != me.slot && !(me.tag == o.tag && abs(me.loc - o.loc) <= 1);
                int pid = 0;
                if (take) {
                    pid = pair_index(me.type, o.type, b.types);
                    const int pt = pairs_all[b.pair_base + pid].type;
                    take = pt >= 1 && pt <= 3;
                    if (EXACT && take) {
                        take = o.cell >= 0;               // FillGrid does not hash a bead outside the grid
                        if (take && (pt == 1 || pt == 3)) take = len < b.cutoff_sq;
                        if (take && pt == 2) {
                            const double ri = radius[me.bead], si = skin[me.bead];
                            const double rj = radius[o.bead], sj = skin[o.bead];
                            const double min_skin = (si > sj) ? sj : si;
                            const double max_r = (ri > rj) ? ri : rj;
                            const double th = max_r + min_skin;
                            take = len < th * th;
                        }
                    }
                }
                if (take) {
                    const int p = atomicAdd(&w.cnt[L], 1);
                    if (p < K) scratch[row_at(me.bead, p, n_slots, K)] = o.bead | (pid << kPidShift);
                }
            }
        }
        wave_sync_lds();
    }
    const int kc = w.cnt[lane];
    if (live) cnt[i] = min(kc, K);
    const int mx = wave_max(kc);
    const int over = wave_sum(kc > K ? 1 : 0);
    const int mocc = wave_max(occ);
    const int big = wave_sum(occ > 64 ? 1 : 0);
    const int nd = wave_sum(dense ? 1 : 0);
    if (lane == 0) {
        raise_max(&stats[EXACT ? kStatMaxList : kStatMaxCand], mx);
        if (over > 0) atomicAdd(&stats[kStatListOver], over);
        raise_max(&stats[kStatMaxCell], mocc);
        if (big > 0) atomicAdd(&stats[kStatCellsOver], big);
        if (nd > 0) atomicAdd(&stats[kStatDense], nd);
    }
}

// The same search with the block's neighbourhood staged in LDS.  A block of
// kStageBlock consecutive slots covers cells [c_lo, c_hi]; for every (dy, dz)
// of the stencil the cells its beads can reach in that row offset form one
// contiguous cell (so slot) interval [c_lo - wx + off, c_hi + wx + off].
// The block copies these intervals into LDS with coalesced loads (as many
// as fit); each thread then reads a candidate range from LDS when the
// range lies inside its staged interval, and from global memory otherwise
// (rows that wrap around a periodic box, intervals that did not fit).
constexpr int kStageBlock = 128;
constexpr int kStageCap = 832;        // records (48 B each) staged per block
constexpr int kStageRows = 25;        // (2w+1)^2 for w <= 2

template <bool EXACT>
__global__ void __launch_bounds__(kStageBlock) k_search_staged(const int *cell_start, int n_cells,
                                                               Sorted so, const PairC *pairs, const double *radius,
                                                               const double *skin, Box b, Stencil st, double r2,
                                                               int *cnt, int *scratch, int K, int *stats,
                                                               const int *gate)
{
    if (gated_off(gate)) return;
    __shared__ SortedBead rec[kStageCap];
    __shared__ int lo_q[kStageRows], hi_q[kStageRows], base_q[kStageRows + 1];
    const int nh = cell_start[n_cells];
    const int s0 = xcd_block() * kStageBlock;
    if (s0 >= nh) return;
    const int s_end = min(s0 + kStageBlock, nh);
    const int wy = st.w[1], wz = st.w[2], ny = 2 * wy + 1, nq = ny * (2 * wz + 1);
    if ((int)threadIdx.x < nq) {          // one thread per row offset: the interval's slots
        const int q = threadIdx.x;
        const int c_lo = slot_cell(so.bead[s0].cell), c_hi = slot_cell(so.bead[s_end - 1].cell);
        const int dy = q % ny - wy, dz = q / ny - wz;
        const long long off = (long long)dy * b.cx + (long long)dz * b.cxy;
        const long long a = max(0LL, (long long)c_lo - st.w[0] + off);
        const long long e = min((long long)n_cells, (long long)c_hi + st.w[0] + 1 + off);
        int lo = 0, hi = 0;
        if (a < e) {
            lo = cell_start[a];
            hi = cell_start[e];
        }
        lo_q[q] = lo;
        hi_q[q] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {                // staging offsets, as many intervals as fit
        int base = 0;
        for (int q = 0; q < nq; ++q) {
            if (base + (hi_q[q] - lo_q[q]) > kStageCap) hi_q[q] = lo_q[q] = 0;   // read from global memory
            base_q[q] = base;
            base += hi_q[q] - lo_q[q];
        }
        base_q[nq] = base;
    }
    __syncthreads();
    const int total = base_q[nq];
    for (int k = threadIdx.x; k < total; k += kStageBlock) {
        int q = 0;
        while (k >= base_q[q + 1]) ++q;
        rec[k] = so.bead[lo_q[q] + (k - base_q[q])];
    }
    __syncthreads();

    const int s = s0 + threadIdx.x;
    int k = 0, occ = 0;
    bool dense = false;
    if (s < s_end) {
        const SortedBead me = so.bead[s];
        const int i = me.bead;
        const int code = me.cell;
        const int c = slot_cell(code);
        if (cell_start[c] == s) occ = cell_start[c + 1] - s;
        dense = cell_start[c + 1] - cell_start[c] > kMaxCellBeads;
        const int cz = c / b.cxy;
        const int cy = (c - cz * b.cxy) / b.cx;
        const int cx = c - cz * b.cxy - cy * b.cx;
        const double ri = EXACT ? radius[i] : 0.0, si = EXACT ? skin[i] : 0.0;
        const SearchRow row{scratch, i, cell_start[n_cells], K};
        const int per = b.bc == 0 ? 1 : 0;
        const AxisRange ax{cx, b.cx, st.w[0], st.all[0], per}, ay{cy, b.cy, wy, st.all[1], per},
            az{cz, b.cz, wz, st.all[2], per};
        const bool row_run = !st.all[0] && cx - st.w[0] >= 0 && cx + st.w[0] < b.cx;
        for (int dz = az.first(); dz <= az.last() && !(EXACT && code < 0) && !dense; ++dz) {
            const int z = az.at(dz);
            if (z < 0) continue;
            for (int dy = ay.first(); dy <= ay.last(); ++dy) {
                const int y = ay.at(dy);
                if (y < 0) continue;
                const int rbase = y * b.cx + z * b.cxy;
                // staged interval of this row offset (none for axes visited whole)
                const int q = (st.all[1] || st.all[2]) ? -1 : (dz + wz) * ny + (dy + wy);
                const int slo = q >= 0 ? lo_q[q] : 0, shi = q >= 0 ? hi_q[q] : 0;
                const int sbase = q >= 0 ? base_q[q] : 0;
                auto scan = [&](int lo, int hi) {
                    if (lo >= hi) return;
                    if (lo >= slo && hi <= shi)
                        scan_range<EXACT>(lo, hi, s, me, ri, si, rec + sbase, slo, pairs, radius, skin, b, r2, row,
                                          k);
                    else
                        scan_range<EXACT>(lo, hi, s, me, ri, si, so.bead, 0, pairs, radius, skin, b, r2, row, k);
                };
                if (row_run) {
                    scan(cell_start[rbase + cx - st.w[0]], cell_start[rbase + cx + st.w[0] + 1]);
                } else {
                    for (int dx = ax.first(); dx <= ax.last(); ++dx) {
                        const int x = ax.at(dx);
                        if (x < 0) continue;
                        scan(cell_start[rbase + x], cell_start[rbase + x + 1]);
                    }
                }
            }
        }
        cnt[i] = min(k, K);   // the row holds at most K; the count past it goes to the statistics
    }
    const int mx = wave_max(k);
    const int over = wave_sum(k > K ? 1 : 0);
    const int mocc = wave_max(occ);
    const int big = wave_sum(occ > 64 ? 1 : 0);
    const int nd = wave_sum(dense ? 1 : 0);
    if ((threadIdx.x & 63) == 0) {
        raise_max(&stats[EXACT ? kStatMaxList : kStatMaxCand], mx);
        if (over > 0) atomicAdd(&stats[kStatListOver], over);
        raise_max(&stats[kStatMaxCell], mocc);
        if (big > 0) atomicAdd(&stats[kStatCellsOver], big);
        if (nd > 0) atomicAdd(&stats[kStatDense], nd);
    }
}

// Scratch rows (one per bead) -> the candidate CSR, each bead's
// entries in ascending order.  The order of beads inside a cell follows the
// atomic counting sort and is arbitrary; ordering the entries makes the
// lists, and so the contact-force summation order, deterministic.
// With cnt_out (the final list): also the list length of every bead.
// A wave's beads are consecutive, so their lists form one range of the CSR:
// when it fits kCompactStage entries, each bead sorts its row into the
// wave's LDS and the wave writes the range with consecutive lanes on
// consecutive entries; otherwise each bead writes its own list.
constexpr int kCompactStage = 1024;

// Sorts bead i's row of m entries (m > 0) into out[0..m).
__device__ __forceinline__ void sort_row(const int *scratch, int i, int m, int n, int K, int *out)
{
    if (m <= kRowHead) {
        // the head: two aligned 16-byte loads (entries past m are masked)
        const int4 h0 = *(const int4 *)(scratch + (size_t)i * kRowHead);
        const int4 h1 = *(const int4 *)(scratch + (size_t)i * kRowHead + 4);
        const int hv[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
        int v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = k < m ? hv[k] : 0x7fffffff;
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
        const int v = scratch[row_at(i, k, n, K)];
        int j = k - 1;
        while (j >= 0 && out[j] > v) {
            out[j + 1] = out[j];
            --j;
        }
        out[j + 1] = v;
    }
}

__global__ void __launch_bounds__(kBlock) k_compact(const int *cnt, const int *cand_start, const int *scratch, int K,
                                                    int n, int *cand, int cap, int *cnt_out, int *stats,
                                                    Snapshot snap, const int *gate)
{
    if (gated_off(gate)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {         // the build's statistics
        stats[kStatCandBuilds] += 1;
        stats[kStatCandMax] = max(stats[kStatCandMax], cand_start[n]);
    }
    {
        // the positions the candidates were built at (the displacement check's reference)
        const int b = blockIdx.x * kBlock + threadIdx.x;
        if (snap.rx && b < n) {
            snap.rx[b] = snap.px[b];
            snap.ry[b] = snap.py[b];
            snap.rz[b] = snap.pz[b];
        }
    }
    __shared__ int stage[kBlock / 64][kCompactStage];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int lane = threadIdx.x & 63;
    if (i - lane >= n) return;                         // wave-uniform: no bead in this wave
    const bool live = i < n;
    int m = live ? min(cnt[i], K) : 0;
    const int base = live ? cand_start[i] : 0;
    if (base < 0 || m > cap - base) m = 0;   // capacity exceeded: the host reports the run invalid
    if (live && cnt_out) cnt_out[i] = m;
    const int wbase = __shfl(base, 0, 64);             // lane 0 holds a bead
    const int span = wave_max(m > 0 ? base + m : wbase) - wbase;   // every written entry lies below cap
    if (span > kCompactStage) {                        // wave-uniform
        if (m > 0) sort_row(scratch, i, m, n, K, cand + base);
        return;
    }
    int *w = stage[threadIdx.x >> 6];
    if (m > 0) sort_row(scratch, i, m, n, K, w + (base - wbase));
    wave_sync_lds();
    for (int k = lane; k < span; k += 64) cand[wbase + k] = w[k];
}

// Sets *need when some bead has moved more than sqrt(lim2) from its position
// at the last candidate build (minimum image in a periodic box; NaN counts as
// moved).
template <bool ENS>
__global__ void __launch_bounds__(kBlock) k_disp(const double *px, const double *py, const double *pz,
                                                 const double *rx, const double *ry, const double *rz, int n, Box bs,
                                                 Box gs, const int *net, const Box *boxes, CellPack pack, double lim2,
                                                 int *need, CellBead *cb, int *key)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    bool moved = false;
    if (i < n) {
        // an ensemble's bead: its network's box and grid (cells local to it in cb)
        const Box &b = net_box<ENS>(bs, boxes, net, i);
        const Box &g = ENS ? b : gs;
        const double x = px[i], y = py[i], z = pz[i];
        double dx = x - rx[i], dy = y - ry[i], dz = z - rz[i];
        if (b.bc == 0) {
            dx = dx > 0.5 * b.lx ? dx - b.lx : (dx < -0.5 * b.lx ? dx + b.lx : dx);
            dy = dy > 0.5 * b.ly ? dy - b.ly : (dy < -0.5 * b.ly ? dy + b.ly : dy);
            dz = dz > 0.5 * b.lz ? dz - b.lz : (dz < -0.5 * b.lz ? dz + b.lz : dz);
        }
        moved = !(dx * dx + dy * dy + dz * dz <= lim2);
        // FillGrid's cell (kernel.cl:493-499), once per bead for the filter
        const int cx = cell_coord(x, g.sx, g.cx), cy = cell_coord(y, g.sy, g.cy), cz = cell_coord(z, g.sz, g.cz);
        const bool in = cx >= 0 && cx < g.cx && cy >= 0 && cy < g.cy && cz >= 0 && cz < g.cz;
        cb[i] = CellBead{(float)x, (float)y, (float)z, in ? (cx | (cy << pack.sy) | (cz << pack.sz)) : -1};
        key[i] = in ? (ENS ? g.cell_base : 0) + cx + cy * g.cx + cz * g.cxy : -1;
    }
    if (__any(moved) && (threadIdx.x & 63) == 0) *need = 1;
}

__global__ void __launch_bounds__(kBlock) k_snapshot(const double *px, const double *py, const double *pz, double *rx,
                                                     double *ry, double *rz, int n, const int *gate)
{
    if (gated_off(gate)) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    rx[i] = px[i];
    ry[i] = py[i];
    rz[i] = pz[i];
}

// The exact listing rule applied to the candidates: bead i's listed partners
// (in candidate order, so ascending) to list[cand_start[i] ...] and their
// number to cnt[i].  The candidates of a wave are evaluated round-robin by
// all its lanes (each lane every 64th entry) against the owning bead's
// record in LDS, the verdicts pass through LDS, a running count over them
// gives every accepted entry its place in its bead's list, and the wave
// writes the pass's accepted entries with consecutive lanes on consecutive
// entries.
//
// A partner is read as its 16-byte CellBead.  With d the rounded coordinate
// differences, |d - exact| <= delta per axis, the squared distance len from
// them differs from the fp64 one by at most 2 sqrt(3) delta sqrt(len_exact)
// + 3 delta^2 (+ its own rounding).  So, for a threshold T = root_t^2 with
// root_t >= 2 delta: len > 4T means beyond T exactly too, and otherwise a
// pair more than 12 delta (2 root_t + delta) + 4e-14 T from T is decided by
// len; the few within that band (and, in a periodic box, pairs that may wrap)
// are decided by the fp64 positions.  The verdicts are those of the fp64
// rule.

// A bead of a wave as the filter reads it: its rounded position (the
// CellBead's), its index, its cell and eb = cand_start - (its first entry in
// the wave's entries), so that entry k of the wave is candidate eb + k.  Its
// fp64 position, radius and skin are read from the bead arrays only by the
// verdicts that need them (the fp64 decision, a penalty pair).
struct alignas(16) FilterOwner {
    float x, y, z;
    int idx;
    int cx, cy, cz, eb;
    int net;                    // the bead's network in an ensemble
};

__device__ __forceinline__ bool needs_fp64(double len, double t, double root_t, double delta)
{
    if (root_t < 2.0 * delta) return true;
    return len <= 4.0 * t && fabs(len - t) <= 12.0 * delta * (2.0 * root_t + delta) + 4e-14 * t;
}

template <bool ENS>
__global__ void __launch_bounds__(kBlock) k_filter(FilterArgs f)
{
    __shared__ __align__(16) int verdict[kBlock / 64][kWaveChunk];
    __shared__ __align__(16) int before[kBlock / 64][kWaveChunk + 4];   // accepted entries of the pass before each slot
    __shared__ unsigned owner_words[kBlock / 64][kWaveChunk / 4];
    __shared__ FilterOwner fo[kBlock];
    __shared__ int ptype[kPairCap];          // the pair table's types (entries a list addresses: at most kPairCap)
    if (!ENS)
        for (int k = threadIdx.x; k < kPairCap; k += kBlock) ptype[k] = k < f.n_pairs ? f.pairs[k].type : 0;
    __syncthreads();
    const int wv = threadIdx.x >> 6;
    int *acc = verdict[wv];
    int *pre = before[wv];
    unsigned char *own = reinterpret_cast<unsigned char *>(owner_words[wv]);
    FilterOwner *ow = fo + wv * 64;
    const int lane = threadIdx.x & 63;
    const int i = xcd_block() * kBlock + threadIdx.x;
    const bool live = i < f.n;
    const int ic = live ? i : f.n - 1;
    const CellBead mine = f.cb[ic];
    const int mcell = mine.cell;
    const bool in = mcell >= 0;
    const int e0 = f.cand_start[ic];
    const int e1 = f.cand_start[ic + 1];
    // past the capacity: the host reports the run invalid
    const int cnt = (live && in && e0 >= 0 && e1 >= e0 && e1 <= f.cap) ? e1 - e0 : 0;
    int incl = cnt;
    incl = wave_incl_scan(incl);
    const int off = incl - cnt;
    const int total = __shfl(incl, 63, 64);
    const int my = (1 << (f.pack.sz - f.pack.sy)) - 1, mx = (1 << f.pack.sy) - 1;
    // the wave's accepted entries go end to end from the start of its
    // candidates (dense: the bead kernels read a wave's lists as one run)
    const int wb = __shfl(e0, 0, 64);
    int kept = 0, wkept = 0;
    if (total > 0) {                                     // wave-uniform
        fo[threadIdx.x] = FilterOwner{mine.x, mine.y, mine.z, ic, mcell & mx, (mcell >> f.pack.sy) & my,
                                      mcell >> f.pack.sz, e0 - off, ENS ? f.net[ic] : 0};
        for (int base = 0; base < total; base += kWaveChunk) {
            mark_owner(off, cnt, base, own);             // also publishes the records
            const int n = (total - base) < kWaveChunk ? total - base : kWaveChunk;
            // every candidate of the pass, then every partner record, before
            // any verdict: the loads of the pass are in flight together
            int ent[kChunkRounds];
#pragma unroll
            for (int r = 0; r < kChunkRounds; ++r) {
                const int k = r * 64 + lane;
                ent[r] = k < n ? f.cand[ow[own[k]].eb + base + k] : 0;
            }
            CellBead q[kChunkRounds];
#pragma unroll
            for (int r = 0; r < kChunkRounds; ++r) {
                const int j = ent[r] & kIdxMask;
                // entries of an overflowed build may be stale: never follow one outside the arrays
                const bool valid = r * 64 + lane < n && j < f.n && (unsigned)(ent[r] >> kPidShift) < (unsigned)f.n_pairs;
                if (!valid) ent[r] = -1;
                q[r] = f.cb[valid ? j : 0];
            }
#pragma unroll
            for (int r = 0; r < kChunkRounds; ++r) {
                if (r * 64 >= n) break;                  // wave-uniform
                int v = -1;
                if (ent[r] >= 0 && q[r].cell >= 0) {
                    const FilterOwner &o = ow[own[r * 64 + lane]];
                    // the owner's box, grid and pair table (an ensemble's: its network's)
                    const Box &b = ENS ? f.boxes[o.net] : f.list;
                    const Box &g = ENS ? b : f.grid;
                    const bool per = b.bc == 0;
                    const int j = ent[r] & kIdxMask;
                    const int qc = q[r].cell;
                    const int ax = abs((qc & mx) - o.cx), ay = abs(((qc >> f.pack.sy) & my) - o.cy),
                              az = abs((qc >> f.pack.sz) - o.cz);
                    bool adj = (ax <= 1 || (per && ax == g.cx - 1)) && (ay <= 1 || (per && ay == g.cy - 1)) &&
                               (az <= 1 || (per && az == g.cz - 1));
                    const int pid = ent[r] >> kPidShift;
                    if (ENS && pid >= b.types * (b.types + 1) / 2) adj = false;   // not an entry of its table
                    if (adj) {
                        const int pt = ENS ? f.pairs[b.pair_base + pid].type
                                           : (pid < kPairCap ? ptype[pid] : f.pairs[pid].type);
                        double t = -1.0, root_t = 0.0;
                        if (pt == 1 || pt == 3) {
                            t = b.cutoff_sq;
                            root_t = ENS ? sqrt(b.cutoff_sq) : f.cutoff;
                        } else if (pt == 2) {            // kernel.cl:632-659
                            const double ri = f.radius[o.idx], si = f.skin[o.idx];
                            const double rj = f.radius[j], sj = f.skin[j];
                            const double min_skin = (si > sj) ? sj : si;
                            const double max_r = (ri > rj) ? ri : rj;
                            root_t = max_r + min_skin;
                            t = root_t * root_t;
                        }
                        if (t >= 0.0) {
                            const double dx = (double)o.x - (double)q[r].x;
                            const double dy = (double)o.y - (double)q[r].y;
                            const double dz = (double)o.z - (double)q[r].z;
                            const double len = dx * dx + dy * dy + dz * dz;
                            bool take = len < t;
                            if (needs_fp64(len, t, root_t, f.delta) || (per && !(len < f.hbs_guard))) {
                                const int oi = o.idx;
                                const double le =
                                    list_len2(f.px[oi] - f.px[j], f.py[oi] - f.py[j], f.pz[oi] - f.pz[j], b);
                                take = le < t;
                            }
                            if (take) v = ent[r];
                        }
                    }
                }
                acc[r * 64 + lane] = v;
            }
            wave_sync_lds();
            // accepted entries before each slot: a lane's four slots, then
            // the lanes before it
            const int4 v4 = *reinterpret_cast<const int4 *>(acc + 4 * lane);
            const int t0 = 4 * lane;
            const int a0 = (t0 < n && v4.x >= 0) ? 1 : 0, a1 = (t0 + 1 < n && v4.y >= 0) ? 1 : 0;
            const int a2 = (t0 + 2 < n && v4.z >= 0) ? 1 : 0, a3 = (t0 + 3 < n && v4.w >= 0) ? 1 : 0;
            const int sum4 = a0 + a1 + a2 + a3;
            const int incl4 = wave_incl_scan(sum4);
            const int ex = incl4 - sum4;
            *reinterpret_cast<int4 *>(pre + t0) = make_int4(ex, ex + a0, ex + a0 + a1, ex + a0 + a1 + a2);
            if (lane == 63) pre[kWaveChunk] = incl4;
            wave_sync_lds();
            // this bead's count so far
            {
                const int s0 = min(max(off - base, 0), n), s1 = min(max(off + cnt - base, 0), n);
                kept += pre[s1] - pre[s0];
            }
#pragma unroll
            for (int r = 0; r < kChunkRounds; ++r) {
                if (r * 64 >= n) break;                  // wave-uniform
                const int t = r * 64 + lane;
                const int v = acc[t];
                if (t < n && v >= 0) f.list_out[wb + wkept + pre[t]] = v;
            }
            wkept += pre[kWaveChunk];
            wave_sync_lds();
        }
    }
    {
        // each bead's list starts after the wave's earlier beads' lists
        const int ik = wave_incl_scan(kept);
        if (live) {
            f.cnt[i] = kept;
            f.list_start[i] = wb + ik - kept;
        }
        if (i == f.n - 1) f.list_start[f.n] = f.cand_start[f.n];   // the extent the lists lie in
    }
    const int mxk = wave_max(kept);
    if (lane == 0) raise_max(&f.stats[kStatMaxList], mxk);
    // the candidates are current again: no kernel of this rebuild reads the flag after the compaction
    if (f.need && blockIdx.x == 0 && threadIdx.x == 0) *f.need = 0;
}

// ---- device-wide exclusive scan (n ints -> n+1 offsets) ----------------------------

constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ int block_excl_scan(int v, int *lds, int &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    x = wave_incl_scan(x);
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

// A block's tile of kScanTile elements is read as kScanItems / 4 rounds of
// kBlock 16-byte chunks, consecutive threads on consecutive chunks: in round
// j thread t holds elements tile + 4 (j kBlock + t) .. + 3 (in and out are
// device allocations, so 16-byte aligned).
constexpr int kScanRounds = kScanItems / 4;
__device__ __forceinline__ int4 scan_load4(const int *in, int idx, int n)
{
    if (idx + 3 < n) return *(const int4 *)(in + idx);
    int4 v{0, 0, 0, 0};
    if (idx < n) v.x = in[idx];
    if (idx + 1 < n) v.y = in[idx + 1];
    if (idx + 2 < n) v.z = in[idx + 2];
    return v;
}

__global__ void __launch_bounds__(kBlock) k_scan_reduce(const int *in, int n, int *bsum, const int *gate)
{
    if (gated_off(gate)) return;
    __shared__ int lds[kBlock / 64];
    int s = 0;
#pragma unroll
    for (int j = 0; j < kScanRounds; ++j) {
        const int4 v = scan_load4(in, blockIdx.x * kScanTile + (j * kBlock + (int)threadIdx.x) * 4, n);
        s += v.x + v.y + v.z + v.w;
    }
    int total;
    block_excl_scan(s, lds, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(const int *in, int n, const int *bsum, int nb, int *out,
                                                       const int *gate)
{
    if (gated_off(gate)) return;
    __shared__ int lds[kBlock / 64];
    int4 v[kScanRounds];
#pragma unroll
    for (int j = 0; j < kScanRounds; ++j)
        v[j] = scan_load4(in, blockIdx.x * kScanTile + (j * kBlock + (int)threadIdx.x) * 4, n);
    // this block's offset: the sum of the block totals before it (and, for
    // block 0, of all of them: the scan's total)
    const int upto = blockIdx.x == 0 ? nb : (int)blockIdx.x;
    int part = 0;
    for (int k = threadIdx.x; k < upto; k += kBlock) part += bsum[k];
    int carry;
    block_excl_scan(part, lds, carry);
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) out[n] = carry;
        carry = 0;
    }
#pragma unroll
    for (int j = 0; j < kScanRounds; ++j) {
        const int idx = blockIdx.x * kScanTile + (j * kBlock + (int)threadIdx.x) * 4;
        int total;
        const int e = carry + block_excl_scan(v[j].x + v[j].y + v[j].z + v[j].w, lds, total);
        const int4 o{e, e + v[j].x, e + v[j].x + v[j].y, e + v[j].x + v[j].y + v[j].z};
        if (idx + 3 < n) {
            *(int4 *)(out + idx) = o;
        } else {
            if (idx < n) out[idx] = o.x;
            if (idx + 1 < n) out[idx + 1] = o.y;
            if (idx + 2 < n) out[idx + 2] = o.z;
        }
        carry += total;
    }
}

// ---- explicit instantiation through launchers ----------------------------------------

static inline int blocks(int n) { return (n + kBlock - 1) / kBlock; }

template <typename T, typename P>
void Launch<T, P>::advance_position(const Args<T, P> &a, int n_active, hipStream_t s)
{
    if (n_active > 0) k_advance_position<T, P><<<blocks(n_active), kBlock, 0, s>>>(a, n_active);
}
template <typename T, typename P>
void Launch<T, P>::spring_terms(const Args<T, P> &a, int total, hipStream_t s)
{
    if (total > 0) k_spring_terms<T, P><<<blocks(total), kBlock, 0, s>>>(a, total);
}
template <typename T, typename P>
void Launch<T, P>::angle_terms(const Args<T, P> &a, int total, hipStream_t s)
{
    if (total > 0) k_angle_terms<T, P><<<blocks(total), kBlock, 0, s>>>(a, total);
}
template <typename T, typename P>
void Launch<T, P>::resolve_springs(const Args<T, P> &a, int total, hipStream_t s)
{
    spring_terms(a, total, s);
    k_gather_terms<T, P><<<blocks(a.n), kBlock, 0, s>>>(a, a.sp_start, a.sp_term);
}
template <typename T, typename P>
void Launch<T, P>::resolve_angles(const Args<T, P> &a, int total, hipStream_t s)
{
    angle_terms(a, total, s);
    k_gather_terms<T, P><<<blocks(a.n), kBlock, 0, s>>>(a, a.an_start, a.an_term);
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
    if (a.net) k_drift_first<T, P, true><<<blocks(a.n), kBlock, 0, s>>>(a);
    else k_drift_first<T, P, false><<<blocks(a.n), kBlock, 0, s>>>(a);
}
template <typename T, typename P, int M, int BLK>
static void launch_fused_mask(const Args<T, P> &a, hipStream_t s)
{
    constexpr bool spr = (M & 1) != 0, ang = (M & 2) != 0, next = (M & 4) != 0, last = (M & 8) != 0;
    constexpr bool ens = (M & 32) != 0;
    if (M & 16) {
        constexpr int per_block = kChainOwned * (BLK / 64);
        const size_t dyn = a.ms_code ? sizeof(MassSet<T>) * (size_t)a.n_ms : 0;
        k_bead_step_chain<T, P, spr, ang, next, last, ens, BLK><<<(a.n + per_block - 1) / per_block, BLK, dyn, s>>>(a);
    } else {
        k_bead_step<T, P, spr, ang, next, last, ens, BLK><<<(a.n + BLK - 1) / BLK, BLK, 0, s>>>(a);
    }
}

template <typename T, typename P, int BLK, int... M>
static void dispatch_fused(const Args<T, P> &a, int mask, hipStream_t s, std::integer_sequence<int, M...>)
{
    using Fn = void (*)(const Args<T, P> &, hipStream_t);
    static constexpr Fn table[] = {&launch_fused_mask<T, P, M, BLK>...};
    table[mask](a, s);
}

template <typename T, typename P>
void Launch<T, P>::bead_step(const Args<T, P> &a, bool spr, bool ang, bool next, bool last, bool chain, int block,
                             hipStream_t s)
{
    const int mask =
        (spr ? 1 : 0) | (ang ? 2 : 0) | (next ? 4 : 0) | (last ? 8 : 0) | (chain ? 16 : 0) | (a.net ? 32 : 0);
    if (block == 128) dispatch_fused<T, P, 128>(a, mask, s, std::make_integer_sequence<int, 64>{});
    else if (block == 512) dispatch_fused<T, P, 512>(a, mask, s, std::make_integer_sequence<int, 64>{});
    else dispatch_fused<T, P, 256>(a, mask, s, std::make_integer_sequence<int, 64>{});
}
template struct Launch<double, double>;
template struct Launch<float, double>;

void chain_status(const int *ch_spring, const int *ch_angle, const int *sp_status, const int *an_status,
                  unsigned *bond_code, int n, hipStream_t s)
{
    if (n > 0) k_chain_status<<<blocks(n), kBlock, 0, s>>>(ch_spring, ch_angle, sp_status, an_status, bond_code, n);
}
void fill_int(int *p, int v, int n, hipStream_t s)
{
    if (n > 0) k_fill_int<<<blocks(n), kBlock, 0, s>>>(p, v, n);
}
void add_int(int *p, int v, hipStream_t s) { k_add_int<<<1, 1, 0, s>>>(p, v); }

// Plain streaming copy, 16 B per lane and iteration, grid-stride.
__global__ void __launch_bounds__(kBlock) k_copy_probe(const int4 *src, int4 *dst, size_t n)
{
    for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (size_t)gridDim.x * kBlock) dst[k] = src[k];
}

void copy_probe(const void *src, void *dst, size_t n16, hipStream_t s)
{
    k_copy_probe<<<4096, kBlock, 0, s>>>((const int4 *)src, (int4 *)dst, n16);
}
void fill_gated(int *p, int v, int n, const int *gate, hipStream_t s)
{
    if (n > 0) k_fill_gated<<<blocks(n), kBlock, 0, s>>>(p, v, n, gate);
}
void set_int(int *p, int v, const int *gate, hipStream_t s) { k_set_int<<<1, 1, 0, s>>>(p, v, gate); }
void hash(const double *px, const double *py, const double *pz, int n, const Box &b, const int *net,
          const Box *boxes, int *key, int *ckey, int *rank, int *count, const int *gate, hipStream_t s)
{
    if (net) k_hash<true><<<blocks(n), kBlock, 0, s>>>(px, py, pz, n, b, net, boxes, key, ckey, rank, count, gate);
    else k_hash<false><<<blocks(n), kBlock, 0, s>>>(px, py, pz, n, b, nullptr, nullptr, key, ckey, rank, count, gate);
}
void place(const int *key, const int *ckey, const int *rank, const int *cell_start, const double *px,
           const double *py, const double *pz, const int *type, const int *tag, const int *loc, int n,
           const Sorted &so, int *count, int n_count, const int *gate, hipStream_t s)
{
    k_place<<<blocks(std::max(n, 1)), kBlock, 0, s>>>(key, ckey, rank, cell_start, px, py, pz, type, tag, loc, n, so,
                                                       count, n_count, gate);
}
template <bool EXACT>
static void search_t(const int *cell_start, int n_cells, const Sorted &so, const PairC *pairs, const double *radius,
                     const double *skin, const Box &b, const int *net, const Box *boxes, const Stencil &st, double r2,
                     bool ens_grids3, int *cnt, int *scratch, int K, int *stats, int n, bool staged, bool flat,
                     const int *gate, hipStream_t s)
{
    const bool stencil3 = st.w[0] == 1 && st.w[1] == 1 && st.w[2] == 1 && !st.all[0] && !st.all[1] && !st.all[2];
    if (!net && flat && !staged && stencil3)
        k_search_flat<EXACT><<<blocks(n), kBlock, 0, s>>>(cell_start, n_cells, so, pairs, radius, skin, b, r2, cnt,
                                                          scratch, K, stats, gate);
    else if (net && flat && stencil3 && ens_grids3)
        k_search_ens_wave<EXACT><<<blocks(n), kBlock, 0, s>>>(cell_start, n_cells, so, pairs, radius, skin, net,
                                                              boxes, r2, cnt, scratch, K, stats, gate);
    else if (net)
        k_search<EXACT, true><<<blocks(n), kBlock, 0, s>>>(cell_start, n_cells, so, pairs, radius, skin, b, net,
                                                           boxes, st, r2, cnt, scratch, K, stats, gate);
    else if (staged && st.w[1] <= 2 && st.w[2] <= 2)
        k_search_staged<EXACT><<<(n + kStageBlock - 1) / kStageBlock, kStageBlock, 0, s>>>(
            cell_start, n_cells, so, pairs, radius, skin, b, st, r2, cnt, scratch, K, stats, gate);
    else
        k_search<EXACT, false><<<blocks(n), kBlock, 0, s>>>(cell_start, n_cells, so, pairs, radius, skin, b,
                                                            nullptr, nullptr, st, r2, cnt, scratch, K, stats, gate);
}
void search(const int *cell_start, int n_cells, const Sorted &so, const PairC *pairs, const double *radius,
            const double *skin, const Box &b, const int *net, const Box *boxes, const Stencil &st, double r2,
            bool ens_grids3, bool exact, int *cnt, int *scratch, int K, int *stats, int n, bool staged, bool flat,
            const int *gate, hipStream_t s)
{
    if (exact)
        search_t<true>(cell_start, n_cells, so, pairs, radius, skin, b, net, boxes, st, r2, ens_grids3, cnt, scratch,
                       K, stats, n, staged, flat, gate, s);
    else
        search_t<false>(cell_start, n_cells, so, pairs, radius, skin, b, net, boxes, st, r2, ens_grids3, cnt, scratch,
                        K, stats,
                        n, staged, flat, gate, s);
}
void compact(const int *cnt, const int *cand_start, const int *scratch, int K, int n, int *cand,
             int cap, int *cnt_out, int *stats, const Snapshot &snap, const int *gate, hipStream_t s)
{
    if (n > 0)
        k_compact<<<blocks(n), kBlock, 0, s>>>(cnt, cand_start, scratch, K, n, cand, cap, cnt_out, stats, snap, gate);
}
void disp_check(const double *px, const double *py, const double *pz, const double *rx, const double *ry,
                const double *rz, int n, const Box &b, const Box &grid, const int *net, const Box *boxes,
                CellPack pack, double lim2, int *need, CellBead *cb, int *key, hipStream_t s)
{
    if (net)
        k_disp<true><<<blocks(n), kBlock, 0, s>>>(px, py, pz, rx, ry, rz, n, b, grid, net, boxes, pack, lim2, need,
                                                  cb, key);
    else
        k_disp<false><<<blocks(n), kBlock, 0, s>>>(px, py, pz, rx, ry, rz, n, b, grid, nullptr, nullptr, pack, lim2,
                                                   need, cb, key);
}
void snapshot(const double *px, const double *py, const double *pz, double *rx, double *ry, double *rz, int n,
              const int *gate, hipStream_t s)
{
    k_snapshot<<<blocks(n), kBlock, 0, s>>>(px, py, pz, rx, ry, rz, n, gate);
}
void filter(const FilterArgs &f, hipStream_t s)
{
    if (f.net) k_filter<true><<<blocks(f.n), kBlock, 0, s>>>(f);
    else k_filter<false><<<blocks(f.n), kBlock, 0, s>>>(f);
}
void copy_bounded(const int *src, int *dst, const int *count, int cap, hipStream_t s)
{
    if (cap > 0) k_copy_bounded<<<std::min(blocks(cap), 4096), kBlock, 0, s>>>(src, dst, count, cap);
}
int scan_blocks_needed(int n) { return (n + kScanTile - 1) / kScanTile; }
void exclusive_scan(const int *in, int n, int *out, int *tmp, const int *gate, hipStream_t s)
{
    if (((uintptr_t)in | (uintptr_t)out) & 15) throw std::invalid_argument("exclusive_scan: unaligned buffers");
    const int nb = scan_blocks_needed(n > 0 ? n : 1);
    k_scan_reduce<<<nb, kBlock, 0, s>>>(in, n, tmp, gate);
    k_scan_apply<<<nb, kBlock, 0, s>>>(in, n, tmp, nb, out, gate);
}

}  // namespace fnk
