// adjoint.hip - reverse-mode derivative of the FibreNet step (gfx950).
//
// The step as the engine runs it (per bead, P = position, W = velocity_old,
// F' = force of the previous step, coefficients lam = lambda*dt/m,
// h = (dt/2)/m, h2 = (dt^2/2)/m):
//     x' = x + dt W + h2 F'                 drift (kernel.cl:209-223)
//     V  = W + lam F'                       predicted velocity (kernel.cl:205-207)
//     F  = D V + G(x')                      damping, gravity, walls, bonds, contacts
//     W' = W + h (F + F')                   kick (kernel.cl:459-461)
// With the contact lists and bond statuses of the forward run held fixed,
// the adjoint of one step, given the adjoints X, Wb of x', W' and Fb of F
// (Fb already holding the loss seed and h Wb), is
//     X  += J(x') Fb                        J = dG/dx (symmetric: every force
//                                            is central or derives from a
//                                            potential)
//     Vb  = D Fb
//     F'b = h Wb + lam Vb + h2 X
//     Wb' = Wb + Vb + dt X
// and the adjoint of a force scale s (force = s * unit force) gains Fb . F.
// J(x') Fb is the directional derivative of every force along Fb, evaluated
// here with dual numbers in the same per-bead gathers as the forces.
//
// Arithmetic: fp64 in both builds.  Force values (the scale adjoints) use
// the forward's own rounding where it matters: spring and contact lengths
// are the float-rounded native_sqrt of the forward (the extension l0 - l and
// the overlap r1 + r2 - r are differences of nearly equal numbers), angle
// values come from the forward's angle_force.  Derivatives are those of
// the exact expressions at the recorded positions.
#include <algorithm>

#include "kernels.h"
#include "engine.h"

namespace fnk {

// A dual number v + d eps (eps^2 = 0).
struct Dual {
    double v, d;
};
__device__ __forceinline__ Dual cst(double v) { return Dual{v, 0.0}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return Dual{-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.v * b.d + a.d * b.v}; }
__device__ __forceinline__ Dual operator*(double s, Dual a) { return Dual{s * a.v, s * a.d}; }
// 1/x and 1/sqrt(x) (x > 0 finite) from the hardware estimates refined by
// one Newton step: within a few ulps, which is all a derivative needs, at a
// third of the correctly rounded expansions' instructions.
__device__ __forceinline__ double rcp_d(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double rsq_d(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    return y * fma(-0.5 * x * y, y, 1.5);
}

// a / b through one reciprocal of b (every adjoint path shares it).
__device__ __forceinline__ Dual operator/(Dual a, Dual b)
{
    const double r = rcp_d(b.v);
    const double q = a.v * r;
    return Dual{q, (a.d - q * b.d) * r};
}
__device__ __forceinline__ Dual dsqrt(Dual a)
{
    const double r = sqrt(a.v);
    return Dual{r, r > 0.0 ? a.d / (2.0 * r) : 0.0};
}
__device__ __forceinline__ Dual datan2(Dual y, Dual x)
{
    const double n = x.v * x.v + y.v * y.v;
    return Dual{atan2(y.v, x.v), n > 0.0 ? (x.v * y.d - y.v * x.d) / n : 0.0};
}
// |r| whose value is the forward's float-rounded length (native_sqrt) and
// whose derivative is the exact one.
__device__ __forceinline__ Dual flen(Dual r2)
{
    return Dual{(double)native_sqrt(r2.v), r2.v > 0.0 ? r2.d * (0.5 * rsq_d(r2.v)) : 0.0};
}

struct DV3 {
    Dual x, y, z;
};
__device__ __forceinline__ DV3 dcross(const DV3 &a, const DV3 &b)
{
    return DV3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ Dual ddot(const DV3 &a, const DV3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ DV3 dscale(Dual s, const DV3 &a) { return DV3{s * a.x, s * a.y, s * a.z}; }

// Spring force on p1 for d = x_p1 - x_p2: k (l0 - l) d / l with the
// forward's rounded length (kernel.cl:384-437).
__device__ __forceinline__ DV3 spring_dual(double k, double l0, const DV3 &d)
{
    const Dual l = flen(ddot(d, d));
    if (l.v == 0.0) return DV3{cst(0), cst(0), cst(0)};
    return dscale((k * (cst(l0) - l)) / l, d);
}

// Angle forces on b1 (pa) and b3 (pc) from ba = x1 - x2, bc = x3 - x2: the
// negative gradient of k (theta - theta0)^2 (kernel.cl:289-348 in exact
// arithmetic).  With c = ba x bc and phi = pi - theta = atan2(|c|, -ba.bc)
// the force on b1 is
//     [2k phi/|c| - 2k (pi - theta0)/|c|] (ba x c) / |ba|^2
// and on b3 the same with (c x bc) / |bc|^2.  phi/|c| is a smooth function
// of |c|^2 (atan(t)/t is even in t), summed as its series for a nearly
// straight triple, so the derivative stays accurate where |c| is rounding
// noise; the second term, zero for theta0 = pi, is singular there and is
// kept only where |c|^2 > 1e-10 |ba|^2 |bc|^2 (below that the forward's
// own angle is dominated by rounding).
__device__ __forceinline__ void angle_dual(double k, double th0, const DV3 &ba, const DV3 &bc, DV3 &pa, DV3 &pc)
{
    const DV3 none{cst(0), cst(0), cst(0)};
    pa = pc = none;
    const DV3 c = dcross(ba, bc);
    const Dual c2 = ddot(c, c), la2 = ddot(ba, ba), lc2 = ddot(bc, bc), dp = ddot(ba, bc);
    if (la2.v == 0.0 || lc2.v == 0.0) return;
    Dual g;                                    // phi / |c|
    if (dp.v < 0.0) {
        const Dual m = -dp;
        const Dual u = c2 / (m * m);           // tan^2 phi
        Dual at;                               // atan(t) / t, t^2 = u
        if (u.v < 1e-4) {
            at = cst(1.0) + u * (cst(-1.0 / 3.0) + u * (cst(0.2) + u * (cst(-1.0 / 7.0) + u * cst(1.0 / 9.0))));
        } else {
            const Dual t = dsqrt(u);
            at = datan2(t, cst(1.0)) / t;
        }
        g = at / m;
    } else {
        if (c2.v == 0.0) return;               // folded onto itself: no direction
        const Dual cn = dsqrt(c2);
        g = datan2(cn, -dp) / cn;
    }
    Dual coef = (2.0 * k) * g;
    if (c2.v > 1e-10 * la2.v * lc2.v) coef = coef - cst(2.0 * k * (M_PI - th0)) / dsqrt(c2);
    pa = dscale(coef / la2, dcross(ba, c));
    pc = dscale(coef / lc2, dcross(c, bc));
}

// Central contact force on bead 1 for r = x1 - x2 (minimum image applied)
// as the forward evaluates it (contact_force): lj |F|/r along r (kernel.cl:
// 875-919), penalty (922-957), two-step harmonic (960-999), with the
// forward's rounded length.  Called only for pairs that exerted a force.
__device__ __forceinline__ DV3 contact_dual(const PairC &pr, double r1, double r2, const DV3 &r)
{
    const Dual len2 = ddot(r, r);
    const Dual rr = flen(len2);
    Dual coef;
    if (pr.type == 1) {
        const Dual il2 = cst(1.0) / len2;      // one reciprocal for both ratios
        const Dual fr2 = (pr.c1 * pr.c1) * il2;
        const Dual fr6 = fr2 * fr2 * fr2;
        coef = (48.0 * pr.c2) * fr6 * (fr6 - cst(0.5)) * il2 - cst(pr.c4);
    } else if (pr.type == 2) {
        coef = pr.c1 * (cst(r1 + r2) - rr);
    } else {
        coef = pr.c1 * (cst(pr.c2) - rr) + pr.c5 * (cst(pr.c6) - rr);
    }
    if (rr.v == 0.0) return DV3{cst(0), cst(0), cst(0)};
    return dscale(coef / rr, r);
}

__device__ __forceinline__ void add_d(double3 &z, const DV3 &f, double sgn)
{
    z.x += sgn * f.x.d;
    z.y += sgn * f.y.d;
    z.z += sgn * f.z.d;
}
__device__ __forceinline__ double dot_v(const double3 &u, const DV3 &f)
{
    return u.x * f.x.v + u.y * f.y.v + u.z * f.z.v;
}
__device__ __forceinline__ DV3 dvec(double3 x, double3 u)
{
    return DV3{Dual{x.x, u.x}, Dual{x.y, u.y}, Dual{x.z, u.z}};
}
__device__ __forceinline__ double3 sub3(double3 a, double3 b) { return double3{a.x - b.x, a.y - b.y, a.z - b.z}; }

// Spring terms of one adjoint step (thread per spring): J Fb of the spring
// force at both ends' incidence slots (zero for a spring that did not act)
// and the spring's scale adjoint Fb . F.  The force on p1 is F(x_p1 - x_p2),
// on p2 its negation.
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_adj_springs(AdjArgs<T, P> a)
{
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= a.springs) return;
    double3 j{0.0, 0.0, 0.0};
    if (s < a.n_springs && a.sp_on[s] == 1 && (a.sp_type[s] == 0 || a.sp_type[s] == 1)) {
        const int p1 = a.sp_p1[s], p2 = a.sp_p2[s];
        const double3 u = sub3(a.fb[p1], a.fb[p2]);
        const double3 d{(double)(a.ix[p1] - a.ix[p2]), (double)(a.iy[p1] - a.iy[p2]), (double)(a.iz[p1] - a.iz[p2])};
        const DV3 F = spring_dual((double)a.sp_k[s], a.sp_l0[s], dvec(d, u));
        j = double3{F.x.d, F.y.d, F.z.d};
        a.g_spring[s] += dot_v(u, F);
    }
    a.sp_jv[a.sp_slot[2 * s]] = j;
    a.sp_jv[a.sp_slot[2 * s + 1]] = double3{-j.x, -j.y, -j.z};
}

// Angle terms of one adjoint step (thread per angle): J Fb of pa at the b1
// slot, of pc at the b3 slot, of -pa-pc at the b2 slot, and the angle's
// scale adjoint from the forward's own force values.
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_adj_angles(AdjArgs<T, P> a)
{
    const int q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= a.angles) return;
    double3 ja{0.0, 0.0, 0.0}, jc{0.0, 0.0, 0.0};
    if (q < a.n_angles && a.an_on[q] == 1) {
        const int b1 = a.an_b1[q], b2 = a.an_b2[q], b3 = a.an_b3[q];
        const double3 x2{(double)a.ix[b2], (double)a.iy[b2], (double)a.iz[b2]};
        const double3 ba{(double)a.ix[b1] - x2.x, (double)a.iy[b1] - x2.y, (double)a.iz[b1] - x2.z};
        const double3 bc{(double)a.ix[b3] - x2.x, (double)a.iy[b3] - x2.y, (double)a.iz[b3] - x2.z};
        const double3 u1 = a.fb[b1], u2 = a.fb[b2], u3 = a.fb[b3];
        const double3 da = sub3(u1, u2), dc = sub3(u3, u2);
        DV3 pa, pc;
        angle_dual((double)a.an_k[q], (double)a.an_th0[q], dvec(ba, da), dvec(bc, dc), pa, pc);
        ja = double3{pa.x.d, pa.y.d, pa.z.d};
        jc = double3{pc.x.d, pc.y.d, pc.z.d};
        V3<T> fa, fc;                             // the forward's own values
        angle_force(AngleP<T>{1, a.an_type[q], a.an_th0[q], a.an_k[q]}, a.ix[b1], a.iy[b1], a.iz[b1], a.ix[b2],
                    a.iy[b2], a.iz[b2], a.ix[b3], a.iy[b3], a.iz[b3], fa, fc);
        a.g_angle[q] += da.x * (double)fa.x + da.y * (double)fa.y + da.z * (double)fa.z + dc.x * (double)fc.x +
                        dc.y * (double)fc.y + dc.z * (double)fc.z;
    }
    a.an_jv[a.an_slot[3 * q]] = ja;
    a.an_jv[a.an_slot[3 * q + 1]] = jc;
    a.an_jv[a.an_slot[3 * q + 2]] = double3{-ja.x - jc.x, -ja.y - jc.y, -ja.z - jc.z};
}

__device__ __forceinline__ void add_slots(double3 &z, const double3 *jv, int e0, int e1)
{
    for (int e = e0; e < e1; ++e) {
        const double3 t = jv[e];
        z.x += t.x;
        z.y += t.y;
        z.z += t.z;
    }
}

// Pair-table entries the bead kernel keeps in LDS, and pair entries whose
// scale adjoints it sums per block in LDS (more go per bead to global memory).
constexpr int kAdjPairCap = 32;
constexpr int kAdjPairsLds = 8;

// The bead part of one adjoint step (thread per bead).  Reads the positions
// at which the step's forces were evaluated, the bond terms at the bead's
// incidence slots, the step's contact list and Fb (adjoint of this step's
// force, seed and h Wb included); updates X, Wb, writes Fb of the previous
// step (seed and h Wb of that step included) to fb_next, and adds the
// contacts' scale adjoints: LDSG, per block row (g_pair[block * n_pairs +
// entry], summed over the block in a fixed order); otherwise per bead row.
template <typename T, typename P, bool LDSG, bool ENS>
__global__ void __launch_bounds__(kBlock) k_adjoint_step(AdjArgs<T, P> a)
{
    __shared__ PairC prs[ENS ? 1 : kAdjPairCap];
    __shared__ double gl[LDSG ? kAdjPairsLds * kBlock : 1];
    const int np = a.n_pairs;
    if (!ENS)
        for (int k = threadIdx.x; k < min(np, kAdjPairCap); k += kBlock) prs[k] = a.pairs[k];
    if (LDSG)
        for (int k = 0; k < np; ++k) gl[k * kBlock + threadIdx.x] = 0.0;
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < a.n) {
        const double3 fb = a.fb[i];
        double3 z{0.0, 0.0, 0.0};
        if (a.springs) add_slots(z, a.sp_jv, a.sp_start[i], a.sp_start[i + 1]);
        if (a.angles) add_slots(z, a.an_jv, a.an_start[i], a.an_start[i + 1]);
        // contacts, from the bead's own side of each listed pair
        const Box &bf = ENS ? a.boxes[a.net[i]] : a.force;
        const PairC *table = ENS ? a.pairs + bf.pair_base : prs;
        const P xi = a.px[i], yi = a.py[i], zi = a.pz[i];
        const T ri = a.radius[i];
        for (int e = a.nb_start[i], e1 = e + a.nb_cnt[i]; e < e1; ++e) {
            const int ent = a.nb_list[e];
            const int j = ent & kIdxMask, pid = ent >> kPidShift;
            const PairC pr = table[pid];
            P rx = xi - a.px[j], ry = yi - a.py[j], rz = zi - a.pz[j];
            P len = rx * rx + ry * ry + rz * rz;
            if (len > (P)bf.hbs) {               // ResolveContacts' minimum image (kernel.cl:844-870)
                if (rx * rx > (P)bf.hbs) rx += (rx > (P)0) ? (P)-bf.lx : (P)bf.lx;
                if (ry * ry > (P)bf.hbs) ry += (ry > (P)0) ? (P)-bf.ly : (P)bf.ly;
                if (rz * rz > (P)bf.hbs) rz += (rz > (P)0) ? (P)-bf.lz : (P)bf.lz;
                len = rx * rx + ry * ry + rz * rz;
            }
            V3<T> t;
            if (!contact_force(pr, bf, rx, ry, rz, len, ri, a.radius, j, t)) continue;   // the forward's mask
            const double rj = pr.type == 2 ? (double)a.radius[j] : 0.0;
            const DV3 F = contact_dual(pr, (double)ri, rj,
                                       dvec(double3{(double)rx, (double)ry, (double)rz}, sub3(fb, a.fb[j])));
            add_d(z, F, 1.0);
            if (LDSG) gl[pid * kBlock + threadIdx.x] += dot_v(fb, F);
            else a.g_pair[(size_t)i * np + pid] += dot_v(fb, F);
        }
        // walls of a finite box (kernel.cl:252-261): diagonal
        const Box &bd = ENS ? bf : a.drift;
        if (bd.bc == 1) {
            const double k = -bd.wall_k;
            if (xi < (P)0 || xi > (P)bd.lx) z.x += k * fb.x;
            if (yi < (P)0 || yi > (P)bd.ly) z.y += k * fb.y;
            if (zi < (P)0 || zi > (P)bd.lz) z.z += k * fb.z;
        }
        double3 X = a.xb[i];
        X.x += z.x;
        X.y += z.y;
        X.z += z.z;
        const double lam = a.lam[i], h = a.hlf[i], h2 = a.hlf2[i];
        const double dt = (double)(T)bd.dt, D = (double)(T)bd.damping;
        const double3 vb{D * fb.x, D * fb.y, D * fb.z};
        double3 W = a.wb[i];
        const double3 fprev{h * W.x + lam * vb.x + h2 * X.x, h * W.y + lam * vb.y + h2 * X.y,
                            h * W.z + lam * vb.z + h2 * X.z};
        W.x += vb.x + dt * X.x;
        W.y += vb.y + dt * X.y;
        W.z += vb.z + dt * X.z;
        a.xb[i] = X;
        a.wb[i] = W;
        // Fb of the previous step: its own force adjoint, its kick's h Wb and its seed
        const double seed = (a.seed_prev && a.group[i]) ? a.seed_prev[ENS ? a.net[i] : 0] : 0.0;
        a.fb_next[i] = double3{fprev.x + h * W.x + seed, fprev.y + h * W.y, fprev.z + h * W.z};
    }
    if (LDSG) {
        for (int k = 0; k < np; ++k)
            for (int w = kBlock / 2; w > 0; w >>= 1) {
                __syncthreads();
                if ((int)threadIdx.x < w) gl[k * kBlock + threadIdx.x] += gl[k * kBlock + threadIdx.x + w];
            }
        __syncthreads();
        if ((int)threadIdx.x < np) a.g_pair[(size_t)blockIdx.x * np + threadIdx.x] += gl[threadIdx.x * kBlock];
    }
}

// A lane's bead as the contact passes read it: position and force adjoint.
struct AdjOwner {
    double x, y, z, fx, fy, fz;
};

// A wave's LDS for its contacts in k_adjoint_step_flat / _chain: the J Fb
// terms of one pass, the lane that owns each entry, the wave's scale
// adjoints, and every lane's bead (position, force adjoint, radius).  Four
// blocks of four waves fit a CU's LDS in the fp32 build, as many as the
// registers allow.
template <typename T>
struct AdjWaveLds {
    double3 term[kWaveChunk];
    unsigned char own[kWaveChunk];
    double g[kAdjPairsLds];
    AdjOwner ow[64];
    T orad[64];
};

// k_adjoint_step for a plain network with at most kAdjPairsLds pair-table
// entries, its contacts evaluated by the whole wave as in the forward's
// contacts_finish: the wave's lists laid end to end, each lane evaluating
// every 64th entry (the forward's mask, then the dual force) and leaving its
// J Fb term in LDS, each bead then adding its own terms in list order.  The
// scale adjoints Fb . F accumulate per evaluating lane and pair entry, are
// summed over the wave by a fixed shuffle tree and over the block's waves in
// order, into one row per block as in the LDSG form.
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) k_adjoint_step_flat(AdjArgs<T, P> a)
{
    __shared__ PairC prs[kAdjPairsLds];
    __shared__ PairTab<T> prf[kAdjPairsLds];          // the forward's form, for its mask
    __shared__ AdjWaveLds<T> wl[kBlock / 64];
    const int np = a.n_pairs;
    for (int k = threadIdx.x; k < min(np, kAdjPairsLds); k += kBlock) {
        prs[k] = a.pairs[k];
        prf[k] = to_tab<T>(a.pairs[k]);
    }
    __syncthreads();
    AdjWaveLds<T> &w = wl[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < a.n;
    const int ic = live ? i : a.n - 1;
    const double3 fb = a.fb[ic];
    double3 z{0.0, 0.0, 0.0};
    if (live && a.springs) add_slots(z, a.sp_jv, a.sp_start[i], a.sp_start[i + 1]);
    if (live && a.angles) add_slots(z, a.an_jv, a.an_start[i], a.an_start[i + 1]);
    const Box &bf = a.force;
    const P xi = a.px[ic], yi = a.py[ic], zi = a.pz[ic];
    const T ri = a.radius[ic];
    const int e0 = a.nb_start[ic], cnt = live ? a.nb_cnt[ic] : 0;
    int incl = cnt;                                   // inclusive scan of the counts over the wave
    incl = wave_incl_scan(incl);
    const int off = incl - cnt;
    const int total = __shfl(incl, 63, 64);
    double g[kAdjPairsLds];
#pragma unroll
    for (int p = 0; p < kAdjPairsLds; ++p) g[p] = 0.0;
    w.ow[lane] = AdjOwner{(double)xi, (double)yi, (double)zi, fb.x, fb.y, fb.z};
    w.orad[lane] = ri;
    for (int base = 0; base < total; base += kWaveChunk) {           // wave-uniform
        const int n = (total - base) < kWaveChunk ? total - base : kWaveChunk;
        mark_owner(off, cnt, base, w.own);
        for (int r = 0; r * 64 < n; ++r) {
            const int k = r * 64 + lane;
            int L;
            const int e = owner_entry(k, k < n, e0, off - base, w.own, L);
            const AdjOwner o = w.ow[L];
            const P xL = (P)o.x, yL = (P)o.y, zL = (P)o.z;
            const T rL = w.orad[L];
            const double3 fL{o.fx, o.fy, o.fz};
            double3 t{-0.0, -0.0, -0.0};
            if (k < n) {
                const int ent = a.nb_list[e];
                const int j = ent & kIdxMask, pid = ent >> kPidShift;
                const PairC pr = prs[pid];
                P rx = xL - a.px[j], ry = yL - a.py[j], rz = zL - a.pz[j];
                P len = rx * rx + ry * ry + rz * rz;
                if (len > (P)bf.hbs) {           // ResolveContacts' minimum image (kernel.cl:844-870)
                    if (rx * rx > (P)bf.hbs) rx += (rx > (P)0) ? (P)-bf.lx : (P)bf.lx;
                    if (ry * ry > (P)bf.hbs) ry += (ry > (P)0) ? (P)-bf.ly : (P)bf.ly;
                    if (rz * rz > (P)bf.hbs) rz += (rz > (P)0) ? (P)-bf.lz : (P)bf.lz;
                    len = rx * rx + ry * ry + rz * rz;
                }
                V3<T> tf;
                if (contact_force(prf[pid], bf, rx, ry, rz, len, rL, a.radius, j, tf)) {   // the forward's mask
                    const double rj = pr.type == 2 ? (double)a.radius[j] : 0.0;
                    const DV3 F = contact_dual(pr, (double)rL, rj,
                                               dvec(double3{(double)rx, (double)ry, (double)rz}, sub3(fL, a.fb[j])));
                    t = double3{F.x.d, F.y.d, F.z.d};
                    const double gv = dot_v(fL, F);
#pragma unroll
                    for (int p = 0; p < kAdjPairsLds; ++p)
                        if (p == pid) g[p] += gv;
                }
            }
            if (k < n) w.term[k] = t;
        }
        wave_sync_lds();
        const int k0 = off > base ? off : base;
        const int k1 = (off + cnt < base + n) ? off + cnt : base + n;
        for (int k = k0; k < k1; ++k) {
            const double3 t = w.term[k - base];
            z.x += t.x;
            z.y += t.y;
            z.z += t.z;
        }
        wave_sync_lds();
    }
    if (live) {
        // walls of a finite box (kernel.cl:252-261): diagonal
        const Box &bd = a.drift;
        if (bd.bc == 1) {
            const double k = -bd.wall_k;
            if (xi < (P)0 || xi > (P)bd.lx) z.x += k * fb.x;
            if (yi < (P)0 || yi > (P)bd.ly) z.y += k * fb.y;
            if (zi < (P)0 || zi > (P)bd.lz) z.z += k * fb.z;
        }
        double3 X = a.xb[i];
        X.x += z.x;
        X.y += z.y;
        X.z += z.z;
        const double lam = a.lam[i], h = a.hlf[i], h2 = a.hlf2[i];
        const double dt = (double)(T)bd.dt, D = (double)(T)bd.damping;
        const double3 vb{D * fb.x, D * fb.y, D * fb.z};
        double3 W = a.wb[i];
        const double3 fprev{h * W.x + lam * vb.x + h2 * X.x, h * W.y + lam * vb.y + h2 * X.y,
                            h * W.z + lam * vb.z + h2 * X.z};
        W.x += vb.x + dt * X.x;
        W.y += vb.y + dt * X.y;
        W.z += vb.z + dt * X.z;
        a.xb[i] = X;
        a.wb[i] = W;
        // Fb of the previous step: its own force adjoint, its kick's h Wb and its seed
        const double seed = (a.seed_prev && a.group[i]) ? a.seed_prev[0] : 0.0;
        a.fb_next[i] = double3{fprev.x + h * W.x + seed, fprev.y + h * W.y, fprev.z + h * W.z};
    }
    // scale adjoints: the wave's sum by a fixed tree, then the block's waves in order
#pragma unroll
    for (int p = 0; p < kAdjPairsLds; ++p) {
        double v = g[p];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) w.g[p] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < np) {
        double sum = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) sum += wl[q].g[threadIdx.x];
        a.g_pair[(size_t)blockIdx.x * np + threadIdx.x] += sum;
    }
}

// g[p] += v for the pair entry p of a lane's run of equal entries.
__device__ __forceinline__ void add_pair(double (&g)[kAdjPairsLds], int p, double v)
{
#pragma unroll
    for (int q = 0; q < kAdjPairsLds; ++q)
        if (q == p) g[q] += v;
}

// The wave's sums of g[0..7] over its 64 lanes, in 10 exchanges instead of
// 8 x 6: each exchange hands half of the remaining entries to the partner
// lane, so that after the exchanges across lane bits 5, 4 and 3 a lane
// holds entry p = lane >> 3 summed over 8 lanes, and the exchanges across
// bits 2, 1 and 0 complete it.  Lane 8p returns the total of entry p.
static_assert(kAdjPairsLds == 8, "wave_pair_sums halves 8 entries three times");
__device__ __forceinline__ double wave_pair_sums(const double (&g)[kAdjPairsLds], int lane)
{
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
    double h4[4], h2[2];
#pragma unroll
    for (int m = 0; m < 4; ++m)
        h4[m] = (b5 ? g[m + 4] : g[m]) + __shfl_xor(b5 ? g[m] : g[m + 4], 32, 64);
#pragma unroll
    for (int m = 0; m < 2; ++m)
        h2[m] = (b4 ? h4[m + 2] : h4[m]) + __shfl_xor(b4 ? h4[m] : h4[m + 2], 16, 64);
    double v = (b3 ? h2[1] : h2[0]) + __shfl_xor(b3 ? h2[0] : h2[1], 8, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 1, 64);
    return v;
}

// The adjoint step of a single network in the chain layout with at most
// kAdjPairsLds pair entries, in one launch: each lane evaluates the spring
// to its right and the angle centred on it as k_adj_springs / k_adj_angles
// do, the terms reach the neighbouring lanes by shuffles, and each owned
// bead adds them in its incidence-slot order (spring (i-1, i), spring
// (i, i+1); angles at i-1, i, i+1) with the slot path's values, so the bead
// adjoints equal k_adj_springs + k_adj_angles + k_adjoint_step_flat bit for
// bit.  Contacts, walls and the update as in k_adjoint_step_flat; a wave
// owns kChainOwned beads with kHalo lanes of neighbours on each side.
//
// A lane issues the global loads of its bead before its first global store
// (the compiler keeps loads behind stores through other pointers): the
// bead's own values, then its bonds' parameters and scale adjoints at once,
// and in the contact passes a partner's position and force adjoint
// together.  Spring (i, i+1) and the angle at i belong to bead i alone.
//
// Four waves per SIMD, as the LDS allows in the fp32 build (one 8-byte
// register spill); the fp64 build's larger pair table leaves three, so the
// target is not met there and the compiler's note about it is silenced.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wpass-failed"
template <typename T, typename P>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) k_adjoint_step_chain(AdjArgs<T, P> a)
{
    __shared__ PairC prs[kAdjPairsLds];
    __shared__ PairTab<T> prf[kAdjPairsLds];          // the forward's form, for its mask
    __shared__ AdjWaveLds<T> wl[kBlock / 64];
    const int np = a.n_pairs;
    for (int k = threadIdx.x; k < min(np, kAdjPairsLds); k += kBlock) {
        prs[k] = a.pairs[k];
        prf[k] = to_tab<T>(a.pairs[k]);
    }
    AdjWaveLds<T> &w = wl[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int i = (blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6)) * kChainOwned + lane - kHalo;
    const bool in = i >= 0 && i < a.n;
    const bool own = in && lane >= kHalo && lane < 64 - kHalo;
    const int ic = in ? i : (i < 0 ? 0 : a.n - 1);
    const double3 fb = a.fb[ic];
    const P xi = a.ix[ic], yi = a.iy[ic], zi = a.iz[ic];
    const bool alias = a.ix == a.px;                  // images are the positions
    const P pxi = alias ? xi : a.px[ic], pyi = alias ? yi : a.py[ic], pzi = alias ? zi : a.pz[ic];
    const T ri = a.radius[ic];
    const int e0 = a.nb_start[ic], cnt = own ? a.nb_cnt[ic] : 0;
    const int s = (in && a.springs) ? a.ch_spring[i] : -1;
    const int q = (in && a.angles) ? a.ch_angle[i] : -1;
    const unsigned code = (in && a.bond_code) ? a.bond_code[i] : 0u;
    double3 X{0.0, 0.0, 0.0}, W{0.0, 0.0, 0.0};
    T lam = (T)0, h = (T)0, h2 = (T)0;
    bool grp = false;
    if (own) {
        X = a.xb[i];
        W = a.wb[i];
        lam = a.lam[i];
        h = a.hlf[i];
        h2 = a.hlf2[i];
        grp = a.seed_prev && a.group[i];
    }
    // the bonds: spring (i, i+1) and the angle at i
    const bool s_ok = s >= 0 && s < a.n_springs, q_ok = q >= 0 && q < a.n_angles;
    int s_on = 0, s_ty = -1, q_on = 0, q_ty = 0;
    double s_l0 = 0.0, gs = 0.0, ga = 0.0;
    T s_k = (T)0, q_th0 = (T)0, q_k = (T)0;
    if (s_ok) {
        s_on = a.sp_on[s];
        if (a.bond_code) {                            // the sets hold the bond arrays' values exactly
            const SpringSet &ss = a.sp_sets[code & kNoSet];
            s_ty = ss.ty;
            s_k = (T)ss.k;
            s_l0 = ss.l0;
        } else {
            s_ty = a.sp_type[s];
            s_k = a.sp_k[s];
            s_l0 = a.sp_l0[s];
        }
        if (own) gs = a.g_spring[s];
    }
    if (q_ok) {
        q_on = a.an_on[q];
        if (a.bond_code) {
            const AngleSet &as = a.an_sets[(code >> 16) & kNoSet];
            q_ty = as.ty;
            q_th0 = (T)as.th0;
            q_k = (T)as.k;
        } else {
            q_ty = a.an_type[q];
            q_th0 = a.an_th0[q];
            q_k = a.an_k[q];
        }
        if (own) ga = a.g_angle[q];
    }
    // LDS: the pair table, and the lane's bead for the contact passes
    w.ow[lane] = AdjOwner{(double)pxi, (double)pyi, (double)pzi, fb.x, fb.y, fb.z};
    w.orad[lane] = ri;
    __syncthreads();
    // neighbours i + 1 and i - 1 (meaningless in the outermost lanes, whose
    // results no owned lane uses)
    const P xn = lane_next(xi), yn = lane_next(yi), zn = lane_next(zi);
    const P xp = lane_prev(xi), yp = lane_prev(yi), zp = lane_prev(zi);
    const double3 fbn{lane_next(fb.x), lane_next(fb.y), lane_next(fb.z)};
    const double3 fbp{lane_prev(fb.x), lane_prev(fb.y), lane_prev(fb.z)};
    // spring (i, i+1): J Fb on p1 = i (p2 receives its negation)
    double3 jr{0.0, 0.0, 0.0};
    const bool s_act = s_on == 1 && (s_ty == 0 || s_ty == 1);
    if (s_act) {
        const double3 u = sub3(fb, fbn);
        const double3 d{(double)(xi - xn), (double)(yi - yn), (double)(zi - zn)};
        const DV3 F = spring_dual((double)s_k, s_l0, dvec(d, u));
        jr = double3{F.x.d, F.y.d, F.z.d};
        gs += dot_v(u, F);
    }
    // angle (i-1, i, i+1): J Fb on b1 (ja), b3 (jc) and b2 (-ja - jc)
    double3 ja{0.0, 0.0, 0.0}, jc{0.0, 0.0, 0.0};
    if (q_on == 1) {
        const double3 x2{(double)xi, (double)yi, (double)zi};
        const double3 ba{(double)xp - x2.x, (double)yp - x2.y, (double)zp - x2.z};
        const double3 bc{(double)xn - x2.x, (double)yn - x2.y, (double)zn - x2.z};
        const double3 da = sub3(fbp, fb), dc = sub3(fbn, fb);
        DV3 pa, pc;
        angle_dual((double)q_k, (double)q_th0, dvec(ba, da), dvec(bc, dc), pa, pc);
        ja = double3{pa.x.d, pa.y.d, pa.z.d};
        jc = double3{pc.x.d, pc.y.d, pc.z.d};
        if (own) {
            V3<T> fa, fc;                             // the forward's own values
            angle_force(AngleP<T>{1, q_ty, q_th0, q_k}, xp, yp, zp, xi, yi, zi, xn, yn, zn, fa, fc);
            ga += da.x * (double)fa.x + da.y * (double)fa.y + da.z * (double)fa.z +
                  dc.x * (double)fc.x + dc.y * (double)fc.y + dc.z * (double)fc.z;
        }
    }
    if (own && s_act) a.g_spring[s] = gs;
    if (own && q_on == 1) a.g_angle[q] = ga;
    const int hsl = lane_prev(s >= 0 ? 1 : 0);
    const int hql = lane_prev(q >= 0 ? 1 : 0), hqn = lane_next(q >= 0 ? 1 : 0);
    const double3 jl{lane_prev(jr.x), lane_prev(jr.y), lane_prev(jr.z)};
    const double3 jcl{lane_prev(jc.x), lane_prev(jc.y), lane_prev(jc.z)};
    const double3 jan{lane_next(ja.x), lane_next(ja.y), lane_next(ja.z)};
    double3 z{0.0, 0.0, 0.0};
    if (a.springs) {
        if (hsl) {              // bead i is p2 of spring (i-1, i)
            z.x += -jl.x;
            z.y += -jl.y;
            z.z += -jl.z;
        }
        if (s >= 0) {           // p1 of spring (i, i+1)
            z.x += jr.x;
            z.y += jr.y;
            z.z += jr.z;
        }
    }
    if (a.angles) {
        if (hql) {              // b3 of the angle at i-1
            z.x += jcl.x;
            z.y += jcl.y;
            z.z += jcl.z;
        }
        if (q >= 0) {           // b2 of the angle at i
            z.x += -ja.x - jc.x;
            z.y += -ja.y - jc.y;
            z.z += -ja.z - jc.z;
        }
        if (hqn) {              // b1 of the angle at i+1
            z.x += jan.x;
            z.y += jan.y;
            z.z += jan.z;
        }
    }
    // contacts, flattened over the wave as in k_adjoint_step_flat
    const Box &bf = a.force;
    int incl = cnt;
    incl = wave_incl_scan(incl);
    const int off = incl - cnt;
    const int total = __shfl(incl, 63, 64);
    double g[kAdjPairsLds];
#pragma unroll
    for (int p = 0; p < kAdjPairsLds; ++p) g[p] = 0.0;
    int gp = -1;                // the pair entry of the lane's current run of terms, and its sum
    double gacc = 0.0;
    for (int base = 0; base < total; base += kWaveChunk) {           // wave-uniform
        const int n = (total - base) < kWaveChunk ? total - base : kWaveChunk;
        mark_owner(off, cnt, base, w.own);
        for (int r = 0; r * 64 < n; ++r) {
            const int k = r * 64 + lane;
            int L;
            const int e = owner_entry(k, k < n, e0, off - base, w.own, L);
            const AdjOwner o = w.ow[L];
            const P xL = (P)o.x, yL = (P)o.y, zL = (P)o.z;
            const T rL = w.orad[L];
            const double3 fL{o.fx, o.fy, o.fz};
            double3 t{-0.0, -0.0, -0.0};
            if (k < n) {
                const int ent = a.nb_list[e];
                const int j = ent & kIdxMask, pid = ent >> kPidShift;
                const P pxj = a.px[j], pyj = a.py[j], pzj = a.pz[j];
                const double3 fbj = a.fb[j];
                const PairC pr = prs[pid];
                P rx = xL - pxj, ry = yL - pyj, rz = zL - pzj;
                P len = rx * rx + ry * ry + rz * rz;
                if (len > (P)bf.hbs) {           // ResolveContacts' minimum image (kernel.cl:844-870)
                    if (rx * rx > (P)bf.hbs) rx += (rx > (P)0) ? (P)-bf.lx : (P)bf.lx;
                    if (ry * ry > (P)bf.hbs) ry += (ry > (P)0) ? (P)-bf.ly : (P)bf.ly;
                    if (rz * rz > (P)bf.hbs) rz += (rz > (P)0) ? (P)-bf.lz : (P)bf.lz;
                    len = rx * rx + ry * ry + rz * rz;
                }
                V3<T> tf;
                if (contact_force(prf[pid], bf, rx, ry, rz, len, rL, a.radius, j, tf)) {   // the forward's mask
                    const double rj = pr.type == 2 ? (double)a.radius[j] : 0.0;
                    const DV3 F = contact_dual(pr, (double)rL, rj,
                                               dvec(double3{(double)rx, (double)ry, (double)rz}, sub3(fL, fbj)));
                    t = double3{F.x.d, F.y.d, F.z.d};
                    const double gv = dot_v(fL, F);
                    if (pid != gp) {
                        if (gp >= 0) add_pair(g, gp, gacc);
                        gp = pid;
                        gacc = 0.0;
                    }
                    gacc += gv;
                }
            }
            if (k < n) w.term[k] = t;
        }
        wave_sync_lds();
        const int k0 = off > base ? off : base;
        const int k1 = (off + cnt < base + n) ? off + cnt : base + n;
        for (int k = k0; k < k1; ++k) {
            const double3 t = w.term[k - base];
            z.x += t.x;
            z.y += t.y;
            z.z += t.z;
        }
        wave_sync_lds();
    }
    if (own) {
        // walls of a finite box (kernel.cl:252-261): diagonal
        const Box &bd = a.drift;
        if (bd.bc == 1) {
            const double k = -bd.wall_k;
            if (pxi < (P)0 || pxi > (P)bd.lx) z.x += k * fb.x;
            if (pyi < (P)0 || pyi > (P)bd.ly) z.y += k * fb.y;
            if (pzi < (P)0 || pzi > (P)bd.lz) z.z += k * fb.z;
        }
        X.x += z.x;
        X.y += z.y;
        X.z += z.z;
        const double dt = (double)(T)bd.dt, D = (double)(T)bd.damping;
        const double3 vb{D * fb.x, D * fb.y, D * fb.z};
        const double hd = h, lamd = lam, h2d = h2;
        const double3 fprev{hd * W.x + lamd * vb.x + h2d * X.x, hd * W.y + lamd * vb.y + h2d * X.y,
                            hd * W.z + lamd * vb.z + h2d * X.z};
        W.x += vb.x + dt * X.x;
        W.y += vb.y + dt * X.y;
        W.z += vb.z + dt * X.z;
        a.xb[i] = X;
        a.wb[i] = W;
        const double seed = grp ? a.seed_prev[0] : 0.0;
        a.fb_next[i] = double3{fprev.x + hd * W.x + seed, fprev.y + hd * W.y, fprev.z + hd * W.z};
    }
    if (gp >= 0) add_pair(g, gp, gacc);
    {
        const double v = wave_pair_sums(g, lane);
        if ((lane & 7) == 0) w.g[lane >> 3] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < np) {
        double sum = 0.0;
        for (int qq = 0; qq < kBlock / 64; ++qq) sum += wl[qq].g[threadIdx.x];
        a.g_pair[(size_t)blockIdx.x * np + threadIdx.x] += sum;
    }
}

#pragma clang diagnostic pop

// Blocks of the chain-layout adjoint kernel for n beads.
static inline int chain_adj_blocks(int n)
{
    constexpr int per_block = kChainOwned * (kBlock / 64);
    return (n + per_block - 1) / per_block;
}

bool adjoint_one_launch(bool chain, bool batched, int np) { return chain && !batched && np <= kAdjPairsLds; }

int adjoint_pair_rows(int n, int np, bool batched, bool chain)
{
    if (batched || np > kAdjPairsLds) return n;
    return chain ? chain_adj_blocks(n) : (n + kBlock - 1) / kBlock;
}

// Positions (and images, when ix is not null) and bond statuses of one step
// into its record: rec_pos = x, y, z[, ix, iy, iz] (n each), rec_st = spring
// then angle flags (1 where the status is 1: the bond acts in the step).
__global__ void __launch_bounds__(kBlock) k_record_step(const double *px, const double *py, const double *pz,
                                                        const double *ix, const double *iy, const double *iz,
                                                        const int *sp_st, int ns, const int *an_st, int na, int n,
                                                        double *rec_pos, unsigned char *rec_st)
{
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < n) {
        rec_pos[k] = px[k];
        rec_pos[(size_t)n + k] = py[k];
        rec_pos[2 * (size_t)n + k] = pz[k];
        if (ix) {
            rec_pos[3 * (size_t)n + k] = ix[k];
            rec_pos[4 * (size_t)n + k] = iy[k];
            rec_pos[5 * (size_t)n + k] = iz[k];
        }
    }
    if (k < ns) rec_st[k] = sp_st[k] == 1 ? 1 : 0;
    if (k < na) rec_st[ns + k] = an_st[k] == 1 ? 1 : 0;
}

void record_step(const double *const p[3], const double *const im[3], const int *sp_st, int ns, const int *an_st,
                 int na, int n, double *rec_pos, unsigned char *rec_st, hipStream_t s)
{
    const int m = std::max(n, std::max(ns, na));
    if (m > 0)
        k_record_step<<<(m + kBlock - 1) / kBlock, kBlock, 0, s>>>(p[0], p[1], p[2], im ? im[0] : nullptr,
                                                                  im ? im[1] : nullptr, im ? im[2] : nullptr, sp_st,
                                                                  ns, an_st, na, n, rec_pos, rec_st);
}

// Fb of the last step: its seed and h Wb.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_adjoint_seed(double3 *fb, const double3 *wb, const T *hlf,
                                                         const int *group, const double *seed, const int *net, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double h = (double)hlf[i];
    const double3 W = wb[i];
    const double sd = (seed && group[i]) ? seed[net ? net[i] : 0] : 0.0;
    fb[i] = double3{h * W.x + sd, h * W.y, h * W.z};
}

// Block sum in a fixed order (tree over the block's threads).
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double part[kBlock];
    part[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    return part[0];
}

// First pass of group_force: block b sums the entries k = b, b + B, ... of
// its threads' strided share (B = gridDim.x), into part[b].
template <typename T>
__global__ void __launch_bounds__(kBlock) k_group_partial(const T *fx, const int *idx, const int *off, double *part)
{
    const int net = blockIdx.y, k0 = off[net], k1 = off[net + 1];
    double s = 0.0;
    for (int k = k0 + blockIdx.x * kBlock + threadIdx.x; k < k1; k += gridDim.x * kBlock) s += (double)fx[idx[k]];
    const double t = block_sum(s);
    if (threadIdx.x == 0) part[(size_t)net * gridDim.x + blockIdx.x] = t;
}

// Second pass of both reductions: out[c] = sum over b < nb of part[c * nb + b].
__global__ void __launch_bounds__(kBlock) k_sum_parts(const double *part, int nb, double *out)
{
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) s += part[(size_t)blockIdx.x * nb + b];
    const double t = block_sum(s);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// First pass of sum_columns: block (b, c) sums column c over rows b, b + B, ...
__global__ void __launch_bounds__(kBlock) k_column_partial(const double *g, int n, int np, double *part)
{
    const int c = blockIdx.y;
    double s = 0.0;
    for (int r = blockIdx.x * kBlock + threadIdx.x; r < n; r += gridDim.x * kBlock) s += g[(size_t)r * np + c];
    const double t = block_sum(s);
    if (threadIdx.x == 0) part[(size_t)c * gridDim.x + blockIdx.x] = t;
}

static int sum_blocks(int n) { return std::max(1, std::min(kSumBlocks, (n + kBlock - 1) / kBlock)); }

// Block (k, t) of sum_network_columns.
__global__ void __launch_bounds__(kBlock) k_network_columns(const double *g, int cols, const int *bead_off,
                                                            const int *pair_base, double *out)
{
    const int k = blockIdx.x, t = blockIdx.y;
    if (t >= pair_base[k + 1] - pair_base[k]) return;          // block-uniform
    double s = 0.0;
    for (int i = bead_off[k] + threadIdx.x; i < bead_off[k + 1]; i += kBlock) s += g[(size_t)i * cols + t];
    const double v = block_sum(s);
    if (threadIdx.x == 0) out[pair_base[k] + t] = v;
}

void sum_network_columns(const double *g, int cols, const int *bead_off, const int *pair_base, int n_net,
                         double *out, hipStream_t s)
{
    if (n_net > 0 && cols > 0) k_network_columns<<<dim3(n_net, cols), kBlock, 0, s>>>(g, cols, bead_off, pair_base, out);
}

void sum_columns(const double *g, int n, int np, double *part, double *out, hipStream_t s)
{
    if (np <= 0) return;
    const int nb = sum_blocks(n);
    k_column_partial<<<dim3(nb, np), kBlock, 0, s>>>(g, n, np, part);
    k_sum_parts<<<np, kBlock, 0, s>>>(part, nb, out);
}

template <typename T, typename P>
void adjoint_step(const AdjArgs<T, P> &a, hipStream_t s)
{
    if (a.ch_spring && !a.net && a.n_pairs <= kAdjPairsLds) {
        if (a.n > 0) k_adjoint_step_chain<T, P><<<chain_adj_blocks(a.n), kBlock, 0, s>>>(a);
        return;
    }
    if (a.springs > 0) k_adj_springs<T, P><<<(a.springs + kBlock - 1) / kBlock, kBlock, 0, s>>>(a);
    if (a.angles > 0) k_adj_angles<T, P><<<(a.angles + kBlock - 1) / kBlock, kBlock, 0, s>>>(a);
    if (a.n <= 0) return;
    const int nb = (a.n + kBlock - 1) / kBlock;
    if (a.net) k_adjoint_step<T, P, false, true><<<nb, kBlock, 0, s>>>(a);
    else if (a.n_pairs <= kAdjPairsLds) k_adjoint_step_flat<T, P><<<nb, kBlock, 0, s>>>(a);
    else k_adjoint_step<T, P, false, false><<<nb, kBlock, 0, s>>>(a);
}
template <typename T>
void adjoint_seed(double3 *fb, const double3 *wb, const T *hlf, const int *group, const double *seed,
                  const int *net, int n, hipStream_t s)
{
    if (n > 0) k_adjoint_seed<T><<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(fb, wb, hlf, group, seed, net, n);
}
template <typename T>
void group_force(const T *fx, const int *idx, const int *off, int n_net, int max_len, double *part, double *out,
                 hipStream_t s)
{
    if (n_net <= 0) return;
    const int nb = sum_blocks(max_len);
    k_group_partial<T><<<dim3(nb, n_net), kBlock, 0, s>>>(fx, idx, off, part);
    k_sum_parts<<<n_net, kBlock, 0, s>>>(part, nb, out);
}

template void adjoint_step<double, double>(const AdjArgs<double, double> &, hipStream_t);
template void adjoint_step<float, double>(const AdjArgs<float, double> &, hipStream_t);
template void adjoint_seed<double>(double3 *, const double3 *, const double *, const int *, const double *,
                                   const int *, int, hipStream_t);
template void adjoint_seed<float>(double3 *, const double3 *, const float *, const int *, const double *, const int *,
                                  int, hipStream_t);
template void group_force<double>(const double *, const int *, const int *, int, int, double *, double *,
                                  hipStream_t);
template void group_force<float>(const float *, const int *, const int *, int, int, double *, double *, hipStream_t);

}  // namespace fnk
