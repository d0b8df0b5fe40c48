// kernels.h - device code of the FibreNet step for gfx950.
//
// Layout: structure of arrays, one array per coordinate.  T is the type of
// velocities, forces and force arithmetic; P the type of positions (fp64 in
// both builds, so that the drift of a pulled bead at x~1000 survives).
//
// Determinism: each spring and each angle is evaluated once per step by one
// thread, which writes its term (and owns any status change); every bead then
// gathers the terms acting on it (incident springs, incident angles, its
// two-sided contact list) in a fixed order and owns its own write; no float
// atomics.  The two beads of a contact evaluate it with identical operands, so
// they see bit-identical, opposite contributions.
//
// Operation order and rounding of each term follow kernel.cl so that the fp64
// build reproduces the reference's arithmetic; mad() is an explicit fma() and
// nothing else is contracted (the library is built with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>

namespace fnk {

constexpr int kBlock = 256;

// Grid / box parameters taken from one bound copy of SimulationSettings.  In
// an ensemble every network has its own, with the offsets of its cells in the
// ensemble's cell numbering and of its entries in the concatenated pair table.
struct Box {
    int bc;                 // 0 periodic, 1 finite
    int cx, cy, cz, cxy;    // grid cells
    int types;              // number of bead types
    int cell_base, pair_base;
    double lx, ly, lz;      // box size
    double sx, sy, sz;      // grid step
    double wall_k, hbs, cutoff_sq, damping, gx, gy, gz, dt;
};

// One bead in cell order: everything the neighbour search reads per
// candidate, in three 16-byte loads (and written by the placement as one
// 48-byte record): position, tag, loc, type, the slot's cell (c, or ~c for a
// bead outside the grid, placed by its clamped position) and the bead index.
struct alignas(16) SortedBead {
    double x, y, z;
    int tag, loc, type, cell, bead, pad;
};
static_assert(sizeof(SortedBead) == 48, "three 16-byte loads per sorted record");

// Bead records copied into cell order at a rebuild (the neighbour search
// reads them contiguously).
struct Sorted {
    SortedBead *bead;
};

// One pair-table entry (Pair.h:5-16): comparisons use the fp64 values.
struct PairC {
    int type;
    double c1, c2, c3, c4, c5, c6;
};

// The same entry in the form the fp32 contact terms use: the float
// conversions and products that every contact would otherwise repeat, and c3
// rounded up to a float, so that r < c3f holds exactly when (double)r < c3
// for a float r.  The two-step harmonic's bound stays fp64 (it is compared
// with the fp64 squared distance).
struct alignas(16) PairF {
    int type;
    float lj_s2, lj_e48, c3f;   // ((float)c1)^2, 48 (float)c2, the least float >= c3
    float c4f, c1f, c2f, c5f;
    float c6f, pad;
    double c4;
};

// The least float not below v (v rounded to nearest, then one float up when
// that fell below v).
__device__ __forceinline__ float float_at_least(double v)
{
    float f = (float)v;
    if ((double)f < v) {
        const int bits = __float_as_int(f);
        f = f == 0.0f ? __int_as_float(1) : __int_as_float(f > 0.0f ? bits + 1 : bits - 1);
    }
    return f;
}

__device__ __forceinline__ PairF pair_f(const PairC &p)
{
    PairF q;
    q.type = p.type;
    q.c1f = (float)p.c1;
    q.c2f = (float)p.c2;
    q.c4f = (float)p.c4;
    q.c5f = (float)p.c5;
    q.c6f = (float)p.c6;
    q.lj_s2 = q.c1f * q.c1f;
    q.lj_e48 = 48.0f * q.c2f;
    q.c3f = float_at_least(p.c3);
    q.pad = 0.0f;
    q.c4 = p.c4;
    return q;
}

// The pair-table entry a build's bead kernels keep in LDS: PairF for the
// fp32 build, PairC for the fp64 one.
template <typename T>
struct PairTabOf { using type = PairC; };
template <>
struct PairTabOf<float> { using type = PairF; };
template <typename T>
using PairTab = typename PairTabOf<T>::type;
template <typename T>
__device__ __forceinline__ PairTab<T> to_tab(const PairC &p)
{
    if constexpr (sizeof(T) == sizeof(float)) return pair_f(p);
    else return p;
}

template <typename T>
struct V3 { T x, y, z; };

// A distinct parameter set of the chain layout's springs (type, rest length,
// E, failure stress, stiffness as the engine's T widened to fp64) and angles
// (type, rest angle and stiffness as T widened).
struct SpringSet {
    int ty;
    double l0, c1, c2, k;
};
struct AngleSet {
    int ty;
    double th0, k;
};
// A bead's chain bonds in one word: bits 0-14 the set of spring (i, i+1),
// bit 15 that spring's status == 1; bits 16-30 the set of the angle centred
// on i, bit 31 its status == 1.  kNoSet in a set field: no such bond.
constexpr unsigned kNoSet = 0x7FFF;
constexpr unsigned kSpringOn = 1u << 15;
constexpr unsigned kAngleOn = 1u << 31;

// A distinct set of a bead's mass coefficients and radius, as the engine's T
// holds them: lambda*dt/m, (dt/2)/m, (dt^2/2)/m and r.
template <typename T>
struct MassSet {
    T lam, hlf, hlf2, r;
};
constexpr int kMassSets = 256;   // a one-byte code per bead

template <typename T, typename P>
struct Args {
    int n;                                  // beads
    P *px, *py, *pz;                        // positions (current)
    P *ix, *iy, *iz;                        // images (current)
    P *qx, *qy, *qz;                        // positions (next, fused look-ahead drift)
    P *jx, *jy, *jz;                        // images (next)
    T *vx, *vy, *vz;                        // predicted velocity
    T *wx, *wy, *wz;                        // velocity_old
    T *fx, *fy, *fz;                        // force in   (granular: f, fused: f of the previous step)
    T *gx, *gy, *gz;                        // force out  (granular: f_old, fused: f of this step)
    const T *lam, *hlf, *hlf2;              // lambda*dt/m, (dt/2)/m, (dt^2/2)/m
    const int *type, *tag, *loc;
    const T *radius;
    // springs
    const int *sp_start, *sp_p1, *sp_p2, *sp_type;
    int *sp_status, *sp_break;
    const double *sp_l0, *sp_c1, *sp_c2;     // fp64: the failure test is decided as the reference decides it
    const T *sp_k;
    int n_springs;
    // angles
    const int *an_start, *an_b1, *an_b2, *an_b3, *an_s1, *an_s2, *an_type;
    int *an_status;
    const T *an_th0, *an_k;
    int n_angles;
    // Bond terms of the current step, stored at the beads' incidence slots
    // (bead i's spring terms are sp_term[sp_start[i] .. sp_start[i+1]), its
    // angle terms an_term[an_start[i] .. an_start[i+1])).  sp_slot[2s+0/1]:
    // slots of spring s at p1 / p2; an_slot[3q+0/1/2]: slots of angle q at
    // b1 / b3 / b2.
    V3<T> *sp_term, *an_term;
    const int *sp_slot, *an_slot;
    // Chain layout (fibres as consecutive bead runs): ch_spring[i] is the
    // spring (p1 i, p2 i+1), ch_angle[i] the angle (b1 i-1, b2 i, b3 i+1),
    // -1 where there is none.
    const int *ch_spring, *ch_angle;
    // The chain layout's bond parameters by bead, when the bonds have at
    // most kNoSet distinct parameter sets of each kind (null otherwise): one
    // word per bead (kNoSet above), whose status bits copy sp_status /
    // an_status (refreshed at the start of every chunk, written with them on
    // a change).
    unsigned *bond_code;
    const SpringSet *sp_sets;
    const AngleSet *an_sets;
    // Each bead's mass coefficients and radius as a one-byte code into at
    // most kMassSets distinct sets (null: read the per-bead arrays).
    const unsigned char *ms_code;
    const MassSet<T> *ms_sets;
    int n_ms;
    // contacts: bead i's partners are nb_list[nb_start[i] .. nb_start[i] + nb_cnt[i])
    const int *nb_start, *nb_cnt, *nb_list;
    const PairC *pairs;
    int n_pairs;
    // settings of AdvancePosition (integrator, walls) and ResolveContacts (minimum image)
    Box drift, force;
    // ensemble: the network of every bead and the Box of every network
    const int *net;
    const Box *boxes;
    int *step;                              // device step counter
    int step_off;
};

// Chain layout: a wave owns kChainOwned consecutive beads, with kHalo lanes
// on either side holding their neighbours.
constexpr int kHalo = 2;
constexpr int kChainOwned = 64 - 2 * kHalo;

// Bits of a contact-list entry: partner index | pair-table index << kPidShift.
constexpr int kPidShift = 26;
constexpr int kIdxMask = (1 << kPidShift) - 1;

__device__ __forceinline__ int pair_index(int t1, int t2, int n)
{
    int lo = t1 < t2 ? t1 : t2;
    int hi = t1 < t2 ? t2 : t1;
    return n * lo + hi - (lo * (lo + 1)) / 2;
}

// native_sqrt(double) as clang-based OpenCL resolves it: the float overload.
__device__ __forceinline__ float native_sqrt(double x) { return __builtin_sqrtf((float)x); }

// The hardware's fp32 square root and reciprocal (v_sqrt_f32, v_rcp_f32, 1 ulp)
// for the fp32 build's contact and angle arithmetic: a few instructions where
// the correctly rounded forms take fifteen.  The fp64 build and the spring
// break decision keep the correctly rounded operations.
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ double tsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float tsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double tacos(double x) { return acos(x); }
__device__ __forceinline__ float tacos(float x) { return acosf(x); }
__device__ __forceinline__ double tfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float tfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// AdvancePosition's drift for one bead (kernel.cl:200-223 and 241-250).
template <typename T, typename P>
__device__ __forceinline__ void drift(const Box &b, T fx, T fy, T fz, T wx, T wy, T wz, T lam, T h2, P &px,
                                      P &py, P &pz, P &ix, P &iy, P &iz, V3<T> &vp)
{
    vp.x = tfma(fx, lam, wx);
    vp.y = tfma(fy, lam, wy);
    vp.z = tfma(fz, lam, wz);
    const T dt = (T)b.dt;
    T ax = wx * dt, ay = wy * dt, az = wz * dt;
    ax = tfma(fx, h2, ax);
    ay = tfma(fy, h2, ay);
    az = tfma(fz, h2, az);
    px += (P)ax;
    py += (P)ay;
    pz += (P)az;
    ix += (P)ax;
    iy += (P)ay;
    iz += (P)az;
    if (b.bc == 0) {
        if (pz < (P)0) pz = (P)b.lz;
        if (py < (P)0) py = (P)b.ly;
        if (px < (P)0) px = (P)b.lx;
        if (pz > (P)b.lz) pz = (P)0;
        if (py > (P)b.ly) py = (P)0;
        if (px > (P)b.lx) px = (P)0;
    }
}

// Force at the start of the step: damping * v_pred + gravity, plus the wall
// penalty of a finite box (kernel.cl:231-238, 252-261).
template <typename T, typename P>
__device__ __forceinline__ V3<T> base_force(const Box &b, const V3<T> &vp, P px, P py, P pz)
{
    V3<T> f;
    const T d = (T)b.damping;
    f.x = d * vp.x;
    f.y = d * vp.y;
    f.z = d * vp.z;
    f.x += (T)b.gx;
    f.y += (T)b.gy;
    f.z += (T)b.gz;
    if (b.bc == 1) {
        const double k = b.wall_k;
        if (pz < (P)0) f.z += (T)((double)pz * k * -1.0);
        if (py < (P)0) f.y += (T)((double)py * k * -1.0);
        if (px < (P)0) f.x += (T)((double)px * k * -1.0);
        if (pz > (P)b.lz) f.z += (T)(((double)pz - b.lz) * k * -1.0);
        if (py > (P)b.ly) f.y += (T)(((double)py - b.ly) * k * -1.0);
        if (px > (P)b.lx) f.x += (T)(((double)px - b.lx) * k * -1.0);
    }
    return f;
}

// Parameters of one spring, loaded together (one round trip to memory).
template <typename T>
struct SpringP {
    int ty, st;
    double l0, c1, c2;
    T k;
};

template <typename T, typename P>
__device__ __forceinline__ SpringP<T> load_spring(const Args<T, P> &a, int s)
{
    return SpringP<T>{a.sp_type[s], a.sp_status[s], a.sp_l0[s], a.sp_c1[s], a.sp_c2[s], a.sp_k[s]};
}

// Spring with parameters p and image difference d = image(p1) - image(p2)
// (kernel.cl:371-417).  Returns false when the spring is inactive or breaks
// now; otherwise F is the force on p1 (p2 receives -F).  `broke` reports a
// break decided by this evaluation.
template <typename T, typename P>
__device__ __forceinline__ bool spring_force(const SpringP<T> &p, P dx, P dy, P dz, V3<T> &F, bool &broke)
{
    broke = false;
    const int ty = p.ty;
    if (!(ty == 0 || ty == 1) || p.st != 1) return false;
    const P d2 = dx * dx + dy * dy + dz * dz;
    // length, extension and the HR_FL stress in fp64 from the float-rounded
    // length (native_sqrt): the break decision is bit-exact in both builds
    double l = (double)native_sqrt((double)d2);
    const double factor = p.l0 - l;
    if (l != 0.0) l = 1.0 / l;
    if (ty == 1) {
        const double stress = -p.c1 * factor * l;
        if (fabs(stress) > p.c2) {
            broke = true;
            return false;
        }
    }
    const T il = (T)l;
    const T ux = (T)dx * il, uy = (T)dy * il, uz = (T)dz * il;
    const T fk = (T)factor * p.k;
    F.x = ux * fk;
    F.y = uy * fk;
    F.z = uz * fk;
    return true;
}

// Spring s evaluated from the current images of its beads.
template <typename T, typename P>
__device__ __forceinline__ bool spring_eval(const Args<T, P> &a, int s, V3<T> &F, bool &broke)
{
    const int p1 = a.sp_p1[s], p2 = a.sp_p2[s];
    return spring_force(load_spring(a, s), a.ix[p1] - a.ix[p2], a.iy[p1] - a.iy[p2], a.iz[p1] - a.iz[p2], F, broke);
}

// Perpendicular force directions of an angle bond scaled to fa, fc
// (kernel.cl:321-348) in the arithmetic type R.
template <typename R, typename T>
__device__ __forceinline__ void angle_dirs(R bax, R bay, R baz, R bcx, R bcy, R bcz, R fa, R fc, V3<T> &pa_out,
                                           V3<T> &pc_out)
{
    const R cbx = bcx * (R)-1, cby = bcy * (R)-1, cbz = bcz * (R)-1;
    const R cx = bay * bcz - baz * bcy;
    const R cy = baz * bcx - bax * bcz;
    const R cz = bax * bcy - bay * bcx;
    R pax = bay * cz - baz * cy, pay = baz * cx - bax * cz, paz = bax * cy - bay * cx;
    R ln = tsqrt(pax * pax + pay * pay + paz * paz);
    if (ln != (R)0) ln = (R)1 / ln;
    ln *= fa;
    pa_out.x = (T)(pax * ln);
    pa_out.y = (T)(pay * ln);
    pa_out.z = (T)(paz * ln);
    R pcx = cby * cz - cbz * cy, pcy = cbz * cx - cbx * cz, pcz = cbx * cy - cby * cx;
    ln = tsqrt(pcx * pcx + pcy * pcy + pcz * pcz);
    if (ln != (R)0) ln = (R)1 / ln;
    ln *= fc;
    pc_out.x = (T)(pcx * ln);
    pc_out.y = (T)(pcy * ln);
    pc_out.z = (T)(pcz * ln);
}

// Angle bond forces on b1 (pa) and b3 (pc); b2 receives -pa-pc (kernel.cl:289-360).
// The angle is acos of the cosine, ill-conditioned at the straight fibre
// (d theta = d cos / sin theta), so cosine, lengths and acos are fp64 in both
// builds.  The directions use T, except where the reference clamps a cosine
// rounded beyond -1 or 1 (kernel.cl:311-312): on such an exactly straight
// segment the force (|F| from the clamped 0.999999) points along the fp64
// cross product of the bond vectors, which is rounding noise that only the
// fp64 evaluation reproduces.
// acos(d) for the cosines of a straight or folded triple, |d| >= 0.999998
// (including the clamped +-0.999999): with s = sqrt((1 - |d|) / 2),
// acos(|d|) = 2 asin(s) = 2 s (1 + s^2/6 + 3 s^4/40 + ...), and the first
// omitted term is below 5e-20 relative for s^2 <= 1e-6.  Other cosines (and
// NaN) take the library acos.
__device__ __forceinline__ double acos_near_unit(double d)
{
    const double ad = fabs(d);
    if (!(ad >= 0.999998)) return acos(d);
    const double s = __builtin_sqrt((1.0 - ad) * 0.5);
    const double s2 = s * s;
    const double a = 2.0 * s * (1.0 + s2 * (1.0 / 6.0 + s2 * 0.075));
    return d < 0.0 ? (3.141592653589793 - a) + 1.2246467991473532e-16 : a;
}

// fp32 factor f / |v| for a vector with |v|^2 = n2 given in fp64 (0 for a
// zero vector, as the reference's normalisation leaves it).
__device__ __forceinline__ float unit_scale(double n2, float f)
{
    const float n = (float)n2;
    if (n >= 1.17549435e-38f) return f * __builtin_amdgcn_rsqf(n);
    return n2 > 0.0 ? (float)((double)f / __builtin_sqrt(n2)) : 0.0f;
}

// Parameters of one angle, loaded together.
template <typename T>
struct AngleP {
    int st, ty;
    T th0, k;
};

template <typename T, typename P>
__device__ __forceinline__ AngleP<T> load_angle(const Args<T, P> &a, int q)
{
    return AngleP<T>{a.an_status[q], a.an_type[q], a.an_th0[q], a.an_k[q]};
}

// Angle with parameters g from the images of b1 (x1..), b2 (tx..) and b3 (x3..).
template <typename T, typename P>
__device__ __forceinline__ void angle_force(const AngleP<T> &g, P x1, P y1, P z1, P tx, P ty, P tz, P x3, P y3,
                                            P z3, V3<T> &pa_out, V3<T> &pc_out)
{
    const double bax = (double)(x1 - tx), bay = (double)(y1 - ty), baz = (double)(z1 - tz);
    const double bcx = (double)(x3 - tx), bcy = (double)(y3 - ty), bcz = (double)(z3 - tz);
    if constexpr (sizeof(T) < sizeof(double)) {
        // fp32 build: the angle as atan2(|ba x bc|, ba . bc), well conditioned
        // at any angle, and the perpendicular directions from the same cross
        // product (|ba x (ba x bc)| = |ba| |ba x bc|) ...
        const T fx = (T)bax, fy = (T)bay, fz = (T)baz, gx = (T)bcx, gy = (T)bcy, gz = (T)bcz;
        const T cx = fy * gz - fz * gy, cy = fz * gx - fx * gz, cz = fx * gy - fy * gx;
        const T c2 = cx * cx + cy * cy + cz * cz;
        const T la2 = fx * fx + fy * fy + fz * fz, lc2 = gx * gx + gy * gy + gz * gz;
        bool clamped = false, neg = false;
        if (!(c2 > (T)1e-10 * la2 * lc2)) {
            // ... except on a straight or folded triple whose fp64 cosine the
            // reference clamps (rounded beyond -1 or 1, kernel.cl:311-312)
            const double l_ba = tsqrt(bax * bax + bay * bay + baz * baz);
            const double l_bc = tsqrt(bcx * bcx + bcy * bcy + bcz * bcz);
            const double dotp = (bax * bcx + bay * bcy + baz * bcz) / (l_ba * l_bc);
            clamped = dotp < -1.0 || dotp > 1.0;
            neg = dotp < 0.0;
        }
        if (!clamped) {
            if (!(c2 > (T)1e-30)) {               // no fp32 direction: no force
                pa_out = pc_out = V3<T>{(T)0, (T)0, (T)0};
                return;
            }
            const float cn = hw_sqrt((float)c2);
            const T ang = atan2f(cn, fx * gx + fy * gy + fz * gz);
            const T q = (T)-2 * g.k * (ang - g.th0) * hw_rcp(cn);
            const T sa = q * hw_rcp((float)la2), sc = q * hw_rcp((float)lc2);
            pa_out.x = (fy * cz - fz * cy) * sa;
            pa_out.y = (fz * cx - fx * cz) * sa;
            pa_out.z = (fx * cy - fy * cx) * sa;
            pc_out.x = (gz * cy - gy * cz) * sc;
            pc_out.y = (gx * cz - gz * cx) * sc;
            pc_out.z = (gy * cx - gx * cy) * sc;
            return;
        }
        // the clamped cosine: the angle acos(-+0.999999) (its series), the
        // perpendicular directions from the fp64 cross products (their
        // rounding noise on an exactly straight segment is the direction)
        // normalised in fp32
        const double l_ba = tsqrt(bax * bax + bay * bay + baz * baz);
        const double l_bc = tsqrt(bcx * bcx + bcy * bcy + bcz * bcz);
        const T mag = (T)(-2.0 * (double)g.k * (acos_near_unit(neg ? -0.999999 : 0.999999) - (double)g.th0));
        const double cbx = -bcx, cby = -bcy, cbz = -bcz;
        const double dx = bay * bcz - baz * bcy, dy = baz * bcx - bax * bcz, dz = bax * bcy - bay * bcx;
        const double pax = bay * dz - baz * dy, pay = baz * dx - bax * dz, paz = bax * dy - bay * dx;
        const double pcx = cby * dz - cbz * dy, pcy = cbz * dx - cbx * dz, pcz = cbx * dy - cby * dx;
        const T sa = unit_scale(pax * pax + pay * pay + paz * paz, mag * hw_rcp((float)l_ba));
        const T sc = unit_scale(pcx * pcx + pcy * pcy + pcz * pcz, mag * hw_rcp((float)l_bc));
        pa_out = V3<T>{(T)pax * sa, (T)pay * sa, (T)paz * sa};
        pc_out = V3<T>{(T)pcx * sc, (T)pcy * sc, (T)pcz * sc};
        return;
    }
    const double l_ba = tsqrt(bax * bax + bay * bay + baz * baz);
    const double l_bc = tsqrt(bcx * bcx + bcy * bcy + bcz * bcz);
    double dotp = (bax * bcx + bay * bcy + baz * bcz) / (l_ba * l_bc);
    const bool clamped = dotp < -1.0 || dotp > 1.0;
    if (dotp < -1.0) dotp = -0.999999;
    if (dotp > 1.0) dotp = 0.999999;
    const double ang = tacos(dotp);
    const double mag = -2.0 * (double)g.k * (ang - (double)g.th0);
    const double fa = mag / l_ba;
    const double fc = mag / l_bc;
    if (sizeof(T) == sizeof(double) || clamped)
        angle_dirs<double, T>(bax, bay, baz, bcx, bcy, bcz, fa, fc, pa_out, pc_out);
    else
        angle_dirs<T, T>((T)bax, (T)bay, (T)baz, (T)bcx, (T)bcy, (T)bcz, (T)fa, (T)fc, pa_out, pc_out);
}

// Angle q evaluated from the current images of its beads.
template <typename T, typename P>
__device__ __forceinline__ void angle_eval(const Args<T, P> &a, int q, V3<T> &pa_out, V3<T> &pc_out)
{
    const int b1 = a.an_b1[q], b2 = a.an_b2[q], b3 = a.an_b3[q];
    angle_force(load_angle(a, q), a.ix[b1], a.iy[b1], a.iz[b1], a.ix[b2], a.iy[b2], a.iz[b2], a.ix[b3], a.iy[b3],
                a.iz[b3], pa_out, pc_out);
}

// Adds the bond terms stored at slots [e0, e1) in slot order (the order in
// which the reference's bond kernels reach bead i: kernel.cl:350-360,
// 426-436).
template <typename T>
__device__ __forceinline__ void add_terms(const V3<T> *term, int e0, int e1, V3<T> &f)
{
    for (int e = e0; e < e1; ++e) {
        const V3<T> t = term[e];
        f.x += t.x;
        f.y += t.y;
        f.z += t.z;
    }
}

// fp32 build: force on bead 1 of a listed pair from r = x1 - x_j at its
// minimum image (len = |r|^2), bead 1's radius r1 and the radii table; false
// when the pair exerts no force this step.  The terms of contact_force below
// through the hardware square root and reciprocal, one reciprocal of r^2
// shared by the Lennard-Jones ratios and 1/r = r / r^2.
template <typename P>
__device__ __forceinline__ bool contact_force_f(const PairF &pr, const Box &b, P rx, P ry, P rz, P len, float r1,
                                                const float *radius, int j, V3<float> &t)
{
    float coef, r, inv;
    if (pr.type == 1) {                           // Lennard-Jones, kernel.cl:875-919
        if (!((double)len < b.cutoff_sq)) return false;
        const float r2 = (float)len;
        r = hw_sqrt(r2);
        if (!(r < pr.c3f)) return false;
        const float rr = hw_rcp(r2);
        const float fr2 = pr.lj_s2 * rr;
        const float fr6 = fr2 * fr2 * fr2;
        coef = pr.lj_e48 * fr6 * (fr6 - 0.5f) * rr;
        coef -= pr.c4f;
        inv = r * rr;
    } else if (pr.type == 2) {                    // penalty, kernel.cl:922-957
        r = hw_sqrt((float)len);
        coef = (r1 + radius[j]) - r;
        if (!(coef > 0.0f)) return false;
        coef *= pr.c1f;
        inv = hw_rcp(r);
    } else if (pr.type == 3) {                    // two-step harmonic, kernel.cl:960-999
        if (!((double)len < pr.c4)) return false;
        r = hw_sqrt((float)len);
        coef = (pr.c2f - r) * pr.c1f;
        coef += (pr.c6f - r) * pr.c5f;
        inv = hw_rcp(r);
    } else {
        return false;
    }
    if (r == 0.0f) inv = 0.0f;
    t.x = coef * ((float)rx * inv);
    t.y = coef * ((float)ry * inv);
    t.z = coef * ((float)rz * inv);
    return true;
}

// Force on bead 1 of a listed pair from r = x1 - x_j at its minimum image
// (len = |r|^2), bead 1's radius r1 and the radii table; false when the pair
// exerts no force this step.
template <typename T, typename P>
__device__ __forceinline__ bool contact_force(const PairC &pr, const Box &b, P rx, P ry, P rz, P len, T r1,
                                              const T *radius, int j, V3<T> &t)
{
    if constexpr (sizeof(T) == sizeof(float)) {
        return contact_force_f(pair_f(pr), b, rx, ry, rz, len, r1, radius, j, t);
    } else {
        T coef;
        T r;
        if (pr.type == 1) {                       // Lennard-Jones, kernel.cl:875-919
            if (!((double)len < b.cutoff_sq)) return false;
            r = (T)native_sqrt((double)len);
            if (!((double)r < pr.c3)) return false;
            const T r2 = (T)len;
            const T c1 = (T)pr.c1;
            const T sigma2 = c1 * c1;
            const T fr2 = sigma2 / r2;
            const T fr6 = fr2 * fr2 * fr2;
            coef = (T)48 * (T)pr.c2 * fr6 * (fr6 - (T)0.5) / r2;
            coef -= (T)pr.c4;
        } else if (pr.type == 2) {                // penalty, kernel.cl:922-957
            r = (T)native_sqrt((double)len);
            coef = (r1 + radius[j]) - r;
            if (!(coef > (T)0)) return false;
            coef *= (T)pr.c1;
        } else if (pr.type == 3) {                // two-step harmonic, kernel.cl:960-999
            if (!((double)len < pr.c4)) return false;
            r = (T)native_sqrt((double)len);
            coef = ((T)pr.c2 - r) * (T)pr.c1;
            coef += ((T)pr.c6 - r) * (T)pr.c5;
        } else {
            return false;
        }
        const T inv = (r != (T)0) ? (T)1 / r : (T)0;
        t.x = coef * ((T)rx * inv);
        t.y = coef * ((T)ry * inv);
        t.z = coef * ((T)rz * inv);
        return true;
    }
}
template <typename P>
__device__ __forceinline__ bool contact_force(const PairF &pr, const Box &b, P rx, P ry, P rz, P len, float r1,
                                              const float *radius, int j, V3<float> &t)
{
    return contact_force_f(pr, b, rx, ry, rz, len, r1, radius, j, t);
}

// Force on bead 1 (position x1.., radius r1) from the contact-list entry
// `ent` (kernel.cl:833-1000 seen from either bead of a pair: the reaction on
// the partner is the exact negation of the owner's term, so each side
// evaluates rij = p_1 - p_j).  Returns false when the pair exerts no force
// this step.
template <typename T, typename P>
__device__ __forceinline__ bool contact_term(const Args<T, P> &a, const Box &b, P x1, P y1, P z1, T r1, int ent,
                                             const PairC *table, V3<T> &t)
{
    const int j = ent & kIdxMask;
    const PairC pr = table[ent >> kPidShift];
    P rx = x1 - a.px[j], ry = y1 - a.py[j], rz = z1 - a.pz[j];
    P len = rx * rx + ry * ry + rz * rz;
    if (len > (P)b.hbs) {
        if (rx * rx > (P)b.hbs) rx += (rx > (P)0) ? (P)-b.lx : (P)b.lx;
        if (ry * ry > (P)b.hbs) ry += (ry > (P)0) ? (P)-b.ly : (P)b.ly;
        if (rz * rz > (P)b.hbs) rz += (rz > (P)0) ? (P)-b.lz : (P)b.lz;
        len = rx * rx + ry * ry + rz * rz;
    }
    return contact_force(pr, b, rx, ry, rz, len, r1, a.radius, j, t);
}

// Adds the contact forces of bead i's two-sided list in list order.
template <typename T, typename P>
__device__ __forceinline__ void gather_contacts(const Args<T, P> &a, int i, V3<T> &f, const PairC *table)
{
    const int e0 = a.nb_start[i], e1 = e0 + a.nb_cnt[i];
    if (e0 == e1) return;
    const P x1 = a.px[i], y1 = a.py[i], z1 = a.pz[i];
    const T r1 = a.radius[i];
    for (int e = e0; e < e1; ++e) {
        V3<T> t;
        if (!contact_term(a, a.force, x1, y1, z1, r1, a.nb_list[e], table, t)) continue;
        f.x += t.x;
        f.y += t.y;
        f.z += t.z;
    }
}

// Contact-list entries evaluated per pass by one wave, and the passes of 64.
constexpr int kWaveChunk = 256;
constexpr int kChunkRounds = kWaveChunk / 64;

// The bead of one lane as the wave's contact passes read it: image-free
// position, radius and network.
template <typename T>
struct alignas(16) ContactOwner {
    double x, y, z;
    T r;
    int net;
};

// One contact term in LDS, padded to 16 (fp32) or 32 (fp64) bytes so that a
// lane reads or writes it with one LDS instruction.
template <typename T>
struct alignas(4 * sizeof(T)) TermSlot {
    T x, y, z, pad;
};

// A wave's LDS for its contacts: the terms of one pass, the lane that owns
// each entry of it, and every lane's bead.
template <typename T>
struct WaveContactLds {
    TermSlot<T> term[kWaveChunk];
    unsigned char own[kWaveChunk];
    ContactOwner<T> ow[64];
};

// Values of the neighbouring lane through the DPP whole-wave shift by one
// (a VALU move, no LDS round trip): lane_next(v) is v of lane + 1,
// lane_prev(v) v of lane - 1; the last (first) lane keeps its own value, as
// __shfl_down(v, 1) / __shfl_up(v, 1) do.  All 64 lanes must be active.
constexpr int kDppNext = 0x130;   // wave_shl:1
constexpr int kDppPrev = 0x138;   // wave_shr:1
__device__ __forceinline__ int lane_next(int v) { return __builtin_amdgcn_update_dpp(v, v, kDppNext, 0xf, 0xf, false); }
__device__ __forceinline__ int lane_prev(int v) { return __builtin_amdgcn_update_dpp(v, v, kDppPrev, 0xf, 0xf, false); }
__device__ __forceinline__ float lane_next(float v) { return __int_as_float(lane_next(__float_as_int(v))); }
__device__ __forceinline__ float lane_prev(float v) { return __int_as_float(lane_prev(__float_as_int(v))); }
__device__ __forceinline__ double lane_next(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = lane_next((int)(unsigned)(b & 0xffffffffLL)), hi = lane_next((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_prev(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = lane_prev((int)(unsigned)(b & 0xffffffffLL)), hi = lane_prev((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Inclusive prefix sum of v over the 64 lanes: Hillis-Steele steps inside
// each row of 16 lanes (DPP row shifts by 1, 2, 4, 8), then the sum of the
// rows before (row broadcasts of lanes 15 and 31).  All 64 lanes must be
// active.
__device__ __forceinline__ int wave_incl_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);   // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);   // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);   // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);   // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);   // row_bcast:15 into rows 1 and 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);   // row_bcast:31 into rows 2 and 3
    return v;
}

// Inclusive maximum of v over the 64 lanes (v >= 0), by the same DPP steps
// as wave_incl_scan.  All 64 lanes must be active.
__device__ __forceinline__ int wave_incl_max(int v)
{
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));   // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));   // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));   // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));   // row_shr:8
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));   // row_bcast:15 into rows 1 and 3
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));   // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The first pass's list entries of a wave, loaded ahead of their use.
struct ContactAhead {
    int off, total;            // this lane's first entry in the wave's entries; their number
    int ent[kChunkRounds];     // entry k = r * 64 + lane of the first pass (0 past the end)
};

// Entry k of the pass starting at `base`: its owner lane (from the LDS map)
// and its position in the contact list.
__device__ __forceinline__ int owner_entry(int k, bool ok, int e0, int off, const unsigned char *own, int &L)
{
    L = ok ? own[k] : 0;
    return __shfl(e0, L, 64) + (k - __shfl(off, L, 64));
}

// Writes to `own` (kWaveChunk bytes, 4-byte aligned, in LDS) the lane that
// owns each entry of [base, base + kWaveChunk), where lane L owns entries
// [off_L, off_L + cnt_L) and the offsets are the exclusive scan of the counts
// (every lane of the wave calls it).  Each lane with entries marks the slot
// its entries start at with L + 1, and a running maximum over the slots (a
// lane's four slots, then the lanes by DPP) carries each mark to the slots
// after it; the owner of slot base itself, when its entries start before
// base, is the last lane that starts at or before base.  A fixed number of
// steps, whatever the counts.  Ends with the map visible to the wave.
__device__ __forceinline__ void mark_owner(int off, int cnt, int base, unsigned char *own)
{
    const int lane = threadIdx.x & 63;
    unsigned *words = reinterpret_cast<unsigned *>(own);
    const unsigned long long before = __ballot(cnt > 0 && off <= base);
    const int first = before ? 64 - __clzll(before) : 0;       // its lane + 1 (0: none)
    words[lane] = 0u;
    wave_sync_lds();
    const int s = off - base;
    if (cnt > 0 && s > 0 && s < kWaveChunk) own[s] = (unsigned char)(lane + 1);
    wave_sync_lds();
    const unsigned q = words[lane];                              // slots 4 lane .. 4 lane + 3
    const int m0 = (int)(q & 255u);
    const int m1 = max(m0, (int)((q >> 8) & 255u));
    const int m2 = max(m1, (int)((q >> 16) & 255u));
    const int m3 = max(m2, (int)(q >> 24));
    int carry = lane_prev(wave_incl_max(m3));                    // the lanes before
    if (lane == 0) carry = 0;
    carry = max(carry, first);
    const unsigned o0 = (unsigned)(max(carry, m0) - 1) & 255u, o1 = (unsigned)(max(carry, m1) - 1) & 255u;
    const unsigned o2 = (unsigned)(max(carry, m2) - 1) & 255u, o3 = (unsigned)(max(carry, m3) - 1) & 255u;
    words[lane] = o0 | (o1 << 8) | (o2 << 16) | (o3 << 24);
    wave_sync_lds();
}

// Starts a wave's contacts (every lane of the wave calls it; lane L's bead
// has list entries [e0, e0 + cnt), cnt 0 for lanes without a bead): the
// counts are scanned over the wave, the first pass's owner map goes to LDS
// and its list entries are loaded, so that their latency passes while the
// caller does other work.
template <typename T, typename P>
__device__ __forceinline__ ContactAhead contacts_ahead(const Args<T, P> &a, int e0, int cnt, WaveContactLds<T> &w)
{
    const int lane = threadIdx.x & 63;
    int incl = cnt;                                   // inclusive scan of the counts over the wave
    incl = wave_incl_scan(incl);
    ContactAhead c;
    c.off = incl - cnt;
    c.total = __shfl(incl, 63, 64);
    mark_owner(c.off, cnt, 0, w.own);
    const int n0 = c.total < kWaveChunk ? c.total : kWaveChunk;
#pragma unroll
    for (int r = 0; r < kChunkRounds; ++r) {
        const int k = r * 64 + lane;
        int L;
        const int e = owner_entry(k, k < n0, e0, c.off, w.own, L);
        c.ent[r] = k < n0 ? a.nb_list[e] : 0;
    }
    return c;
}

// The contact term of entry `ent` for bead 1 at x1.. with radius r1 and the
// partner at xj..: ResolveContacts' minimum image (kernel.cl:844-870), then
// contact_force.  The same operations as contact_term.
template <typename T, typename P, typename E>
__device__ __forceinline__ bool contact_term_at(const Args<T, P> &a, const Box &b, P x1, P y1, P z1, T r1, int ent,
                                                P xj, P yj, P zj, const E *table, V3<T> &t)
{
    const int j = ent & kIdxMask;
    const E pr = table[ent >> kPidShift];
    P rx = x1 - xj, ry = y1 - yj, rz = z1 - zj;
    P len = rx * rx + ry * ry + rz * rz;
    if (len > (P)b.hbs) {
        if (rx * rx > (P)b.hbs) rx += (rx > (P)0) ? (P)-b.lx : (P)b.lx;
        if (ry * ry > (P)b.hbs) ry += (ry > (P)0) ? (P)-b.ly : (P)b.ly;
        if (rz * rz > (P)b.hbs) rz += (rz > (P)0) ? (P)-b.lz : (P)b.lz;
        len = rx * rx + ry * ry + rz * rz;
    }
    return contact_force(pr, b, rx, ry, rz, len, r1, a.radius, j, t);
}

// contact_term_at for a pair table whose addressed entries are all one
// Lennard-Jones entry pr (fp32 build): the same operations in the same
// order, with the entry's tests folded into one mask instead of branches.
template <typename P>
__device__ __forceinline__ V3<float> lj_uniform_term(const PairF &pr, const Box &b, P x1, P y1, P z1, P xj, P yj,
                                                     P zj, bool valid)
{
    P rx = x1 - xj, ry = y1 - yj, rz = z1 - zj;
    P len = rx * rx + ry * ry + rz * rz;
    if (len > (P)b.hbs) {
        if (rx * rx > (P)b.hbs) rx += (rx > (P)0) ? (P)-b.lx : (P)b.lx;
        if (ry * ry > (P)b.hbs) ry += (ry > (P)0) ? (P)-b.ly : (P)b.ly;
        if (rz * rz > (P)b.hbs) rz += (rz > (P)0) ? (P)-b.lz : (P)b.lz;
        len = rx * rx + ry * ry + rz * rz;
    }
    const float r2 = (float)len;
    const float r = hw_sqrt(r2);
    const bool on = valid && (double)len < b.cutoff_sq && r < pr.c3f;
    const float rr = hw_rcp(r2);
    const float fr2 = pr.lj_s2 * rr;
    const float fr6 = fr2 * fr2 * fr2;
    float coef = pr.lj_e48 * fr6 * (fr6 - 0.5f) * rr;
    coef -= pr.c4f;
    float inv = r * rr;
    if (r == 0.0f) inv = 0.0f;
    V3<float> t;
    t.x = on ? coef * ((float)rx * inv) : -0.0f;
    t.y = on ? coef * ((float)ry * inv) : -0.0f;
    t.z = on ? coef * ((float)rz * inv) : -0.0f;
    return t;
}

// Whether the entries [0, na) of a pair table in LDS (na <= 64) are all
// one Lennard-Jones entry as the fp32 contact term reads it; wave-uniform,
// every lane of the wave calls it.  The fp64 build has no such path.
__device__ __forceinline__ bool lj_uniform_wave(const PairF *tab, int na)
{
    const int lane = threadIdx.x & 63;
    const PairF p0 = tab[0];
    bool ok = true;
    if (lane < na) {
        const PairF p = tab[lane];
        ok = p.type == p0.type && p.lj_s2 == p0.lj_s2 && p.lj_e48 == p0.lj_e48 && p.c3f == p0.c3f && p.c4f == p0.c4f;
    }
    return na >= 1 && p0.type == 1 && __ballot(!ok) == 0;
}
__device__ __forceinline__ bool lj_uniform_wave(const PairC *, int) { return false; }

// One pass of a wave's contacts: entries [base, base + n) with their list
// entries ent[r] at k = base + r * 64 + lane, owner lanes in w.own and the
// owners' beads in w.ow; the terms go to w.term (an entry without force
// stores -0.0), and each lane adds its own entries in list order.  Every
// partner position of the pass is loaded before the terms are evaluated.
template <typename T, typename P, bool ENS, typename E>
__device__ __forceinline__ void contacts_pass(const Args<T, P> &a, const int (&ent)[kChunkRounds], int base, int n,
                                              int off, int cnt, V3<T> &f, const E *table, WaveContactLds<T> &w,
                                              bool lj_uni)
{
    const int lane = threadIdx.x & 63;
    P xj[kChunkRounds], yj[kChunkRounds], zj[kChunkRounds];
#pragma unroll
    for (int r = 0; r < kChunkRounds; ++r) {
        xj[r] = yj[r] = zj[r] = (P)0;
        if (r * 64 + lane < n) {
            const int j = ent[r] & kIdxMask;
            xj[r] = a.px[j];
            yj[r] = a.py[j];
            zj[r] = a.pz[j];
        }
    }
    bool done = false;
    if constexpr (!ENS && sizeof(T) == sizeof(float)) {
        if (lj_uni) {                                  // block-uniform
            const PairF pr = table[0];
#pragma unroll
            for (int r = 0; r < kChunkRounds; ++r) {
                if (r * 64 >= n) break;               // wave-uniform
                const int k = r * 64 + lane;
                const int L = k < n ? w.own[k] : 0;
                const ContactOwner<T> o = w.ow[L];
                const V3<float> t = lj_uniform_term(pr, a.force, (P)o.x, (P)o.y, (P)o.z, xj[r], yj[r], zj[r], k < n);
                w.term[k] = TermSlot<T>{t.x, t.y, t.z, (T)0};
            }
            done = true;
        }
    }
#pragma unroll
    for (int r = 0; r < kChunkRounds; ++r) {
        if (done || r * 64 >= n) break;               // block- / wave-uniform
        const int k = r * 64 + lane;
        const int L = k < n ? w.own[k] : 0;
        const ContactOwner<T> o = w.ow[L];
        V3<T> t{(T)-0.0, (T)-0.0, (T)-0.0};
        if (ENS) {
            const Box &bk = a.boxes[o.net];
            if (k < n && !contact_term_at(a, bk, (P)o.x, (P)o.y, (P)o.z, o.r, ent[r], xj[r], yj[r], zj[r],
                                          a.pairs + bk.pair_base, t))
                t = V3<T>{(T)-0.0, (T)-0.0, (T)-0.0};
        } else {
            if (k < n && !contact_term_at(a, a.force, (P)o.x, (P)o.y, (P)o.z, o.r, ent[r], xj[r], yj[r], zj[r], table,
                                          t))
                t = V3<T>{(T)-0.0, (T)-0.0, (T)-0.0};
        }
        w.term[k] = TermSlot<T>{t.x, t.y, t.z, (T)0};
    }
    wave_sync_lds();
    // Each lane adds its entries [k0, k1) in list order, four per iteration:
    // the four reads are issued together, and a slot past k1 adds -0.0,
    // which leaves every sum (including +0.0) unchanged bit for bit.
    const int k0 = off > base ? off : base;
    const int k1 = (off + cnt < base + n) ? off + cnt : base + n;
    for (int k = k0; k < k1; k += 4) {
        TermSlot<T> t[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) t[u] = w.term[min(k + u - base, kWaveChunk - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool on = k + u < k1;
            f.x += on ? t[u].x : (T)-0.0;
            f.y += on ? t[u].y : (T)-0.0;
            f.z += on ? t[u].z : (T)-0.0;
        }
    }
    wave_sync_lds();
}

// Finishes a wave's contacts started by contacts_ahead (every lane calls
// it): the first pass with the entries loaded ahead, then further passes of
// kWaveChunk entries for a wave with more.  The wave's lists are laid end to
// end and each lane evaluates every 64th entry, so a wave costs (its total
// entries / 64) rounds instead of its longest list; each bead then adds its
// own terms in list order, which equals the per-bead loop (gather_contacts)
// bit for bit.  ENS: the lane's bead belongs to network `netw` (its Box and
// pair table); otherwise a.force and `table`.  lj_uni (block-uniform): every
// entry the lists address is table[0], a Lennard-Jones entry.
template <typename T, typename P, bool ENS, typename E>
__device__ __forceinline__ void contacts_finish(const Args<T, P> &a, const ContactAhead &c, int e0, int cnt, T r1,
                                                P x1, P y1, P z1, V3<T> &f, const E *table, WaveContactLds<T> &w,
                                                int netw = 0, bool lj_uni = false)
{
    static_assert(sizeof(P) == sizeof(double), "positions are fp64 in both builds");
    const int lane = threadIdx.x & 63;
    if (c.total == 0) return;                         // wave-uniform
    w.ow[lane] = ContactOwner<T>{(double)x1, (double)y1, (double)z1, r1, netw};
    wave_sync_lds();
    contacts_pass<T, P, ENS>(a, c.ent, 0, c.total < kWaveChunk ? c.total : kWaveChunk, c.off, cnt, f, table, w,
                             lj_uni);
    for (int base = kWaveChunk; base < c.total; base += kWaveChunk) {
        const int n = (c.total - base) < kWaveChunk ? c.total - base : kWaveChunk;
        mark_owner(c.off, cnt, base, w.own);
        int ent[kChunkRounds];
#pragma unroll
        for (int r = 0; r < kChunkRounds; ++r) {
            const int k = r * 64 + lane;
            int L;
            const int e = owner_entry(k, k < n, e0, c.off - base, w.own, L);
            ent[r] = k < n ? a.nb_list[e] : 0;
        }
        contacts_pass<T, P, ENS>(a, ent, base, n, c.off, cnt, f, table, w, lj_uni);
    }
}

}  // namespace fnk
