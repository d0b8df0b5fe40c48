// kernels.h - device code of the FibreNet step for gfx950.
//
// Layout: structure of arrays, one array per coordinate.  T is the type of
// velocities, forces and force arithmetic; P the type of positions (fp64 in
// both builds, so that the drift of a pulled bead at x~1000 survives).
//
// Determinism: every bead gathers the terms acting on it (incident springs,
// incident angles, its two-sided contact list) in a fixed order and owns its
// own write; no float atomics.  A bond shared by two or three beads is
// evaluated by each of them with identical operands and operation order, so
// the partners see bit-identical, opposite contributions.
//
// Operation order and rounding of each term follow kernel.cl so that the fp64
// build reproduces the reference's arithmetic; mad() is an explicit fma() and
// nothing else is contracted (the library is built with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>

namespace fnk {

constexpr int kBlock = 256;

// Grid / box parameters taken from one bound copy of SimulationSettings.
struct Box {
    int bc;                 // 0 periodic, 1 finite
    int cx, cy, cz, cxy;    // grid cells
    int types;              // number of bead types
    double lx, ly, lz;      // box size
    double sx, sy, sz;      // grid step
    double wall_k, hbs, cutoff_sq, damping, gx, gy, gz, dt;
};

// Bead data copied into cell order at a rebuild (the neighbour search reads
// these contiguously) and the inverse permutation bead -> slot.
struct Sorted {
    double *x, *y, *z;
    int *tag, *loc, *type, *slot_of;
};

// One pair-table entry (Pair.h:5-16): comparisons use the fp64 values.
struct PairC {
    int type;
    double c1, c2, c3, c4, c5, c6;
};

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
    const int *sp_start, *sp_ent, *sp_p1, *sp_p2, *sp_type;
    int *sp_status, *sp_break;
    const double *sp_l0, *sp_c1, *sp_c2;     // fp64: the failure test is decided as the reference decides it
    const T *sp_k;
    int n_springs;
    // angles
    const int *an_start, *an_ent, *an_b1, *an_b2, *an_b3, *an_s1, *an_s2, *an_type;
    int *an_status;
    const T *an_th0, *an_k;
    int n_angles;
    // index of the owner entry of the same bond: spring p2 entry -> p1 entry,
    // angle b1/b3 entry -> b2 entry (bead-incidence lists)
    const int *sp_twin, *an_owner;
    // contacts
    const int *nb_start, *nb_list;
    const PairC *pairs;
    int n_pairs;
    // settings of AdvancePosition (integrator, walls) and ResolveContacts (minimum image)
    Box drift, force;
    int *step;                              // device step counter
    int step_off;
};

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

__device__ __forceinline__ double tsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float tsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double tacos(double x) { return acos(x); }
__device__ __forceinline__ float tacos(float x) { return acosf(x); }
__device__ __forceinline__ double tfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float tfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T>
struct V3 { T x, y, z; };

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

// Spring s evaluated from the current images (kernel.cl:371-417).  Returns
// false when the spring is inactive or breaks now; otherwise F is the force on
// p1 (p2 receives -F).  `broke` reports a break decided by this evaluation.
template <typename T, typename P>
__device__ __forceinline__ bool spring_eval(const Args<T, P> &a, int s, V3<T> &F, bool &broke)
{
    broke = false;
    const int ty = a.sp_type[s];
    if (!(ty == 0 || ty == 1) || a.sp_status[s] != 1) return false;
    const int p1 = a.sp_p1[s], p2 = a.sp_p2[s];
    P dx = a.ix[p1], dy = a.iy[p1], dz = a.iz[p1];
    dx -= a.ix[p2];
    dy -= a.iy[p2];
    dz -= a.iz[p2];
    const P d2 = dx * dx + dy * dy + dz * dz;
    // length, extension and the HR_FL stress in fp64 from the float-rounded
    // length (native_sqrt): the break decision is bit-exact in both builds
    double l = (double)native_sqrt((double)d2);
    const double factor = a.sp_l0[s] - l;
    if (l != 0.0) l = 1.0 / l;
    if (ty == 1) {
        const double stress = -a.sp_c1[s] * factor * l;
        if (fabs(stress) > a.sp_c2[s]) {
            broke = true;
            return false;
        }
    }
    const T il = (T)l;
    const T ux = (T)dx * il, uy = (T)dy * il, uz = (T)dz * il;
    const T fk = (T)factor * a.sp_k[s];
    F.x = ux * fk;
    F.y = uy * fk;
    F.z = uz * fk;
    return true;
}

// Status of spring s after this step's spring pass (for the HR_FL angle rule),
// decided from the same evaluation every bead performs.
template <typename T, typename P>
__device__ __forceinline__ int spring_status_now(const Args<T, P> &a, int s)
{
    if (a.sp_status[s] == 0) return 0;
    V3<T> F;
    bool broke;
    spring_eval(a, s, F, broke);
    return broke ? 0 : a.sp_status[s];
}

// Angle bond forces on b1 (pa) and b3 (pc); b2 receives -pa-pc (kernel.cl:289-360).
// Evaluated in fp64 in both builds and rounded to T at the end: the angle is
// acos of the cosine, ill-conditioned at the straight fibre (d theta =
// d cos / sin theta), and on an exactly straight segment the reference's
// clamp of a cosine rounded below -1 to -0.999999 produces a finite force
// whose direction comes from the fp64 cross product of the bond vectors.
template <typename T, typename P>
__device__ __forceinline__ void angle_eval(const Args<T, P> &a, int q, V3<T> &pa_out, V3<T> &pc_out)
{
    const int b1 = a.an_b1[q], b2 = a.an_b2[q], b3 = a.an_b3[q];
    const P tx = a.ix[b2], ty = a.iy[b2], tz = a.iz[b2];
    const double bax = (double)(a.ix[b1] - tx), bay = (double)(a.iy[b1] - ty), baz = (double)(a.iz[b1] - tz);
    const double bcx = (double)(a.ix[b3] - tx), bcy = (double)(a.iy[b3] - ty), bcz = (double)(a.iz[b3] - tz);
    const double cbx = bcx * -1.0, cby = bcy * -1.0, cbz = bcz * -1.0;
    const double l_ba = tsqrt(bax * bax + bay * bay + baz * baz);
    const double l_bc = tsqrt(bcx * bcx + bcy * bcy + bcz * bcz);
    double dotp = (bax * bcx + bay * bcy + baz * bcz) / (l_ba * l_bc);
    if (dotp < -1.0) dotp = -0.999999;
    if (dotp > 1.0) dotp = 0.999999;
    const double ang = tacos(dotp);
    const double mag = -2.0 * (double)a.an_k[q] * (ang - (double)a.an_th0[q]);
    const double fa = mag / l_ba;
    const double fc = mag / l_bc;
    const double cx = bay * bcz - baz * bcy;
    const double cy = baz * bcx - bax * bcz;
    const double cz = bax * bcy - bay * bcx;
    double pax = bay * cz - baz * cy, pay = baz * cx - bax * cz, paz = bax * cy - bay * cx;
    double ln = tsqrt(pax * pax + pay * pay + paz * paz);
    if (ln != 0.0) ln = 1.0 / ln;
    ln *= fa;
    pa_out.x = (T)(pax * ln);
    pa_out.y = (T)(pay * ln);
    pa_out.z = (T)(paz * ln);
    double pcx = cby * cz - cbz * cy, pcy = cbz * cx - cbx * cz, pcz = cbx * cy - cby * cx;
    ln = tsqrt(pcx * pcx + pcy * pcy + pcz * pcz);
    if (ln != 0.0) ln = 1.0 / ln;
    ln *= fc;
    pc_out.x = (T)(pcx * ln);
    pc_out.y = (T)(pcy * ln);
    pc_out.z = (T)(pcz * ln);
}

// One evaluated bond term as the beads of the bond consume it.
template <typename T>
struct SpTerm { T x, y, z; int on; };
template <typename T>
struct AnTerm { T ax, ay, az, cx, cy, cz; int on; };

// Spring s evaluated by its owner (the p1 bead), which also records a break.
template <typename T, typename P>
__device__ __forceinline__ SpTerm<T> spring_owner(const Args<T, P> &a, int s)
{
    V3<T> F;
    bool broke;
    SpTerm<T> t;
    t.on = spring_eval(a, s, F, broke) ? 1 : 0;
    t.x = F.x; t.y = F.y; t.z = F.z;
    if (broke) {
        a.sp_status[s] = 0;
        a.sp_break[s] = *a.step + a.step_off;
    }
    return t;
}

// Spring s evaluated by its p2 bead (no status write).
template <typename T, typename P>
__device__ __forceinline__ SpTerm<T> spring_partner(const Args<T, P> &a, int s)
{
    V3<T> F;
    bool broke;
    SpTerm<T> t;
    t.on = spring_eval(a, s, F, broke) ? 1 : 0;
    t.x = F.x; t.y = F.y; t.z = F.z;
    return t;
}

template <typename T>
__device__ __forceinline__ void add_spring(V3<T> &f, const SpTerm<T> &t, bool p2)
{
    if (!t.on) return;
    if (p2) {
        f.x += t.x * (T)-1;
        f.y += t.y * (T)-1;
        f.z += t.z * (T)-1;
    } else {
        f.x += t.x;
        f.y += t.y;
        f.z += t.z;
    }
}

// Adds the springs incident to bead i to f (kernel.cl:419-437 seen from bead i).
template <typename T, typename P>
__device__ __forceinline__ void gather_springs(const Args<T, P> &a, int i, V3<T> &f)
{
    const int e0 = a.sp_start[i], e1 = a.sp_start[i + 1];
    for (int e = e0; e < e1; ++e) {
        const int ent = a.sp_ent[e];
        const int s = ent >> 1;
        if (s >= a.n_springs) continue;
        add_spring(f, (ent & 1) ? spring_partner(a, s) : spring_owner(a, s), ent & 1);
    }
}

// Angle q with this step's HR_FL rule (kernel.cl:274-286): an HR_FL angle dies
// with either spring.  `recheck` re-derives this step's spring statuses (fused
// path, where the spring pass runs in the same launch); the owner (b2) records
// the death.
template <typename T, typename P, bool recheck>
__device__ __forceinline__ AnTerm<T> angle_term(const Args<T, P> &a, int q, bool owner)
{
    AnTerm<T> t;
    int st = a.an_status[q];
    if (a.an_type[q] == 1 && st == 1) {
        const int s1 = a.an_s1[q], s2 = a.an_s2[q];
        const int st1 = recheck ? spring_status_now(a, s1) : a.sp_status[s1];
        const int st2 = recheck ? spring_status_now(a, s2) : a.sp_status[s2];
        if (st1 == 0 || st2 == 0) {
            st = 0;
            if (owner) a.an_status[q] = 0;
        }
    }
    t.on = st == 1 ? 1 : 0;
    if (t.on) {
        V3<T> pa, pc;
        angle_eval(a, q, pa, pc);
        t.ax = pa.x; t.ay = pa.y; t.az = pa.z;
        t.cx = pc.x; t.cy = pc.y; t.cz = pc.z;
    }
    return t;
}

template <typename T>
__device__ __forceinline__ void add_angle(V3<T> &f, const AnTerm<T> &t, int role)
{
    if (!t.on) return;
    if (role == 0) {
        f.x += t.ax;
        f.y += t.ay;
        f.z += t.az;
    } else if (role == 2) {
        f.x += t.cx;
        f.y += t.cy;
        f.z += t.cz;
    } else {
        f.x += -t.ax - t.cx;
        f.y += -t.ay - t.cy;
        f.z += -t.az - t.cz;
    }
}

// Adds the angle bonds incident to bead i.
template <typename T, typename P, bool recheck>
__device__ __forceinline__ void gather_angles(const Args<T, P> &a, int i, V3<T> &f)
{
    const int e0 = a.an_start[i], e1 = a.an_start[i + 1];
    for (int e = e0; e < e1; ++e) {
        const int ent = a.an_ent[e];
        const int q = ent >> 2;
        const int role = ent & 3;
        if (q >= a.n_angles) continue;
        add_angle(f, angle_term<T, P, recheck>(a, q, role == 1), role);
    }
}

// Adds the contact forces of bead i's two-sided list (kernel.cl:833-1000 seen
// from either bead of a pair: the reaction on the partner is the exact
// negation of the owner's term, so each side evaluates rij = p_i - p_j).
template <typename T, typename P>
__device__ __forceinline__ void gather_contacts(const Args<T, P> &a, int i, V3<T> &f, const PairC *table)
{
    const int e0 = a.nb_start[i], e1 = a.nb_start[i + 1];
    if (e0 == e1) return;
    const P x1 = a.px[i], y1 = a.py[i], z1 = a.pz[i];
    const T r1 = a.radius[i];
    const Box &b = a.force;
    for (int e = e0; e < e1; ++e) {
        const int ent = a.nb_list[e];
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
        T coef;
        T r;
        if (pr.type == 1) {                           // Lennard-Jones, kernel.cl:875-919
            if (!((double)len < b.cutoff_sq)) continue;
            r = (T)native_sqrt((double)len);
            if (!((double)r < pr.c3)) continue;
            const T r2 = (T)len;
            const T c1 = (T)pr.c1;
            const T sigma2 = c1 * c1;
            const T fr2 = sigma2 / r2;
            const T fr6 = fr2 * fr2 * fr2;
            coef = (T)48 * (T)pr.c2 * fr6 * (fr6 - (T)0.5) / r2;
            coef -= (T)pr.c4;
        } else if (pr.type == 2) {                    // penalty, kernel.cl:922-957
            r = (T)native_sqrt((double)len);
            coef = (r1 + a.radius[j]) - r;
            if (!(coef > (T)0)) continue;
            coef *= (T)pr.c1;
        } else if (pr.type == 3) {                    // two-step harmonic, kernel.cl:960-999
            if (!((double)len < pr.c4)) continue;
            r = (T)native_sqrt((double)len);
            coef = ((T)pr.c2 - r) * (T)pr.c1;
            coef += ((T)pr.c6 - r) * (T)pr.c5;
        } else {
            continue;
        }
        const T inv = (r != (T)0) ? (T)1 / r : (T)0;
        f.x += coef * ((T)rx * inv);
        f.y += coef * ((T)ry * inv);
        f.z += coef * ((T)rz * inv);
    }
}

}  // namespace fnk
