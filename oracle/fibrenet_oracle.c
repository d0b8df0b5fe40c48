/*
 * fibrenet_oracle.c - CPU restatement of FibreNet's particle-dynamics step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP engine is
 * checked against and the scalar CPU baseline bench.py times.  Nothing in the
 * product (fibrenet_amd/, include/) links, loads or calls it.
 *
 * It restates, in plain C and double precision, the ten kernels of the
 * reference's kernel.cl plus the host-side grid geometry and the step cadence
 * of SimulationManager::RunSimulationLoop.  Every function cites the lines it
 * follows.  Execution is sequential in work-item order, which is one valid
 * OpenCL execution of the reference (no kernel uses barriers or local
 * memory); it makes the CAS-atomic accumulation order (kernel.cl:174-193) and
 * the atomic_add slot order of FillGrid/BuildContactList deterministic.
 *
 * Arithmetic choices the reference leaves to the OpenCL implementation, fixed
 * here (and mirrored by the engine):
 *   - mad(a,b,c) (kernel.cl:205-215,459-461) is a fused multiply-add, fma().
 *     No other expression is contracted (built with -ffp-contract=off).
 *   - native_sqrt(double) (kernel.cl:384,880,924,964) binds the float overload
 *     under clang-based OpenCL compilers: the squared length is rounded to
 *     float, square-rooted in float and widened back to double.
 *   - acos (kernel.cl:314) and pow (kernel.cl:650,766) are the C library's.
 *
 * Parity pins (tests/test_oracle.py): closed forms for the integrator, springs,
 * spring failure, angles, lj/hs/st contacts, cell hashing, periodic reset and
 * wall penalty; Newton's-third-law and momentum invariants; a brute-force
 * O(N^2) contact set over the same cell restriction; the rebuild cadence of
 * or_run_loop (a pair is listed exactly at the first rebuild after it comes
 * within the cutoff, the counter restarting with every call); the chunks and
 * resets of or_minimize (closed-form displacement under constant gravity);
 * README fixtures.
 *
 * Parity unpinned: the bit-level results of a particular OpenCL
 * implementation - its acos and pow ulps, its native_sqrt binding, its
 * atomic summation order and bucket slot order.  No reference test or
 * fixture fixes them.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fibrenet_oracle.h"

/* Work-item execution.  The sequential build runs work-items in order and
 * adds plainly.  The timing twin (built with -DOR_MT -fopenmp as
 * liboracle_mt.so) runs each kernel's work-items on all host threads in
 * chunks of 64 (one OpenCL work-group of the reference's local size,
 * OpenclManager.cpp:219-285), adds shared forces with the compare-and-swap
 * loop of kernel.cl:174-193 and takes list slots with atomic increments
 * (kernel.cl:497-503, 632-637), as a CPU OpenCL device executes kernel.cl. */
#ifdef OR_MT
#define OR_PARALLEL_FOR _Pragma("omp parallel for schedule(dynamic, 64)")
static void or_atomic_add(double *p, double v)
{
    unsigned long long *u = (unsigned long long *)p;
    unsigned long long old = __atomic_load_n(u, __ATOMIC_RELAXED), nw;
    do {
        double d;
        memcpy(&d, &old, sizeof d);
        d += v;
        memcpy(&nw, &d, sizeof nw);
    } while (!__atomic_compare_exchange_n(u, &old, nw, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED));
}
#define OR_ADD(p, v) or_atomic_add((p), (v))
#define OR_FETCH_INC(p) __atomic_fetch_add((p), 1, __ATOMIC_RELAXED)
#else
#define OR_PARALLEL_FOR
#define OR_ADD(p, v) (*(p) += (v))
#define OR_FETCH_INC(p) ((*(p))++)
#endif

/* kernel.cl:9 GETADRESS: symmetric upper-triangular index of a type pair. */
static int or_pair_index(int t1, int t2, int n)
{
    int lo = t1 < t2 ? t1 : t2;
    int hi = t1 < t2 ? t2 : t1;
    return n * lo + hi - (lo * (lo + 1)) / 2;
}

/* native_sqrt(double) as resolved by clang-based OpenCL (float overload). */
static double or_native_sqrt(double x) { return (double)sqrtf((float)x); }

/* kernel.cl:195-265 AdvancePosition. */
void or_advance_position(or_bead *beads, const or_info *info, const or_settings *s, int n)
{
    OR_PARALLEL_FOR
    for (int i = 0; i < n; ++i) {
        or_bead b = beads[i];
        double lam = info[i].lambda_dt_mass_inv;
        double h2 = info[i].half_dt2_mass_inv;
        or_vec3 inc = b.velocity_old;
        b.velocity.x = fma(b.force.x, lam, inc.x);
        b.velocity.y = fma(b.force.y, lam, inc.y);
        b.velocity.z = fma(b.force.z, lam, inc.z);
        inc.x *= s->dt;
        inc.y *= s->dt;
        inc.z *= s->dt;
        inc.x = fma(b.force.x, h2, inc.x);
        inc.y = fma(b.force.y, h2, inc.y);
        inc.z = fma(b.force.z, h2, inc.z);
        b.position.x += inc.x;
        b.position.y += inc.y;
        b.position.z += inc.z;
        b.image_position.x += inc.x;
        b.image_position.y += inc.y;
        b.image_position.z += inc.z;
        b.force_old = b.force;
        b.force.x = s->damping * b.velocity.x;
        b.force.y = s->damping * b.velocity.y;
        b.force.z = s->damping * b.velocity.z;
        b.force.x += s->gravity.x;
        b.force.y += s->gravity.y;
        b.force.z += s->gravity.z;
        if (s->bc_type == 0) {      /* kernel.cl:241-250: reset to the opposite face */
            if (b.position.z < 0.0) b.position.z = s->size.z;
            if (b.position.y < 0.0) b.position.y = s->size.y;
            if (b.position.x < 0.0) b.position.x = s->size.x;
            if (b.position.z > s->size.z) b.position.z = 0.0;
            if (b.position.y > s->size.y) b.position.y = 0.0;
            if (b.position.x > s->size.x) b.position.x = 0.0;
        }
        if (s->bc_type == 1) {      /* kernel.cl:252-261: wall penalty */
            if (b.position.z < 0.0) b.force.z += b.position.z * s->wall_k * -1.0;
            if (b.position.y < 0.0) b.force.y += b.position.y * s->wall_k * -1.0;
            if (b.position.x < 0.0) b.force.x += b.position.x * s->wall_k * -1.0;
            if (b.position.z > s->size.z) b.force.z += (b.position.z - s->size.z) * s->wall_k * -1.0;
            if (b.position.y > s->size.y) b.force.y += (b.position.y - s->size.y) * s->wall_k * -1.0;
            if (b.position.x > s->size.x) b.force.x += (b.position.x - s->size.x) * s->wall_k * -1.0;
        }
        beads[i] = b;
    }
}

/* kernel.cl:366-441 ResolveSprings.  break_step (may be NULL) records the step
 * index at which a spring's status went to 0. */
void or_resolve_springs(or_bead *beads, or_spring *springs, int n_springs, int step, int *break_step)
{
    OR_PARALLEL_FOR
    for (int s = 0; s < n_springs; ++s) {
        or_spring sp = springs[s];
        if (!(sp.type == 0 || sp.type == 1) || sp.status != 1) continue;
        or_vec3 ds = beads[sp.p1].image_position;
        or_vec3 t = beads[sp.p2].image_position;
        ds.x -= t.x;
        ds.y -= t.y;
        ds.z -= t.z;
        double l = or_native_sqrt(ds.x * ds.x + ds.y * ds.y + ds.z * ds.z);
        double factor = sp.rest_length - l;
        if (l != 0.0) l = 1.0 / l;
        ds.x *= l;
        ds.y *= l;
        ds.z *= l;
        if (sp.type == 1) {
            double stress = -sp.c1 * factor * l;
            if (fabs(stress) > sp.c2) {
                sp.status = 0;
                springs[s].status = 0;
                if (break_step) break_step[s] = step;
            }
        }
        if (sp.status == 1) {
            factor *= sp.stiffness;
            ds.x *= factor;
            ds.y *= factor;
            ds.z *= factor;
            OR_ADD(&beads[sp.p1].force.x, ds.x);
            OR_ADD(&beads[sp.p1].force.y, ds.y);
            OR_ADD(&beads[sp.p1].force.z, ds.z);
            ds.x *= -1.0;
            ds.y *= -1.0;
            ds.z *= -1.0;
            OR_ADD(&beads[sp.p2].force.x, ds.x);
            OR_ADD(&beads[sp.p2].force.y, ds.y);
            OR_ADD(&beads[sp.p2].force.z, ds.z);
        }
    }
}

/* kernel.cl:268-363 ResolveAngles. */
void or_resolve_angles(or_angle *angles, const or_spring *springs, or_bead *beads, int n_angles)
{
    OR_PARALLEL_FOR
    for (int a = 0; a < n_angles; ++a) {
        or_angle ab = angles[a];
        if (ab.type == 1) {
            if (springs[ab.s1].status == 0) { ab.status = 0; angles[a].status = 0; }
            if (springs[ab.s2].status == 0) { ab.status = 0; angles[a].status = 0; }
        }
        if (ab.status != 1) continue;
        or_vec3 t = beads[ab.b2].image_position;
        or_vec3 ba = beads[ab.b1].image_position;
        ba.x -= t.x; ba.y -= t.y; ba.z -= t.z;
        or_vec3 bc = beads[ab.b3].image_position;
        bc.x -= t.x; bc.y -= t.y; bc.z -= t.z;
        or_vec3 cb = bc;
        cb.x *= -1.0; cb.y *= -1.0; cb.z *= -1.0;
        double l_ba = sqrt(ba.x * ba.x + ba.y * ba.y + ba.z * ba.z);
        double l_bc = sqrt(bc.x * bc.x + bc.y * bc.y + bc.z * bc.z);
        double dotp = (ba.x * bc.x + ba.y * bc.y + ba.z * bc.z) / (l_ba * l_bc);
        if (dotp < -1.0) dotp = -0.999999;
        if (dotp > 1.0) dotp = 0.999999;
        double angle = acos(dotp);
        double magnitude = -2.0 * ab.c1 * (angle - ab.angle);
        double fa = magnitude / l_ba;
        double fc = magnitude / l_bc;
        t.x = ba.y * bc.z - ba.z * bc.y;
        t.y = ba.z * bc.x - ba.x * bc.z;
        t.z = ba.x * bc.y - ba.y * bc.x;
        or_vec3 pa, pc;
        pa.x = (ba.y * t.z - ba.z * t.y);
        pa.y = (ba.z * t.x - ba.x * t.z);
        pa.z = (ba.x * t.y - ba.y * t.x);
        double ln = sqrt(pa.x * pa.x + pa.y * pa.y + pa.z * pa.z);
        if (ln != 0.0) ln = 1.0 / ln;
        ln *= fa;
        pa.x *= ln; pa.y *= ln; pa.z *= ln;
        pc.x = (cb.y * t.z - cb.z * t.y);
        pc.y = (cb.z * t.x - cb.x * t.z);
        pc.z = (cb.x * t.y - cb.y * t.x);
        ln = sqrt(pc.x * pc.x + pc.y * pc.y + pc.z * pc.z);
        if (ln != 0.0) ln = 1.0 / ln;
        ln *= fc;
        pc.x *= ln; pc.y *= ln; pc.z *= ln;
        OR_ADD(&beads[ab.b1].force.x, pa.x);
        OR_ADD(&beads[ab.b1].force.y, pa.y);
        OR_ADD(&beads[ab.b1].force.z, pa.z);
        OR_ADD(&beads[ab.b3].force.x, pc.x);
        OR_ADD(&beads[ab.b3].force.y, pc.y);
        OR_ADD(&beads[ab.b3].force.z, pc.z);
        OR_ADD(&beads[ab.b2].force.x, -pa.x - pc.x);
        OR_ADD(&beads[ab.b2].force.y, -pa.y - pc.y);
        OR_ADD(&beads[ab.b2].force.z, -pa.z - pc.z);
    }
}

/* kernel.cl:444-464 AdvanceVelocity. */
void or_advance_velocity(or_bead *beads, const or_info *info, int n)
{
    OR_PARALLEL_FOR
    for (int i = 0; i < n; ++i) {
        double h = info[i].half_dt_mass_inv;
        or_vec3 v = beads[i].velocity_old;
        or_vec3 f = beads[i].force;
        f.x += beads[i].force_old.x;
        f.y += beads[i].force_old.y;
        f.z += beads[i].force_old.z;
        v.x = fma(f.x, h, v.x);
        v.y = fma(f.y, h, v.y);
        v.z = fma(f.z, h, v.z);
        beads[i].velocity_old = v;
    }
}

/* kernel.cl:468-483 ResetBeadEnergy. */
void or_reset_bead_energy(or_bead *beads, int n)
{
    OR_PARALLEL_FOR
    for (int i = 0; i < n; ++i) {
        beads[i].force.x = beads[i].force.y = beads[i].force.z = 0.0;
        beads[i].velocity.x = beads[i].velocity.y = beads[i].velocity.z = 0.0;
        beads[i].force_old.x = beads[i].force_old.y = beads[i].force_old.z = 0.0;
        beads[i].velocity_old.x = beads[i].velocity_old.y = beads[i].velocity_old.z = 0.0;
    }
}

/* kernel.cl:508-515 ResetGrid and kernel.cl:518-525 ResetContactList. */
void or_reset_counts(int *head, int n) { memset(head, 0, sizeof(int) * (size_t)n); }

/* Cell coordinates of a position, kernel.cl:493-495. */
static void or_cell_of(const or_vec3 *p, const or_settings *s, int *cx, int *cy, int *cz)
{
    *cx = (int)floor(p->x / s->step.x);
    *cy = (int)floor(p->y / s->step.y);
    *cz = (int)floor(p->z / s->step.z);
}

/* kernel.cl:487-506 FillGrid.  cell_of_bead (may be NULL) receives the cell
 * index of every bead, -1 for beads outside the grid. */
void or_fill_grid(const or_bead *beads, int *head, int *tail, const or_settings *s, int n, int *cell_of_bead)
{
    OR_PARALLEL_FOR
    for (int i = 0; i < n; ++i) {
        int cx, cy, cz;
        or_cell_of(&beads[i].position, s, &cx, &cy, &cz);
        int cell = -1;
        if (cx >= 0 && cx < s->cx && cy >= 0 && cy < s->cy && cz >= 0 && cz < s->cz) {
            cell = cx + cy * s->cx + cz * s->cxy;
            int old = OR_FETCH_INC(&head[cell]);
            if (old < OR_GRID_TAIL_SIZE) tail[(size_t)cell * OR_GRID_TAIL_SIZE + old] = i;
        }
        if (cell_of_bead) cell_of_bead[i] = cell;
    }
}

/* Pair predicate and append of BuildContactList, kernel.cl:574-659 (intra-cell)
 * and kernel.cl:690-775 (half-shell), which are the same code. */
static void or_try_pair(int i, int j, const or_bead *beads, const or_info *info, const or_pair *table,
                        const or_settings *s, int *nh_head, int *nh_tail, int *pair_ids)
{
    const or_info *a = &info[i];
    const or_info *b = &info[j];
    int collidable = (a->tag != b->tag) || (abs(a->loc_id - b->loc_id) > 1);
    if (!collidable) return;
    or_vec3 rij;
    rij.x = beads[i].position.x - beads[j].position.x;
    rij.y = beads[i].position.y - beads[j].position.y;
    rij.z = beads[i].position.z - beads[j].position.z;
    double length = rij.x * rij.x + rij.y * rij.y + rij.z * rij.z;
    if (s->bc_type == 0 && length > s->half_box_sq) {
        if (rij.x * rij.x > s->half_box_sq) rij.x += (rij.x > 0.0) ? -s->size.x : s->size.x;
        if (rij.y * rij.y > s->half_box_sq) rij.y += (rij.y > 0.0) ? -s->size.y : s->size.y;
        if (rij.z * rij.z > s->half_box_sq) rij.z += (rij.z > 0.0) ? -s->size.z : s->size.z;
        length = rij.x * rij.x + rij.y * rij.y + rij.z * rij.z;
    }
    int pid = or_pair_index(a->type, b->type, s->types_num);
    int ptype = table[pid].type;
    int take = 0;
    if (ptype == 1 || ptype == 3) take = length < s->cutoff_sq;
    if (ptype == 2) {
        double min_skin = (a->skin_radius > b->skin_radius) ? b->skin_radius : a->skin_radius;
        double max_radius = (a->radius > b->radius) ? a->radius : b->radius;
        take = length < pow(max_radius + min_skin, 2.0);
    }
    if (!take) return;
    int k = OR_FETCH_INC(&nh_head[i]);
    if (k < OR_BEAD_NHOOD_SIZE) {
        nh_tail[(size_t)i * OR_BEAD_NHOOD_SIZE + k] = j;
        pair_ids[(size_t)i * OR_BEAD_NHOOD_SIZE + k] = pid;
    }
}

/* Note on x-add vs x-sub in the minimum image above: kernel.cl:602-608 writes
 * "if (r > 0) r -= L; else r += L"; r + (-L) and r - L are the same IEEE
 * operation, so the restatement is exact. */

/* kernel.cl:532-789 BuildContactList, cells visited in work-item order. */
void or_build_contact_list(const int *head, const int *tail, const int *cell_nh, int *nh_head, int *nh_tail,
                           const or_bead *beads, const or_info *info, int *pair_ids, const or_pair *table,
                           const or_settings *s, int n_cells)
{
    OR_PARALLEL_FOR
    for (int c = 0; c < n_cells; ++c) {
        int nc = head[c];
        if (nc <= 0) continue;
        if (nc > OR_GRID_TAIL_SIZE) nc = OR_GRID_TAIL_SIZE;
        const int *own = &tail[(size_t)c * OR_GRID_TAIL_SIZE];
        for (int ii = 0; ii < nc - 1; ++ii)
            for (int jj = ii + 1; jj < nc; ++jj)
                or_try_pair(own[ii], own[jj], beads, info, table, s, nh_head, nh_tail, pair_ids);
        for (int q = 0; q < 13; ++q) {
            int nb = cell_nh[c * 13 + q];
            if (nb == -1) continue;
            int nn = head[nb];
            if (nn <= 0) continue;
            if (nn > OR_GRID_TAIL_SIZE) nn = OR_GRID_TAIL_SIZE;
            const int *other = &tail[(size_t)nb * OR_GRID_TAIL_SIZE];
            for (int jj = 0; jj < nn; ++jj)
                for (int kk = 0; kk < nc; ++kk)
                    or_try_pair(own[kk], other[jj], beads, info, table, s, nh_head, nh_tail, pair_ids);
        }
    }
}

/* kernel.cl:792-1011 ResolveContacts. */
void or_resolve_contacts(const int *nh_head, const int *nh_tail, or_bead *beads, const or_info *info,
                         const int *pair_ids, const or_pair *table, const or_settings *s, int n)
{
    OR_PARALLEL_FOR
    for (int i = 0; i < n; ++i) {
        int m = nh_head[i];
        if (m <= 0) continue;
        if (m > OR_BEAD_NHOOD_SIZE) m = OR_BEAD_NHOOD_SIZE;
        int has_contact = 0;
        double r1 = info[i].radius;
        or_vec3 acc = {0.0, 0.0, 0.0, 0.0};
        or_vec3 p1 = beads[i].position;
        for (int e = 0; e < m; ++e) {
            int j = nh_tail[(size_t)i * OR_BEAD_NHOOD_SIZE + e];
            or_vec3 p2 = beads[j].position;
            or_vec3 rij;
            rij.x = p1.x - p2.x;
            rij.y = p1.y - p2.y;
            rij.z = p1.z - p2.z;
            double length = rij.x * rij.x + rij.y * rij.y + rij.z * rij.z;
            if (length > s->half_box_sq) {  /* kernel.cl:844-870, no bc_type test */
                if (rij.x * rij.x > s->half_box_sq) rij.x += (rij.x > 0.0) ? -s->size.x : s->size.x;
                if (rij.y * rij.y > s->half_box_sq) rij.y += (rij.y > 0.0) ? -s->size.y : s->size.y;
                if (rij.z * rij.z > s->half_box_sq) rij.z += (rij.z > 0.0) ? -s->size.z : s->size.z;
                length = rij.x * rij.x + rij.y * rij.y + rij.z * rij.z;
            }
            const or_pair *pr = &table[pair_ids[(size_t)i * OR_BEAD_NHOOD_SIZE + e]];
            double fx = 0.0, fy = 0.0, fz = 0.0;
            int hit = 0;
            if (pr->type == 1 && length < s->cutoff_sq) {          /* kernel.cl:875-919 */
                double r2 = length;
                double r = or_native_sqrt(length);
                if (r < pr->c3) {
                    double sigma2 = pr->c1 * pr->c1;
                    double fr2 = sigma2 / r2;
                    double fr6 = fr2 * fr2 * fr2;
                    double fpr = 48.0 * pr->c2 * fr6 * (fr6 - 0.5) / r2;
                    fpr -= pr->c4;
                    double inv = (r != 0.0) ? 1.0 / r : 0.0;
                    fx = fpr * (rij.x * inv);
                    fy = fpr * (rij.y * inv);
                    fz = fpr * (rij.z * inv);
                    hit = 1;
                }
            }
            if (pr->type == 2) {                                   /* kernel.cl:922-957 */
                double r = or_native_sqrt(length);
                double pen = (r1 + info[j].radius) - r;
                if (pen > 0.0) {
                    double inv = (r != 0.0) ? 1.0 / r : 0.0;
                    pen *= pr->c1;
                    fx = pen * (rij.x * inv);
                    fy = pen * (rij.y * inv);
                    fz = pen * (rij.z * inv);
                    hit = 1;
                }
            }
            if (pr->type == 3 && length < pr->c4) {               /* kernel.cl:960-999 */
                double r = or_native_sqrt(length);
                double factor = (pr->c2 - r) * pr->c1;
                factor += (pr->c6 - r) * pr->c5;
                double inv = (r != 0.0) ? 1.0 / r : 0.0;
                fx = factor * (rij.x * inv);
                fy = factor * (rij.y * inv);
                fz = factor * (rij.z * inv);
                hit = 1;
            }
            if (hit) {
                acc.x += fx;
                acc.y += fy;
                acc.z += fz;
                OR_ADD(&beads[j].force.x, -fx);
                OR_ADD(&beads[j].force.y, -fy);
                OR_ADD(&beads[j].force.z, -fz);
                has_contact = 1;
            }
        }
        if (has_contact) {
            OR_ADD(&beads[i].force.x, acc.x);
            OR_ADD(&beads[i].force.y, acc.y);
            OR_ADD(&beads[i].force.z, acc.z);
        }
    }
}

/* Wrap or clip one neighbour coordinate; SimulationManager.cpp:2221-2240.
 * Returns -1 when a finite box leaves the grid. */
static int or_nb_coord(int v, int n, int periodic)
{
    if (periodic) {
        if (v < 0) v = n - 1;
        if (v >= n) v = 0;
        return v;
    }
    return (v < 0 || v >= n) ? -1 : v;
}

static int or_nb_cell(int x, int y, int z, const or_settings *s)
{
    int periodic = s->bc_type == 0;
    int nx = or_nb_coord(x, s->cx, periodic);
    int ny = or_nb_coord(y, s->cy, periodic);
    int nz = or_nb_coord(z, s->cz, periodic);
    if (s->bc_type != 0 && s->bc_type != 1) return -1;   /* table left at -1 */
    if (nx < 0 || ny < 0 || nz < 0) return -1;
    return nx + ny * s->cx + nz * s->cx * s->cy;
}

/* SimulationManager.cpp:2207-2301: the 13-cell half shell of every cell, in
 * the reference's order: nine cells at dz=-1 (dx outer, dy inner), three at
 * dx=+1, dz=0 (dy ascending), then (0,-1,0). */
void or_grid_neighbours(const or_settings *s, int *nh)
{
    for (int z = 0; z < s->cz; ++z)
        for (int y = 0; y < s->cy; ++y)
            for (int x = 0; x < s->cx; ++x) {
                int c = x + y * s->cx + z * s->cx * s->cy;
                int *out = &nh[(size_t)c * 13];
                int k = 0;
                for (int dx = -1; dx < 2; ++dx)
                    for (int dy = -1; dy < 2; ++dy) out[k++] = or_nb_cell(x + dx, y + dy, z - 1, s);
                for (int dy = -1; dy < 2; ++dy) out[k++] = or_nb_cell(x + 1, y + dy, z, s);
                out[k++] = or_nb_cell(x, y - 1, z, s);
            }
}


/* SimulationManager.cpp:151-181 RunSimulationLoop: the rebuild counter restarts
 * at zero on every call, so the grid is rebuilt at the first step of each call
 * and every grid_update_freq steps after it. */
int or_run_loop(or_state *st, int loop_size, int grid_update_freq)
{
    if (loop_size <= 0) return -1;
    int cnt = 0;
    for (int step = 0; step < loop_size; ++step) {
        if (cnt == grid_update_freq) cnt = 0;
        or_advance_position(st->beads, st->info, &st->settings, st->n_beads);
        if (st->n_springs > 0) or_resolve_springs(st->beads, st->springs, st->n_springs, st->step, st->break_step);
        if (st->n_angles > 0) or_resolve_angles(st->angles, st->springs, st->beads, st->n_angles);
        if (cnt == 0) {
            or_reset_counts(st->grid_head, st->n_cells);
            or_reset_counts(st->nh_head, st->n_beads);
            or_fill_grid(st->beads, st->grid_head, st->grid_tail, &st->settings, st->n_beads, st->cell_of_bead);
            or_build_contact_list(st->grid_head, st->grid_tail, st->cell_nh, st->nh_head, st->nh_tail, st->beads,
                                  st->info, st->pair_ids, st->pairs, &st->settings, st->n_cells);
        }
        or_resolve_contacts(st->nh_head, st->nh_tail, st->beads, st->info, st->pair_ids, st->pairs,
                            &st->settings, st->n_beads);
        or_advance_velocity(st->beads, st->info, st->n_beads);
        st->step++;
        cnt++;
    }
    return 0;
}

/* SimulationManager.cpp:111-143 MinimizeSystemEnergy (the loop count is the
 * integer quotient; the remainder of steps is dropped). */
int or_minimize(or_state *st, int steps, int sub_steps, int reset, int grid_update_freq)
{
    if (steps < 0) return -1;
    if (reset) or_reset_bead_energy(st->beads, st->n_beads);
    if (steps == 0) return 0;
    int loops = steps / sub_steps;
    for (int p = 0; p < loops; ++p) {
        if (or_run_loop(st, sub_steps, grid_update_freq)) return -1;
        if (reset) or_reset_bead_energy(st->beads, st->n_beads);
    }
    return 0;
}

/* Size of the records, for the Python side to check its dtypes against. */
int or_sizeof(int what)
{
    switch (what) {
    case 0: return (int)sizeof(or_bead);
    case 1: return (int)sizeof(or_info);
    case 2: return (int)sizeof(or_spring);
    case 3: return (int)sizeof(or_angle);
    case 4: return (int)sizeof(or_pair);
    case 5: return (int)sizeof(or_settings);
    case 6: return (int)sizeof(or_state);
    default: return -1;
    }
}
