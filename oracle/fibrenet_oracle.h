/*
 * fibrenet_oracle.h - records and run state shared by the sequential oracle
 * (fibrenet_oracle.c) and its multi-threaded timing twin
 * (fibrenet_oracle_mt.c).
 *
 * TEST INFRASTRUCTURE ONLY: nothing in the product includes or links it.
 */
#ifndef FIBRENET_ORACLE_H
#define FIBRENET_ORACLE_H

#define OR_GRID_TAIL_SIZE 64    /* kernel.cl:10 */
#define OR_BEAD_NHOOD_SIZE 256  /* kernel.cl:11 */

/* Records restated from kernel.cl:13-156 (AoS, natural alignment). */
typedef struct { double x, y, z, w; } or_vec3;
typedef struct { or_vec3 position, image_position, velocity, force, velocity_old, force_old; } or_bead;
typedef struct {
    int type, tag, loc_id;
    double mass, mass_inv, lambda_dt_mass_inv, half_dt_mass_inv, half_dt2_mass_inv;
    double radius, radius_sq, skin_radius, skin_radius_sq;
} or_info;
typedef struct {
    int type, tag, status, p1, p2;
    double rest_length, stiffness, damping, c1, c2;
} or_spring;
typedef struct {
    int type, tag, status, b1, b2, b3, s1, s2;
    double angle, current_angle, c1, c2, c3;
} or_angle;
typedef struct { int type, p1, p2; double c1, c2, c3, c4, c5, c6; } or_pair;
typedef struct {
    int types_num, bc_type;
    double wall_k, cutoff, cutoff_sq, unit, dt, half_dt, half_dt2, lambda_dt, gamma, sigma, damping;
    int cx, cy, cz, cxy;
    double half_box_sq;
    or_vec3 step, size, gravity;
} or_settings;

/* Whole state of one oracle run. */
typedef struct {
    or_bead *beads;
    or_info *info;
    or_spring *springs;
    or_angle *angles;
    or_pair *pairs;
    or_settings settings;
    int n_beads, n_springs, n_angles, n_cells;
    int *cell_nh;       /* n_cells*13                     */
    int *grid_head;     /* n_cells                        */
    int *grid_tail;     /* n_cells*64                     */
    int *nh_head;       /* n_beads                        */
    int *nh_tail;       /* n_beads*256                    */
    int *pair_ids;      /* n_beads*256                    */
    int *cell_of_bead;  /* n_beads, last rebuild          */
    int *break_step;    /* n_springs, -1 = intact         */
    int step;           /* global step counter            */
} or_state;

#endif /* FIBRENET_ORACLE_H */
