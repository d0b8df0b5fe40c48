/*
 * fibrenet_types.h - host wire format of the FibreNet particle-dynamics step.
 *
 * These are the array-of-structures records that FibreNet's host code builds
 * and uploads, laid out byte-for-byte as the reference declares them, so that
 * a caller holding FibreNet's own std::vector<Bead> (etc.) can pass its
 * storage straight through the C ABI in fibrenet.h.  The engine converts them
 * to its internal structure-of-arrays layout on upload and back on read.
 *
 *   fn_vector3        <- Vector3.h:3-6,           kernel.cl:13-16      (32 B)
 *   fn_bead           <- Bead.h:6-14,             kernel.cl:19-27      (192 B)
 *   fn_bead_info      <- BeadInfo.h:6-20,         kernel.cl:30-45      (88 B)
 *   fn_spring         <- SpringStructure.h:6-18,  kernel.cl:127-139    (64 B)
 *   fn_angle          <- AngleBond.h:5-20,        kernel.cl:141-156    (72 B)
 *   fn_pair           <- Pair.h:5-16,             kernel.cl:113-124    (64 B)
 *   fn_settings       <- SimulationSettings.h:6-29, kernel.cl:74-97    (216 B)
 *   fn_grid_cell_nh   <- BufferChunk.h:6-9,       kernel.cl:49-52      (52 B)
 *   fn_grid_cell      <- BufferChunk.h:22-25,     kernel.cl:69-72      (256 B)
 *   fn_particle_nh    <- BufferChunk.h:14-17,     kernel.cl:56-59      (1024 B)
 *
 * All members use natural alignment (the reference relies on identical
 * host/OpenCL-C layout of these structs); the static asserts pin the sizes.
 */
#ifndef FIBRENET_TYPES_H
#define FIBRENET_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

/* Capacities of the reference's fixed-size grid buckets and contact lists
 * (kernel.cl:10-11, BufferChunk.h:16,24). */
#define FN_GRID_TAIL_SIZE 64
#define FN_BEAD_NHOOD_SIZE 256

typedef struct fn_vector3 { double x, y, z, w; } fn_vector3;

typedef struct fn_bead {
    fn_vector3 position;        /* wrapped position (contacts, grid)          */
    fn_vector3 image_position;  /* unwrapped position (springs, angles)       */
    fn_vector3 velocity;        /* predicted velocity of the last drift       */
    fn_vector3 force;           /* force of the current step                  */
    fn_vector3 velocity_old;    /* velocity advanced by the closing kick      */
    fn_vector3 force_old;       /* force of the previous step                 */
} fn_bead;

typedef struct fn_bead_info {
    int    type;                /* index into the declared type list          */
    int    tag;                 /* group id (fibre index)                     */
    int    loc_id;              /* position along the group                   */
    double mass;
    double mass_inv;
    double lambda_dt_mass_inv;  /* lambda*dt/m                                */
    double half_dt_mass_inv;    /* (dt/2)/m                                   */
    double half_dt2_mass_inv;   /* (dt*dt/2)/m                                */
    double radius;
    double radius_sq;
    double skin_radius;
    double skin_radius_sq;
} fn_bead_info;

typedef struct fn_spring {
    int    type;                /* 0 = HR, 1 = HR_FL (breakable)              */
    int    tag;
    int    status;              /* 1 = intact, 0 = broken                     */
    int    p1, p2;
    double rest_length;
    double stiffness;
    double damping;             /* parsed, never used by the step             */
    double c1;                  /* HR_FL: Young's modulus                     */
    double c2;                  /* HR_FL: failure stress                      */
} fn_spring;

typedef struct fn_angle {
    int    type;                /* 0 = HR, 1 = HR_FL (dies with its springs)  */
    int    tag;
    int    status;
    int    b1, b2, b3;
    int    s1, s2;              /* springs b1-b2 and b2-b3                    */
    double angle;               /* rest angle                                 */
    double current_angle;       /* unused by the step                         */
    double c1;                  /* stiffness                                  */
    double c2, c3;              /* unused by the step                         */
} fn_angle;

typedef struct fn_pair {
    int    type;                /* 1 = lj, 2 = hs, 3 = st                     */
    int    p1, p2;              /* the two bead types                         */
    double c1, c2, c3, c4, c5, c6;
} fn_pair;

typedef struct fn_settings {
    int        types_num;
    int        bc_type;         /* 0 = periodic, 1 = finite box               */
    double     wall_contact_stiffness;
    double     cutoff;
    double     cutoff_sq;
    double     box_unit_cell;
    double     dt;
    double     half_dt;
    double     half_dt2;
    double     lambda_dt;
    double     gamma;
    double     sigma;
    double     damping;         /* already negated by the loader              */
    int        cells_x, cells_y, cells_z, cells_xy;
    double     half_box_sq;     /* (min(L)/2)^2                               */
    fn_vector3 grid_step;
    fn_vector3 box_size;
    fn_vector3 gravity;
} fn_settings;

typedef struct fn_grid_cell_nh { int n[13]; } fn_grid_cell_nh;
typedef struct fn_grid_cell { int n[FN_GRID_TAIL_SIZE]; } fn_grid_cell;
typedef struct fn_particle_nh { int n[FN_BEAD_NHOOD_SIZE]; } fn_particle_nh;

#ifdef __cplusplus
}
static_assert(sizeof(fn_vector3) == 32, "fn_vector3 layout");
static_assert(sizeof(fn_bead) == 192, "fn_bead layout");
static_assert(sizeof(fn_bead_info) == 88, "fn_bead_info layout");
static_assert(sizeof(fn_spring) == 64, "fn_spring layout");
static_assert(sizeof(fn_angle) == 72, "fn_angle layout");
static_assert(sizeof(fn_pair) == 64, "fn_pair layout");
static_assert(sizeof(fn_settings) == 216, "fn_settings layout");
static_assert(sizeof(fn_grid_cell_nh) == 52, "fn_grid_cell_nh layout");
#endif

#endif /* FIBRENET_TYPES_H */
