/*
 * grains_oracle.h -- FP64 CPU oracle of the DEM time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared
 * against: a plain-C restatement of the reference's CPU step
 * (Grains/Base/src/GrainsCPU.cpp:46-70 and everything it calls).  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load
 * it; the product library never links or calls it.
 *
 * The API mirrors include/grains_hip.h call for call (go_* <-> grains_*) so a
 * test drives both with the same scene.  The state is kept the reference's
 * way: a row-major rotation matrix per particle (Transform3), converted to a
 * quaternion every step by Quaternion(Matrix3) (Quaternion.cpp:125-169).
 *
 * Parity pins, per function, are listed in the header comment of
 * grains_oracle.c.
 */
#ifndef GRAINS_ORACLE_H
#define GRAINS_ORACLE_H

#include "../include/grains_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct go_world go_world;

go_world* go_create(const grains_config_t* cfg);
void go_destroy(go_world* w);
const char* go_last_error(const go_world* w);

int go_set_shapes(go_world* w, const grains_shape_t* shapes, uint32_t n_shapes,
                  const double* vertices, uint32_t n_vertices);
int go_set_contact_tables(go_world* w, const grains_hooke_t* tables,
                          uint32_t n_materials, uint32_t n_tables);
int go_set_sims(go_world* w, const grains_sim_t* sims, uint32_t n_sims);
int go_set_obstacles(go_world* w, const uint32_t* shape, const uint32_t* sim,
                     const double* transforms, uint32_t n_obstacles);
int go_set_particles(go_world* w, const uint32_t* shape, const uint32_t* sim,
                     uint64_t n);
int go_set_state(go_world* w, const double* pos, const double* quat,
                 const double* vel, const double* omega, uint64_t n);
int go_get_state(go_world* w, double* pos, double* quat, double* vel,
                 double* omega, uint64_t n);
int go_set_transforms(go_world* w, const double* transforms, uint64_t n);
int go_get_transforms(go_world* w, double* transforms, uint64_t n);

/* n_threads <= 0: all OpenMP threads. */
int go_step(go_world* w, uint32_t n_steps, int n_threads);

int go_get_cell_keys(go_world* w, uint32_t* keys, uint32_t* sorted, uint64_t n);
int go_get_cell_layout(go_world* w, uint32_t sim, uint32_t* ncell3,
                       double* cell_size, uint32_t* cell_base);
int go_get_pairs(go_world* w, int which, uint32_t* ij, uint64_t cap, uint64_t* m);
int go_get_contact_count(go_world* w, int32_t* counts, uint64_t n);
int go_get_torce(go_world* w, double* force, double* torque, uint64_t n);
int go_get_stats(go_world* w, grains_stats_t* stats);

/* ---- primitives, for unit tests ------------------------------------------ */
/* Mass properties of a shape: out = volume, mass, I[6], I^-1[6], R_c
 * (reference formula), bounding box[3], exact R_c  (19 doubles). */
int go_shape_props(const grains_shape_t* s, const double* vertices, double* out);
/* Support point of the full shape (REF semantics) in its own frame. */
int go_support(const grains_shape_t* s, const double* vertices,
               const double* dir, double* out);
/* LinkedCell: number of cells per direction (LinkedCell.cpp:23-29). */
void go_cell_counts(const double* min, const double* max, double h, uint32_t* n3);
/* 27-neighbour table, UINT32_MAX off grid (LinkedCell.cpp:35-72). */
void go_neighbor_table(const uint32_t* n3, uint32_t* table);
/* Cell hash of points (LinkedCell.cpp:117-149) on the grid [min, min + n*h];
 * returns the number of points outside it (their indices are clamped). */
uint64_t go_cell_hash(const double* min, double h, const uint32_t* n3,
                      const double* pos, uint64_t n, uint32_t* keys);
/* GJK distance of a pair given two 12-double transforms; flags as in
 * grains_config_t.  Returns the distance; witness points in world frame. */
double go_gjk(const grains_shape_t* a, const grains_shape_t* b,
              const double* vertices, const double* ta, const double* tb,
              uint32_t flags, double rel_tol, double* pa, double* pb,
              int* iters);
/* Full narrow phase of one pair (closestPointsRigidBodies,
 * CollisionDetection.cpp:183-284).  out = contact point[3], contact
 * vector[3], overlap.  Returns 1 on contact. */
int go_closest_points(const grains_shape_t* a, const grains_shape_t* b,
                      const double* vertices, const double* ta,
                      const double* tb, uint32_t flags, double rel_tol, double* out);
/* Hooke force on A for one contact (HookeContactForceModel.cpp:91-186).
 * ci = cp[3], cv[3], overlap; out = force[3], torque[3]. */
void go_hooke(const grains_hooke_t* p, const double* ci, const double* rel_vel,
              const double* rel_ang_vel, double m1, double m2,
              const double* origin_a, double* out);
/* Quaternion(Matrix3) (Quaternion.cpp:125-169): q = (x,y,z,w). */
void go_quat_from_matrix(const double* m9, double* q);
/* Rotation matrix of a unit quaternion (x,y,z,w). */
void go_matrix_from_quat(const double* q, double* m9);
/* Transform3::setBasis(aX, aY, aZ) (Transform3.cpp:151-162), radians. */
void go_basis_from_angles(double ax, double ay, double az, double* m9);

#ifdef __cplusplus
}
#endif

#endif
