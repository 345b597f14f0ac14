/*
 * axsq_quest.h -- C names of the quest signed-distance interface in libaxsq.so.
 *
 * The QUEST_signed_distance_* functions keep the names, arguments and meaning
 * of the reference's C wrapper for a build without MPI
 * (src/axom/quest/interface/c_fortran/wrapQUEST.h:84-129, over
 * src/axom/quest/interface/signed_distance.{hpp,cpp}): one process-wide
 * query, options set before init, evaluation after it. Every query runs on
 * the MI355X whatever execution space is set (CPU, OpenMP and GPU are
 * accepted and recorded).
 *
 * Misuse (an option set after init, evaluate before init, a null buffer):
 * where the reference aborts through SLIC_ERROR, these functions change
 * nothing, set the text that axsq_last_error() returns (it is empty after a
 * call that succeeded), and return -1 (init), NaN (evaluate) or nothing.
 *
 * The axsq_signed_distance_* functions add what the C wrapper lacks: init from
 * a mesh held in arrays (signed_distance_init(const mint::Mesh*),
 * signed_distance.hpp:137), the batched evaluate over SoA coordinates
 * (signed_distance.hpp:290-294), and the choice of traversal precision.
 * Their pointers may be host or device memory.
 */
#ifndef AXSQ_QUEST_H
#define AXSQ_QUEST_H

#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#include "axsq.h"

#ifdef __cplusplus
extern "C" {
#endif

/* axom::quest::SignedDistExec (wrapQUEST.h:33-38) */
enum QUEST_SignedDistExec {
  QUEST_SignedDistExec_CPU = 0,
  QUEST_SignedDistExec_OpenMP = 1,
  QUEST_SignedDistExec_GPU = 2
};

/* signed_distance_init(file): reads the STL file as quest::STLReader does
 * (axsq_stl_read) and builds the query. 0 on success, -1 on failure
 * (signed_distance.cpp:29-30,109-160). */
int QUEST_signed_distance_init_serial(const char* file);
/* The same for a Fortran string of file_len characters (trailing blanks
 * dropped). */
int QUEST_signed_distance_init_serial_bufferify(char* file, int file_len);
bool QUEST_signed_distance_initialized(void);
/* Bounds of the mesh's nodes (signed_distance.cpp:222-232); lo and hi hold
 * three doubles each. */
void QUEST_signed_distance_get_mesh_bounds(double* lo, double* hi);
void QUEST_signed_distance_set_dimension(int dim); /* only 3 */
void QUEST_signed_distance_set_closed_surface(bool status);
void QUEST_signed_distance_set_compute_signs(bool computeSign);
void QUEST_signed_distance_set_allocator(int allocatorID);
void QUEST_signed_distance_set_verbose(bool status);
void QUEST_signed_distance_use_shared_memory(bool status); /* no MPI-3: ignored */
void QUEST_signed_distance_set_execution_space(int execSpace);
/* signed_distance_evaluate(x, y, z) (signed_distance.hpp:248) */
double QUEST_signed_distance_evaluate_0(double x, double y, double z);
/* signed_distance_evaluate(x, y, z, cp..., n...) (signed_distance.hpp:266-274):
 * also the closest surface point and the unit (pseudo-)normal there. */
double QUEST_signed_distance_evaluate_1(double x, double y, double z, double* cp_x,
                                        double* cp_y, double* cp_z, double* n_x, double* n_y,
                                        double* n_z);
void QUEST_signed_distance_finalize(void);

/* signed_distance_init(const mint::Mesh* m) for a triangle mesh given as SoA
 * node coordinates (nnodes) and int32 connectivity (3 * ncells). 0 or -1. */
int axsq_signed_distance_init_mesh(const double* x, const double* y, const double* z,
                                   int64_t nnodes, const int32_t* conn, int64_t ncells);
/* signed_distance_evaluate(x, y, z, npoints, phi) (signed_distance.hpp:290-294).
 * AXSQ_OK or a negative status. */
int axsq_signed_distance_evaluate_batch(const double* x, const double* y, const double* z,
                                        int npoints, double* phi);
/* AXSQ_FP64 (default) or AXSQ_FP32_TRAVERSAL; an option, set before init. */
void axsq_signed_distance_set_precision(int fp_mode);

#ifdef __cplusplus
}
#endif
#endif
