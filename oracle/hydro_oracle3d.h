/*
 * hydro_oracle3d.h -- CPU oracle of the hydro3d time step.
 *
 * TEST INFRASTRUCTURE ONLY: the 3D restatement of the reference algorithm
 * (hydro<geom3d::MeshStructured<double>>, hydro3d.cpp:17-20), used as the
 * parity checker of hg3_* (include/hydro_gpu3d.h).  Only tests/ load it;
 * the product library never links or calls it.
 */
#ifndef HYDRO_ORACLE3D_H_
#define HYDRO_ORACLE3D_H_

#include "../include/hydro_gpu3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or3_ctx or3_ctx;

int or3_create(const hg3_config* cfg, or3_ctx** out, char* err, size_t errlen);
void or3_destroy(or3_ctx* ctx);
int or3_num_cells(const or3_ctx* ctx);
int or3_num_faces(const or3_ctx* ctx);
int or3_get_field(or3_ctx* ctx, int field, double* host, size_t n);
int or3_set_field(or3_ctx* ctx, int field, const double* host, size_t n);
int or3_step(or3_ctx* ctx, hg_step_stats* stats);

#ifdef __cplusplus
}
#endif

#endif
