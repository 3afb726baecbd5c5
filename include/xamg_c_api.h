/*
 * XAMG C API on one AMD Instinct MI355X (gfx950).
 *
 * Drop-in replacement for the reference's api/c/xamg_c_api.h (lines 26-123): the
 * same symbol names, handle types and C signatures.  Objects live in a process-wide
 * table indexed by size_t handles that are never reused (xamg_c_api.cpp:64-85);
 * *_destroy releases one slot, XAMG_finalize releases all.  Matrices and distributed
 * vectors are held in HBM; fill/upload calls copy from caller memory and download
 * copies out (xamg_c_api.cpp:378-443).
 *
 * Differences a caller can observe:
 *  - nv is a run-time value in {1, 2, 4, 8, 16} instead of the compile-time CONST_NV;
 *    XAMG_get_nv() returns the value of the XAMG_NV environment variable (default 1).
 *  - Every function returns 0 on success.  On failure it returns a nonzero code (the
 *    reference asserts instead); xamgx_last_error() (xamg_ext.h) gives the message.
 *  - One process drives one GPU: XAMG_id_get_proc() is 0, XAMG_id_get_nprocs() is 1
 *    and XAMG_id_get_comm_ptr() returns NULL.
 *  - Not thread-safe, like the reference.
 */
#ifndef XAMG_C_API_H_
#define XAMG_C_API_H_

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t XAMG_param_id;
typedef size_t XAMG_param_override_id;
typedef size_t XAMG_part_id;
typedef size_t XAMG_vector_id;
typedef size_t XAMG_csr_matrix_id;
typedef size_t XAMG_matrix_id;
typedef size_t XAMG_solver_id;

typedef enum { XAMG_MEM_LOCAL, XAMG_MEM_SHARED, XAMG_MEM_DISTRIBUTED } vector_alloc_mode;

/* xamg_c_api.h:43-44 -> XAMG::init / finalize (init.h:54-80) */
int XAMG_init(int argc, char **argv, char *conf);
int XAMG_finalize();

/* xamg_c_api.h:46 */
int XAMG_get_nv();

/* xamg_c_api.h:48-52 (MPI identity; one GPU = one rank) */
int XAMG_id_get_proc();
int XAMG_id_get_nprocs();
void *XAMG_id_get_comm_ptr();
#define XAMG_id_get_comm() (*((MPI_Comm *)XAMG_id_get_comm_ptr()))

/* xamg_c_api.h:54-79 parameter dictionary (src/param) */
int XAMG_param_create(XAMG_param_id *);
int XAMG_param_add_value_s(char *solver_type, char *key, char *value, XAMG_param_id pid);
int XAMG_param_add_value_i(char *solver_type, char *key, uint32_t value, XAMG_param_id pid);
int XAMG_param_add_value_d(char *solver_type, char *key, double value, XAMG_param_id pid);
int XAMG_param_add_value_b(char *solver_type, char *key, uint8_t value, XAMG_param_id pid);

int XAMG_param_forced_change_value_s(char *solver_type, char *key, char *value, XAMG_param_id pid);
int XAMG_param_forced_change_value_i(char *solver_type, char *key, uint32_t value,
                                     XAMG_param_id pid);
int XAMG_param_forced_change_value_d(char *solver_type, char *key, double value, XAMG_param_id pid);
int XAMG_param_forced_change_value_b(char *solver_type, char *key, uint8_t value,
                                     XAMG_param_id pid);

int XAMG_param_change_value_s(char *solver_type, char *key, char *value, XAMG_param_id pid);
int XAMG_param_change_value_i(char *solver_type, char *key, uint32_t value, XAMG_param_id pid);
int XAMG_param_change_value_d(char *solver_type, char *key, double value, XAMG_param_id pid);
int XAMG_param_change_value_b(char *solver_type, char *key, uint8_t value, XAMG_param_id pid);

int XAMG_param_override_create(XAMG_param_override_id *);
int XAMG_param_override_add_value(char *key, char *val, XAMG_param_override_id ovid);
int XAMG_param_override_apply(char *solver_type, XAMG_param_id pid, XAMG_param_override_id ovid);
int XAMG_param_override_destroy(XAMG_param_override_id);

void XAMG_param_set_default(XAMG_param_id);
int XAMG_param_print(XAMG_param_id);
int XAMG_param_destroy(XAMG_param_id);

/* xamg_c_api.h:81-84 row partition (src/part; one block on one GPU) */
int XAMG_part_create(XAMG_part_id *);
int XAMG_part_construct(int, XAMG_part_id);
int XAMG_part_get_numa_size(XAMG_part_id);
int XAMG_part_destroy(XAMG_part_id);

/* xamg_c_api.h:86-96 vectors: [size][nv] row-interleaved FP64, in HBM unless LOCAL */
int XAMG_vector_create(XAMG_vector_id *, vector_alloc_mode);
int XAMG_vector_offset(uint64_t, XAMG_vector_id);
int XAMG_vector_alloc_d(int, uint16_t, XAMG_vector_id);
int XAMG_vector_set_part(XAMG_part_id, XAMG_vector_id);
int XAMG_vector_set_val_d(double, XAMG_vector_id);
int XAMG_vector_download_d(double *, uint64_t, uint64_t, uint16_t, XAMG_vector_id);
int XAMG_vector_upload_v(XAMG_vector_id, XAMG_vector_id);
int XAMG_vector_upload_d(double *, uint64_t, uint64_t, uint16_t, XAMG_vector_id);
int XAMG_vector_destroy(XAMG_vector_id);

/* xamg_c_api.h:98-103 host CSR block (int32 input copied to u32, csr_matrix.h:47-559) */
int XAMG_csr_matrix_create(XAMG_csr_matrix_id *);
int XAMG_csr_matrix_alloc(uint64_t, uint64_t, uint64_t, XAMG_csr_matrix_id);
int XAMG_csr_matrix_fill(int *, int *, double *, XAMG_matrix_id);
int XAMG_csr_matrix_offset(int, int, XAMG_matrix_id);
uint32_t XAMG_csr_matrix_get_nrows(XAMG_csr_matrix_id);
int XAMG_csr_matrix_destroy(XAMG_csr_matrix_id);

/* xamg_c_api.h:105-108 device matrix (construct uploads the CSR to HBM) */
int XAMG_matrix_create(XAMG_matrix_id *);
int XAMG_matrix_set_part(XAMG_part_id, XAMG_matrix_id);
int XAMG_matrix_construct(XAMG_csr_matrix_id mcsrid, XAMG_matrix_id mid);
int XAMG_matrix_destroy(XAMG_matrix_id);

/* xamg_c_api.h:110-120 solvers (construct_solver_hierarchy + base_solver::solve) */
int XAMG_solver_create1(XAMG_matrix_id, XAMG_param_id, XAMG_solver_id *);
int XAMG_solver_create2(XAMG_matrix_id, XAMG_param_id, XAMG_vector_id, XAMG_vector_id,
                        XAMG_solver_id *);
int XAMG_solver_renew_params(XAMG_param_id, XAMG_solver_id);
int XAMG_solver_solve1(uint16_t, XAMG_solver_id);
int XAMG_solver_solve2(uint16_t, XAMG_vector_id, XAMG_vector_id, XAMG_solver_id);
int XAMG_solver_print_info(XAMG_solver_id);
int XAMG_solver_get_convergence_info(int *, uint16_t, XAMG_solver_id);
int XAMG_solver_get_iter_info(int *, XAMG_solver_id);
int XAMG_solver_destroy(XAMG_solver_id);

/* xamg_c_api.h:122-123 */
double XAMG_timer();
int XAMG_barrier();

#ifdef __cplusplus
}
#endif

#endif /* XAMG_C_API_H_ */
