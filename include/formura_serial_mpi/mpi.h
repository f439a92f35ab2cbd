/* Serial MPI stand-in for the single-GPU step libraries.
 *
 * The generated ABI (reference: src/Formura/MPICxx/Translate.hs:1000-1139)
 * includes <mpi.h> for `MPI_Comm` in `struct Formura_Navigator` and in
 * `Formura_Init`.  A one-GPU library has exactly one rank, so this header
 * provides that type and the handful of calls the reference drivers make
 * (examples/pearson-3d-main.cpp: MPI_Init, MPI_Comm_rank, MPI_Barrier,
 * MPI_Finalize) as single-rank no-ops.  On a machine with MPI, put the real
 * <mpi.h> first on the include path instead.
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

typedef int MPI_Comm;
#define MPI_COMM_WORLD ((MPI_Comm)0)
#define MPI_SUCCESS 0

static inline int MPI_Init(int *argc, char ***argv) { (void)argc; (void)argv; return MPI_SUCCESS; }
static inline int MPI_Finalize(void) { return MPI_SUCCESS; }
static inline int MPI_Comm_rank(MPI_Comm c, int *r) { (void)c; *r = 0; return MPI_SUCCESS; }
static inline int MPI_Comm_size(MPI_Comm c, int *s) { (void)c; *s = 1; return MPI_SUCCESS; }
static inline int MPI_Barrier(MPI_Comm c) { (void)c; return MPI_SUCCESS; }

#ifdef __cplusplus
}
#endif
