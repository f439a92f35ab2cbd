/* Serial MPI stand-in for the single-GPU step libraries.
 *
 * The generated ABI (reference: src/Formura/MPICxx/Translate.hs:1000-1139)
 * includes <mpi.h> for `MPI_Comm` in `struct Formura_Navigator` and in
 * `Formura_Init`.  A one-GPU library has exactly one rank, so this header
 * provides that type and the handful of calls the reference drivers make
 * (examples/pearson-3d-main.cpp: MPI_Init, MPI_Comm_rank, MPI_Barrier,
 * MPI_Finalize) as single-rank no-ops.  On a machine with MPI, put the real
 * <mpi.h> first on the include path instead.
 *
 * It also delivers the self-messages of a single rank, for C generated with
 * the reference's region plan (facet exchange,
 * MPICxx/Translate.hs:798-836): MPI_Irecv posts a receive under its tag,
 * the MPI_Isend with the same tag copies its bytes into it, MPI_Wait returns.
 * A send with no posted receive aborts (that generated code always posts
 * the receive first).
 */
#pragma once
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
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

typedef int MPI_Datatype;
typedef int MPI_Request;
typedef struct { int MPI_SOURCE, MPI_TAG, MPI_ERROR; } MPI_Status;
#define MPI_BYTE ((MPI_Datatype)1)
#define MPI_STATUS_IGNORE ((MPI_Status *)0)

enum { FMR_MPI_SLOTS = 256 };
static struct { void *buf; int count; int tag; int live; } fmr_mpi_posted[FMR_MPI_SLOTS];

static inline int MPI_Irecv(void *buf, int count, MPI_Datatype t, int src, int tag, MPI_Comm c, MPI_Request *req) {
  (void)t; (void)src; (void)c;
  for (int i = 0; i < FMR_MPI_SLOTS; ++i)
    if (!fmr_mpi_posted[i].live) {
      fmr_mpi_posted[i].buf = buf;
      fmr_mpi_posted[i].count = count;
      fmr_mpi_posted[i].tag = tag;
      fmr_mpi_posted[i].live = 1;
      *req = i;
      return MPI_SUCCESS;
    }
  fprintf(stderr, "serial MPI: too many posted receives\n");
  abort();
}

static inline int MPI_Isend(const void *buf, int count, MPI_Datatype t, int dest, int tag, MPI_Comm c,
                            MPI_Request *req) {
  (void)t; (void)dest; (void)c;
  for (int i = 0; i < FMR_MPI_SLOTS; ++i)
    if (fmr_mpi_posted[i].live && fmr_mpi_posted[i].tag == tag) {
      if (fmr_mpi_posted[i].count != count) {
        fprintf(stderr, "serial MPI: message of %d bytes for a receive of %d (tag %d)\n", count,
                fmr_mpi_posted[i].count, tag);
        abort();
      }
      memcpy(fmr_mpi_posted[i].buf, buf, (size_t)count);
      fmr_mpi_posted[i].live = 0;
      *req = i;
      return MPI_SUCCESS;
    }
  fprintf(stderr, "serial MPI: send with tag %d has no posted receive\n", tag);
  abort();
}

static inline int MPI_Wait(MPI_Request *req, MPI_Status *st) { (void)req; (void)st; return MPI_SUCCESS; }

#ifdef __cplusplus
}
#endif
