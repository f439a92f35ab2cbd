// Driver for a generated 3D Gray-Scott step library, written against the
// Formura_Navigator ABI only (the contract of the reference drivers,
// examples/pearson-3d-main.cpp:160-245): Formura_Init, initial condition
// written into the state arrays on [lower, upper), Formura_Forward in the time
// loop, binary monitor slices between calls.  The same source links against
// the HIP library or the CPU library of the same program.
//
//   usage: main T_MAX T_MONITOR OUTDIR [SEED]
//
// Monitor files OUTDIR/monitorX-%06d-%d.txt and monitorY-..., as the
// reference writes them (examples/pearson-3d-main.cpp:116-157): a header of
// six int32 (global position of the lower corner x, y, z, then the extents),
// then FP64 planes: U then V of the x = mid plane (monitorX) or y = mid plane
// (monitorY), rows of z.
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/time.h>

#include GS3D_HEADER

static uint64_t sm_state;
static double urand() {   // SplitMix64 -> U[0,1)
  uint64_t z = (sm_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static double wctime() {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return (double)tv.tv_sec + (double)tv.tv_usec * 1e-6;
}

// u = 1, v = 0 except a centred cube of side N/8 with u = 0.5, v = 0.25;
// uniform noise in [-0.01, 0.01) on u and v everywhere
static void init() {
  const int side_x = NX / 8, side_y = NY / 8, side_z = NZ / 8;
  for (int ix = navi.lower_x; ix < navi.upper_x; ++ix)
    for (int iy = navi.lower_y; iy < navi.upper_y; ++iy)
      for (int iz = navi.lower_z; iz < navi.upper_z; ++iz) {
        const int x = navi.offset_x + ix, y = navi.offset_y + iy, z = navi.offset_z + iz;
        const int in = x >= NX / 2 - side_x / 2 && x < NX / 2 - side_x / 2 + side_x && y >= NY / 2 - side_y / 2 &&
                       y < NY / 2 - side_y / 2 + side_y && z >= NZ / 2 - side_z / 2 &&
                       z < NZ / 2 - side_z / 2 + side_z;
        U[ix][iy][iz] = (in ? 0.5 : 1.0) + 0.01 * (2.0 * urand() - 1.0);
        V[ix][iy][iz] = (in ? 0.25 : 0.0) + 0.01 * (2.0 * urand() - 1.0);
      }
}

static void write_monitor(const char* outdir) {
  int pos[6] = {navi.offset_x + navi.lower_x, navi.offset_y + navi.lower_y, navi.offset_z + navi.lower_z,
                navi.upper_x - navi.lower_x, navi.upper_y - navi.lower_y, navi.upper_z - navi.lower_z};
  const int zs = pos[5];
  char fn[512];
  snprintf(fn, sizeof fn, "%s/monitorX-%06d-%d.txt", outdir, navi.time_step, navi.mpi_my_rank);
  FILE* fp = fopen(fn, "wb");
  if (!fp) { perror(fn); exit(1); }
  fwrite(pos, sizeof(int), 6, fp);
  const int x = navi.lower_x + pos[3] / 2;
  for (int y = navi.lower_y; y < navi.upper_y; ++y) fwrite(U[x][y] + navi.lower_z, sizeof(double), zs, fp);
  for (int y = navi.lower_y; y < navi.upper_y; ++y) fwrite(V[x][y] + navi.lower_z, sizeof(double), zs, fp);
  fclose(fp);
  snprintf(fn, sizeof fn, "%s/monitorY-%06d-%d.txt", outdir, navi.time_step, navi.mpi_my_rank);
  fp = fopen(fn, "wb");
  if (!fp) { perror(fn); exit(1); }
  fwrite(pos, sizeof(int), 6, fp);
  const int y = navi.lower_y + pos[4] / 2;
  for (int xx = navi.lower_x; xx < navi.upper_x; ++xx) fwrite(U[xx][y] + navi.lower_z, sizeof(double), zs, fp);
  for (int xx = navi.lower_x; xx < navi.upper_x; ++xx) fwrite(V[xx][y] + navi.lower_z, sizeof(double), zs, fp);
  fclose(fp);
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s T_MAX T_MONITOR OUTDIR [SEED]\n", argv[0]);
    return 2;
  }
  const int t_max = atoi(argv[1]), t_monitor = atoi(argv[2]);
  const char* outdir = argv[3];
  sm_state = argc > 4 ? strtoull(argv[4], NULL, 10) : 1;
  MPI_Init(&argc, &argv);
  Formura_Init(&navi, MPI_COMM_WORLD);
  init();
  int last_monitor = -t_monitor;
  double t0 = wctime(), elapsed = 0.0;
  for (;;) {
    if (navi.time_step >= last_monitor + t_monitor) {
      write_monitor(outdir);
      last_monitor += t_monitor;
    }
    if (navi.time_step >= t_max) break;
    const double t = wctime();
    Formura_Forward(&navi);
    MPI_Barrier(MPI_COMM_WORLD);
    elapsed += wctime() - t;
  }
  printf("steps %d  Formura_Forward wall %.3f s  %.3f GLUP/s (host arrays, transfers included)\n", navi.time_step,
         elapsed, (double)navi.time_step * NX * NY * NZ / elapsed / 1e9);
  (void)t0;
  MPI_Finalize();
  return 0;
}
