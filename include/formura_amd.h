/* formura_amd.h -- C ABI of a formura_amd step library (one MI355X).
 *
 * Every library generated from a `.fmr` program exports two groups of
 * symbols.  Plain C types only: pointers, ints, doubles.
 *
 * 1. The reference ABI, declared per program in the generated `<name>.h`
 *    (layout of src/Formura/MPICxx/Translate.hs:1000-1139):
 *
 *      extern double U[SX][SY][SZ]; ...   state arrays, interior [lower, upper)
 *      extern struct Formura_Navigator navi;
 *      int Formura_Init (struct Formura_Navigator *navi, MPI_Comm comm);
 *          replaces the emitted Formura_Init (MPICxx/Translate.hs:1070-1092):
 *          sets time_step = 0 and lower/upper/offset per axis, and also
 *          allocates the device state.  Returns 0.
 *      int Formura_Forward (struct Formura_Navigator *navi);
 *          replaces the emitted Formura_Forward (MPICxx/Translate.hs:1106-1131):
 *          advances the state arrays by monitorInterval2 steps (the monitor
 *          interval rounded up to a multiple of 2T), adds that to time_step,
 *          returns 0.  Synchronous: it uploads the interior, runs on the GPU
 *          and downloads the interior before returning.  HIP errors abort
 *          with a message.
 *
 * 2. Device-resident entry points declared below.  They have no counterpart
 *    in the reference, whose state lives only in host arrays; they let a
 *    driver keep the state in HBM across many steps (a reference driver
 *    would call Formura_Forward in its time loop instead,
 *    examples/pearson-3d-main.cpp:189-240).
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

struct Formura_Navigator;

/* Copy the interior of the host state arrays to the device state. */
int Formura_Upload(struct Formura_Navigator *navi);

/* Copy the device state into the interior of the host state arrays. */
int Formura_Download(struct Formura_Navigator *navi);

/* Advance the device state by nsteps (>= 0) steps, any count; returns after
 * the GPU has finished.  time_step += nsteps. */
int Formura_Forward_Device(struct Formura_Navigator *navi, int nsteps);

/* As Formura_Forward_Device; returns the GPU time in milliseconds measured
 * with HIP events on the library's stream around the launches. */
float Formura_Forward_Device_Timed(struct Formura_Navigator *navi, int nsteps);

/* Breakdown of the last Formura_Forward_Device_Timed call: the selected
 * kernel's launches run first (main_launches of them, main_ms in total), then
 * the steps left over when nsteps is not a multiple of its fusion depth T. */
int Formura_Last_Timing(float *main_ms, int *main_launches, float *total_ms);

/* Select the kernel: kind 0 = one site per thread (any stencil), kind 2 =
 * the stage kernels of a multi-stage program (the reference's loop nest per
 * manifest node, MPICxx/Translate.hs:671-737; `rows` = wavefronts per
 * workgroup, 4 (also for 0) or 16; other arguments ignored),
 * kind 1 = temporally blocked march with T fused steps, `rows` y-rows per lane,
 * `waves` wavefronts per tile, input planes prefetched `pf` (1 or 2) planes
 * ahead, compiled for at least `occ` waves per SIMD; kinds 3 and 4 = the same
 * march with the x-difference carried from plane to plane and one z-site per
 * lane (3), or two consecutive z-sites per lane (4); kind 5 = kind 3 with
 * (T + 1) / 2 more rows on the first and the last wave of a tile, which skip
 * the ghost rows no later level needs.  Returns 0, or -1 if that variant is not
 * compiled into the library. */
int Formura_Set_Kernel(int kind, int T, int rows, int waves, int pf, int occ);

/* Work split of the march kernel.  lx > 0: x is cut into chunks of lx
 * planes, one workgroup per (chunk, tile); -300 <= lx < -200: x is cut into
 * the most chunks k per tile whose k x tiles workgroups are resident at once
 * with (-lx - 200) waves per SIMD, one workgroup per (chunk, tile) (whole
 * runs); lx < -300: whole runs with (-lx - 300) waves per SIMD, and the
 * resident workgroups they leave idle take the last planes of every tile,
 * the runs shortened so that no workgroup marches longer than needed;
 * -200 <= lx < -100: (-lx - 100) waves per SIMD of the device, x cut into the
 * chunks that give each workgroup one run of its own (balanced runs); other
 * lx < 0: x is kept whole and -lx waves per SIMD of the device share the
 * (tile, plane) items evenly.  Returns 0, or -1 for 0. */
int Formura_Set_March_Chunk(int lx);

/* Number of manifest nodes the stage kernels (kind 2) evaluate per step; 0
 * for a program whose only manifest nodes are the state updates. */
int Formura_Num_Stages(void);

/* Number of stage kernels launched per step: manifest nodes that do not read
 * each other's arrays share one kernel (at most 16 per kernel). */
int Formura_Num_Stage_Kernels(void);

/* Current kernel selection. */
int Formura_Get_Kernel(int *kind, int *T, int *rows, int *waves, int *pf, int *occ, int *lx);

/* Writes up to max_variants (kind, T, rows, waves, pf, occ) tuples of the
 * compiled march variants into out[6*i .. 6*i+5]; returns the number
 * written. */
int Formura_List_Kernels(int *out, int max_variants);

/* on != 0 (the default): the launch sequence of a Formura_Forward_Device*
 * call is captured in a hipGraph once per (kernel, work split, step count,
 * starting buffer) and replayed; 0: launched eagerly (and the graphs are
 * released).  Returns 0. */
int Formura_Set_Graphs(int on);

/* With graphs on: capture and instantiate now the graphs that a following
 * Formura_Forward_Device*(navi, nsteps) from the current state will replay,
 * without running them or changing the state, so that the one-time capture
 * is not part of that call.  Returns 0, or -1 if nsteps < 0.  No counterpart
 * in the reference (its Formura_Forward launches nothing). */
int Formura_Prepare_Device(int nsteps);

/* Per-launch scalar parameters (programs compiled with named parameters);
 * returns 0, or -1 if np exceeds 16. */
int Formura_Set_Params(const double *p, int np);

/* Device address of the current buffer of state field f (in `step` argument
 * order), or NULL.  Valid until the next advance. */
void *Formura_Device_Field(int f);

/* Size in bytes of one device element: 8 (FP64) or 4 (FP32 bulk). */
int Formura_Device_Elem_Size(void);

/* FP64 reduction of device field f: res[0] = sum, res[1] = sum of squares
 * (res must hold 3 doubles).  Deterministic.  Returns 0 or -1. */
int Formura_Reduce_Device(int f, double *res);

/* FP64 comparison of device field f with another device array of the same
 * shape (element size 8 or 4): res[0] = sum (f - ref)^2, res[1] = sum ref^2,
 * res[2] = max |f - ref|.  Returns 0 or -1. */
int Formura_Diff_Device(int f, const void *ref, int ref_elem_size, double *res);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the comparison of Formura_Diff_Device, for two caller-owned
 * device arrays a and ref of n elements each (element sizes 8 or 4),
 * res (DEVICE memory, 3 doubles) = {sum (a - ref)^2, sum ref^2,
 * max |a - ref|}, enqueued on `stream` (a hipStream_t; NULL: the library's
 * stream) and not waited for.  It allocates nothing: its partial sums live
 * in a static 24 KiB device array of the library, one per GPU, held until
 * the process ends, so calls that overlap on one GPU race.  Same kernels
 * and summation order as Formura_Diff_Device: the same bits for the same
 * data.  Returns 0, or -1 for a null pointer, an element size other than 4
 * or 8, n < 0 or n > 2^36 (nothing is enqueued then). */
int Formura_Diff_Arrays_NoAlloc(const void *a, int a_elem_size, const void *ref, int ref_elem_size, long long n,
                                double *res, void *stream);

#ifdef __cplusplus
}
#endif
