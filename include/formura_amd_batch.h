/* formura_amd_batch.h -- C ABI of a batched step library (one MI355X).
 *
 * A batched library is generated from a 2D `.fmr` program with some of its
 * named scalars (for Gray-Scott: F, k, dt, cu, cv) turned into per-instance
 * parameters.  It advances many independent periodic instances of uneven
 * size together: instances are grouped into size buckets, each bucket is
 * advanced by persistent launches of the fused (T-step) 2D march kernel over
 * a tile list covering all its instances, the buckets run on parallel
 * branches of one hipGraph.
 *
 * None of these entry points has a counterpart in the reference, whose
 * generator emits one fixed-size, fixed-parameter program per compile
 * (src/Formura/MPICxx/Translate.hs:973-1139); a caller of the reference would
 * build and run one program per instance.  All functions return 0 (or a
 * non-negative count / time) on success and -1 on a bad argument; HIP errors
 * abort with a message.  Call from one host thread.
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

int Formura_Batch_Num_Fields(void);   /* state fields per instance */
int Formura_Batch_Num_Params(void);   /* parameters per instance */
int Formura_Batch_Elem_Size(void);    /* 4 (FP32 bulk) or 8 (FP64) */

/* Allocate n instances: instance i has an nx[i] x ny[i] periodic grid (y
 * contiguous) and parameters params[i*np .. i*np+np-1]. */
int Formura_Batch_Init(int n, const int *nx, const int *ny, const double *params, int np);

/* Release every device resource; Formura_Batch_Init may be called again. */
int Formura_Batch_Destroy(void);

/* Writes up to max_variants (T, points per lane) pairs of the compiled march
 * variants into out[2*i], out[2*i+1]; returns the number written. */
int Formura_Batch_List_Variants(int *out, int max_variants);

/* Group the instances into nbuckets buckets (bucket_of_instance[i]) and give
 * bucket b the variant (T, points per lane, march chunk) =
 * variant[3*b .. 3*b+2].  Keeps the instances' current state.  When every
 * bucket has the same T (and at most two lane widths are used), the plan is
 * fused. */
int Formura_Batch_Plan(int nbuckets, const int *bucket_of_instance, const int *variant);

/* Copy field f of instance i from / to a host array of nx*ny doubles. */
int Formura_Batch_Set_State(int i, int f, const double *src);
int Formura_Batch_Get_State(int i, int f, double *dst);

/* Advance every instance by nsteps; use_graph != 0 captures the launch
 * sequence in a hipGraph on first use (per step count and starting parity)
 * and replays it.  With a fused plan (every bucket of one depth T, see
 * Formura_Batch_Fused) each T-block is one persistent launch over the tiles
 * of all instances; otherwise the buckets run on parallel branches.  Returns
 * GPU milliseconds (HIP events around the work). */
float Formura_Batch_Advance(int nsteps, int use_graph);

/* 1 if the current plan is fused -- one kernel launch advances every
 * instance of every bucket by T steps, each tile with its bucket's lane
 * width and march chunk -- with *T its depth and *mask its lane widths
 * (bit 1: one point per lane, 2: two, 4: four); 0 if the buckets' depths
 * differ (per-bucket launches), -1 before Formura_Batch_Init. */
int Formura_Batch_Fused(int *T, int *mask);

/* Number of nodes of the hipGraph captured last by Formura_Batch_Advance
 * (a fused plan: one kernel node per T-block). */
int Formura_Batch_Last_Graph_Nodes(void);

/* Advance only the instances of bucket b (for autotuning); GPU ms. */
float Formura_Batch_Advance_Bucket(int b, int nsteps);

/* The plain method: one single-site-per-thread launch per instance per step.
 * GPU ms. */
float Formura_Batch_Advance_Plain(int nsteps);

/* Device address of field f of instance i (current buffer), or NULL. */
void *Formura_Batch_Device_Field(int i, int f);

/* FP64 comparison of field f of instance i with a device array of the same
 * shape (element size 8 or 4): res[0] = sum d^2, res[1] = sum ref^2,
 * res[2] = max |d|. */
int Formura_Batch_Diff(int i, int f, const void *ref, int ref_elem_size, double *res);

#ifdef __cplusplus
}
#endif
