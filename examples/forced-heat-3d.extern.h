/* Extern functions of forced-heat-3d.fmr, for the GPU library (host and
 * device) and the CPU library alike.  Only + - * / so that both compute the
 * same bits. */
#pragma once
#ifndef FMR_HD
#define FMR_HD
#endif
static inline FMR_HD double damp(double x) { return x / (1.0 + x * x); }
#ifdef __cplusplus
static inline FMR_HD float damp(float x) { return x / (1.0f + x * x); }   /* FP32 bulk libraries */
#endif
