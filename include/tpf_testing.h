/*
 * tpf_testing.h -- host-only test hooks of libtpf (not part of the product
 * path; no counterpart in the reference).  They run scalar host models of
 * the device algorithms so the precomputed tables can be checked on a
 * machine without a GPU.  Arrays here are HOST pointers.
 */
#ifndef TPF_TESTING_H
#define TPF_TESTING_H

#include "tpf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Scalar model of the CCD line kernel for one axis (same tables, same
 * partitioning) on host arrays of the plan's full shape. */
TPF_API tpf_status tpf_test_ccd_host_model(tpf_ccd_plan* plan, const double* f,
                                           double* d1, double* d2, int axis);

/* Workgroup order of a fused multi-axis CCD launch: for the nslots (<= 3)
 * axis slots with counts[i] workgroups each (bit i of strided_mask: slot i
 * is a strided axis, whose workgroups are renumbered XCD by XCD), writes
 * slot[b] and the slot-local workgroup local[b] of every launch workgroup b. */
TPF_API tpf_status tpf_test_ccd_block_map(int nslots, const unsigned* counts,
                                          unsigned strided_mask, int* slot, unsigned* local);

#ifdef __cplusplus
}
#endif
#endif /* TPF_TESTING_H */
