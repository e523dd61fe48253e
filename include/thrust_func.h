/*
 * thrust_func.h -- drop-in declarations of the reference's Thrust helpers
 * (p_seg_tree/util/thrust_func.h:4-12, defined in p_seg_tree/thrust_func.cu:19-42),
 * defined in libsegclip (csrc/sc_thrust.hip) on its own HIP radix sort and
 * scan.  C++ linkage and overloads exactly as the reference declares them
 * (REAL = double, LONG = long).  Device variants take HIP device pointers,
 * sort / scan in place on the default stream and return after the work is
 * done; a failure throws std::runtime_error, as Thrust throws
 * thrust::system_error.  thrustDeviceSort / thrustHostSort are declared but
 * never defined by the reference (thrust_func.cu:34-36,43-45), so they are
 * not declared here.
 */
#ifndef SEGCLIP_THRUST_FUNC_H
#define SEGCLIP_THRUST_FUNC_H

void thrustDeviceStableSortByKey(double *arr, int *arrIdx, int size); /* thrust_func.cu:19-21 */
void thrustDeviceStableSortByKey(int *arr, int *arrIdx, int size);    /* thrust_func.cu:22-24 */
void thrustDeviceStableSortByKey(long *arr, int *arrIdx, int size);   /* thrust_func.cu:25-27 */
void thrustDeviceExclusiveScan(int *arr, int size);                   /* thrust_func.cu:28-30 */
void thrustDeviceExclusiveScan(long *arr, int size);                  /* thrust_func.cu:31-33 */
void thrustHostStableSortByKey(double *arr, int *arrIdx, int size);   /* thrust_func.cu:40-42 (host) */

#endif
