// Single translation unit of libgodar_amd.so.
#include "radix_sort.hip"
#include "godar_kernels.hip"
#include "godar_api.hip"
