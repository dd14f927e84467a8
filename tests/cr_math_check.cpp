// Host build of godar_amd/csrc/cr_math.h for tests/test_cr_math.py: the same
// source the device compiles, evaluated on the CPU over arrays.
#include "cr_math.h"

extern "C" {
void crm_exp_n(const double* x, double* y, int n) { for (int i = 0; i < n; ++i) y[i] = crm::cr_exp(x[i]); }
void crm_log_n(const double* x, double* y, int n) { for (int i = 0; i < n; ++i) y[i] = crm::cr_log(x[i]); }
void crm_asin_n(const double* x, double* y, int n) { for (int i = 0; i < n; ++i) y[i] = crm::cr_asin(x[i]); }
void crm_acos_n(const double* x, double* y, int n) { for (int i = 0; i < n; ++i) y[i] = crm::cr_acos(x[i]); }
}
