// Solver entry points with the reference's signatures (src/cdmf.h:4-10).
// All three run on the HIP device through the C ABI of cdmf_hip.h; the
// kernel-file argument of the OpenCL drivers is accepted and ignored.
#ifndef CDMF_HOST_CDMF_H
#define CDMF_HOST_CDMF_H

#include "tools.h"

void cdmf_ocl_01(SparseMatrix& R, MatData& W_c, MatData& H_c, TestData& T, parameter& param, char filename[]);

void cdmf_ocl_02(SparseMatrix& R, MatData& W_c, MatData& H_c, TestData& T, parameter& param, char filename[]);

void cdmf_csr5(SparseMatrix& R, MatData& W, MatData& H, parameter& param, char filename[]);

// Variant-explicit form used by the three above (variant: 1, 2 or 3).
void cdmf_hip(SparseMatrix& R, MatData& W, MatData& H, TestData* T, parameter& param, int variant);

#endif
