// Host tools with the reference's names and behaviour (src/tools.h:19-55):
// loading, initialisation, command line, comparison and RMSE reporting.
#ifndef CDMF_HOST_TOOLS_H
#define CDMF_HOST_TOOLS_H

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "pmf.h"
#include "pmf_util.h"

void load(const char* srcdir, SparseMatrix& R, TestData& T);

void initial_col(MatData& X, unsigned k, unsigned n);

void exit_with_help();

parameter parse_command_line(int argc, char** argv);

void golden_compare(MatData W, MatData W_ref, unsigned k, unsigned m);

void calculate_rmse(const MatData& W_c, const MatData& H_c, const char* srcdir, unsigned k);

double calculate_rmse_directly(MatData& W, MatData& H, TestData& T, unsigned rank, bool ifALS);

bool read_factor_file(const char* fname, MatData& X, unsigned k, long n);

#endif
