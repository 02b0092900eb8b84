// Host containers with the public interface of the reference's
// SparseMatrix / TestData (src/pmf_util.h:13-161): both compressed layouts of
// R with u32 pointers and indices and float values, the shallow transpose,
// and the COO test set.  Storage comes from the data layer (cdmf_data.h).
#ifndef CDMF_HOST_PMF_UTIL_H
#define CDMF_HOST_PMF_UTIL_H

#include <memory>
#include <string>
#include <vector>

#include "cdmf_data.h"
#include "pmf.h"

using VecData = std::vector<VALUE_TYPE>;
using MatData = std::vector<VecData>;

class SparseMatrix {
public:
    long rows = 0, cols = 0;
    unsigned long nnz = 0, max_row_nnz = 0, max_col_nnz = 0;

    // Takes the arrays of a loaded or generated data set (ownership moves).
    void adopt(cdmf_dataset& d);

    SparseMatrix get_shallow_transpose() const;

    unsigned* get_csc_col_ptr() const { return csc_col_ptr_.get(); }
    unsigned* get_csc_row_indx() const { return csc_row_indx_.get(); }
    VALUE_TYPE* get_csc_val() const { return csc_val_.get(); }
    unsigned* get_csr_col_indx() const { return csr_col_indx_.get(); }
    unsigned* get_csr_row_ptr() const { return csr_row_ptr_.get(); }
    VALUE_TYPE* get_csr_val() const { return csr_val_.get(); }

private:
    std::shared_ptr<unsigned> csc_col_ptr_, csr_row_ptr_;
    std::shared_ptr<VALUE_TYPE> csr_val_, csc_val_;
    std::shared_ptr<unsigned> csc_row_indx_, csr_col_indx_;
};

class TestData {
public:
    unsigned long rows = 0, cols = 0, nnz = 0;

    void adopt(cdmf_dataset& d);

    unsigned* getTestCol() const { return test_col.get(); }
    unsigned* getTestRow() const { return test_row.get(); }
    VALUE_TYPE* getTestVal() const { return test_val.get(); }

private:
    std::shared_ptr<unsigned> test_row, test_col;
    std::shared_ptr<VALUE_TYPE> test_val;
};

#endif
